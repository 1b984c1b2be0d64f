// MI355X (gfx950) kernels of the Davidson sigma step and the C-ABI around them.
//
// One H.C is a list of GEMM pairs (the flattened two-operator TensorMultiply
// of operatorfunctions.C:485-537), grouped by the V block they update:
//   W = alpha_a op(A) C        (first GEMM, operatorfunctions.C:512-516)
//   V += alpha W op(B)^T       (second GEMM, operatorfunctions.C:530)
// k_sigma fuses both: a workgroup owns one V tile and walks a slice of that
// block's pair list in order; for every pair it builds W in LDS chunk by
// chunk (MFMA over the A.C contraction) and immediately contracts it with
// B into the V accumulator, so W never travels to HBM.  Tile classes: basic
// tiles (up to 64 rows, 4 wavefronts of 32 x 32 or 16 x 64), and wide
// 16-row tiles of 256, 384 (4 wavefronts) or 768 columns (8 wavefronts in
// one row), so a wide V block forms each W column once per row panel;
// v_mfma_f64_16x16x4f64 throughout.  k_sigma_bf is the same on pre-split
// BF16 slices (v_mfma_f32_16x16x32_bf16, 256- and 512-wide tiles), k_small
// gives V blocks up to 32 x 32 one wavefront each.  Long pair lists are cut
// into work units whose partial tiles k_reduce adds in list order, so V is
// reduced without atomics and bitwise reproducibly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/block_sigma.h"

namespace {

thread_local std::string g_err;
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
#define HIPCHK(x)                                                                             \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) return fail(BSG_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int TK = 16;  // K depth of one LDS slice

typedef double dbl4 __attribute__((ext_vector_type(4)));

struct TileRef {
  int32_t item, tm, tn, pad;
};

// Several right-hand sides (roots) in one launch: root r reads c + r
// wstride, accumulates into v + r wstride, keeps its partial tiles at part +
// r pstride and its BF16 C slices at cslices + r cstride.
struct RootSet {
  int32_t n;
  int32_t pad;
  int64_t wstride, pstride, cstride;
};
// Workgroup L of a launch over rs.n roots -> (tile t, root): the roots of one
// tile are dispatched next to each other, eight workgroups apart, so tile t
// runs on XCD t mod 8 (workgroup L runs on XCD L mod 8) for every root, and
// the roots share the tile's operator blocks in that XCD's L2.  The launch
// has roundup(ntiles, 8) rs.n workgroups.
__device__ inline int root_of(int L, int nroot, int& t) {
  const int g = L / (8 * nroot), rem = L - g * 8 * nroot;
  t = g * 8 + (rem & 7);
  return rem >> 3;
}



// A work unit: pairs [first, first+count) of one V block.  p_off < 0: the
// unit owns the whole pair list and accumulates into V; otherwise it writes
// its partial tile to P[p_off...] for k_reduce.
struct Seg {
  int32_t vb, first, count, pad;
  int64_t p_off;
};
// Ordered reduction of the partial tiles of one V block slice:
// V[v_off + e] += sum_s P[p_off + s*mn + e] for e in [e0, e0+len), len <= 256.
struct RedItem {
  int64_t v_off, p_off;
  int32_t nseg, mn, e0, len;
};

__global__ __launch_bounds__(256) void k_reduce(const RedItem* __restrict__ items, const double* __restrict__ part,
                                                double* __restrict__ v, RootSet rs) {
  const RedItem it = items[blockIdx.x];
  part += blockIdx.y * rs.pstride;
  v += blockIdx.y * rs.wstride;
  const int e = threadIdx.x;
  if (e >= it.len) return;
  const int64_t x = it.e0 + e;
  double s = 0.0;
  for (int q = 0; q < it.nseg; ++q) s += part[it.p_off + (int64_t)q * it.mn + x];
  v[it.v_off + x] += s;
}

// ---------------------------------------------------------------- FP64 grouped kernel
// One launch covers every V block.  A workgroup of 4 wavefronts owns a V
// panel tile: rows [row0, row0 + TM) and one column group of width JC (the
// whole block when n <= JC), with the waves arranged WR x WC (WR WC = 4),
// each wave owning NR 16-row blocks and PB = 2 16-wide column blocks (wc,
// wc + WC).  Per tile geometry (WR, NR) is chosen by the host per V block:
// TM = 16 NR WR <= 64, JC = 32 WC.  Two waves per SIMD (register budget 256).  Each W block is formed once per tile and
// padding is 16-granular.
//   GEMM 1: W chunk (TM x JC of W) = alpha_a op(A) C: op(A) staged through LDS
//           in 32-deep K chunks (double-buffered, shared by the waves of a
//           wave row), C column blocks straight into MFMA B operands; the
//           next chunk's loads are issued before the current chunk's MFMAs.
//   GEMM 2: V += alpha W B': the W chunk in LDS is the MFMA A operand, B'
//           column blocks go straight into B operands, 32 K deep per batch.
// Operand loads are raw buffer loads: masked lanes use an out-of-range offset
// and read zero, so no select follows a load.
constexpr int PB = 2;      // column blocks per wave
constexpr int PKC = 32;    // K depth of one staged op(A) chunk
constexpr int GTM = 64;    // largest tile height
constexpr int GP = GTM + 1;        // LDS pitch of k-major rows (doubles)
constexpr int GWS = 16 * PB * 4 * 33;  // W chunk doubles: max over geometries of JC (TM + 1)
constexpr uint32_t OOB = 0x80000000u;

__device__ inline __amdgpu_buffer_rsrc_t block_rsrc(const double* p, int64_t n) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n * 8), 0x00020000);
}
__device__ inline double bload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
// voff: per-lane byte offset (OOB for a masked lane); soff: uniform byte offset
__device__ inline double bload2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// ---- split precision: three BF16 slices on the fly ----------------------
// An FP64 operand x is rounded to FP32 (error <= 2^-24 |x|) and split into
// BF16 slices h + m + l that represent that FP32 value exactly (8 + 8 + 8
// significant bits); BF16 keeps FP32's exponent range, so no block scale is
// needed.  A product keeps the six slice products of order <= 2^-16 (i+j <=
// 2), summed smallest first in FP32 by v_mfma_f32_16x16x32_bf16 (exact BF16
// products) for one 32-deep K group, and each group's FP32 result is added
// into FP64 accumulators.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ inline void split_f32(double x, __bf16& h, __bf16& m, __bf16& l) {
  const float f = (float)x;
  h = (__bf16)f;
  const float r = f - (float)h;
  m = (__bf16)r;
  l = (__bf16)(r - (float)m);
}
// acc[r][t] += A(row blocks of this wave, 4-deep K step at k) * b[t] for the
// first nv column blocks and the first nr row blocks; A k-major in LDS.
template <int NRM, int NB>
__device__ inline void mma_step(int nv, int nr, dbl4 (&acc)[NRM][NB], const double* Al, int P, int arow, int k, int li,
                                int lk, const double (&b)[NB]) {
  const double a0 = Al[(k + lk) * P + arow + li];
  double a1 = 0.0;
  if (NRM > 1 && nr > 1) a1 = Al[(k + lk) * P + arow + 16 + li];
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    if (t < nv) {
      acc[0][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b[t], acc[0][t], 0, 0, 0);
      if constexpr (NRM > 1)
        if (nr > 1) acc[1][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b[t], acc[1][t], 0, 0, 0);
    }
  }
}
// S K steps (4 S deep) from k0; steps at or past kend skipped, column slots
// past nv skipped (uniform branches)
template <int NRM, int NB, int S = 4>
__device__ inline void mma_run(int nv, int nr, dbl4 (&acc)[NRM][NB], const double* Al, int P, int arow, int k0, int kend,
                               int li, int lk, const double (&b)[S][NB]) {
  if (nv <= 0) return;
#pragma unroll
  for (int s = 0; s < S; ++s)
    if (k0 + 4 * s < kend) mma_step<NRM, NB>(nv, nr, acc, Al, P, arow, k0 + 4 * s, li, lk, b[s]);
}

// Tile geometry code in TileRef::pad: bits 0-1 log2(WR), bit 2 (NR - 1);
// NRM: row blocks per wave the accumulators hold (NR <= NRM).
// NV: V column blocks per wave (the V tile is 16 NV WC wide); the W chunk of
// GEMM 1 is PB blocks per wave (16 PB WC wide), so a V tile wider than the W
// chunk forms every W column once and contracts it into NV / PB chunks' worth
// of V columns.
// NWV: wavefronts per workgroup; with 8 they form one wave row of 8 columns
// of waves (a 16 x 768 tile for NV = 6 with the registers of the 384-wide
// one), so V blocks up to 768 wide form each W column once per row panel.
template <int NRM, int NV, int PBW = PB, int NWV = 4>
__global__ __launch_bounds__(64 * NWV) __attribute__((amdgpu_waves_per_eu(NV == 4 ? 3 : 2))) void k_sigma(
    const bsg_vblock_t* __restrict__ vbs, const bsg_pair_t* __restrict__ prs, const Seg* __restrict__ segs,
    const TileRef* __restrict__ tiles, int ntiles, const double* __restrict__ arena, const double* __restrict__ c,
    double* __restrict__ v, double* __restrict__ part, RootSet rs) {
  static_assert(NV % PBW == 0, "V tile width is a whole number of W chunks");
  static_assert(NWV == 4 || (NWV == 8 && NRM == 1 && PBW == PB), "eight waves: one wave row, 16-row tiles");
  constexpr int NT = 64 * NWV;
  // GEMM 1 uses two op(A) chunk buffers, GEMM 2 the W chunk; they share the
  // LDS (barriers separate the phases).  Both k-major: X[k * P + row].
  // W chunk: JC (TM + 1) doubles, JC = 16 PBW WC
  constexpr int WSD = NWV == 8 ? 16 * PBW * 8 * 17 : (PBW == PB ? GWS : 16 * PBW * 4 * 17);
  constexpr int LDSD = 2 * PKC * GP > WSD ? 2 * PKC * GP : WSD;
  __shared__ double lds[LDSD];
  double* const As0 = lds;
  double* const As1 = lds + PKC * GP;
  double* const Ws = lds;
  int tid;
  const int root = root_of(blockIdx.x, rs.n, tid);
  if (tid >= ntiles) return;
  c += root * rs.wstride;
  v += root * rs.wstride;
  part += root * rs.pstride;
  const TileRef tr = tiles[tid];
  if (tr.item < 0) return;  // padding of the XCD-interleaved order
  const Seg sg = segs[tr.item];
  const bsg_vblock_t vb = vbs[sg.vb];
  const int lwr = tr.pad & 3, nr = min(((tr.pad >> 2) & 1) + 1, NRM);
  const int WR = NWV == 8 ? 1 : 1 << lwr, WC = NWV == 8 ? 8 : 4 >> lwr;
  const int TM = 16 * nr * WR, JC = 16 * PBW * WC, JV = 16 * NV * WC, P = TM + 1;
  const int row0 = tr.tm * TM, colg = tr.tn * JV;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wr = w / WC, wc = w % WC;
  const int arow = wr * nr * 16;  // first panel row of this wave
  const int li = lane & 15, lk = lane >> 4;
  // row blocks of this wave inside the block (0, 1 or 2)
  int nrw = (vb.m - row0 - arow + 15) / 16;
  nrw = nrw < 0 ? 0 : (nrw > nr ? nr : nrw);
  int nvb = 0;  // V column blocks of this wave that exist
  for (int t = 0; t < NV; ++t)
    if (colg + (wc + WC * t) * 16 < vb.n) nvb = t + 1;
  if (nrw == 0) nvb = 0;
  dbl4 accV[NRM][NV];
#pragma unroll
  for (int r = 0; r < NRM; ++r)
#pragma unroll
    for (int t = 0; t < NV; ++t) accV[r][t] = dbl4{0.0, 0.0, 0.0, 0.0};
  int abuf = 0;
  for (int q = 0; q < sg.count; ++q) {
    const bsg_pair_t pr = prs[sg.first + q];
    const __amdgpu_buffer_rsrc_t rC = block_rsrc(c + pr.c_off, (int64_t)pr.k1 * pr.k2);
    if (pr.a_off < 0 && pr.b_off < 0) {
      // identity A and B': V += alpha C
#pragma unroll
      for (int r = 0; r < NRM; ++r)
#pragma unroll
        for (int t = 0; t < NV; ++t) {
          const int col = colg + (wc + WC * t) * 16 + li;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = row0 + arow + r * 16 + lk + 4 * i;
            const bool ok = r < nr && row < vb.m && col < vb.n;
            accV[r][t][i] += pr.alpha * bload(rC, ok ? (uint32_t)(row * pr.k2 + col) * 8u : OOB);
          }
        }
      continue;
    }
    for (int j0 = 0; j0 < pr.k2; j0 += JC) {
      // identity B': only the W chunks of this tile's V columns
      if (pr.b_off < 0 && (j0 < colg || j0 >= colg + JV)) continue;
      const int jw = min(JC, pr.k2 - j0);        // W columns in this chunk
      int nwb = 0;                               // W column blocks of this wave in the chunk
      for (int t = 0; t < PBW; ++t)
        if ((wc + WC * t) * 16 < jw) nwb = t + 1;
      if (nrw == 0) nwb = 0;
      if (pr.a_off < 0) {
        // identity A: W = alpha C (panel rows of C)
        for (int e = threadIdx.x; e < TM * JC; e += NT) {
          const int j = e % JC, i = e / JC;
          const int gi = row0 + i;
          const bool ok = gi < vb.m && j < jw;
          Ws[j * P + i] = pr.alpha * bload(rC, ok ? (uint32_t)(gi * pr.k2 + j0 + j) * 8u : OOB);
        }
      } else {
        const __amdgpu_buffer_rsrc_t rA = block_rsrc(arena + pr.a_off, (int64_t)vb.m * pr.k1);
        dbl4 accW[NRM][PBW];
#pragma unroll
        for (int r = 0; r < NRM; ++r)
#pragma unroll
          for (int t = 0; t < PBW; ++t) accW[r][t] = dbl4{0.0, 0.0, 0.0, 0.0};
        // Pipeline over 32-deep chunks, each consumed in two 16-deep halves:
        // the op(A) elements of a chunk and the C operands of its first half
        // are loaded during the previous chunk's second half, the C operands
        // of the second half during the first half.  Per-lane offsets are
        // loop-invariant; the K position goes in the uniform offset.  Loads
        // past the end of a block read zero (buffer range check); op(A)
        // elements loaded past k1 inside the block only meet zero C rows.
        // wide tiles are always one 16-row block per wave in one wave row
        constexpr int AE = (NV > PB ? 16 : GTM) * PKC / NT;
        uint32_t aoff[AE];
#pragma unroll
        for (int u = 0; u < AE; ++u) {
          const int e = threadIdx.x + u * NT;
          const int i = pr.a_trans ? e % TM : e / PKC, p = pr.a_trans ? e / TM : e % PKC;
          const int gi = row0 + i;
          aoff[u] = e < TM * PKC ? (uint32_t)(pr.a_trans ? p * vb.m + gi : gi * pr.k1 + p) * 8u : OOB;
        }
        const uint32_t astep = (uint32_t)(pr.a_trans ? vb.m : 1) * 8u;  // bytes per K index
        uint32_t coff[PBW];
#pragma unroll
        for (int t = 0; t < PBW; ++t) {
          const int col = j0 + (wc + WC * t) * 16 + li;
          coff[t] = (t < nwb && col < pr.k2) ? (uint32_t)(lk * pr.k2 + col) * 8u : OOB;
        }
        const uint32_t crow = (uint32_t)pr.k2 * 8u;  // bytes per C row
        // C operands in batches of S1 K steps (8 deep for the 256-wide W
        // chunks: register budget); each batch is loaded while the previous
        // one runs, and the last batch of a chunk overlaps the next chunk's
        // op(A) and first C loads
        constexpr int S1 = PBW > PB ? 2 : 4;
        double an[AE], bn[S1][PBW];
        auto load_a = [&](int k0) {
#pragma unroll
          for (int u = 0; u < AE; ++u)
            if (u * NT < TM * PKC) an[u] = bload2(rA, aoff[u], (uint32_t)k0 * astep);
        };
        auto load_c = [&](int k0, int kh) {  // C rows k0 + kh .. k0 + kh + 4 S1 - 1
#pragma unroll
          for (int s4 = 0; s4 < S1; ++s4)
#pragma unroll
            for (int t = 0; t < PBW; ++t) bn[s4][t] = bload2(rC, coff[t], (uint32_t)(k0 + kh + 4 * s4) * crow);
        };
        load_a(0);
        load_c(0, 0);
        for (int k0 = 0; k0 < pr.k1; k0 += PKC) {
          const int kc = min(PKC, pr.k1 - k0);
          double* Ab = abuf ? As1 : As0;
          abuf ^= 1;
#pragma unroll
          for (int u = 0; u < AE; ++u) {
            const int e = threadIdx.x + u * NT;
            if (u * NT < TM * PKC && e < TM * PKC) {
              const int i = pr.a_trans ? e % TM : e / PKC, p = pr.a_trans ? e / TM : e % PKC;
              Ab[p * P + i] = pr.alpha_a * an[u];
            }
          }
          double bc[S1][PBW];
#pragma unroll
          for (int s4 = 0; s4 < S1; ++s4)
#pragma unroll
            for (int t = 0; t < PBW; ++t) bc[s4][t] = bn[s4][t];
          __syncthreads();
          if constexpr (S1 == 4) {
            // two 16-deep halves
            if (kc > 16) load_c(k0, 16);
            mma_run<NRM, PBW>(nwb, nrw, accW, Ab, P, arow, 0, kc, li, lk, bc);
            if (kc > 16) {
#pragma unroll
              for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
                for (int t = 0; t < PBW; ++t) bc[s4][t] = bn[s4][t];
            }
            if (k0 + PKC < pr.k1) {
              load_a(k0 + PKC);
              load_c(k0 + PKC, 0);
            }
            if (kc > 16) mma_run<NRM, PBW>(nwb, nrw, accW, Ab, P, arow, 16, kc, li, lk, bc);
          } else {
            for (int kh = 0; kh < kc; kh += 4 * S1) {
              const bool more = kh + 4 * S1 < kc;
              if (more) {
                load_c(k0, kh + 4 * S1);
              } else if (k0 + PKC < pr.k1) {
                load_a(k0 + PKC);
                load_c(k0 + PKC, 0);
              }
              mma_run<NRM, PBW, S1>(nwb, nrw, accW, Ab, P, arow, kh, kc, li, lk, bc);
              if (more) {
#pragma unroll
                for (int s4 = 0; s4 < S1; ++s4)
#pragma unroll
                  for (int t = 0; t < PBW; ++t) bc[s4][t] = bn[s4][t];
              }
            }
          }
        }
        __syncthreads();  // op(A) buffers done before W (same LDS) is written
        if (pr.b_off < 0) {
          // identity B': the W blocks of this wave are V blocks t + ch PBW
          // of the tile, ch = the chunk's position in the tile
          const int ch = (j0 - colg) / JC;
#pragma unroll
          for (int cc = 0; cc < NV / PBW; ++cc)
            if (cc == ch)
#pragma unroll
              for (int r = 0; r < NRM; ++r)
#pragma unroll
                for (int t = 0; t < PBW; ++t)
#pragma unroll
                  for (int i = 0; i < 4; ++i) accV[r][cc * PBW + t][i] += pr.alpha * accW[r][t][i];
          continue;
        }
#pragma unroll
        for (int r = 0; r < NRM; ++r)
#pragma unroll
          for (int t = 0; t < PBW; ++t) {
            const int cb = wc + WC * t;
            if (r < nr)
#pragma unroll
              for (int i = 0; i < 4; ++i) Ws[(cb * 16 + li) * P + arow + r * 16 + lk + 4 * i] = pr.alpha * accW[r][t][i];
          }
      }
      __syncthreads();
      // GEMM 2: V[tile] += W chunk * B'[j0:j0+jw, tile columns]; W columns
      // past jw up to the next multiple of 4 are zero.  B' operands of the
      // next batch (16 deep, 8 for the widest tiles: register budget) are
      // loaded while the current one runs.
      {
        constexpr int S2 = NV >= 6 || PBW > PB ? 2 : 4;
        const __amdgpu_buffer_rsrc_t rB = block_rsrc(arena + pr.b_off, (int64_t)vb.n * pr.k2);
        // B'(k, col) at k * ks + col * cs inside the block
        const int ks = pr.b_trans ? 1 : vb.n, cs = pr.b_trans ? pr.k2 : 1;
        // per-lane offsets loop-invariant, K position in the uniform offset;
        // rows past k2 read zero or meet zero W columns
        uint32_t boff[NV];
#pragma unroll
        for (int t = 0; t < NV; ++t) {
          const int col = colg + (wc + WC * t) * 16 + li;
          boff[t] = (t < nvb && col < vb.n) ? (uint32_t)(lk * ks + col * cs) * 8u : OOB;
        }
        const uint32_t bstep = (uint32_t)ks * 8u;
        double bn[S2][NV];
        auto load_b = [&](int kg) {
#pragma unroll
          for (int s4 = 0; s4 < S2; ++s4)
#pragma unroll
            for (int t = 0; t < NV; ++t) bn[s4][t] = bload2(rB, boff[t], (uint32_t)(j0 + kg + 4 * s4) * bstep);
        };
        load_b(0);
        for (int kg = 0; kg < jw; kg += 4 * S2) {
          double bb[S2][NV];
#pragma unroll
          for (int s4 = 0; s4 < S2; ++s4)
#pragma unroll
            for (int t = 0; t < NV; ++t) bb[s4][t] = bn[s4][t];
          if (kg + 4 * S2 < jw) load_b(kg + 4 * S2);
          mma_run<NRM, NV, S2>(nvb, nrw, accV, Ws, P, arow, kg, jw, li, lk, bb);
        }
      }
      __syncthreads();
    }
  }
  const bool direct = sg.p_off < 0;
  double* out = direct ? v + vb.v_off : part + sg.p_off;
#pragma unroll
  for (int r = 0; r < NRM; ++r)
#pragma unroll
    for (int t = 0; t < NV; ++t) {
      const int col = colg + (wc + WC * t) * 16 + li;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = row0 + arow + r * 16 + lk + 4 * i;
        if (r < nr && row < vb.m && col < vb.n) {
          if (direct)
            out[(size_t)row * vb.n + col] += accV[r][t][i];
          else
            out[(size_t)row * vb.n + col] = accV[r][t][i];
        }
      }
    }
}

// ---------------------------------------------------------------- split-precision grouped kernel
// The grouped kernel's tiles and work units with every operand already in
// three BF16 slices (x = h + m + l, the exact FP32 rounding of x):
//   op(A) blocks, row-major with K contiguous   (built once per arena)
//   B' blocks, [V column][K] with K contiguous  (built once per arena)
//   C blocks transposed, [k2][k1]               (built per application)
// each slice padded to a multiple of 8 along K and zero there.  A lane's 8
// K slots are then one 16-byte load per slice (K = 8 lk .. 8 lk + 7 for lane
// (li, lk), the same bijection for both MFMA operands).  GEMM 1 runs in the
// transposed orientation W^T = C^T op(A)^T so that a lane's result registers
// hold 4 consecutive W columns of one W row: they are scaled, split and
// written to LDS as one 8-byte store per slice, the layout GEMM 2 reads as
// its A operand.  Identity-B' pairs keep the direct orientation, whose result
// layout is the V accumulators'.  Each 32-deep K group is 6 slice products
// (i + j <= 2, smallest first) in FP32, added into FP64 accumulators.
typedef int i32x4 __attribute__((ext_vector_type(4)));
struct PairS {
  int64_t a, b, c;  // BF16 element offsets of the op(A), B' and C^T slice 0 (-1: identity)
};
struct SliceItem {
  // element (r, k) of the sliced matrix is src[trans ? k rows + r : r cols + k]
  // for k < cols, zero for cols <= k < pitch; slice s at dst + s rows pitch
  int64_t src, dst;
  int32_t rows, cols, pitch, trans;
};
struct Frag {
  bf16x8 h, m, l;
};
__device__ inline bf16x8 bload8(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ inline __amdgpu_buffer_rsrc_t slice_rsrc(const __bf16* p, int64_t elems) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(elems * 2), 0x00020000);
}
__device__ inline void mma6(dbl4& acc, const Frag& a, const Frag& b) {
  f32x4 x = {0.f, 0.f, 0.f, 0.f};
  x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.l, b.h, x, 0, 0, 0);
  x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, b.m, x, 0, 0, 0);
  x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.l, x, 0, 0, 0);
  x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, b.h, x, 0, 0, 0);
  x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.m, x, 0, 0, 0);
  x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.h, x, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] += (double)x[i];
}

__global__ __launch_bounds__(256) void k_slices(const SliceItem* __restrict__ items, const double* __restrict__ src,
                                                __bf16* __restrict__ dst, int64_t src_stride, int64_t dst_stride) {
  const SliceItem it = items[blockIdx.x];
  src += blockIdx.y * src_stride;
  dst += blockIdx.y * dst_stride;
  const int64_t tot = (int64_t)it.rows * it.pitch;
  for (int64_t e = threadIdx.x; e < tot; e += 256) {
    const int r = (int)(e / it.pitch), k = (int)(e % it.pitch);
    double x = 0.0;
    if (k < it.cols) x = src[it.src + (it.trans ? (int64_t)k * it.rows + r : (int64_t)r * it.cols + k)];
    __bf16 h, m, l;
    split_f32(x, h, m, l);
    dst[it.dst + e] = h;
    dst[it.dst + tot + e] = m;
    dst[it.dst + 2 * tot + e] = l;
  }
}

constexpr int SAP = 40;                 // LDS pitch (BF16) of a staged op(A) row: 32 K + 8
constexpr int SAB = 3 * GTM * SAP;      // one staged op(A) chunk, 3 slices
constexpr int SWMAX = 3 * 64 * 72;      // W chunk slices: max over geometries of 3 TM (JC + 8)

// NRM, NV: as for the FP64 kernel (row blocks per wave the accumulators
// hold, V column blocks per wave).  NWV: wavefronts per workgroup; with 8
// they form one wave row of 8 columns of waves (16 x 512 for NV = 4, the
// registers of the 256-wide tile), so V blocks up to 512 wide form each W
// column once per row panel.
template <int NRM, int NV, int NWV = 4>
__global__ __launch_bounds__(64 * NWV) __attribute__((amdgpu_waves_per_eu(2))) void k_sigma_bf(
    const bsg_vblock_t* __restrict__ vbs, const bsg_pair_t* __restrict__ prs, const PairS* __restrict__ pss,
    const Seg* __restrict__ segs, const TileRef* __restrict__ tiles, int ntiles, const __bf16* __restrict__ slices,
    const __bf16* __restrict__ cslices, const double* __restrict__ c, double* __restrict__ v,
    double* __restrict__ part, RootSet rs) {
  // two op(A) chunk buffers for GEMM 1, the W chunk for GEMM 2 (aliased,
  // barriers separate the phases)
  constexpr int LDSE = 2 * SAB > SWMAX ? 2 * SAB : SWMAX;
  __shared__ __attribute__((aligned(16))) __bf16 lds[LDSE];
  int tid;
  const int root = root_of(blockIdx.x, rs.n, tid);
  if (tid >= ntiles) return;
  c += root * rs.wstride;
  v += root * rs.wstride;
  part += root * rs.pstride;
  cslices += root * rs.cstride;
  const TileRef tr = tiles[tid];
  if (tr.item < 0) return;
  const Seg sg = segs[tr.item];
  const bsg_vblock_t vb = vbs[sg.vb];
  static_assert(NV % PB == 0, "V tile width is a whole number of W chunks");
  static_assert(NWV == 4 || (NWV == 8 && NRM == 1), "eight waves: one wave row, 16-row tiles");
  constexpr int NT = 64 * NWV;
  const int lwr = tr.pad & 3, nr = min(((tr.pad >> 2) & 1) + 1, NRM);
  const int WR = NWV == 8 ? 1 : 1 << lwr, WC = NWV == 8 ? 8 : 4 >> lwr;
  const int TM = 16 * nr * WR, JC = 16 * PB * WC, JV = 16 * NV * WC, WP = JC + 8;
  const int row0 = tr.tm * TM, colg = tr.tn * JV;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wr = w / WC, wc = w % WC;
  const int arow = wr * nr * 16;
  const int li = lane & 15, lk = lane >> 4;
  int nrw = (vb.m - row0 - arow + 15) / 16;
  nrw = nrw < 0 ? 0 : (nrw > nr ? nr : nrw);
  int nvb = 0;
  for (int t = 0; t < NV; ++t)
    if (colg + (wc + WC * t) * 16 < vb.n) nvb = t + 1;
  if (nrw == 0) nvb = 0;
  dbl4 accV[NRM][NV];
#pragma unroll
  for (int r = 0; r < NRM; ++r)
#pragma unroll
    for (int t = 0; t < NV; ++t) accV[r][t] = dbl4{0.0, 0.0, 0.0, 0.0};
  // op(A) staging: piece u (8 BF16) of a chunk is slice u / (4 TM), row
  // (u / 4) % TM, K quarter u % 4; three pieces per thread at most (one on
  // the wide tiles, which are 16 rows)
  constexpr int SPT = ((NV > PB ? 16 : GTM) * 12 + NT - 1) / NT;
  int abuf = 0;
  for (int q = 0; q < sg.count; ++q) {
    const bsg_pair_t pr = prs[sg.first + q];
    const PairS ps = pss[sg.first + q];
    const int k1p = (pr.k1 + 7) & ~7, k2p = (pr.k2 + 7) & ~7;
    if (pr.a_off < 0 && pr.b_off < 0) {  // V += alpha C
      const __amdgpu_buffer_rsrc_t rC = block_rsrc(c + pr.c_off, (int64_t)pr.k1 * pr.k2);
#pragma unroll
      for (int r = 0; r < NRM; ++r)
#pragma unroll
        for (int t = 0; t < NV; ++t) {
          const int col = colg + (wc + WC * t) * 16 + li;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = row0 + arow + r * 16 + 4 * lk + i;
            const bool ok = r < nr && row < vb.m && col < vb.n;
            accV[r][t][i] += pr.alpha * bload(rC, ok ? (uint32_t)(row * pr.k2 + col) * 8u : OOB);
          }
        }
      continue;
    }
    for (int j0 = 0; j0 < pr.k2; j0 += JC) {
      // identity B': only the W chunks of this tile's V columns
      if (pr.b_off < 0 && (j0 < colg || j0 >= colg + JV)) continue;
      const int jw = min(JC, pr.k2 - j0);
      int nwb = 0;
      for (int t = 0; t < PB; ++t)
        if ((wc + WC * t) * 16 < jw) nwb = t + 1;
      if (nrw == 0) nwb = 0;
      if (pr.a_off < 0) {
        const __amdgpu_buffer_rsrc_t rC = block_rsrc(c + pr.c_off, (int64_t)pr.k1 * pr.k2);
        // W = alpha C (panel rows), split into the W slices
        for (int e = threadIdx.x; e < TM * JC; e += NT) {
          const int j = e % JC, i = e / JC;
          const int gi = row0 + i;
          const bool ok = gi < vb.m && j < jw;
          __bf16 h, m, l;
          split_f32(pr.alpha * bload(rC, ok ? (uint32_t)(gi * pr.k2 + j0 + j) * 8u : OOB), h, m, l);
          lds[i * WP + j] = h;
          lds[TM * WP + i * WP + j] = m;
          lds[2 * TM * WP + i * WP + j] = l;
        }
      } else {
        const __amdgpu_buffer_rsrc_t rA = slice_rsrc(slices + ps.a, 3 * (int64_t)vb.m * k1p);
        const __amdgpu_buffer_rsrc_t rC = slice_rsrc(cslices + ps.c, 3 * (int64_t)pr.k2 * k1p);
        const uint32_t aslice = (uint32_t)vb.m * k1p * 2u, cslice = (uint32_t)pr.k2 * k1p * 2u;  // bytes per slice
        dbl4 accW[NRM][PB];
#pragma unroll
        for (int r = 0; r < NRM; ++r)
#pragma unroll
          for (int t = 0; t < PB; ++t) accW[r][t] = dbl4{0.0, 0.0, 0.0, 0.0};
        uint32_t aoff[SPT];
        int alds[SPT], aq[SPT];
#pragma unroll
        for (int x = 0; x < SPT; ++x) {
          const int u = threadIdx.x + NT * x;
          const int s = u / (4 * TM), i = (u >> 2) % TM, kq = u & 3;
          const bool ok = u < 12 * TM && row0 + i < vb.m;
          aoff[x] = ok ? (uint32_t)s * aslice + (uint32_t)((row0 + i) * k1p + 8 * kq) * 2u : OOB;
          alds[x] = u < 12 * TM ? (s * GTM + i) * SAP + 8 * kq : -1;
          aq[x] = 8 * kq;
        }
        uint32_t coff[PB];
#pragma unroll
        for (int t = 0; t < PB; ++t) {
          const int j = j0 + (wc + WC * t) * 16 + li;
          coff[t] = (t < nwb && j < pr.k2) ? (uint32_t)(j * k1p + 8 * lk) * 2u : OOB;
        }
        i32x4 an[SPT];
        Frag cn[PB];
        auto load_a = [&](int k0) {
#pragma unroll
          for (int x = 0; x < SPT; ++x)
            an[x] = __builtin_amdgcn_raw_buffer_load_b128(rA, k0 + aq[x] < k1p ? aoff[x] : OOB, (uint32_t)k0 * 2u, 0);
        };
        auto load_c = [&](int k0) {
          const bool in = k0 + 8 * lk < k1p;
#pragma unroll
          for (int t = 0; t < PB; ++t) {
            const uint32_t o = in ? coff[t] : OOB;
            cn[t].h = bload8(rC, o, (uint32_t)k0 * 2u);
            cn[t].m = bload8(rC, o, cslice + (uint32_t)k0 * 2u);
            cn[t].l = bload8(rC, o, 2u * cslice + (uint32_t)k0 * 2u);
          }
        };
        load_a(0);
        load_c(0);
        for (int k0 = 0; k0 < pr.k1; k0 += PKC) {
          __bf16* Ab = lds + abuf * SAB;
          abuf ^= 1;
#pragma unroll
          for (int x = 0; x < SPT; ++x)
            if (alds[x] >= 0) *(i32x4*)(Ab + alds[x]) = an[x];
          Frag cc[PB];
#pragma unroll
          for (int t = 0; t < PB; ++t) cc[t] = cn[t];
          __syncthreads();
          if (k0 + PKC < pr.k1) {
            load_a(k0 + PKC);
            load_c(k0 + PKC);
          }
          Frag af[NRM];
#pragma unroll
          for (int r = 0; r < NRM; ++r) {
            if (r >= nr) {
              af[r] = af[0];
              continue;
            }
            const int o = (arow + r * 16 + li) * SAP + 8 * lk;
            af[r].h = *(const bf16x8*)(Ab + o);
            af[r].m = *(const bf16x8*)(Ab + GTM * SAP + o);
            af[r].l = *(const bf16x8*)(Ab + 2 * GTM * SAP + o);
          }
#pragma unroll
          for (int t = 0; t < PB; ++t)
            if (t < nwb)
#pragma unroll
              for (int r = 0; r < NRM; ++r)
                if (r < nrw) {
                  if (pr.b_off < 0)
                    mma6(accW[r][t], af[r], cc[t]);  // W block: (row 4 lk + i, col li)
                  else
                    mma6(accW[r][t], cc[t], af[r]);  // W^T block: (W col 4 lk + i, W row li)
                }
        }
        __syncthreads();  // op(A) buffers done before the W slices (same LDS) are written
        const double aw = pr.alpha * pr.alpha_a;
        if (pr.b_off < 0) {
          // identity B': the W blocks of this wave are V blocks t + ch PB
          const int ch = (j0 - colg) / JC;
#pragma unroll
          for (int cc = 0; cc < NV / PB; ++cc)
            if (cc == ch)
#pragma unroll
              for (int r = 0; r < NRM; ++r)
#pragma unroll
                for (int t = 0; t < PB; ++t)
#pragma unroll
                  for (int i = 0; i < 4; ++i) accV[r][cc * PB + t][i] += aw * accW[r][t][i];
          continue;
        }
        // lane holds W(arow + 16 r + li, 16 cb + 4 lk + i), i < 4
#pragma unroll
        for (int r = 0; r < NRM; ++r)
          if (r < nr)
#pragma unroll
            for (int t = 0; t < PB; ++t) {
              const int cb = wc + WC * t;
              __bf16 hs[4], ms[4], ls[4];
#pragma unroll
              for (int i = 0; i < 4; ++i) split_f32(aw * accW[r][t][i], hs[i], ms[i], ls[i]);
              const int o = (arow + r * 16 + li) * WP + cb * 16 + 4 * lk;
              typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
              *(bf16x4*)(lds + o) = bf16x4{hs[0], hs[1], hs[2], hs[3]};
              *(bf16x4*)(lds + TM * WP + o) = bf16x4{ms[0], ms[1], ms[2], ms[3]};
              *(bf16x4*)(lds + 2 * TM * WP + o) = bf16x4{ls[0], ls[1], ls[2], ls[3]};
            }
      }
      __syncthreads();
      // GEMM 2: V[tile] += W chunk B'[j0 + .., tile columns], 32 deep per group
      {
        const __amdgpu_buffer_rsrc_t rB = slice_rsrc(slices + ps.b, 3 * (int64_t)vb.n * k2p);
        const uint32_t bslice = (uint32_t)vb.n * k2p * 2u;
        uint32_t boff[NV];
#pragma unroll
        for (int t = 0; t < NV; ++t) {
          const int col = colg + (wc + WC * t) * 16 + li;
          boff[t] = (t < nvb && col < vb.n) ? (uint32_t)(col * k2p + 8 * lk) * 2u : OOB;
        }
        Frag bn[NV];
        auto load_b = [&](int kg) {
          const bool in = j0 + kg + 8 * lk < k2p;
          const uint32_t so = (uint32_t)(j0 + kg) * 2u;
#pragma unroll
          for (int t = 0; t < NV; ++t) {
            const uint32_t o = in ? boff[t] : OOB;
            bn[t].h = bload8(rB, o, so);
            bn[t].m = bload8(rB, o, bslice + so);
            bn[t].l = bload8(rB, o, 2u * bslice + so);
          }
        };
        load_b(0);
        for (int kg = 0; kg < jw; kg += 32) {
          Frag bc[NV];
#pragma unroll
          for (int t = 0; t < NV; ++t) bc[t] = bn[t];
          if (kg + 32 < jw) load_b(kg + 32);
          Frag af[NRM];
#pragma unroll
          for (int r = 0; r < NRM; ++r) {
            if (r >= nr) {
              af[r] = af[0];
              continue;
            }
            const int o = (arow + r * 16 + li) * WP + kg + 8 * lk;
            af[r].h = *(const bf16x8*)(lds + o);
            af[r].m = *(const bf16x8*)(lds + TM * WP + o);
            af[r].l = *(const bf16x8*)(lds + 2 * TM * WP + o);
          }
#pragma unroll
          for (int r = 0; r < NRM; ++r)
            if (r < nrw)
#pragma unroll
              for (int t = 0; t < NV; ++t)
                if (t < nvb) mma6(accV[r][t], af[r], bc[t]);
        }
      }
      __syncthreads();
    }
  }
  const bool direct = sg.p_off < 0;
  double* out = direct ? v + vb.v_off : part + sg.p_off;
#pragma unroll
  for (int r = 0; r < NRM; ++r)
#pragma unroll
    for (int t = 0; t < NV; ++t) {
      const int col = colg + (wc + WC * t) * 16 + li;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = row0 + arow + r * 16 + 4 * lk + i;
        if (r < nr && row < vb.m && col < vb.n) {
          if (direct)
            out[(size_t)row * vb.n + col] += accV[r][t][i];
          else
            out[(size_t)row * vb.n + col] = accV[r][t][i];
        }
      }
    }
}

// ---------------------------------------------------------------- FP64 small-block kernel
// V blocks of at most 32 x 32 (2 x 2 blocks of 16) are latency-bound in the
// grouped kernel: their pairs are tiny and every pair pays the barriers and
// memory round trips of a cooperative workgroup.  Here each wavefront owns a
// whole small V block for one work unit and walks its pairs alone -- no LDS,
// no barriers -- and a workgroup's 4 waves run 4 unrelated units.  A pair is
// fused in registers by forming W^T = C^T op(A)^T instead of W: register i of
// a v_mfma_f64_16x16x4f64 result holds (lane/16 + 4 i, lane%16), i.e. for W^T
// the element W(lane%16, lane/16 + 4 i), which is exactly the A operand of
// the second product's K step i.  Identity-B pairs use the direct W = op(A) C
// orientation, whose result layout matches the V accumulators.
// Small blocks run in FP64 in both precision modes: their MFMAs are not
// what bounds them (operand latency is), so BF16 slices would only add the
// slicing work, and FP64 is exact.
// SB: 16-blocks per dimension of the V blocks an instantiation takes: SB = 1
// for blocks (and contraction depths) up to 16, whose few registers leave
// room for four waves per SIMD to hide the operand latency, SB = 2 for the
// rest up to 32 x 32 (two waves per SIMD).
template <int SB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SB == 1 ? 4 : 2))) void k_small(
    const bsg_vblock_t* __restrict__ vbs, const bsg_pair_t* __restrict__ prs, const Seg* __restrict__ segs,
    const int4* __restrict__ units, int nunits, const double* __restrict__ arena, const double* __restrict__ c,
    double* __restrict__ v, double* __restrict__ part, RootSet roots) {
  constexpr int KF = 4 * SB;  // 4-deep K steps of a dimension <= 16 SB
  int u;
  const int root = root_of(blockIdx.x, roots.n, u);
  if (u >= nunits) return;
  c += root * roots.wstride;
  v += root * roots.wstride;
  part += root * roots.pstride;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int4 un = units[u];
  const int si = w == 0 ? un.x : (w == 1 ? un.y : (w == 2 ? un.z : un.w));
  if (si < 0) return;  // no unit for this wave (waves never synchronise here)
  const Seg sg = segs[si];
  const bsg_vblock_t vb = vbs[sg.vb];
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  // row of accumulator register i inside its block: lk + 4 i
  const int nrb = (vb.m + 15) / 16, nnb = (vb.n + 15) / 16;
  dbl4 accV[SB][SB];
#pragma unroll
  for (int r = 0; r < SB; ++r)
#pragma unroll
    for (int t = 0; t < SB; ++t) accV[r][t] = dbl4{0.0, 0.0, 0.0, 0.0};
  // The plan sends here only pairs with k1, k2 <= 16 SB and a non-identity A,
  // or identity A and B' (the rest of a small block's pairs run in k_sigma).
  // The next pair's descriptor is requested before this pair's operands, so
  // its latency hides behind this pair's loads and MFMAs (units hold at
  // least one pair).
  bsg_pair_t nxt = prs[sg.first];
  for (int q = 0; q < sg.count; ++q) {
    const bsg_pair_t pr = nxt;
    if (q + 1 < sg.count) nxt = prs[sg.first + q + 1];
    const __amdgpu_buffer_rsrc_t rC = block_rsrc(c + pr.c_off, (int64_t)pr.k1 * pr.k2);
    const uint32_t crow = (uint32_t)pr.k2 * 8u;
    if (pr.a_off < 0) {  // identity A and B': V += alpha C
#pragma unroll
      for (int r = 0; r < SB; ++r)
#pragma unroll
        for (int t = 0; t < SB; ++t)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = r * 16 + lk + 4 * i, col = t * 16 + li;
            const bool ok = row < vb.m && col < vb.n;
            accV[r][t][i] += pr.alpha * bload(rC, ok ? (uint32_t)(row * pr.k2 + col) * 8u : OOB);
          }
      continue;
    }
    const __amdgpu_buffer_rsrc_t rA = block_rsrc(arena + pr.a_off, (int64_t)vb.m * pr.k1);
    // op(A)(row, k) at row * ar + k * ak inside the block; rows past m masked,
    // k past k1 meets zero C rows (out of the C block's range).  Every operand
    // of the pair is requested before the first MFMA.
    const int nk1 = (pr.k1 + 3) / 4;
    const int ar = pr.a_trans ? 1 : pr.k1, ak = pr.a_trans ? vb.m : 1;
    const double aw = pr.alpha_a * pr.alpha;
    double a[SB][KF];  // a[r][s]: op(A)(r*16 + li, 4 s + lk)
#pragma unroll
    for (int r = 0; r < SB; ++r) {
      const uint32_t off = (r < nrb && r * 16 + li < vb.m) ? (uint32_t)((r * 16 + li) * ar + lk * ak) * 8u : OOB;
#pragma unroll
      for (int s4 = 0; s4 < KF; ++s4) a[r][s4] = s4 < nk1 ? bload2(rA, off, (uint32_t)(s4 * ak) * 32u) : 0.0;
    }
    // C operands: ct[jb][s] = C(4 s + lk, jb*16 + li), zero past k2
    const int njb = (pr.k2 + 15) / 16;
    double ct[SB][KF];
#pragma unroll
    for (int jb = 0; jb < SB; ++jb) {
      const int col = jb * 16 + li;
      const uint32_t coff = (jb < njb && col < pr.k2) ? (uint32_t)(lk * pr.k2 + col) * 8u : OOB;
#pragma unroll
      for (int s4 = 0; s4 < KF; ++s4) ct[jb][s4] = (jb < njb && s4 < nk1) ? bload2(rC, coff, (uint32_t)(4 * s4) * crow) : 0.0;
    }
    if (pr.b_off < 0) {
      // identity B': V(:, block t) += alpha alpha_a op(A) C(:, block t) (k2 = n)
#pragma unroll
      for (int t = 0; t < SB; ++t) {
        if (t >= nnb) continue;
        dbl4 accW[SB][1];
#pragma unroll
        for (int r = 0; r < SB; ++r) accW[r][0] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < KF; ++s4)
          if (s4 < nk1)
#pragma unroll
            for (int r = 0; r < SB; ++r)
              if (r < nrb) accW[r][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r][s4], ct[t][s4], accW[r][0], 0, 0, 0);
#pragma unroll
        for (int r = 0; r < SB; ++r)
#pragma unroll
          for (int i = 0; i < 4; ++i) accV[r][t][i] += aw * accW[r][0][i];
      }
      continue;
    }
    const __amdgpu_buffer_rsrc_t rB = block_rsrc(arena + pr.b_off, (int64_t)vb.n * pr.k2);
    const int ks = pr.b_trans ? 1 : vb.n, cs = pr.b_trans ? pr.k2 : 1;
    double b[SB][4][SB];  // b[jb][i][t]: B'(jb*16 + 4 i + lk, t*16 + li)
#pragma unroll
    for (int t = 0; t < SB; ++t) {
      const uint32_t boff = (t < nnb && t * 16 + li < vb.n) ? (uint32_t)(lk * ks + (t * 16 + li) * cs) * 8u : OOB;
#pragma unroll
      for (int jb = 0; jb < SB; ++jb)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          b[jb][i][t] = jb < njb ? bload2(rB, boff, (uint32_t)((jb * 4 + i) * ks) * 32u) : 0.0;
    }
#pragma unroll
    for (int jb = 0; jb < SB; ++jb) {
      if (jb >= njb) continue;
      // W^T block row jb: register i of lane (li, lk) is W(li, jb*16 + lk + 4 i)
      dbl4 accT[SB];
#pragma unroll
      for (int r = 0; r < SB; ++r) accT[r] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s4 = 0; s4 < KF; ++s4)
        if (s4 < nk1)
#pragma unroll
          for (int r = 0; r < SB; ++r)
            if (r < nrb) accT[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(ct[jb][s4], a[r][s4], accT[r], 0, 0, 0);
      // GEMM 2 over K = jb*16 + 4 i + lk, A operand = register i of accT
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < SB; ++r)
#pragma unroll
          for (int t = 0; t < SB; ++t)
            if (r < nrb && t < nnb)
              accV[r][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(aw * accT[r][i], b[jb][i][t], accV[r][t], 0, 0, 0);
    }
  }
  const bool direct = sg.p_off < 0;
  double* out = direct ? v + vb.v_off : part + sg.p_off;
#pragma unroll
  for (int r = 0; r < SB; ++r)
#pragma unroll
    for (int t = 0; t < SB; ++t) {
      const int col = t * 16 + li;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = r * 16 + lk + 4 * i;
        if (row < vb.m && col < vb.n) {
          if (direct)
            out[(size_t)row * vb.n + col] += accV[r][t][i];
          else
            out[(size_t)row * vb.n + col] = accV[r][t][i];
        }
      }
    }
}

// ---------------------------------------------------------------- FP64 per-wave tiles
// Medium V blocks are latency-bound in the grouped kernel too: a pair's
// MFMAs are few next to its operand round trips, and a 4-wave tile runs one
// chain of round trips for 4 waves.  Here, as in the small-block kernel, each
// wavefront owns one tile of a V block (16 rows x 16 NB columns) for one work
// unit and walks the unit's pairs alone -- no LDS, no barriers, four
// unrelated tiles per workgroup -- so a CU keeps four times as many operand
// streams in flight.  A pair's op(A) row panel (16 rows x k1, k1 <= 4 KA)
// sits in registers; W^T = C^T op(A)^T is formed 16 W columns at a time and
// its result registers are at once GEMM 2's A operand against the matching
// 16 rows of B'.  On the 32-wide tiles (NB = 2) the C operands of the next
// 16 W columns are requested before the MFMAs of the current ones (two
// register sets, used in turn); the 64-wide tiles (NB = 4, which form each
// W column once for twice the V columns) spend that room on B' and keep one
// set.  The B' operands of a column group are requested before its GEMM 1
// MFMAs, and the next pair's descriptor before this pair's operands.
// Identity-B' pairs use the direct orientation W = op(A) C, whose result
// layout is the V accumulators'; identity-A pairs load W = alpha C straight
// into GEMM 2's operand layout.  Operands past a block's rows, columns or
// contraction depth read as zero (buffer range checks and masked lanes), and
// every C column past k2 is masked, so B' rows past k2 only meet zero W.
// 4-deep K steps of op(A) held in registers: k1 <= 64 on the 32-wide tiles,
// k1 <= 80 on the 64-wide ones (their single C register set leaves room)
constexpr int wave_ka(int nb) { return nb == 4 ? 20 : 16; }
template <int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_wave(
    const bsg_vblock_t* __restrict__ vbs, const bsg_pair_t* __restrict__ prs, const Seg* __restrict__ segs,
    const TileRef* __restrict__ tiles, int ntiles, const double* __restrict__ arena, const double* __restrict__ c,
    double* __restrict__ v, double* __restrict__ part, RootSet rs) {
  constexpr int KA = wave_ka(NB);
  int g;
  const int root = root_of(blockIdx.x, rs.n, g);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tid = g * 4 + w;
  if (tid >= ntiles) return;
  const TileRef tr = tiles[tid];
  if (tr.item < 0) return;  // padding: no tile for this wave (waves never synchronise here)
  c += root * rs.wstride;
  v += root * rs.wstride;
  part += root * rs.pstride;
  const Seg sg = segs[tr.item];
  const bsg_vblock_t vb = vbs[sg.vb];
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int row0 = tr.tm * 16, col0 = tr.tn * 16 * NB;
  const bool rowok = row0 + li < vb.m;  // op(A) / W row li of the tile inside the block
  dbl4 accV[NB];  // register i of lane (li, lk): V(row0 + lk + 4 i, col0 + 16 t + li)
#pragma unroll
  for (int t = 0; t < NB; ++t) accV[t] = dbl4{0.0, 0.0, 0.0, 0.0};
  bsg_pair_t nxt = prs[sg.first];
  for (int q = 0; q < sg.count; ++q) {
    const bsg_pair_t pr = nxt;
    if (q + 1 < sg.count) nxt = prs[sg.first + q + 1];
    const __amdgpu_buffer_rsrc_t rC = block_rsrc(c + pr.c_off, (int64_t)pr.k1 * pr.k2);
    const uint32_t crow = (uint32_t)pr.k2 * 8u;
    if (pr.a_off < 0 && pr.b_off < 0) {  // identity A and B': V += alpha C
#pragma unroll
      for (int t = 0; t < NB; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = row0 + lk + 4 * i, col = col0 + t * 16 + li;
          const bool ok = row < vb.m && col < vb.n;
          accV[t][i] += pr.alpha * bload(rC, ok ? (uint32_t)(row * pr.k2 + col) * 8u : OOB);
        }
      continue;
    }
    const int nk1 = (pr.k1 + 3) / 4;
    double a[KA];  // a[s] = op(A)(row0 + li, 4 s + lk)
    if (pr.a_off >= 0) {
      const __amdgpu_buffer_rsrc_t rA = block_rsrc(arena + pr.a_off, (int64_t)vb.m * pr.k1);
      const int ar = pr.a_trans ? 1 : pr.k1, ak = pr.a_trans ? vb.m : 1;
      const uint32_t off = rowok ? (uint32_t)((row0 + li) * ar + lk * ak) * 8u : OOB;
#pragma unroll
      for (int s = 0; s < KA; ++s) a[s] = bload2(rA, s < nk1 ? off : OOB, (uint32_t)(s * ak) * 32u);
    }
    const double aw = pr.alpha_a * pr.alpha;
    if (pr.b_off < 0) {
      // identity B' (k2 = n): V(:, block t) += alpha alpha_a op(A) C(:, block t)
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const int col = col0 + t * 16 + li;
        if (col0 + t * 16 >= vb.n) break;
        const uint32_t coff = col < pr.k2 ? (uint32_t)(lk * pr.k2 + col) * 8u : OOB;
        double ct[KA];
#pragma unroll
        for (int s = 0; s < KA; ++s) ct[s] = bload2(rC, s < nk1 ? coff : OOB, (uint32_t)(4 * s) * crow);
        dbl4 accW = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < KA; s4 += 4)
          if (s4 < nk1)
#pragma unroll
            for (int s = s4; s < s4 + 4; ++s) accW = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], ct[s], accW, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) accV[t][i] += aw * accW[i];
      }
      continue;
    }
    // B'(k, col) at k ks + col cs inside the block; per-lane offsets of the
    // tile's columns, the K position in the uniform offset
    const __amdgpu_buffer_rsrc_t rB = block_rsrc(arena + pr.b_off, (int64_t)vb.n * pr.k2);
    const int ks = pr.b_trans ? 1 : vb.n, cs = pr.b_trans ? pr.k2 : 1;
    uint32_t boff[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const int col = col0 + t * 16 + li;
      boff[t] = col < vb.n ? (uint32_t)(lk * ks + col * cs) * 8u : OOB;
    }
    const int njb = (pr.k2 + 15) / 16;
    if (pr.a_off < 0) {
      // identity A (k1 = m): W(row0 + li, 16 jb + 4 i + lk) = alpha C(...)
      for (int jb = 0; jb < njb; ++jb) {
        double wa[4], b[4][NB];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int k = jb * 16 + 4 * i + lk;
          wa[i] = pr.alpha * bload(rC, rowok && k < pr.k2 ? (uint32_t)((row0 + li) * pr.k2 + k) * 8u : OOB);
#pragma unroll
          for (int t = 0; t < NB; ++t) b[i][t] = bload2(rB, boff[t], (uint32_t)((jb * 16 + 4 * i) * ks) * 8u);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int t = 0; t < NB; ++t) accV[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(wa[i], b[i][t], accV[t], 0, 0, 0);
      }
      continue;
    }
    // W^T 16 columns at a time: ct[s] = C(4 s + lk, 16 jb + li) (columns past
    // k2 masked), b[i][t] = B'(16 jb + 4 i + lk, col0 + 16 t + li); the B'
    // operands of a column group are requested before its GEMM 1 MFMAs
    auto load_jb = [&](int jb, double (&ct)[KA]) {
      const int col = jb * 16 + li;
      const uint32_t coff = col < pr.k2 ? (uint32_t)(lk * pr.k2 + col) * 8u : OOB;
#pragma unroll
      for (int s = 0; s < KA; ++s) ct[s] = bload2(rC, s < nk1 ? coff : OOB, (uint32_t)(4 * s) * crow);
    };
    auto run_jb = [&](int jb, const double (&ct)[KA]) {
      double b[4][NB];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int t = 0; t < NB; ++t) b[i][t] = bload2(rB, boff[t], (uint32_t)((jb * 16 + 4 * i) * ks) * 8u);
      // register i of lane (li, lk): W(row0 + li, 16 jb + lk + 4 i)
      dbl4 accT = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s4 = 0; s4 < KA; s4 += 4)
        if (s4 < nk1)
#pragma unroll
          for (int s = s4; s < s4 + 4; ++s) accT = __builtin_amdgcn_mfma_f64_16x16x4f64(ct[s], a[s], accT, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double wv = aw * accT[i];
#pragma unroll
        for (int t = 0; t < NB; ++t) accV[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(wv, b[i][t], accV[t], 0, 0, 0);
      }
    };
    if constexpr (NB <= 2) {
      double ct0[KA], ct1[KA];
      load_jb(0, ct0);
      for (int jb = 0; jb < njb; jb += 2) {
        if (jb + 1 < njb) load_jb(jb + 1, ct1);
        run_jb(jb, ct0);
        if (jb + 1 >= njb) break;
        if (jb + 2 < njb) load_jb(jb + 2, ct0);
        run_jb(jb + 1, ct1);
      }
    } else {  // one register set: the wider tile's B' operands take the room
      for (int jb = 0; jb < njb; ++jb) {
        double ct[KA];
        load_jb(jb, ct);
        run_jb(jb, ct);
      }
    }
  }
  const bool direct = sg.p_off < 0;
  double* out = direct ? v + vb.v_off : part + sg.p_off;
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int col = col0 + t * 16 + li;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = row0 + lk + 4 * i;
      if (row < vb.m && col < vb.n) {
        if (direct)
          out[(size_t)row * vb.n + col] += accV[t][i];
        else
          out[(size_t)row * vb.n + col] = accV[t][i];
      }
    }
  }
}



// Tile geometry of the grouped kernels for an m x n V block: waves WR x WC
// (WR WC = 4), NR 16-row blocks per wave (TM = 16 NR WR <= 64 rows), a W
// chunk of JC = 16 PB WC columns and a V tile of JV = 16 NV WC columns.
// NV = PB is the basic tile; the wide tiles (WR = 1, NR = 1, NV = 4 or 6)
// cover up to 256 or 384 V columns, so every W column of a pair is formed
// once per row panel instead of once per column group.  cls: the FP64
// kernel instantiation (0: <2, 2>, 1: <1, 4>, 2: <1, 6>); the split-precision
// kernel runs every tile class as basic tiles (JV = JC).  Chosen to minimise
// the MFMA issue time of the block's tiles (per-wave MFMAs per pair, summed
// over tiles), which counts row padding, idle column slots and GEMM 1
// repeated per column group.
constexpr int NCLS = 4;
struct Geom {
  int lwr, nr, nv, tm, jc, jv, code, cls;
  int wc = 4, nw = 4;  // wave columns, waves per workgroup
};
// MFMAs of all waves of all tiles of a V block, in units of a 4-wave tile's
inline long geom_cost(const Geom& g, long m, long n, long k) {
  auto cdiv = [](long a, long b) { return (a + b - 1) / b; };
  const int wc = g.wc;
  const long panels = cdiv(m, g.tm), groups = cdiv(n, g.jv);
  const long wblk = cdiv(cdiv(k, 16), wc);  // W column blocks per wave (all chunks)
  const long vblk = std::max<long>(1, std::min<long>(g.nv, cdiv(std::min<long>(n, g.jv), 16 * wc)));
  return panels * groups * g.nr * (wblk + vblk) * cdiv(k, 4) * g.nw / 4;
}
inline Geom make_geom(int lwr, int nr, int nv) {
  Geom g;
  g.lwr = lwr;
  g.nr = nr;
  g.nv = nv;
  g.tm = 16 * nr * (1 << lwr);
  g.jc = 16 * PB * (4 >> lwr);
  g.jv = 16 * nv * (4 >> lwr);
  g.code = lwr | ((nr - 1) << 2);
  g.cls = nv / 2 - 1;
  g.wc = 4 >> lwr;
  return g;
}
// the 8-wave tiles: 16 rows, 8 wave columns of nv V column blocks (768 wide
// for the FP64 kernel's nv = 6, tile class 3; 512 for the split kernel's
// nv = 4, split tile class 2)
inline Geom make_geom_w8(int nv) {
  Geom g = make_geom(0, 1, nv);
  g.wc = 8;
  g.nw = 8;
  g.jc = 16 * PB * 8;
  g.jv = 16 * nv * 8;
  g.code = nv == 6 ? 8 : 16;
  g.cls = nv == 6 ? 3 : 2;
  return g;
}
// Measured tile choices (bsg_set_tile_table): for V blocks with m in
// [m_lo, m_hi] and n in [n_lo, n_hi], the index of the geometry to use in
// pick_geom's candidate list (FP64 and split-precision lists; -1: the
// model's choice).  Set before plans are created; read when they are.
struct TileRule {
  int32_t m_lo, m_hi, n_lo, n_hi, fp64, split;
};
std::vector<TileRule> g_tile_rules;
unsigned g_tile_version = 0;  // bumped by bsg_set_tile_table (pick_geom's cache key)
std::vector<Geom> geom_candidates(int max_nv) {
  std::vector<Geom> cand;
  for (int lwr = 0; lwr <= 2; ++lwr)
    for (int nr = 1; nr <= 2; ++nr) cand.push_back(make_geom(lwr, nr, PB));
  for (int nv = 4; nv <= max_nv; nv += 2) cand.push_back(make_geom(0, 1, nv));
  cand.push_back(make_geom_w8(max_nv >= 6 ? 6 : 4));
  return cand;
}
Geom pick_geom_uncached(int m, int n, int max_nv);
// pick_geom per V block shape, remembered per thread (plans repeat shapes)
inline Geom pick_geom(int m, int n, int max_nv) {
  thread_local std::unordered_map<uint64_t, Geom> cache;
  thread_local unsigned version = ~0u;
  if (version != g_tile_version) {
    cache.clear();
    version = g_tile_version;
  }
  const uint64_t key = ((uint64_t)(uint32_t)m << 33) | ((uint64_t)(uint32_t)n << 1) | (max_nv >= 6 ? 1u : 0u);
  auto f = cache.find(key);
  if (f != cache.end()) return f->second;
  const Geom g = pick_geom_uncached(m, n, max_nv);
  cache.emplace(key, g);
  return g;
}
Geom pick_geom_uncached(int m, int n, int max_nv) {
  // typical contraction depth: the block's own dimensions
  const long k = std::max(m, n);
  // the two candidate lists, built once (pick_geom runs per V block)
  static const std::vector<Geom> cand_fp64 = geom_candidates(6), cand_split = geom_candidates(4);
  const std::vector<Geom>& cand = max_nv >= 6 ? cand_fp64 : cand_split;
  for (const TileRule& r : g_tile_rules)
    if (m >= r.m_lo && m <= r.m_hi && n >= r.n_lo && n <= r.n_hi) {
      const int idx = max_nv >= 6 ? r.fp64 : r.split;
      if (idx >= 0 && idx < (int)cand.size() && cand[idx].tm <= GTM) return cand[idx];
      break;
    }
  Geom best = cand[0];
  long bc = -1;
  for (const Geom& g : cand) {
    if (g.tm > GTM) continue;
    const long cst = geom_cost(g, m, n, k);
    if (bc < 0 || cst < bc || (cst == bc && g.tm > best.tm)) {
      bc = cst;
      best = g;
    }
  }
  return best;
}
// MFMA instructions of one pair in the small-block kernel plus a latency
// allowance (segment cost; the wave works alone)
inline long small_cost(int m, int n, const bsg_pair_t& pr) {
  auto cdiv = [](long a, long b) { return (a + b - 1) / b; };
  const long nrb = cdiv(m, 16), nnb = cdiv(n, 16), njb = cdiv(pr.k2, 16);
  const long g1 = pr.a_off < 0 ? 0 : (pr.b_off < 0 ? nnb : njb) * cdiv(pr.k1, 4) * nrb;
  const long g2 = pr.b_off < 0 ? 0 : njb * 4 * nrb * nnb;
  return g1 + g2 + 24;
}
// Per-wave tiles: 16 x 32 for V blocks up to 32 columns, 16 x 64 above
// (wave_nb).  In the FP64 column of the tile table the candidate index after
// the geometries selects them for a V block's pairs within their contraction
// depth (k1 <= 4 wave_ka, or identity A); the block's other pairs run on the
// geometry the model picks.  Without a rule, blocks stay on the grouped
// kernel.
inline int wave_nb(int n) { return n <= 32 ? 2 : 4; }
inline int wave_candidate() {
  static const int n = (int)geom_candidates(6).size();
  return n;
}
inline bool wave_block(int m, int n) {
  if (std::max(m, n) <= 32) return false;  // the small-block kernel's
  for (const TileRule& r : g_tile_rules)
    if (m >= r.m_lo && m <= r.m_hi && n >= r.n_lo && n <= r.n_hi) return r.fp64 == wave_candidate();
  return false;
}
// MFMA instructions of one pair in a per-wave tile plus a latency allowance
// (segment cost)
inline long wave_cost(int n, const bsg_pair_t& pr) {
  auto cdiv = [](long a, long b) { return (a + b - 1) / b; };
  const long nnb = std::min<long>(wave_nb(n), cdiv(n, 16)), njb = cdiv(pr.k2, 16), nk1 = cdiv(pr.k1, 4);
  if (pr.a_off < 0) return (pr.b_off < 0 ? 0 : njb * 4 * nnb) + 24;
  return (pr.b_off < 0 ? nnb * nk1 : njb * (nk1 + 4 * nnb)) + 24;
}
// MFMA instructions per wavefront of one pair in a tile (segment cost)
inline long pair_cost(const Geom& g, int n, const bsg_pair_t& pr) {
  auto cdiv = [](long a, long b) { return (a + b - 1) / b; };
  const int wc = g.wc;
  const long kw = pr.b_off < 0 ? std::min<long>(pr.k2, g.jv) : pr.k2;  // W columns formed per tile
  const long nch = cdiv(kw, g.jc);
  const long g1 = pr.a_off < 0 ? 0 : g.nr * cdiv(cdiv(kw, 16), wc) * cdiv(pr.k1, 4);
  const long g2 =
      pr.b_off < 0 ? 0 : g.nr * std::min<long>(g.nv, cdiv(cdiv(std::min<long>(n, g.jv), 16), wc)) * cdiv(pr.k2, 4);
  return g1 + g2 + 16 * nch;
}

// Launch order of the grouped kernel's tiles.  Workgroup i runs on XCD
// i mod 8, each XCD with its own L2: the tiles of one work unit share their
// C and B' blocks, so a unit's tiles go to one XCD, units are spread over
// the XCDs by cost (costliest first, to the least loaded XCD), each XCD's
// queue runs costliest units first, and the queues are interleaved
// (padding tiles with item -1 keep the positions aligned).
constexpr int NXCD = 8;
// a work unit's tiles: segment seg, ntm x ntn tiles of geometry code
struct GUnit {
  long cost = 0;
  int seg = 0, ntm = 0, ntn = 0, code = 0;
  long ntiles() const { return (long)ntm * ntn; }
};
inline std::vector<TileRef> xcd_order(std::vector<GUnit>& units) {
  // (work, index) pairs sorted by work descending, index ascending on ties
  std::vector<std::pair<long, int>> ord(units.size());
  for (size_t i = 0; i < ord.size(); ++i) ord[i] = {units[i].cost * units[i].ntiles(), (int)i};
  std::sort(ord.begin(), ord.end(), [](const std::pair<long, int>& a, const std::pair<long, int>& b) {
    return a.first != b.first ? a.first > b.first : a.second < b.second;
  });
  std::vector<std::vector<TileRef>> q(NXCD);
  for (auto& v : q) v.reserve(units.size() / NXCD + 16);
  long load[NXCD] = {};
  for (const auto& o : ord) {
    int x = 0;
    for (int y = 1; y < NXCD; ++y)
      if (load[y] < load[x]) x = y;
    const GUnit& g = units[o.second];
    load[x] += o.first;
    for (int tm = 0; tm < g.ntm; ++tm)
      for (int tn = 0; tn < g.ntn; ++tn) q[x].push_back({g.seg, tm, tn, g.code});
  }
  size_t len = 0;
  for (auto& v : q) len = std::max(len, v.size());
  std::vector<TileRef> out;
  out.reserve(len * NXCD);
  for (size_t i = 0; i < len; ++i)
    for (int x = 0; x < NXCD; ++x) out.push_back(i < q[x].size() ? q[x][i] : TileRef{-1, 0, 0, 0});
  return out;
}

// ---------------------------------------------------------------- vector kernels
constexpr int RB = 256, RMAXB = 1024;

__global__ __launch_bounds__(RB) void k_dot_partial(const double* __restrict__ x, const double* __restrict__ y,
                                                    int64_t n, double* __restrict__ part) {
  __shared__ double sh[RB];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) s += x[i] * y[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(RB) void k_sum_partial(const double* __restrict__ part, int np, double* __restrict__ out) {
  __shared__ double sh[RB];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += RB) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}
__global__ void k_axpy(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}
__global__ void k_scal(double a, double* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= a;
}
// the same with the scalar in device memory: x *= 1/sqrt(s[0]) (only if
// s[0] > floor), y += -s[0] x
__global__ void k_scal_rsqrt_dev(const double* __restrict__ s, double floor, double* __restrict__ x, int64_t n) {
  const double n2 = s[0];
  if (!(n2 > floor)) return;
  const double a = 1.0 / sqrt(n2);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= a;
}
// y -= (num[0] / den[0]) x (Olsen's step with both dot products on the device)
__global__ void k_axpy_ratio_dev(const double* __restrict__ num, const double* __restrict__ den,
                                 const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  const double a = -(num[0] / den[0]);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}
__global__ void k_axpy_neg_dev(const double* __restrict__ s, const double* __restrict__ x, double* __restrict__ y,
                               int64_t n) {
  const double a = -s[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}
__global__ void k_precond(double* __restrict__ x, const double* __restrict__ d, double e, double shift, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double den = e - d[i];
    if (fabs(den) > 1e-12) x[i] /= (den + shift);
  }
}
__global__ void k_copy(const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}
// one workgroup per chunk of at most COPY_CHUNK doubles
constexpr int64_t COPY_CHUNK = 16384;
__global__ __launch_bounds__(256) void k_copy_chunks(const bsg_copy_seg_t* __restrict__ chunks) {
  const bsg_copy_seg_t c = chunks[blockIdx.x];
  for (int64_t i = threadIdx.x; i < c.n; i += 256) c.dst[i] = c.src[i];
}

// Composite-operator build: groups of scaled block copies.  Group g owns the
// destination rectangle dst[r ld + c] (r < rows, c < cols) and the run of
// sources [first, first + count); its elements are numbered from
// estart[g] in one flat index space, 256 per workgroup, so small rectangles
// share workgroups.  Each element sums its sources in list order,
// (alpha a) beta per source, and adds the sum to the destination (zero
// beforehand): bitwise the host build's accumulation.
__global__ __launch_bounds__(256) void k_build(const bsg_build_group_t* __restrict__ groups,
                                               const int64_t* __restrict__ estart, int ngroups,
                                               const bsg_build_src_t* __restrict__ srcs, int64_t nelem) {
  // no fused multiply-adds: the host build rounds every product and sum
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= nelem) return;
  // last group whose start is <= e
  int lo = 0, hi = ngroups - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (estart[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  const bsg_build_group_t g = groups[lo];
  const int64_t x = e - estart[lo];
  const int r = (int)(x / g.cols), c = (int)(x % g.cols);
  double s = 0.0;
  for (int j = 0; j < g.count; ++j) {
    const bsg_build_src_t sj = srcs[g.first + j];
    const double a = sj.trans ? sj.src[(int64_t)c * sj.ld + r] : sj.src[(int64_t)r * sj.ld + c];
    s += (sj.alpha * a) * sj.beta;
  }
  g.dst[(int64_t)r * g.ld + c] += s;
}

// Gram matrix tile of two vector sets (mx, my <= GMAX vectors of length n,
// stride ld): per-block partial sums G[i my + j] = sum_x X_i[x] Y_j[x] over
// the block's chunks, then an ordered sum over blocks (deterministic).
// Larger sets (up to GRAM_MAX vectors) run as GMAX x GMAX tiles.
constexpr int GMAX = 32, GCH = 64, GBLK = 256, GRAM_MAX = 128;
__global__ __launch_bounds__(256) void k_gram_partial(const double* __restrict__ X, const double* __restrict__ Y, int mx,
                                                      int my, int64_t n, int64_t ld, double* __restrict__ part) {
  __shared__ double xs[GMAX][GCH + 1], ys[GMAX][GCH + 1];
  double acc[GMAX * GMAX / 256];
  const int npair = mx * my;
#pragma unroll
  for (int u = 0; u < GMAX * GMAX / 256; ++u) acc[u] = 0.0;
  for (int64_t x0 = (int64_t)blockIdx.x * GCH; x0 < n; x0 += (int64_t)gridDim.x * GCH) {
    for (int e = threadIdx.x; e < GMAX * GCH; e += 256) {
      const int i = e / GCH, x = e % GCH;
      const bool ok = x0 + x < n;
      if (i < mx) xs[i][x] = ok ? X[(int64_t)i * ld + x0 + x] : 0.0;
      if (i < my) ys[i][x] = ok ? Y[(int64_t)i * ld + x0 + x] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < GMAX * GMAX / 256; ++u) {
      const int pidx = threadIdx.x + 256 * u;
      if (pidx < npair) {
        const int i = pidx / my, j = pidx % my;
        double s = 0.0;
        for (int x = 0; x < GCH; ++x) s += xs[i][x] * ys[j][x];
        acc[u] += s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < GMAX * GMAX / 256; ++u) {
    const int pidx = threadIdx.x + 256 * u;
    if (pidx < npair) part[(int64_t)blockIdx.x * npair + pidx] = acc[u];
  }
}
// out[(i) ldo + j] = sum over blocks of the tile's partials (i < mx, j < my)
__global__ __launch_bounds__(256) void k_gram_sum(const double* __restrict__ part, int nblk, int mx, int my,
                                                  double* __restrict__ out, int ldo) {
  const int pidx = blockIdx.x * 256 + threadIdx.x;
  const int npair = mx * my;
  if (pidx >= npair) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * npair + pidx];
  out[(int64_t)(pidx / my) * ldo + pidx % my] = s;
}
// Y_j[x] (+)= sum_{i < m} coef[i ldc + j] X_i[x] for j < mo (m, mo <= GMAX);
// accumulate = 0 overwrites Y
__global__ __launch_bounds__(256) void k_lincomb(const double* __restrict__ X, int m, int64_t n, int64_t ld,
                                                 const double* __restrict__ coef, int ldc, int mo,
                                                 double* __restrict__ Y, int64_t ldy, int accumulate) {
  __shared__ double cs[GMAX * GMAX];
  for (int e = threadIdx.x; e < m * mo; e += 256) cs[e] = coef[(int64_t)(e / mo) * ldc + e % mo];
  __syncthreads();
  for (int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (int64_t)gridDim.x * 256) {
    double xv[GMAX];
#pragma unroll
    for (int i = 0; i < GMAX; ++i) xv[i] = i < m ? X[(int64_t)i * ld + x] : 0.0;
    for (int j = 0; j < mo; ++j) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < GMAX; ++i)
        if (i < m) s += xv[i] * cs[i * mo + j];
      if (accumulate)
        Y[(int64_t)j * ldy + x] += s;
      else
        Y[(int64_t)j * ldy + x] = s;
    }
  }
}

// diag(H) by V block: item i covers elements [e0, e0 + 256) of the block of
// its term group; every thread sums the group's terms in list order for its
// element (alpha_t A_ii B_jj, identity partners read as 1).
struct DiagGroup {
  int32_t first, count;
};
struct DiagItem {
  int32_t group, e0;
};
__global__ __launch_bounds__(256) void k_diag(const bsg_diag_t* __restrict__ terms, const DiagGroup* __restrict__ groups,
                                              const DiagItem* __restrict__ items, const double* __restrict__ arena,
                                              double* __restrict__ diag) {
  const DiagItem it = items[blockIdx.x];
  const DiagGroup g = groups[it.group];
  const bsg_diag_t t0 = terms[g.first];
  const int64_t e = (int64_t)it.e0 + threadIdx.x;
  if (e >= (int64_t)t0.m * t0.n) return;
  const int i = (int)(e / t0.n), j = (int)(e % t0.n);
  double s = 0.0;
  for (int q = 0; q < g.count; ++q) {
    const bsg_diag_t t = terms[g.first + q];
    const double a = t.a_off < 0 ? 1.0 : arena[t.a_off + (int64_t)i * (t.m + 1)];
    const double b = t.b_off < 0 ? 1.0 : arena[t.b_off + (int64_t)j * (t.n + 1)];
    s += t.alpha * a * b;
  }
  diag[t0.v_off + e] += s;
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(2048, (n + 255) / 256)); }

}  // namespace

// ---------------------------------------------------------------- static scratch
// Partial tiles of bsg_sigma_apply_noalloc: a file-scope device array, one
// copy per GPU the library runs on, held until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): the address comes from hipGetSymbolAddress, once
// per device with that device current, and is kept per device.
__device__ __align__(256) double scratch[BSG_NOALLOC_SCRATCH_DOUBLES];
namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};
// this device's scratch address (device current); nullptr with g_err set on failure
double* scratch_address(int device) {
  if (device < 0 || device >= MAX_DEVICES) {
    g_err = "static scratch: device index out of range";
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[device];
}
}  // namespace

// ---------------------------------------------------------------- context and plan
struct bsg_context {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream (default priority) on which the small-block kernel runs
  // beside the grouped kernel; fork/join events tie it to `stream`
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  double* red = nullptr;  // RMAXB partials + 4 result slots
  double* gpart = nullptr;  // Gram partials (GBLK GMAX^2) + result (GRAM_MAX^2) + coefficients (GRAM_MAX^2)
  bsg_copy_seg_t* copy_list = nullptr;  // chunk list of bsg_copy_segments (grown on demand)
  size_t copy_cap = 0;
  double* coef_pin = nullptr;    // page-locked staging of bsg_lincomb's coefficients (GRAM_MAX^2)
  hipEvent_t coef_done = nullptr;  // its last copy to the device
  double* scratch = nullptr;  // this device's copy of the static scratch (bsg_sigma_apply_noalloc)
  // operator-slice buffer lent to one split-precision plan at a time (grown
  // on demand), so successive superblocks do not allocate and free it
  __bf16* abs_buf = nullptr;
  size_t abs_cap = 0;  // elements
  bool abs_leased = false;
};

struct PlanChunk {
  bsg_vblock_t* d_vb = nullptr;
  bsg_pair_t* d_pr = nullptr;
  void* d_blob = nullptr;  // one device allocation holding the arrays below (not d_ps)
  PairS* d_ps = nullptr;  // slice offsets of the grouped kernel's pairs (split precision, built lazily)
  // host copies for the slice recipes: pairs in device order, the V block of
  // each grouped-kernel pair (-1: small-block kernel), the V blocks
  std::vector<bsg_pair_t> h_perm;
  std::vector<int> h_vbi;
  std::vector<bsg_vblock_t> h_vb;
  Seg* d_seg = nullptr;
  TileRef* d_g[NCLS] = {};  // FP64 grouped-kernel tiles per kernel class
  int n_g[NCLS] = {};
  TileRef* d_gbf[3] = {};  // split-precision tiles (basic, 256 wide, 512 wide)
  int n_gbf[3] = {};
  int4* d_s[2] = {nullptr, nullptr};  // FP64 small-block units, 4 per workgroup (SB = 2, SB = 1)
  int n_s[2] = {0, 0};
  TileRef* d_w[2] = {};  // FP64 per-wave tiles (16 x 32, 16 x 64), 4 per workgroup (item -1: none), a multiple of 32
  int n_w[2] = {};
  RedItem* d_red[2] = {nullptr, nullptr};  // after the FP64 / the split-precision kernels
  int n_red[2] = {0, 0};
};

struct bsg_plan {
  bsg_context* ctx = nullptr;
  std::vector<PlanChunk> chunks;
  double* part = nullptr;    // partial tiles of `batch` roots, part_doubles apart
  int64_t part_doubles = 0;  // partial tiles of split V blocks per root (largest chunk)
  int32_t batch = 1;         // roots per launch (bsg_plan_set_batch)
  bool cs_ready = false;     // d_cs holds C slices for `batch` roots
  bool slices_built = false; // slice recipes and pair slice offsets (first bsg_plan_split_arena)
  bool abs_from_ctx = false; // d_abs is the context's lent buffer
  int64_t wave_size = 0;
  int64_t arena_extent = 0;  // doubles of operator arena the pairs reach
  double flops = 0.0, flops1 = 0.0, flops2 = 0.0;
  double flops_split = 0.0;  // of flops: the grouped kernels' pairs (BF16 slices in split mode)
  // split precision: the arena bsg_plan_split_arena sliced, the slice
  // recipes of the operator blocks (once per arena) and of the C blocks
  // (every application) and their BF16 buffers
  const double* split_src = nullptr;
  std::vector<SliceItem> ab_items, c_items;
  int64_t ab_elems = 0, c_elems = 0;
  SliceItem *d_ab_items = nullptr, *d_c_items = nullptr;
  __bf16 *d_abs = nullptr, *d_cs = nullptr;
  // one captured application (bsg_plan_capture): its executable graph and
  // what it was captured for
  hipGraphExec_t graph = nullptr;
  const double *g_arena = nullptr, *g_c = nullptr;
  double* g_v = nullptr;
  int32_t g_nvec = 0, g_precision = -1;
};

extern "C" {

const char* bsg_last_error(void) { return g_err.c_str(); }

int bsg_init(bsg_context** ctx, int device) {
  if (!ctx) return fail(BSG_ERR_ARG, "bsg_init: null ctx");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(BSG_ERR_ARG, "bsg_init: no such device");
  HIPCHK(hipSetDevice(device));
  auto* c = new bsg_context();
  c->device = device;
  hipError_t e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) {
    delete c;
    return fail(BSG_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  e = hipMalloc(&c->red, sizeof(double) * (RMAXB + 4));
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(BSG_ERR_NOMEM, "bsg_init: reduction buffer");
  }
  if (hipMalloc(&c->gpart, sizeof(double) * ((size_t)GBLK * GMAX * GMAX + 2 * GRAM_MAX * GRAM_MAX)) != hipSuccess) {
    (void)hipFree(c->red);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(BSG_ERR_NOMEM, "bsg_init: Gram scratch");
  }
  c->scratch = scratch_address(device);
  if (!c->scratch) {
    const std::string m = g_err;
    (void)hipFree(c->gpart);
    (void)hipFree(c->red);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(BSG_ERR_HIP, "bsg_init: " + m);
  }
  if (hipStreamCreate(&c->side) != hipSuccess || hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->coef_done, hipEventDisableTiming) != hipSuccess ||
      hipHostMalloc((void**)&c->coef_pin, sizeof(double) * GRAM_MAX * GRAM_MAX, hipHostMallocDefault) != hipSuccess) {
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->coef_done) (void)hipEventDestroy(c->coef_done);
    if (c->coef_pin) (void)hipHostFree(c->coef_pin);
    (void)hipFree(c->gpart);
    (void)hipFree(c->red);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(BSG_ERR_HIP, "bsg_init: side stream");
  }
  *ctx = c;
  return BSG_OK;
}

int bsg_free(bsg_context* ctx) {
  if (!ctx) return BSG_OK;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(ctx->red));
  if (ctx->copy_list) HIPCHK(hipFree(ctx->copy_list));
  HIPCHK(hipStreamSynchronize(ctx->side));
  HIPCHK(hipFree(ctx->gpart));
  if (ctx->abs_buf) HIPCHK(hipFree(ctx->abs_buf));
  HIPCHK(hipHostFree(ctx->coef_pin));
  HIPCHK(hipEventDestroy(ctx->coef_done));
  HIPCHK(hipEventDestroy(ctx->fork));
  HIPCHK(hipEventDestroy(ctx->join));
  HIPCHK(hipStreamDestroy(ctx->side));
  HIPCHK(hipStreamDestroy(ctx->stream));
  delete ctx;
  return BSG_OK;
}

int bsg_synchronize(bsg_context* ctx) {
  if (!ctx) return fail(BSG_ERR_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BSG_OK;
}

int bsg_malloc(bsg_context* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return fail(BSG_ERR_ARG, "bsg_malloc: null argument");
  HIPCHK(hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(dptr, std::max<size_t>(bytes, 8));
  if (e != hipSuccess) return fail(BSG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return BSG_OK;
}
int bsg_mfree(bsg_context* ctx, void* dptr) {
  if (!ctx) return fail(BSG_ERR_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(dptr));
  return BSG_OK;
}
int bsg_memcpy_h2d(bsg_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return fail(BSG_ERR_ARG, "null ctx");
  if (bytes == 0) return BSG_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BSG_OK;
}
int bsg_host_alloc(bsg_context* ctx, size_t bytes, void** hptr) {
  if (!ctx || !hptr) return fail(BSG_ERR_ARG, "bsg_host_alloc: null argument");
  hipError_t e = hipHostMalloc(hptr, std::max<size_t>(bytes, 8), hipHostMallocDefault);
  if (e != hipSuccess) return fail(BSG_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  return BSG_OK;
}
int bsg_host_free(bsg_context* ctx, void* hptr) {
  if (!ctx) return fail(BSG_ERR_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipHostFree(hptr));
  return BSG_OK;
}
int bsg_memcpy_h2d_async(bsg_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return fail(BSG_ERR_ARG, "null ctx");
  if (bytes == 0) return BSG_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return BSG_OK;
}
int bsg_memcpy_d2h(bsg_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return fail(BSG_ERR_ARG, "null ctx");
  if (bytes == 0) return BSG_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BSG_OK;
}
int bsg_memset_zero(bsg_context* ctx, void* dst, size_t bytes) {
  if (!ctx) return fail(BSG_ERR_ARG, "null ctx");
  if (bytes == 0) return BSG_OK;
  HIPCHK(hipMemsetAsync(dst, 0, bytes, ctx->stream));
  return BSG_OK;
}

int bsg_plan_create(bsg_context* ctx, const bsg_chunk_t* chunks, int32_t n_chunks, int64_t wave_size,
                    bsg_plan** plan) {
  if (!ctx || !plan || (n_chunks > 0 && !chunks) || wave_size < 0) return fail(BSG_ERR_ARG, "bsg_plan_create: bad args");
  HIPCHK(hipSetDevice(ctx->device));
  auto* p = new bsg_plan();
  p->ctx = ctx;
  p->wave_size = wave_size;
  int64_t pmax = 0;
  const long SEG_COST = 2048;        // MFMAs per wavefront of a work unit before its pair list is split
  const long SEG_COST_SMALL = 1024;  // the same for the per-wave small-block units
  const long SEG_COST_WAVE = 2048;   // and for the per-wave tiles
  auto cleanup = [&](int code, const std::string& m) {
    bsg_plan_destroy(p);
    return fail(code, m);
  };
  for (int ci = 0; ci < n_chunks; ++ci) {
    const bsg_chunk_t& ch = chunks[ci];
    PlanChunk pc;
    // The V blocks are independent: contiguous ranges of them (about equal
    // in pairs) are planned on host threads into local lists with local
    // indices, which are then concatenated in V block order (the result does
    // not depend on the number of threads).
    struct Local {
      std::vector<GUnit> gunits[NCLS], gunits_bf[3];  // the tiles of each work unit (FP64, split)
      std::vector<bsg_pair_t> perm;                   // pairs in device order (grouped per V block as given)
      std::vector<int> h_vbi;
      std::vector<std::pair<long, int>> sunits[2];  // (cost, segment) of small-block units
      std::vector<GUnit> wunits[2];                 // per-wave tile units (cost per tile), 16 x 32 and 16 x 64
      std::vector<Seg> segs;
      std::vector<RedItem> reds[2];  // ordered reductions after the FP64 / the split-precision kernels
      int64_t poff = 0, arena_extent = 0;
      double flops = 0.0, flops1 = 0.0, flops2 = 0.0, flops_split = 0.0;
      std::string err;
      std::vector<int> fast, slow;  // per V block, reused
      std::vector<std::pair<int, int>> units;
      std::vector<char> ukind;  // per unit: 0 grouped kernel, 1 small-block kernel, 2 per-wave tiles
    };
    auto plan_vblock = [&](int i, Local& L) -> bool {
      std::vector<bsg_pair_t>& perm = L.perm;
      std::vector<Seg>& segs = L.segs;
      int64_t& poff = L.poff;
      auto& fast = L.fast;
      auto& slow = L.slow;
      auto& units = L.units;
      auto& ukind = L.ukind;
      const auto& vb = ch.vblocks[i];
      if (vb.m <= 0 || vb.n <= 0 || vb.v_off < 0 || vb.v_off + (int64_t)vb.m * vb.n > wave_size || vb.first < 0 ||
          vb.count < 0 || vb.first + vb.count > ch.n_pairs) {
        L.err = "bsg_plan_create: vblock out of range";
        return false;
      }
      const Geom gm = pick_geom(vb.m, vb.n, 6);     // FP64 tiles
      const Geom gs = pick_geom(vb.m, vb.n, 4);     // split-precision tiles
      // small blocks (up to 32 x 32) go to the per-wave kernel, except pairs
      // whose C block is larger than that (or whose A is the identity and B
      // is not): those run in the grouped kernel with its K pipeline
      const bool small_vb = std::max(vb.m, vb.n) <= 32;
      // medium blocks the tile table puts on per-wave tiles (FP64 only): the
      // pairs within the per-wave kernel's contraction depth go there
      const bool wave_vb = !small_vb && wave_block(vb.m, vb.n);
      // tiny: the block and every small-kernel pair's C block within 16 x 16
      bool tiny = std::max(vb.m, vb.n) <= 16;
      fast.clear();  // pair indices (absolute) in term order
      slow.clear();
      for (int q = 0; q < vb.count; ++q) {
        const auto& pr = ch.pairs[vb.first + q];
        if (pr.k1 <= 0 || pr.k2 <= 0 || pr.c_off < 0 || pr.c_off + (int64_t)pr.k1 * pr.k2 > wave_size ||
            (pr.a_off < 0 && pr.k1 != vb.m) || (pr.b_off < 0 && pr.k2 != vb.n)) {
          L.err = "bsg_plan_create: pair out of range";
          return false;
        }
        if (pr.a_off >= 0) L.arena_extent = std::max(L.arena_extent, pr.a_off + (int64_t)vb.m * pr.k1);
        if (pr.b_off >= 0) L.arena_extent = std::max(L.arena_extent, pr.b_off + (int64_t)vb.n * pr.k2);
        if (pr.a_off >= 0) {
          L.flops += 2.0 * vb.m * pr.k1 * pr.k2;
          L.flops1 += 2.0 * vb.m * pr.k1 * pr.k2;
        }
        if (pr.b_off >= 0) {
          L.flops += 2.0 * vb.m * pr.k2 * vb.n;
          L.flops2 += 2.0 * vb.m * pr.k2 * vb.n;
        }
        const bool f = small_vb ? (pr.a_off >= 0 && pr.k1 <= 32 && pr.k2 <= 32) || (pr.a_off < 0 && pr.b_off < 0)
                                : wave_vb && (pr.a_off < 0 || pr.k1 <= 4 * wave_ka(wave_nb(vb.n)));
        (f ? fast : slow).push_back(vb.first + q);
        if (!f || wave_vb)  // split mode: every pair of a per-wave block runs on BF16 slices
          L.flops_split += (pr.a_off >= 0 ? 2.0 * vb.m * pr.k1 * pr.k2 : 0.0) +
                           (pr.b_off >= 0 ? 2.0 * vb.m * pr.k2 * vb.n : 0.0);
        if (f && small_vb && (pr.k1 > 16 || pr.k2 > 16)) tiny = false;
      }
      // the block's pairs in device order: the small-kernel (per-wave
      // kernel) ones, then the rest
      const int base = (int)perm.size();
      for (int x : fast) {
        perm.push_back(ch.pairs[x]);
        L.h_vbi.push_back(wave_vb ? i : -1);  // the small-block kernel reads no slices
      }
      for (int x : slow) {
        perm.push_back(ch.pairs[x]);
        L.h_vbi.push_back(i);
      }
      // cut each list into work units of about SEG_COST(_SMALL) MFMAs per
      // wave.  A block on wide tiles is cut twice: for the FP64 wide tiles
      // and for the split-precision kernel's basic tiles (mode 1 and 2, each
      // with its own partial tiles and reduction); other blocks share one
      // cut (mode 0).  Wide blocks have no small-kernel pairs.
      const int64_t mn = (int64_t)vb.m * vb.n;
      // nf: the first nf pairs go to the kernel of kind fk (1 small-block, 2
      // per-wave tiles), the rest to the grouped kernel on geometry g
      auto emit = [&](const Geom& g, int mode, int nf, int fk) {
        units.clear();  // (first, count) in device order
        ukind.clear();
        auto cut = [&](int first, int n, int kind) {
          const long lim = kind == 1 ? SEG_COST_SMALL : (kind == 2 ? SEG_COST_WAVE : SEG_COST);
          int start = 0;
          long cost = 0;
          for (int q = 0; q < n; ++q) {
            const auto& pr = perm[first + q];
            cost += kind == 1 ? small_cost(vb.m, vb.n, pr) : (kind == 2 ? wave_cost(vb.n, pr) : pair_cost(g, vb.n, pr));
            if (cost >= lim && q + 1 < n) {
              units.push_back({first + start, q + 1 - start});
              ukind.push_back((char)kind);
              start = q + 1;
              cost = 0;
            }
          }
          if (n > start) {
            units.push_back({first + start, n - start});
            ukind.push_back((char)kind);
          }
        };
        const int ntot = (int)(fast.size() + slow.size());
        cut(base, nf, fk);
        cut(base + nf, ntot - nf, 0);
        const int nseg = (int)units.size();
        for (int u = 0; u < nseg; ++u) {
          Seg sg;
          sg.vb = i;
          sg.first = units[u].first;
          sg.count = units[u].second;
          sg.pad = 0;
          sg.p_off = nseg == 1 ? -1 : poff + (int64_t)u * mn;
          int si = (int)segs.size();
          segs.push_back(sg);
          if (ukind[u] == 1) {
            long sc = 0;
            for (int q = 0; q < sg.count; ++q) sc += small_cost(vb.m, vb.n, perm[sg.first + q]);
            L.sunits[tiny ? 1 : 0].push_back({sc, si});
            continue;
          }
          if (ukind[u] == 2) {
            GUnit gu;
            gu.cost = 0;
            for (int q = 0; q < sg.count; ++q) gu.cost += wave_cost(vb.n, perm[sg.first + q]);
            gu.seg = si;
            gu.ntm = (vb.m + 15) / 16;
            const int nb = wave_nb(vb.n);
            gu.ntn = (vb.n + 16 * nb - 1) / (16 * nb);
            gu.code = 0;
            L.wunits[nb == 4].push_back(gu);
            continue;
          }
          long ucost = 0;
          for (int q = 0; q < sg.count; ++q) ucost += pair_cost(g, vb.n, perm[sg.first + q]);
          GUnit gu;
          gu.cost = ucost;
          gu.seg = si;
          gu.ntm = (vb.m + g.tm - 1) / g.tm;
          gu.ntn = (vb.n + g.jv - 1) / g.jv;
          gu.code = g.code;
          if (mode != 2) L.gunits[g.cls].push_back(gu);
          if (mode != 1 && g.cls < 3) L.gunits_bf[g.cls].push_back(gu);
        }
        if (nseg > 1) {
          for (int64_t e0 = 0; e0 < mn; e0 += 256) {
            RedItem r;
            r.v_off = vb.v_off;
            r.p_off = poff;
            r.nseg = nseg;
            r.mn = (int32_t)mn;
            r.e0 = (int32_t)e0;
            r.len = (int32_t)std::min<int64_t>(256, mn - e0);
            if (mode != 2) L.reds[0].push_back(r);
            if (mode != 1) L.reds[1].push_back(r);
          }
          poff += nseg * mn;
        }
      };
      const int nfast = (int)fast.size();
      if (wave_vb) {
        emit(gm, 1, nfast, 2);
        emit(gs, 2, 0, 0);
      } else if (gm.code == gs.code && gm.nv == gs.nv) {
        emit(gm, 0, nfast, 1);
      } else {
        emit(gm, 1, nfast, 1);
        emit(gs, 2, nfast, 1);
      }
      return true;
    };
    // ranges of V blocks with about equal pair counts, one per thread
    int nth = (int)std::min<long>(std::max(1u, std::thread::hardware_concurrency()), 16);
    if (ch.n_pairs < 20000) nth = 1;
    nth = std::max(1, std::min(nth, ch.n_vblocks));
    std::vector<int> bounds(nth + 1, ch.n_vblocks);
    bounds[0] = 0;
    {
      int64_t acc = 0, tot = 0;
      for (int i = 0; i < ch.n_vblocks; ++i) tot += std::max(0, ch.vblocks[i].count) + 1;
      int t = 1;
      for (int i = 0; i < ch.n_vblocks && t < nth; ++i) {
        acc += std::max(0, ch.vblocks[i].count) + 1;
        if (acc * nth >= tot * t) bounds[t++] = i + 1;
      }
    }
    std::vector<Local> loc(nth);
    auto work = [&](int t) {
      for (int i = bounds[t]; i < bounds[t + 1]; ++i)
        if (!plan_vblock(i, loc[t])) return;
    };
    if (nth == 1) {
      work(0);
    } else {
      std::vector<std::thread> th;
      for (int t = 1; t < nth; ++t) th.emplace_back(work, t);
      work(0);
      for (auto& x : th) x.join();
    }
    for (auto& L : loc)
      if (!L.err.empty()) return cleanup(BSG_ERR_ARG, L.err);
    // concatenate with the local indices shifted
    std::vector<GUnit> gunits[NCLS], gunits_bf[3];
    std::vector<bsg_pair_t> perm;
    std::vector<std::pair<long, int>> sunits[2];
    std::vector<GUnit> wunits[2];
    std::vector<Seg> segs;
    std::vector<RedItem> reds[2];
    int64_t poff = 0;
    for (auto& L : loc) {
      const int pb = (int)perm.size(), sb = (int)segs.size();
      const int64_t ob = poff;
      p->flops += L.flops;
      p->flops1 += L.flops1;
      p->flops2 += L.flops2;
      p->flops_split += L.flops_split;
      p->arena_extent = std::max(p->arena_extent, L.arena_extent);
      perm.insert(perm.end(), L.perm.begin(), L.perm.end());
      pc.h_vbi.insert(pc.h_vbi.end(), L.h_vbi.begin(), L.h_vbi.end());
      for (Seg sg : L.segs) {
        sg.first += pb;
        if (sg.p_off >= 0) sg.p_off += ob;
        segs.push_back(sg);
      }
      for (int k = 0; k < 2; ++k)
        for (auto x : L.sunits[k]) sunits[k].push_back({x.first, x.second + sb});
      for (int k = 0; k < 2; ++k)
        for (GUnit g : L.wunits[k]) {
          g.seg += sb;
          wunits[k].push_back(g);
        }
      for (int k = 0; k < NCLS; ++k)
        for (GUnit g : L.gunits[k]) {
          g.seg += sb;
          gunits[k].push_back(g);
        }
      for (int k = 0; k < 3; ++k)
        for (GUnit g : L.gunits_bf[k]) {
          g.seg += sb;
          gunits_bf[k].push_back(g);
        }
      for (int k = 0; k < 2; ++k)
        for (RedItem r : L.reds[k]) {
          r.p_off += ob;
          reds[k].push_back(r);
        }
      poff += L.poff;
    }
    pmax = std::max(pmax, poff);
    // every array of the chunk goes into one device allocation with one copy
    // (256-byte aligned parts): one hipMalloc / hipFree per chunk
    std::vector<std::pair<void**, size_t>> parts;  // (device pointer to set, offset in the blob)
    std::vector<char> blob;
    blob.reserve(sizeof(bsg_vblock_t) * ch.n_vblocks + sizeof(bsg_pair_t) * perm.size() + sizeof(Seg) * segs.size() +
                 sizeof(RedItem) * (reds[0].size() + reds[1].size()) +
                 sizeof(int4) * (sunits[0].size() + sunits[1].size() + 8) / 4 + (1 << 16));
    auto up = [&](void** d, const void* h, size_t bytes) {
      if (bytes == 0) return;
      const size_t off = (blob.size() + 255) & ~(size_t)255;
      blob.resize(off + bytes);
      std::memcpy(blob.data() + off, h, bytes);
      parts.push_back({d, off});
    };
    up((void**)&pc.d_vb, ch.vblocks, sizeof(bsg_vblock_t) * ch.n_vblocks);
    up((void**)&pc.d_pr, perm.data(), sizeof(bsg_pair_t) * perm.size());
    up((void**)&pc.d_seg, segs.data(), sizeof(Seg) * segs.size());
    up((void**)&pc.d_red[0], reds[0].data(), sizeof(RedItem) * reds[0].size());
    up((void**)&pc.d_red[1], reds[1].data(), sizeof(RedItem) * reds[1].size());
    std::vector<TileRef> gtiles[NCLS];
    for (int k = 0; k < NCLS; ++k) gtiles[k] = xcd_order(gunits[k]);
    std::vector<TileRef> gtiles_bf[3] = {xcd_order(gunits_bf[0]), xcd_order(gunits_bf[1]), xcd_order(gunits_bf[2])};
    // small-block units, costliest first, four per workgroup
    std::vector<int4> sq[2];
    for (int k = 0; k < 2; ++k) {
      std::sort(sunits[k].begin(), sunits[k].end(), [](const std::pair<long, int>& x, const std::pair<long, int>& y) {
        return x.first != y.first ? x.first > y.first : x.second < y.second;
      });
      for (size_t i = 0; i < sunits[k].size(); i += 4) {
        int e[4];
        for (int j = 0; j < 4; ++j) e[j] = i + j < sunits[k].size() ? sunits[k][i + j].second : -1;
        sq[k].push_back(make_int4(e[0], e[1], e[2], e[3]));
      }
      up((void**)&pc.d_s[k], sq[k].data(), sizeof(int4) * sq[k].size());
      pc.n_s[k] = (int)sq[k].size();
    }
    for (int k = 0; k < NCLS; ++k) {
      up((void**)&pc.d_g[k], gtiles[k].data(), sizeof(TileRef) * gtiles[k].size());
      pc.n_g[k] = (int)gtiles[k].size();
    }
    // per-wave tiles of each width: units (costliest first) go to the least
    // loaded of the 8 XCDs, a unit's tiles in consecutive wave slots of that
    // XCD's workgroups (they share the unit's C and B' blocks in its L2);
    // workgroup 8 j + x is XCD x's j-th
    for (int k = 0; k < 2; ++k) {
      std::vector<std::pair<long, int>> order;
      for (int u = 0; u < (int)wunits[k].size(); ++u)
        order.push_back({wunits[k][u].cost * wunits[k][u].ntm * wunits[k][u].ntn, u});
      std::sort(order.begin(), order.end(), [](const std::pair<long, int>& x, const std::pair<long, int>& y) {
        return x.first != y.first ? x.first > y.first : x.second < y.second;
      });
      std::vector<TileRef> q[8];
      long load[8] = {};
      for (const auto& o : order) {
        const int x = (int)(std::min_element(load, load + 8) - load);
        const GUnit& gu = wunits[k][o.second];
        for (int tm = 0; tm < gu.ntm; ++tm)
          for (int tn = 0; tn < gu.ntn; ++tn) q[x].push_back(TileRef{gu.seg, tm, tn, 0});
        load[x] += o.first;
      }
      size_t nwg = 0;
      for (auto& e : q) nwg = std::max(nwg, (e.size() + 3) / 4);
      std::vector<TileRef> wt(nwg * 32, TileRef{-1, 0, 0, 0});
      for (int x = 0; x < 8; ++x)
        for (size_t e = 0; e < q[x].size(); ++e) wt[((e / 4) * 8 + x) * 4 + e % 4] = q[x][e];
      up((void**)&pc.d_w[k], wt.data(), sizeof(TileRef) * wt.size());
      pc.n_w[k] = (int)wt.size();
    }
    for (int k = 0; k < 3; ++k) {
      up((void**)&pc.d_gbf[k], gtiles_bf[k].data(), sizeof(TileRef) * gtiles_bf[k].size());
      pc.n_gbf[k] = (int)gtiles_bf[k].size();
    }
    if (!blob.empty()) {
      char* d = nullptr;
      if (hipMalloc((void**)&d, blob.size()) != hipSuccess ||
          hipMemcpy(d, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        p->chunks.push_back(std::move(pc));
        return cleanup(BSG_ERR_NOMEM, "bsg_plan_create: plan arrays");
      }
      pc.d_blob = d;
      for (auto& x : parts) *x.first = d + x.second;
    }
    pc.n_red[0] = (int)reds[0].size();
    pc.n_red[1] = (int)reds[1].size();
    pc.h_perm = std::move(perm);
    pc.h_vb.assign(ch.vblocks, ch.vblocks + ch.n_vblocks);
    p->chunks.push_back(std::move(pc));
  }
  p->part_doubles = pmax;
  if (pmax > 0) {
    hipError_t e = hipMalloc(&p->part, sizeof(double) * pmax * p->batch);
    if (e != hipSuccess) return cleanup(BSG_ERR_NOMEM, "bsg_plan_create: partial-tile scratch");
  }
  *plan = p;
  return BSG_OK;
}

int bsg_plan_destroy(bsg_plan* p) {
  if (!p) return BSG_OK;
  if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
  for (auto& c : p->chunks) {
    (void)hipFree(c.d_blob);  // every array but the lazily built slice offsets
    (void)hipFree(c.d_ps);
  }
  (void)hipFree(p->part);
  (void)hipFree(p->d_ab_items);
  (void)hipFree(p->d_c_items);
  if (p->abs_from_ctx)
    p->ctx->abs_leased = false;
  else
    (void)hipFree(p->d_abs);
  (void)hipFree(p->d_cs);
  if (p->graph) (void)hipGraphExecDestroy(p->graph);
  delete p;
  return BSG_OK;
}

int64_t bsg_plan_arena_extent(const bsg_plan* p) { return p ? p->arena_extent : 0; }
int64_t bsg_plan_partial_doubles(const bsg_plan* p) { return p ? p->part_doubles : 0; }

}  // extern "C"

namespace {
// open-addressing tables (no allocation per entry), grown by doubling at
// half load so that their size follows the distinct blocks, not the pairs:
// keys of the operator blocks (offset, orientation, rows, cols) and of the
// C blocks (offset)
struct SliceTable {
  struct E {
    int64_t off, key, val;  // off < 0: empty
  };
  std::vector<E> e;
  size_t n = 0;
  explicit SliceTable(size_t hint) {  // hint: entries expected at most
    size_t cap = 4096;
    while (cap < hint) cap <<= 1;
    e.assign(cap, E{-1, 0, 0});
  }
  static size_t hash(int64_t off, int64_t key) {
    uint64_t h = (uint64_t)off * 0x9E3779B97F4A7C15ull ^ (uint64_t)key * 0xC2B2AE3D27D4EB4Full;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    return (size_t)(h ^ (h >> 29));
  }
  E& slot(int64_t off, int64_t key) {
    const size_t mask = e.size() - 1;
    for (size_t h = hash(off, key) & mask;; h = (h + 1) & mask)
      if (e[h].off < 0 || (e[h].off == off && e[h].key == key)) return e[h];
  }
  // the value of (off, key), made by make() on first sight
  template <class F>
  int64_t get(int64_t off, int64_t key, F&& make) {
    if (2 * (n + 1) > e.size()) {
      std::vector<E> old(2 * e.size(), E{-1, 0, 0});
      old.swap(e);
      for (const E& x : old)
        if (x.off >= 0) slot(x.off, x.key) = x;
    }
    E& x = slot(off, key);
    if (x.off < 0) {
      x = E{off, key, make()};
      ++n;
    }
    return x.val;
  }
};
// BF16 slice recipes of a plan, one per distinct (operator block,
// orientation) and per C block its grouped-kernel pairs read, and the pairs'
// slice offsets; made on the first split-precision set-up, so FP64-only
// plans (the rotations, FP64 solves) never pay for them.
int build_slices(bsg_plan* p) {
  // the operator table sized by its references (most operator blocks of a
  // superblock are referenced once) and its item list reserved likewise;
  // the C blocks are few (one per wavefunction block)
  size_t n_ab = 0;
  for (auto& pc : p->chunks)
    for (size_t j = 0; j < pc.h_perm.size(); ++j)
      if (pc.h_vbi[j] >= 0) n_ab += (pc.h_perm[j].a_off >= 0) + (pc.h_perm[j].b_off >= 0);
  SliceTable abt(n_ab), ct(0);
  p->ab_items.reserve(n_ab);
  auto slice_of = [&](std::vector<SliceItem>& items, int64_t& elems, int64_t src, int rows, int cols, int trans) {
    SliceItem it;
    it.src = src;
    it.dst = elems;
    it.rows = rows;
    it.cols = cols;
    it.pitch = (cols + 7) & ~7;
    it.trans = trans;
    items.push_back(it);
    elems += 3 * (int64_t)rows * it.pitch;
    return it.dst;
  };
  auto ab_slice = [&](int64_t off, int trans, int rows, int cols) {
    const int64_t key = ((int64_t)rows << 33) | ((int64_t)cols << 1) | trans;
    return abt.get(off, key, [&] { return slice_of(p->ab_items, p->ab_elems, off, rows, cols, trans); });
  };
  auto c_slice = [&](int64_t off, int k1, int k2) {
    // C^T: (j, p) = C[p k2 + j]
    return ct.get(off, 0, [&] { return slice_of(p->c_items, p->c_elems, off, k2, k1, 1); });
  };
  for (auto& pc : p->chunks) {
    std::vector<PairS> ps(pc.h_perm.size(), PairS{-1, -1, -1});
    for (size_t j = 0; j < pc.h_perm.size(); ++j) {
      if (pc.h_vbi[j] < 0) continue;
      const bsg_vblock_t& vb = pc.h_vb[pc.h_vbi[j]];
      const bsg_pair_t& pr = pc.h_perm[j];
      if (pr.a_off >= 0) {
        ps[j].a = ab_slice(pr.a_off, pr.a_trans, vb.m, pr.k1);  // op(A)(i, p), K = p contiguous
        ps[j].c = c_slice(pr.c_off, pr.k1, pr.k2);
      }
      if (pr.b_off >= 0) ps[j].b = ab_slice(pr.b_off, pr.b_trans ? 0 : 1, vb.n, pr.k2);  // B'(k, col) as [col][k]
    }
    if (!ps.empty()) {
      HIPCHK(hipMalloc((void**)&pc.d_ps, sizeof(PairS) * ps.size()));
      HIPCHK(hipMemcpy(pc.d_ps, ps.data(), sizeof(PairS) * ps.size(), hipMemcpyHostToDevice));
    }
  }
  for (auto* v : {&p->ab_items, &p->c_items})
    if (!v->empty()) {
      SliceItem** d = v == &p->ab_items ? &p->d_ab_items : &p->d_c_items;
      if (hipMalloc((void**)d, sizeof(SliceItem) * v->size()) != hipSuccess ||
          hipMemcpy(*d, v->data(), sizeof(SliceItem) * v->size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(BSG_ERR_NOMEM, "bsg_plan_split_arena: slice recipes");
    }
  p->slices_built = true;
  return BSG_OK;
}
}  // namespace

extern "C" {

int bsg_plan_split_arena(bsg_context* ctx, bsg_plan* p, const double* arena, int64_t n) {
  if (!ctx || !p || !arena || n < 0) return fail(BSG_ERR_ARG, "bsg_plan_split_arena: bad args");
  if (p->ctx != ctx) return fail(BSG_ERR_ARG, "bsg_plan_split_arena: plan belongs to another context");
  if (n < p->arena_extent) return fail(BSG_ERR_ARG, "bsg_plan_split_arena: arena smaller than the plan's extent");
  HIPCHK(hipSetDevice(ctx->device));
  if (!p->slices_built) {
    const int rc = build_slices(p);
    if (rc != BSG_OK) return rc;
  }
  // operator slices for the grouped kernel (the small-block kernel splits
  // in registers), and room for the C slices of one launch's roots
  if (!p->d_abs && p->ab_elems > 0) {
    if (!ctx->abs_leased) {
      if (ctx->abs_cap < (size_t)p->ab_elems) {
        (void)hipFree(ctx->abs_buf);
        ctx->abs_buf = nullptr;
        ctx->abs_cap = 0;
        const size_t cap = (size_t)p->ab_elems + (size_t)p->ab_elems / 4;
        if (hipMalloc((void**)&ctx->abs_buf, 2 * cap) != hipSuccess)
          return fail(BSG_ERR_NOMEM, "bsg_plan_split_arena: operator slices");
        ctx->abs_cap = cap;
      }
      p->d_abs = ctx->abs_buf;
      ctx->abs_leased = true;
      p->abs_from_ctx = true;
    } else {
      hipError_t e = hipMalloc((void**)&p->d_abs, 2 * p->ab_elems);
      if (e != hipSuccess) return fail(BSG_ERR_NOMEM, "bsg_plan_split_arena: operator slices");
    }
  }
  if (!p->cs_ready && p->c_elems > 0) {
    hipError_t e = hipMalloc((void**)&p->d_cs, 2 * p->c_elems * p->batch);
    if (e != hipSuccess) return fail(BSG_ERR_NOMEM, "bsg_plan_split_arena: wavefunction slices");
  }
  p->cs_ready = true;
  if (!p->ab_items.empty())
    hipLaunchKernelGGL(k_slices, dim3(p->ab_items.size()), dim3(256), 0, ctx->stream, p->d_ab_items, arena, p->d_abs,
                       (int64_t)0, (int64_t)0);
  HIPCHK(hipGetLastError());
  p->split_src = arena;
  return BSG_OK;
}

int bsg_set_tile_table(const int32_t* rules, int32_t n) {
  if (n < 0 || (n > 0 && !rules)) return fail(BSG_ERR_ARG, "bsg_set_tile_table: bad args");
  std::vector<TileRule> t;
  const int nfp = (int)geom_candidates(6).size(), nsp = (int)geom_candidates(4).size();
  for (int32_t i = 0; i < n; ++i) {
    const int32_t* e = rules + 6 * i;
    TileRule r{e[0], e[1], e[2], e[3], e[4], e[5]};
    if (r.m_lo > r.m_hi || r.n_lo > r.n_hi || r.fp64 < -1 || r.fp64 > nfp || r.split < -1 || r.split >= nsp)
      return fail(BSG_ERR_ARG, "bsg_set_tile_table: bad rule");
    t.push_back(r);
  }
  g_tile_rules = t;
  ++g_tile_version;
  return BSG_OK;
}
int bsg_tile_candidates(int32_t precision) {
  // FP64: the geometries, then the per-wave kernel
  return precision == BSG_SPLIT_BF16 ? (int)geom_candidates(4).size() : (int)geom_candidates(6).size() + 1;
}

int bsg_plan_set_batch(bsg_context* ctx, bsg_plan* p, int32_t nroot) {
  if (!ctx || !p || nroot < 1) return fail(BSG_ERR_ARG, "bsg_plan_set_batch: bad args");
  if (p->ctx != ctx) return fail(BSG_ERR_ARG, "bsg_plan_set_batch: plan belongs to another context");
  if (nroot == p->batch) return BSG_OK;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));  // no queued work reads the old buffers
  if (p->graph) {  // it reads the buffers replaced below
    HIPCHK(hipGraphExecDestroy(p->graph));
    p->graph = nullptr;
  }
  if (p->part_doubles > 0) {
    double* np = nullptr;
    if (hipMalloc(&np, sizeof(double) * p->part_doubles * nroot) != hipSuccess)
      return fail(BSG_ERR_NOMEM, "bsg_plan_set_batch: partial-tile scratch");
    (void)hipFree(p->part);
    p->part = np;
  }
  if (p->cs_ready && p->c_elems > 0) {
    __bf16* nc = nullptr;
    if (hipMalloc((void**)&nc, 2 * p->c_elems * nroot) != hipSuccess)
      return fail(BSG_ERR_NOMEM, "bsg_plan_set_batch: wavefunction slices");
    (void)hipFree(p->d_cs);
    p->d_cs = nc;
  }
  p->batch = nroot;
  return BSG_OK;
}

double bsg_plan_flops(const bsg_plan* p) { return p ? p->flops : 0.0; }
double bsg_plan_split_share(const bsg_plan* p) { return p && p->flops > 0.0 ? p->flops_split / p->flops : 0.0; }
int bsg_plan_phase_flops(const bsg_plan* p, double* f1, double* f2) {
  if (!p || !f1 || !f2) return fail(BSG_ERR_ARG, "bsg_plan_phase_flops: bad args");
  *f1 = p->flops1;
  *f2 = p->flops2;
  return BSG_OK;
}

}  // extern "C"

namespace {
// Launch one chunk for rs.n roots: the grouped kernels on the context
// stream with the side-stream kernels beside them, then the ordered
// reduction.  ev (optional) receives 5 events: before the fused kernels,
// around the side stream's kernels, before and after the reduction.
// launches (optional) counts the kernel dispatches.
int enqueue_chunk(bsg_context* ctx, const bsg_plan* p, const PlanChunk& ch, const double* arena, const double* c,
                  double* v, hipEvent_t* ev, bool split, double* part, const RootSet& rs, long* launches = nullptr) {
  hipStream_t st = ctx->stream;
  auto grid = [&](int n) { return dim3((unsigned)(((n + 7) / 8) * 8 * rs.n)); };
  long nl = 0;
  if (ev) HIPCHK(hipEventRecord(ev[0], st));
  // the side stream runs the small-block kernel and the basic tiles beside
  // the wide tiles on the context stream; they write disjoint V blocks and
  // partial tiles, and the reduction waits for both
  const int nbasic = split ? ch.n_gbf[0] : ch.n_g[0];
  const bool side = ch.n_s[0] > 0 || ch.n_s[1] > 0 || nbasic > 0;
  if (side) {
    HIPCHK(hipEventRecord(ctx->fork, st));
    HIPCHK(hipStreamWaitEvent(ctx->side, ctx->fork, 0));
    if (ev) HIPCHK(hipEventRecord(ev[1], ctx->side));
    if (ch.n_s[0] > 0) {
      hipLaunchKernelGGL((k_small<2>), grid(ch.n_s[0]), dim3(256), 0, ctx->side, ch.d_vb, ch.d_pr, ch.d_seg, ch.d_s[0],
                         ch.n_s[0], arena, c, v, part, rs);
      ++nl;
    }
    if (ch.n_s[1] > 0) {
      hipLaunchKernelGGL((k_small<1>), grid(ch.n_s[1]), dim3(256), 0, ctx->side, ch.d_vb, ch.d_pr, ch.d_seg, ch.d_s[1],
                         ch.n_s[1], arena, c, v, part, rs);
      ++nl;
    }
    if (nbasic > 0) {
      if (split)
        hipLaunchKernelGGL((k_sigma_bf<2, 2>), grid(nbasic), dim3(256), 0, ctx->side, ch.d_vb, ch.d_pr, ch.d_ps,
                           ch.d_seg, ch.d_gbf[0], nbasic, p->d_abs, p->d_cs, c, v, part, rs);
      else
        hipLaunchKernelGGL((k_sigma<2, 2>), grid(nbasic), dim3(256), 0, ctx->side, ch.d_vb, ch.d_pr, ch.d_seg,
                           ch.d_g[0], nbasic, arena, c, v, part, rs);
      ++nl;
    }
    if (ev) HIPCHK(hipEventRecord(ev[2], ctx->side));
    HIPCHK(hipEventRecord(ctx->join, ctx->side));
  }
  if (split) {
    if (ch.n_gbf[2] > 0) {
      hipLaunchKernelGGL((k_sigma_bf<1, 4, 8>), grid(ch.n_gbf[2]), dim3(512), 0, st, ch.d_vb, ch.d_pr, ch.d_ps,
                         ch.d_seg, ch.d_gbf[2], ch.n_gbf[2], p->d_abs, p->d_cs, c, v, part, rs);
      ++nl;
    }
    if (ch.n_gbf[1] > 0) {
      hipLaunchKernelGGL((k_sigma_bf<1, 4>), grid(ch.n_gbf[1]), dim3(256), 0, st, ch.d_vb, ch.d_pr, ch.d_ps, ch.d_seg,
                         ch.d_gbf[1], ch.n_gbf[1], p->d_abs, p->d_cs, c, v, part, rs);
      ++nl;
    }
  } else {
    if (ch.n_w[1] > 0) {
      hipLaunchKernelGGL((k_wave<4>), grid(ch.n_w[1] / 4), dim3(256), 0, st, ch.d_vb, ch.d_pr, ch.d_seg, ch.d_w[1],
                         ch.n_w[1], arena, c, v, part, rs);
      ++nl;
    }
    if (ch.n_w[0] > 0) {
      hipLaunchKernelGGL((k_wave<2>), grid(ch.n_w[0] / 4), dim3(256), 0, st, ch.d_vb, ch.d_pr, ch.d_seg, ch.d_w[0],
                         ch.n_w[0], arena, c, v, part, rs);
      ++nl;
    }
    if (ch.n_g[3] > 0) {
      hipLaunchKernelGGL((k_sigma<1, 6, PB, 8>), grid(ch.n_g[3]), dim3(512), 0, st, ch.d_vb, ch.d_pr, ch.d_seg,
                         ch.d_g[3], ch.n_g[3], arena, c, v, part, rs);
      ++nl;
    }
    if (ch.n_g[2] > 0) {
      hipLaunchKernelGGL((k_sigma<1, 6>), grid(ch.n_g[2]), dim3(256), 0, st, ch.d_vb, ch.d_pr, ch.d_seg, ch.d_g[2],
                         ch.n_g[2], arena, c, v, part, rs);
      ++nl;
    }
    if (ch.n_g[1] > 0) {
      hipLaunchKernelGGL((k_sigma<1, 4>), grid(ch.n_g[1]), dim3(256), 0, st, ch.d_vb, ch.d_pr, ch.d_seg, ch.d_g[1],
                         ch.n_g[1], arena, c, v, part, rs);
      ++nl;
    }
  }
  if (side) HIPCHK(hipStreamWaitEvent(st, ctx->join, 0));
  if (ev) HIPCHK(hipEventRecord(ev[3], st));
  if (ch.n_red[split] > 0) {
    hipLaunchKernelGGL(k_reduce, dim3(ch.n_red[split], rs.n), dim3(256), 0, st, ch.d_red[split], part, v, rs);
    ++nl;
  }
  if (ev) HIPCHK(hipEventRecord(ev[4], st));
  HIPCHK(hipGetLastError());
  if (launches) *launches += nl;
  return BSG_OK;
}

int check_precision(const bsg_plan* p, const double* arena, int32_t precision) {
  if (precision == BSG_FP64) return BSG_OK;
  if (precision != BSG_SPLIT_BF16) return fail(BSG_ERR_ARG, "unknown precision mode");
  if (p->split_src != arena || !p->slices_built || (p->ab_elems > 0 && !p->d_abs) || (p->c_elems > 0 && !p->d_cs))
    return fail(BSG_ERR_ARG, "split precision needs bsg_plan_split_arena on this arena first");
  return BSG_OK;
}
// v += H c for nvec vectors, at most `group` roots per launch, with the
// partial tiles of split V blocks in part (roots part_doubles apart; written
// before they are read within each launch, so no clearing).  launches
// (optional) counts the kernel dispatches; ev (optional) receives 5 events
// per chunk launch.
int sigma_apply_with(bsg_context* ctx, bsg_plan* p, const double* arena, const double* c, double* v, int32_t nvec,
                     int32_t precision, double* part, int group, long* launches = nullptr, hipEvent_t* ev = nullptr) {
  int rc = check_precision(p, arena, precision);
  if (rc) return rc;
  const bool split = precision == BSG_SPLIT_BF16;
  group = std::max(1, std::min(group, p->batch));
  for (int iv = 0; iv < nvec; iv += group) {
    RootSet rs;
    rs.n = std::min(group, nvec - iv);
    rs.pad = 0;
    rs.wstride = p->wave_size;
    rs.pstride = p->part_doubles;
    rs.cstride = p->c_elems;
    const double* cv = c + (size_t)iv * p->wave_size;
    double* vv = v + (size_t)iv * p->wave_size;
    if (split && !p->c_items.empty()) {
      hipLaunchKernelGGL(k_slices, dim3(p->c_items.size(), rs.n), dim3(256), 0, ctx->stream, p->d_c_items, cv,
                         p->d_cs, rs.wstride, rs.cstride);
      if (launches) ++*launches;
    }
    for (auto& ch : p->chunks) {
      if ((rc = enqueue_chunk(ctx, p, ch, arena, cv, vv, ev, split, part, rs, launches))) return rc;
      if (ev) ev += 5;
    }
  }
  return BSG_OK;
}
}  // namespace

extern "C" {

int bsg_sigma_apply(bsg_context* ctx, bsg_plan* p, const double* arena, const double* c, double* v, int32_t nvec,
                    int32_t precision) {
  if (!ctx || !p || !arena || !c || !v || nvec <= 0) return fail(BSG_ERR_ARG, "bsg_sigma_apply: bad args");
  if (p->ctx != ctx) return fail(BSG_ERR_ARG, "bsg_sigma_apply: plan belongs to another context");
  return sigma_apply_with(ctx, p, arena, c, v, nvec, precision, p->part, p->batch);
}

int bsg_plan_capture(bsg_context* ctx, bsg_plan* p, const double* arena, const double* c, double* v, int32_t nvec,
                     int32_t precision) {
  if (!ctx || !p || !arena || !c || !v || nvec <= 0) return fail(BSG_ERR_ARG, "bsg_plan_capture: bad args");
  if (p->ctx != ctx) return fail(BSG_ERR_ARG, "bsg_plan_capture: plan belongs to another context");
  int rc = check_precision(p, arena, precision);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (p->graph) {
    HIPCHK(hipGraphExecDestroy(p->graph));
    p->graph = nullptr;
    p->g_nvec = 0;
  }
  // the launches of one application on the context stream (the side stream
  // joins through the fork/join events and is captured with it); no memset
  // or allocation is recorded
  HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  rc = sigma_apply_with(ctx, p, arena, c, v, nvec, precision, p->part, p->batch);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(BSG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e2 != hipSuccess) return fail(BSG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
  p->graph = x;
  p->g_arena = arena;
  p->g_c = c;
  p->g_v = v;
  p->g_nvec = nvec;
  p->g_precision = precision;
  return BSG_OK;
}

int bsg_plan_replay(bsg_context* ctx, bsg_plan* p) {
  if (!ctx || !p || !p->graph) return fail(BSG_ERR_ARG, "bsg_plan_replay: no captured application");
  if (p->ctx != ctx) return fail(BSG_ERR_ARG, "bsg_plan_replay: plan belongs to another context");
  // the slices the split kernels read must still be those of the captured arena
  int rc = check_precision(p, p->g_arena, p->g_precision);
  if (rc) return rc;
  HIPCHK(hipGraphLaunch(p->graph, ctx->stream));
  return BSG_OK;
}

int bsg_sigma_apply_noalloc(bsg_context* ctx, bsg_plan* p, const double* arena, const double* c, double* v,
                            int32_t nvec, int32_t precision) {
  if (!ctx || !p || !arena || !c || !v || nvec <= 0) return fail(BSG_ERR_ARG, "bsg_sigma_apply_noalloc: bad args");
  if (p->ctx != ctx) return fail(BSG_ERR_ARG, "bsg_sigma_apply_noalloc: plan belongs to another context");
  if (!ctx->scratch) return fail(BSG_ERR_ARG, "bsg_sigma_apply_noalloc: no scratch address for this device");
  if (p->part_doubles > BSG_NOALLOC_SCRATCH_DOUBLES)
    return fail(BSG_ERR_ARG, "bsg_sigma_apply_noalloc: the plan needs " + std::to_string(p->part_doubles) +
                                 " doubles of partial tiles, more than the " +
                                 std::to_string((long long)BSG_NOALLOC_SCRATCH_DOUBLES) + " of the static scratch");
  // as many roots per launch as the scratch holds partial tiles for
  const int group = p->part_doubles > 0 ? (int)std::min<int64_t>(p->batch, BSG_NOALLOC_SCRATCH_DOUBLES / p->part_doubles)
                                        : p->batch;
  return sigma_apply_with(ctx, p, arena, c, v, nvec, precision, ctx->scratch, group);
}

int bsg_sigma_profile(bsg_context* ctx, bsg_plan* p, const double* arena, const double* c, double* v, int32_t nvec,
                      int32_t reps, int32_t precision, double* wall_ms, double* kernel_ms, long* launches) {
  if (!ctx || !p || !arena || !c || !v || nvec <= 0 || reps <= 0 || !wall_ms || !kernel_ms || !launches)
    return fail(BSG_ERR_ARG, "bsg_sigma_profile: bad args");
  int rc = check_precision(p, arena, precision);
  if (rc) return rc;
  const bool split = precision == BSG_SPLIT_BF16;
  const size_t nch = p->chunks.size();
  const int ngroups = (nvec + p->batch - 1) / p->batch;
  constexpr size_t per = 5;  // events per chunk launch (enqueue_chunk)
  std::vector<hipEvent_t> ev((size_t)reps * ngroups * nch * per + 2);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  long disp = 0;
  HIPCHK(hipEventRecord(ev[0], ctx->stream));
  for (int r = 0; r < reps; ++r)
    if ((rc = sigma_apply_with(ctx, p, arena, c, v, nvec, precision, p->part, p->batch, &disp,
                               &ev[1 + (size_t)r * ngroups * nch * per])))
      return rc;
  const size_t q = 1 + (size_t)reps * ngroups * nch * per;
  HIPCHK(hipEventRecord(ev[q], ctx->stream));
  HIPCHK(hipEventSynchronize(ev[q]));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[q]));
  *wall_ms = ms;
  for (int i = 0; i < 3; ++i) {
    kernel_ms[i] = 0.0;
    launches[i] = 0;
  }
  size_t k = 1;
  for (int r = 0; r < reps * ngroups; ++r)
    for (auto& ch : p->chunks) {
      const bool sidek = ch.n_s[0] > 0 || ch.n_s[1] > 0 || (split ? ch.n_gbf[0] : ch.n_g[0]) > 0;
      // [0] fork to join of all fused kernels, [1] reduction, [2] the side
      // stream's kernels alone
      const int from[3] = {0, 3, 1}, to[3] = {3, 4, 2};
      const bool present[3] = {true, ch.n_red[split] > 0, sidek};
      for (int i = 0; i < 3; ++i) {
        if (!present[i]) continue;
        HIPCHK(hipEventElapsedTime(&ms, ev[k + from[i]], ev[k + to[i]]));
        kernel_ms[i] += ms;
      }
      k += per;
    }
  // kernel dispatches of all applications (every kernel launch counted)
  launches[0] = disp;
  for (auto& e : ev) HIPCHK(hipEventDestroy(e));
  return BSG_OK;
}

int bsg_diag_apply(bsg_context* ctx, const bsg_diag_t* terms, int32_t n, const double* arena, double* diag) {
  if (!ctx || (n > 0 && (!terms || !diag))) return fail(BSG_ERR_ARG, "bsg_diag_apply: bad args");
  if (n == 0) return BSG_OK;
  // group terms by target block, keeping their order inside a group
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return terms[a].v_off < terms[b].v_off; });
  std::vector<bsg_diag_t> sorted(n);
  std::vector<DiagGroup> groups;
  for (int i = 0; i < n; ++i) {
    sorted[i] = terms[idx[i]];
    if (i == 0 || sorted[i].v_off != sorted[i - 1].v_off) groups.push_back({i, 0});
    else if (sorted[i].m != sorted[i - 1].m || sorted[i].n != sorted[i - 1].n)
      return fail(BSG_ERR_ARG, "bsg_diag_apply: terms of one block disagree on its shape");
    groups.back().count++;
  }
  std::vector<DiagItem> items;
  for (size_t g = 0; g < groups.size(); ++g) {
    const int64_t mn = (int64_t)sorted[groups[g].first].m * sorted[groups[g].first].n;
    for (int64_t e0 = 0; e0 < mn; e0 += 256) items.push_back({(int32_t)g, (int32_t)e0});
  }
  if (items.empty()) return BSG_OK;
  bsg_diag_t* dt = nullptr;
  DiagGroup* dg = nullptr;
  DiagItem* di = nullptr;
  HIPCHK(hipMalloc(&dt, sizeof(bsg_diag_t) * n));
  HIPCHK(hipMalloc(&dg, sizeof(DiagGroup) * groups.size()));
  HIPCHK(hipMalloc(&di, sizeof(DiagItem) * items.size()));
  HIPCHK(hipMemcpyAsync(dt, sorted.data(), sizeof(bsg_diag_t) * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(dg, groups.data(), sizeof(DiagGroup) * groups.size(), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(di, items.data(), sizeof(DiagItem) * items.size(), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_diag, dim3(items.size()), dim3(256), 0, ctx->stream, dt, dg, di, arena, diag);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(dt));
  HIPCHK(hipFree(dg));
  HIPCHK(hipFree(di));
  return BSG_OK;
}

int bsg_dot(bsg_context* ctx, const double* x, const double* y, int64_t n, double* result) {
  if (!ctx || !x || !y || !result || n < 0) return fail(BSG_ERR_ARG, "bsg_dot: bad args");
  int nb = (int)std::max<int64_t>(1, std::min<int64_t>(RMAXB, (n + RB - 1) / RB));
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(RB), 0, ctx->stream, x, y, n, ctx->red);
  hipLaunchKernelGGL(k_sum_partial, dim3(1), dim3(RB), 0, ctx->stream, ctx->red, nb, ctx->red + RMAXB);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(result, ctx->red + RMAXB, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BSG_OK;
}
int bsg_gs_orthonormalise(bsg_context* ctx, double* r, const double* B, int32_t m, int64_t ld, int64_t n,
                          int32_t normalise) {
  if (!ctx || !r || n < 0 || m < 0 || (m > 0 && (!B || ld < n))) return fail(BSG_ERR_ARG, "bsg_gs_orthonormalise: bad args");
  if (n == 0) return BSG_OK;
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(RMAXB, (n + RB - 1) / RB));
  double* s = ctx->red + RMAXB;
  auto dot = [&](const double* x, const double* y) {
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(RB), 0, ctx->stream, x, y, n, ctx->red);
    hipLaunchKernelGGL(k_sum_partial, dim3(1), dim3(RB), 0, ctx->stream, ctx->red, nb, s);
  };
  for (int32_t j = 0; j < m; ++j) {
    const double* bj = B + (int64_t)j * ld;
    dot(r, r);
    hipLaunchKernelGGL(k_scal_rsqrt_dev, dim3(grid_for(n)), dim3(256), 0, ctx->stream, s, -1.0, r, n);
    dot(r, bj);
    hipLaunchKernelGGL(k_axpy_neg_dev, dim3(grid_for(n)), dim3(256), 0, ctx->stream, s, bj, r, n);
  }
  if (normalise) {
    dot(r, r);
    hipLaunchKernelGGL(k_scal_rsqrt_dev, dim3(grid_for(n)), dim3(256), 0, ctx->stream, s, 1e-15, r, n);
  }
  HIPCHK(hipGetLastError());
  return BSG_OK;
}
int bsg_axpy(bsg_context* ctx, double a, const double* x, double* y, int64_t n) {
  if (!ctx || !x || !y || n < 0) return fail(BSG_ERR_ARG, "bsg_axpy: bad args");
  if (n == 0) return BSG_OK;
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, ctx->stream, a, x, y, n);
  HIPCHK(hipGetLastError());
  return BSG_OK;
}
int bsg_gram(bsg_context* ctx, const double* X, const double* Y, int32_t m, int64_t n, int64_t ld, double* out) {
  if (!ctx || !X || !Y || !out || m <= 0 || m > GRAM_MAX || n < 0 || ld < n)
    return fail(BSG_ERR_ARG, "bsg_gram: bad args");
  const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(GBLK, (n + GCH - 1) / GCH));
  double* part = ctx->gpart;
  double* res = ctx->gpart + (size_t)GBLK * GMAX * GMAX;
  // GMAX x GMAX tiles in stream order (the partials are reused tile by tile)
  for (int i0 = 0; i0 < m; i0 += GMAX)
    for (int j0 = 0; j0 < m; j0 += GMAX) {
      const int mx = std::min(GMAX, m - i0), my = std::min(GMAX, m - j0);
      hipLaunchKernelGGL(k_gram_partial, dim3(nblk), dim3(256), 0, ctx->stream, X + (int64_t)i0 * ld,
                         Y + (int64_t)j0 * ld, mx, my, n, ld, part);
      hipLaunchKernelGGL(k_gram_sum, dim3((mx * my + 255) / 256), dim3(256), 0, ctx->stream, part, nblk, mx, my,
                         res + (int64_t)i0 * m + j0, m);
    }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, res, sizeof(double) * m * m, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BSG_OK;
}
int bsg_lincomb(bsg_context* ctx, const double* X, int32_t m, int64_t n, int64_t ld, const double* coef, int32_t mo,
                double* Y, int64_t ldy) {
  if (!ctx || !X || !coef || !Y || m <= 0 || m > GRAM_MAX || mo <= 0 || mo > GRAM_MAX || n < 0 || ld < n || ldy < n)
    return fail(BSG_ERR_ARG, "bsg_lincomb: bad args");
  double* cd = ctx->gpart + (size_t)GBLK * GMAX * GMAX + (size_t)GRAM_MAX * GRAM_MAX;
  // coefficients through page-locked staging, so the call returns without
  // waiting for the kernels (the staging is reused once its last copy is done)
  HIPCHK(hipEventSynchronize(ctx->coef_done));
  std::memcpy(ctx->coef_pin, coef, sizeof(double) * m * mo);
  HIPCHK(hipMemcpyAsync(cd, ctx->coef_pin, sizeof(double) * m * mo, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipEventRecord(ctx->coef_done, ctx->stream));
  // GMAX x GMAX coefficient tiles; the first row tile of each output tile
  // overwrites Y, the others add in order
  for (int j0 = 0; j0 < mo; j0 += GMAX)
    for (int i0 = 0; i0 < m; i0 += GMAX)
      hipLaunchKernelGGL(k_lincomb, dim3(grid_for(n)), dim3(256), 0, ctx->stream, X + (int64_t)i0 * ld,
                         std::min(GMAX, m - i0), n, ld, cd + (int64_t)i0 * mo + j0, mo, std::min(GMAX, mo - j0),
                         Y + (int64_t)j0 * ldy, ldy, i0 > 0 ? 1 : 0);
  HIPCHK(hipGetLastError());
  return BSG_OK;
}
int bsg_copy(bsg_context* ctx, const double* x, double* y, int64_t n) {
  if (!ctx || !x || !y || n < 0) return fail(BSG_ERR_ARG, "bsg_copy: bad args");
  if (n == 0) return BSG_OK;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(256), 0, ctx->stream, x, y, n);
  HIPCHK(hipGetLastError());
  return BSG_OK;
}
int bsg_copy_segments(bsg_context* ctx, const bsg_copy_seg_t* segs, int32_t n_segs) {
  if (!ctx || n_segs < 0 || (n_segs > 0 && !segs)) return fail(BSG_ERR_ARG, "bsg_copy_segments: bad args");
  std::vector<bsg_copy_seg_t> chunks;
  for (int32_t i = 0; i < n_segs; ++i) {
    const bsg_copy_seg_t& s = segs[i];
    if (s.n < 0 || (s.n > 0 && (!s.src || !s.dst))) return fail(BSG_ERR_ARG, "bsg_copy_segments: bad segment");
    for (int64_t x = 0; x < s.n; x += COPY_CHUNK) chunks.push_back({s.src + x, s.dst + x, std::min(COPY_CHUNK, s.n - x)});
  }
  if (chunks.empty()) return BSG_OK;
  if (chunks.size() > (size_t)0x7fffffff) return fail(BSG_ERR_ARG, "bsg_copy_segments: too many chunks");
  HIPCHK(hipSetDevice(ctx->device));
  // every call returns with its copies done, so no queued work reads the list
  if (ctx->copy_cap < chunks.size()) {
    if (ctx->copy_list) HIPCHK(hipFree(ctx->copy_list));
    ctx->copy_list = nullptr;
    ctx->copy_cap = 0;
    const size_t cap = chunks.size() + chunks.size() / 2;
    HIPCHK(hipMalloc((void**)&ctx->copy_list, sizeof(bsg_copy_seg_t) * cap));
    ctx->copy_cap = cap;
  }
  HIPCHK(hipMemcpyAsync(ctx->copy_list, chunks.data(), sizeof(bsg_copy_seg_t) * chunks.size(), hipMemcpyHostToDevice,
                        ctx->stream));
  hipLaunchKernelGGL(k_copy_chunks, dim3((unsigned)chunks.size()), dim3(256), 0, ctx->stream, ctx->copy_list);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BSG_OK;
}
int bsg_build_apply(bsg_context* ctx, const bsg_build_group_t* groups, int32_t ngroups, const bsg_build_src_t* srcs,
                    int32_t nsrcs) {
  if (!ctx || ngroups < 0 || nsrcs < 0 || (ngroups > 0 && (!groups || !srcs))) return fail(BSG_ERR_ARG, "bsg_build_apply: bad args");
  if (ngroups == 0) return BSG_OK;
  std::vector<int64_t> estart(ngroups);
  int64_t nelem = 0;
  for (int32_t g = 0; g < ngroups; ++g) {
    const bsg_build_group_t& gr = groups[g];
    if (!gr.dst || gr.rows <= 0 || gr.cols <= 0 || gr.ld < gr.cols || gr.first < 0 || gr.count < 0 ||
        gr.first + gr.count > nsrcs)
      return fail(BSG_ERR_ARG, "bsg_build_apply: bad group");
    estart[g] = nelem;
    nelem += (int64_t)gr.rows * gr.cols;
  }
  for (int32_t j = 0; j < nsrcs; ++j)
    if (!srcs[j].src || srcs[j].ld <= 0) return fail(BSG_ERR_ARG, "bsg_build_apply: bad source");
  if (nelem == 0) return BSG_OK;
  if ((nelem + 255) / 256 > 0x7fffffffLL) return fail(BSG_ERR_ARG, "bsg_build_apply: too many elements");
  HIPCHK(hipSetDevice(ctx->device));
  bsg_build_group_t* dg = nullptr;
  bsg_build_src_t* ds = nullptr;
  int64_t* de = nullptr;
  HIPCHK(hipMalloc(&dg, sizeof(bsg_build_group_t) * ngroups));
  HIPCHK(hipMalloc(&ds, sizeof(bsg_build_src_t) * std::max(1, nsrcs)));
  HIPCHK(hipMalloc(&de, sizeof(int64_t) * ngroups));
  HIPCHK(hipMemcpyAsync(dg, groups, sizeof(bsg_build_group_t) * ngroups, hipMemcpyHostToDevice, ctx->stream));
  if (nsrcs > 0) HIPCHK(hipMemcpyAsync(ds, srcs, sizeof(bsg_build_src_t) * nsrcs, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(de, estart.data(), sizeof(int64_t) * ngroups, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_build, dim3((unsigned)((nelem + 255) / 256)), dim3(256), 0, ctx->stream, dg, de, ngroups, ds,
                     nelem);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(dg));
  HIPCHK(hipFree(ds));
  HIPCHK(hipFree(de));
  return BSG_OK;
}
int bsg_scal(bsg_context* ctx, double a, double* x, int64_t n) {
  if (!ctx || !x || n < 0) return fail(BSG_ERR_ARG, "bsg_scal: bad args");
  if (n == 0) return BSG_OK;
  hipLaunchKernelGGL(k_scal, dim3(grid_for(n)), dim3(256), 0, ctx->stream, a, x, n);
  HIPCHK(hipGetLastError());
  return BSG_OK;
}
int bsg_precondition(bsg_context* ctx, double* x, const double* d, double e, double shift, int64_t n) {
  if (!ctx || !x || !d || n < 0) return fail(BSG_ERR_ARG, "bsg_precondition: bad args");
  if (n == 0) return BSG_OK;
  hipLaunchKernelGGL(k_precond, dim3(grid_for(n)), dim3(256), 0, ctx->stream, x, d, e, shift, n);
  HIPCHK(hipGetLastError());
  return BSG_OK;
}
int bsg_olsen_precondition(bsg_context* ctx, double* r, const double* c, const double* d, double e, double shift,
                           double* work, int64_t n) {
  if (!ctx || !r || !c || !d || !work || n < 0) return fail(BSG_ERR_ARG, "bsg_olsen_precondition: bad args");
  if (n == 0) return BSG_OK;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(256), 0, ctx->stream, c, work, n);
  int rc;
  if ((rc = bsg_precondition(ctx, work, d, e, shift, n))) return rc;
  // the two dot products stay in device memory: no host round trip
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(RMAXB, (n + RB - 1) / RB));
  double* s = ctx->red + RMAXB;
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(RB), 0, ctx->stream, work, r, n, ctx->red);
  hipLaunchKernelGGL(k_sum_partial, dim3(1), dim3(RB), 0, ctx->stream, ctx->red, nb, s + 1);
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(RB), 0, ctx->stream, c, work, n, ctx->red);
  hipLaunchKernelGGL(k_sum_partial, dim3(1), dim3(RB), 0, ctx->stream, ctx->red, nb, s + 2);
  hipLaunchKernelGGL(k_axpy_ratio_dev, dim3(grid_for(n)), dim3(256), 0, ctx->stream, s + 1, s + 2, c, r, n);
  HIPCHK(hipGetLastError());
  return bsg_precondition(ctx, r, d, e, shift, n);
}

}  // extern "C"
