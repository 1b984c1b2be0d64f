// MI355X (gfx950) kernels of the Davidson sigma step and the C-ABI around them.
//
// One H.C is a list of sector GEMM pairs (the flattened two-operator
// TensorMultiply of operatorfunctions.C:485-537):
//   phase 1:  W_t   = alpha_t op(A_t) C_t            (one GEMM per item)
//   phase 2:  V_blk += sum_c alpha_c S_c op(B_c)^T   (all contributions of a
//                                                      V block, in list order)
// Both phases are grouped GEMMs over ragged shapes, tiled 64x64 per 256-thread
// workgroup; each wavefront owns a 32x32 quarter computed with the FP64
// matrix-core instruction v_mfma_f64_16x16x4f64.  Operands are staged through
// LDS in 16-deep K slices.  A phase-2 tile belongs to exactly one workgroup,
// which walks that block's contributions in order, so V is reduced without
// atomics and the result is bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// Tile shapes: a workgroup of WR x WC wavefronts owns a (32 WR) x (32 WC)
// output tile and every wavefront a 32x32 quarter of it, i.e. 2x2 blocks of
// 16x16 accumulated with v_mfma_f64_16x16x4f64.  Two shapes are built:
// 64x64 (4 waves, for large sector blocks) and 32x32 (1 wave, for the many
// small ones); the host picks per GEMM the shape with the least padding.
// Lane layout of v_mfma_f64_16x16x4f64 (checked on gfx950): A element
// (lane%16, lane/16), B element (lane/16, lane%16), result register i holds
// (lane/16 + 4*i, lane%16).
template <int WR, int WC>
struct Shape {
  static constexpr int TM = 32 * WR, TN = 32 * WC, NT = 64 * WR * WC;
  static constexpr int LDA = TK + 1;  // LDS row pitch of the A slice (doubles)
  static constexpr int LDB = TN + 2;  // LDS row pitch of the B slice
};

struct Acc {
  dbl4 r[2][2];
};

__device__ inline void acc_zero(Acc& a) {
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) a.r[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
}

// Stage op(X) rows [row0,row0+TM) x cols [k0,k0+TK) into As (zero outside m x k).
// trans=0: X is m x k row-major (ld = k); trans=1: X is k x m row-major (ld = m).
template <class S>
__device__ inline void load_a(double* As, const double* X, int trans, int m, int k, int row0, int k0, double alpha) {
  if (!trans) {
    for (int e = threadIdx.x; e < S::TM * TK; e += S::NT) {
      int i = e / TK, p = e % TK;
      int gi = row0 + i, gp = k0 + p;
      As[i * S::LDA + p] = (gi < m && gp < k) ? alpha * X[(size_t)gi * k + gp] : 0.0;
    }
  } else {
    for (int e = threadIdx.x; e < S::TM * TK; e += S::NT) {
      int p = e / S::TM, i = e % S::TM;
      int gi = row0 + i, gp = k0 + p;
      As[i * S::LDA + p] = (gi < m && gp < k) ? alpha * X[(size_t)gp * m + gi] : 0.0;
    }
  }
}

// Stage B' rows [k0,k0+TK) x cols [col0,col0+TN) into Bs (zero outside k x n).
// trans=0: B' stored k x n row-major (ld = ldb); trans=1: stored n x k (ld = k).
template <class S>
__device__ inline void load_b(double* Bs, const double* Y, int trans, int k, int n, int ldb, int k0, int col0) {
  if (!trans) {
    for (int e = threadIdx.x; e < TK * S::TN; e += S::NT) {
      int p = e / S::TN, j = e % S::TN;
      int gp = k0 + p, gj = col0 + j;
      Bs[p * S::LDB + j] = (gp < k && gj < n) ? Y[(size_t)gp * ldb + gj] : 0.0;
    }
  } else {
    for (int e = threadIdx.x; e < TK * S::TN; e += S::NT) {
      int j = e / TK, p = e % TK;
      int gp = k0 + p, gj = col0 + j;
      Bs[p * S::LDB + j] = (gp < k && gj < n) ? Y[(size_t)gj * k + gp] : 0.0;
    }
  }
}

template <int WR, int WC>
__device__ inline void wave_origin(int& wr, int& wc) {
  const int wave = threadIdx.x >> 6;
  wr = (wave / WC) * 32;
  wc = (wave % WC) * 32;
}

template <int WR, int WC>
__device__ inline void mma_slice(Acc& acc, const double* As, const double* Bs) {
  using S = Shape<WR, WC>;
  const int lane = threadIdx.x & 63;
  int wr, wc;
  wave_origin<WR, WC>(wr, wc);
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < TK; kk += 4) {
    double a0 = As[(wr + li) * S::LDA + kk + lk];
    double a1 = As[(wr + 16 + li) * S::LDA + kk + lk];
    double b0 = Bs[(kk + lk) * S::LDB + wc + li];
    double b1 = Bs[(kk + lk) * S::LDB + wc + 16 + li];
    acc.r[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc.r[0][0], 0, 0, 0);
    acc.r[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc.r[0][1], 0, 0, 0);
    acc.r[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc.r[1][0], 0, 0, 0);
    acc.r[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc.r[1][1], 0, 0, 0);
  }
}

// (row, col) inside the workgroup tile of accumulator element (bi, bj, i)
template <int WR, int WC>
__device__ inline void acc_pos(int bi, int bj, int i, int& row, int& col) {
  const int lane = threadIdx.x & 63;
  int wr, wc;
  wave_origin<WR, WC>(wr, wc);
  row = wr + bi * 16 + (lane >> 4) + 4 * i;
  col = wc + bj * 16 + (lane & 15);
}

template <int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC) void k_phase1(const bsg_gemm1_t* __restrict__ items,
                                                         const TileRef* __restrict__ tiles, int ntiles,
                                                         const double* __restrict__ arena,
                                                         const double* __restrict__ c, double* __restrict__ w) {
  using S = Shape<WR, WC>;
  __shared__ double As[S::TM * S::LDA];
  __shared__ double Bs[TK * S::LDB];
  const int tid = blockIdx.x;
  if (tid >= ntiles) return;
  const TileRef tr = tiles[tid];
  const bsg_gemm1_t it = items[tr.item];
  const int row0 = tr.tm * S::TM, col0 = tr.tn * S::TN;
  const double* A = arena + it.a_off;
  const double* C = c + it.c_off;
  Acc acc;
  acc_zero(acc);
  for (int k0 = 0; k0 < it.k; k0 += TK) {
    load_a<S>(As, A, it.a_trans, it.m, it.k, row0, k0, it.alpha);
    load_b<S>(Bs, C, 0, it.k, it.n, it.n, k0, col0);
    __syncthreads();
    mma_slice<WR, WC>(acc, As, Bs);
    __syncthreads();
  }
  double* W = w + it.w_off;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int i = 0; i < 4; ++i) {
        int r, cc;
        acc_pos<WR, WC>(bi, bj, i, r, cc);
        r += row0;
        cc += col0;
        if (r < it.m && cc < it.n) W[(size_t)r * it.n + cc] = acc.r[bi][bj][i];
      }
}

// A phase-2 work unit: contributions [first, first+count) of one V block.
// p_off < 0: the unit owns the whole contribution list and accumulates into V;
// otherwise it writes its partial tile to P[p_off...] for k_reduce.
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

template <int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC) void k_phase2(const bsg_vblock_t* __restrict__ vbs,
                                                         const bsg_contrib_t* __restrict__ cts,
                                                         const Seg* __restrict__ segs,
                                                         const TileRef* __restrict__ tiles, int ntiles,
                                                         const double* __restrict__ arena,
                                                         const double* __restrict__ c,
                                                         const double* __restrict__ w, double* __restrict__ v,
                                                         double* __restrict__ part) {
  using S = Shape<WR, WC>;
  __shared__ double As[S::TM * S::LDA];
  __shared__ double Bs[TK * S::LDB];
  const int tid = blockIdx.x;
  if (tid >= ntiles) return;
  const TileRef tr = tiles[tid];
  const Seg sg = segs[tr.item];
  const bsg_vblock_t vb = vbs[sg.vb];
  const int row0 = tr.tm * S::TM, col0 = tr.tn * S::TN;
  Acc acc;
  acc_zero(acc);
  for (int q = 0; q < sg.count; ++q) {
    const bsg_contrib_t ct = cts[sg.first + q];
    const double* Sx = (ct.src_kind == 0 ? w : c) + ct.src_off;
    if (ct.b_off < 0) {  // identity partner: V += alpha * S (k == n)
      for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj)
          for (int i = 0; i < 4; ++i) {
            int r, cc;
            acc_pos<WR, WC>(bi, bj, i, r, cc);
            r += row0;
            cc += col0;
            if (r < vb.m && cc < vb.n) acc.r[bi][bj][i] += ct.alpha * Sx[(size_t)r * ct.k + cc];
          }
      continue;
    }
    const double* B = arena + ct.b_off;
    for (int k0 = 0; k0 < ct.k; k0 += TK) {
      load_a<S>(As, Sx, 0, vb.m, ct.k, row0, k0, ct.alpha);
      load_b<S>(Bs, B, ct.b_trans, ct.k, vb.n, vb.n, k0, col0);
      __syncthreads();
      mma_slice<WR, WC>(acc, As, Bs);
      __syncthreads();
    }
  }
  const bool direct = sg.p_off < 0;
  double* out = direct ? v + vb.v_off : part + sg.p_off;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int i = 0; i < 4; ++i) {
        int r, cc;
        acc_pos<WR, WC>(bi, bj, i, r, cc);
        r += row0;
        cc += col0;
        if (r < vb.m && cc < vb.n) {
          if (direct)
            out[(size_t)r * vb.n + cc] += acc.r[bi][bj][i];
          else
            out[(size_t)r * vb.n + cc] = acc.r[bi][bj][i];
        }
      }
}

__global__ __launch_bounds__(256) void k_reduce(const RedItem* __restrict__ items, const double* __restrict__ part,
                                                double* __restrict__ v) {
  const RedItem it = items[blockIdx.x];
  const int e = threadIdx.x;
  if (e >= it.len) return;
  const int64_t x = it.e0 + e;
  double s = 0.0;
  for (int q = 0; q < it.nseg; ++q) s += part[it.p_off + (int64_t)q * it.mn + x];
  v[it.v_off + x] += s;
}

// tile-shape choice: the shape with the smaller padded area (ties: 64x64)
inline int pick_shape(int m, int n) {
  long a64 = (long)((m + 63) / 64) * ((n + 63) / 64) * 4096;
  long a32 = (long)((m + 31) / 32) * ((n + 31) / 32) * 1024;
  return a32 < a64 ? 1 : 0;
}
inline int shape_tm(int s) { return s ? 32 : 64; }

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

struct DiagGroup {
  int32_t first, count;
};
__global__ __launch_bounds__(256) void k_diag(const bsg_diag_t* __restrict__ terms, const DiagGroup* __restrict__ groups,
                                              const double* __restrict__ arena, double* __restrict__ diag) {
  const DiagGroup g = groups[blockIdx.x];
  for (int q = 0; q < g.count; ++q) {
    const bsg_diag_t t = terms[g.first + q];
    const int64_t tot = (int64_t)t.m * t.n;
    for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
      int i = (int)(e / t.n), j = (int)(e % t.n);
      double a = t.a_off < 0 ? 1.0 : arena[t.a_off + (int64_t)i * (t.m + 1)];
      double b = t.b_off < 0 ? 1.0 : arena[t.b_off + (int64_t)j * (t.n + 1)];
      diag[t.v_off + e] += t.alpha * a * b;
    }
    __syncthreads();
  }
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(2048, (n + 255) / 256)); }

}  // namespace

// ---------------------------------------------------------------- context and plan
struct bsg_context {
  int device = 0;
  hipStream_t stream = nullptr;
  double* red = nullptr;  // RMAXB partials + 1 result
};

struct PlanChunk {
  bsg_gemm1_t* d_g1 = nullptr;
  TileRef* d_t1[2] = {nullptr, nullptr};  // per tile shape (0: 64x64, 1: 32x32)
  int n_t1[2] = {0, 0};
  bsg_vblock_t* d_vb = nullptr;
  bsg_contrib_t* d_ct = nullptr;
  Seg* d_seg = nullptr;
  TileRef* d_t2[2] = {nullptr, nullptr};
  int n_t2[2] = {0, 0};
  RedItem* d_red = nullptr;
  int n_red = 0;
};

struct bsg_plan {
  bsg_context* ctx = nullptr;
  std::vector<PlanChunk> chunks;
  double* w = nullptr;
  double* part = nullptr;
  int64_t wave_size = 0;
  double flops = 0.0, flops1 = 0.0, flops2 = 0.0;
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
  e = hipMalloc(&c->red, sizeof(double) * (RMAXB + 1));
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(BSG_ERR_NOMEM, "bsg_init: reduction buffer");
  }
  *ctx = c;
  return BSG_OK;
}

int bsg_free(bsg_context* ctx) {
  if (!ctx) return BSG_OK;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(ctx->red));
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
  int64_t wmax = 0, pmax = 0;
  const int KSEG = 512;  // contraction depth per phase-2 work unit
  auto cleanup = [&](int code, const std::string& m) {
    bsg_plan_destroy(p);
    return fail(code, m);
  };
  for (int ci = 0; ci < n_chunks; ++ci) {
    const bsg_chunk_t& ch = chunks[ci];
    PlanChunk pc;
    // validate and tile
    std::vector<TileRef> t1[2], t2[2];
    for (int i = 0; i < ch.n_gemm1; ++i) {
      const auto& g = ch.gemm1[i];
      if (g.m <= 0 || g.n <= 0 || g.k <= 0 || g.w_off < 0 || g.w_off + (int64_t)g.m * g.n > ch.w_size ||
          g.c_off < 0 || g.c_off + (int64_t)g.k * g.n > wave_size || g.a_off < 0)
        return cleanup(BSG_ERR_ARG, "bsg_plan_create: gemm1 item out of range");
      const int sh = pick_shape(g.m, g.n), T = shape_tm(sh);
      for (int tm = 0; tm < (g.m + T - 1) / T; ++tm)
        for (int tn = 0; tn < (g.n + T - 1) / T; ++tn) t1[sh].push_back({i, tm, tn, 0});
      p->flops += 2.0 * g.m * g.n * g.k;
      p->flops1 += 2.0 * g.m * g.n * g.k;
    }
    std::vector<Seg> segs;
    std::vector<RedItem> reds;
    int64_t poff = 0;
    for (int i = 0; i < ch.n_vblocks; ++i) {
      const auto& vb = ch.vblocks[i];
      if (vb.m <= 0 || vb.n <= 0 || vb.v_off < 0 || vb.v_off + (int64_t)vb.m * vb.n > wave_size || vb.first < 0 ||
          vb.count < 0 || vb.first + vb.count > ch.n_contribs)
        return cleanup(BSG_ERR_ARG, "bsg_plan_create: vblock out of range");
      for (int q = 0; q < vb.count; ++q) {
        const auto& ct = ch.contribs[vb.first + q];
        int64_t lim = ct.src_kind == 0 ? ch.w_size : wave_size;
        if (ct.k <= 0 || ct.src_off < 0 || ct.src_off + (int64_t)vb.m * ct.k > lim || (ct.b_off < 0 && ct.k != vb.n))
          return cleanup(BSG_ERR_ARG, "bsg_plan_create: contribution out of range");
        if (ct.b_off >= 0) {
          p->flops += 2.0 * vb.m * vb.n * ct.k;
          p->flops2 += 2.0 * vb.m * vb.n * ct.k;
        }
      }
      // split the contribution list into units of about KSEG contraction depth
      std::vector<std::pair<int, int>> units;
      int start = 0, depth = 0;
      for (int q = 0; q < vb.count; ++q) {
        const auto& ct = ch.contribs[vb.first + q];
        depth += ct.b_off < 0 ? TK : ct.k;
        if (depth >= KSEG && q + 1 < vb.count) {
          units.push_back({start, q + 1 - start});
          start = q + 1;
          depth = 0;
        }
      }
      units.push_back({start, vb.count - start});
      const int nseg = (int)units.size();
      const int64_t mn = (int64_t)vb.m * vb.n;
      for (int u = 0; u < nseg; ++u) {
        Seg sg;
        sg.vb = i;
        sg.first = vb.first + units[u].first;
        sg.count = units[u].second;
        sg.pad = 0;
        sg.p_off = nseg == 1 ? -1 : poff + (int64_t)u * mn;
        int si = (int)segs.size();
        segs.push_back(sg);
        const int sh = pick_shape(vb.m, vb.n), T = shape_tm(sh);
        for (int tm = 0; tm < (vb.m + T - 1) / T; ++tm)
          for (int tn = 0; tn < (vb.n + T - 1) / T; ++tn) t2[sh].push_back({si, tm, tn, 0});
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
          reds.push_back(r);
        }
        poff += nseg * mn;
      }
    }
    wmax = std::max(wmax, ch.w_size);
    pmax = std::max(pmax, poff);
    auto up = [&](void** d, const void* h, size_t bytes) -> int {
      if (bytes == 0) return BSG_OK;
      HIPCHK(hipMalloc(d, bytes));
      HIPCHK(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
      return BSG_OK;
    };
    int rc = BSG_OK;
    if ((rc = up((void**)&pc.d_g1, ch.gemm1, sizeof(bsg_gemm1_t) * ch.n_gemm1)) ||
        (rc = up((void**)&pc.d_t1[0], t1[0].data(), sizeof(TileRef) * t1[0].size())) ||
        (rc = up((void**)&pc.d_t1[1], t1[1].data(), sizeof(TileRef) * t1[1].size())) ||
        (rc = up((void**)&pc.d_vb, ch.vblocks, sizeof(bsg_vblock_t) * ch.n_vblocks)) ||
        (rc = up((void**)&pc.d_ct, ch.contribs, sizeof(bsg_contrib_t) * ch.n_contribs)) ||
        (rc = up((void**)&pc.d_seg, segs.data(), sizeof(Seg) * segs.size())) ||
        (rc = up((void**)&pc.d_red, reds.data(), sizeof(RedItem) * reds.size())) ||
        (rc = up((void**)&pc.d_t2[0], t2[0].data(), sizeof(TileRef) * t2[0].size())) ||
        (rc = up((void**)&pc.d_t2[1], t2[1].data(), sizeof(TileRef) * t2[1].size()))) {
      p->chunks.push_back(pc);
      std::string m = g_err;
      return cleanup(rc, m);
    }
    for (int sh = 0; sh < 2; ++sh) {
      pc.n_t1[sh] = (int)t1[sh].size();
      pc.n_t2[sh] = (int)t2[sh].size();
    }
    pc.n_red = (int)reds.size();
    p->chunks.push_back(pc);
  }
  if (wmax > 0) {
    hipError_t e = hipMalloc(&p->w, sizeof(double) * wmax);
    if (e != hipSuccess) return cleanup(BSG_ERR_NOMEM, "bsg_plan_create: W scratch");
  }
  if (pmax > 0) {
    hipError_t e = hipMalloc(&p->part, sizeof(double) * pmax);
    if (e != hipSuccess) return cleanup(BSG_ERR_NOMEM, "bsg_plan_create: partial-tile scratch");
  }
  *plan = p;
  return BSG_OK;
}

int bsg_plan_destroy(bsg_plan* p) {
  if (!p) return BSG_OK;
  if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
  for (auto& c : p->chunks) {
    (void)hipFree(c.d_g1);
    (void)hipFree(c.d_t1[0]);
    (void)hipFree(c.d_t1[1]);
    (void)hipFree(c.d_vb);
    (void)hipFree(c.d_ct);
    (void)hipFree(c.d_seg);
    (void)hipFree(c.d_red);
    (void)hipFree(c.d_t2[0]);
    (void)hipFree(c.d_t2[1]);
  }
  (void)hipFree(p->w);
  (void)hipFree(p->part);
  delete p;
  return BSG_OK;
}

double bsg_plan_flops(const bsg_plan* p) { return p ? p->flops : 0.0; }
int bsg_plan_phase_flops(const bsg_plan* p, double* f1, double* f2) {
  if (!p || !f1 || !f2) return fail(BSG_ERR_ARG, "bsg_plan_phase_flops: bad args");
  *f1 = p->flops1;
  *f2 = p->flops2;
  return BSG_OK;
}

}  // extern "C"

namespace {
// Launch one chunk: phase 1 (both shapes), phase 2 (both shapes), ordered
// reduction.  ev (optional) receives events before phase 1, phase 2, the
// reduction and after it.
int enqueue_chunk(hipStream_t st, const bsg_plan* p, const PlanChunk& ch, const double* arena, const double* c,
                  double* v, hipEvent_t* ev) {
  if (ev) HIPCHK(hipEventRecord(ev[0], st));
  if (ch.n_t1[0] > 0)
    hipLaunchKernelGGL((k_phase1<2, 2>), dim3(ch.n_t1[0]), dim3(256), 0, st, ch.d_g1, ch.d_t1[0], ch.n_t1[0], arena, c,
                       p->w);
  if (ch.n_t1[1] > 0)
    hipLaunchKernelGGL((k_phase1<1, 1>), dim3(ch.n_t1[1]), dim3(64), 0, st, ch.d_g1, ch.d_t1[1], ch.n_t1[1], arena, c,
                       p->w);
  if (ev) HIPCHK(hipEventRecord(ev[1], st));
  if (ch.n_t2[0] > 0)
    hipLaunchKernelGGL((k_phase2<2, 2>), dim3(ch.n_t2[0]), dim3(256), 0, st, ch.d_vb, ch.d_ct, ch.d_seg, ch.d_t2[0],
                       ch.n_t2[0], arena, c, p->w, v, p->part);
  if (ch.n_t2[1] > 0)
    hipLaunchKernelGGL((k_phase2<1, 1>), dim3(ch.n_t2[1]), dim3(64), 0, st, ch.d_vb, ch.d_ct, ch.d_seg, ch.d_t2[1],
                       ch.n_t2[1], arena, c, p->w, v, p->part);
  if (ev) HIPCHK(hipEventRecord(ev[2], st));
  if (ch.n_red > 0) hipLaunchKernelGGL(k_reduce, dim3(ch.n_red), dim3(256), 0, st, ch.d_red, p->part, v);
  if (ev) HIPCHK(hipEventRecord(ev[3], st));
  HIPCHK(hipGetLastError());
  return BSG_OK;
}
}  // namespace

extern "C" {

int bsg_sigma_apply(bsg_context* ctx, bsg_plan* p, const double* arena, const double* c, double* v, int32_t nvec,
                    int32_t precision) {
  if (!ctx || !p || !arena || !c || !v || nvec <= 0) return fail(BSG_ERR_ARG, "bsg_sigma_apply: bad args");
  if (p->ctx != ctx) return fail(BSG_ERR_ARG, "bsg_sigma_apply: plan belongs to another context");
  if (precision != BSG_FP64) return fail(BSG_ERR_UNSUPPORTED, "bsg_sigma_apply: only BSG_FP64 is implemented");
  for (int iv = 0; iv < nvec; ++iv) {
    const double* cv = c + (size_t)iv * p->wave_size;
    double* vv = v + (size_t)iv * p->wave_size;
    for (auto& ch : p->chunks) {
      int rc = enqueue_chunk(ctx->stream, p, ch, arena, cv, vv, nullptr);
      if (rc) return rc;
    }
  }
  return BSG_OK;
}

int bsg_sigma_profile(bsg_context* ctx, bsg_plan* p, const double* arena, const double* c, double* v, int32_t reps,
                      double* wall_ms, double* kernel_ms, long* launches) {
  if (!ctx || !p || !arena || !c || !v || reps <= 0 || !wall_ms || !kernel_ms || !launches)
    return fail(BSG_ERR_ARG, "bsg_sigma_profile: bad args");
  const size_t nch = p->chunks.size();
  std::vector<hipEvent_t> ev((size_t)reps * nch * 4 + 2);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], ctx->stream));
  size_t q = 1;
  for (int r = 0; r < reps; ++r)
    for (auto& ch : p->chunks) {
      int rc = enqueue_chunk(ctx->stream, p, ch, arena, c, v, &ev[q]);
      if (rc) return rc;
      q += 4;
    }
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
  for (int r = 0; r < reps; ++r)
    for (auto& ch : p->chunks) {
      const int present[3] = {(ch.n_t1[0] > 0) + (ch.n_t1[1] > 0), (ch.n_t2[0] > 0) + (ch.n_t2[1] > 0),
                              ch.n_red > 0};
      for (int i = 0; i < 3; ++i) {
        HIPCHK(hipEventElapsedTime(&ms, ev[k + i], ev[k + i + 1]));
        if (present[i]) {
          kernel_ms[i] += ms;
          launches[i] += present[i];
        }
      }
      k += 4;
    }
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
    groups.back().count++;
  }
  bsg_diag_t* dt = nullptr;
  DiagGroup* dg = nullptr;
  HIPCHK(hipMalloc(&dt, sizeof(bsg_diag_t) * n));
  HIPCHK(hipMalloc(&dg, sizeof(DiagGroup) * groups.size()));
  HIPCHK(hipMemcpyAsync(dt, sorted.data(), sizeof(bsg_diag_t) * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(dg, groups.data(), sizeof(DiagGroup) * groups.size(), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_diag, dim3(groups.size()), dim3(256), 0, ctx->stream, dt, dg, arena, diag);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(dt));
  HIPCHK(hipFree(dg));
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
int bsg_axpy(bsg_context* ctx, double a, const double* x, double* y, int64_t n) {
  if (!ctx || !x || !y || n < 0) return fail(BSG_ERR_ARG, "bsg_axpy: bad args");
  if (n == 0) return BSG_OK;
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, ctx->stream, a, x, y, n);
  HIPCHK(hipGetLastError());
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
  double num = 0.0, den = 0.0;
  if ((rc = bsg_dot(ctx, work, r, n, &num))) return rc;
  if ((rc = bsg_dot(ctx, c, work, n, &den))) return rc;
  if ((rc = bsg_axpy(ctx, -num / den, c, r, n))) return rc;
  return bsg_precondition(ctx, r, d, e, shift, n);
}

}  // extern "C"
