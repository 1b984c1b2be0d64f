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

constexpr int TM = 64, TN = 64, TK = 16, NT = 256;
constexpr int LDA = TK + 1;   // LDS row pitch of the A slice (doubles)
constexpr int LDB = TN + 2;   // LDS row pitch of the B slice

typedef double dbl4 __attribute__((ext_vector_type(4)));

struct TileRef {
  int32_t item, tm, tn, pad;
};

// Accumulator fragment of one wavefront: 2x2 blocks of 16x16.  Lane layout of
// v_mfma_f64_16x16x4f64 (checked on gfx950): A element (lane%16, lane/16),
// B element (lane/16, lane%16), result register i holds (lane/16 + 4*i, lane%16).
struct Acc {
  dbl4 r[2][2];
};

__device__ inline void acc_zero(Acc& a) {
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) a.r[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
}

// Stage op(X) rows [row0,row0+TM) x cols [k0,k0+TK) into As (zero outside m x k).
// trans=0: X is m x k row-major (ld = k); trans=1: X is k x m row-major (ld = m).
__device__ inline void load_a(double* As, const double* X, int trans, int m, int k, int row0, int k0, double alpha) {
  const int t = threadIdx.x;
  if (!trans) {
    const int p = t & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int i = (t >> 4) + 16 * r;
      int gi = row0 + i, gp = k0 + p;
      As[i * LDA + p] = (gi < m && gp < k) ? alpha * X[(size_t)gi * k + gp] : 0.0;
    }
  } else {
    const int i = t & 63;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int p = (t >> 6) + 4 * r;
      int gi = row0 + i, gp = k0 + p;
      As[i * LDA + p] = (gi < m && gp < k) ? alpha * X[(size_t)gp * m + gi] : 0.0;
    }
  }
}

// Stage B' rows [k0,k0+TK) x cols [col0,col0+TN) into Bs (zero outside k x n).
// trans=0: B' stored k x n row-major (ld = ldb); trans=1: stored n x k (ld = k).
__device__ inline void load_b(double* Bs, const double* Y, int trans, int k, int n, int ldb, int k0, int col0) {
  const int t = threadIdx.x;
  if (!trans) {
    const int j = t & 63;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int p = (t >> 6) + 4 * r;
      int gp = k0 + p, gj = col0 + j;
      Bs[p * LDB + j] = (gp < k && gj < n) ? Y[(size_t)gp * ldb + gj] : 0.0;
    }
  } else {
    const int p = t & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int j = (t >> 4) + 16 * r;
      int gp = k0 + p, gj = col0 + j;
      Bs[p * LDB + j] = (gp < k && gj < n) ? Y[(size_t)gj * k + gp] : 0.0;
    }
  }
}

__device__ inline void mma_slice(Acc& acc, const double* As, const double* Bs) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < TK; kk += 4) {
    double a0 = As[(wr + li) * LDA + kk + lk];
    double a1 = As[(wr + 16 + li) * LDA + kk + lk];
    double b0 = Bs[(kk + lk) * LDB + wc + li];
    double b1 = Bs[(kk + lk) * LDB + wc + 16 + li];
    acc.r[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc.r[0][0], 0, 0, 0);
    acc.r[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc.r[0][1], 0, 0, 0);
    acc.r[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc.r[1][0], 0, 0, 0);
    acc.r[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc.r[1][1], 0, 0, 0);
  }
}

// (row, col) inside the 64x64 tile of accumulator element (bi, bj, i)
__device__ inline void acc_pos(int bi, int bj, int i, int& row, int& col) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  row = (wave >> 1) * 32 + bi * 16 + (lane >> 4) + 4 * i;
  col = (wave & 1) * 32 + bj * 16 + (lane & 15);
}

__global__ __launch_bounds__(NT) void k_phase1(const bsg_gemm1_t* __restrict__ items,
                                               const TileRef* __restrict__ tiles, int ntiles,
                                               const double* __restrict__ arena, const double* __restrict__ c,
                                               double* __restrict__ w) {
  __shared__ double As[TM * LDA];
  __shared__ double Bs[TK * LDB];
  const int tid = blockIdx.x;
  if (tid >= ntiles) return;
  const TileRef tr = tiles[tid];
  const bsg_gemm1_t it = items[tr.item];
  const int row0 = tr.tm * TM, col0 = tr.tn * TN;
  const double* A = arena + it.a_off;
  const double* C = c + it.c_off;
  Acc acc;
  acc_zero(acc);
  for (int k0 = 0; k0 < it.k; k0 += TK) {
    load_a(As, A, it.a_trans, it.m, it.k, row0, k0, it.alpha);
    load_b(Bs, C, 0, it.k, it.n, it.n, k0, col0);
    __syncthreads();
    mma_slice(acc, As, Bs);
    __syncthreads();
  }
  double* W = w + it.w_off;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int i = 0; i < 4; ++i) {
        int r, cc;
        acc_pos(bi, bj, i, r, cc);
        r += row0;
        cc += col0;
        if (r < it.m && cc < it.n) W[(size_t)r * it.n + cc] = acc.r[bi][bj][i];
      }
}

__global__ __launch_bounds__(NT) void k_phase2(const bsg_vblock_t* __restrict__ vbs,
                                               const bsg_contrib_t* __restrict__ cts,
                                               const TileRef* __restrict__ tiles, int ntiles,
                                               const double* __restrict__ arena, const double* __restrict__ c,
                                               const double* __restrict__ w, double* __restrict__ v) {
  __shared__ double As[TM * LDA];
  __shared__ double Bs[TK * LDB];
  const int tid = blockIdx.x;
  if (tid >= ntiles) return;
  const TileRef tr = tiles[tid];
  const bsg_vblock_t vb = vbs[tr.item];
  const int row0 = tr.tm * TM, col0 = tr.tn * TN;
  Acc acc;
  acc_zero(acc);
  for (int q = 0; q < vb.count; ++q) {
    const bsg_contrib_t ct = cts[vb.first + q];
    const double* S = (ct.src_kind == 0 ? w : c) + ct.src_off;
    if (ct.b_off < 0) {  // identity partner: V += alpha * S (k == n)
      for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj)
          for (int i = 0; i < 4; ++i) {
            int r, cc;
            acc_pos(bi, bj, i, r, cc);
            r += row0;
            cc += col0;
            if (r < vb.m && cc < vb.n) acc.r[bi][bj][i] += ct.alpha * S[(size_t)r * ct.k + cc];
          }
      continue;
    }
    const double* B = arena + ct.b_off;
    for (int k0 = 0; k0 < ct.k; k0 += TK) {
      load_a(As, S, 0, vb.m, ct.k, row0, k0, ct.alpha);
      load_b(Bs, B, ct.b_trans, ct.k, vb.n, vb.n, k0, col0);
      __syncthreads();
      mma_slice(acc, As, Bs);
      __syncthreads();
    }
  }
  double* V = v + vb.v_off;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int i = 0; i < 4; ++i) {
        int r, cc;
        acc_pos(bi, bj, i, r, cc);
        r += row0;
        cc += col0;
        if (r < vb.m && cc < vb.n) V[(size_t)r * vb.n + cc] += acc.r[bi][bj][i];
      }
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
  TileRef* d_t1 = nullptr;
  int n_t1 = 0;
  bsg_vblock_t* d_vb = nullptr;
  bsg_contrib_t* d_ct = nullptr;
  TileRef* d_t2 = nullptr;
  int n_t2 = 0;
};

struct bsg_plan {
  bsg_context* ctx = nullptr;
  std::vector<PlanChunk> chunks;
  double* w = nullptr;
  int64_t wave_size = 0;
  double flops = 0.0;
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
  int64_t wmax = 0;
  auto cleanup = [&](int code, const std::string& m) {
    bsg_plan_destroy(p);
    return fail(code, m);
  };
  for (int ci = 0; ci < n_chunks; ++ci) {
    const bsg_chunk_t& ch = chunks[ci];
    PlanChunk pc;
    // validate and tile
    std::vector<TileRef> t1, t2;
    for (int i = 0; i < ch.n_gemm1; ++i) {
      const auto& g = ch.gemm1[i];
      if (g.m <= 0 || g.n <= 0 || g.k <= 0 || g.w_off < 0 || g.w_off + (int64_t)g.m * g.n > ch.w_size ||
          g.c_off < 0 || g.c_off + (int64_t)g.k * g.n > wave_size || g.a_off < 0)
        return cleanup(BSG_ERR_ARG, "bsg_plan_create: gemm1 item out of range");
      for (int tm = 0; tm < (g.m + TM - 1) / TM; ++tm)
        for (int tn = 0; tn < (g.n + TN - 1) / TN; ++tn) t1.push_back({i, tm, tn, 0});
      p->flops += 2.0 * g.m * g.n * g.k;
    }
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
        if (ct.b_off >= 0) p->flops += 2.0 * vb.m * vb.n * ct.k;
      }
      for (int tm = 0; tm < (vb.m + TM - 1) / TM; ++tm)
        for (int tn = 0; tn < (vb.n + TN - 1) / TN; ++tn) t2.push_back({i, tm, tn, 0});
    }
    wmax = std::max(wmax, ch.w_size);
    auto up = [&](void** d, const void* h, size_t bytes) -> int {
      if (bytes == 0) return BSG_OK;
      HIPCHK(hipMalloc(d, bytes));
      HIPCHK(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
      return BSG_OK;
    };
    int rc = BSG_OK;
    if ((rc = up((void**)&pc.d_g1, ch.gemm1, sizeof(bsg_gemm1_t) * ch.n_gemm1)) ||
        (rc = up((void**)&pc.d_t1, t1.data(), sizeof(TileRef) * t1.size())) ||
        (rc = up((void**)&pc.d_vb, ch.vblocks, sizeof(bsg_vblock_t) * ch.n_vblocks)) ||
        (rc = up((void**)&pc.d_ct, ch.contribs, sizeof(bsg_contrib_t) * ch.n_contribs)) ||
        (rc = up((void**)&pc.d_t2, t2.data(), sizeof(TileRef) * t2.size()))) {
      p->chunks.push_back(pc);
      std::string m = g_err;
      return cleanup(rc, m);
    }
    pc.n_t1 = (int)t1.size();
    pc.n_t2 = (int)t2.size();
    p->chunks.push_back(pc);
  }
  if (wmax > 0) {
    hipError_t e = hipMalloc(&p->w, sizeof(double) * wmax);
    if (e != hipSuccess) return cleanup(BSG_ERR_NOMEM, "bsg_plan_create: W scratch");
  }
  *plan = p;
  return BSG_OK;
}

int bsg_plan_destroy(bsg_plan* p) {
  if (!p) return BSG_OK;
  if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
  for (auto& c : p->chunks) {
    (void)hipFree(c.d_g1);
    (void)hipFree(c.d_t1);
    (void)hipFree(c.d_vb);
    (void)hipFree(c.d_ct);
    (void)hipFree(c.d_t2);
  }
  (void)hipFree(p->w);
  delete p;
  return BSG_OK;
}

double bsg_plan_flops(const bsg_plan* p) { return p ? p->flops : 0.0; }

int bsg_sigma_apply(bsg_context* ctx, bsg_plan* p, const double* arena, const double* c, double* v, int32_t nvec,
                    int32_t precision) {
  if (!ctx || !p || !arena || !c || !v || nvec <= 0) return fail(BSG_ERR_ARG, "bsg_sigma_apply: bad args");
  if (p->ctx != ctx) return fail(BSG_ERR_ARG, "bsg_sigma_apply: plan belongs to another context");
  if (precision != BSG_FP64) return fail(BSG_ERR_UNSUPPORTED, "bsg_sigma_apply: only BSG_FP64 is implemented");
  for (int iv = 0; iv < nvec; ++iv) {
    const double* cv = c + (size_t)iv * p->wave_size;
    double* vv = v + (size_t)iv * p->wave_size;
    for (auto& ch : p->chunks) {
      if (ch.n_t1 > 0)
        hipLaunchKernelGGL(k_phase1, dim3(ch.n_t1), dim3(NT), 0, ctx->stream, ch.d_g1, ch.d_t1, ch.n_t1, arena, cv, p->w);
      if (ch.n_t2 > 0)
        hipLaunchKernelGGL(k_phase2, dim3(ch.n_t2), dim3(NT), 0, ctx->stream, ch.d_vb, ch.d_ct, ch.d_t2, ch.n_t2, arena,
                           cv, p->w, vv);
    }
  }
  HIPCHK(hipGetLastError());
  return BSG_OK;
}

int bsg_sigma_profile(bsg_context* ctx, bsg_plan* p, const double* arena, const double* c, double* v, int32_t reps,
                      double* wall_ms, double* kernel_ms, long* launches) {
  if (!ctx || !p || !arena || !c || !v || reps <= 0 || !wall_ms || !kernel_ms || !launches)
    return fail(BSG_ERR_ARG, "bsg_sigma_profile: bad args");
  const size_t nch = p->chunks.size();
  std::vector<hipEvent_t> ev((size_t)reps * nch * 3 + 2);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  size_t q = 0;
  HIPCHK(hipEventRecord(ev[q++], ctx->stream));
  for (int r = 0; r < reps; ++r)
    for (auto& ch : p->chunks) {
      HIPCHK(hipEventRecord(ev[q++], ctx->stream));
      if (ch.n_t1 > 0)
        hipLaunchKernelGGL(k_phase1, dim3(ch.n_t1), dim3(NT), 0, ctx->stream, ch.d_g1, ch.d_t1, ch.n_t1, arena, c, p->w);
      HIPCHK(hipEventRecord(ev[q++], ctx->stream));
      if (ch.n_t2 > 0)
        hipLaunchKernelGGL(k_phase2, dim3(ch.n_t2), dim3(NT), 0, ctx->stream, ch.d_vb, ch.d_ct, ch.d_t2, ch.n_t2, arena,
                           c, p->w, v);
      HIPCHK(hipEventRecord(ev[q++], ctx->stream));
    }
  HIPCHK(hipEventRecord(ev[q++], ctx->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventSynchronize(ev[q - 1]));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[q - 1]));
  *wall_ms = ms;
  kernel_ms[0] = kernel_ms[1] = 0.0;
  launches[0] = launches[1] = 0;
  size_t k = 1;
  for (int r = 0; r < reps; ++r)
    for (auto& ch : p->chunks) {
      HIPCHK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      if (ch.n_t1 > 0) {
        kernel_ms[0] += ms;
        launches[0]++;
      }
      HIPCHK(hipEventElapsedTime(&ms, ev[k + 1], ev[k + 2]));
      if (ch.n_t2 > 0) {
        kernel_ms[1] += ms;
        launches[1]++;
      }
      k += 3;
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
