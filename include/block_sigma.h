/* C-ABI of the MI355X Davidson sigma path (libblock_sigma.so).
 *
 * Plain pointers and sizes only.  Every entry point returns BSG_OK (0) or a
 * negative error code; bsg_last_error() gives the message of the last failure
 * on the calling thread.  All work of one context runs on that context's
 * single HIP stream, in call order (the single-stream contract).
 *
 * Reference interfaces these replace (sanshar/Block 1.1.1):
 *   multiply_h::operator()(Wavefunction& c, Wavefunction& v)   davidson.h:22-29, davidson.C:17-22
 *     -> bsg_sigma_apply  (v += H.c, accumulate semantics as in
 *        operatorfunctions.C:530 and spinblock.C:722-789)
 *   multiply_h(const SpinBlock& b, bool onedot) constructor      davidson.h:26
 *     -> bsg_plan_create (the superblock flattened into GEMM pairs by the host
 *        work-list compiler; operator blocks resident in HBM)
 *   SpinBlock::diagonalH(DiagonalMatrix&)                         spinblock.h:240, spinblock.C:855-899
 *     -> bsg_diag_apply
 *   DotProduct / ScaleAdd / Scale                                 BaseOperator.h:287-291, BaseOperator.C:240-288
 *     -> bsg_dot / bsg_axpy / bsg_scal
 *   Linear::precondition / olsenPrecondition                      linear.C:27-60
 *     -> bsg_precondition / bsg_olsen_precondition
 */
#ifndef BLOCK_SIGMA_H
#define BLOCK_SIGMA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BSG_OK = 0,
  BSG_ERR_ARG = -1,
  BSG_ERR_HIP = -2,
  BSG_ERR_NOMEM = -3,
  BSG_ERR_UNSUPPORTED = -4
};

enum { BSG_FP64 = 0, BSG_SPLIT_BF16 = 1 };

typedef struct bsg_context bsg_context;
typedef struct bsg_plan bsg_plan;

/* One GEMM pair of the flattened TensorMultiply (operatorfunctions.C:485-537)
 * contributing to a V block (m x n):
 *   W = alpha_a * op(A) (m x k1) * C (k1 x k2),   V += alpha * W * B' (k2 x n)
 * A lives in the operator arena at a_off, stored m x k1 row-major (a_trans=0)
 * or k1 x m (a_trans=1); a_off < 0 means A is the identity (W = C, k1 == m).
 * C is the wavefunction block at c_off, k1 x k2 row-major.  B' = op(B)^T from
 * the arena at b_off, stored k2 x n (b_trans=0) or n x k2 (b_trans=1);
 * b_off < 0 means B' is the identity (k2 == n).  W never leaves the chip. */
typedef struct {
  int64_t a_off, c_off, b_off;
  int32_t k1, k2, a_trans, b_trans;
  double alpha_a, alpha;
} bsg_pair_t;

/* V block at v_off (m x n row-major) with pairs [first, first+count), summed
 * in list order (deterministic). */
typedef struct {
  int64_t v_off;
  int32_t m, n, first, count;
} bsg_vblock_t;

/* One chunk of the work list. */
typedef struct {
  const bsg_vblock_t* vblocks;
  int32_t n_vblocks;
  const bsg_pair_t* pairs;
  int32_t n_pairs;
} bsg_chunk_t;

/* Diagonal term: diag[v_off + i*n + j] += alpha * A[a_off + i*(m+1)] * B[b_off + j*(n+1)]
 * for i < m, j < n (a_off/b_off < 0: identity). */
typedef struct {
  int64_t v_off, a_off, b_off;
  int32_t m, n;
  double alpha;
} bsg_diag_t;

const char* bsg_last_error(void);

int bsg_init(bsg_context** ctx, int device);
int bsg_free(bsg_context* ctx);
int bsg_synchronize(bsg_context* ctx);

/* device memory owned by the context's caller */
int bsg_malloc(bsg_context* ctx, size_t bytes, void** dptr);
int bsg_mfree(bsg_context* ctx, void* dptr);
int bsg_memcpy_h2d(bsg_context* ctx, void* dst, const void* src, size_t bytes);
int bsg_memcpy_d2h(bsg_context* ctx, void* dst, const void* src, size_t bytes);
int bsg_memset_zero(bsg_context* ctx, void* dst, size_t bytes);
/* Page-locked host staging for operator uploads (no counterpart in the
 * reference, whose operators never leave host memory), and an upload that
 * returns before the copy ends: src must be page-locked and unchanged until
 * bsg_synchronize. */
int bsg_host_alloc(bsg_context* ctx, size_t bytes, void** hptr);
int bsg_host_free(bsg_context* ctx, void* hptr);
int bsg_memcpy_h2d_async(bsg_context* ctx, void* dst, const void* src, size_t bytes);

/* Build a device plan from host work-list chunks (copied; the host arrays may
 * be freed afterwards).  The plan owns the partial-tile scratch of the
 * ordered split reduction. */
int bsg_plan_create(bsg_context* ctx, const bsg_chunk_t* chunks, int32_t n_chunks, int64_t wave_size,
                    bsg_plan** plan);
int bsg_plan_destroy(bsg_plan* plan);
/* Flops (algorithmic, GEMM pairs) of one application, for reporting. */
double bsg_plan_flops(const bsg_plan* plan);
/* The same split by GEMM: first (W = op(A) C) and second (V += W op(B)^T). */
int bsg_plan_phase_flops(const bsg_plan* plan, double* phase1, double* phase2);
/* Share of the flops that BSG_SPLIT_BF16 runs on BF16 slices (the grouped
 * kernels' pairs; the pairs of V blocks up to 32 x 32 stay FP64, except those
 * with an identity A and an operator B).  No counterpart in the
 * reference; the device Davidson skips the split phase of its precision
 * ladder below a share of one half. */
double bsg_plan_split_share(const bsg_plan* plan);

/* Doubles of operator arena the plan's pairs reach (the arena passed to
 * bsg_sigma_apply must hold at least this many). */
int64_t bsg_plan_arena_extent(const bsg_plan* plan);

/* Split precision set-up: checks that the arena holds the plan's extent and
 * enables BSG_SPLIT_BF16 for this arena.  On its first call for a plan it
 * allocates the plan's BF16 slice buffers: three slices x = x0 + x1 + x2 of
 * the FP32 rounding of x (exact) for every operator block the grouped kernel
 * reads, in each orientation it reads it (6 B per element and orientation,
 * rows padded to a multiple of 8; about 0.75x the FP64 arena), and the
 * transposed slices of the wavefunction's C blocks (6 B per element), which
 * every split-precision application rewrites.  It then slices the arena's
 * operator blocks; call again after the arena moves or its contents change.
 * The small-block kernel's pairs (V blocks of at most 32 x 32 whose A is an
 * operator, or whose A and B are identities) run in FP64 in both modes (it
 * is bound by operand latency, not by its MFMAs).
 * No counterpart in the reference (its sigma is FP64 only). */
int bsg_plan_split_arena(bsg_context* ctx, bsg_plan* plan, const double* arena, int64_t n);

/* Measured tile choices (no counterpart in the reference): n rules of six
 * int32 (m_lo, m_hi, n_lo, n_hi, fp64, split); a V block of m x n inside a
 * rule's ranges gets tile geometry number fp64 (FP64 kernels) or split
 * (split-precision kernels) of the candidate list, -1 leaving the MFMA-count
 * model's choice; the first matching rule applies.  Process-wide, read when
 * plans are created (n = 0 clears).  bsg_tile_candidates gives the list's
 * length for a precision.  For FP64 its last entry is not a geometry but the
 * per-wave tile kernel (one wavefront per 16 x 32 V tile, 16 x 64 for blocks
 * wider than 32 columns; no LDS or barriers)
 * for a block's pairs with k1 <= 64 (80 on the 64-wide tiles) or an identity
 * A; the block's other pairs
 * keep the model's geometry, and split mode keeps the grouped BF16 kernel
 * for the whole block.  Blocks of at most 32 x 32 stay on the small-block
 * kernel whatever the rule.  The tuner that measures them is bench.py --tune;
 * block_amd/tile_table.txt holds the committed result. */
int bsg_set_tile_table(const int32_t* rules, int32_t n);
int bsg_tile_candidates(int32_t precision);

/* Roots per launch (default 1): bsg_sigma_apply runs up to nroot vectors in
 * one set of kernel launches (every tile once per root, a tile's roots on one
 * XCD so they share its operator blocks in L2), the block Davidson's new
 * vectors of one iteration (linear.C:234-258) and state-averaged roots
 * together.  Reallocates the plan's partial tiles and C slices for nroot
 * vectors (after synchronising the context stream). */
int bsg_plan_set_batch(bsg_context* ctx, bsg_plan* plan, int32_t nroot);

/* v += H c for nvec vectors stored back to back (stride wave_size doubles),
 * in launches of up to the plan's batch size (bsg_plan_set_batch).
 * arena: device operator arena; c, v: device pointers.
 * precision: BSG_FP64, or BSG_SPLIT_BF16 (BF16 slice products on MFMA with
 * FP32 accumulation per 32-deep step and FP64 sums for the grouped kernels'
 * pairs, FP64 for the small-block kernel's; needs bsg_plan_split_arena on the same
 * arena first; ~1e-7 relative error). */
int bsg_sigma_apply(bsg_context* ctx, bsg_plan* plan, const double* arena, const double* c, double* v, int32_t nvec,
                    int32_t precision);

/* One application captured as a HIP graph (no counterpart in the reference):
 * bsg_plan_capture records the launches of bsg_sigma_apply for these exact
 * arena, c, v, nvec and precision (both streams, no memset, no allocation)
 * into the plan, replacing any earlier capture; bsg_plan_replay launches it
 * on the context stream (v += H c again; the caller clears v beforehand if it
 * wants v = H c).  The caller keeps the buffers in place between replays; a
 * capture is dropped by bsg_plan_set_batch.  One graph launch instead of the
 * application's kernel launches and stream events. */
int bsg_plan_capture(bsg_context* ctx, bsg_plan* plan, const double* arena, const double* c, double* v, int32_t nvec,
                     int32_t precision);
int bsg_plan_replay(bsg_context* ctx, bsg_plan* plan);

/* Doubles of partial-tile scratch one application of the plan uses (V blocks
 * whose pair list is cut into several work units; 0 if none). */
int64_t bsg_plan_partial_doubles(const bsg_plan* plan);

/* bsg_sigma_apply without any allocation, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid it).  No counterpart in the reference.  The partial tiles live
 * in a static device array of BSG_NOALLOC_SCRATCH_DOUBLES doubles (64 MiB),
 * one per GPU, held until the process ends; a plan whose
 * bsg_plan_partial_doubles exceeds it is rejected with BSG_ERR_ARG.  Results
 * are bitwise those of bsg_sigma_apply.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other contexts' streams, other
 * threads) race and must be serialised by the caller.  The plan itself (and,
 * for BSG_SPLIT_BF16, bsg_plan_split_arena's slice buffers) is set up
 * beforehand by the allocating calls. */
#define BSG_NOALLOC_SCRATCH_DOUBLES (8LL * 1024 * 1024)
int bsg_sigma_apply_noalloc(bsg_context* ctx, bsg_plan* plan, const double* arena, const double* c, double* v,
                            int32_t nvec, int32_t precision);

/* Timing helper: runs `reps` applications of the sigma to nvec vectors
 * (c, v: nvec vectors back to back, roots batched as bsg_sigma_apply does)
 * on the context stream with HIP events around every chunk launch.
 * wall_ms: whole loop; kernel_ms[0..2]: summed time of the fused GEMM-pair
 * kernels (grouped kernels on the context stream and the small-block and
 * basic-tile kernels beside them, fork to join), of the partial-tile
 * reduction and of the side stream's kernels alone; launches[0]: kernel
 * dispatches of all applications (every launch counted), launches[1..2]: 0. */
int bsg_sigma_profile(bsg_context* ctx, bsg_plan* plan, const double* arena, const double* c, double* v, int32_t nvec,
                      int32_t reps, int32_t precision, double* wall_ms, double* kernel_ms, long* launches);

/* diag += diagonal terms (device pointers). */
int bsg_diag_apply(bsg_context* ctx, const bsg_diag_t* terms_host, int32_t n_terms, const double* arena,
                   double* diag);

/* Davidson vector kernels on device vectors of length n. */
int bsg_dot(bsg_context* ctx, const double* x, const double* y, int64_t n, double* result_host);
int bsg_axpy(bsg_context* ctx, double a, const double* x, double* y, int64_t n);
int bsg_scal(bsg_context* ctx, double a, double* x, int64_t n);
/* y = x (the Wavefunction copies of linear.C:241-244) */
int bsg_copy(bsg_context* ctx, const double* x, double* y, int64_t n);

/* Many device-to-device copies in one launch: seg[i].dst[0..n) =
   seg[i].src[0..n) (device pointers; segments must not overlap).  No
   counterpart in the reference: the sweep moves operators between device
   buffers one operator block list at a time (operators still in the arena
   to the rotation buffer, operators uploaded ahead of use to the arena),
   hundreds per step, which as single copies cost a launch each.  Runs on
   the context's stream after the work queued there and returns when the
   copies are done. */
typedef struct {
  const double* src;
  double* dst;
  int64_t n;
} bsg_copy_seg_t;
int bsg_copy_segments(bsg_context* ctx, const bsg_copy_seg_t* segs_host, int32_t n_segs);

/* Composite-operator build in device memory (replaces the host's
 * TensorProduct / TensorProductElement for a block (x) dot operator,
 * operatorfunctions.C:146-254, through MatrixTensorProduct's scaled block
 * copies, MatrixBLAS.C:138-162, and the operator-vector sums of the
 * complementary operators, Operators.C:1059-1196,1472-1616,1905-1966):
 * group g adds into the destination rectangle dst[r*ld + c] (r < rows,
 * c < cols) the sum over its sources j in [first, first+count) of
 * (alpha_j * op(src_j)[r][c]) * beta_j, op(src)[r][c] = trans ? src[c*ld_j +
 * r] : src[r*ld_j + c], summed in list order.  Rectangles of different groups
 * must not overlap.  Device pointers; returns when the build is done. */
typedef struct {
  double* dst;
  int32_t ld, rows, cols, first, count, pad;
} bsg_build_group_t;
typedef struct {
  const double* src;
  int32_t ld, trans;
  double alpha, beta;
} bsg_build_src_t;
int bsg_build_apply(bsg_context* ctx, const bsg_build_group_t* groups_host, int32_t n_groups,
                    const bsg_build_src_t* srcs_host, int32_t n_srcs);
/* Subspace matrix of the block Davidson (linear.C:265-270): out_host[i*m+j]
 * = <X_i, Y_j> for m <= 128 device vectors of length n stored with stride ld
 * (the reference's subspace holds max(20, nroots+20) vectors, input.C:885). */
int bsg_gram(bsg_context* ctx, const double* X, const double* Y, int32_t m, int64_t n, int64_t ld, double* out_host);
/* Ritz rotation (linear.C:279-294): Y_j = sum_i coef_host[i*mo+j] X_i for
 * j < mo, m and mo <= 128; X, Y device vector sets (strides ld, ldy; must not
 * overlap).  The coefficients are copied to page-locked staging at once, so
 * coef_host may be reused on return; the kernels finish in stream order. */
int bsg_lincomb(bsg_context* ctx, const double* X, int32_t m, int64_t n, int64_t ld, const double* coef_host,
                int32_t mo, double* Y, int64_t ldy);
/* x[i] /= (e - diag[i] + shift) where |e - diag[i]| > 1e-12 (linear.C:27-47) */
int bsg_precondition(bsg_context* ctx, double* x, const double* diag, double e, double shift, int64_t n);
/* Olsen step (linear.C:50-60): r -= (<C~,r>/<C,C~>) C with C~ = precond(C),
 * then r = precond(r).  work: device scratch of n doubles.  The two dot
 * products stay in device memory (no host round trip); returns without
 * waiting for the kernels. */
int bsg_olsen_precondition(bsg_context* ctx, double* r, const double* c, const double* diag, double e, double shift,
                           double* work, int64_t n);
/* Gram-Schmidt of the new Davidson vector (linear.C:358-366 and the
 * Normalise that follows): for j < m, r /= |r|, r -= <r, b_j> b_j; then, if
 * normalise != 0, r is normalised unless |r|^2 <= 1e-15 (BaseOperator.C:
 * 282-287; a caller that projects r further before that Normalise passes 0).
 * B: m device vectors at stride ld.  The scalars stay in device memory, so
 * the loop runs without a host round trip per step; same reductions and
 * order as bsg_dot / bsg_scal / bsg_axpy. */
int bsg_gs_orthonormalise(bsg_context* ctx, double* r, const double* B, int32_t m, int64_t ld, int64_t n,
                          int32_t normalise);

#ifdef __cplusplus
}
#endif
#endif
