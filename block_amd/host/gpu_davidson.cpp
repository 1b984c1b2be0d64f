// Linear::block_davidson (linear.C:179-385) on device vectors.  Same control
// flow as blk::block_davidson (host restatement in davidson.cpp): Gram-Schmidt
// of the guesses, sigma for new vectors only, subspace diagonalisation,
// Ritz rotation, residual of the first unconverged root, Olsen
// preconditioner, convergence on the squared residual norm, deflation and
// re-orthogonalised expansion, plus the precision ladder.  Every vector of
// length N (b, sigma, residual, diagonal) stays in HBM and is handled by the
// bsg_* vector kernels; only the m x m subspace matrix (m <= 128) comes to the
// host for the eigen-decomposition.
#include <chrono>
#include <cmath>

#include "davidson.h"
#include "gpu_sigma.h"

namespace blk {

namespace {
struct DevBuf {
  bsg_context* ctx;
  double* p = nullptr;
  DevBuf(bsg_context* c, size_t n) : ctx(c) { bsg_check(bsg_malloc(ctx, sizeof(double) * std::max<size_t>(1, n), (void**)&p), "davidson buffer"); }
  ~DevBuf() {
    if (p) bsg_mfree(ctx, p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};
}  // namespace

void block_davidson_device(GpuSigma& h, std::vector<std::vector<double>>& b, std::vector<double>& h_diag,
                           double normtol, bool useprecond, int deflation_min, int deflation_max, DavidsonStats* st,
                           int max_iter, double ladder, const std::vector<std::vector<double>>* lower,
                           double min_split_share) {
  bsg_context* ctx = h.ctx();
  const int64_t N = h.wave_size();
  const int nroots = (int)b.size();
  // deflation keeps at least every root's vector
  deflation_min = std::max(deflation_min, nroots);
  deflation_max = std::max(deflation_max, deflation_min + 1);
  const int cap = deflation_max + 1;
  if (cap > kDeviceDavidsonMaxVectors)
    throw std::runtime_error("block_davidson_device: subspace larger than " +
                             std::to_string(kDeviceDavidsonMaxVectors) + " vectors");
  const int nlow = lower ? (int)lower->size() : 0;
  const double levelshift = 0.0;
  DevBuf B(ctx, (size_t)cap * N), S(ctx, (size_t)cap * N), T(ctx, (size_t)cap * N), R(ctx, N), Wk(ctx, N),
      D(ctx, N);
  double* Bp = B.p;
  double* Sp = S.p;
  double* Tp = T.p;
  auto vec = [&](double* base, int i) { return base + (size_t)i * N; };
  auto dot = [&](const double* x, const double* y) {
    double r = 0.0;
    bsg_check(bsg_dot(ctx, x, y, N, &r), "dot");
    return r;
  };
  auto normalise = [&](double* x) {  // skipped below NUMERICAL_ZERO (BaseOperator.C:282-287)
    const double n2 = dot(x, x);
    if (n2 > 1e-15) bsg_check(bsg_scal(ctx, 1.0 / std::sqrt(n2), x, N), "scal");
  };
  for (int i = 0; i < nroots; ++i)
    bsg_check(bsg_memcpy_h2d(ctx, vec(Bp, i), b[i].data(), sizeof(double) * N), "guess upload");
  bsg_check(bsg_memcpy_h2d(ctx, D.p, h_diag.data(), sizeof(double) * N), "diagonal upload");
  // lower states (linear.C:203-208,311-316,369-375): x -= <x,l>/<l,l> l in turn
  DevBuf Lw(ctx, (size_t)std::max(1, nlow) * N);
  std::vector<double> lnorm2(nlow);
  for (int i = 0; i < nlow; ++i) {
    if ((int64_t)(*lower)[i].size() != N) throw std::runtime_error("block_davidson_device: lower state of wrong size");
    bsg_check(bsg_memcpy_h2d(ctx, vec(Lw.p, i), (*lower)[i].data(), sizeof(double) * N), "lower state upload");
    lnorm2[i] = dot(vec(Lw.p, i), vec(Lw.p, i));
  }
  auto project_lower = [&](double* x) {
    for (int i = 0; i < nlow; ++i)
      bsg_check(bsg_axpy(ctx, -dot(x, vec(Lw.p, i)) / lnorm2[i], vec(Lw.p, i), x, N), "lower projection");
  };
  // orthonormalise the guesses (linear.C:190-200)
  for (int i = 0; i < nroots; ++i) {
    for (int j = 0; j < i; ++j) bsg_check(bsg_axpy(ctx, -dot(vec(Bp, j), vec(Bp, i)), vec(Bp, j), vec(Bp, i), N), "axpy");
    normalise(vec(Bp, i));
  }
  if (nlow > 0) {
    project_lower(vec(Bp, 0));
    normalise(vec(Bp, 0));
  }
  // the ladder's split phase only where split mode moves most of the work
  // onto BF16 slices: elsewhere the split sigma is mostly the FP64 small-block
  // kernel, and the restart at the switch would cost more than it saves
  bool split = ladder > 0.0 && h.split_share() >= min_split_share;
  h.set_precision(split ? BSG_SPLIT_BF16 : BSG_FP64);
  // the first iteration's (and a ladder restart's) nroots sigmas in one launch set
  h.set_batch(nroots);
  int m = nroots, nsig = 0, converged_roots = 0, iter = 0;
  std::vector<double> G, sub, eval, coef;
  while (true) {
    ++iter;
    if (st) st->iterations = iter;
    auto t0 = std::chrono::steady_clock::now();
    // sigma for the new vectors only (linear.C:234-258), batched
    if (m > nsig) {
      bsg_check(bsg_memset_zero(ctx, vec(Sp, nsig), sizeof(double) * N * (m - nsig)), "sigma clear");
      h.apply_device(vec(Bp, nsig), vec(Sp, nsig), m - nsig);
      if (st) {
        st->sigma_calls += m - nsig;
        if (split) st->sigma_calls_split += m - nsig;
      }
    }
    bsg_check(bsg_synchronize(ctx), "sigma");
    if (st) st->sigma_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    nsig = m;
    // subspace matrix h(i, j) = <b_i, sigma_j>, j <= i (linear.C:265-270)
    G.assign((size_t)m * m, 0.0);
    bsg_check(bsg_gram(ctx, Bp, Sp, m, N, N, G.data()), "gram");
    sub.assign((size_t)m * m, 0.0);
    eval.assign(m, 0.0);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j <= i; ++j) sub[(size_t)i * m + j] = sub[(size_t)j * m + i] = G[(size_t)i * m + j];
    syev(m, sub.data(), eval.data());  // eigenvectors in columns
    // Ritz vectors b_k = sum_i alpha(i, k) b_i, same for sigma (linear.C:279-294)
    bsg_check(bsg_lincomb(ctx, Bp, m, N, N, sub.data(), m, Tp, N), "ritz b");
    std::swap(Bp, Tp);
    bsg_check(bsg_lincomb(ctx, Sp, m, N, N, sub.data(), m, Tp, N), "ritz sigma");
    std::swap(Sp, Tp);
    // residual of the first unconverged root (linear.C:298-317)
    for (int i = 0; i < converged_roots; ++i) {
      bsg_check(bsg_copy(ctx, vec(Sp, i), R.p, N), "copy");
      bsg_check(bsg_axpy(ctx, -eval[i], vec(Bp, i), R.p, N), "axpy");
      if (dot(R.p, R.p) > normtol) {
        converged_roots = i;
        continue;
      }
    }
    const int k = converged_roots;
    bsg_check(bsg_copy(ctx, vec(Sp, k), R.p, N), "copy");
    bsg_check(bsg_axpy(ctx, -eval[k], vec(Bp, k), R.p, N), "axpy");
    project_lower(R.p);
    const double rnorm = dot(R.p, R.p);
    if (split && (rnorm < ladder * ladder || rnorm < normtol)) {
      // ladder step: continue from the current Ritz vectors with FP64 sigmas
      h.set_precision(BSG_FP64);
      split = false;
      if (st) st->ladder_switches++;
      m = nroots;
      nsig = 0;
      converged_roots = 0;
      continue;
    }
    if (useprecond)
      bsg_check(bsg_olsen_precondition(ctx, R.p, vec(Bp, k), D.p, eval[k], levelshift, Wk.p, N), "olsen");
    if (rnorm < normtol || iter >= max_iter) {
      ++converged_roots;
      if (converged_roots == nroots || iter >= max_iter) {
        for (int i = 0; i < std::min(m, (int)h_diag.size()); ++i) h_diag[i] = eval[i];
        for (int i = 0; i < nroots; ++i)
          bsg_check(bsg_memcpy_d2h(ctx, b[i].data(), vec(Bp, i), sizeof(double) * N), "ritz download");
        break;
      }
    } else {
      if (m >= deflation_max) {  // deflation (linear.C:352-357)
        m = deflation_min;
        nsig = std::min(nsig, m);
      }
      // linear.C:358-366 and the Normalise after it, scalars on the device
      if (nlow > 0) {
        // the lower-state projection sits between the loop and the Normalise
        bsg_check(bsg_gs_orthonormalise(ctx, R.p, Bp, m, N, N, 0), "orthogonalise");
        project_lower(R.p);
        normalise(R.p);
      } else {
        bsg_check(bsg_gs_orthonormalise(ctx, R.p, Bp, m, N, N, 1), "orthonormalise");
      }
      bsg_check(bsg_copy(ctx, R.p, vec(Bp, m), N), "copy");
      ++m;
    }
  }
  bsg_check(bsg_synchronize(ctx), "davidson");
}

}  // namespace blk
