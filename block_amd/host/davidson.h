// Linear::block_davidson restated (linear.C:179-385) behind the same functor
// boundary as the reference: Davidson_functor / multiply_h (davidson.h:16-29).
#pragma once
#include <vector>

#include "superblock.h"

namespace blk {

// v = H c for one vector (v is overwritten; the reference clears sigma and
// multiply_h accumulates into it, linear.C:239-253).
struct DavidsonFunctor {
  virtual ~DavidsonFunctor() = default;
  virtual void operator()(const double* c, double* v) = 0;
  // Arithmetic of the following sigmas: 0 = FP64, 1 = split BF16.  Returns
  // false when the functor has no such mode (the default: FP64 only).
  virtual bool set_precision(int mode) { return mode == 0; }
  // diag(H) in wavefunction order into d, when the functor computes it
  // itself (SpinBlock::diagonalH); false: the caller computes it
  virtual bool diagonal(std::vector<double>& d) { return false; }
  // runs the whole Davidson solve itself (device-resident vectors) with the
  // arguments of block_davidson; false: the caller runs block_davidson
  virtual bool davidson(std::vector<std::vector<double>>& b, std::vector<double>& h_diag, double normtol,
                        bool useprecond, int deflation_min, int deflation_max, struct DavidsonStats* st, int max_iter,
                        double ladder, const std::vector<std::vector<double>>* lower) {
    return false;
  }
};

struct DavidsonStats {
  int iterations = 0;
  int sigma_calls = 0;
  int sigma_calls_split = 0;  // of which in split precision
  int ladder_switches = 0;
  double sigma_seconds = 0.0;  // the reference's hmultiply timer (linear.C:220,259)
  double setup_seconds = 0.0;  // sigma functor construction (work list, operator upload)
  double diag_seconds = 0.0;   // diagonal of H (preconditioner)
};

// Restatement of block_davidson with its control flow: Gram-Schmidt of the
// guesses, sigma for new vectors only, subspace diagonalisation, Ritz
// rotation, residual of the first unconverged root, Olsen preconditioner,
// convergence on the squared residual norm, deflation and re-orthogonalised
// expansion.  On return b holds the Ritz vectors and h_diag[0..nroots) the
// eigenvalues.
//
// lower (state-specific excited states, the reference's lowerStates): when
// given and not empty, the first guess, every residual and every new
// subspace vector are projected off these vectors (x -= <x,l>/<l,l> l for
// each in turn, linear.C:203-208,311-316,369-375), so the solve converges to
// the lowest states orthogonal to them.
//
// Precision ladder (ladder > 0 and a functor with split precision): sigmas
// start in split precision; once the residual norm of the current root falls
// below `ladder`, or it meets normtol while still in split precision, the
// solver keeps the current Ritz vectors, drops the rest of the subspace and
// continues with FP64 sigmas, so convergence is always decided on FP64
// sigmas.  Subspace algebra, orthogonalisation and residuals are FP64
// throughout.
void block_davidson(std::vector<std::vector<double>>& b, std::vector<double>& h_diag, double normtol,
                    DavidsonFunctor& h, bool useprecond, int deflation_min, int deflation_max, DavidsonStats* st,
                    int max_iter = 1000, double ladder = 0.0,
                    const std::vector<std::vector<double>>* lower = nullptr);

// vector kernels of BaseOperator.C:240-288 / linear.C:27-60 on host vectors
double ddot(const std::vector<double>& x, const std::vector<double>& y);
void daxpy(double a, const std::vector<double>& x, std::vector<double>& y);
void dscal(double a, std::vector<double>& x);
void normalise(std::vector<double>& x);
void precondition(std::vector<double>& x, double e, const std::vector<double>& diag, double shift);
void olsen_precondition(std::vector<double>& r, const std::vector<double>& c0, double e,
                        const std::vector<double>& diag, double shift);

}  // namespace blk
