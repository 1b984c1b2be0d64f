#include "davidson.h"

#include <chrono>

namespace blk {

double ddot(const std::vector<double>& x, const std::vector<double>& y) {
  double s = 0.0;
  const size_t n = x.size();
#pragma omp parallel for reduction(+ : s) if (n > 65536)
  for (size_t i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}
void daxpy(double a, const std::vector<double>& x, std::vector<double>& y) {
  const size_t n = x.size();
#pragma omp parallel for if (n > 65536)
  for (size_t i = 0; i < n; ++i) y[i] += a * x[i];
}
void dscal(double a, std::vector<double>& x) {
  const size_t n = x.size();
#pragma omp parallel for if (n > 65536)
  for (size_t i = 0; i < n; ++i) x[i] *= a;
}
// Normalise skips vectors whose squared norm is below NUMERICAL_ZERO
// (BaseOperator.C:282-287)
void normalise(std::vector<double>& x) {
  double n2 = ddot(x, x);
  if (n2 > 1e-15) dscal(1.0 / std::sqrt(n2), x);
}
void precondition(std::vector<double>& x, double e, const std::vector<double>& diag, double shift) {
  for (size_t i = 0; i < x.size(); ++i)
    if (std::fabs(e - diag[i]) > 1e-12) x[i] /= (e - diag[i] + shift);
}
void olsen_precondition(std::vector<double>& r, const std::vector<double>& c0, double e,
                        const std::vector<double>& diag, double shift) {
  std::vector<double> c0p = c0;
  precondition(c0p, e, diag, shift);
  double num = ddot(c0p, r), den = ddot(c0, c0p);
  daxpy(-num / den, c0, r);
  precondition(r, e, diag, shift);
}

void block_davidson(std::vector<std::vector<double>>& b, std::vector<double>& h_diag, double normtol,
                    DavidsonFunctor& h, bool useprecond, int deflation_min, int deflation_max, DavidsonStats* st,
                    int max_iter, double ladder, const std::vector<std::vector<double>>* lower) {
  const double levelshift = 0.0;
  // x -= <x,l>/<l,l> l for every lower state in turn
  auto project_lower = [&](std::vector<double>& x) {
    if (!lower) return;
    for (const auto& l : *lower) daxpy(-ddot(x, l) / ddot(l, l), l, x);
  };
  bool split = ladder > 0.0 && h.set_precision(1);
  const int nroots = (int)b.size();
  // deflation keeps at least every root's vector
  deflation_min = std::max(deflation_min, nroots);
  deflation_max = std::max(deflation_max, deflation_min + 1);
  // orthonormalise the guesses (linear.C:190-200)
  for (int i = 0; i < nroots; ++i) {
    for (int j = 0; j < i; ++j) daxpy(-ddot(b[j], b[i]), b[j], b[i]);
    normalise(b[i]);
  }
  if (lower && !lower->empty()) {  // linear.C:203-208
    project_lower(b[0]);
    normalise(b[0]);
  }
  std::vector<std::vector<double>> sigma;
  int converged_roots = 0;
  int iter = 0;
  const std::vector<double> diag0 = h_diag;
  while (true) {
    ++iter;
    if (st) st->iterations = iter;
    auto t0 = std::chrono::steady_clock::now();
    for (size_t i = sigma.size(); i < b.size(); ++i) {
      sigma.emplace_back(b[i].size(), 0.0);
      h(b[i].data(), sigma[i].data());
      if (st) {
        st->sigma_calls++;
        if (split) st->sigma_calls_split++;
      }
    }
    if (st) st->sigma_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    const int m = (int)b.size();
    std::vector<double> sub((size_t)m * m), eval(m);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j <= i; ++j) sub[(size_t)i * m + j] = sub[(size_t)j * m + i] = ddot(b[i], sigma[j]);
    syev(m, sub.data(), eval.data());  // alpha(i,j) = sub[i*m+j], eigenvectors in columns
    // Ritz vectors (linear.C:279-294)
    {
      std::vector<std::vector<double>> btmp = b, stmp = sigma;
      for (int i = 0; i < m; ++i) {
        dscal(sub[(size_t)i * m + i], b[i]);
        dscal(sub[(size_t)i * m + i], sigma[i]);
      }
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
          if (i != j) {
            daxpy(sub[(size_t)i * m + j], btmp[i], b[j]);
            daxpy(sub[(size_t)i * m + j], stmp[i], sigma[j]);
          }
    }
    // residual of the first unconverged root (linear.C:298-317)
    std::vector<double> r;
    for (int i = 0; i < converged_roots; ++i) {
      r = sigma[i];
      daxpy(-eval[i], b[i], r);
      if (ddot(r, r) > normtol) {
        converged_roots = i;
        continue;
      }
    }
    r = sigma[converged_roots];
    daxpy(-eval[converged_roots], b[converged_roots], r);
    project_lower(r);  // linear.C:311-316
    double rnorm = ddot(r, r);
    if (split && (rnorm < ladder * ladder || rnorm < normtol)) {
      // ladder step: continue from the current Ritz vectors with FP64 sigmas
      h.set_precision(0);
      split = false;
      if (st) st->ladder_switches++;
      b.resize(nroots);
      sigma.clear();
      converged_roots = 0;
      continue;
    }
    if (useprecond) olsen_precondition(r, b[converged_roots], eval[converged_roots], diag0, levelshift);
    if (rnorm < normtol || iter >= max_iter) {
      ++converged_roots;
      if (converged_roots == nroots || iter >= max_iter) {
        for (int i = 0; i < std::min((int)b.size(), (int)h_diag.size()); ++i) h_diag[i] = eval[i];
        b.resize(nroots);
        break;
      }
    } else {
      if ((int)b.size() >= deflation_max) {  // deflation (linear.C:352-357)
        b.resize(deflation_min);
        sigma.resize(deflation_min);
      }
      for (size_t j = 0; j < b.size(); ++j) {  // linear.C:358-366
        double nrm = ddot(r, r);
        dscal(1.0 / std::sqrt(nrm), r);
        daxpy(-ddot(r, b[j]), b[j], r);
      }
      project_lower(r);  // linear.C:369-375
      normalise(r);
      b.push_back(r);
    }
  }
}

}  // namespace blk
