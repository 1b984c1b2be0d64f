// Fiedler-vector orbital ordering (fiedler_reorder, fiedler.C:47-81, on the
// interaction matrix genetic::ReadIntegral builds, genetic/ReadIntegral.C:
// 11-62): K(i, j) = |(ij|ij)| + 1e-7 |t_ij|, the graph Laplacian of K, and the
// orbitals sorted by their component of the eigenvector of its second
// smallest eigenvalue (ties in orbital order).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "spinblock.h"

namespace blk {

std::vector<int> fiedler_order_matrix(int n, const double* m) {
  std::vector<double> lap((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    double rs = 0.0;
    for (int j = 0; j < n; ++j) rs += std::fabs(m[(size_t)i * n + j]);
    for (int j = 0; j < n; ++j) lap[(size_t)i * n + j] = -std::fabs(m[(size_t)i * n + j]);
    lap[(size_t)i * n + i] += rs;
  }
  std::vector<double> w(n);
  syev(n, lap.data(), w.data());  // ascending eigenvalues, eigenvectors in columns
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  if (n < 2) return order;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return lap[(size_t)a * n + 1] < lap[(size_t)b * n + 1]; });
  return order;
}

std::vector<int> fiedler_order(const Integrals& I) {
  const int n = I.norb;
  // an FCIDUMP lists each distinct integral once; (ij|ij) is the entry the
  // reference's reader accumulates for the pair
  std::vector<double> K((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      K[(size_t)i * n + j] = std::fabs(I.v(i, j, i, j)) + 1e-7 * std::fabs(I.t(i, j));
  return fiedler_order_matrix(n, K.data());
}

}  // namespace blk
