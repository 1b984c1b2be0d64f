#include "guess.h"

#include <cmath>

namespace blk {

namespace {
bool same_flat(const StateInfo& a, const StateInfo& b) { return a.q == b.q && a.dim == b.dim; }
int parity_sign(int x) { return (x & 1) ? -1 : 1; }
}  // namespace

bool same_sectors(const StateInfo& a, const StateInfo& b) {
  if (!same_flat(a, b)) return false;
  if (!a.left != !b.left || !a.right != !b.right) return false;
  if (a.left && !same_flat(*a.left, *b.left)) return false;
  if (a.right && !same_flat(*a.right, *b.right)) return false;
  if (a.subs.size() != b.subs.size()) return false;
  for (size_t s = 0; s < a.subs.size(); ++s) {
    if (a.subs[s].size() != b.subs[s].size()) return false;
    for (size_t i = 0; i < a.subs[s].size(); ++i) {
      const auto &x = a.subs[s][i], &y = b.subs[s][i];
      if (x.l != y.l || x.r != y.r || x.off != y.off) return false;
    }
  }
  return true;
}

// GuessWave::transform_previous_wavefunction, two-dot branch
// (guess_wavefunction.C:578-611): TransformLeftBlock (:17-31),
// onedot_shufflesysdot (:434-485 with the 6j of :200-255), TransformRightBlock
// (:33-51).
bool guess_transform(const WaveLayout& w0, const std::vector<std::vector<double>>& c0, const RotRecord& UL,
                     const RotRecord& UE, const WaveLayout& w1, std::vector<std::vector<double>>& out) {
  const StateInfo &L0 = *w0.L, &R0 = *w0.R, &L1 = *w1.L, &R1 = *w1.R;
  if (w0.target != w1.target) return false;
  if (!L0.left || !R0.left || !L1.left || !R1.left) return false;
  if (!same_sectors(L0, *UL.composite) || !same_flat(*L1.left, *UL.kept)) return false;
  if (!same_flat(*R0.left, *UE.kept) || !same_sectors(R1, *UE.composite)) return false;
  if (!same_flat(*R0.right, *L1.right)) return false;  // the dot that moves from R to L
  const StateInfo &S1 = *L1.left, &D = *R0.right, &E = *R0.left;
  for (int b = 0; b < D.nsec(); ++b)
    if (D.dim[b] != 1) return false;  // single-orbital dot: one state per sector
  std::vector<int> inv(L0.nsec(), -1);
  for (int a = 0; a < (int)UL.oldidx.size(); ++a) inv[UL.oldidx[a]] = a;
  const int J = w0.target.s2;
  std::vector<std::vector<double>> res(c0.size(), std::vector<double>(w1.size, 0.0));
  std::vector<double> T1;
  for (size_t root = 0; root < c0.size(); ++root) {
    if ((long)c0[root].size() != w0.size) return false;
    double* o = res[root].data();
    for (const auto& bk : w0.blocks) {
      const int a = inv[bk.l];
      if (a < 0) continue;
      const int ma = S1.dim[a], d = bk.rows, cols = bk.cols;
      // left rotation: T1 = UL[l]^T C(l, r)   (ma x cols)
      T1.assign((size_t)ma * cols, 0.0);
      dgemm(true, false, ma, cols, d, 1.0, UL.rot[bk.l].data(), ma, c0[root].data() + bk.off, cols, 0.0, T1.data(),
            cols);
      const int A = S1.q[a].s2, CB = R0.q[bk.r].s2;
      for (const auto& sub : R0.subs[bk.r]) {
        const int c = sub.l, b = sub.r;
        const int C = E.q[c].s2, B = D.q[b].s2;
        // (c b)_CB -> (b c)_CB: fermion sign of b passing c, CG symmetry
        // (getCommuteParity, BaseOperator.C:20-53)
        const double p1 = (E.q[c].n & 1) && (D.q[b].n & 1) ? -1.0 : 1.0;
        const double swap = p1 * parity_sign((C + B - CB) / 2);
        const int QE = UE.oldidx[c], dQE = R1.dim[QE], dc = E.dim[c];
        const double* rotE = UE.rot[QE].data();  // dQE x dc
        for (const auto& xs : L1.prod_to_sector[(size_t)a * D.nsec() + b]) {
          const int x = xs.first, xoff = xs.second;
          const int blk = w1.find(x, QE);
          if (blk < 0) continue;
          const int AB = L1.q[x].s2;
          double coef = sixj(A, B, AB, C, J, CB);
          if (coef == 0.0) continue;
          coef *= std::sqrt((AB + 1.0) * (CB + 1.0)) * parity_sign((A + B + J + C) / 2) * swap;
          // right expansion: rows xoff.. of block (x, QE) += coef T1(:, c part) UE[QE]^T
          dgemm(false, true, ma, dQE, dc, coef, T1.data() + sub.off, cols, rotE, dc, 1.0,
                o + w1.blocks[blk].off + (size_t)xoff * dQE, dQE);
        }
      }
    }
  }
  out.swap(res);
  return true;
}

// transpose_previous_wavefunction (guess_wavefunction.C:55-139) for a turn
// whose new halves are the old ones swapped.
bool guess_transpose(const WaveLayout& w0, const std::vector<std::vector<double>>& c0, const WaveLayout& w1,
                     std::vector<std::vector<double>>& out) {
  if (w0.target != w1.target) return false;
  if (!same_sectors(*w1.L, *w0.R) || !same_sectors(*w1.R, *w0.L)) return false;
  const int J = w0.target.s2;
  std::vector<std::vector<double>> res(c0.size(), std::vector<double>(w1.size, 0.0));
  for (size_t root = 0; root < c0.size(); ++root) {
    if ((long)c0[root].size() != w0.size) return false;
    for (const auto& bk : w1.blocks) {
      const int ob = w0.find(bk.r, bk.l);
      if (ob < 0) continue;
      const Quantum &qx = w1.L->q[bk.l], &qy = w1.R->q[bk.r];
      const double sign = ((qx.n & 1) && (qy.n & 1) ? -1.0 : 1.0) * parity_sign((qx.s2 + qy.s2 - J) / 2);
      const double* src = c0[root].data() + w0.blocks[ob].off;  // bk.cols x bk.rows
      double* dst = res[root].data() + bk.off;
      for (int i = 0; i < bk.rows; ++i)
        for (int j = 0; j < bk.cols; ++j) dst[(size_t)i * bk.cols + j] = sign * src[(size_t)j * bk.rows + i];
    }
  }
  out.swap(res);
  return true;
}

}  // namespace blk
