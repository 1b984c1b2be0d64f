#include "superblock.h"

#include <algorithm>

namespace blk {

static const double SQ2 = std::sqrt(2.0);

SuperBlock make_superblock(const Integrals& I, BlockPtr L, BlockPtr R, const Quantum& target) {
  SuperBlock sb;
  sb.L = std::move(L);
  sb.R = std::move(R);
  sb.target = target;
  sb.ecore = I.ecore;
  sb.wave = make_wave_layout(sb.L->si, sb.R->si, target);
  build_sigma_terms(I, sb);
  return sb;
}

void build_sigma_terms(const Integrals& I, SuperBlock& sb) {
  const SpinBlock& L = *sb.L;
  const SpinBlock& R = *sb.R;
  auto& T = sb.terms;
  T.clear();
  auto push = [&](OpView a, OpView b, double c, int kind) { T.push_back(SigmaTerm{a, b, c, kind}); };
  // H_L (x) 1 and 1 (x) H_R (spinblock.C:742-747), E_core (:735-740)
  push(V(*L.ham), V(*R.ident), 1.0, 0);
  push(V(*L.ident), V(*R.ham), 1.0, 0);
  if (sb.ecore != 0.0) push(V(*L.ident), V(*R.ident), sb.ecore, 0);
  // CCDcomp (x) C in both directions (opxop::cxcddcomp, opxop.C:232-285)
  for (int q : R.sites) {
    const SparseOp* x = L.ccd[q].get();
    const SparseOp* c = R.cre[q].get();
    if (!x || !c) continue;
    push(V(*x), Vd(*c), -SQ2, 1);
    push(Vd(*x), V(*c), -SQ2, 1);
  }
  for (int q : L.sites) {
    const SparseOp* c = L.cre[q].get();
    const SparseOp* x = R.ccd[q].get();
    if (!x || !c) continue;
    push(Vd(*c), V(*x), -SQ2, 1);
    push(V(*c), Vd(*x), -SQ2, 1);
  }
  // two-index pairs of the loop block (opxop::cdxcdcomp / ddxcccomp, opxop.C:155-228)
  for (int i : L.sites)
    for (int j : L.sites) {
      if (j > i) continue;
      for (int K : {0, 2}) {
        const SparseOp* cd = L.get(L.cd, i, j, K);
        const SparseOp* q = R.get(R.cdc, i, j, K);
        if (cd && q) {
          push(V(*cd), V(*q), 1.0, 2);
          if (i != j) push(Vd(*cd), Vd(*q), 1.0, 2);
        }
        const SparseOp* cc = L.get(L.cc, i, j, K);
        const SparseOp* p = R.get(R.ddc, i, j, K);
        if (cc && p) {
          double f = (i == j) ? 1.0 : 2.0;
          push(V(*cc), V(*p), f, 3);
          push(Vd(*cc), Vd(*p), f, 3);
        }
      }
    }
}

std::vector<double> diagonal_h(const SuperBlock& sb) {
  const WaveLayout& W = *sb.wave;
  const StateInfo &ls = *W.L, &rs = *W.R;
  std::vector<double> e(W.size, 0.0);
  const int St = W.target.s2;
  for (const auto& t : sb.terms) {
    if (t.a.dn() != 0 || t.b.dn() != 0 || t.a.dirrep() != 0 || t.b.dirrep() != 0) continue;
    const int k = t.a.k2();
    for (const auto& blk : W.blocks) {
      int ia = t.a.find(blk.l, blk.l), ib = t.b.find(blk.r, blk.r);
      if (ia < 0 || ib < 0) continue;
      const Quantum &ql = ls.q[blk.l], &qr = rs.q[blk.r];
      double f = t.coef * wninej(ql.s2, qr.s2, St, k, k, 0, ql.s2, qr.s2, St);
      if (t.b.fermion() && (ql.n & 1)) f = -f;
      f *= t.a.scale(blk.l, blk.l) * t.b.scale(blk.r, blk.r);
      if (f == 0.0) continue;
      const double* A = t.a.op->blk(ia);
      const double* B = t.b.op->blk(ib);
      const int dl = blk.rows, dr = blk.cols;
      double* out = e.data() + blk.off;
      for (int il = 0; il < dl; ++il) {
        double av = f * A[(size_t)il * dl + il];
        for (int ir = 0; ir < dr; ++ir) out[(size_t)il * dr + ir] += av * B[(size_t)ir * dr + ir];
      }
    }
  }
  return e;
}

double sigma_flops(const SuperBlock& sb) {
  // walks the operators' allowed blocks instead of all sector pairs: the
  // work is proportional to the GEMM pairs, not to terms x sectors^2
  const WaveLayout& W = *sb.wave;
  const StateInfo &ls = *W.L, &rs = *W.R;
  double fl = 0.0;
  std::vector<std::vector<int>> acol(ls.nsec()), brow(rs.nsec());
  for (const auto& t : sb.terms) {
    bool aid = (t.a.op == sb.L->ident.get()), bid = (t.b.op == sb.R->ident.get());
    if (aid && bid) {
      fl += 2.0 * W.size;
      continue;
    }
    // view blocks (l, l') of a by l', (r, r') of b by r'
    for (auto& v : acol) v.clear();
    for (auto& v : brow) v.clear();
    for (const auto& bk : t.a.op->pat->blocks) {
      const int l = t.a.dag ? bk.c : bk.r, lp = t.a.dag ? bk.r : bk.c;
      acol[lp].push_back(l);
    }
    for (const auto& bk : t.b.op->pat->blocks) {
      const int r = t.b.dag ? bk.c : bk.r, rp = t.b.dag ? bk.r : bk.c;
      brow[rp].push_back(r);
    }
    for (const auto& cb : W.blocks) {  // c block (lQ', rQ')
      const int lp = cb.l, rp = cb.r;
      const double dlp = ls.dim[lp], drp = rs.dim[rp];
      for (int l : acol[lp]) {
        const double dl = ls.dim[l];
        if (!aid) fl += 2.0 * dl * dlp * drp;
        if (bid) continue;
        for (int r : brow[rp])
          if (W.find(l, r) >= 0) fl += 2.0 * dl * drp * rs.dim[r];
      }
    }
  }
  return fl;
}

}  // namespace blk
