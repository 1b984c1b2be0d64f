// Direct CPU restatement of multiply_h -> SpinBlock::multiplyH -> opxop ->
// TensorMultiply.  TEST INFRASTRUCTURE ONLY (see oracle.h).
#include <cstring>
#include <omp.h>

#include "oracle.h"

namespace oracle {
using namespace blk;

namespace {
const double SQ2 = std::sqrt(2.0);

// element (i,j) of view block (r,c); stored block id sb
inline double vel(const OpView& o, int sbid, int i, int j, int rows, int cols) {
  const double* d = o.op->blk(sbid);
  return o.dag ? d[(size_t)j * rows + i] : d[(size_t)i * cols + j];
}

// C[m x n] += alpha * A[m x k] * B[k x n], row-major with leading dims
void mm_acc(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb, double* C,
            int ldc) {
  for (int i = 0; i < m; ++i) {
    double* c = C + (size_t)i * ldc;
    for (int p = 0; p < k; ++p) {
      double a = alpha * A[(size_t)i * lda + p];
      if (a == 0.0) continue;
      const double* b = B + (size_t)p * ldb;
      for (int j = 0; j < n; ++j) c[j] += a * b[j];
    }
  }
}

// the sector products: the oracle's own loops, or (fast, the timed CPU
// baseline only) the host library's blocked DGEMM
bool g_fast = false;
inline void gemm_acc(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb, double* C,
                     int ldc) {
  if (g_fast)
    dgemm(false, false, m, n, k, alpha, A, lda, B, ldb, 1.0, C, ldc);
  else
    mm_acc(m, n, k, alpha, A, lda, B, ldb, C, ldc);
}

// v += scale * [a(L) (x) b(R)]^0 c   -- operatorfunctions.C:485-537
void tensor_multiply(const SuperBlock& sb, const OpView& a, const OpView& b, double scale, const double* c,
                     double* v) {
  const WaveLayout& W = *sb.wave;
  const StateInfo &ls = *W.L, &rs = *W.R;
  const int St = W.target.s2;
  std::vector<double> abuf, m, bt;
  for (int lQ = 0; lQ < ls.nsec(); ++lQ)
    for (int rQp = 0; rQp < rs.nsec(); ++rQp)
      for (int lQp = 0; lQp < ls.nsec(); ++lQp) {
        int ia = a.find(lQ, lQp);
        int ic = W.find(lQp, rQp);
        if (ia < 0 || ic < 0) continue;
        const int dl = ls.dim[lQ], dlp = ls.dim[lQp], drp = rs.dim[rQp];
        // m = a(lQ,lQ') * c(lQ',rQ')  (beta = 0, operatorfunctions.C:512-516)
        abuf.resize((size_t)dl * dlp);
        for (int i = 0; i < dl; ++i)
          for (int j = 0; j < dlp; ++j) abuf[(size_t)i * dlp + j] = vel(a, ia, i, j, dl, dlp);
        m.assign((size_t)dl * drp, 0.0);
        gemm_acc(dl, drp, dlp, a.scale(lQ, lQp), abuf.data(), dlp, c + W.blocks[ic].off, drp, m.data(), drp);
        for (int rQ = 0; rQ < rs.nsec(); ++rQ) {
          int iv = W.find(lQ, rQ);
          int ib = b.find(rQ, rQp);
          if (iv < 0 || ib < 0) continue;
          const int dr = rs.dim[rQ];
          double f = scale * wninej(ls.q[lQp].s2, rs.q[rQp].s2, St, a.k2(), b.k2(), 0, ls.q[lQ].s2, rs.q[rQ].s2, St);
          if (b.fermion() && (ls.q[lQp].n & 1)) f = -f;
          f *= b.scale(rQ, rQp);
          if (f == 0.0) continue;
          // v(lQ,rQ) += f * m * b(rQ,rQ')^T
          bt.resize((size_t)drp * dr);
          for (int j = 0; j < drp; ++j)
            for (int r = 0; r < dr; ++r) bt[(size_t)j * dr + r] = vel(b, ib, r, j, dr, drp);
          gemm_acc(dl, dr, drp, f, m.data(), drp, bt.data(), dr, v + W.blocks[iv].off, dr);
        }
      }
}

struct Pair {
  OpView a, b;
  double coef;
};

// the term set of SpinBlock::multiplyH, enumerated by operator arrays
std::vector<Pair> enumerate(const SuperBlock& sb) {
  const SpinBlock& L = *sb.L;  // loop block
  const SpinBlock& R = *sb.R;  // other block
  std::vector<Pair> P;
  if (std::fabs(sb.ecore) > 1e-20) P.push_back({V(*L.ident), V(*R.ident), sb.ecore});
  P.push_back({V(*L.ham), V(*R.ident), 1.0});
  P.push_back({V(*L.ident), V(*R.ham), 1.0});
  // opxop::cxcddcomp with otherblock = left, loop over CRE of the right block
  for (int q = 0; q < R.norb; ++q) {
    if (!R.cre[q] || !L.ccd[q]) continue;
    P.push_back({V(*L.ccd[q]), Vd(*R.cre[q]), -SQ2});
    P.push_back({Vd(*L.ccd[q]), V(*R.cre[q]), -SQ2});
  }
  // opxop::cxcddcomp with otherblock = right, loop over CRE of the left block
  for (int q = 0; q < L.norb; ++q) {
    if (!L.cre[q] || !R.ccd[q]) continue;
    P.push_back({Vd(*L.cre[q]), V(*R.ccd[q]), -SQ2});
    P.push_back({V(*L.cre[q]), Vd(*R.ccd[q]), -SQ2});
  }
  // opxop::cdxcdcomp and ddxcccomp over the loop block's pair operators
  const int n = L.norb;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j)
      for (int K : {0, 2}) {
        size_t id = L.pidx(i, j, K);
        if (L.cd[id] && R.cdc[id]) {
          P.push_back({V(*L.cd[id]), V(*R.cdc[id]), 1.0});
          if (i != j) P.push_back({Vd(*L.cd[id]), Vd(*R.cdc[id]), 1.0});
        }
        if (L.cc[id] && R.ddc[id]) {
          double f = (i == j) ? 1.0 : 2.0;
          P.push_back({V(*L.cc[id]), V(*R.ddc[id]), f});
          P.push_back({Vd(*L.cc[id]), Vd(*R.ddc[id]), f});
        }
      }
  return P;
}
}  // namespace

void oracle_set_blocked_gemm(bool on) { g_fast = on; }

void oracle_multiply_h(const SuperBlock& sb, const double* c, double* v, int nthreads) {
  auto P = enumerate(sb);
  const long N = sb.wave->size;
  if (nthreads <= 1) {
    for (const auto& p : P) tensor_multiply(sb, p.a, p.b, p.coef, c, v);
    return;
  }
  // per-thread sigma copies summed at the end (distribute.h:128-224)
  std::vector<std::vector<double>> part(nthreads, std::vector<double>(N, 0.0));
#pragma omp parallel num_threads(nthreads)
  {
    int t = omp_get_thread_num();
#pragma omp for schedule(static, 1)  // fixed pair -> thread map: reproducible sums
    for (long i = 0; i < (long)P.size(); ++i) tensor_multiply(sb, P[i].a, P[i].b, P[i].coef, c, part[t].data());
  }
  for (int t = 0; t < nthreads; ++t)
    for (long i = 0; i < N; ++i) v[i] += part[t][i];
}

double oracle_sigma_flops(const SuperBlock& sb) {
  // every (term, l, l', r, r') quadruple checked against the operators'
  // allowed blocks and the wavefunction's: the definition, by brute force
  const WaveLayout& W = *sb.wave;
  const StateInfo &ls = *W.L, &rs = *W.R;
  double fl = 0.0;
  for (const auto& t : sb.terms) {
    const bool aid = t.a.op == sb.L->ident.get(), bid = t.b.op == sb.R->ident.get();
    if (aid && bid) {
      fl += 2.0 * W.size;
      continue;
    }
    for (const auto& cb : W.blocks)
      for (int l = 0; l < ls.nsec(); ++l) {
        if (t.a.find(l, cb.l) < 0) continue;
        if (!aid) fl += 2.0 * ls.dim[l] * ls.dim[cb.l] * rs.dim[cb.r];
        if (bid) continue;
        for (int r = 0; r < rs.nsec(); ++r)
          if (W.find(l, r) >= 0 && t.b.find(r, cb.r) >= 0) fl += 2.0 * ls.dim[l] * rs.dim[cb.r] * rs.dim[r];
      }
  }
  return fl;
}

std::vector<double> oracle_diagonal(const SuperBlock& sb) {
  auto P = enumerate(sb);
  const WaveLayout& W = *sb.wave;
  const StateInfo &ls = *W.L, &rs = *W.R;
  const int St = W.target.s2;
  std::vector<double> e(W.size, 0.0);
  for (const auto& p : P) {
    for (const auto& bk : W.blocks) {
      int ia = p.a.find(bk.l, bk.l), ib = p.b.find(bk.r, bk.r);
      if (ia < 0 || ib < 0) continue;
      double f = p.coef * wninej(ls.q[bk.l].s2, rs.q[bk.r].s2, St, p.a.k2(), p.b.k2(), 0, ls.q[bk.l].s2,
                                 rs.q[bk.r].s2, St);
      if (p.b.fermion() && (ls.q[bk.l].n & 1)) f = -f;
      f *= p.a.scale(bk.l, bk.l) * p.b.scale(bk.r, bk.r);
      for (int i = 0; i < bk.rows; ++i)
        for (int r = 0; r < bk.cols; ++r)
          e[bk.off + (size_t)i * bk.cols + r] +=
              f * vel(p.a, ia, i, i, bk.rows, bk.rows) * vel(p.b, ib, r, r, bk.cols, bk.cols);
    }
  }
  return e;
}

}  // namespace oracle
