// Determinant full CI, brute force: the exact answer every untruncated
// spin-adapted DMRG superblock must reproduce.  TEST INFRASTRUCTURE ONLY.
//
// Determinants |I_a I_b> (alpha creators first), MS = 0.  sigma uses the
// standard E_pq form  H = sum_pq k_pq E_pq + 1/2 sum_pqrs (pq|rs) E_pq E_rs
// with k_pq = t_pq - 1/2 sum_r (pr|rq); the spin penalty adds mu * S^2 with
// S^2 = S_- S_+ on MS = 0.  Lowest roots by a Davidson iteration with the
// diagonal preconditioner.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <unordered_map>

#include "oracle.h"

namespace oracle {
using namespace blk;

namespace {
struct Strings {
  int n, ne;
  std::vector<uint32_t> s;
  std::unordered_map<uint32_t, int> idx;
  Strings(int n_, int ne_) : n(n_), ne(ne_) {
    if (ne < 0 || ne > n) return;
    for (uint32_t x = 0; x < (1u << n); ++x)
      if (__builtin_popcount(x) == ne) {
        idx[x] = (int)s.size();
        s.push_back(x);
      }
  }
  int size() const { return (int)s.size(); }
};
inline int below(uint32_t x, int p) { return __builtin_popcount(x & ((1u << p) - 1)); }
// a+_p a_q on string x; returns false if zero, else y and sign
inline bool excite(uint32_t x, int p, int q, uint32_t& y, int& sign) {
  if (!(x & (1u << q))) return false;
  uint32_t t = x & ~(1u << q);
  int s = below(x, q);
  if (t & (1u << p)) return false;
  s += below(t, p);
  y = t | (1u << p);
  sign = (s & 1) ? -1 : 1;
  return true;
}
struct Exc {
  int J, pq, sign;
};
std::vector<std::vector<Exc>> exc_list(const Strings& S) {
  std::vector<std::vector<Exc>> L(S.size());
  for (int I = 0; I < S.size(); ++I)
    for (int p = 0; p < S.n; ++p)
      for (int q = 0; q < S.n; ++q) {
        uint32_t y;
        int sg;
        if (excite(S.s[I], p, q, y, sg)) L[I].push_back({S.idx.at(y), p * S.n + q, sg});
      }
  return L;
}

struct Fci {
  const Integrals& I;
  int n, na, nb;
  double mu;
  Strings A, B, Ap, Bm;  // alpha/beta strings and the S_+ intermediate space
  std::vector<std::vector<Exc>> ea, eb;
  std::vector<double> k1, vpq;  // k_pq, (pq|rs) as n^2 x n^2
  Fci(const Integrals& I_, int nelec, double mu_)
      : I(I_), n(I_.norb), na(nelec / 2), nb(nelec - nelec / 2), mu(mu_), A(n, na), B(n, nb), Ap(n, na + 1),
        Bm(n, nb - 1) {
    ea = exc_list(A);
    eb = exc_list(B);
    k1.assign(n * n, 0.0);
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q) {
        double s = I.t(p, q);
        for (int r = 0; r < n; ++r) s -= 0.5 * I.v(p, r, r, q);
        k1[p * n + q] = s;
      }
    vpq.assign((size_t)n * n * n * n, 0.0);
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q)
        for (int r = 0; r < n; ++r)
          for (int s = 0; s < n; ++s) vpq[(size_t)(p * n + q) * n * n + r * n + s] = I.v(p, q, r, s);
  }
  long dim() const { return (long)A.size() * B.size(); }

  void sigma(const double* c, double* out) const {
    const int na_ = A.size(), nb_ = B.size(), nn = n * n;
    const long N = dim();
    std::vector<double> D((size_t)nn * N, 0.0);
    // D_pq = E_pq c
#pragma omp parallel for schedule(dynamic)
    for (int Ia = 0; Ia < na_; ++Ia) {
      for (int Ib = 0; Ib < nb_; ++Ib) {
        double cv = c[(long)Ia * nb_ + Ib];
        if (cv == 0.0) continue;
        for (const auto& e : ea[Ia]) D[(size_t)e.pq * N + (long)e.J * nb_ + Ib] += e.sign * cv;
      }
    }
#pragma omp parallel for schedule(dynamic)
    for (int Ia = 0; Ia < na_; ++Ia)
      for (int Ib = 0; Ib < nb_; ++Ib) {
        double cv = c[(long)Ia * nb_ + Ib];
        if (cv == 0.0) continue;
        for (const auto& e : eb[Ib]) D[(size_t)e.pq * N + (long)Ia * nb_ + e.J] += e.sign * cv;
      }
    // G_rs = 1/2 sum_pq (rs|pq) D_pq + k_rs (one-body) carried as E_rs G_rs
    std::vector<double> G((size_t)nn * N, 0.0);
#pragma omp parallel for schedule(static)
    for (long x = 0; x < N; ++x)
      for (int rs = 0; rs < nn; ++rs) {
        double s = 0.0;
        for (int pq = 0; pq < nn; ++pq) s += vpq[(size_t)rs * nn + pq] * D[(size_t)pq * N + x];
        G[(size_t)rs * N + x] = 0.5 * s;
      }
    for (int rs = 0; rs < nn; ++rs)
      for (long x = 0; x < N; ++x) G[(size_t)rs * N + x] += k1[rs] * c[x];
    std::fill(out, out + N, 0.0);
    // out += sum_rs E_rs G_rs  (gather form: out(J) from sources I with E_rs I = J)
    // E_rs acting: out(Ja,Ib) += sign * G_rs(Ia,Ib) for a+_r a_s Ia = Ja
    std::vector<double> tmp(N, 0.0);
    for (int Ia = 0; Ia < na_; ++Ia)
      for (const auto& e : ea[Ia])
        for (int Ib = 0; Ib < nb_; ++Ib)
          out[(long)e.J * nb_ + Ib] += e.sign * G[(size_t)e.pq * N + (long)Ia * nb_ + Ib];
    for (int Ia = 0; Ia < na_; ++Ia)
      for (int Ib = 0; Ib < nb_; ++Ib)
        for (const auto& e : eb[Ib]) out[(long)Ia * nb_ + e.J] += e.sign * G[(size_t)e.pq * N + (long)Ia * nb_ + Ib];
    if (mu != 0.0) spin_penalty(c, out);
  }

  // out += mu * S_- S_+ c
  void spin_penalty(const double* c, double* out) const {
    if (nb == 0) return;
    const int na_ = A.size(), nb_ = B.size(), nap = Ap.size(), nbm = Bm.size();
    std::vector<double> t((size_t)nap * nbm, 0.0);
    // S_+ = sum_p a+_{p,up} a_{p,dn}
    for (int Ia = 0; Ia < na_; ++Ia)
      for (int Ib = 0; Ib < nb_; ++Ib) {
        double cv = c[(long)Ia * nb_ + Ib];
        if (cv == 0.0) continue;
        uint32_t xa = A.s[Ia], xb = B.s[Ib];
        for (int p = 0; p < n; ++p) {
          if (!(xb & (1u << p)) || (xa & (1u << p))) continue;
          int s = na + below(xb, p);           // a_{p,dn}: pass alpha string and lower betas
          uint32_t yb = xb & ~(1u << p);
          s += below(xa, p);                   // a+_{p,up}
          uint32_t ya = xa | (1u << p);
          t[(size_t)Ap.idx.at(ya) * nbm + Bm.idx.at(yb)] += ((s & 1) ? -1.0 : 1.0) * cv;
        }
      }
    // S_- = sum_p a+_{p,dn} a_{p,up}
    for (int Ja = 0; Ja < nap; ++Ja)
      for (int Jb = 0; Jb < nbm; ++Jb) {
        double tv = t[(size_t)Ja * nbm + Jb];
        if (tv == 0.0) continue;
        uint32_t xa = Ap.s[Ja], xb = Bm.s[Jb];
        for (int p = 0; p < n; ++p) {
          if (!(xa & (1u << p)) || (xb & (1u << p))) continue;
          int s = below(xa, p);
          uint32_t ya = xa & ~(1u << p);
          s += na + below(xb, p);  // alpha string now holds na electrons
          uint32_t yb = xb | (1u << p);
          out[(long)A.idx.at(ya) * nb_ + B.idx.at(yb)] += mu * ((s & 1) ? -1.0 : 1.0) * tv;
        }
      }
  }

  std::vector<double> diag() const {
    const int na_ = A.size(), nb_ = B.size();
    std::vector<double> d(dim());
    for (int Ia = 0; Ia < na_; ++Ia)
      for (int Ib = 0; Ib < nb_; ++Ib) {
        uint32_t xa = A.s[Ia], xb = B.s[Ib];
        double e = 0.0;
        std::vector<int> oa, ob;
        for (int p = 0; p < n; ++p) {
          if (xa & (1u << p)) oa.push_back(p);
          if (xb & (1u << p)) ob.push_back(p);
        }
        for (int p : oa) e += I.t(p, p);
        for (int p : ob) e += I.t(p, p);
        for (int p : oa)
          for (int q : oa) e += 0.5 * (I.v(p, p, q, q) - I.v(p, q, q, p));
        for (int p : ob)
          for (int q : ob) e += 0.5 * (I.v(p, p, q, q) - I.v(p, q, q, p));
        for (int p : oa)
          for (int q : ob) e += I.v(p, p, q, q);
        int open_b = 0;
        for (int p : ob)
          if (!(xa & (1u << p))) ++open_b;
        d[(long)Ia * nb_ + Ib] = e + mu * open_b;
      }
    return d;
  }
};

double dotv(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
}  // namespace

FciResult fci_lowest(const Integrals& I, int nelec, int nroots, double s2_penalty, double tol, int irrep) {
  Fci F(I, nelec, s2_penalty);
  const long N = F.dim();
  auto dg = F.diag();
  if (irrep >= 0) {  // restrict to determinants of the target irrep (H preserves it)
    const int nb_ = F.B.size();
    for (long x = 0; x < N; ++x) {
      uint32_t occ_a = F.A.s[x / nb_], occ_b = F.B.s[x % nb_];
      int g = 0;
      for (int p = 0; p < I.norb; ++p) {
        if (occ_a & (1u << p)) g ^= I.orbsym[p];
        if (occ_b & (1u << p)) g ^= I.orbsym[p];
      }
      if (g != irrep) dg[x] = 1e30;
    }
  }
  std::vector<long> order(N);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](long a, long b) { return dg[a] < dg[b]; });
  std::vector<std::vector<double>> b, s;
  int nguess = std::min<long>(N, std::max(nroots + 4, 2 * nroots));
  for (int g = 0; g < nguess; ++g) {
    if (dg[order[g]] > 1e29) break;
    std::vector<double> x(N, 0.0);
    x[order[g]] = 1.0;
    b.push_back(x);
  }
  FciResult res;
  std::vector<double> evals;
  std::vector<std::vector<double>> ritz;
  for (int iter = 0; iter < 400; ++iter) {
    // orthonormalise new vectors and compute their sigma
    for (size_t i = s.size(); i < b.size(); ++i) {
      for (size_t j = 0; j < i; ++j) {
        double o = dotv(b[j], b[i]);
        for (long k = 0; k < N; ++k) b[i][k] -= o * b[j][k];
      }
      double nn = std::sqrt(dotv(b[i], b[i]));
      for (long k = 0; k < N; ++k) b[i][k] /= nn;
      std::vector<double> si(N);
      F.sigma(b[i].data(), si.data());
      s.push_back(si);
    }
    int m = (int)b.size();
    std::vector<double> h((size_t)m * m), w(m);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) h[(size_t)i * m + j] = dotv(b[i], s[j]);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < i; ++j) h[(size_t)i * m + j] = h[(size_t)j * m + i] = 0.5 * (h[(size_t)i * m + j] + h[(size_t)j * m + i]);
    syev(m, h.data(), w.data());
    std::vector<std::vector<double>> newb;
    bool conv = true;
    ritz.assign(nroots, std::vector<double>(N, 0.0));
    std::vector<std::vector<double>> rs(nroots, std::vector<double>(N, 0.0));
    for (int r = 0; r < nroots && r < m; ++r) {
      for (int i = 0; i < m; ++i) {
        double a = h[(size_t)i * m + r];
        for (long k = 0; k < N; ++k) {
          ritz[r][k] += a * b[i][k];
          rs[r][k] += a * s[i][k];
        }
      }
      std::vector<double> res_(N);
      for (long k = 0; k < N; ++k) res_[k] = rs[r][k] - w[r] * ritz[r][k];
      double rn = std::sqrt(dotv(res_, res_));
      if (rn > tol) {
        conv = false;
        for (long k = 0; k < N; ++k) {
          double den = w[r] - dg[k];
          res_[k] = (dg[k] > 1e29) ? 0.0 : res_[k] / (std::fabs(den) > 1e-8 ? den : 1e-8);
        }
        newb.push_back(res_);
      }
    }
    evals.assign(w.begin(), w.begin() + std::min(nroots, m));
    if (conv) break;
    if (b.size() + newb.size() > 60) {  // restart from the Ritz vectors
      b = ritz;
      s = rs;
    }
    for (auto& x : newb) {
      for (auto& y : b) {
        double o = dotv(y, x);
        for (long k = 0; k < N; ++k) x[k] -= o * y[k];
      }
      double nn = std::sqrt(dotv(x, x));
      if (nn > 1e-10) {
        for (auto& v : x) v /= nn;
        b.push_back(x);
      }
    }
  }
  // report energies without the penalty: E = <H> = lambda - mu <S^2>
  for (int r = 0; r < (int)evals.size(); ++r) {
    double s2 = 0.0;
    if (s2_penalty != 0.0) {
      Fci F0(I, nelec, 1.0);
      std::vector<double> out(N, 0.0);
      F0.spin_penalty(ritz[r].data(), out.data());
      s2 = dotv(ritz[r], out) / dotv(ritz[r], ritz[r]);
    }
    res.energies.push_back(evals[r] - s2_penalty * s2 + I.ecore);
    res.s2.push_back(s2);
  }
  return res;
}

}  // namespace oracle

namespace oracle {
// dense spectrum of H (no penalty) in the MS=0 determinant space; debugging aid
std::vector<double> fci_dense_spectrum(const blk::Integrals& I, int nelec, double mu) {
  Fci F(I, nelec, mu);
  long N = F.dim();
  std::vector<double> H((size_t)N * N), e(N), v(N), w(N);
  for (long i = 0; i < N; ++i) {
    std::fill(e.begin(), e.end(), 0.0);
    e[i] = 1.0;
    F.sigma(e.data(), v.data());
    for (long j = 0; j < N; ++j) H[(size_t)j * N + i] = v[j];
  }
  blk::syev(N, H.data(), w.data());
  for (auto& x : w) x += I.ecore;
  return w;
}
}  // namespace oracle
