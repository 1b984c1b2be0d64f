// SU(2) coupling coefficients from the Racah closed forms (the published
// formulas; see e.g. Varshalovich et al., "Quantum Theory of Angular
// Momentum", ch. 8-10).  Plays the role of new_anglib.C / couplingCoeffs.C in
// the reference: clebsch (new_anglib.C:134), six_j (:73), nine_j (:19) and the
// normalised Ninej (couplingCoeffs.h:88-95).  All arguments are doubled spins.
#include <unordered_map>

#include "blk.h"

namespace blk {
namespace {

struct FactTable {
  long double f[256];
  FactTable() {
    f[0] = 1.0L;
    for (int i = 1; i < 256; ++i) f[i] = f[i - 1] * (long double)i;
  }
};
const FactTable& facts() {
  static FactTable t;
  return t;
}
inline long double fct(int n) { return facts().f[n]; }

// Delta(abc) for doubled args (integer sum assumed)
inline long double tri_coef(int a, int b, int c) {
  return fct((a + b - c) / 2) * fct((a - b + c) / 2) * fct((-a + b + c) / 2) / fct((a + b + c) / 2 + 1);
}

}  // namespace

double cg(int j1, int m1, int j2, int m2, int J, int M) {
  if (m1 + m2 != M) return 0.0;
  if (!triangle(j1, j2, J)) return 0.0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(M) > J) return 0.0;
  if (((j1 + m1) & 1) || ((j2 + m2) & 1) || ((J + M) & 1)) return 0.0;
  long double pre = (long double)(J + 1) * tri_coef(j1, j2, J);
  pre *= fct((j1 + m1) / 2) * fct((j1 - m1) / 2) * fct((j2 + m2) / 2) * fct((j2 - m2) / 2) * fct((J + M) / 2) *
         fct((J - M) / 2);
  long double s = 0.0L;
  for (int k = 0;; ++k) {
    int a1 = (j1 + j2 - J) / 2 - k, a2 = (j1 - m1) / 2 - k, a3 = (j2 + m2) / 2 - k;
    int a4 = (J - j2 + m1) / 2 + k, a5 = (J - j1 - m2) / 2 + k;
    if (a1 < 0 || a2 < 0 || a3 < 0) break;
    if (a4 < 0 || a5 < 0) continue;
    long double d = fct(k) * fct(a1) * fct(a2) * fct(a3) * fct(a4) * fct(a5);
    s += ((k & 1) ? -1.0L : 1.0L) / d;
  }
  return (double)(std::sqrt(pre) * s);
}

double sixj(int a, int b, int c, int d, int e, int f) {
  if (!triangle(a, b, c) || !triangle(a, e, f) || !triangle(d, b, f) || !triangle(d, e, c)) return 0.0;
  long double pre = tri_coef(a, b, c) * tri_coef(a, e, f) * tri_coef(d, b, f) * tri_coef(d, e, c);
  int t0 = std::max(std::max((a + b + c) / 2, (a + e + f) / 2), std::max((d + b + f) / 2, (d + e + c) / 2));
  int t1 = std::min(std::min((a + b + d + e) / 2, (a + c + d + f) / 2), (b + c + e + f) / 2);
  long double s = 0.0L;
  for (int t = t0; t <= t1; ++t) {
    long double den = fct(t - (a + b + c) / 2) * fct(t - (a + e + f) / 2) * fct(t - (d + b + f) / 2) *
                      fct(t - (d + e + c) / 2) * fct((a + b + d + e) / 2 - t) * fct((a + c + d + f) / 2 - t) *
                      fct((b + c + e + f) / 2 - t);
    s += ((t & 1) ? -1.0L : 1.0L) * fct(t + 1) / den;
  }
  return (double)(std::sqrt(pre) * s);
}

double ninej(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
  if (!triangle(a, b, c) || !triangle(d, e, f) || !triangle(g, h, i) || !triangle(a, d, g) || !triangle(b, e, h) ||
      !triangle(c, f, i))
    return 0.0;
  int xlo = std::max(std::max(std::abs(h - d), std::abs(b - f)), std::abs(a - i));
  int xhi = std::min(std::min(h + d, b + f), a + i);
  double s = 0.0;
  for (int x = xlo; x <= xhi; x += 2) {
    double t = (x + 1) * sixj(a, b, c, f, i, x) * sixj(d, e, f, b, x, h) * sixj(g, h, i, x, a, d);
    s += (x & 1) ? -t : t;
  }
  return s;
}

namespace {
struct Key9 {
  uint64_t k;
  bool operator==(const Key9& o) const { return k == o.k; }
};
struct Hash9 {
  size_t operator()(const Key9& x) const { return std::hash<uint64_t>()(x.k); }
};
}  // namespace

double wninej(int jlk, int jrk, int jk, int ka, int kb, int K, int jlb, int jrb, int jb) {
  // all arguments are small non-negative ints (< 128)
  uint64_t key = 0;
  for (int v : {jlk, jrk, jk, ka, kb, K, jlb, jrb, jb}) key = (key << 7) | (uint64_t)(v & 127);
  thread_local std::unordered_map<Key9, double, Hash9> cache;
  auto it = cache.find(Key9{key});
  if (it != cache.end()) return it->second;
  double w = std::sqrt((double)(jlb + 1) * (jrb + 1) * (jk + 1) * (K + 1)) * ninej(jlk, jrk, jk, ka, kb, K, jlb, jrb, jb);
  cache[Key9{key}] = w;
  return w;
}

double dagger_phase(int J2, int Jp2, int k2) {
  uint64_t key = ((uint64_t)J2 << 16) | ((uint64_t)Jp2 << 8) | (uint64_t)k2;
  thread_local std::unordered_map<uint64_t, double> cache;
  {
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  double ph = 0.0;
  bool found = false;
  for (int M = -J2; M <= J2 && !found; M += 2)
    for (int q = -k2; q <= k2 && !found; q += 2) {
      int Mp = M - q;
      if (std::abs(Mp) > Jp2) continue;
      double den = cg(Jp2, Mp, k2, q, J2, M);
      if (std::fabs(den) < 1e-12) continue;
      double num = cg(J2, M, k2, -q, Jp2, Mp);
      int e = (k2 - q) / 2;
      ph = ((e & 1) ? -1.0 : 1.0) * num / den;
      found = true;
    }
  if (!found) throw std::runtime_error("dagger_phase: spins do not couple");
  cache[key] = ph;
  return ph;
}

}  // namespace blk
