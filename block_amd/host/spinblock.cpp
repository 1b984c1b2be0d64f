// Dot-site operators, composite-block construction (BuildSumBlock and the
// Op::build recursions) and renormalisation.
#include "spinblock.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <utility>

namespace blk {

static const double SQ2 = std::sqrt(2.0), SQ3 = std::sqrt(3.0), SQ32 = std::sqrt(1.5);

// ============================================================== integrals
Integrals Integrals::reordered(const std::vector<int>& perm) const {
  Integrals o = *this;
  size_t n = norb;
  for (size_t a = 0; a < n; ++a) {
    o.orbsym[a] = orbsym[perm[a]];
    for (size_t b = 0; b < n; ++b) o.h1[a * n + b] = h1[perm[a] * n + perm[b]];
  }
  for (size_t a = 0; a < n; ++a)
    for (size_t b = 0; b < n; ++b)
      for (size_t c = 0; c < n; ++c)
        for (size_t d = 0; d < n; ++d)
          o.h2[((a * n + b) * n + c) * n + d] = h2[((perm[a] * n + perm[b]) * n + perm[c]) * n + perm[d]];
  return o;
}

Integrals read_fcidump(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open FCIDUMP " + path);
  std::string header, line;
  // header runs until a line containing &END or /
  while (std::getline(f, line)) {
    header += line + " ";
    std::string u = line;
    for (auto& ch : u) ch = toupper(ch);
    if (u.find("&END") != std::string::npos || u.find("/") != std::string::npos) break;
  }
  for (auto& ch : header)
    if (ch == ',' || ch == '=' || ch == '&') ch = ' ';
  std::istringstream hs(header);
  std::string tok;
  Integrals I;
  std::vector<std::string> toks;
  while (hs >> tok) toks.push_back(tok);
  for (size_t i = 0; i < toks.size(); ++i) {
    std::string u = toks[i];
    for (auto& ch : u) ch = toupper(ch);
    if (u == "NORB") I.norb = std::stoi(toks[++i]);
    else if (u == "NELEC") I.nelec = std::stoi(toks[++i]);
    else if (u == "MS2") I.ms2 = std::stoi(toks[++i]);
    else if (u == "ISYM") I.isym = std::stoi(toks[++i]) - 1;
    else if (u == "ORBSYM") {
      while (i + 1 < toks.size() && std::isdigit((unsigned char)toks[i + 1][0])) I.orbsym.push_back(std::stoi(toks[++i]) - 1);
    }
  }
  if (I.norb <= 0) throw std::runtime_error("FCIDUMP: NORB missing");
  if ((int)I.orbsym.size() != I.norb) I.orbsym.assign(I.norb, 0);
  if (I.isym < 0) I.isym = 0;
  size_t n = I.norb;
  I.h1.assign(n * n, 0.0);
  I.h2.assign(n * n * n * n, 0.0);
  double val;
  int i, j, k, l;
  while (f >> val >> i >> j >> k >> l) {
    if (i == 0 && j == 0 && k == 0 && l == 0) {
      I.ecore = val;
    } else if (k == 0 && l == 0) {
      --i, --j;
      I.h1[i * n + j] = I.h1[j * n + i] = val;
    } else {
      --i, --j, --k, --l;
      int idx[8][4] = {{i, j, k, l}, {j, i, k, l}, {i, j, l, k}, {j, i, l, k},
                       {k, l, i, j}, {l, k, i, j}, {k, l, j, i}, {l, k, j, i}};
      for (auto& q : idx) I.h2[((q[0] * n + q[1]) * n + q[2]) * n + q[3]] = val;
    }
  }
  return I;
}

void write_fcidump(const std::string& path, const Integrals& I) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fprintf(f, " &FCI NORB=%d,NELEC=%d,MS2=%d,\n  ORBSYM=", I.norb, I.nelec, I.ms2);
  for (int a = 0; a < I.norb; ++a) std::fprintf(f, "%d,", I.orbsym[a] + 1);
  std::fprintf(f, "\n  ISYM=%d,\n &END\n", I.isym + 1);
  int n = I.norb;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k < n; ++k)
        for (int l = 0; l <= k; ++l) {
          if (i * (i + 1) / 2 + j < k * (k + 1) / 2 + l) continue;
          double v = I.v(i, j, k, l);
          if (v != 0.0) std::fprintf(f, "%23.16e %4d %4d %4d %4d\n", v, i + 1, j + 1, k + 1, l + 1);
        }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j)
      if (I.t(i, j) != 0.0) std::fprintf(f, "%23.16e %4d %4d %4d %4d\n", I.t(i, j), i + 1, j + 1, 0, 0);
  std::fprintf(f, "%23.16e %4d %4d %4d %4d\n", I.ecore, 0, 0, 0, 0);
  std::fclose(f);
}

// ============================================================== screening
bool Screen::cd_interacts(const Integrals& I, int i, int j, const std::vector<int>& other) const {
  if (other.empty()) return true;
  for (int k : other)
    for (int l : other)
      if (std::fabs(I.v(i, j, k, l)) >= thresh || std::fabs(I.v(i, l, k, j)) >= thresh) return true;
  return false;
}
bool Screen::dd_interacts(const Integrals& I, int i, int j, const std::vector<int>& other) const {
  if (other.empty()) return true;
  for (int k : other)
    for (int l : other)
      if (std::fabs(I.v(i, k, j, l)) >= thresh) return true;
  return false;
}

std::function<bool(const Quantum&)> keep_for(const Quantum& target, int nrem) {
  return [target, nrem](const Quantum& q) {
    int np = target.n - q.n;
    if (np < 0 || np > 2 * nrem) return false;
    if (nrem == 0) return q == target;
    int smax = std::min(np, 2 * nrem - np);
    for (int sp = np & 1; sp <= smax; sp += 2)
      if (triangle(q.s2, sp, target.s2)) return true;
    return false;
  };
}

size_t SpinBlock::nops() const {
  size_t c = 0;
  for (auto* v : {&cre, &ccd, &cd, &cc, &cdc, &ddc})
    for (auto& p : *v) c += p ? 1 : 0;
  return c + 2;
}
size_t SpinBlock::nbytes() const {
  size_t c = 0;
  for (auto* v : {&cre, &ccd, &cd, &cc, &cdc, &ddc})
    for (auto& p : *v) c += p ? p->data.size() * 8 : 0;
  return c + (ham ? ham->data.size() * 8 : 0) + (ident ? ident->data.size() * 8 : 0);
}

static void init_arrays(SpinBlock& b, int norb) {
  b.norb = norb;
  b.inblock.assign(norb, 0);
  for (int s : b.sites) b.inblock[s] = 1;
  b.cre.assign(norb, nullptr);
  b.ccd.assign(norb, nullptr);
  size_t np = (size_t)norb * norb * 2;
  b.cd.assign(np, nullptr);
  b.cc.assign(np, nullptr);
  b.cdc.assign(np, nullptr);
  b.ddc.assign(np, nullptr);
}

// ============================================================== dot block
namespace {
using Mat4 = std::array<double, 16>;
Mat4 zero4() { Mat4 m; m.fill(0.0); return m; }
Mat4 mul4(const Mat4& a, const Mat4& b) {
  Mat4 c = zero4();
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k)
      for (int j = 0; j < 4; ++j) c[i * 4 + j] += a[i * 4 + k] * b[k * 4 + j];
  return c;
}
Mat4 tr4(const Mat4& a) {
  Mat4 c;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) c[i * 4 + j] = a[j * 4 + i];
  return c;
}
void axpy4(double s, const Mat4& x, Mat4& y) {
  for (int i = 0; i < 16; ++i) y[i] += s * x[i];
}
// Sz-resolved spherical tensor on the single-site Fock space
// {|0>, a+_up|0>, a+_dn|0>, a+_up a+_dn|0>}: comp[q2 + k2] = component q.
struct T4 {
  int k2 = 0;
  std::vector<Mat4> comp;
  explicit T4(int k = 0) : k2(k), comp(k + 1, zero4()) {}
  Mat4& at(int q2) { return comp[(q2 + k2) / 2]; }
  const Mat4& at(int q2) const { return comp[(q2 + k2) / 2]; }
};
T4 couple4(const T4& a, const T4& b, int K) {
  T4 c(K);
  for (int Q = -K; Q <= K; Q += 2)
    for (int qa = -a.k2; qa <= a.k2; qa += 2) {
      int qb = Q - qa;
      if (std::abs(qb) > b.k2) continue;
      double w = cg(a.k2, qa, b.k2, qb, K, Q);
      if (w == 0.0) continue;
      axpy4(w, mul4(a.at(qa), b.at(qb)), c.at(Q));
    }
  return c;
}
T4 lin4(double x, const T4& a, double y, const T4& b) {
  T4 c(a.k2);
  for (size_t i = 0; i < c.comp.size(); ++i) {
    axpy4(x, a.comp[i], c.comp[i]);
    axpy4(y, b.comp[i], c.comp[i]);
  }
  return c;
}
T4 scal4(double x, const T4& a) { return lin4(x, a, 0.0, a); }

// Fock index of (dot sector, 2M)
int fock_index(int sec, int m2) {
  if (sec == 0) return 0;
  if (sec == 1) return m2 > 0 ? 1 : 2;
  return 3;
}

// Reduced (Wigner-Eckart) form of an explicit dot tensor; checks that the
// explicit components really obey the theorem.
OpPtr reduce4(const SIPtr& si, const T4& t, int dn, int dirrep, bool fermion) {
  auto op = std::make_shared<SparseOp>(get_pattern(si, dn, t.k2, dirrep), fermion);
  const auto& P = *op->pat;
  std::vector<char> covered(16 * t.comp.size(), 0);
  for (size_t ib = 0; ib < P.blocks.size(); ++ib) {
    const auto& b = P.blocks[ib];
    int Sr = si->q[b.r].s2, Sc = si->q[b.c].s2;
    double red = 0.0;
    bool have = false;
    for (int mc = -Sc; mc <= Sc; mc += 2)
      for (int q = -t.k2; q <= t.k2; q += 2) {
        int mr = mc + q;
        if (std::abs(mr) > Sr) continue;
        double w = cg(Sc, mc, t.k2, q, Sr, mr);
        int fr = fock_index(b.r, mr), fc = fock_index(b.c, mc);
        double val = t.at(q)[fr * 4 + fc];
        covered[((q + t.k2) / 2) * 16 + fr * 4 + fc] = 1;
        if (std::fabs(w) < 1e-12) {
          if (std::fabs(val) > 1e-12) throw std::runtime_error("reduce4: component outside Wigner-Eckart support");
          continue;
        }
        if (!have) {
          red = val / w;
          have = true;
        } else if (std::fabs(val - w * red) > 1e-12) {
          throw std::runtime_error("reduce4: explicit operator is not a spherical tensor");
        }
      }
    op->blk((int)ib)[0] = red;
  }
  for (size_t qi = 0; qi < t.comp.size(); ++qi)
    for (int e = 0; e < 16; ++e)
      if (!covered[qi * 16 + e] && std::fabs(t.comp[qi][e]) > 1e-12)
        throw std::runtime_error("reduce4: element outside the allowed sector blocks");
  return op;
}
}  // namespace

BlockPtr make_dot(const Integrals& I, int d, const Screen& scr) {
  // dot operators are written element by element on the host
  if (build_recording()) throw std::runtime_error("make_dot: not inside a build recording");
  const int n = I.norb;
  const int g = I.orbsym[d];
  auto si = std::make_shared<StateInfo>();
  si->q = {Quantum{0, 0, 0}, Quantum{1, 1, g}, Quantum{2, 0, 0}};
  si->dim = {1, 1, 1};
  si->finalize();
  auto b = std::make_shared<SpinBlock>();
  b->sites = {d};
  init_arrays(*b, n);
  b->si = si;
  b->has_normal = b->has_comp = true;

  // explicit single-site operators
  Mat4 cu = zero4(), cdn = zero4();
  cu[1 * 4 + 0] = 1.0;   // a+_up |0> = |up>
  cu[3 * 4 + 2] = 1.0;   // a+_up |dn> = |updn>
  cdn[2 * 4 + 0] = 1.0;  // a+_dn |0> = |dn>
  cdn[3 * 4 + 1] = -1.0; // a+_dn |up> = -|updn>
  Mat4 au = tr4(cu), ad = tr4(cdn);
  T4 C(1), D(1), Id(0), N(0), Nud(0);
  C.at(1) = cu;
  C.at(-1) = cdn;
  D.at(1) = ad;           // D_{+1/2} = a_dn
  for (int i = 0; i < 16; ++i) D.at(-1)[i] = -au[i];  // D_{-1/2} = -a_up
  for (int i = 0; i < 4; ++i) Id.at(0)[i * 4 + i] = 1.0;
  N.at(0) = mul4(cu, au);
  axpy4(1.0, mul4(cdn, ad), N.at(0));
  Nud.at(0) = mul4(mul4(cu, au), mul4(cdn, ad));
  T4 CD0 = couple4(C, D, 0), CD1 = couple4(C, D, 2), CC0 = couple4(C, C, 0), DD0 = couple4(D, D, 0);
  T4 CN(1);
  for (int q : {-1, 1}) CN.at(q) = mul4(C.at(q), N.at(0));

  b->ident = reduce4(si, Id, 0, 0, false);
  b->cre[d] = reduce4(si, C, 1, g, true);
  b->extra["des"] = reduce4(si, D, -1, g, true);
  b->extra["n"] = reduce4(si, N, 0, 0, false);
  b->extra["cn"] = reduce4(si, CN, 1, g, true);
  b->extra["dd0"] = reduce4(si, DD0, -2, 0, false);
  b->cd[b->pidx(d, d, 0)] = reduce4(si, CD0, 0, 0, false);
  b->cd[b->pidx(d, d, 2)] = reduce4(si, CD1, 0, 0, false);
  b->cc[b->pidx(d, d, 0)] = reduce4(si, CC0, 2, 0, false);
  b->ham = reduce4(si, lin4(I.t(d, d), N, I.v(d, d, d, d), Nud), 0, 0, false);

  std::vector<int> me{d};
  // complementary two-index operators for every pair outside the dot
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      if (i == d || j == d) continue;
      int gij = irrep_mul(I.orbsym[i], I.orbsym[j]);
      if (scr.cd_interacts(I, i, j, me)) {
        double x0 = 2.0 * I.v(i, j, d, d) - I.v(i, d, d, j);
        double x1 = SQ3 * I.v(i, d, d, j);
        if (gij == 0) {
          b->cdc[b->pidx(i, j, 0)] = reduce4(si, scal4(x0, CD0), 0, 0, false);
          b->cdc[b->pidx(i, j, 2)] = reduce4(si, scal4(x1, CD1), 0, 0, false);
        }
      }
      if (scr.dd_interacts(I, i, j, me) && gij == 0) {
        // P^0_ij = w_0 (id|jd) [D_d D_d]^0 ; the K=1 part vanishes on one site
        b->ddc[b->pidx(i, j, 0)] = reduce4(si, scal4(-0.5 * I.v(i, d, j, d), DD0), -2, 0, false);
      }
    }
  // X_q(d) = t_dq/2 C_d + (dq|dd) C_d n_d
  for (int q = 0; q < n; ++q) {
    if (q == d) continue;
    if (I.orbsym[q] != g) {
      b->ccd[q] = std::make_shared<SparseOp>(get_pattern(si, 1, 1, I.orbsym[q]), true);
      continue;
    }
    b->ccd[q] = reduce4(si, lin4(0.5 * I.t(d, q), C, I.v(d, q, d, d), CN), 1, g, true);
  }
  return b;
}

// ============================================================== composite blocks
namespace {
struct Builder {
  const Integrals& I;
  const SpinBlock& B1;
  const SpinBlock& dot;
  int d;
  SIPtr si;
  const SparseOp& I1() const { return *B1.ident; }
  const SparseOp& Id() const { return *dot.ident; }
  const SparseOp& Cd() const { return *dot.cre[d]; }
  const SparseOp& Dd() const { return *dot.extra.at("des"); }

  OpPtr make(int dn, int k2, int irrep, bool fermion) const {
    return std::make_shared<SparseOp>(get_pattern(si, dn, k2, irrep), fermion);
  }
  // sum_k x_k C_k(B1) restricted to orbitals of the given irrep
  OpPtr cre_comb(const std::vector<double>& x, int irrep) const {
    OpPtr out;
    for (int k : B1.sites) {
      double c = x[k];
      if (c == 0.0) continue;
      if (I.orbsym[k] != irrep) {
        if (std::fabs(c) > 1e-10) throw std::runtime_error("integral breaks the point-group symmetry");
        continue;
      }
      const SparseOp& ck = *B1.cre[k];
      if (!out) out = std::make_shared<SparseOp>(ck.pat, true);
      out->axpy(c, ck);
    }
    return out;
  }

  OpPtr cre(int i) const {
    auto c = make(1, 1, I.orbsym[i], true);
    if (i == d)
      tensor_product(V(I1()), V(Cd()), *c, 1.0);
    else
      tensor_product(V(*B1.cre[i]), V(Id()), *c, 1.0);
    return c;
  }

  // [C_i D_j]^K or [C_i C_j]^K for i>=j in B
  OpPtr pair_normal(bool isCC, int i, int j, int K) const {
    int irr = irrep_mul(I.orbsym[i], I.orbsym[j]);
    auto c = make(isCC ? 2 : 0, K, irr, false);
    const auto& v1 = isCC ? B1.cc : B1.cd;
    const auto& vd = isCC ? dot.cc : dot.cd;
    bool i1 = B1.in(i), j1 = B1.in(j);
    double sgnK = (K == 2) ? -1.0 : 1.0;
    if (i1 && j1) {
      const SparseOp* o = B1.get(v1, i, j, K);
      if (!o) return nullptr;
      tensor_product(V(*o), V(Id()), *c, 1.0);
    } else if (i == d && j == d) {
      const SparseOp* o = dot.get(vd, d, d, K);
      if (!o) return nullptr;
      tensor_product(V(I1()), V(*o), *c, 1.0);
    } else if (i1) {  // (i in B1, j = d)
      tensor_product(V(*B1.cre[i]), isCC ? V(Cd()) : V(Dd()), *c, 1.0);
    } else {          // (i = d, j in B1): [X_d Y_j]^K = (-1)^K [Y_j(B1) X_d]^K
      if (isCC)
        tensor_product(V(*B1.cre[j]), V(Cd()), *c, sgnK);
      else
        tensor_product(Vd(*B1.cre[j]), V(Cd()), *c, sgnK);
    }
    return c;
  }

  // Q^K_ij (CDcomp) for i>=j not in B
  OpPtr cdcomp(int i, int j, int K) const {
    int irr = irrep_mul(I.orbsym[i], I.orbsym[j]);
    auto c = make(0, K, irr, false);
    bool any = false;
    if (const SparseOp* o = B1.get(B1.cdc, i, j, K)) {
      tensor_product(V(*o), V(Id()), *c, 1.0);
      any = true;
    }
    if (const SparseOp* o = dot.get(dot.cdc, i, j, K)) {
      tensor_product(V(I1()), V(*o), *c, 1.0);
      any = true;
    }
    std::vector<double> x1(I.norb, 0.0), x2(I.norb, 0.0);
    for (int k : B1.sites) {
      if (K == 0) {
        x1[k] = 2.0 * I.v(i, j, k, d) - I.v(i, d, k, j);  // coefficient of [C_k D_d]
        x2[k] = 2.0 * I.v(i, j, d, k) - I.v(i, k, d, j);  // coefficient of [C_d D_k]
      } else {
        x1[k] = SQ3 * I.v(i, d, k, j);
        x2[k] = SQ3 * I.v(i, k, d, j);
      }
    }
    int need = irrep_mul(irr, I.orbsym[d]);
    if (auto s1 = cre_comb(x1, need)) {
      tensor_product(V(*s1), V(Dd()), *c, 1.0);
      any = true;
    }
    if (auto s2 = cre_comb(x2, need)) {
      tensor_product(Vd(*s2), V(Cd()), *c, (K == 2) ? -1.0 : 1.0);
      any = true;
    }
    return any ? c : nullptr;
  }

  // P^K_ij (DDcomp) for i>=j not in B
  OpPtr ddcomp(int i, int j, int K) const {
    int irr = irrep_mul(I.orbsym[i], I.orbsym[j]);
    auto c = make(-2, K, irr, false);
    bool any = false;
    if (const SparseOp* o = B1.get(B1.ddc, i, j, K)) {
      tensor_product(V(*o), V(Id()), *c, 1.0);
      any = true;
    }
    if (const SparseOp* o = dot.get(dot.ddc, i, j, K)) {
      tensor_product(V(I1()), V(*o), *c, 1.0);
      any = true;
    }
    double w = (K == 0) ? -0.5 : 0.5 * SQ3;
    double sK = (K == 2) ? -1.0 : 1.0;
    std::vector<double> y(I.norb, 0.0);
    for (int s : B1.sites) y[s] = w * (I.v(i, d, j, s) + sK * I.v(i, s, j, d));
    if (auto ys = cre_comb(y, irrep_mul(irr, I.orbsym[d]))) {
      tensor_product(Vd(*ys), V(Dd()), *c, 1.0);
      any = true;
    }
    return any ? c : nullptr;
  }

  // view of Q^K_{ab} on B1 for an arbitrary ordered pair, via Q_ba = (-1)^K Q_ab^dag
  bool q_view(int a, int b, int K, OpView& out) const {
    double sK = (K == 2) ? -1.0 : 1.0;
    if (a >= b) {
      const SparseOp* o = B1.get(B1.cdc, a, b, K);
      if (!o) return false;
      out = V(*o);
    } else {
      const SparseOp* o = B1.get(B1.cdc, b, a, K);
      if (!o) return false;
      out = Vd(*o, sK);
    }
    return true;
  }
  // view of (P^K_{ab})^dag on B1, via P_ba = (-1)^K P_ab
  bool pdag_view(int a, int b, int K, OpView& out) const {
    double sK = (K == 2) ? -1.0 : 1.0;
    const SparseOp* o = (a >= b) ? B1.get(B1.ddc, a, b, K) : B1.get(B1.ddc, b, a, K);
    if (!o) return false;
    out = Vd(*o, a >= b ? 1.0 : sK);
    return true;
  }

  // X_q (CCDcomp) for q not in B
  OpPtr ccdcomp(int q) const {
    auto c = make(1, 1, I.orbsym[q], true);
    if (const SparseOp* o = B1.ccd[q].get()) tensor_product(V(*o), V(Id()), *c, 1.0);
    if (const SparseOp* o = dot.ccd[q].get()) tensor_product(V(I1()), V(*o), *c, 1.0);
    for (int K : {0, 2}) {
      OpView qv, pv;
      if (q_view(d, q, K, qv)) tensor_product(qv, V(Cd()), *c, -1.0 / SQ2);
      if (pdag_view(d, q, K, pv)) tensor_product(pv, V(Dd()), *c, SQ2);
    }
    std::vector<double> z1(I.norb, 0.0), z2(I.norb, 0.0);
    for (int p : B1.sites) {
      z1[p] = I.v(p, q, d, d);
      z2[p] = I.v(d, q, p, d);
    }
    int gq = I.orbsym[q];
    if (auto Z1 = cre_comb(z1, gq)) tensor_product(V(*Z1), V(*dot.extra.at("n")), *c, 1.0);
    if (auto Z2 = cre_comb(z2, gq)) {
      tensor_product(V(*Z2), V(*dot.get(dot.cd, d, d, 0)), *c, 1.0 / SQ2);
      tensor_product(V(*Z2), V(*dot.get(dot.cd, d, d, 2)), *c, SQ32);
      tensor_product(Vd(*Z2), V(*dot.get(dot.cc, d, d, 0)), *c, 1.0 / SQ2);
    }
    return c;
  }

  OpPtr ham() const {
    auto c = make(0, 0, 0, false);
    tensor_product(V(*B1.ham), V(Id()), *c, 1.0);
    tensor_product(V(I1()), V(*dot.ham), *c, 1.0);
    // three indices in B1, one on the dot
    if (const SparseOp* x = B1.ccd[d].get()) {
      tensor_product(V(*x), V(Dd()), *c, -SQ2);
      tensor_product(Vd(*x), V(Cd()), *c, -SQ2);
    }
    // one index in B1, three on the dot: X_q(d) = t_dq/2 C_d + (dq|dd) C_d n_d
    std::vector<double> zt(I.norb, 0.0), zv(I.norb, 0.0);
    for (int q : B1.sites) {
      zt[q] = 0.5 * I.t(d, q);
      zv[q] = I.v(d, q, d, d);
    }
    int gd = I.orbsym[d];
    const SparseOp& cn = *dot.extra.at("cn");
    if (auto Zt = cre_comb(zt, gd)) {
      tensor_product(Vd(*Zt), V(Cd()), *c, -SQ2);
      tensor_product(V(*Zt), Vd(Cd()), *c, -SQ2);
    }
    if (auto Zv = cre_comb(zv, gd)) {
      tensor_product(Vd(*Zv), V(cn), *c, -SQ2);
      tensor_product(V(*Zv), Vd(cn), *c, -SQ2);
    }
    // two and two: loop block is the dot, pair (d,d)
    for (int K : {0, 2}) {
      const SparseOp* q = B1.get(B1.cdc, d, d, K);
      if (q) tensor_product(V(*q), V(*dot.get(dot.cd, d, d, K)), *c, 1.0);
    }
    if (const SparseOp* p = B1.get(B1.ddc, d, d, 0)) {
      const SparseOp& cc = *dot.get(dot.cc, d, d, 0);
      tensor_product(V(*p), V(cc), *c, 1.0);
      tensor_product(Vd(*p), Vd(cc), *c, 1.0);
    }
    return c;
  }
};
}  // namespace

namespace {
struct Task {
  int kind, i, j, K;  // kind 0 cd, 1 cc, 2 cdc, 3 ddc, 4 ccd
};
void run_tasks(const Builder& bd, SpinBlock& b, const std::vector<Task>& tasks,
               const std::function<OpPtr(OpPtr)>& post) {
  std::string err;
  BuildSession::Impl* const session = current_build_session();  // workers record into the caller's
#pragma omp parallel for schedule(dynamic)
  for (long t = 0; t < (long)tasks.size(); ++t) {
    BuildSession::Impl* const outer = current_build_session();
    enter_build_session(session);
    try {
    const Task& k = tasks[t];
    OpPtr o;
    switch (k.kind) {
      case 0: o = bd.pair_normal(false, k.i, k.j, k.K); break;
      case 1: o = bd.pair_normal(true, k.i, k.j, k.K); break;
      case 2: o = bd.cdcomp(k.i, k.j, k.K); break;
      case 3: o = bd.ddcomp(k.i, k.j, k.K); break;
      default: o = bd.ccdcomp(k.i); break;
    }
    if (o && post) o = post(o);
    switch (k.kind) {
      case 0: b.cd[b.pidx(k.i, k.j, k.K)] = o; break;
      case 1: b.cc[b.pidx(k.i, k.j, k.K)] = o; break;
      case 2: b.cdc[b.pidx(k.i, k.j, k.K)] = o; break;
      case 3: b.ddc[b.pidx(k.i, k.j, k.K)] = o; break;
      default: b.ccd[k.i] = o; break;
    }
    } catch (const std::exception& e) {
#pragma omp critical(blk_task_err)
      if (err.empty()) err = e.what();
    }
    enter_build_session(outer);
  }
  if (!err.empty()) throw std::runtime_error(err);
}
std::vector<Task> comp_tasks(const Integrals& I, const SpinBlock& b, const std::vector<int>& outside,
                             const Screen& scr) {
  std::vector<Task> tasks;
  for (int i : outside)
    for (int j : outside) {
      if (j > i) continue;
      bool cdk = scr.cd_interacts(I, i, j, b.sites), ddk = scr.dd_interacts(I, i, j, b.sites);
      for (int K : {0, 2}) {
        if (cdk) tasks.push_back({2, i, j, K});
        if (ddk && !(i == j && K == 2)) tasks.push_back({3, i, j, K});
      }
    }
  return tasks;
}
}  // namespace

BlockPtr build_sum_block(const Integrals& I, const SpinBlock& B1, const SpinBlock& dot, const Quantum& target,
                         int nrem, const BuildSpec& spec, const Screen& scr) {
  if (dot.sites.size() != 1) throw std::runtime_error("build_sum_block: second block must be a single site");
  const int n = I.norb;
  const int d = dot.sites[0];
  auto si = make_product(B1.si, dot.si, keep_for(target, nrem));
  auto b = std::make_shared<SpinBlock>();
  b->sites = B1.sites;
  b->sites.push_back(d);
  std::sort(b->sites.begin(), b->sites.end());
  init_arrays(*b, n);
  b->si = si;
  b->has_normal = spec.normal;
  b->has_comp = spec.comp;
  Builder bd{I, B1, dot, d, si};
  std::vector<int> outside;
  for (int i = 0; i < n; ++i)
    if (!b->in(i)) outside.push_back(i);

  b->ident = bd.make(0, 0, 0, false);
  tensor_product(V(*B1.ident), V(*dot.ident), *b->ident, 1.0);
  for (int i : b->sites) b->cre[i] = bd.cre(i);
  std::vector<Task> tasks;
  if (spec.normal) {
    for (int i : b->sites)
      for (int j : b->sites) {
        if (j > i) continue;
        bool cdk = scr.cd_interacts(I, i, j, outside), ddk = scr.dd_interacts(I, i, j, outside);
        for (int K : {0, 2}) {
          if (cdk) tasks.push_back({0, i, j, K});
          if (ddk && !(i == j && K == 2)) tasks.push_back({1, i, j, K});
        }
      }
  }
  if (spec.comp) {
    auto ct = comp_tasks(I, *b, outside, scr);
    tasks.insert(tasks.end(), ct.begin(), ct.end());
  }
  if (spec.ccd)
    for (int q : outside) tasks.push_back({4, q, 0, 0});
  run_tasks(bd, *b, tasks, nullptr);
  b->ham = bd.ham();
  return b;
}

// ============================================================== renormalisation
Rotation make_rotation_info(const StateInfo& os, const std::vector<std::vector<double>>& rot,
                            const std::vector<int>& mkeep) {
  Rotation R;
  auto ns = std::make_shared<StateInfo>();
  for (int s = 0; s < os.nsec(); ++s)
    if (mkeep[s] > 0) {
      R.oldidx.push_back(s);
      ns->q.push_back(os.q[s]);
      ns->dim.push_back(mkeep[s]);
    }
  ns->finalize();
  R.ns = ns;
  R.old = &os;
  R.rot = &rot;
  return R;
}

OpPtr rotate_op(const SparseOp& o, const Rotation& R) {
  const StateInfo& os = *R.old;
  auto r = std::make_shared<SparseOp>(get_pattern(R.ns, o.dn(), o.k2(), o.dirrep()), o.fermion);
  const auto& P = *r->pat;
  std::vector<double> tmp;
  for (size_t ib = 0; ib < P.blocks.size(); ++ib) {
    const auto& bk = P.blocks[ib];
    int orow = R.oldidx[bk.r], ocol = R.oldidx[bk.c];
    int obi = o.pat->find(orow, ocol);
    if (obi < 0) continue;
    int dr = os.dim[orow], dc = os.dim[ocol], mr = bk.rows, mc = bk.cols;
    // tmp = O(orow,ocol) R_c  (dr x mc);  new = R_r^T tmp
    tmp.assign((size_t)dr * mc, 0.0);
    dgemm(false, false, dr, mc, dc, 1.0, o.blk(obi), dc, (*R.rot)[ocol].data(), mc, 0.0, tmp.data(), mc);
    dgemm(true, false, mr, mc, dr, 1.0, (*R.rot)[orow].data(), mr, tmp.data(), mc, 0.0, r->blk((int)ib), mc);
  }
  return r;
}

BlockPtr renormalise(const SpinBlock& B, const std::vector<std::vector<double>>& rot, const std::vector<int>& mkeep) {
  return renormalise(B, make_rotation_info(*B.si, rot, mkeep));
}

std::vector<OpPtr> OpRotator::rotate(const std::vector<const SparseOp*>& ops, const Rotation& R) const {
  std::vector<OpPtr> out(ops.size());
  std::string err;
#pragma omp parallel for schedule(dynamic)
  for (long w = 0; w < (long)ops.size(); ++w) {
    try {
      out[w] = rotate_op(*ops[w], R);
    } catch (const std::exception& e) {
#pragma omp critical(blk_task_err)
      if (err.empty()) err = e.what();
    }
  }
  if (!err.empty()) throw std::runtime_error(err);
  return out;
}

BlockPtr renormalise(const SpinBlock& B, const Rotation& R, const OpRotator* rotator) {
  if (R.old != B.si.get()) throw std::runtime_error("renormalise: rotation built for another block");
  auto nb = std::make_shared<SpinBlock>();
  nb->norb = B.norb;
  nb->sites = B.sites;
  nb->inblock = B.inblock;
  nb->si = R.ns;
  nb->has_normal = B.has_normal;
  nb->has_comp = B.has_comp;
  std::vector<std::pair<const std::vector<OpPtr>*, std::vector<OpPtr>*>> lists = {
      {&B.cre, &nb->cre}, {&B.ccd, &nb->ccd}, {&B.cd, &nb->cd}, {&B.cc, &nb->cc}, {&B.cdc, &nb->cdc}, {&B.ddc, &nb->ddc}};
  std::vector<const SparseOp*> src = {B.ident.get(), B.ham.get()};
  std::vector<OpPtr*> dst = {&nb->ident, &nb->ham};
  for (auto& l : lists) {
    l.second->assign(l.first->size(), nullptr);
    for (size_t i = 0; i < l.first->size(); ++i)
      if ((*l.first)[i]) {
        src.push_back((*l.first)[i].get());
        dst.push_back(&(*l.second)[i]);
      }
  }
  const OpRotator host;
  std::vector<OpPtr> rot = (rotator ? rotator : &host)->rotate(src, R);
  for (size_t w = 0; w < src.size(); ++w) *dst[w] = rot[w];
  return nb;
}

std::atomic<double> g_comp_build_seconds{0.0};
double comp_build_seconds() { return g_comp_build_seconds.load(); }

namespace {
// slots of out that the tasks fill, in task order
template <class SB>
std::vector<decltype(&std::declval<SB&>().cd[0])> task_slots(const std::vector<Task>& tasks, SB& out) {
  std::vector<decltype(&out.cd[0])> slot;
  for (const auto& k : tasks)
    slot.push_back(k.kind == 0   ? &out.cd[out.pidx(k.i, k.j, k.K)]
                   : k.kind == 1 ? &out.cc[out.pidx(k.i, k.j, k.K)]
                   : k.kind == 2 ? &out.cdc[out.pidx(k.i, k.j, k.K)]
                   : k.kind == 3 ? &out.ddc[out.pidx(k.i, k.j, k.K)]
                                 : &out.ccd[k.i]);
  return slot;
}
void add_seconds(std::atomic<double>& a, double x) {
  double cur = a.load();
  while (!a.compare_exchange_weak(cur, cur + x)) {
  }
}
}  // namespace

struct CompOpsAhead {
  const StateInfo* si = nullptr;  // the composite basis they are built in
  std::vector<Task> tasks;
  SpinBlock ops;  // the built operators, in the slots of their tasks
};

std::shared_ptr<CompOpsAhead> build_comp_ops(const Integrals& I, const SpinBlock& B, const SpinBlock& B1,
                                             const SpinBlock& dot, const Screen& scr) {
  auto a = std::make_shared<CompOpsAhead>();
  a->si = B.si.get();
  Builder bd{I, B1, dot, dot.sites[0], B.si};
  std::vector<int> outside;
  for (int i = 0; i < I.norb; ++i)
    if (!B.in(i)) outside.push_back(i);
  a->tasks = comp_tasks(I, B, outside, scr);
  a->ops.sites = B.sites;
  init_arrays(a->ops, I.norb);
  const auto t0 = std::chrono::steady_clock::now();
  run_tasks(bd, a->ops, a->tasks, nullptr);
  add_seconds(g_comp_build_seconds, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  return a;
}

std::vector<OpPtr> comp_ops_list(const CompOpsAhead& ahead) {
  std::vector<OpPtr> v;
  for (auto p : task_slots(ahead.tasks, ahead.ops))
    if (*p) v.push_back(*p);
  return v;
}

void add_rotated_comp_ops(const CompOpsAhead& ahead, const Rotation& R, SpinBlock& out, const OpRotator* rotator) {
  if (R.old != ahead.si) throw std::runtime_error("add_rotated_comp_ops: operators built for another block");
  const auto from = task_slots(ahead.tasks, ahead.ops);
  const auto to = task_slots(ahead.tasks, out);
  std::vector<const SparseOp*> src;
  std::vector<OpPtr*> dst;
  for (size_t t = 0; t < from.size(); ++t)
    if (*from[t]) {
      src.push_back(from[t]->get());
      dst.push_back(to[t]);
    }
  const OpRotator host;
  std::vector<OpPtr> rot = (rotator ? rotator : &host)->rotate(src, R);
  for (size_t w = 0; w < dst.size(); ++w) *dst[w] = rot[w];
  out.has_comp = true;
}

void add_rotated_comp_ops(const Integrals& I, const SpinBlock& B, const SpinBlock& B1, const SpinBlock& dot,
                          const Screen& scr, const Rotation& R, SpinBlock& out, const OpRotator* rotator) {
  const int d = dot.sites[0];
  Builder bd{I, B1, dot, d, B.si};
  std::vector<int> outside;
  for (int i = 0; i < I.norb; ++i)
    if (!B.in(i)) outside.push_back(i);
  auto tasks = comp_tasks(I, B, outside, scr);
  if (!rotator) {
    // built and rotated one at a time: only one composite operator per thread
    run_tasks(bd, out, tasks, [&](OpPtr o) { return rotate_op(*o, R); });
  } else {
    // all composite operators built first, then rotated as one batch
    const auto t0 = std::chrono::steady_clock::now();
    run_tasks(bd, out, tasks, nullptr);
    add_seconds(g_comp_build_seconds, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    std::vector<OpPtr*> slot;
    for (OpPtr* p : task_slots(tasks, out))
      if (*p) slot.push_back(p);
    std::vector<const SparseOp*> src;
    for (OpPtr* p : slot) src.push_back(p->get());
    std::vector<OpPtr> rot = rotator->rotate(src, R);
    for (size_t w = 0; w < slot.size(); ++w) *slot[w] = rot[w];
  }
  out.has_comp = true;
}

}  // namespace blk
