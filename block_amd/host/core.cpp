// State bookkeeping, block-sparse operator patterns, the composite-operator
// tensor product and small dense kernels of the host library.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <mutex>
#include <tuple>

#include "blk.h"

namespace blk {

// ---------------------------------------------------------------- StateInfo
int StateInfo::find(const Quantum& x) const {
  auto it = std::lower_bound(q.begin(), q.end(), x);
  if (it == q.end() || *it != x) return -1;
  return (int)(it - q.begin());
}

void StateInfo::finalize() {
  off.assign(q.size(), 0);
  total = 0;
  for (size_t i = 0; i < q.size(); ++i) {
    off[i] = total;
    total += dim[i];
  }
}

SIPtr make_product(const SIPtr& l, const SIPtr& r, const std::function<bool(const Quantum&)>& keep) {
  // collect (l, r, S) combinations by total quanta (StateInfo.C:123-340)
  std::map<Quantum, std::vector<StateInfo::Sub>> coll;
  for (int a = 0; a < l->nsec(); ++a)
    for (int b = 0; b < r->nsec(); ++b) {
      const Quantum &qa = l->q[a], &qb = r->q[b];
      for (int s2 = std::abs(qa.s2 - qb.s2); s2 <= qa.s2 + qb.s2; s2 += 2) {
        Quantum t{qa.n + qb.n, s2, irrep_mul(qa.irrep, qb.irrep)};
        if (!keep(t)) continue;
        coll[t].push_back(StateInfo::Sub{a, b, 0});
      }
    }
  auto si = std::make_shared<StateInfo>();
  si->left = l;
  si->right = r;
  si->prod_to_sector.assign((size_t)l->nsec() * r->nsec(), {});
  for (auto& kv : coll) {
    int sec = (int)si->q.size();
    int off = 0;
    for (auto& s : kv.second) {
      s.off = off;
      off += l->dim[s.l] * r->dim[s.r];
      si->prod_to_sector[(size_t)s.l * r->nsec() + s.r].push_back({sec, s.off});
    }
    si->q.push_back(kv.first);
    si->dim.push_back(off);
    si->subs.push_back(kv.second);
  }
  si->finalize();
  return si;
}

// ---------------------------------------------------------------- patterns
int OpPattern::find(int r, int c) const {
  for (int i = row_start[r]; i < row_start[r + 1]; ++i)
    if (blocks[i].c == c) return i;
  return -1;
}

PatPtr get_pattern(const SIPtr& si, int dn, int k2, int dirrep) {
  static std::mutex mu;
  static std::map<std::tuple<const StateInfo*, int, int, int>, std::weak_ptr<const OpPattern>> cache;
  auto key = std::make_tuple(si.get(), dn, k2, dirrep);
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(key);
  if (it != cache.end()) {
    if (auto p = it->second.lock()) return p;
  }
  auto p = std::make_shared<OpPattern>();
  p->si = si;
  p->dn = dn;
  p->k2 = k2;
  p->dirrep = dirrep;
  int ns = si->nsec();
  p->row_start.assign(ns + 1, 0);
  long off = 0;
  for (int r = 0; r < ns; ++r) {
    p->row_start[r] = (int)p->blocks.size();
    for (int c = 0; c < ns; ++c)
      if (p->allowed(si->q[r], si->q[c])) {
        p->blocks.push_back({r, c, si->dim[r], si->dim[c], off});
        off += (long)si->dim[r] * si->dim[c];
      }
  }
  p->row_start[ns] = (int)p->blocks.size();
  p->size = off;
  // drop expired entries now and then
  if (cache.size() > 4096) {
    for (auto i = cache.begin(); i != cache.end();) i = i->second.expired() ? cache.erase(i) : std::next(i);
  }
  cache[key] = p;
  return p;
}

// ---------------------------------------------------------------- device-only operators
namespace {
HostMaterialiser g_materialiser = nullptr;
}
void set_host_materialiser(HostMaterialiser f) { g_materialiser = f; }
void materialise_host(const std::vector<SparseOp*>& ops) {
  std::vector<SparseOp*> need;
  for (SparseOp* o : ops)
    if (o && !o->host_valid) need.push_back(o);
  if (need.empty()) return;
  if (!g_materialiser) throw std::runtime_error("device-only operators and no GPU layer to download them");
  g_materialiser(need);
}

// ---------------------------------------------------------------- build recording
namespace {
struct Recorder {
  std::vector<BuildItem> items;
  std::vector<OpPtr> keep;
};
}  // namespace
struct BuildSession::Impl {
  uint64_t id = 0;  // unique over the process (a later session may reuse this address)
  std::mutex mu;
  std::vector<std::unique_ptr<Recorder>> recs;
};
namespace {
std::atomic<uint64_t> g_session_ids{0};
// the session the calling thread records into, and its recorder there
// (identified by the session's id, never by its address)
thread_local BuildSession::Impl* tl_session = nullptr;
thread_local uint64_t tl_rec_id = 0;
thread_local Recorder* tl_rec = nullptr;
Recorder* recorder() {
  BuildSession::Impl* s = tl_session;
  if (!s) return nullptr;
  if (tl_rec_id != s->id) {
    auto r = std::make_unique<Recorder>();
    Recorder* raw = r.get();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->recs.push_back(std::move(r));
    }
    tl_rec = raw;
    tl_rec_id = s->id;
  }
  return tl_rec;
}
void keep_written(Recorder* r, const SparseOp& o) {
  // operators are owned by shared pointers; keep the written ones alive so
  // no later allocation can reuse an address an item refers to
  r->keep.push_back(std::const_pointer_cast<SparseOp>(o.shared_from_this()));
}
}  // namespace
BuildSession::BuildSession() : impl(new Impl) {
  if (tl_session) throw std::runtime_error("BuildSession: a session is already open on this thread");
  impl->id = ++g_session_ids;
  tl_session = impl.get();
}
BuildSession::~BuildSession() {
  if (tl_session == impl.get()) tl_session = nullptr;
}
std::vector<const BuildItem*> BuildSession::item_ptrs() const {
  std::vector<const BuildItem*> out;
  std::lock_guard<std::mutex> lk(impl->mu);
  size_t n = 0;
  for (const auto& r : impl->recs) n += r->items.size();
  out.reserve(n);
  for (const auto& r : impl->recs)
    for (const BuildItem& it : r->items) out.push_back(&it);
  return out;
}
std::vector<OpPtr> BuildSession::written() const {
  std::vector<OpPtr> out;
  std::lock_guard<std::mutex> lk(impl->mu);
  for (const auto& r : impl->recs) out.insert(out.end(), r->keep.begin(), r->keep.end());
  return out;
}
bool build_recording() { return tl_session != nullptr; }
BuildSession::Impl* current_build_session() { return tl_session; }
void enter_build_session(BuildSession::Impl* s) { tl_session = s; }

SparseOp::SparseOp(PatPtr p, bool f) : pat(std::move(p)), fermion(f) {
  if (tl_session)
    host_valid = false;  // built in HBM from the recorded items
  else
    data.assign(pat->size, 0.0);
}

void SparseOp::axpy(double a, const SparseOp& x) {
  if (x.pat.get() != pat.get()) throw std::runtime_error("SparseOp::axpy: pattern mismatch");
  if (Recorder* r = recorder()) {
    if (a != 0.0 && pat->size > 0) {
      if (pat->size > 0x7fffffffL) throw std::runtime_error("SparseOp::axpy: operator too large to record");
      r->items.push_back({this, &x, 0, 0, (int32_t)pat->size, (int32_t)pat->size, 1, (int32_t)pat->size, 0, 0, a, 1.0});
      keep_written(r, *this);
    }
    return;
  }
  if (!host_valid || !x.host_valid) throw std::runtime_error("SparseOp::axpy: operator not in host memory");
  const double* xs = x.data.data();
  double* ys = data.data();
  for (size_t i = 0; i < data.size(); ++i) ys[i] += a * xs[i];
}

double OpView::scale(int r, int c) const {
  if (!dag) return coef;
  const StateInfo& s = si();
  return coef * dagger_phase(s.q[r].s2, s.q[c].s2, op->k2());
}

// ---------------------------------------------------------------- tensor product
void tensor_product(const OpView& a, const OpView& b, SparseOp& c, double scale) {
  if (scale == 0.0) return;
  const StateInfo& cs = *c.pat->si;
  const StateInfo& ls = *cs.left;
  const StateInfo& rs = *cs.right;
  if (&a.si() != &ls || &b.si() != &rs) throw std::runtime_error("tensor_product: operand block mismatch");
  const int ka = a.k2(), kb = b.k2(), K = c.k2();
  if (a.dn() + b.dn() != c.dn() || irrep_mul(a.dirrep(), b.dirrep()) != c.dirrep() || !triangle(ka, kb, K))
    throw std::runtime_error("tensor_product: delta quanta do not combine");
  const auto& blocks = c.pat->blocks;
  if (recorder()) {
    // composite (block (x) dot) operators: every sub-block term is a scaled
    // copy of one block of a, the dot factor a 1 x 1 block of b
    if (!b.op->host_valid) throw std::runtime_error("tensor_product: dot operator not in host memory");
    Recorder* r = recorder();
    bool any = false;
    for (int ib = 0; ib < (int)blocks.size(); ++ib) {
      const auto& B = blocks[ib];
      const Quantum &Q = cs.q[B.r], &Qp = cs.q[B.c];
      for (const auto& s : cs.subs[B.r])
        for (const auto& sp : cs.subs[B.c]) {
          int ia = a.find(s.l, sp.l);
          if (ia < 0) continue;
          int ibb = b.find(s.r, sp.r);
          if (ibb < 0) continue;
          const Quantum &ql = ls.q[s.l], &qlp = ls.q[sp.l], &qr = rs.q[s.r], &qrp = rs.q[sp.r];
          double f = scale * wninej(qlp.s2, qrp.s2, Qp.s2, ka, kb, K, ql.s2, qr.s2, Q.s2);
          if (f == 0.0) continue;
          if (b.fermion() && (qlp.n & 1)) f = -f;
          f *= a.scale(s.l, sp.l) * b.scale(s.r, sp.r);
          const int dl = ls.dim[s.l], dlp = ls.dim[sp.l], dr = rs.dim[s.r], drp = rs.dim[sp.r];
          if (dr != 1 || drp != 1) throw std::runtime_error("tensor_product: recording needs a single-site second factor");
          const double bval = b.op->blk(ibb)[0];
          if (bval == 0.0 || dl == 0 || dlp == 0) continue;
          BuildItem it;
          it.dst = &c;
          it.src = a.op;
          it.dst_off = B.off + (int64_t)s.off * B.cols + sp.off;
          it.dst_ld = B.cols;
          it.src_off = a.op->pat->blocks[ia].off;
          it.src_ld = a.dag ? dl : dlp;  // the stored block is dlp x dl when a is the adjoint view
          it.rows = dl;
          it.cols = dlp;
          it.trans = a.dag ? 1 : 0;
          it.level = 1;
          it.alpha = f;
          it.beta = bval;
          r->items.push_back(it);
          any = true;
        }
    }
    if (any) keep_written(r, c);
    return;
  }
  if (!a.op->host_valid || !b.op->host_valid || !c.host_valid)
    throw std::runtime_error("tensor_product: operator not in host memory");
#pragma omp parallel for schedule(dynamic)
  for (int ib = 0; ib < (int)blocks.size(); ++ib) {
    const auto& B = blocks[ib];
    double* cd = c.data.data() + B.off;
    const int ccols = B.cols;
    const Quantum &Q = cs.q[B.r], &Qp = cs.q[B.c];
    for (const auto& s : cs.subs[B.r])
      for (const auto& sp : cs.subs[B.c]) {
        int ia = a.find(s.l, sp.l);
        if (ia < 0) continue;
        int ibb = b.find(s.r, sp.r);
        if (ibb < 0) continue;
        const Quantum &ql = ls.q[s.l], &qlp = ls.q[sp.l], &qr = rs.q[s.r], &qrp = rs.q[sp.r];
        double f = scale * wninej(qlp.s2, qrp.s2, Qp.s2, ka, kb, K, ql.s2, qr.s2, Q.s2);
        if (f == 0.0) continue;
        if (b.fermion() && (qlp.n & 1)) f = -f;
        f *= a.scale(s.l, sp.l) * b.scale(s.r, sp.r);
        const int dl = ls.dim[s.l], dlp = ls.dim[sp.l], dr = rs.dim[s.r], drp = rs.dim[sp.r];
        const double* ad = a.op->blk(ia);
        const double* bd = b.op->blk(ibb);
        for (int il = 0; il < dl; ++il)
          for (int ilp = 0; ilp < dlp; ++ilp) {
            double av = a.dag ? ad[(size_t)ilp * dl + il] : ad[(size_t)il * dlp + ilp];
            if (av == 0.0) continue;
            av *= f;
            for (int ir = 0; ir < dr; ++ir) {
              double* crow = cd + (size_t)(s.off + il * dr + ir) * ccols + sp.off + ilp * drp;
              if (b.dag) {
                for (int irp = 0; irp < drp; ++irp) crow[irp] += av * bd[(size_t)irp * dr + ir];
              } else {
                const double* brow = bd + (size_t)ir * drp;
                for (int irp = 0; irp < drp; ++irp) crow[irp] += av * brow[irp];
              }
            }
          }
      }
  }
}

// ---------------------------------------------------------------- dense kernels
static void gemm_nn(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb, double* C,
                    int ldc) {
  const int KB = 128, NB = 512;
#pragma omp parallel for schedule(static) if ((long)m * n * k > 2000000)
  for (int i = 0; i < m; ++i) {
    for (int p0 = 0; p0 < k; p0 += KB) {
      int p1 = std::min(k, p0 + KB);
      for (int j0 = 0; j0 < n; j0 += NB) {
        int j1 = std::min(n, j0 + NB);
        double* c = C + (size_t)i * ldc;
        for (int p = p0; p < p1; ++p) {
          double a = alpha * A[(size_t)i * lda + p];
          if (a == 0.0) continue;
          const double* b = B + (size_t)p * ldb;
          for (int j = j0; j < j1; ++j) c[j] += a * b[j];
        }
      }
    }
  }
}

void dgemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
           double beta, double* C, int ldc) {
  if (beta != 1.0)
    for (int i = 0; i < m; ++i) {
      double* c = C + (size_t)i * ldc;
      if (beta == 0.0)
        std::memset(c, 0, sizeof(double) * n);
      else
        for (int j = 0; j < n; ++j) c[j] *= beta;
    }
  if (m == 0 || n == 0 || k == 0 || alpha == 0.0) return;
  std::vector<double> at, bt;
  if (ta) {  // op(A) = A^T, A is k x m
    at.resize((size_t)m * k);
    for (int p = 0; p < k; ++p)
      for (int i = 0; i < m; ++i) at[(size_t)i * k + p] = A[(size_t)p * lda + i];
    A = at.data();
    lda = k;
  }
  if (tb) {  // op(B) = B^T, B is n x k
    bt.resize((size_t)k * n);
    for (int j = 0; j < n; ++j)
      for (int p = 0; p < k; ++p) bt[(size_t)p * n + j] = B[(size_t)j * ldb + p];
    B = bt.data();
    ldb = n;
  }
  gemm_nn(m, n, k, alpha, A, lda, B, ldb, C, ldc);
}

void syev(int n, double* a, double* w) {
  if (n == 0) return;
  if (n == 1) {
    w[0] = a[0];
    a[0] = 1.0;
    return;
  }
  std::vector<double> e(n, 0.0);
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
  // Householder reduction to tridiagonal form (tred2)
  for (int i = n - 1; i > 0; --i) {
    int l = i - 1;
    double h = 0.0, scale = 0.0;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += std::fabs(A(i, k));
      if (scale == 0.0) {
        e[i] = A(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          A(i, k) /= scale;
          h += A(i, k) * A(i, k);
        }
        double f = A(i, l);
        double g = (f >= 0.0) ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        A(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          A(j, i) = A(i, j) / h;
          g = 0.0;
          for (int k = 0; k <= j; ++k) g += A(j, k) * A(i, k);
          for (int k = j + 1; k <= l; ++k) g += A(k, j) * A(i, k);
          e[j] = g / h;
          f += e[j] * A(i, j);
        }
        double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          f = A(i, j);
          e[j] = g = e[j] - hh * f;
          for (int k = 0; k <= j; ++k) A(j, k) -= (f * e[k] + g * A(i, k));
        }
      }
    } else {
      e[i] = A(i, l);
    }
    w[i] = h;
  }
  w[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    int l = i - 1;
    if (w[i] != 0.0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0.0;
        for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
        for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
      }
    }
    w[i] = A(i, i);
    A(i, i) = 1.0;
    for (int j = 0; j <= l; ++j) A(j, i) = A(i, j) = 0.0;
  }
  // implicit QL on the tridiagonal matrix (tql2 / tqli)
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(w[m]) + std::fabs(w[m + 1]);
        if (std::fabs(e[m]) <= 1e-15 * dd) break;
      }
      if (m != l) {
        if (iter++ == 200) throw std::runtime_error("syev: QL iteration did not converge");
        double g = (w[l + 1] - w[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = w[m] - w[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = std::hypot(f, g));
          if (r == 0.0) {
            w[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = w[i + 1] - p;
          r = (w[i] - g) * s + 2.0 * c * b;
          w[i + 1] = g + (p = s * r);
          g = c * r - b;
          for (int k = 0; k < n; ++k) {
            f = A(k, i + 1);
            A(k, i + 1) = s * A(k, i) + c * f;
            A(k, i) = c * A(k, i) - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        w[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  // sort ascending (selection sort keeps it simple; n is modest)
  for (int i = 0; i < n - 1; ++i) {
    int k = i;
    for (int j = i + 1; j < n; ++j)
      if (w[j] < w[k]) k = j;
    if (k != i) {
      std::swap(w[i], w[k]);
      for (int r = 0; r < n; ++r) std::swap(A(r, i), A(r, k));
    }
  }
}

// ---------------------------------------------------------------- wavefunction layout
std::shared_ptr<const WaveLayout> make_wave_layout(const SIPtr& L, const SIPtr& R, const Quantum& target) {
  auto w = std::make_shared<WaveLayout>();
  w->L = L;
  w->R = R;
  w->target = target;
  w->index.assign((size_t)L->nsec() * R->nsec(), -1);
  long off = 0;
  for (int l = 0; l < L->nsec(); ++l)
    for (int r = 0; r < R->nsec(); ++r) {
      const Quantum &ql = L->q[l], &qr = R->q[r];
      if (ql.n + qr.n != target.n) continue;
      if (!triangle(ql.s2, qr.s2, target.s2)) continue;
      if (irrep_mul(ql.irrep, qr.irrep) != target.irrep) continue;
      w->index[(size_t)l * R->nsec() + r] = (int)w->blocks.size();
      w->blocks.push_back({l, r, L->dim[l], R->dim[r], off});
      off += (long)L->dim[l] * R->dim[r];
    }
  w->size = off;
  return w;
}

}  // namespace blk
