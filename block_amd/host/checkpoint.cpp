#include "checkpoint.h"

#include <cstdint>
#include <cstring>
#include <fstream>
#include <unordered_map>

namespace blk {

namespace {
constexpr uint64_t kMagic = 0x31544b4843475342ull;  // format tag of this file type, version 1

struct Out {
  std::ofstream f;
  template <class T>
  void pod(const T& x) {
    f.write(reinterpret_cast<const char*>(&x), sizeof(T));
  }
  template <class T>
  void vec(const std::vector<T>& v) {
    pod<uint64_t>(v.size());
    if (!v.empty()) f.write(reinterpret_cast<const char*>(v.data()), sizeof(T) * v.size());
  }
  void str(const std::string& s) {
    pod<uint64_t>(s.size());
    f.write(s.data(), (std::streamsize)s.size());
  }
};
struct In {
  std::ifstream f;
  template <class T>
  T pod() {
    T x;
    if (!f.read(reinterpret_cast<char*>(&x), sizeof(T))) throw std::runtime_error("checkpoint: truncated file");
    return x;
  }
  template <class T>
  std::vector<T> vec() {
    const uint64_t n = pod<uint64_t>();
    if (n > (1ull << 40)) throw std::runtime_error("checkpoint: corrupt length");
    std::vector<T> v(n);
    if (n && !f.read(reinterpret_cast<char*>(v.data()), sizeof(T) * n))
      throw std::runtime_error("checkpoint: truncated file");
    return v;
  }
  std::string str() {
    const uint64_t n = pod<uint64_t>();
    if (n > (1ull << 32)) throw std::runtime_error("checkpoint: corrupt length");
    std::string s(n, '\0');
    if (n && !f.read(&s[0], (std::streamsize)n)) throw std::runtime_error("checkpoint: truncated file");
    return s;
  }
};

// Shared objects (state infos, operators) are written once, into tables, and
// referred to by index; children precede parents.
struct Writer {
  Out& o;
  std::unordered_map<const StateInfo*, int> si_id;
  std::vector<const StateInfo*> sis;
  std::unordered_map<const SparseOp*, int> op_id;
  std::vector<const SparseOp*> ops;
  int si(const SIPtr& p) {
    if (!p) return -1;
    auto f = si_id.find(p.get());
    if (f != si_id.end()) return f->second;
    si(p->left);
    si(p->right);
    const int id = (int)sis.size();
    sis.push_back(p.get());
    si_id[p.get()] = id;
    return id;
  }
  int op(const OpPtr& p) {
    if (!p) return -1;
    auto f = op_id.find(p.get());
    if (f != op_id.end()) return f->second;
    si(p->pat->si);
    const int id = (int)ops.size();
    ops.push_back(p.get());
    op_id[p.get()] = id;
    return id;
  }
};

void collect(Writer& w, const BlockPtr& b) {
  if (!b) return;
  w.si(b->si);
  w.op(b->ham);
  w.op(b->ident);
  for (auto* l : {&b->cre, &b->ccd, &b->cd, &b->cc, &b->cdc, &b->ddc})
    for (const auto& o : *l) w.op(o);
  for (const auto& kv : b->extra) w.op(kv.second);
  if (b->rotrec) {
    w.si(b->rotrec->composite);
    w.si(b->rotrec->kept);
  }
}

void write_block(Out& o, Writer& w, const BlockPtr& b) {
  o.pod<uint8_t>(b ? 1 : 0);
  if (!b) return;
  o.pod<int32_t>(b->norb);
  o.vec(b->sites);
  o.vec(b->inblock);
  o.pod<int32_t>(w.si_id.at(b->si.get()));
  o.pod<int32_t>(w.op(b->ham));
  o.pod<int32_t>(w.op(b->ident));
  for (auto* l : {&b->cre, &b->ccd, &b->cd, &b->cc, &b->cdc, &b->ddc}) {
    o.pod<uint64_t>(l->size());
    std::vector<int64_t> ent;  // (position, operator) pairs of the present entries
    for (size_t i = 0; i < l->size(); ++i)
      if ((*l)[i]) {
        ent.push_back((int64_t)i);
        ent.push_back(w.op((*l)[i]));
      }
    o.vec(ent);
  }
  o.pod<uint8_t>(b->has_normal);
  o.pod<uint8_t>(b->has_comp);
  o.pod<uint64_t>(b->extra.size());
  for (const auto& kv : b->extra) {
    o.str(kv.first);
    o.pod<int32_t>(w.op(kv.second));
  }
  o.pod<uint8_t>(b->rotrec ? 1 : 0);
  if (b->rotrec) {
    const RotRecord& r = *b->rotrec;
    o.pod<int32_t>(w.si_id.at(r.composite.get()));
    o.pod<int32_t>(w.si_id.at(r.kept.get()));
    o.vec(r.oldidx);
    o.pod<uint64_t>(r.rot.size());
    for (const auto& m : r.rot) o.vec(m);
  }
}

BlockPtr read_block(In& in, const std::vector<SIPtr>& sis, const std::vector<OpPtr>& ops) {
  if (!in.pod<uint8_t>()) return nullptr;
  auto at_si = [&](int32_t i) -> SIPtr {
    if (i < 0 || i >= (int32_t)sis.size()) throw std::runtime_error("checkpoint: bad state-info index");
    return sis[i];
  };
  auto at_op = [&](int64_t i) -> OpPtr {
    if (i < 0) return nullptr;
    if (i >= (int64_t)ops.size()) throw std::runtime_error("checkpoint: bad operator index");
    return ops[i];
  };
  auto b = std::make_shared<SpinBlock>();
  b->norb = in.pod<int32_t>();
  b->sites = in.vec<int>();
  b->inblock = in.vec<char>();
  b->si = at_si(in.pod<int32_t>());
  b->ham = at_op(in.pod<int32_t>());
  b->ident = at_op(in.pod<int32_t>());
  for (auto* l : {&b->cre, &b->ccd, &b->cd, &b->cc, &b->cdc, &b->ddc}) {
    l->assign(in.pod<uint64_t>(), nullptr);
    const auto ent = in.vec<int64_t>();
    for (size_t e = 0; e + 1 < ent.size(); e += 2) {
      if (ent[e] < 0 || ent[e] >= (int64_t)l->size()) throw std::runtime_error("checkpoint: bad operator slot");
      (*l)[ent[e]] = at_op(ent[e + 1]);
    }
  }
  b->has_normal = in.pod<uint8_t>() != 0;
  b->has_comp = in.pod<uint8_t>() != 0;
  const uint64_t ne = in.pod<uint64_t>();
  for (uint64_t e = 0; e < ne; ++e) {
    std::string k = in.str();
    b->extra[k] = at_op(in.pod<int32_t>());
  }
  if (in.pod<uint8_t>()) {
    auto r = std::make_shared<RotRecord>();
    r->composite = at_si(in.pod<int32_t>());
    r->kept = at_si(in.pod<int32_t>());
    r->oldidx = in.vec<int>();
    r->rot.resize(in.pod<uint64_t>());
    for (auto& m : r->rot) m = in.vec<double>();
    b->rotrec = r;
  }
  return b;
}
}  // namespace

void save_sweep_state(const std::string& path, const SweepState& s) {
  Out o;
  const std::string tmp = path + ".tmp";
  o.f.open(tmp, std::ios::binary | std::ios::trunc);
  if (!o.f) throw std::runtime_error("checkpoint: cannot write " + tmp);
  Writer w{o};
  for (const auto& b : s.fwd) collect(w, b);
  for (const auto& b : s.bwd) collect(w, b);
  if (s.wave) {
    w.si(s.wave->L);
    w.si(s.wave->R);
  }
  o.pod(kMagic);
  // state infos
  o.pod<uint64_t>(w.sis.size());
  for (const StateInfo* p : w.sis) {
    std::vector<int32_t> qs;
    for (const auto& q : p->q) {
      qs.push_back(q.n);
      qs.push_back(q.s2);
      qs.push_back(q.irrep);
    }
    o.vec(qs);
    o.vec(p->dim);
    o.pod<int32_t>(p->left ? w.si_id.at(p->left.get()) : -1);
    o.pod<int32_t>(p->right ? w.si_id.at(p->right.get()) : -1);
    o.pod<uint64_t>(p->subs.size());
    for (const auto& v : p->subs) {
      std::vector<int32_t> t;
      for (const auto& x : v) {
        t.push_back(x.l);
        t.push_back(x.r);
        t.push_back(x.off);
      }
      o.vec(t);
    }
    o.pod<uint64_t>(p->prod_to_sector.size());
    for (const auto& v : p->prod_to_sector) {
      std::vector<int32_t> t;
      for (const auto& x : v) {
        t.push_back(x.first);
        t.push_back(x.second);
      }
      o.vec(t);
    }
  }
  // operators: pattern key and data (device-only ones downloaded first)
  {
    std::vector<SparseOp*> dev;
    for (const SparseOp* p : w.ops)
      if (!p->host_valid) dev.push_back(const_cast<SparseOp*>(p));
    materialise_host(dev);
  }
  o.pod<uint64_t>(w.ops.size());
  for (const SparseOp* p : w.ops) {
    o.pod<int32_t>(w.si_id.at(p->pat->si.get()));
    o.pod<int32_t>(p->pat->dn);
    o.pod<int32_t>(p->pat->k2);
    o.pod<int32_t>(p->pat->dirrep);
    o.pod<uint8_t>(p->fermion);
    o.vec(p->data);
  }
  // blocks
  o.pod<uint64_t>(s.fwd.size());
  for (const auto& b : s.fwd) write_block(o, w, b);
  o.pod<uint64_t>(s.bwd.size());
  for (const auto& b : s.bwd) write_block(o, w, b);
  // last superblock and its wavefunctions
  o.pod<uint8_t>(s.wave ? 1 : 0);
  if (s.wave) {
    o.pod<int32_t>(w.si_id.at(s.wave->L.get()));
    o.pod<int32_t>(w.si_id.at(s.wave->R.get()));
    o.pod<int32_t>(s.wave->target.n);
    o.pod<int32_t>(s.wave->target.s2);
    o.pod<int32_t>(s.wave->target.irrep);
  }
  o.pod<uint64_t>(s.b.size());
  for (const auto& v : s.b) o.vec(v);
  // sweep records
  o.pod<int32_t>(s.next_sweep);
  o.pod<double>(s.prev_full);
  o.pod<uint8_t>(s.res.converged);
  o.vec(s.res.energies);
  o.pod<uint64_t>(s.res.sweeps.size());
  for (const auto& r : s.res.sweeps) {
    o.pod<int32_t>(r.sweep);
    o.pod<uint8_t>(r.forward);
    o.pod<int32_t>(r.M);
    o.vec(r.energies);
    const double d[] = {r.max_discarded, r.seconds, r.sigma_seconds, r.sigma_flops, r.build_seconds,
                        r.davidson_seconds, r.renorm_seconds, r.setup_seconds, r.diag_seconds, r.guess_seconds};
    for (double x : d) o.pod(x);
    const int64_t c[] = {r.sigma_calls, r.sigma_calls_split, r.guessed_steps};
    for (int64_t x : c) o.pod(x);
  }
  o.str(s.rng);
  o.pod(kMagic);
  o.f.close();
  if (!o.f) throw std::runtime_error("checkpoint: write failed for " + tmp);
  if (std::rename(tmp.c_str(), path.c_str()) != 0) throw std::runtime_error("checkpoint: cannot replace " + path);
}

SweepState load_sweep_state(const std::string& path) {
  In in;
  in.f.open(path, std::ios::binary);
  if (!in.f) throw std::runtime_error("checkpoint: cannot read " + path);
  if (in.pod<uint64_t>() != kMagic) throw std::runtime_error("checkpoint: not a sweep-state file: " + path);
  std::vector<SIPtr> sis(in.pod<uint64_t>());
  for (size_t i = 0; i < sis.size(); ++i) {
    auto p = std::make_shared<StateInfo>();
    const auto qs = in.vec<int32_t>();
    if (qs.size() % 3) throw std::runtime_error("checkpoint: corrupt quanta");
    for (size_t j = 0; j + 3 <= qs.size(); j += 3) p->q.push_back({qs[j], qs[j + 1], qs[j + 2]});
    p->dim = in.vec<int>();
    if (p->dim.size() != p->q.size()) throw std::runtime_error("checkpoint: corrupt state info");
    const int32_t l = in.pod<int32_t>(), r = in.pod<int32_t>();
    if (l >= (int32_t)i || r >= (int32_t)i) throw std::runtime_error("checkpoint: bad state-info order");
    if (l >= 0) p->left = sis[l];
    if (r >= 0) p->right = sis[r];
    p->subs.resize(in.pod<uint64_t>());
    for (auto& v : p->subs) {
      const auto t = in.vec<int32_t>();
      if (t.size() % 3) throw std::runtime_error("checkpoint: corrupt sub-blocks");
      for (size_t j = 0; j + 3 <= t.size(); j += 3) v.push_back({t[j], t[j + 1], t[j + 2]});
    }
    p->prod_to_sector.resize(in.pod<uint64_t>());
    for (auto& v : p->prod_to_sector) {
      const auto t = in.vec<int32_t>();
      for (size_t j = 0; j + 1 < t.size(); j += 2) v.push_back({t[j], t[j + 1]});
    }
    p->finalize();
    sis[i] = p;
  }
  std::vector<OpPtr> ops(in.pod<uint64_t>());
  for (auto& op : ops) {
    const int32_t si = in.pod<int32_t>();
    if (si < 0 || si >= (int32_t)sis.size()) throw std::runtime_error("checkpoint: bad operator state info");
    const int32_t dn = in.pod<int32_t>(), k2 = in.pod<int32_t>(), dirrep = in.pod<int32_t>();
    const bool fermion = in.pod<uint8_t>() != 0;
    op = std::make_shared<SparseOp>(get_pattern(sis[si], dn, k2, dirrep), fermion);
    auto data = in.vec<double>();
    if (data.size() != op->data.size()) throw std::runtime_error("checkpoint: operator size does not match its pattern");
    op->data.swap(data);
  }
  SweepState s;
  s.fwd.resize(in.pod<uint64_t>());
  for (auto& b : s.fwd) b = read_block(in, sis, ops);
  s.bwd.resize(in.pod<uint64_t>());
  for (auto& b : s.bwd) b = read_block(in, sis, ops);
  if (in.pod<uint8_t>()) {
    const int32_t L = in.pod<int32_t>(), R = in.pod<int32_t>();
    if (L < 0 || R < 0 || L >= (int32_t)sis.size() || R >= (int32_t)sis.size())
      throw std::runtime_error("checkpoint: bad superblock");
    Quantum t;
    t.n = in.pod<int32_t>();
    t.s2 = in.pod<int32_t>();
    t.irrep = in.pod<int32_t>();
    s.wave = make_wave_layout(sis[L], sis[R], t);
  }
  s.b.resize(in.pod<uint64_t>());
  for (auto& v : s.b) {
    v = in.vec<double>();
    if (s.wave && (long)v.size() != s.wave->size) throw std::runtime_error("checkpoint: wavefunction size mismatch");
  }
  s.next_sweep = in.pod<int32_t>();
  s.prev_full = in.pod<double>();
  s.res.converged = in.pod<uint8_t>() != 0;
  s.res.energies = in.vec<double>();
  s.res.sweeps.resize(in.pod<uint64_t>());
  for (auto& r : s.res.sweeps) {
    r.sweep = in.pod<int32_t>();
    r.forward = in.pod<uint8_t>() != 0;
    r.M = in.pod<int32_t>();
    r.energies = in.vec<double>();
    double* d[] = {&r.max_discarded, &r.seconds, &r.sigma_seconds, &r.sigma_flops, &r.build_seconds,
                   &r.davidson_seconds, &r.renorm_seconds, &r.setup_seconds, &r.diag_seconds, &r.guess_seconds};
    for (double* x : d) *x = in.pod<double>();
    r.sigma_calls = in.pod<int64_t>();
    r.sigma_calls_split = in.pod<int64_t>();
    r.guessed_steps = in.pod<int64_t>();
  }
  s.rng = in.str();
  if (in.pod<uint64_t>() != kMagic) throw std::runtime_error("checkpoint: trailer missing in " + path);
  return s;
}

}  // namespace blk
