#include "gpu_sigma.h"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace blk {

void bsg_check(int rc, const char* what) {
  if (rc != BSG_OK) throw std::runtime_error(std::string(what) + ": " + bsg_last_error());
}

GpuContext::GpuContext(int dev) : device(dev) { bsg_check(bsg_init(&ctx, dev), "bsg_init"); }
GpuContext::~GpuContext() {
  if (!ctx) return;
  for (auto& p : pre)
    if (p.buf && side) bsg_mfree(side->ctx, p.buf);
  side.reset();
  if (arena) bsg_mfree(ctx, arena);
  if (work) bsg_mfree(ctx, work);
  if (staging) bsg_host_free(ctx, staging);
  bsg_free(ctx);
}
double* GpuContext::ensure_staging(size_t n) {
  if (staging_cap < n) {
    if (staging) bsg_check(bsg_host_free(ctx, staging), "staging free");
    staging = nullptr;
    staging_cap = 0;
    const size_t cap = n + n / 4;
    bsg_check(bsg_host_alloc(ctx, sizeof(double) * cap, (void**)&staging), "staging");
    staging_cap = cap;
  }
  return staging;
}
double* GpuContext::ensure_work(size_t n) {
  if (work_cap < n) {
    if (work) bsg_check(bsg_mfree(ctx, work), "work free");
    work = nullptr;
    work_cap = 0;
    const size_t cap = n + n / 4;
    bsg_check(bsg_malloc(ctx, sizeof(double) * cap, (void**)&work), "rotation scratch");
    work_cap = cap;
  }
  return work;
}

// Upload operators to dst + offset: copied into page-locked staging by all
// threads in batches of about 64 MB, each batch sent while the next one is
// copied; device memory between the operators is left as it is.  n: doubles
// spanned by the offsets.
void upload_ops(GpuContext& g, const std::vector<std::pair<const SparseOp*, int64_t>>& ops, double* dst, size_t n) {
  // device-resident and preloaded operators: device copies, in one launch
  std::vector<std::pair<const SparseOp*, int64_t>> rest;
  std::vector<bsg_copy_seg_t> segs;
  for (const auto& kv : ops) {
    if (kv.first->dev) {
      segs.push_back({kv.first->dev, dst + kv.second, (int64_t)kv.first->size()});
      continue;
    }
    if (!kv.first->host_valid) throw std::runtime_error("upload_ops: operator in neither host nor device memory");
    bool done = false;
    for (auto& p : g.pre) {
      auto f = p.at.find(kv.first);
      if (f == p.at.end()) continue;
      segs.push_back({p.buf + f->second, dst + kv.second, (int64_t)kv.first->size()});
      p.at.erase(f);
      done = true;
      break;
    }
    if (!done) rest.push_back(kv);
  }
  bsg_check(bsg_copy_segments(g.ctx, segs.data(), (int32_t)segs.size()), "preloaded operator copy");
  for (auto& p : g.pre)
    if (p.at.empty()) p.held.clear();
  double* staging = g.ensure_staging(std::max<size_t>(1, n));
  struct Piece {
    const double* src;
    int64_t off, n;
  };
  std::vector<Piece> pieces;
  constexpr int64_t PIECE = 1 << 19;  // doubles (4 MB)
  for (const auto& kv : rest)
    for (int64_t x = 0; x < (int64_t)kv.first->data.size(); x += PIECE)
      pieces.push_back({kv.first->data.data() + x, kv.second + x, std::min<int64_t>(PIECE, kv.first->data.size() - x)});
  std::sort(pieces.begin(), pieces.end(), [](const Piece& a, const Piece& b) { return a.off < b.off; });
  constexpr int64_t BATCH = 1 << 23;  // doubles (64 MB)
  for (size_t p0 = 0; p0 < pieces.size();) {
    // a batch is a gap-free run of pieces: the range sent is exactly theirs
    size_t p1 = p0;
    int64_t len = 0;
    while (p1 < pieces.size() && (len == 0 || (len + pieces[p1].n <= BATCH && pieces[p1].off == pieces[p0].off + len)))
      len += pieces[p1++].n;
#pragma omp parallel for schedule(dynamic)
    for (long i = (long)p0; i < (long)p1; ++i)
      std::memcpy(staging + pieces[i].off, pieces[i].src, sizeof(double) * pieces[i].n);
    const int64_t b0 = pieces[p0].off, b1 = pieces[p1 - 1].off + pieces[p1 - 1].n;
    bsg_check(bsg_memcpy_h2d_async(g.ctx, dst + b0, staging + b0, sizeof(double) * (b1 - b0)), "operator upload");
    p0 = p1;
  }
  bsg_check(bsg_synchronize(g.ctx), "operator upload");
}

namespace {
SetupTimes g_setup;
double secs() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }
}  // namespace
const SetupTimes& gpu_setup_times() { return g_setup; }

WorkList compile_worklist(const SuperBlock& sb) {
  WorkList wl;
  const WaveLayout& W = *sb.wave;
  const StateInfo &ls = *W.L, &rs = *W.R;
  const int St = W.target.s2;
  const SparseOp* idL = sb.L->ident.get();
  const SparseOp* idR = sb.R->ident.get();

  // operator arena: every distinct operator the term list references, the
  // L operators first, then the R ones (each half one contiguous run, so an
  // upload that skips one half's preloaded operators stays in large batches)
  for (int side = 0; side < 2; ++side)
    for (const auto& t : sb.terms) {
      const SparseOp* o = side ? t.b.op : t.a.op;
      if (!wl.arena_off.count(o)) {
        wl.arena_off[o] = wl.arena_size;
        wl.arena_size += (int64_t)o->size();
      }
    }

  // pairs per V block (l, r), in term order and, within a term, in C block
  // order; V blocks are independent, so they are built in parallel
  std::vector<std::vector<int>> cblocks(ls.nsec());  // C blocks by left sector
  for (size_t ic = 0; ic < W.blocks.size(); ++ic) cblocks[W.blocks[ic].l].push_back((int)ic);
  std::vector<int64_t> toff_a(sb.terms.size()), toff_b(sb.terms.size());
  for (size_t it = 0; it < sb.terms.size(); ++it) {
    toff_a[it] = wl.arena_off.at(sb.terms[it].a.op);
    toff_b[it] = wl.arena_off.at(sb.terms[it].b.op);
  }
  // each term's a-view blocks by view row l: (l', block id), l' ascending
  // (pattern blocks are stored in (row, col) order), so a V block visits
  // only the allowed (l, l') instead of every l'
  const int nls = ls.nsec();
  std::vector<std::vector<int>> arow_start(sb.terms.size());
  std::vector<std::vector<std::pair<int, int>>> arow(sb.terms.size());
#pragma omp parallel for schedule(dynamic)
  for (long it = 0; it < (long)sb.terms.size(); ++it) {
    const OpView& a = sb.terms[it].a;
    const auto& blks = a.op->pat->blocks;
    std::vector<int>& st = arow_start[it];
    st.assign(nls + 1, 0);
    for (const auto& bk : blks) ++st[(a.dag ? bk.c : bk.r) + 1];
    for (int q = 0; q < nls; ++q) st[q + 1] += st[q];
    std::vector<int> pos(st.begin(), st.end() - 1);
    arow[it].resize(blks.size());
    for (size_t ib = 0; ib < blks.size(); ++ib) {
      const int row = a.dag ? blks[ib].c : blks[ib].r, col = a.dag ? blks[ib].r : blks[ib].c;
      arow[it][pos[row]++] = {col, (int)ib};
    }
  }
  std::vector<std::vector<bsg_pair_t>> per_v(W.blocks.size());
  std::vector<std::vector<WorkList::Prov>> per_p(W.blocks.size());
#pragma omp parallel for schedule(dynamic)
  for (long iv = 0; iv < (long)W.blocks.size(); ++iv) {
    const int l = W.blocks[iv].l, r = W.blocks[iv].r;
    for (size_t it = 0; it < sb.terms.size(); ++it) {
      const auto& t = sb.terms[it];
      const bool aid = (t.a.op == idL), bid = (t.b.op == idR);
      for (int e = arow_start[it][l]; e < arow_start[it][l + 1]; ++e) {
        const int lp = arow[it][e].first, ia = arow[it][e].second;
        for (int ic : cblocks[lp]) {
          const auto& cb = W.blocks[ic];
          const int rp = cb.r;
          const int ib = t.b.find(r, rp);
          if (ib < 0) continue;
          double f = t.coef * wninej(ls.q[lp].s2, rs.q[rp].s2, St, t.a.k2(), t.b.k2(), 0, ls.q[l].s2, rs.q[r].s2, St);
          if (t.b.fermion() && (ls.q[lp].n & 1)) f = -f;
          f *= t.b.scale(r, rp);
          if (f == 0.0) continue;
          bsg_pair_t p;
          p.a_off = aid ? -1 : toff_a[it] + t.a.op->pat->blocks[ia].off;
          p.c_off = cb.off;
          p.b_off = bid ? -1 : toff_b[it] + t.b.op->pat->blocks[ib].off;
          p.k1 = ls.dim[lp];
          p.k2 = rs.dim[rp];
          p.a_trans = t.a.dag ? 1 : 0;
          p.b_trans = t.b.dag ? 0 : 1;
          p.alpha_a = aid ? 1.0 : t.a.scale(l, lp);
          p.alpha = f;
          per_v[iv].push_back(p);
          per_p[iv].push_back({(int)it, l, lp, r, rp});
        }
      }
    }
  }
  // every operator block a pair names must lie inside the arena
  for (size_t iv = 0; iv < per_v.size(); ++iv)
    for (const auto& p : per_v[iv]) {
      const int64_t m = W.blocks[iv].rows, n = W.blocks[iv].cols;
      if ((p.a_off >= 0 && p.a_off + m * p.k1 > wl.arena_size) || (p.b_off >= 0 && p.b_off + n * p.k2 > wl.arena_size))
        throw std::runtime_error("compile_worklist: operator block outside the arena");
    }
  for (size_t iv = 0; iv < per_v.size(); ++iv) {
    if (per_v[iv].empty()) continue;
    const auto& b = W.blocks[iv];
    wl.vb.push_back({b.off, b.rows, b.cols, (int32_t)wl.pairs.size(), (int32_t)per_v[iv].size()});
    wl.pairs.insert(wl.pairs.end(), per_v[iv].begin(), per_v[iv].end());
    wl.prov.insert(wl.prov.end(), per_p[iv].begin(), per_p[iv].end());
  }

  // diagonal terms (SpinBlock::diagonalH)
  for (const auto& t : sb.terms) {
    if (t.a.dn() != 0 || t.b.dn() != 0 || t.a.dirrep() != 0 || t.b.dirrep() != 0) continue;
    const int k = t.a.k2();
    for (const auto& bk : W.blocks) {
      int ia = t.a.find(bk.l, bk.l), ib = t.b.find(bk.r, bk.r);
      if (ia < 0 || ib < 0) continue;
      const Quantum &ql = ls.q[bk.l], &qr = rs.q[bk.r];
      double f = t.coef * wninej(ql.s2, qr.s2, St, k, k, 0, ql.s2, qr.s2, St);
      if (t.b.fermion() && (ql.n & 1)) f = -f;
      f *= t.a.scale(bk.l, bk.l) * t.b.scale(bk.r, bk.r);
      if (f == 0.0) continue;
      bsg_diag_t d;
      d.v_off = bk.off;
      d.a_off = wl.arena_off[t.a.op] + t.a.op->pat->blocks[ia].off;
      d.b_off = wl.arena_off[t.b.op] + t.b.op->pat->blocks[ib].off;
      d.m = bk.rows;
      d.n = bk.cols;
      d.alpha = f;
      wl.diag.push_back(d);
    }
  }
  return wl;
}

GpuSigma::GpuSigma(GpuContext& g, const SuperBlock& sb, int max_vec) : g_(g) {
  double t0 = secs();
  wave_size_ = sb.wave->size;
  WorkList wl = compile_worklist(sb);
  n_pairs_ = wl.pairs.size();
  arena_size_ = wl.arena_size;
  diag_terms_ = std::move(wl.diag);
  double t1 = secs();
  const size_t need = (size_t)std::max<int64_t>(1, arena_size_);
  if (!g_.arena_leased) {
    if (g_.arena_cap < need) {
      if (g_.arena) bsg_check(bsg_mfree(g_.ctx, g_.arena), "arena free");
      g_.arena = nullptr;
      g_.arena_cap = 0;
      const size_t cap = need + need / 4;  // headroom for the next, larger superblocks
      bsg_check(bsg_malloc(g_.ctx, sizeof(double) * cap, (void**)&g_.arena), "arena");
      g_.arena_cap = cap;
    }
    d_arena_ = g_.arena;
    g_.arena_leased = true;
  } else {
    bsg_check(bsg_malloc(g_.ctx, sizeof(double) * need, (void**)&d_arena_), "arena");
    owns_arena_ = true;
  }
  std::vector<std::pair<const SparseOp*, int64_t>> ops(wl.arena_off.begin(), wl.arena_off.end());
  g_.resident.clear();
  upload_ops(g_, ops, d_arena_, (size_t)arena_size_);
  if (!owns_arena_) g_.resident = wl.arena_off;
  double t2 = secs();
  max_vec_ = std::max(1, max_vec);
  size_t wbytes = sizeof(double) * std::max<long>(1, wave_size_) * max_vec_;
  bsg_check(bsg_malloc(g_.ctx, wbytes, (void**)&d_c_), "c buffer");
  bsg_check(bsg_malloc(g_.ctx, wbytes, (void**)&d_v_), "v buffer");
  vb_ = std::move(wl.vb);
  pairs_ = std::move(wl.pairs);
  replan();
  flops_ = bsg_plan_flops(plan_);
  double t3 = secs();
  g_setup.compile += t1 - t0;
  g_setup.upload += t2 - t1;
  g_setup.plan += t3 - t2;
}

GpuSigma::~GpuSigma() {
  if (plan_) bsg_plan_destroy(plan_);
  if (d_arena_) {
    if (owns_arena_)
      bsg_mfree(g_.ctx, d_arena_);
    else
      g_.arena_leased = false;
  }
  if (d_c_) bsg_mfree(g_.ctx, d_c_);
  if (d_v_) bsg_mfree(g_.ctx, d_v_);
}

void GpuSigma::set_precision(int mode) {
  if (mode != BSG_FP64 && mode != BSG_SPLIT_BF16) throw std::runtime_error("GpuSigma: unknown precision mode");
  if (mode == BSG_SPLIT_BF16 && !split_ready_) {
    bsg_check(bsg_plan_split_arena(g_.ctx, plan_, d_arena_, arena_size_), "split arena");
    split_ready_ = true;
  }
  precision_ = mode;
}

void GpuSigma::apply_device(const double* dc, double* dv, int nvec) {
  if (!use_graph_ || nvec > max_vec_) {
    bsg_check(bsg_sigma_apply(g_.ctx, plan_, d_arena_, dc, dv, nvec, precision_), "sigma");
    return;
  }
  // one graph launch on the handle's own c and v buffers (captured once per
  // precision and vector count): c copied in, v cleared outside the graph,
  // the result added to dv
  const int64_t n = (int64_t)wave_size_ * nvec;
  if (graph_precision_ != precision_ || graph_nvec_ != nvec) {
    bsg_check(bsg_plan_capture(g_.ctx, plan_, d_arena_, d_c_, d_v_, nvec, precision_), "sigma capture");
    graph_precision_ = precision_;
    graph_nvec_ = nvec;
  }
  const bool own = dc == d_c_ && dv == d_v_;  // already the captured buffers (v += H c in place)
  if (!own) {
    bsg_check(bsg_copy(g_.ctx, dc, d_c_, n), "sigma input");
    bsg_check(bsg_memset_zero(g_.ctx, d_v_, sizeof(double) * n), "sigma clear");
  }
  bsg_check(bsg_plan_replay(g_.ctx, plan_), "sigma replay");
  if (!own) bsg_check(bsg_axpy(g_.ctx, 1.0, d_v_, dv, n), "sigma output");
}

void GpuSigma::set_graph(bool on) {
  use_graph_ = on;
  graph_precision_ = -1;
  graph_nvec_ = 0;
}

void GpuSigma::replan() {
  if (plan_) bsg_check(bsg_plan_destroy(plan_), "plan destroy");
  plan_ = nullptr;
  bsg_chunk_t chunk;
  chunk.vblocks = vb_.data();
  chunk.n_vblocks = (int32_t)vb_.size();
  chunk.pairs = pairs_.data();
  chunk.n_pairs = (int32_t)pairs_.size();
  bsg_check(bsg_plan_create(g_.ctx, &chunk, 1, wave_size_, &plan_), "plan");
  batch_ = 1;
  split_ready_ = false;
  graph_precision_ = -1;
  graph_nvec_ = 0;
  if (max_vec_ > 1) set_batch(max_vec_);
  if (precision_ == BSG_SPLIT_BF16) set_precision(BSG_SPLIT_BF16);
}

void GpuSigma::set_batch(int nroot) {
  nroot = std::max(1, nroot);
  if (nroot == batch_) return;
  bsg_check(bsg_plan_set_batch(g_.ctx, plan_, nroot), "batch");
  batch_ = nroot;
  graph_precision_ = -1;  // the plan dropped its capture
  graph_nvec_ = 0;
}

void GpuSigma::apply(const double* c_host, double* v_host, int nvec) {
  if (nvec < 1 || nvec > max_vec_) throw std::runtime_error("GpuSigma::apply: more vectors than the buffers hold");
  size_t bytes = sizeof(double) * wave_size_ * nvec;
  bsg_check(bsg_memcpy_h2d(g_.ctx, d_c_, c_host, bytes), "c upload");
  bsg_check(bsg_memset_zero(g_.ctx, d_v_, bytes), "v clear");
  apply_device(d_c_, d_v_, nvec);
  bsg_check(bsg_memcpy_d2h(g_.ctx, v_host, d_v_, bytes), "v download");
}

std::vector<double> GpuSigma::diagonal() {
  size_t bytes = sizeof(double) * wave_size_;
  bsg_check(bsg_memset_zero(g_.ctx, d_v_, bytes), "diag clear");
  bsg_check(bsg_diag_apply(g_.ctx, diag_terms_.data(), (int32_t)diag_terms_.size(), d_arena_, d_v_), "diag");
  std::vector<double> out(wave_size_);
  bsg_check(bsg_memcpy_d2h(g_.ctx, out.data(), d_v_, bytes), "diag download");
  return out;
}

void preload_ops(GpuContext& g, int slot, const std::vector<OpPtr>& ops) {
  GpuContext::Preload& p = g.pre[slot];
  p.at.clear();
  p.held = ops;
  if (!g.side) g.side.reset(new GpuContext(g.device));
  std::vector<std::pair<const SparseOp*, int64_t>> at;
  int64_t n = 0;
  for (const OpPtr& o : ops) {
    at.push_back({o.get(), n});
    n += (int64_t)o->size();
  }
  if (p.cap < (size_t)n) {
    if (p.buf) bsg_check(bsg_mfree(g.side->ctx, p.buf), "preload free");
    p.buf = nullptr;
    p.cap = 0;
    const size_t cap = (size_t)n + (size_t)n / 4;
    bsg_check(bsg_malloc(g.side->ctx, sizeof(double) * cap, (void**)&p.buf), "preload buffer");
    p.cap = cap;
  }
  upload_ops(*g.side, at, p.buf, (size_t)n);
  for (const auto& x : at) p.at[x.first] = x.second;
}

std::vector<OpPtr> block_ops(const SpinBlock& b) {
  std::vector<OpPtr> v;
  for (const OpPtr& o : {b.ident, b.ham})
    if (o) v.push_back(o);
  for (const auto* l : {&b.cre, &b.ccd, &b.cd, &b.cc, &b.cdc, &b.ddc})
    for (const OpPtr& o : *l)
      if (o) v.push_back(o);
  return v;
}

std::vector<OpPtr> GpuRotator::rotate(const std::vector<const SparseOp*>& ops, const Rotation& R) const {
  const double t0 = secs();
  const StateInfo& os = *R.old;
  const auto& rot = *R.rot;
  // rotation matrices (the plan's operator arena): d_q x m_q row-major per old sector
  std::vector<int64_t> uoff(os.nsec(), -1);
  int64_t usize = 0;
  for (int q = 0; q < os.nsec(); ++q)
    if (!rot[q].empty()) {
      uoff[q] = usize;
      usize += (int64_t)rot[q].size();
    }
  // one device buffer: old operators, then the rotated ones, then the matrices
  std::vector<OpPtr> out(ops.size());
  std::vector<std::pair<const SparseOp*, int64_t>> src;
  int64_t soff = 0;
  for (const SparseOp* o : ops) {
    src.push_back({o, soff});
    soff += (int64_t)o->size();
  }
  std::vector<int64_t> ooff(ops.size());
  int64_t osize = 0;
  for (size_t w = 0; w < ops.size(); ++w) {
    const SparseOp& o = *ops[w];
    out[w] = device_out_ ? std::make_shared<SparseOp>(get_pattern(R.ns, o.dn(), o.k2(), o.dirrep()), o.fermion,
                                                      SparseOp::DeviceOnly{})
                         : std::make_shared<SparseOp>(get_pattern(R.ns, o.dn(), o.k2(), o.dirrep()), o.fermion);
    ooff[w] = soff + osize;
    osize += (int64_t)out[w]->size();
  }
  const int64_t total = soff + osize;
  if (osize == 0) return out;
  std::vector<bsg_vblock_t> vb;
  std::vector<bsg_pair_t> pr;
  for (size_t w = 0; w < ops.size(); ++w) {
    const SparseOp& o = *ops[w];
    const auto& P = *out[w]->pat;
    for (const auto& bk : P.blocks) {
      const int orow = R.oldidx[bk.r], ocol = R.oldidx[bk.c];
      const int obi = o.pat->find(orow, ocol);
      if (obi < 0 || bk.rows == 0 || bk.cols == 0) continue;
      // new(r, c) = R_r^T O(orow, ocol) R_c: op(A) = R_r^T (stored d_r x m_r),
      // C = the old block, B' = R_c (d_c x m_c, k-major)
      bsg_pair_t p;
      p.a_off = uoff[orow];
      p.c_off = src[w].second + o.pat->blocks[obi].off;
      p.b_off = uoff[ocol];
      p.k1 = os.dim[orow];
      p.k2 = os.dim[ocol];
      p.a_trans = 1;
      p.b_trans = 0;
      p.alpha_a = 1.0;
      p.alpha = 1.0;
      vb.push_back({ooff[w] + bk.off, bk.rows, bk.cols, (int32_t)pr.size(), 1});
      pr.push_back(p);
    }
  }
  const double t1 = secs();
  double* base = g_.ensure_work((size_t)(total + usize));
  double* ubuf = base + total;
  // operators still in the arena are copied on the device, the rest uploaded
  // (preloaded ones copied on the device by upload_ops)
  std::vector<std::pair<const SparseOp*, int64_t>> host;
  std::vector<bsg_copy_seg_t> segs;
  for (const auto& x : src) {
    auto f = g_.resident.find(x.first);
    if (f != g_.resident.end() && g_.arena && !g_.arena_leased)
      segs.push_back({g_.arena + f->second, base + x.second, (int64_t)x.first->size()});
    else
      host.push_back(x);  // upload_ops copies device-resident ones on the device
  }
  bsg_check(bsg_copy_segments(g_.ctx, segs.data(), (int32_t)segs.size()), "operator copy");
  upload_ops(g_, host, base, (size_t)soff);
  {
    std::vector<double> u((size_t)std::max<int64_t>(1, usize));
    for (int q = 0; q < os.nsec(); ++q)
      if (uoff[q] >= 0) std::copy(rot[q].begin(), rot[q].end(), u.begin() + uoff[q]);
    bsg_check(bsg_memcpy_h2d(g_.ctx, ubuf, u.data(), sizeof(double) * usize), "rotation upload");
  }
  bsg_check(bsg_memset_zero(g_.ctx, base + soff, sizeof(double) * osize), "rotation clear");
  double t2 = secs(), t3 = t2;
  if (!pr.empty()) {
    bsg_chunk_t chunk;
    chunk.vblocks = vb.data();
    chunk.n_vblocks = (int32_t)vb.size();
    chunk.pairs = pr.data();
    chunk.n_pairs = (int32_t)pr.size();
    bsg_plan* plan = nullptr;
    bsg_check(bsg_plan_create(g_.ctx, &chunk, 1, total, &plan), "rotation plan");
    t3 = secs();
    const int rc = bsg_sigma_apply(g_.ctx, plan, ubuf, base, base, 1, BSG_FP64);
    bsg_plan_destroy(plan);  // synchronises the stream
    bsg_check(rc, "rotation");
  }
  const double t4 = secs();
  g_setup.rot_pairs += t1 - t0;
  g_setup.rot_in += t2 - t1;
  g_setup.rot_plan += t3 - t2;
  g_setup.rot_run += t4 - t3;
  struct OutTimer {
    double t;
    ~OutTimer() { g_setup.rot_out += secs() - t; }
  } out_timer{t4};
  if (device_out_) {
    // the rotated operators stay in HBM: one buffer for the batch
    double* nb = nullptr;
    auto owner = device_buffer(g_, (size_t)osize, &nb);
    const bsg_copy_seg_t seg{base + soff, nb, osize};
    bsg_check(bsg_copy_segments(g_.ctx, &seg, 1), "rotated operators");
    for (size_t w = 0; w < ops.size(); ++w) {
      out[w]->dev = nb + (ooff[w] - soff);
      out[w]->dev_owner = owner;
    }
    return out;
  }
  // rotated operators back through the staging
  double* staging = g_.ensure_staging((size_t)osize);
  bsg_check(bsg_memcpy_d2h(g_.ctx, staging, base + soff, sizeof(double) * osize), "rotation download");
#pragma omp parallel for schedule(dynamic)
  for (long w = 0; w < (long)ops.size(); ++w)
    std::memcpy(out[w]->data.data(), staging + (ooff[w] - soff), sizeof(double) * out[w]->data.size());
  return out;
}

}  // namespace blk
