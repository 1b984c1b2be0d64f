#include "gpu_sigma.h"

#include <algorithm>
#include <cstring>

namespace blk {

void bsg_check(int rc, const char* what) {
  if (rc != BSG_OK) throw std::runtime_error(std::string(what) + ": " + bsg_last_error());
}

GpuContext::GpuContext(int device) { bsg_check(bsg_init(&ctx, device), "bsg_init"); }
GpuContext::~GpuContext() {
  if (ctx) bsg_free(ctx);
}

WorkList compile_worklist(const SuperBlock& sb, int64_t w_budget) {
  WorkList wl;
  const WaveLayout& W = *sb.wave;
  const StateInfo &ls = *W.L, &rs = *W.R;
  const int St = W.target.s2;
  const SparseOp* idL = sb.L->ident.get();
  const SparseOp* idR = sb.R->ident.get();

  // operator arena: every distinct operator the term list references
  for (const auto& t : sb.terms)
    for (const SparseOp* o : {t.a.op, t.b.op})
      if (!wl.arena_off.count(o)) {
        wl.arena_off[o] = wl.arena_size;
        wl.arena_size += (int64_t)o->data.size();
      }

  std::vector<std::vector<bsg_contrib_t>> per_v(W.blocks.size());
  std::vector<bsg_gemm1_t> g1;
  int64_t wcur = 0;
  auto flush = [&]() {
    std::vector<bsg_vblock_t> vbs;
    std::vector<bsg_contrib_t> cts;
    for (size_t iv = 0; iv < per_v.size(); ++iv) {
      if (per_v[iv].empty()) continue;
      const auto& b = W.blocks[iv];
      vbs.push_back({b.off, b.rows, b.cols, (int32_t)cts.size(), (int32_t)per_v[iv].size()});
      cts.insert(cts.end(), per_v[iv].begin(), per_v[iv].end());
      per_v[iv].clear();
    }
    if (vbs.empty() && g1.empty()) return;
    wl.g1.push_back(std::move(g1));
    wl.vb.push_back(std::move(vbs));
    wl.ct.push_back(std::move(cts));
    wl.wsize.push_back(wcur);
    g1.clear();
    wcur = 0;
  };

  for (const auto& t : sb.terms) {
    const bool aid = (t.a.op == idL), bid = (t.b.op == idR);
    const int64_t aoff = wl.arena_off[t.a.op], boff = wl.arena_off[t.b.op];
    for (size_t ic = 0; ic < W.blocks.size(); ++ic) {  // c block (lQ', rQ')
      const auto& cb = W.blocks[ic];
      const int lp = cb.l, rp = cb.r;
      for (int l = 0; l < ls.nsec(); ++l) {
        int ia = t.a.find(l, lp);
        if (ia < 0) continue;
        const int dl = ls.dim[l], dlp = ls.dim[lp], drp = rs.dim[rp];
        int32_t src_kind;
        int64_t src_off;
        if (aid) {
          src_kind = 1;
          src_off = cb.off;
        } else {
          src_kind = 0;
          src_off = wcur;
          bsg_gemm1_t g;
          g.a_off = aoff + t.a.op->pat->blocks[ia].off;
          g.c_off = cb.off;
          g.w_off = wcur;
          g.m = dl;
          g.n = drp;
          g.k = dlp;
          g.a_trans = t.a.dag ? 1 : 0;
          g.alpha = t.a.scale(l, lp);
          g1.push_back(g);
          wcur += (int64_t)dl * drp;
        }
        for (int r = 0; r < rs.nsec(); ++r) {
          int iv = W.find(l, r);
          if (iv < 0) continue;
          int ib = t.b.find(r, rp);
          if (ib < 0) continue;
          double f = t.coef * wninej(ls.q[lp].s2, rs.q[rp].s2, St, t.a.k2(), t.b.k2(), 0, ls.q[l].s2, rs.q[r].s2, St);
          if (t.b.fermion() && (ls.q[lp].n & 1)) f = -f;
          f *= t.b.scale(r, rp);
          if (f == 0.0) continue;
          bsg_contrib_t ct;
          ct.src_off = src_off;
          ct.src_kind = src_kind;
          ct.b_off = bid ? -1 : boff + t.b.op->pat->blocks[ib].off;
          ct.b_trans = t.b.dag ? 0 : 1;
          ct.k = drp;
          ct.pad = 0;
          ct.alpha = f;
          per_v[iv].push_back(ct);
          ++wl.n_pairs;
        }
      }
    }
    if (wcur >= w_budget) flush();
  }
  flush();

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

GpuSigma::GpuSigma(GpuContext& g, const SuperBlock& sb, int64_t w_budget, int max_vec) : g_(g) {
  wave_size_ = sb.wave->size;
  WorkList wl = compile_worklist(sb, w_budget);
  n_pairs_ = wl.n_pairs;
  arena_size_ = wl.arena_size;
  diag_terms_ = std::move(wl.diag);
  bsg_check(bsg_malloc(g_.ctx, sizeof(double) * std::max<int64_t>(1, arena_size_), (void**)&d_arena_), "arena");
  // upload operators
  for (const auto& kv : wl.arena_off)
    bsg_check(bsg_memcpy_h2d(g_.ctx, d_arena_ + kv.second, kv.first->data.data(), sizeof(double) * kv.first->data.size()),
              "arena upload");
  size_t wbytes = sizeof(double) * std::max<long>(1, wave_size_) * std::max(1, max_vec);
  bsg_check(bsg_malloc(g_.ctx, wbytes, (void**)&d_c_), "c buffer");
  bsg_check(bsg_malloc(g_.ctx, wbytes, (void**)&d_v_), "v buffer");
  std::vector<bsg_chunk_t> chunks(wl.g1.size());
  for (size_t i = 0; i < chunks.size(); ++i) {
    chunks[i].gemm1 = wl.g1[i].data();
    chunks[i].n_gemm1 = (int32_t)wl.g1[i].size();
    chunks[i].vblocks = wl.vb[i].data();
    chunks[i].n_vblocks = (int32_t)wl.vb[i].size();
    chunks[i].contribs = wl.ct[i].data();
    chunks[i].n_contribs = (int32_t)wl.ct[i].size();
    chunks[i].w_size = wl.wsize[i];
  }
  bsg_check(bsg_plan_create(g_.ctx, chunks.data(), (int32_t)chunks.size(), wave_size_, &plan_), "plan");
  flops_ = bsg_plan_flops(plan_);
}

GpuSigma::~GpuSigma() {
  if (plan_) bsg_plan_destroy(plan_);
  if (d_arena_) bsg_mfree(g_.ctx, d_arena_);
  if (d_c_) bsg_mfree(g_.ctx, d_c_);
  if (d_v_) bsg_mfree(g_.ctx, d_v_);
}

void GpuSigma::apply_device(const double* dc, double* dv, int nvec) {
  bsg_check(bsg_sigma_apply(g_.ctx, plan_, d_arena_, dc, dv, nvec, BSG_FP64), "sigma");
}

void GpuSigma::apply(const double* c_host, double* v_host) {
  size_t bytes = sizeof(double) * wave_size_;
  bsg_check(bsg_memcpy_h2d(g_.ctx, d_c_, c_host, bytes), "c upload");
  bsg_check(bsg_memset_zero(g_.ctx, d_v_, bytes), "v clear");
  apply_device(d_c_, d_v_, 1);
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

}  // namespace blk
