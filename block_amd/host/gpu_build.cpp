// Composite operators built in HBM: the host records the scaled block copies
// of build_sum_block / build_comp_ops (the TensorProduct recursions of
// Operators.C:624-691,1059-1196,1472-1616,1905-1966,2322-2395 for a block
// (x) dot, operatorfunctions.C:146-254 with MatrixBLAS.C:138-162's copies)
// and bsg_build_apply executes them from the device copies of the block's
// operators, so a sweep's blocks never cross PCIe.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

#include "gpu_sigma.h"

namespace blk {

namespace {
double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }
BuildTimes g_build;
std::mutex g_mat_mu;
GpuContext* g_mat_ctx = nullptr;

// downloads device-only operators into their host storage
void download_ops(const std::vector<SparseOp*>& ops) {
  std::lock_guard<std::mutex> lk(g_mat_mu);
  if (!g_mat_ctx) throw std::runtime_error("materialise_host: no GPU context to download operators");
  const double t0 = now_s();
  for (SparseOp* o : ops) {
    if (o->host_valid) continue;
    if (!o->dev) throw std::runtime_error("materialise_host: operator in neither host nor device memory");
    o->data.assign(o->size(), 0.0);
    bsg_check(bsg_memcpy_d2h(g_mat_ctx->ctx, o->data.data(), o->dev, sizeof(double) * o->size()), "operator download");
    o->host_valid = true;
    g_build.download_bytes += 8.0 * o->size();
  }
  g_build.download += now_s() - t0;
}
}  // namespace

const BuildTimes& gpu_build_times() { return g_build; }

BuildPlan plan_build(const std::vector<const BuildItem*>& items,
                     const std::function<const double*(const SparseOp*)>& addr) {
  // groups: items with the same level and destination rectangle, sources in
  // the order recorded (one recording thread per destination operator).
  // Items are bucketed by destination operator (first appearance, stable),
  // each bucket is ordered by (level, offset) on its own, and the buckets
  // fill their slices of the group and source lists in parallel.  Groups
  // write disjoint rectangles, so their order does not change the result.
  const size_t n = items.size();
  if (n > 0x7fffffff) throw std::runtime_error("plan_build: too many items");
  // destination and source operators numbered once (addresses looked up
  // once per operator, not per item)
  std::unordered_map<const SparseOp*, int> dix, six;
  std::vector<const SparseOp*> dops, sops;
  std::vector<int, UninitAlloc<int>> di(n), si(n);
  const SparseOp *last = nullptr, *lasts = nullptr;
  int lastix = -1, lastsx = -1;
  for (size_t i = 0; i < n; ++i) {
    const BuildItem& it = *items[i];
    if (it.level < 0 || it.level > 1) throw std::runtime_error("plan_build: unknown item level");
    if (it.dst != last) {
      auto f = dix.emplace(it.dst, (int)dops.size());
      if (f.second) dops.push_back(it.dst);
      last = it.dst;
      lastix = f.first->second;
    }
    if (it.src != lasts) {
      auto f = six.emplace(it.src, (int)sops.size());
      if (f.second) sops.push_back(it.src);
      lasts = it.src;
      lastsx = f.first->second;
    }
    di[i] = lastix;
    si[i] = lastsx;
  }
  const int nd = (int)dops.size();
  std::vector<const double*> daddr(nd), saddr(sops.size());
  for (int d = 0; d < nd; ++d) daddr[d] = addr(dops[d]);
  for (size_t k = 0; k < sops.size(); ++k) saddr[k] = addr(sops[k]);
  std::vector<int64_t> start(nd + 1, 0);
  for (size_t i = 0; i < n; ++i) ++start[di[i] + 1];
  for (int d = 0; d < nd; ++d) start[d + 1] += start[d];
  std::vector<int, UninitAlloc<int>> order(n);
  {
    std::vector<int64_t> pos(start.begin(), start.end() - 1);
    for (size_t i = 0; i < n; ++i) order[pos[di[i]]++] = (int)i;
  }
  std::vector<int64_t> ng[2], ns[2];
  for (int l = 0; l < 2; ++l) {
    ng[l].assign(nd + 1, 0);
    ns[l].assign(nd + 1, 0);
  }
  auto same = [&](const BuildItem& x, const BuildItem& y) { return x.level == y.level && x.dst_off == y.dst_off; };
#pragma omp parallel for schedule(dynamic, 8)
  for (int d = 0; d < nd; ++d) {
    std::stable_sort(order.begin() + start[d], order.begin() + start[d + 1], [&](int x, int y) {
      const BuildItem &p = *items[x], &q = *items[y];
      if (p.level != q.level) return p.level < q.level;
      return p.dst_off < q.dst_off;
    });
    for (int64_t k = start[d]; k < start[d + 1];) {
      const BuildItem& h = *items[order[k]];
      int64_t e = k + 1;
      while (e < start[d + 1] && same(*items[order[e]], h)) ++e;
      ++ng[h.level][d + 1];
      ns[h.level][d + 1] += e - k;
      k = e;
    }
  }
  BuildPlan bp;
  for (int l = 0; l < 2; ++l) {
    for (int d = 0; d < nd; ++d) {
      ng[l][d + 1] += ng[l][d];
      ns[l][d + 1] += ns[l][d];
    }
    if (ng[l][nd] > 0x7fffffff || ns[l][nd] > 0x7fffffff) throw std::runtime_error("plan_build: too many items");
    bp.groups[l].resize(ng[l][nd]);
    bp.srcs[l].resize(ns[l][nd]);
  }
  std::string err;
  double bytes = 0.0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : bytes)
  for (int d = 0; d < nd; ++d) {
    int64_t g[2] = {ng[0][d], ng[1][d]}, sp[2] = {ns[0][d], ns[1][d]};
    try {
      for (int64_t k = start[d]; k < start[d + 1];) {
        const BuildItem& h = *items[order[k]];
        const int l = h.level;
        bsg_build_group_t& gr = bp.groups[l][g[l]++];
        gr.dst = const_cast<double*>(daddr[d]) + h.dst_off;
        gr.ld = h.dst_ld;
        gr.rows = h.rows;
        gr.cols = h.cols;
        gr.first = (int32_t)sp[l];
        gr.pad = 0;
        int64_t e = k;
        for (; e < start[d + 1]; ++e) {
          const int ie = order[e];
          const BuildItem& it = *items[ie];
          if (!same(it, h)) break;
          if (it.rows != h.rows || it.cols != h.cols || it.dst_ld != h.dst_ld)
            throw std::runtime_error("plan_build: overlapping destination rectangles");
          bsg_build_src_t& s = bp.srcs[l][sp[l]++];
          s.src = saddr[si[ie]] + it.src_off;
          s.ld = it.src_ld;
          s.trans = it.trans;
          s.alpha = it.alpha;
          s.beta = it.beta;
        }
        gr.count = (int32_t)(e - k);
        bytes += 8.0 * h.rows * h.cols * (gr.count + 2);
        k = e;
      }
    } catch (const std::exception& ex) {
#pragma omp critical(plan_build_err)
      if (err.empty()) err = ex.what();
    }
  }
  if (!err.empty()) throw std::runtime_error(err);
  bp.bytes = bytes;
  return bp;
}

void run_build_host(const BuildPlan& bp) {
  for (int level = 0; level < 2; ++level) {
    const auto& groups = bp.groups[level];
    const auto& srcs = bp.srcs[level];
#pragma omp parallel for schedule(dynamic)
    for (long g = 0; g < (long)groups.size(); ++g) {
      const bsg_build_group_t gr = groups[g];
      for (int r = 0; r < gr.rows; ++r)
        for (int c = 0; c < gr.cols; ++c) {
          double s = 0.0;
          for (int j = 0; j < gr.count; ++j) {
            const bsg_build_src_t& sj = srcs[gr.first + j];
            const double a = sj.trans ? sj.src[(int64_t)c * sj.ld + r] : sj.src[(int64_t)r * sj.ld + c];
            s += (sj.alpha * a) * sj.beta;
          }
          gr.dst[(int64_t)r * gr.ld + c] += s;
        }
    }
  }
}

void execute_build_host(const BuildSession& session, const std::vector<SparseOp*>& targets) {
  const std::vector<const BuildItem*> items = session.item_ptrs();
  const std::vector<OpPtr> written = session.written();
  std::unordered_map<const SparseOp*, double*> where;
  std::vector<std::vector<double>> scratch;
  for (SparseOp* o : targets) {
    if (!o || where.count(o)) continue;
    o->data.assign(o->size(), 0.0);
    o->host_valid = true;
    where[o] = o->data.data();
  }
  for (const OpPtr& o : written) {
    if (where.count(o.get())) continue;
    scratch.emplace_back(o->size(), 0.0);
    where[o.get()] = scratch.back().data();
  }
  auto addr = [&](const SparseOp* o) -> const double* {
    auto f = where.find(o);
    if (f != where.end()) return f->second;
    if (!o->host_valid) throw std::runtime_error("execute_build_host: source operator not in host memory");
    return o->data.data();
  };
  const double tp = now_s();
  const BuildPlan bp = plan_build(items, addr);
  g_build.plan += now_s() - tp;
  g_build.items += (double)items.size();
  run_build_host(bp);
}

void set_materialiser_context(GpuContext* g) {
  std::lock_guard<std::mutex> lk(g_mat_mu);
  g_mat_ctx = g;
  set_host_materialiser(g ? download_ops : nullptr);
}

void make_device_resident(GpuContext& g, const std::vector<OpPtr>& ops) {
  std::vector<std::pair<const SparseOp*, int64_t>> up;
  int64_t total = 0;
  for (const OpPtr& o : ops) {
    if (!o || o->dev || !o->host_valid || o->size() == 0) continue;
    up.push_back({o.get(), total});
    total += o->size();
  }
  if (up.empty()) return;
  double* base = nullptr;
  auto owner = device_buffer(g, (size_t)total, &base);
  upload_ops(g, up, base, (size_t)total);
  bsg_check(bsg_synchronize(g.ctx), "resident upload");
  for (auto& x : up) {
    SparseOp* o = const_cast<SparseOp*>(x.first);
    o->dev = base + x.second;
    o->dev_owner = owner;
    o->data = std::vector<double>();
    o->host_valid = false;
  }
}

std::shared_ptr<void> device_buffer(GpuContext& g, size_t n, double** ptr) {
  double* p = nullptr;
  bsg_check(bsg_malloc(g.ctx, sizeof(double) * std::max<size_t>(1, n), (void**)&p), "device block storage");
  bsg_context* ctx = g.ctx;
  *ptr = p;
  return std::shared_ptr<void>(p, [ctx](void* q) { bsg_mfree(ctx, q); });
}

void execute_build(GpuContext& g, const BuildSession& session, const std::vector<SparseOp*>& targets) {
  const double t0 = now_s();
  const std::vector<const BuildItem*> items = session.item_ptrs();
  const std::vector<OpPtr> written = session.written();
  // destinations: the targets in one buffer they keep, other written
  // operators (linear-combination intermediates) in scratch
  std::unordered_map<const SparseOp*, double*> where;
  int64_t tsize = 0;
  std::vector<std::pair<SparseOp*, int64_t>> tlist;
  for (SparseOp* o : targets) {
    if (!o || where.count(o)) continue;
    where[o] = nullptr;
    tlist.push_back({o, tsize});
    tsize += o->size();
  }
  std::vector<std::pair<const SparseOp*, int64_t>> slist;  // scratch: intermediates, then uploaded sources
  int64_t ssize = 0;
  for (const OpPtr& o : written) {
    if (where.count(o.get())) continue;
    where[o.get()] = nullptr;
    slist.push_back({o.get(), ssize});
    ssize += o->size();
  }
  const int64_t inter = ssize;
  std::vector<std::pair<const SparseOp*, int64_t>> up;  // host-only sources
  const SparseOp* seen = nullptr;
  for (const BuildItem* it : items) {
    const SparseOp* s = it->src;
    if (s == seen) continue;
    seen = s;
    if (where.count(s) || s->dev) continue;
    if (!s->host_valid) throw std::runtime_error("execute_build: source operator in neither host nor device memory");
    where[s] = nullptr;
    up.push_back({s, ssize - inter});
    ssize += s->size();
  }
  double* tbuf = nullptr;
  auto owner = device_buffer(g, (size_t)tsize, &tbuf);
  double* sbuf = nullptr;
  std::shared_ptr<void> sowner;
  if (ssize > 0) sowner = device_buffer(g, (size_t)ssize, &sbuf);
  for (auto& x : tlist) where[x.first] = tbuf + x.second;
  for (auto& x : slist) where[x.first] = sbuf + x.second;
  for (auto& x : up) where[x.first] = sbuf + inter + x.second;
  bsg_check(bsg_memset_zero(g.ctx, tbuf, sizeof(double) * std::max<int64_t>(1, tsize)), "build clear");
  if (inter > 0) bsg_check(bsg_memset_zero(g.ctx, sbuf, sizeof(double) * inter), "build clear");
  if (!up.empty()) {
    // small host operators (the dot's) through a plain copy each
    std::vector<double> h;
    for (auto& x : up) h.insert(h.end(), x.first->data.begin(), x.first->data.end());
    bsg_check(bsg_memcpy_h2d(g.ctx, sbuf + inter, h.data(), sizeof(double) * h.size()), "build sources");
  }
  auto addr = [&](const SparseOp* o) -> const double* {
    auto f = where.find(o);
    if (f != where.end() && f->second) return f->second;
    if (o->dev) return o->dev;
    throw std::runtime_error("execute_build: operator without an address");
  };
  const double tp = now_s();
  const BuildPlan bp = plan_build(items, addr);
  const double ta = now_s();
  for (int level = 0; level < 2; ++level)
    bsg_check(bsg_build_apply(g.ctx, bp.groups[level].data(), (int32_t)bp.groups[level].size(), bp.srcs[level].data(),
                              (int32_t)bp.srcs[level].size()),
              "composite build");
  g_build.plan += ta - tp;
  g_build.apply += now_s() - ta;
  g_build.items += (double)items.size();
  const double bytes = bp.bytes;
  for (auto& x : tlist) {
    x.first->data = std::vector<double>();
    x.first->host_valid = false;
    x.first->dev = tbuf + x.second;
    x.first->dev_owner = owner;
  }
  sowner.reset();  // the scratch (intermediates, uploaded sources) goes now
  g_build.execute += now_s() - t0;
  g_build.bytes += bytes;
}

BlockPtr GpuBlockBuilder::sum_block(const Integrals& I, const SpinBlock& B1, const SpinBlock& dot,
                                    const Quantum& target, int nrem, const BuildSpec& spec, const Screen& scr) const {
  BlockPtr b;
  {
    BuildSession session;
    const double t0 = now_s();
    b = build_sum_block(I, B1, dot, target, nrem, spec, scr);
    g_build.record += now_s() - t0;
    std::vector<SparseOp*> ops;
    for (const OpPtr& o : block_ops(*b)) ops.push_back(o.get());
    execute_build(g_, session, ops);
  }
  return b;
}

std::shared_ptr<CompOpsAhead> GpuBlockBuilder::comp_ops(const Integrals& I, const SpinBlock& B, const SpinBlock& B1,
                                                        const SpinBlock& dot, const Screen& scr) const {
  std::shared_ptr<CompOpsAhead> a;
  {
    BuildSession session;
    const double t0 = now_s();
    a = build_comp_ops(I, B, B1, dot, scr);
    g_build.record += now_s() - t0;
    std::vector<SparseOp*> ops;
    for (const OpPtr& o : comp_ops_list(*a)) ops.push_back(o.get());
    execute_build(g_, session, ops);
  }
  return a;
}

}  // namespace blk
