// extern "C" surface of the host library (libblock_host.so) used by the
// Python layer: integrals, superblock snapshots, the GPU multiply_h and the
// sweep driver running its Davidson sigma on the GPU.  Errors are reported as
// negative return codes / null handles with blkh_last_error() holding the text.
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <string>

#include "checkpoint.h"
#include "gpu_sigma.h"
#include "sweep.h"

using namespace blk;

namespace {
thread_local std::string g_err;
template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
template <class F>
void* guarded_ptr(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

struct GpuSigmaHandle {
  std::unique_ptr<GpuContext> ctx;
  std::unique_ptr<GpuSigma> sig;
};

struct GpuFunctor : DavidsonFunctor {
  GpuSigma sig;
  double min_split_share = 0.0;  // the ladder's split phase below this share is skipped
  GpuFunctor(GpuContext& g, const SuperBlock& sb, double min_share = 0.0) : sig(g, sb), min_split_share(min_share) {}
  void operator()(const double* c, double* v) override { sig.apply(c, v); }
  bool set_precision(int mode) override {
    sig.set_precision(mode);
    return true;
  }
  bool diagonal(std::vector<double>& d) override {
    d = sig.diagonal();
    return true;
  }
  bool davidson(std::vector<std::vector<double>>& b, std::vector<double>& h_diag, double normtol, bool useprecond,
                int deflation_min, int deflation_max, DavidsonStats* st, int max_iter, double ladder,
                const std::vector<std::vector<double>>* lower) override {
    // a subspace beyond the device solve's limit runs in the host Davidson
    // (block_davidson) with this functor's GPU sigma
    const int dmin = std::max(deflation_min, (int)b.size());
    if (std::max(deflation_max, dmin + 1) + 1 > kDeviceDavidsonMaxVectors) return false;
    block_davidson_device(sig, b, h_diag, normtol, useprecond, deflation_min, deflation_max, st, max_iter, ladder,
                          lower, min_split_share);
    return true;
  }
};
}  // namespace

extern "C" {

const char* blkh_last_error(void) { return g_err.c_str(); }

void* blkh_load_fcidump(const char* path) {
  return guarded_ptr([&]() -> void* { return new Integrals(read_fcidump(path)); });
}
void blkh_free_integrals(void* h) { delete static_cast<Integrals*>(h); }
// Fiedler ordering of the orbitals (input.C:2366-2376) and the integrals in a
// given order (new orbital a = old orbital perm[a]).
int blkh_fiedler_order(void* ints, int* order) {
  return guarded([&]() {
    auto o = fiedler_order(*static_cast<Integrals*>(ints));
    std::copy(o.begin(), o.end(), order);
    return 0;
  });
}
int blkh_fiedler_order_matrix(int n, const double* m, int* order) {
  return guarded([&]() {
    auto o = fiedler_order_matrix(n, m);
    std::copy(o.begin(), o.end(), order);
    return 0;
  });
}
void* blkh_reorder_integrals(void* ints, const int* perm) {
  return guarded_ptr([&]() -> void* {
    const Integrals& I = *static_cast<Integrals*>(ints);
    std::vector<int> p(perm, perm + I.norb), seen(I.norb, 0);
    for (int x : p) {
      if (x < 0 || x >= I.norb || seen[x]++) throw std::runtime_error("reorder: not a permutation");
    }
    return new Integrals(I.reordered(p));
  });
}
int blkh_norb(void* h) { return static_cast<Integrals*>(h)->norb; }
int blkh_nelec(void* h) { return static_cast<Integrals*>(h)->nelec; }
int blkh_isym(void* h) { return static_cast<Integrals*>(h)->isym; }
int blkh_ms2(void* h) { return static_cast<Integrals*>(h)->ms2; }
double blkh_ecore(void* h) { return static_cast<Integrals*>(h)->ecore; }

// mid-chain superblock: sys = sites [0, nsys), dots nsys and nsys+1, env the rest,
// blocks grown with a Wilson truncation to M
void* blkh_snapshot(void* ints, int nelec, int twoS, int irrep, int nsys, int M) {
  return guarded_ptr([&]() -> void* {
    const Integrals& I = *static_cast<Integrals*>(ints);
    if (nsys < 1 || nsys + 2 > I.norb - 1) throw std::runtime_error("blkh_snapshot: nsys out of range");
    Screen scr;
    return new SuperBlock(snapshot_superblock(I, Quantum{nelec, twoS, irrep}, nsys, M, scr));
  });
}
// mid-chain superblock from the infinite-system DMRG warm-up (density-matrix
// truncation of both halves to M, Davidson sigma on the GPU)
void* blkh_snapshot_dmrg(void* ints, int nelec, int twoS, int irrep, int nroots, int M, int device, int verbose) {
  return guarded_ptr([&]() -> void* {
    const Integrals& I = *static_cast<Integrals*>(ints);
    DmrgParams P;
    P.nelec = nelec;
    P.twoS = twoS;
    P.irrep = irrep;
    P.nroots = nroots;
    P.schedule = {{0, M, 1e-7, 0.0}};  // noise-free, as in converged sweeps
    P.verbose = verbose;
    GpuContext ctx(device);
    P.rotator = std::make_shared<GpuRotator>(ctx);
    SuperBlock sb = dmrg_snapshot(I, P, [&](const SuperBlock& s) {
      return std::unique_ptr<DavidsonFunctor>(new GpuFunctor(ctx, s));
    });
    return new SuperBlock(std::move(sb));
  });
}
// The two-site superblock of forward step l (sys = sites 0..l, dots l+1 and
// l+2, env = sites l+3..) from the blocks of a saved sweep state
// (checkpoint file of blkh_dmrg_gpu): the workload a sweep at that state's
// bond dimension solves at that step.  l < 0: the middle of the chain.
void* blkh_superblock_from_state(void* ints, const char* path, int nelec, int twoS, int irrep, int l) {
  return guarded_ptr([&]() -> void* {
    const Integrals& I = *static_cast<Integrals*>(ints);
    const int k = I.norb;
    if (l < 0) l = k / 2 - 2;
    if (l < 0 || l + 3 >= k) throw std::runtime_error("blkh_superblock_from_state: step out of range");
    SweepState st = load_sweep_state(path);
    if ((int)st.fwd.size() != k || !st.fwd[l] || !st.bwd[l + 3])
      throw std::runtime_error("blkh_superblock_from_state: the state has no blocks for this step");
    Screen scr;
    const Quantum target{nelec, twoS, irrep};
    BuildSpec sl;
    sl.normal = true;
    sl.comp = false;
    BuildSpec sr;
    sr.normal = false;
    sr.comp = true;
    const int nL = (int)st.fwd[l]->sites.size() + 1, nR = (int)st.bwd[l + 3]->sites.size() + 1;
    BlockPtr L = build_sum_block(I, *st.fwd[l], *make_dot(I, l + 1, scr), target, nR, sl, scr);
    BlockPtr R = build_sum_block(I, *st.bwd[l + 3], *make_dot(I, l + 2, scr), target, nL, sr, scr);
    return new SuperBlock(make_superblock(I, L, R, target));
  });
}
// Kept states (multiplets) and sectors of every block of a saved sweep
// state: fwd_m[l], bwd_m[l] (0 where absent), fwd_sec[l], bwd_sec[l];
// arrays of norb entries.
int blkh_state_block_dims(void* ints, const char* path, int* fwd_m, int* bwd_m, int* fwd_sec, int* bwd_sec) {
  return guarded([&]() {
    const int k = static_cast<Integrals*>(ints)->norb;
    SweepState st = load_sweep_state(path);
    if ((int)st.fwd.size() != k) throw std::runtime_error("blkh_state_block_dims: another number of orbitals");
    for (int l = 0; l < k; ++l) {
      fwd_m[l] = st.fwd[l] ? (int)st.fwd[l]->si->total : 0;
      bwd_m[l] = st.bwd[l] ? (int)st.bwd[l]->si->total : 0;
      fwd_sec[l] = st.fwd[l] ? st.fwd[l]->si->nsec() : 0;
      bwd_sec[l] = st.bwd[l] ? st.bwd[l]->si->nsec() : 0;
    }
    return 0;
  });
}
// The compiled work list of a superblock with its provenance, for checking
// the scalar factors against other conventions.  Returns the pair count
// (call with max 0 to size the arrays).  pair_int[6 p ..]: term, l, l', r,
// r', identity flags (bit 0: A, bit 1: B); pair_dbl[2 p ..]: alpha_a, alpha.
// term_int[7 t ..]: rank of a, rank of b (2k), a adjoint, b adjoint, b
// fermionic, delta N of a, delta N of b; term_dbl[t]: coefficient.
// lq[3 i ..], rq[3 i ..]: (N, 2S, irrep) of the left and right sectors;
// meta[0..4]: number of terms, of left and right sectors, 2S of the target.
long blkh_worklist_provenance(void* sbp, long max, int* pair_int, double* pair_dbl, int* term_int, double* term_dbl,
                              int* lq, int* rq, int* meta) {
  try {
    const SuperBlock& sb = *static_cast<SuperBlock*>(sbp);
    WorkList wl = compile_worklist(sb);
    const long n = (long)wl.pairs.size();
    if (max <= 0) return n;
    if (max < n) throw std::runtime_error("blkh_worklist_provenance: arrays too small");
    for (long p = 0; p < n; ++p) {
      const auto& pv = wl.prov[p];
      const auto& pr = wl.pairs[p];
      int* pi = pair_int + 6 * p;
      pi[0] = pv.term;
      pi[1] = pv.l;
      pi[2] = pv.lp;
      pi[3] = pv.r;
      pi[4] = pv.rp;
      pi[5] = (pr.a_off < 0 ? 1 : 0) | (pr.b_off < 0 ? 2 : 0);
      pair_dbl[2 * p] = pr.alpha_a;
      pair_dbl[2 * p + 1] = pr.alpha;
    }
    for (size_t t = 0; t < sb.terms.size(); ++t) {
      const auto& tm = sb.terms[t];
      int* ti = term_int + 7 * t;
      ti[0] = tm.a.k2();
      ti[1] = tm.b.k2();
      ti[2] = tm.a.dag;
      ti[3] = tm.b.dag;
      ti[4] = tm.b.fermion();
      ti[5] = tm.a.dn();
      ti[6] = tm.b.dn();
      term_dbl[t] = tm.coef;
    }
    const StateInfo &ls = *sb.wave->L, &rs = *sb.wave->R;
    for (int i = 0; i < ls.nsec(); ++i) {
      lq[3 * i] = ls.q[i].n;
      lq[3 * i + 1] = ls.q[i].s2;
      lq[3 * i + 2] = ls.q[i].irrep;
    }
    for (int i = 0; i < rs.nsec(); ++i) {
      rq[3 * i] = rs.q[i].n;
      rq[3 * i + 1] = rs.q[i].s2;
      rq[3 * i + 2] = rs.q[i].irrep;
    }
    meta[0] = (int)sb.terms.size();
    meta[1] = ls.nsec();
    meta[2] = rs.nsec();
    meta[3] = sb.wave->target.s2;
    return n;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
void blkh_free_superblock(void* h) { delete static_cast<SuperBlock*>(h); }
long blkh_wave_size(void* h) { return static_cast<SuperBlock*>(h)->wave->size; }
double blkh_flops(void* h) { return sigma_flops(*static_cast<SuperBlock*>(h)); }
int blkh_nterms(void* h) { return (int)static_cast<SuperBlock*>(h)->terms.size(); }
int blkh_nsectors(void* h, int* nl, int* nr, int* nblocks) {
  auto* sb = static_cast<SuperBlock*>(h);
  *nl = sb->L->si->nsec();
  *nr = sb->R->si->nsec();
  *nblocks = (int)sb->wave->blocks.size();
  return 0;
}
// rows and columns of the wavefunction's blocks (V blocks), at most max of them;
// returns the number of blocks
int blkh_wave_blocks(void* h, int* rows, int* cols, int max) {
  auto* sb = static_cast<SuperBlock*>(h);
  const int n = (int)sb->wave->blocks.size();
  for (int i = 0; i < n && i < max; ++i) {
    rows[i] = sb->wave->blocks[i].rows;
    cols[i] = sb->wave->blocks[i].cols;
  }
  return n;
}
long blkh_operator_bytes(void* h) {
  auto* sb = static_cast<SuperBlock*>(h);
  return (long)(sb->L->nbytes() + sb->R->nbytes());
}
int blkh_diag_cpu(void* h, double* out) {
  return guarded([&]() {
    auto d = diagonal_h(*static_cast<SuperBlock*>(h));
    std::memcpy(out, d.data(), sizeof(double) * d.size());
    return 0;
  });
}

// max_vec: vectors the handle's device buffers hold, and the plan's batch
void* blkh_gpu_sigma_create(void* sb, int device, long max_vec) {
  return guarded_ptr([&]() -> void* {
    auto* h = new GpuSigmaHandle();
    try {
      h->ctx.reset(new GpuContext(device));
      h->sig.reset(new GpuSigma(*h->ctx, *static_cast<SuperBlock*>(sb), (int)std::max(1L, max_vec)));
    } catch (...) {
      delete h;
      throw;
    }
    return h;
  });
}
void blkh_gpu_sigma_free(void* h) { delete static_cast<GpuSigmaHandle*>(h); }
int blkh_gpu_sigma_apply(void* h, const double* c, double* v) {
  return guarded([&]() {
    static_cast<GpuSigmaHandle*>(h)->sig->apply(c, v);
    return 0;
  });
}
int blkh_gpu_sigma_diag(void* h, double* out) {
  return guarded([&]() {
    auto d = static_cast<GpuSigmaHandle*>(h)->sig->diagonal();
    std::memcpy(out, d.data(), sizeof(double) * d.size());
    return 0;
  });
}
double blkh_gpu_sigma_flops(void* h) { return static_cast<GpuSigmaHandle*>(h)->sig->flops(); }
double blkh_gpu_sigma_split_share(void* h) { return static_cast<GpuSigmaHandle*>(h)->sig->split_share(); }
int blkh_gpu_sigma_phase_flops(void* h, double* f1, double* f2) {
  return bsg_plan_phase_flops(static_cast<GpuSigmaHandle*>(h)->sig->plan(), f1, f2);
}
int blkh_gpu_sigma_set_precision(void* h, int mode) {
  return guarded([&]() {
    static_cast<GpuSigmaHandle*>(h)->sig->set_precision(mode);
    return 0;
  });
}
// the C-ABI handles behind a GPU sigma handle (context, plan, operator arena,
// device c and v buffers of wave_size doubles), for callers that drive
// libblock_sigma.so directly
int blkh_gpu_sigma_raw(void* h, void** ctx, void** plan, const double** arena, double** dev_c, double** dev_v) {
  auto* g = static_cast<GpuSigmaHandle*>(h);
  *ctx = g->sig->ctx();
  *plan = g->sig->plan();
  *arena = g->sig->dev_arena();
  *dev_c = g->sig->dev_c();
  *dev_v = g->sig->dev_v();
  return 0;
}
long blkh_gpu_sigma_pairs(void* h) { return (long)static_cast<GpuSigmaHandle*>(h)->sig->n_pairs(); }
long blkh_gpu_sigma_arena_bytes(void* h) { return (long)static_cast<GpuSigmaHandle*>(h)->sig->arena_bytes(); }
// device-resident loop for timing: c is uploaded once, v accumulates `reps` sigmas.
// kernel_ms[0..2] receive the summed fused-kernel / reduction / wavefunction
// split times measured with HIP events on the plan's stream, wall_ms the
// whole loop; the handle's current precision is used.
// c: nvec vectors (nvec <= the handle's max_vec), batched by the plan's
// batch size (blkh_gpu_sigma_set_batch).
int blkh_gpu_sigma_timed(void* h, const double* c, int nvec, int warmup, int reps, double* wall_ms, double* kernel_ms,
                         long* launches) {
  return guarded([&]() {
    auto* g = static_cast<GpuSigmaHandle*>(h);
    GpuSigma& s = *g->sig;
    if (nvec < 1 || nvec > s.max_vec()) throw std::runtime_error("blkh_gpu_sigma_timed: nvec beyond the handle's max_vec");
    size_t bytes = sizeof(double) * s.wave_size() * nvec;
    bsg_check(bsg_memcpy_h2d(s.ctx(), s.dev_c(), c, bytes), "c upload");
    bsg_check(bsg_memset_zero(s.ctx(), s.dev_v(), bytes), "v clear");
    for (int i = 0; i < warmup; ++i) s.apply_device(s.dev_c(), s.dev_v(), nvec);
    bsg_check(bsg_synchronize(s.ctx()), "sync");
    if (s.graph()) {
      // graph replays back to back: wall time between stream synchronisations
      const auto t0 = std::chrono::steady_clock::now();
      for (int i = 0; i < reps; ++i) s.apply_device(s.dev_c(), s.dev_v(), nvec);
      bsg_check(bsg_synchronize(s.ctx()), "sync");
      *wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      for (int i = 0; i < 3; ++i) {
        kernel_ms[i] = 0.0;
        launches[i] = 0;
      }
      launches[0] = reps;  // one graph launch per application
      return 0;
    }
    bsg_check(bsg_sigma_profile(s.ctx(), s.plan(), s.dev_arena(), s.dev_c(), s.dev_v(), nvec, reps, s.precision(),
                                wall_ms, kernel_ms, launches),
              "profile");
    return 0;
  });
}
// the handle's plan built again (tile choices set since with bsg_set_tile_table)
int blkh_gpu_sigma_replan(void* h) {
  return guarded([&]() {
    static_cast<GpuSigmaHandle*>(h)->sig->replan();
    return 0;
  });
}
// sigma applications through a captured HIP graph (1) or launched one by one (0)
int blkh_gpu_sigma_set_graph(void* h, int on) {
  return guarded([&]() {
    static_cast<GpuSigmaHandle*>(h)->sig->set_graph(on != 0);
    return 0;
  });
}
int blkh_gpu_sigma_set_batch(void* h, int nroot) {
  return guarded([&]() {
    static_cast<GpuSigmaHandle*>(h)->sig->set_batch(nroot);
    return 0;
  });
}
// v = H c for nvec host vectors (nvec <= max_vec) in the plan's batches
int blkh_gpu_sigma_apply_n(void* h, const double* c, double* v, int nvec) {
  return guarded([&]() {
    static_cast<GpuSigmaHandle*>(h)->sig->apply(c, v, nvec);
    return 0;
  });
}

// One Davidson solve on a superblock with the GPU sigma (Linear::
// block_davidson, linear.C:179-385): mode 0 runs the device-resident solve
// (vectors in HBM), mode 1 the host control flow (block_davidson) calling
// the same GPU sigma.  guesses: nroots vectors of wave_size (the solve starts
// from them); lower: nlow states to project off (lowerStates, null if 0).
// Deflation as the sweep's (min max(2, nroots), max max(20, nroots + 20)).
// energies[nroots], vecs[nroots * wave_size] (may be null),
// stats[3] = iterations, sigma calls, split-precision sigma calls.
int blkh_davidson(void* sbp, int device, int mode, int nroots, const double* guesses, int nlow, const double* lower,
                  double tol, double ladder, double* energies, double* vecs, int* stats) {
  return guarded([&]() {
    const SuperBlock& sb = *static_cast<SuperBlock*>(sbp);
    const long N = sb.wave->size;
    if (nroots < 1 || nlow < 0 || !guesses || (nlow > 0 && !lower) || !energies || !stats)
      throw std::runtime_error("blkh_davidson: bad arguments");
    GpuContext ctx(device);
    GpuFunctor h(ctx, sb);
    std::vector<std::vector<double>> b(nroots), low(nlow);
    for (int r = 0; r < nroots; ++r) b[r].assign(guesses + (size_t)r * N, guesses + (size_t)(r + 1) * N);
    for (int r = 0; r < nlow; ++r) low[r].assign(lower + (size_t)r * N, lower + (size_t)(r + 1) * N);
    std::vector<double> e;
    h.diagonal(e);
    const int dmin = std::max(2, nroots);
    const int dmax = nroots > 1 ? std::max(20, nroots + 20) : std::max(20, dmin + 1);
    DavidsonStats st;
    bool done = false;
    if (mode == 0) done = h.davidson(b, e, tol, true, dmin, dmax, &st, 200, ladder, nlow ? &low : nullptr);
    if (!done) block_davidson(b, e, tol, h, true, dmin, dmax, &st, 200, ladder, nlow ? &low : nullptr);
    for (int r = 0; r < nroots; ++r) {
      energies[r] = e[r];
      if (vecs) std::copy(b[r].begin(), b[r].end(), vecs + (size_t)r * N);
    }
    stats[0] = st.iterations;
    stats[1] = st.sigma_calls;
    stats[2] = st.sigma_calls_split;
    return done || mode != 0 ? 0 : 1;
  });
}

// Two-site DMRG with the sigma on the GPU.  schedule: n_sched rows of
// (start_sweep, M, davidson_tol, noise).  ladder > 0: split-precision sigma
// until the Davidson residual norm falls below it, FP64 after.  ahead: host
// operator builds beside the Davidson solve (1 on, 0 off, -1 default: on).
// device_blocks: composite operators built in HBM and renormalised blocks
// kept there (no operator crosses PCIe; host builds beside the solve off).
// energies_out[nroots]; sweep_out[max_sweeps * kSweepCols]: one row per sweep
// (columns as write_sweep_records).  Returns the number of rows written.
int blkh_dmrg_gpu(void* ints, int nelec, int twoS, int irrep, int nroots, const double* weights, int n_sched,
                  const double* sched, int max_sweeps, double sweep_tol, double ladder, int ladder_from, int guess,
                  const char* checkpoint, const char* resume, int ahead, int device_blocks, int device, int verbose,
                  double ladder_min_share, double* energies_out, double* sweep_out) {
  return guarded([&]() {
    const Integrals& I = *static_cast<Integrals*>(ints);
    DmrgParams P;
    P.nelec = nelec;
    P.twoS = twoS;
    P.irrep = irrep;
    P.nroots = nroots;
    if (weights) P.weights.assign(weights, weights + nroots);
    for (int i = 0; i < n_sched; ++i)
      P.schedule.push_back({(int)sched[4 * i], (int)sched[4 * i + 1], sched[4 * i + 2], sched[4 * i + 3]});
    P.max_sweeps = max_sweeps;
    P.sweep_tol = sweep_tol;
    P.verbose = verbose;
    P.ladder = ladder;
    P.ladder_from_sweep = ladder_from;
    P.transform_guess = guess != 0;
    if (checkpoint) P.checkpoint = checkpoint;
    if (resume) P.resume = resume;
    P.host_ahead = ahead;
    GpuContext ctx(device);
    // device-only operators are downloaded through ctx while it lives
    struct MatGuard {
      explicit MatGuard(GpuContext* g) { set_materialiser_context(g); }
      ~MatGuard() { set_materialiser_context(nullptr); }
    } mat(&ctx);
    P.rotator = std::make_shared<GpuRotator>(ctx, device_blocks != 0);
    if (device_blocks) {
      P.builder = std::make_shared<GpuBlockBuilder>(ctx);
      P.resident_block = [&ctx](const SpinBlock& b) { make_device_resident(ctx, block_ops(b)); };
    } else {
      P.preload_block = [&ctx](const SpinBlock& b) { preload_ops(ctx, 1, block_ops(b)); };
    }
    auto res = run_dmrg(I, P, [&](const SuperBlock& sb) {
      return std::unique_ptr<DavidsonFunctor>(new GpuFunctor(ctx, sb, ladder_min_share));
    });
    for (int r = 0; r < nroots; ++r) energies_out[r] = res.energies[r];
    return write_sweep_records(res, sweep_out, max_sweeps);
  });
}

// Accumulated recorded-build times of this process: out[6] = record,
// execute, grouping, device, algorithmic bytes, items (BuildTimes).
int blkh_build_times(double* out) {
  const BuildTimes& t = gpu_build_times();
  const double v[6] = {t.record, t.execute, t.plan, t.apply, t.bytes, t.items};
  for (int i = 0; i < 6; ++i) out[i] = v[i];
  return 0;
}

// The recorded build executed on the host (plan_build's groups, k_build's
// arithmetic) against the host build, for a Wilson-grown block sys of nsys
// sites and the dot nsys (normal and complementary sets of sys (x) dot and
// its complementary operators in its own basis): no GPU needed.
int blkh_build_check_cpu(void* ints, int nsys, int M, double* maxdiff, long* nops) {
  return guarded([&]() {
    const Integrals& I = *static_cast<Integrals*>(ints);
    Screen scr;
    const Quantum target{I.nelec, 0, I.isym};
    std::vector<int> order(I.norb);
    for (int i = 0; i < I.norb; ++i) order[i] = i;
    BlockPtr sys = warmup_block(I, order, nsys, target, M, true, scr);
    BlockPtr dot = make_dot(I, nsys, scr);
    const int nrem = I.norb - nsys - 1;
    double worst = 0.0;
    long n = 0;
    auto cmp = [&](const std::vector<OpPtr>& a, const std::vector<OpPtr>& b) {
      if (a.size() != b.size()) throw std::runtime_error("build check: operator sets differ");
      for (size_t i = 0; i < a.size(); ++i) {
        if (a[i]->size() != b[i]->size() || a[i]->pat->blocks.size() != b[i]->pat->blocks.size())
          throw std::runtime_error("build check: patterns differ");
        for (long e = 0; e < a[i]->size(); ++e) worst = std::max(worst, std::fabs(a[i]->data[e] - b[i]->data[e]));
        ++n;
      }
    };
    auto ptrs = [](const std::vector<OpPtr>& v) {
      std::vector<SparseOp*> p;
      for (const OpPtr& o : v) p.push_back(o.get());
      return p;
    };
    for (int comp = 0; comp < 2; ++comp) {
      BuildSpec sp;
      sp.normal = comp == 0;
      sp.comp = comp == 1;
      BlockPtr rec;
      {
        BuildSession session;
        rec = build_sum_block(I, *sys, *dot, target, nrem, sp, scr);
        execute_build_host(session, ptrs(block_ops(*rec)));
      }
      BlockPtr h = build_sum_block(I, *sys, *dot, target, nrem, sp, scr);
      cmp(block_ops(*h), block_ops(*rec));
      if (comp == 0) {
        std::shared_ptr<CompOpsAhead> ra;
        {
          BuildSession session;
          ra = build_comp_ops(I, *h, *sys, *dot, scr);
          execute_build_host(session, ptrs(comp_ops_list(*ra)));
        }
        cmp(comp_ops_list(*build_comp_ops(I, *h, *sys, *dot, scr)), comp_ops_list(*ra));
      }
    }
    *maxdiff = worst;
    *nops = n;
    return 0;
  });
}

// Device builds beside host builds: a Wilson-grown block sys (nsys sites) and
// the dot nsys; sys (x) dot with normal operators, with complementary
// operators, and sys (x) dot's complementary operators in its own basis, each
// built on the host and in HBM; then the same from a device-only sys (its
// operators rotated on the GPU and left there, so the build reads HBM only).
// maxdiff: largest |host - device| element over all operators, nops: the
// number compared.
int blkh_build_check(void* ints, int nsys, int M, int device, double* maxdiff, long* nops) {
  return guarded([&]() {
    const Integrals& I = *static_cast<Integrals*>(ints);
    Screen scr;
    const Quantum target{I.nelec, 0, I.isym};
    std::vector<int> order(I.norb);
    for (int i = 0; i < I.norb; ++i) order[i] = i;
    BlockPtr sys = warmup_block(I, order, nsys, target, M, true, scr);
    BlockPtr dot = make_dot(I, nsys, scr);
    GpuContext ctx(device);
    struct MatGuard {
      explicit MatGuard(GpuContext* g) { set_materialiser_context(g); }
      ~MatGuard() { set_materialiser_context(nullptr); }
    } mat(&ctx);
    GpuBlockBuilder gb(ctx);
    double worst = 0.0;
    long n = 0;
    auto cmp_lists = [&](const std::vector<OpPtr>& a, const std::vector<OpPtr>& b) {
      if (a.size() != b.size()) throw std::runtime_error("build check: operator sets differ");
      std::vector<SparseOp*> dev;
      for (const OpPtr& o : b) dev.push_back(o.get());
      materialise_host(dev);
      for (size_t i = 0; i < a.size(); ++i) {
        if (a[i]->size() != b[i]->size() || a[i]->pat->blocks.size() != b[i]->pat->blocks.size())
          throw std::runtime_error("build check: patterns differ");
        for (long e = 0; e < a[i]->size(); ++e) worst = std::max(worst, std::fabs(a[i]->data[e] - b[i]->data[e]));
        ++n;
      }
    };
    // device builds first (B1 may be device-only), then B1 in host memory for the host builds
    auto check_from = [&](const SpinBlock& B1) {
      const int nrem = I.norb - (int)B1.sites.size() - 1;
      BlockPtr d[2];
      std::shared_ptr<CompOpsAhead> da;
      for (int comp = 0; comp < 2; ++comp) {
        BuildSpec sp;
        sp.normal = comp == 0;
        sp.comp = comp == 1;
        d[comp] = gb.sum_block(I, B1, *dot, target, nrem, sp, scr);
        if (comp == 0) da = gb.comp_ops(I, *d[0], B1, *dot, scr);
      }
      std::vector<SparseOp*> b1;
      for (const OpPtr& o : block_ops(B1)) b1.push_back(o.get());
      materialise_host(b1);
      for (int comp = 0; comp < 2; ++comp) {
        BuildSpec sp;
        sp.normal = comp == 0;
        sp.comp = comp == 1;
        BlockPtr h = build_sum_block(I, B1, *dot, target, nrem, sp, scr);
        cmp_lists(block_ops(*h), block_ops(*d[comp]));
        if (comp == 0) cmp_lists(comp_ops_list(*build_comp_ops(I, *h, B1, *dot, scr)), comp_ops_list(*da));
      }
    };
    check_from(*sys);
    // sys with every operator device-only (an identity rotation on the GPU)
    std::vector<std::vector<double>> rot(sys->si->nsec());
    std::vector<int> keep(sys->si->nsec());
    for (int q = 0; q < sys->si->nsec(); ++q) {
      const int d = sys->si->dim[q];
      keep[q] = d;
      rot[q].assign((size_t)d * d, 0.0);
      for (int i = 0; i < d; ++i) rot[q][(size_t)i * d + i] = 1.0;
    }
    Rotation R = make_rotation_info(*sys->si, rot, keep);
    GpuRotator dev_rot(ctx, true);
    BlockPtr dsys = renormalise(*sys, R, &dev_rot);
    for (const OpPtr& o : block_ops(*dsys))
      if (o->host_valid || !o->dev) throw std::runtime_error("build check: rotated block not device-only");
    check_from(*dsys);
    *maxdiff = worst;
    *nops = n;
    return 0;
  });
}

}  // extern "C"

extern "C" {
// Renormalisation of a Wilson-grown block sys(nsys sites) (x) dot with a
// random density matrix truncated to M, once with the host rotations and
// once with the GPU rotator; returns the largest relative difference of any
// rotated operator (stored and complementary) and the number compared.
int blkh_rotation_check(void* ints, int nsys, int M, int device, double* maxrel, long* nops) {
  return guarded([&]() {
    const Integrals& I = *static_cast<Integrals*>(ints);
    Screen scr;
    const Quantum target{I.nelec, 0, I.isym};
    std::vector<int> order(I.norb);
    for (int i = 0; i < I.norb; ++i) order[i] = i;
    BlockPtr sys = warmup_block(I, order, nsys, target, M, true, scr);
    BlockPtr dot = make_dot(I, nsys, scr);
    BuildSpec sl;
    sl.normal = true;
    sl.comp = false;
    BlockPtr L = build_sum_block(I, *sys, *dot, target, I.norb - nsys - 1, sl, scr);
    std::mt19937_64 rng(7);
    std::normal_distribution<double> gauss(0.0, 1.0);
    std::vector<std::vector<double>> rho(L->si->nsec());
    for (int q = 0; q < L->si->nsec(); ++q) {
      const int d = L->si->dim[q];
      std::vector<double> G((size_t)d * d);
      for (auto& x : G) x = gauss(rng);
      rho[q].assign((size_t)d * d, 0.0);
      dgemm(false, true, d, d, d, 1.0, G.data(), d, G.data(), d, 0.0, rho[q].data(), d);
    }
    std::vector<std::vector<double>> rot;
    std::vector<int> mkeep;
    make_rotation(*L->si, rho, M, rot, mkeep);
    Rotation R = make_rotation_info(*L->si, rot, mkeep);
    GpuContext ctx(device);
    GpuRotator gpu(ctx);
    BlockPtr a = renormalise(*L, R, nullptr);
    add_rotated_comp_ops(I, *L, *sys, *dot, scr, R, *a, nullptr);
    BlockPtr b = renormalise(*L, R, &gpu);
    add_rotated_comp_ops(I, *L, *sys, *dot, scr, R, *b, &gpu);
    double worst = 0.0;
    long n = 0;
    auto cmp = [&](const OpPtr& x, const OpPtr& y) {
      if (!x && !y) return;
      if (!x || !y || x->data.size() != y->data.size()) throw std::runtime_error("rotation check: operator sets differ");
      double num = 0.0, den = 0.0;
      for (size_t e = 0; e < x->data.size(); ++e) {
        num += (x->data[e] - y->data[e]) * (x->data[e] - y->data[e]);
        den += x->data[e] * x->data[e];
      }
      if (den > 0.0) worst = std::max(worst, std::sqrt(num / den));
      ++n;
    };
    cmp(a->ident, b->ident);
    cmp(a->ham, b->ham);
    for (auto m : {&SpinBlock::cre, &SpinBlock::ccd, &SpinBlock::cd, &SpinBlock::cc, &SpinBlock::cdc, &SpinBlock::ddc})
      for (size_t i = 0; i < ((*a).*m).size(); ++i) cmp(((*a).*m)[i], ((*b).*m)[i]);
    *maxrel = worst;
    *nops = n;
    return 0;
  });
}

// SU(2) coefficients (doubled spins), exposed for the closed-form checks
double blkh_cg(int j1, int m1, int j2, int m2, int J, int M) { return cg(j1, m1, j2, m2, J, M); }
double blkh_sixj(int a, int b, int c, int d, int e, int f) { return sixj(a, b, c, d, e, f); }
double blkh_ninej(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
  return ninej(a, b, c, d, e, f, g, h, i);
}
double blkh_dagger_phase(int J2, int Jp2, int k2) { return dagger_phase(J2, Jp2, k2); }
}
