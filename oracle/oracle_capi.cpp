// extern "C" surface of the oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY:
// loaded by tests/, the smoke check and bench.py's cpu_baseline leg.
#include <chrono>
#include <cstring>
#include <string>

#include "../block_amd/host/gpu_sigma.h"
#include "../block_amd/host/sweep.h"
#include "oracle.h"

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
struct OracleFunctor : DavidsonFunctor {
  const SuperBlock& sb;
  int nthreads;
  OracleFunctor(const SuperBlock& s, int nt) : sb(s), nthreads(nt) {}
  void operator()(const double* c, double* v) override {
    std::fill(v, v + sb.wave->size, 0.0);
    oracle::oracle_multiply_h(sb, c, v, nthreads);
  }
};
}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// v = H c by the direct restatement (v overwritten)
int orc_sigma(void* sb, const double* c, double* v, int nthreads) {
  return guarded([&]() {
    auto* s = static_cast<SuperBlock*>(sb);
    std::fill(v, v + s->wave->size, 0.0);
    oracle::oracle_multiply_h(*s, c, v, nthreads);
    return 0;
  });
}
// same, timed (seconds returned through secs)
int orc_sigma_timed(void* sb, const double* c, double* v, int nthreads, double* secs) {
  auto t0 = std::chrono::steady_clock::now();
  int rc = orc_sigma(sb, c, v, nthreads);
  *secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}
// the same with the blocked host DGEMM for the sector products (the CPU
// baseline of the benchmark)
int orc_sigma_fast_timed(void* sb, const double* c, double* v, int nthreads, double* secs) {
  oracle::oracle_set_blocked_gemm(true);
  const int rc = orc_sigma_timed(sb, c, v, nthreads, secs);
  oracle::oracle_set_blocked_gemm(false);
  return rc;
}
double orc_sigma_flops(void* sb) { return oracle::oracle_sigma_flops(*static_cast<SuperBlock*>(sb)); }
int orc_diag(void* sb, double* out) {
  return guarded([&]() {
    auto d = oracle::oracle_diagonal(*static_cast<SuperBlock*>(sb));
    std::memcpy(out, d.data(), sizeof(double) * d.size());
    return 0;
  });
}
// CPU execution of the compiled GPU work list (host arena, same offsets)
int orc_exec_worklist(void* sbp, const double* c, double* v) {
  return guarded([&]() {
    const SuperBlock& sb = *static_cast<SuperBlock*>(sbp);
    WorkList wl = compile_worklist(sb);
    std::vector<double> arena(wl.arena_size);
    for (const auto& kv : wl.arena_off) std::memcpy(arena.data() + kv.second, kv.first->data.data(), 8 * kv.first->data.size());
    const long N = sb.wave->size;
    std::fill(v, v + N, 0.0);
    std::vector<double> W;
    for (const auto& vb : wl.vb)
      for (int q = 0; q < vb.count; ++q) {
        const auto& p = wl.pairs[vb.first + q];
        // W = alpha_a op(A) C   (m x k2)
        W.assign((size_t)vb.m * p.k2, 0.0);
        for (int i = 0; i < vb.m; ++i)
          for (int s = 0; s < p.k1; ++s) {
            double a = p.a_off < 0 ? (i == s ? 1.0 : 0.0)
                                   : p.alpha_a * (p.a_trans ? arena[p.a_off + (size_t)s * vb.m + i]
                                                            : arena[p.a_off + (size_t)i * p.k1 + s]);
            if (a == 0.0) continue;
            for (int j = 0; j < p.k2; ++j) W[(size_t)i * p.k2 + j] += a * c[p.c_off + (size_t)s * p.k2 + j];
          }
        // V += alpha W B'
        for (int i = 0; i < vb.m; ++i)
          for (int j = 0; j < vb.n; ++j) {
            double s = 0.0;
            if (p.b_off < 0) {
              s = W[(size_t)i * p.k2 + j];
            } else {
              for (int t = 0; t < p.k2; ++t) {
                double b = p.b_trans ? arena[p.b_off + (size_t)j * p.k2 + t] : arena[p.b_off + (size_t)t * vb.n + j];
                s += W[(size_t)i * p.k2 + t] * b;
              }
            }
            v[vb.v_off + (size_t)i * vb.n + j] += p.alpha * s;
          }
      }
    return 0;
  });
}

// determinant FCI: lowest nroots of the given total spin sector (penalty on S^2)
int orc_fci(void* ints, int nelec, int nroots, double penalty, int irrep, double* energies, double* s2) {
  return guarded([&]() {
    auto r = oracle::fci_lowest(*static_cast<Integrals*>(ints), nelec, nroots, penalty, 1e-9, irrep);
    for (int i = 0; i < (int)r.energies.size(); ++i) {
      energies[i] = r.energies[i];
      s2[i] = r.s2[i];
    }
    return (int)r.energies.size();
  });
}

// untruncated superblock L = sites [0, nL), R = sites [nL, norb)
void* orc_exact_superblock(void* ints, int nelec, int twoS, int irrep, int nL) {
  try {
    const Integrals& I = *static_cast<Integrals*>(ints);
    const int n = I.norb;
    Screen scr;
    Quantum target{nelec, twoS, irrep};
    BuildSpec sp;
    BlockPtr L = make_dot(I, 0, scr);
    for (int l = 1; l < nL; ++l) L = build_sum_block(I, *L, *make_dot(I, l, scr), target, n - (l + 1), sp, scr);
    BlockPtr R = make_dot(I, n - 1, scr);
    for (int r = n - 2; r >= nL; --r) R = build_sum_block(I, *R, *make_dot(I, r, scr), target, r, sp, scr);
    return new SuperBlock(make_superblock(I, L, R, target));
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

// One Davidson solve (block_davidson, linear.C:179-385) with the oracle
// sigma and diagonal; arguments as blkh_davidson (host control flow).
int orc_davidson(void* sbp, int nroots, const double* guesses, int nlow, const double* lower, double tol,
                 int nthreads, double* energies, double* vecs, int* stats) {
  return guarded([&]() {
    const SuperBlock& sb = *static_cast<SuperBlock*>(sbp);
    const long N = sb.wave->size;
    OracleFunctor h(sb, nthreads);
    std::vector<std::vector<double>> b(nroots), low(nlow);
    for (int r = 0; r < nroots; ++r) b[r].assign(guesses + (size_t)r * N, guesses + (size_t)(r + 1) * N);
    for (int r = 0; r < nlow; ++r) low[r].assign(lower + (size_t)r * N, lower + (size_t)(r + 1) * N);
    std::vector<double> e = oracle::oracle_diagonal(sb);
    const int dmin = std::max(2, nroots);
    const int dmax = nroots > 1 ? std::max(20, nroots + 20) : std::max(20, dmin + 1);
    DavidsonStats st;
    block_davidson(b, e, tol, h, true, dmin, dmax, &st, 200, 0.0, nlow ? &low : nullptr);
    for (int r = 0; r < nroots; ++r) {
      energies[r] = e[r];
      if (vecs) std::copy(b[r].begin(), b[r].end(), vecs + (size_t)r * N);
    }
    stats[0] = st.iterations;
    stats[1] = st.sigma_calls;
    stats[2] = st.sigma_calls_split;
    return 0;
  });
}

// the sweep driver with the oracle sigma (CPU reference run); same outputs as
// blkh_dmrg_gpu (sweep_out rows of kSweepCols)
int orc_dmrg_cpu(void* ints, int nelec, int twoS, int irrep, int nroots, const double* weights, int n_sched,
                 const double* sched, int max_sweeps, double sweep_tol, int guess, const char* checkpoint,
                 const char* resume, int ahead, int nthreads, int verbose, double* energies_out, double* sweep_out) {
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
    P.transform_guess = guess != 0;
    if (checkpoint) P.checkpoint = checkpoint;
    if (resume) P.resume = resume;
    P.host_ahead = ahead;
    auto res = run_dmrg(I, P, [&](const SuperBlock& sb) {
      return std::unique_ptr<DavidsonFunctor>(new OracleFunctor(sb, nthreads));
    });
    for (int r = 0; r < nroots; ++r) energies_out[r] = res.energies[r];
    return write_sweep_records(res, sweep_out, max_sweeps);
  });
}

}  // extern "C"
