#include "sweep.h"
#include "checkpoint.h"
#include "gpu_sigma.h"
#include "guess.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <map>
#include <numeric>
#include <omp.h>
#include <random>
#include <sstream>
#include <thread>

namespace blk {

namespace {
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Hartree-Fock-like occupation from the orbital diagonal (hf_occ integral):
// fill the lowest h_ii first, sharing electrons equally among degenerate levels.
std::vector<double> hf_occupation(const Integrals& I, int nelec) {
  const int n = I.norb;
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return I.t(a, a) < I.t(b, b); });
  std::vector<double> occ(n, 0.0);
  double left = nelec;
  for (size_t g0 = 0; g0 < ord.size() && left > 0;) {
    size_t g1 = g0 + 1;
    while (g1 < ord.size() && std::fabs(I.t(ord[g1], ord[g1]) - I.t(ord[g0], ord[g0])) < 1e-8) ++g1;
    double cap = 2.0 * (g1 - g0);
    double put = std::min(cap, left);
    for (size_t k = g0; k < g1; ++k) occ[ord[k]] = put / (g1 - g0);
    left -= put;
    g0 = g1;
  }
  return occ;
}

// Wilson-type truncation of a composite block: lowest eigenstates of its
// Hamiltonian, with the kept multiplets spread over particle numbers by a
// Gaussian quota centred on the block's Hartree-Fock occupation (width sig),
// so the environment offers the superblock every nearby particle number.
void wilson_rotation(const SpinBlock& B, double nbar, double sig, int M, std::vector<std::vector<double>>& rot,
                     std::vector<int>& mkeep) {
  const StateInfo& si = *B.si;
  const int ns = si.nsec();
  std::vector<std::vector<double>> evec(ns), eval(ns);
  struct Cand {
    double e;
    int q, k;
  };
  std::map<int, std::vector<Cand>> byn;
  for (int q = 0; q < ns; ++q) {
    int d = si.dim[q];
    int ib = B.ham->pat->find(q, q);
    evec[q].assign((size_t)d * d, 0.0);
    eval[q].assign(d, 0.0);
    if (ib >= 0) std::copy(B.ham->blk(ib), B.ham->blk(ib) + (size_t)d * d, evec[q].begin());
    syev(d, evec[q].data(), eval[q].data());
    for (int k = 0; k < d; ++k) byn[si.q[q].n].push_back({eval[q][k], q, k});
  }
  for (auto& kv : byn)
    std::stable_sort(kv.second.begin(), kv.second.end(), [](const Cand& a, const Cand& b) { return a.e < b.e; });
  // quotas: fill by descending weight, one state at a time proportionally
  std::vector<std::pair<double, int>> wn;
  for (auto& kv : byn) wn.push_back({std::exp(-0.5 * (kv.first - nbar) * (kv.first - nbar) / (sig * sig)), kv.first});
  std::map<int, int> taken;
  int total = 0;
  double wsum = 0.0;
  for (auto& x : wn) wsum += x.first;
  for (auto& x : wn) {
    int q = (int)std::floor(M * x.first / wsum);
    q = std::min<int>(q, (int)byn[x.second].size());
    taken[x.second] = q;
    total += q;
  }
  std::sort(wn.begin(), wn.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
    return a.first > b.first;
  });
  bool progress = true;
  while (total < M && progress) {
    progress = false;
    for (auto& x : wn) {
      if (total >= M) break;
      if (taken[x.second] < (int)byn[x.second].size()) {
        taken[x.second]++;
        total++;
        progress = true;
      }
    }
  }
  mkeep.assign(ns, 0);
  std::vector<std::vector<int>> picked(ns);
  for (auto& kv : byn)
    for (int i = 0; i < taken[kv.first]; ++i) picked[kv.second[i].q].push_back(kv.second[i].k);
  rot.assign(ns, {});
  for (int q = 0; q < ns; ++q) {
    std::sort(picked[q].begin(), picked[q].end());
    int d = si.dim[q], m = (int)picked[q].size();
    mkeep[q] = m;
    rot[q].assign((size_t)d * m, 0.0);
    for (int c = 0; c < m; ++c)
      for (int r = 0; r < d; ++r) rot[q][(size_t)r * m + c] = evec[q][(size_t)r * d + picked[q][c]];
  }
}
}  // namespace

double make_rotation(const StateInfo& si, const std::vector<std::vector<double>>& rho, int M,
                     std::vector<std::vector<double>>& rot, std::vector<int>& mkeep) {
  const int ns = si.nsec();
  std::vector<std::vector<double>> evec(ns), eval(ns);
  // weights sorted descending; ties in (sector, index) order (sort_weights)
  std::multimap<double, std::pair<int, int>> wmap;
  double total = 0.0;
  // sectors are independent: largest first over the threads
  std::vector<int> order(ns);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return si.dim[a] > si.dim[b]; });
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < ns; ++i) {
    const int q = order[i], d = si.dim[q];
    evec[q] = rho[q];
    eval[q].assign(d, 0.0);
    syev(d, evec[q].data(), eval[q].data());
  }
  for (int q = 0; q < ns; ++q) {
    int d = si.dim[q];
    for (int k = 0; k < d; ++k) {
      if (eval[q][k] < 1e-14) eval[q][k] = 0.0;
      wmap.insert({eval[q][k], {q, k}});
      total += eval[q][k];
    }
  }
  std::vector<std::vector<int>> picked(ns);
  double kept = 0.0;
  int taken = 0;
  for (auto it = wmap.rbegin(); it != wmap.rend() && taken < M; ++it, ++taken) {
    if (it->first <= 1e-13) continue;  // assign_matrix_by_dm keeps weights > 1e-13 only
    picked[it->second.first].push_back(it->second.second);
    kept += it->first;
  }
  rot.assign(ns, {});
  mkeep.assign(ns, 0);
  for (int q = 0; q < ns; ++q) {
    int d = si.dim[q], m = (int)picked[q].size();
    mkeep[q] = m;
    rot[q].assign((size_t)d * m, 0.0);
    for (int c = 0; c < m; ++c)
      for (int r = 0; r < d; ++r) rot[q][(size_t)r * m + c] = evec[q][(size_t)r * d + picked[q][c]];
  }
  return total - kept;
}

BlockPtr warmup_block(const Integrals& I, const std::vector<int>& order, int nsites, const Quantum& target, int M,
                      bool normal_ops, const Screen& scr) {
  const int k = I.norb;
  auto occ = hf_occupation(I, target.n);
  BlockPtr B = make_dot(I, order[0], scr);
  double nbar = occ[order[0]];
  BuildSpec sp;
  sp.normal = normal_ops;
  for (int s = 1; s < nsites; ++s) {
    const int site = order[s];
    nbar += occ[site];
    auto dot = make_dot(I, site, scr);
    sp.comp = false;
    BlockPtr C = build_sum_block(I, *B, *dot, target, k - (s + 1), sp, scr);
    std::vector<std::vector<double>> rot;
    std::vector<int> mkeep;
    wilson_rotation(*C, nbar, std::max(1.0, 0.5 * std::sqrt((double)(s + 1))), M, rot, mkeep);
    Rotation R = make_rotation_info(*C->si, rot, mkeep);
    BlockPtr nb = renormalise(*C, R);
    add_rotated_comp_ops(I, *C, *B, *dot, scr, R, *nb);
    B = nb;
  }
  return B;
}

SuperBlock snapshot_superblock(const Integrals& I, const Quantum& target, int nsys, int M, const Screen& scr) {
  const int k = I.norb;
  std::vector<int> fwd(k), bwd(k);
  std::iota(fwd.begin(), fwd.end(), 0);
  for (int i = 0; i < k; ++i) bwd[i] = k - 1 - i;
  // sys: sites 0..nsys-1, dots nsys and nsys+1, env: nsys+2..k-1
  BlockPtr sys = warmup_block(I, fwd, nsys, target, M, true, scr);
  BlockPtr env = warmup_block(I, bwd, k - nsys - 2, target, M, false, scr);
  BuildSpec sl;
  sl.normal = true;
  sl.comp = false;
  BuildSpec sr;
  sr.normal = false;
  sr.comp = true;
  BlockPtr L = build_sum_block(I, *sys, *make_dot(I, nsys, scr), target, k - nsys - 1, sl, scr);
  BlockPtr R = build_sum_block(I, *env, *make_dot(I, nsys + 1, scr), target, nsys + 1, sr, scr);
  return make_superblock(I, L, R, target);
}

namespace {
struct Solved {
  std::vector<std::vector<double>> b;  // wave solutions
  std::vector<double> e;               // energies
  DavidsonStats st;
  bool guessed = false;  // started from the previous step's wavefunction
};

// Solver::solve_wavefunction (solver.C:19-143): diagonal, guess, block Davidson
// on_ready: called once the sigma functor and the diagonal are set up, just
// before the Davidson iterations
Solved solve_superblock(const SuperBlock& sb, const DmrgParams& P, double tol, const SigmaFactory& make_sigma,
                        std::mt19937_64& rng, const std::vector<std::vector<double>>* guess, double ladder,
                        const std::function<void()>& on_ready = nullptr) {
  std::normal_distribution<double> gauss(0.0, 1.0);
  const long N = sb.wave->size;
  if (N < P.nroots) throw std::runtime_error("run_dmrg: superblock smaller than the number of roots");
  double t0 = now_s();
  auto hfun = make_sigma(sb);
  double t1 = now_s();
  std::vector<double> diag;
  if (!hfun->diagonal(diag)) diag = diagonal_h(sb);
  double t2 = now_s();
  // guess: lowest diagonal elements with a little seeded noise
  std::vector<long> ord(N);
  std::iota(ord.begin(), ord.end(), 0);
  std::partial_sort(ord.begin(), ord.begin() + P.nroots, ord.end(), [&](long a, long b) { return diag[a] < diag[b]; });
  Solved out;
  out.b.assign(P.nroots, std::vector<double>(N));
  for (int r = 0; r < P.nroots; ++r) {
    for (long x = 0; x < N; ++x) out.b[r][x] = 1e-3 * gauss(rng);
    out.b[r][ord[r]] += 1.0;
  }
  // the previous step's wavefunction carried into this basis, when every
  // root kept most of its norm (guess_wavefunctions, guess_wavefunction.C:275-400)
  if (guess && (int)guess->size() == P.nroots) {
    bool ok = true;
    for (const auto& g : *guess) ok = ok && (long)g.size() == N && ddot(g, g) > 1e-2;
    if (ok) {
      out.b = *guess;
      out.guessed = true;
    }
  }
  out.e = diag;
  if (on_ready) on_ready();
  // with several roots the subspace never deflates below nroots vectors and
  // grows to nroots + 20 (the nroots keyword, input.C:882-885)
  const int dmin = std::max(P.deflation_min, P.nroots);
  const int dmax = P.nroots > 1 ? std::max(P.deflation_max, P.nroots + 20) : std::max(P.deflation_max, dmin + 1);
  if (!hfun->davidson(out.b, out.e, tol, true, dmin, dmax, &out.st, P.davidson_max_iter, ladder, nullptr))
    block_davidson(out.b, out.e, tol, *hfun, true, dmin, dmax, &out.st, P.davidson_max_iter, ladder);
  out.e.resize(P.nroots);
  out.st.setup_seconds = t1 - t0;
  out.st.diag_seconds = t2 - t1;
  return out;
}

// reduced density matrix of the left (left=true) or right half, state
// averaged (makedensitymatrix, density.C:27-90), plus seeded random noise
std::vector<std::vector<double>> density_matrix(const SuperBlock& sb, const Solved& sol,
                                                const std::vector<double>& weights, bool left, double noise,
                                                std::mt19937_64& rng) {
  std::normal_distribution<double> gauss(0.0, 1.0);
  const StateInfo& si = left ? *sb.wave->L : *sb.wave->R;
  std::vector<std::vector<double>> rho(si.nsec());
  for (int q = 0; q < si.nsec(); ++q) rho[q].assign((size_t)si.dim[q] * si.dim[q], 0.0);
  for (size_t r = 0; r < sol.b.size(); ++r)
    for (const auto& bk : sb.wave->blocks) {
      const double* C = sol.b[r].data() + bk.off;
      if (left)
        dgemm(false, true, bk.rows, bk.rows, bk.cols, weights[r], C, bk.cols, C, bk.cols, 1.0, rho[bk.l].data(), bk.rows);
      else
        dgemm(true, false, bk.cols, bk.cols, bk.rows, weights[r], C, bk.cols, C, bk.cols, 1.0, rho[bk.r].data(), bk.cols);
    }
  if (noise > 0.0) {
    // one random state over every sector, normalised, added with total
    // weight `noise` (add_twodot_noise, density.C:92-156: random
    // wavefunctions scaled to unit norm, MultiplyProduct with weight noise):
    // a sector's share grows with its size, so noise does not spread the
    // kept states evenly over the (many, mostly unphysical) small sectors
    int nz = 0;
    for (int q = 0; q < si.nsec(); ++q) nz += si.dim[q] > 0;
    for (int q = 0; q < si.nsec(); ++q) {
      int d = si.dim[q];
      if (d == 0) continue;
      std::vector<double> G((size_t)d * d);
      double z = 0.0;
      for (auto& x : G) {
        x = gauss(rng);
        z += x * x;
      }
      dgemm(false, true, d, d, d, noise / (nz * z), G.data(), d, G.data(), d, 1.0, rho[q].data(), d);
    }
  }
  return rho;
}

// seconds spent in the phases of truncate_block (eigen-decomposition,
// rotation of the stored operators, complementary operators built rotated)
struct RenormTimes {
  double eig = 0, rotate = 0, comp = 0, build = 0;  // build: complementary operators built by the builder
};
RenormTimes g_rt;

// B = B1 (x) dot with the builder of the run (the GPU's) or the host's
BlockPtr sum_block(const DmrgParams& P, const Integrals& I, const SpinBlock& B1, const SpinBlock& dot,
                   const Quantum& target, int nrem, const BuildSpec& spec, const Screen& scr) {
  return P.builder ? P.builder->sum_block(I, B1, dot, target, nrem, spec, scr)
                   : build_sum_block(I, B1, dot, target, nrem, spec, scr);
}

BlockPtr truncate_block(const Integrals& I, const SpinBlock& B, const SpinBlock& B1, const SpinBlock& dot,
                        const std::vector<std::vector<double>>& rho, int M, const Screen& scr, double* disc,
                        const OpRotator* rotator, const CompOpsAhead* ahead = nullptr,
                        const BlockBuilder* builder = nullptr) {
  std::shared_ptr<CompOpsAhead> built;
  if (!B.has_comp && !ahead && builder) {
    const double tb = now_s();
    built = builder->comp_ops(I, B, B1, dot, scr);
    ahead = built.get();
    g_rt.build += now_s() - tb;
  }
  std::vector<std::vector<double>> rot;
  std::vector<int> mkeep;
  double t0 = now_s();
  *disc = make_rotation(*B.si, rho, M, rot, mkeep);
  Rotation Rot = make_rotation_info(*B.si, rot, mkeep);
  double t1 = now_s();
  BlockPtr nb = renormalise(B, Rot, rotator);
  double t2 = now_s();
  if (!B.has_comp) {
    if (ahead)
      add_rotated_comp_ops(*ahead, Rot, *nb, rotator);
    else
      add_rotated_comp_ops(I, B, B1, dot, scr, Rot, *nb, rotator);
  }
  g_rt.eig += t1 - t0;
  g_rt.rotate += t2 - t1;
  g_rt.comp += now_s() - t2;
  auto rec = std::make_shared<RotRecord>();
  rec->composite = B.si;
  rec->kept = Rot.ns;
  rec->oldidx = Rot.oldidx;
  rec->rot = std::move(rot);
  nb->rotrec = rec;
  return nb;
}

// Host work run on a thread of its own beside the Davidson solve; the
// destructor waits for it (an exception of the solve leaves no thread behind).
class HostAhead {
 public:
  template <class F>
  void start(F f) {
    th_ = std::thread([this, f]() {
      try {
        f();
      } catch (...) {
        err_ = std::current_exception();
      }
    });
  }
  // wait; rethrows what the work threw
  void join() {
    if (th_.joinable()) th_.join();
    if (err_) {
      std::exception_ptr e = err_;
      err_ = nullptr;
      std::rethrow_exception(e);
    }
  }
  ~HostAhead() {
    if (th_.joinable()) th_.join();
  }

 private:
  std::thread th_;
  std::exception_ptr err_;
};

// target of a partial chain of kp of the k orbitals (infinite-system growth)
Quantum partial_target(const Quantum& t, int kp, int k) {
  if (kp >= k) return t;
  int n = (int)std::lround((double)t.n * kp / k);
  if ((n & 1) != (t.s2 & 1)) n += (n < t.n) ? 1 : -1;
  n = std::max(n, t.s2);
  n = std::min(n, 2 * kp - t.s2);
  return Quantum{n, t.s2, t.irrep};
}
}  // namespace

// Infinite-system growth from both ends to the middle: every step solves the
// partial superblock sys (x) dot | dot (x) env and truncates both halves by
// their density matrices.  Fills fwd[0..kh-1] and bwd[k-kh..k-1]; returns the
// last (full-chain, mid-point) superblock.
SuperBlock infinite_warmup(const Integrals& I, const DmrgParams& P, const SigmaFactory& make_sigma,
                           std::vector<BlockPtr>& fwd, std::vector<BlockPtr>& bwd, std::mt19937_64& rng,
                           std::vector<std::vector<double>>* last_b = nullptr) {
  const int k = I.norb;
  const Quantum target{P.nelec, P.twoS, P.irrep};
  Screen scr;
  std::vector<double> weights = P.weights;
  if ((int)weights.size() != P.nroots) weights.assign(P.nroots, 1.0 / P.nroots);
  const int M = P.warmup_M > 0 ? P.warmup_M : P.schedule[0].M;
  const double tol = P.schedule[0].davidson_tol, noise = P.schedule[0].noise;
  fwd[0] = make_dot(I, 0, scr);
  bwd[k - 1] = make_dot(I, k - 1, scr);
  SuperBlock last;
  for (int s = 0; 2 * (s + 2) <= k; ++s) {
    const int kp = 2 * (s + 2);
    const Quantum tp = partial_target(target, kp, k);
    auto dot1 = make_dot(I, s + 1, scr);
    auto dot2 = make_dot(I, k - 2 - s, scr);
    BuildSpec sl;
    sl.normal = true;
    sl.comp = false;
    // R carries both sets: its complementary operators serve this superblock,
    // its normal ones the backward sweep that will use it as a system block
    BuildSpec sr;
    sr.normal = true;
    sr.comp = true;
    BlockPtr L = sum_block(P, I, *fwd[s], *dot1, target, k - (s + 2), sl, scr);
    BlockPtr R = sum_block(P, I, *bwd[k - 1 - s], *dot2, target, k - (s + 2), sr, scr);
    SuperBlock sb = make_superblock(I, L, R, tp);
    if (sb.wave->size < P.nroots) throw std::runtime_error("infinite_warmup: partial superblock too small");
    Solved sol = solve_superblock(sb, P, tol, make_sigma, rng, nullptr, P.ladder_from_sweep <= 0 ? P.ladder : 0.0);
    if (P.verbose > 1)
      std::fprintf(stderr, "  warmup step %d  sites %d  N=%d  wave=%ld  E0=%.12f iters=%d\n", s, kp, tp.n,
                   sb.wave->size, sol.e[0], sol.st.iterations);
    double d1 = 0, d2 = 0;
    auto rhoL = density_matrix(sb, sol, weights, true, noise, rng);
    auto rhoR = density_matrix(sb, sol, weights, false, noise, rng);
    fwd[s + 1] = truncate_block(I, *L, *fwd[s], *dot1, rhoL, M, scr, &d1, P.rotator.get(), nullptr, P.builder.get());
    BlockPtr Rn =
        truncate_block(I, *R, *bwd[k - 1 - s], *dot2, rhoR, M, scr, &d2, P.rotator.get(), nullptr, P.builder.get());
    bwd[k - 2 - s] = Rn;
    if (kp == k || 2 * (s + 3) > k) {
      last = sb;
      if (last_b) *last_b = sol.b;
    }
  }
  return last;
}

DmrgResult run_dmrg(const Integrals& I, const DmrgParams& P, const SigmaFactory& make_sigma) {
  const int k = I.norb;
  if (k < 4) throw std::runtime_error("run_dmrg: need at least 4 orbitals for two-site sweeps");
  if (P.schedule.empty()) throw std::runtime_error("run_dmrg: empty schedule");
  const Quantum target{P.nelec, P.twoS, P.irrep};
  Screen scr;
  std::vector<double> weights = P.weights;
  if ((int)weights.size() != P.nroots) weights.assign(P.nroots, 1.0 / P.nroots);
  std::mt19937_64 rng(P.seed);
  DmrgResult res;

  std::vector<BlockPtr> fwd(k), bwd(k);  // fwd[l]: sites 0..l ; bwd[l]: sites l..k-1
  double tw = now_s();
  // the previous step: its superblock layout and wavefunctions, its
  // environment block and the block it made (for the next guess)
  struct Prev {
    std::shared_ptr<const WaveLayout> wave;
    std::vector<std::vector<double>> b;
    BlockPtr env, nsys;
  } prev;
  double prev_full = 1e300;
  int start_sweep = 0;
  if (!P.resume.empty()) {
    // continue from a saved sweep state (blocks, last wavefunction, records)
    SweepState st = load_sweep_state(P.resume);
    if ((int)st.fwd.size() != k || (int)st.bwd.size() != k)
      throw std::runtime_error("run_dmrg: the saved state is for another number of orbitals");
    if (st.next_sweep > P.max_sweeps)
      throw std::runtime_error("run_dmrg: the saved state is past max_sweeps (" + std::to_string(st.next_sweep) +
                               " sweeps done)");
    if ((int)st.b.size() != P.nroots || (int)st.res.energies.size() != P.nroots)
      throw std::runtime_error("run_dmrg: the saved state has another number of roots");
    if (!st.wave || !(st.wave->target == target))
      throw std::runtime_error("run_dmrg: the saved state is for another target (N, 2S, irrep)");
    fwd = std::move(st.fwd);
    bwd = std::move(st.bwd);
    if (P.resident_block)
      for (const auto* side : {&fwd, &bwd})
        for (const BlockPtr& b : *side)
          if (b) P.resident_block(*b);
    prev.wave = st.wave;
    prev.b = std::move(st.b);
    res = std::move(st.res);
    prev_full = st.prev_full;
    start_sweep = st.next_sweep;
    std::istringstream is(st.rng);
    is >> rng;
    if (P.verbose) std::fprintf(stderr, "resumed from %s at sweep %d\n", P.resume.c_str(), start_sweep);
  } else {
    SuperBlock last = infinite_warmup(I, P, make_sigma, fwd, bwd, rng, &prev.b);
    prev.wave = last.wave;
  }
  if (!P.transform_guess) prev.b.clear();
  int fwd_max = 0;
  while (fwd_max + 1 < k && fwd[fwd_max + 1]) ++fwd_max;
  const int l0 = fwd_max - 1;  // first forward step (env = bwd[l0 + 3] exists)
  if (P.verbose && P.resume.empty())
    std::fprintf(stderr, "warmup: infinite-system growth to %d|%d sites, %.2fs\n", fwd_max + 1, k - fwd_max - 1,
                 now_s() - tw);

  for (int sweep = start_sweep; sweep < P.max_sweeps; ++sweep) {
    const ScheduleEntry* S = &P.schedule[0];
    for (const auto& e : P.schedule)
      if (sweep >= e.start_sweep) S = &e;
    const bool forward = (sweep % 2 == 0);
    SweepRecord rec;
    rec.sweep = sweep;
    rec.forward = forward;
    rec.M = S->M;
    rec.energies.assign(P.nroots, 1e300);
    double t0 = now_s();
    const int first = (sweep == 0) ? std::max(0, l0) : 0;
    const int nsteps = k - 3;
    // (not with a builder of its own: the device builds share the solve's context)
    const bool use_ahead = !P.builder && (P.host_ahead > 0 || (P.host_ahead < 0 && P.rotator));
    BlockPtr next_R;  // the next step's R, built during this step's solve
    const SpinBlock* next_env = nullptr;
    int next_dot = -1;
    for (int step = first; step < nsteps; ++step) {
      // forward: sys = 0..l, dots l+1, l+2, env l+3..; backward mirrored
      int l = forward ? step : (k - 1 - step);
      int d1 = forward ? l + 1 : l - 1;
      int d2 = forward ? l + 2 : l - 2;
      BlockPtr sys = forward ? fwd[l] : bwd[l];
      BlockPtr env = forward ? bwd[l + 3] : fwd[l - 3];
      if (!sys || !env) throw std::runtime_error("run_dmrg: missing block");
      auto dot1 = make_dot(I, d1, scr);
      auto dot2 = make_dot(I, d2, scr);
      int nL = (int)sys->sites.size() + 1, nR = (int)env->sites.size() + 1;
      BuildSpec sl;
      sl.normal = true;
      sl.comp = false;
      BuildSpec sr;
      sr.normal = false;
      sr.comp = true;
      double ts = now_s();
      BlockPtr L = sum_block(P, I, *sys, *dot1, target, nR, sl, scr);
      BlockPtr R = (next_R && next_env == env.get() && next_dot == d2)
                       ? next_R
                       : sum_block(P, I, *env, *dot2, target, nL, sr, scr);
      next_R.reset();
      next_env = nullptr;
      SuperBlock sb = make_superblock(I, L, R, target);
      const long N = sb.wave->size;
      const double tbuild = now_s() - ts;
      rec.build_seconds += tbuild;
      // while the solve runs: L's complementary operators in its composite
      // basis (renormalisation input, handed to the rotator, which may upload
      // them beside the sigma kernels) and the next step's R = env' (x) dot
      // (started once the solve's own host set-up -- work list, operator
      // upload through the staging -- is done, so as not to compete with it)
      std::shared_ptr<CompOpsAhead> compL;
      HostAhead ahead;
      std::function<void()> start_ahead;
      if (use_ahead) {
        const int ln = forward ? l + 1 : l - 1;
        const int d2n = forward ? ln + 2 : ln - 2;
        BlockPtr envn = step + 1 < nsteps ? (forward ? bwd[ln + 3] : fwd[ln - 3]) : nullptr;
        start_ahead = [&, envn, d2n]() {
          ahead.start([&, envn, d2n]() {
            // leave two cores to the solve's own host thread
            omp_set_num_threads(std::max(1, omp_get_num_procs() - 2));
            if (!L->has_comp) {
              compL = build_comp_ops(I, *L, *sys, *dot1, scr);
              if (P.rotator) P.rotator->preload(comp_ops_list(*compL));
            }
            if (envn) {
              auto dn = make_dot(I, d2n, scr);
              next_R = build_sum_block(I, *envn, *dn, target, nL + 1, sr, scr);
              next_env = envn.get();
              next_dot = d2n;
              if (P.preload_block) P.preload_block(*next_R);
            }
          });
        };
      }
      ts = now_s();
      // guess: the previous wavefunction in this basis (same superblock after
      // the warm-up, the halves swapped at a turn, else the two-site transform)
      std::vector<std::vector<double>> guess;
      bool have_guess = false;
      if (!prev.b.empty() && prev.wave) {
        const WaveLayout& w0 = *prev.wave;
        const WaveLayout& w1 = *sb.wave;
        if (sweep == 0 && step == first) {
          if (same_sectors(*w0.L, *w1.L) && same_sectors(*w0.R, *w1.R) && w0.target == w1.target) {
            guess = prev.b;
            have_guess = true;
          }
        } else if (step == first) {
          have_guess = guess_transpose(w0, prev.b, w1, guess);
        } else if (prev.nsys && prev.nsys->rotrec && prev.env && prev.env->rotrec) {
          have_guess = guess_transform(w0, prev.b, *prev.nsys->rotrec, *prev.env->rotrec, w1, guess);
        }
      }
      rec.guess_seconds += now_s() - ts;
      Solved sol = solve_superblock(sb, P, S->davidson_tol, make_sigma, rng, have_guess ? &guess : nullptr,
                                    sweep >= P.ladder_from_sweep ? P.ladder : 0.0, start_ahead);
      rec.guessed_steps += sol.guessed;
      rec.davidson_seconds += now_s() - ts - sol.st.sigma_seconds;
      rec.setup_seconds += sol.st.setup_seconds;
      rec.diag_seconds += sol.st.diag_seconds;
      ts = now_s();
      rec.sigma_seconds += sol.st.sigma_seconds;
      rec.sigma_calls += sol.st.sigma_calls;
      rec.sigma_calls_split += sol.st.sigma_calls_split;
      rec.sigma_flops += sol.st.sigma_calls * sigma_flops(sb);
      for (int r = 0; r < P.nroots; ++r) rec.energies[r] = std::min(rec.energies[r], sol.e[r]);
      const RenormTimes rt0 = g_rt;
      double twait = now_s();
      ahead.join();
      twait = now_s() - twait;
      double trho = now_s();
      auto rho = density_matrix(sb, sol, weights, true, S->noise, rng);
      trho = now_s() - trho;
      double disc = 0.0;
      BlockPtr nsys = truncate_block(I, *L, *sys, *dot1, rho, S->M, scr, &disc, P.rotator.get(), compL.get(),
                                     P.builder.get());
      compL.reset();
      rec.max_discarded = std::max(rec.max_discarded, disc);
      rec.renorm_seconds += now_s() - ts;
      if (forward)
        fwd[l + 1] = nsys;
      else
        bwd[l - 1] = nsys;
      if (P.transform_guess) {
        prev.wave = sb.wave;
        prev.b = std::move(sol.b);
        prev.env = env;
        prev.nsys = nsys;
      }
      if (P.verbose > 1)
        std::fprintf(stderr,
                     "  sweep %d step %d  N=%ld secL=%d secR=%d E0=%.12f  iters=%d%s  disc=%.2e  renorm: wait %.3f "
                     "rho %.3f eig %.3f rotate %.3f comp %.3f (built %.3f); build %.3f\n",
                     sweep, step, N, L->si->nsec(), R->si->nsec(), sol.e[0], sol.st.iterations,
                     sol.guessed ? " (guess)" : "", disc, twait, trho, g_rt.eig - rt0.eig, g_rt.rotate - rt0.rotate,
                     g_rt.comp - rt0.comp, g_rt.build - rt0.build, tbuild);
    }
    // the sweep ends at the far edge: the last block is the single-site dot
    if (forward)
      bwd[k - 1] = make_dot(I, k - 1, scr);
    else
      fwd[0] = make_dot(I, 0, scr);
    rec.seconds = now_s() - t0;
    res.sweeps.push_back(rec);
    if (P.verbose > 1) {
      const SetupTimes& st = gpu_setup_times();
      const BuildTimes& bt = gpu_build_times();
      std::fprintf(stderr,
                   "  sigma set-up so far: compile %.2f upload %.2f plan %.2f; renorm so far: eig %.2f rotate %.2f comp "
                   "%.2f (build %.2f); device builds so far: record %.2f execute %.2f [grouping %.2f, device %.2f, %.3g "
                   "items], host copies %.2f s (%.3g B); rotations: pairs %.2f in %.2f plan %.2f run %.2f out %.2f\n",
                   st.compile, st.upload, st.plan, g_rt.eig, g_rt.rotate, g_rt.comp, comp_build_seconds(), bt.record,
                   bt.execute, bt.plan, bt.apply, bt.items, bt.download, bt.download_bytes, st.rot_pairs, st.rot_in,
                   st.rot_plan, st.rot_run, st.rot_out);
    }
    if (P.verbose)
      std::fprintf(stderr,
                   "sweep %d %s M=%d E=%.12f disc=%.2e t=%.2fs (build %.2f, sigma %.2f in %ld calls, davidson %.2f "
                   "[setup %.2f, diag %.2f, guess %.2f for %ld steps], renorm %.2f)\n",
                   sweep, forward ? "fwd" : "bwd", S->M, rec.energies[0], rec.max_discarded, rec.seconds,
                   rec.build_seconds, rec.sigma_seconds, rec.sigma_calls, rec.davidson_seconds, rec.setup_seconds,
                   rec.diag_seconds, rec.guess_seconds, rec.guessed_steps, rec.renorm_seconds);
    res.energies = rec.energies;
    // convergence between consecutive full sweeps once the schedule is exhausted
    if (!forward) {
      double e = rec.energies[0];
      bool last_entry = sweep >= P.schedule.back().start_sweep;
      if (last_entry && std::fabs(e - prev_full) < P.sweep_tol) res.converged = true;
      else prev_full = e;
    }
    if (!P.checkpoint.empty()) {
      SweepState st;
      st.next_sweep = sweep + 1;
      st.fwd = fwd;
      st.bwd = bwd;
      st.wave = prev.wave;
      st.b = prev.b;
      st.res = res;
      st.prev_full = prev_full;
      std::ostringstream os;
      os << rng;
      st.rng = os.str();
      save_sweep_state(P.checkpoint, st);
    }
    if (res.converged) break;
  }
  return res;
}

SuperBlock dmrg_snapshot(const Integrals& I, const DmrgParams& P, const SigmaFactory& make_sigma) {
  const int k = I.norb;
  std::vector<BlockPtr> fwd(k), bwd(k);
  std::mt19937_64 rng(P.seed);
  return infinite_warmup(I, P, make_sigma, fwd, bwd, rng);
}

int write_sweep_records(const DmrgResult& res, double* out, int max_rows) {
  const int n = std::min<int>((int)res.sweeps.size(), std::max(0, max_rows));
  for (int i = 0; i < n; ++i) {
    const auto& s = res.sweeps[i];
    double* o = out + kSweepCols * i;
    o[0] = s.sweep;
    o[1] = s.forward;
    o[2] = s.M;
    o[3] = s.energies[0];
    o[4] = s.max_discarded;
    o[5] = s.seconds;
    o[6] = s.sigma_seconds;
    o[7] = s.sigma_flops;
    o[8] = (double)s.sigma_calls;
    o[9] = (double)s.sigma_calls_split;
    o[10] = s.build_seconds;
    o[11] = s.davidson_seconds;
    o[12] = s.renorm_seconds;
    o[13] = (double)s.guessed_steps;
  }
  return n;
}

}  // namespace blk
