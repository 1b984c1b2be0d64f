// Two-site sweep driver: the caller of the accelerated path.  Restates
// Sweep::do_one / BlockAndDecimate (sweep.C:180-651), the environment warm-up,
// Solver::solve_wavefunction (solver.C:19-143) and the density-matrix
// truncation of makeRotateMatrix (rotationmat.C:60-347) with its own code; the
// Davidson sigma is supplied through a functor factory so the sweep calls the
// GPU multiply_h as a drop-in.
#pragma once
#include <functional>
#include <memory>
#include <string>

#include "davidson.h"

namespace blk {

struct ScheduleEntry {
  int start_sweep;  // first sweep (half-sweep counter) this entry applies to
  int M;
  double davidson_tol;  // on the squared residual norm, as linear.C:301
  double noise;
};

struct DmrgParams {
  int nelec = 0, twoS = 0, irrep = 0;
  int nroots = 1;
  std::vector<double> weights;  // state-average weights (default equal)
  std::vector<ScheduleEntry> schedule;
  int max_sweeps = 20;
  double sweep_tol = 1e-8;  // on the change of the lowest sweep energy between full sweeps
  int warmup_M = 0;         // 0: first schedule M
  int deflation_min = 2, deflation_max = 20;
  int davidson_max_iter = 200;
  double ladder = 0.0;  // > 0: split-precision sigma until the residual norm falls below this (FP64 after)
  int ladder_from_sweep = 0;  // first sweep that uses the ladder (0: also the warm-up)
  bool transform_guess = true;  // start each step from the previous wavefunction in the new basis
  std::shared_ptr<const OpRotator> rotator;  // operator rotations of the renormalisation (null: host threads)
  // composite blocks (sys (x) dot, their complementary operators) built
  // elsewhere, e.g. in HBM from device-resident blocks (null: host threads)
  std::shared_ptr<const BlockBuilder> builder;
  std::string checkpoint;  // file the sweep state is written to after every sweep (empty: none)
  std::string resume;      // sweep-state file to continue from instead of the warm-up (empty: none)
  // build host-side operators that do not depend on the current solve (the
  // new system block's complementary operators, the next step's environment
  // block) on a host thread during the Davidson solve: 1 on, 0 off, -1 on
  // when a rotator is set (GPU runs, whose solve leaves the host idle)
  int host_ahead = -1;
  // called on that thread with the next step's environment-side block once
  // it is built (the GPU runs upload its operators ahead of the next solve)
  std::function<void(const SpinBlock&)> preload_block;
  // called for every block read back by a resume (the GPU runs with device
  // builds move its operators into HBM, where a continuous run keeps them)
  std::function<void(const SpinBlock&)> resident_block;
  uint64_t seed = 12345;
  int verbose = 0;
};

struct SweepRecord {
  int sweep;
  bool forward;
  int M;
  std::vector<double> energies;  // lowest per root over the sweep
  double max_discarded = 0.0;
  double seconds = 0.0;
  double sigma_seconds = 0.0;
  long sigma_calls = 0;
  long sigma_calls_split = 0;
  double sigma_flops = 0.0;
  double build_seconds = 0.0;     // L and R block builds and the superblock term list
  double davidson_seconds = 0.0;  // solve minus its sigma calls (functor set-up, vectors, guesses)
  double renorm_seconds = 0.0;    // density matrix, rotation and renormalised operators
  double setup_seconds = 0.0;     // of davidson_seconds: sigma functor construction
  double diag_seconds = 0.0;      // of davidson_seconds: diagonal of H
  double guess_seconds = 0.0;     // of davidson_seconds: guess wavefunction transforms
  long guessed_steps = 0;         // steps that started from the transformed wavefunction
};


struct DmrgResult {
  std::vector<SweepRecord> sweeps;
  std::vector<double> energies;  // final per-root energies
  bool converged = false;
};

// Columns of the per-sweep records the C entry points write.
constexpr int kSweepCols = 14;
// sweep, forward, M, E0, max discarded weight, seconds, sigma seconds, sigma
// flops, sigma calls, split-precision sigma calls, build seconds, other
// Davidson seconds, renormalisation seconds, steps started from a guess.
// out holds max_rows rows; returns the number of rows written.
int write_sweep_records(const DmrgResult& res, double* out, int max_rows);

using SigmaFactory = std::function<std::unique_ptr<DavidsonFunctor>(const SuperBlock&)>;

DmrgResult run_dmrg(const Integrals& I, const DmrgParams& P, const SigmaFactory& make_sigma);

// Infinite-system growth of both halves to the middle of the chain (the
// warm-up of run_dmrg).  Returns the final mid-chain superblock, whose blocks
// are truncated by real density matrices: the benchmark's workload.
SuperBlock dmrg_snapshot(const Integrals& I, const DmrgParams& P, const SigmaFactory& make_sigma);

// Truncation of one block's density matrix to at most M multiplets
// (diagonalise_dm + sort_weights + assign_matrix_by_dm, rotationmat.C):
// rot[Q] gets the kept eigenvectors as columns, mkeep[Q] their number;
// returns the discarded weight.
double make_rotation(const StateInfo& si, const std::vector<std::vector<double>>& rho, int M,
                     std::vector<std::vector<double>>& rot, std::vector<int>& mkeep);

// Environment warm-up block sequence and mid-chain superblock snapshots used
// by the benchmark: blocks grown from one end with a Wilson-type truncation
// (lowest energy states of the block Hamiltonian, particle number held near
// its Hartree-Fock value).
BlockPtr warmup_block(const Integrals& I, const std::vector<int>& order, int nsites, const Quantum& target, int M,
                      bool normal_ops, const Screen& scr);
SuperBlock snapshot_superblock(const Integrals& I, const Quantum& target, int nsys, int M, const Screen& scr);

}  // namespace blk
