// On-disk sweep state for checkpoint and resume: the renormalised blocks of
// both sweep directions with all their operators and rotations, the last
// solved superblock's layout and wavefunctions, the sweep records and the
// random generator.  It plays the part of Block's per-block files
// (SpinBlock::store / restore, save_load_block.C:23-91), its rotation-matrix
// files (rotationmat.C:21-60) and its sweep state (sweep_params.C:156-185),
// in one binary file of this project's own format (not Boost archives).
#pragma once
#include <string>

#include "sweep.h"

namespace blk {

struct SweepState {
  int next_sweep = 0;                      // first sweep a resumed run performs
  std::vector<BlockPtr> fwd, bwd;          // fwd[l]: sites 0..l, bwd[l]: sites l..k-1 (null where absent)
  std::shared_ptr<const WaveLayout> wave;  // last solved superblock
  std::vector<std::vector<double>> b;      // its wavefunctions (one per root)
  DmrgResult res;
  double prev_full = 1e300;                // last full-sweep energy (convergence test)
  std::string rng;                         // generator state (std::mt19937_64 text form)
};

void save_sweep_state(const std::string& path, const SweepState& s);
SweepState load_sweep_state(const std::string& path);

}  // namespace blk
