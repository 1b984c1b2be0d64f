// Guess wavefunctions for the next superblock of a sweep: the converged
// wavefunction of the previous two-site step carried into the new basis, so
// the Davidson solver starts next to the answer instead of from the diagonal.
// Restates GuessWave::transform_previous_wavefunction (guess_wavefunction.C:
// 524-636: left rotation, onedot_shufflesysdot, right expansion) and the
// half swap at the turn of a sweep (transpose_previous_wavefunction, :55-139).
#pragma once
#include "superblock.h"

namespace blk {

// Same collected sectors (quanta, dimensions and, for composite blocks, the
// (left, right) sub-block lists).
bool same_sectors(const StateInfo& a, const StateInfo& b);

// Previous step, superblock L0 (x) R0 with L0 = sys (x) dot1 and
// R0 = env (x) dot2; next step L1 = sys' (x) dot2, R1 = env'' (x) dot3, where
//   sys' = UL^T (sys (x) dot1)        (the block the previous step made)
//   env  = UE^T (env'' (x) dot3)      (the previous environment's rotation)
// c1 = (UL^T (x) 1) shuffle(c0) (1 (x) UE), dot2 moved from R to L with the
// fermion sign of passing env and the 6j recoupling of
// (a (c b)_CB)_J -> ((a b)_AB c)_J.  Exact when both rotations keep the
// wavefunction's support.  Returns false (out untouched) if the blocks do not
// have that structure.
bool guess_transform(const WaveLayout& w0, const std::vector<std::vector<double>>& c0, const RotRecord& UL,
                     const RotRecord& UE, const WaveLayout& w1, std::vector<std::vector<double>>& out);

// Turn of a sweep: L1 has the sectors of R0 and R1 those of L0; the halves
// swap with sign (-1)^(nL nR) (-1)^((SL + SR - J)/2).
bool guess_transpose(const WaveLayout& w0, const std::vector<std::vector<double>>& c0, const WaveLayout& w1,
                     std::vector<std::vector<double>>& out);

}  // namespace blk
