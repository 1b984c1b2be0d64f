// Superblock L (x) R of the two-site algorithm and the term list of one
// H.C product.  The term list is the flattened form of SpinBlock::multiplyH
// (spinblock.C:722-789): every entry is coef * [a(L) (x) b(R)]^0 applied to the
// wavefunction, exactly what one opxop functor call hands to the two-operator
// TensorMultiply (operatorfunctions.C:485-537).
#pragma once
#include "spinblock.h"

namespace blk {

struct SigmaTerm {
  OpView a, b;      // operator on L, operator on R
  double coef = 1;  // scalar in front of [a (x) b]^0
  int kind = 0;     // 0 H/identity, 1 CCDcomp x C, 2 CD x CDcomp, 3 CC x DDcomp
};

struct SuperBlock {
  BlockPtr L, R;  // L is the loop block (carries CD/CC), R carries the complements
  Quantum target;
  double ecore = 0.0;
  std::shared_ptr<const WaveLayout> wave;
  std::vector<SigmaTerm> terms;
};

// L = sys (x) dot1 with normal operators, R = env (x) dot2 with complementary
// operators, as makeSystemEnvironmentBigBlocks arranges them (sweep.C:91-178).
SuperBlock make_superblock(const Integrals& I, BlockPtr L, BlockPtr R, const Quantum& target);
void build_sigma_terms(const Integrals& I, SuperBlock& sb);

// diag(H_eff) in wavefunction element order (SpinBlock::diagonalH,
// spinblock.C:855-899).
std::vector<double> diagonal_h(const SuperBlock& sb);

// Algorithmic flop count of one H.C over the term list, GEMM pairs
// 2*(dl*dl'*dr' + dl*dr'*dr) with identity partners counted once.
double sigma_flops(const SuperBlock& sb);

}  // namespace blk
