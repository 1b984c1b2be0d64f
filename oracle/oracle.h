// CPU oracle for the Davidson sigma path.  TEST INFRASTRUCTURE ONLY: nothing
// in the product (block_amd/) may call, link or load this code; only tests/,
// the smoke check and bench.py's cpu_baseline leg use it, and only to check
// results against or as the reported CPU baseline.
//
//  * oracle_multiply_h  -- direct restatement of SpinBlock::multiplyH
//    (spinblock.C:722-789) looping over the operator arrays the way the opxop
//    functors do (opxop.C:155-285) and applying each pair with the
//    two-operator TensorMultiply loop nest (operatorfunctions.C:485-537):
//    m = A.C (beta=0) then v += f m B^T.  Naive loops, its own GEMM.
//  * oracle_diagonal    -- SpinBlock::diagonalH (spinblock.C:855-899).
//  * fci_*              -- determinant full CI (brute force) used to pin the
//    whole spin-adapted construction: an untruncated superblock must
//    reproduce the FCI spectrum in the target spin sector.
//
// Parity pinning of each function is described in the test files that use it.
#pragma once
#include <vector>

#include "../block_amd/host/superblock.h"

namespace oracle {

void oracle_multiply_h(const blk::SuperBlock& sb, const double* c, double* v, int nthreads = 1);
// Sector products through the host library's blocked DGEMM instead of the
// oracle's own loops (the timed CPU baseline; the parity tests keep the loops).
void oracle_set_blocked_gemm(bool on);
std::vector<double> oracle_diagonal(const blk::SuperBlock& sb);
// Algorithmic flops of one H.C (2 dl dl' dr' per A product, 2 dl dr' dr per
// B product, identity partners not multiplied, 2 per element for the
// identity-identity term), by brute force over all sector quadruples.
double oracle_sigma_flops(const blk::SuperBlock& sb);

struct FciResult {
  std::vector<double> energies;
  std::vector<double> s2;  // <S^2> of each root
};
// Lowest roots of H + penalty*S^2 in the (nelec, ms2=0) determinant space
// (restricted to one total irrep when irrep >= 0);
// energies are reported without the penalty.
FciResult fci_lowest(const blk::Integrals& I, int nelec, int nroots, double s2_penalty, double tol = 1e-10,
                     int irrep = -1);

}  // namespace oracle
