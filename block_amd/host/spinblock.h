// Blocks, their operator sets and the recursive construction of composite
// (block (x) dot) operators.  Restates SpinBlock (spinblock.h), the operator
// classes of Operators.C and the component choice of
// set_spinblock_components.C:331-800 with its own storage.
//
// Operator content of a block B (orbital indices are global):
//   cre[i]            C_i                   i in B                 (CRE)
//   cd[i,j,K]         [C_i D_j]^K           i>=j in B, K=0,1       (CRE_DES)
//   cc[i,j,K]         [C_i C_j]^K           i>=j in B              (CRE_CRE)
//   cdc[i,j,K]        Q^K_ij  (CDcomp)      i>=j not in B          (CRE_DESCOMP)
//   ddc[i,j,K]        P^K_ij  (DDcomp)      i>=j not in B          (DES_DESCOMP)
//   ccd[q]            X_q     (CCDcomp)     q not in B             (CRE_CRE_DESCOMP)
//   ham, ident
// with D_j = C_j^dagger-tensor and
//   Q^0_ij = sum_{kl in B} (2(ij|kl) - (il|kj)) [C_k D_l]^0
//   Q^1_ij = sum_{kl in B} sqrt(3) (il|kj) [C_k D_l]^1
//   P^K_ij = sum_{qs in B} w_K (iq|js) [D_s D_q]^K,  w_0 = -1/2, w_1 = sqrt(3)/2
//   X_q    = sum_{p in B} t_pq/2 C_p - sqrt(2) sum_{prs in B} (pq|rs) [C_p [C_r D_s]^0]^{1/2}
// so that for a split L|R of the orbitals
//   H = H_L + H_R + E_core
//     - sqrt2 sum_{q in R} ([X_q(L) D_q(R)] + [X_q(L)^dag C_q(R)])
//     - sqrt2 sum_{q in L} ([D_q(L) X_q(R)] + [C_q(L) X_q(R)^dag])
//     + sum_{i>=j in L,K} ([CD^K_ij Q^K_ij] + (i!=j)[CD^K_ij^dag Q^K_ij^dag])
//     + sum_{i>=j in L,K} (2-d_ij) ([CC^K_ij P^K_ij] + [CC^K_ij^dag P^K_ij^dag])
// (all products coupled to total spin 0), the term set of SpinBlock::multiplyH
// (spinblock.C:722-789) and of Ham::build.
#pragma once
#include <array>
#include <string>

#include "blk.h"

namespace blk {

struct Integrals {
  int norb = 0;
  int nelec = 0, ms2 = 0, isym = 0;  // from the FCIDUMP header (isym 0-based)
  std::vector<int> orbsym;           // 0-based irreps
  double ecore = 0.0;
  std::vector<double> h1;  // norb x norb
  std::vector<double> h2;  // norb^4, chemist (ij|kl)
  double t(int i, int j) const { return h1[(size_t)i * norb + j]; }
  double v(int i, int j, int k, int l) const {
    size_t n = norb;
    return h2[((i * n + j) * n + k) * n + l];
  }
  // orbital permutation: new index a holds old orbital perm[a]
  Integrals reordered(const std::vector<int>& perm) const;
};
// FCIDUMP reader/writer (the input side of input.C:1407-1560).
Integrals read_fcidump(const std::string& path);
void write_fcidump(const std::string& path, const Integrals& I);
// Fiedler-vector orbital ordering (fiedler.C:47-81): order[a] = the orbital
// placed at position a (Integrals::reordered's convention).  The matrix form
// takes any symmetric n x n interaction matrix.
std::vector<int> fiedler_order(const Integrals& I);
std::vector<int> fiedler_order_matrix(int n, const double* m);

struct SpinBlock;
using BlockPtr = std::shared_ptr<SpinBlock>;

// Rotation that made a renormalised block from its composite (block (x) dot)
// basis (the rotation matrices Block saves per block, rotationmat.C:21-60):
// kept sector s of `kept` is composite sector oldidx[s]; rot[Q] is the
// d_Q x m_Q row-major matrix of composite sector Q (m_Q = 0: dropped).
struct RotRecord {
  SIPtr composite, kept;
  std::vector<int> oldidx;
  std::vector<std::vector<double>> rot;
};

struct SpinBlock {
  int norb = 0;
  std::vector<int> sites;      // sorted orbital indices
  std::vector<char> inblock;   // norb flags
  SIPtr si;
  OpPtr ham, ident;
  std::vector<OpPtr> cre, ccd;          // [norb]
  std::vector<OpPtr> cd, cc, cdc, ddc;  // [(i*norb+j)*2 + K]
  bool has_normal = false, has_comp = false;
  // single-site helpers kept on dot blocks: "des" (D_d), "n" (n_d),
  // "cn" (C_d n_d), "dd0" ([D_d D_d]^0)
  std::map<std::string, OpPtr> extra;
  // set on renormalised blocks: the rotation from their composite basis
  std::shared_ptr<const RotRecord> rotrec;

  bool in(int i) const { return inblock[i] != 0; }
  size_t pidx(int i, int j, int K) const { return ((size_t)i * norb + j) * 2 + (K >> 1); }  // K doubled (0 or 2)
  const SparseOp* get(const std::vector<OpPtr>& v, int i, int j, int K) const {
    auto& p = v[pidx(i, j, K)];
    return p ? p.get() : nullptr;
  }
  size_t nops() const;
  size_t nbytes() const;
};

// Which two-index operators exist: Block's screening (screen.C:173-290) with
// its NUMERICAL_ZERO threshold.  A normal pair on a block is kept if it
// couples through a non-negligible integral to some pair of the complement; a
// complementary operator is kept if its integral sum over the block is not
// identically zero.
struct Screen {
  double thresh = 1e-15;
  bool cd_interacts(const Integrals& I, int i, int j, const std::vector<int>& other) const;
  bool dd_interacts(const Integrals& I, int i, int j, const std::vector<int>& other) const;
};

// Sector pruning for a block with nrem orbitals outside it and the given
// target: keep a sector only if the complement can complete it to the target.
std::function<bool(const Quantum&)> keep_for(const Quantum& target, int nrem);

// Single-site block with every operator it can carry (the dot block,
// default_op_components for a dot, set_spinblock_components.C:331-390).
BlockPtr make_dot(const Integrals& I, int site, const Screen& scr);

struct BuildSpec {
  bool normal = true;   // build CD/CC normal operators
  bool comp = true;     // build CDcomp/DDcomp complementary operators
  bool ccd = true;      // build CCDcomp for all sites outside
};
// B = B1 (x) dot, the composite block of BuildSumBlock (spinblock.C).
BlockPtr build_sum_block(const Integrals& I, const SpinBlock& B1, const SpinBlock& dot, const Quantum& target,
                         int nrem, const BuildSpec& spec, const Screen& scr);

// Renormalisation: new block with rotated operators O' = R^T O R per sector
// (SparseMatrix::renormalise_transform, BaseOperator.C:341-459).  rot[Q] is
// d_Q x m_Q row-major (m_Q may be 0: sector dropped).
struct Rotation {
  SIPtr ns;                 // kept sectors with their new dimensions
  const StateInfo* old = nullptr;
  std::vector<int> oldidx;  // new sector -> old sector
  const std::vector<std::vector<double>>* rot = nullptr;
};
Rotation make_rotation_info(const StateInfo& os, const std::vector<std::vector<double>>& rot,
                            const std::vector<int>& mkeep);
OpPtr rotate_op(const SparseOp& o, const Rotation& R);
// Rotation of a batch of operators of one block (renormalise_transform for
// all of them).  The default runs rotate_op on the host threads; the sweep
// can hand a device implementation to the renormalisation.
struct OpRotator {
  virtual ~OpRotator() = default;
  virtual std::vector<OpPtr> rotate(const std::vector<const SparseOp*>& ops, const Rotation& R) const;
  // operators a later rotate() will take, handed over early (may be called
  // from another host thread while the sweep's solve runs, never during a
  // rotate(); the default keeps nothing)
  virtual void preload(const std::vector<OpPtr>& ops) const {}
};
BlockPtr renormalise(const SpinBlock& B, const std::vector<std::vector<double>>& rot, const std::vector<int>& mkeep);
BlockPtr renormalise(const SpinBlock& B, const Rotation& R, const OpRotator* rotator = nullptr);
// Build the complementary operators of the composite block B = B1 (x) dot one
// at a time and store them rotated into out, the streaming form of
// build_and_renormalise_transform (BaseOperator.C:341-459).
void add_rotated_comp_ops(const Integrals& I, const SpinBlock& B, const SpinBlock& B1, const SpinBlock& dot,
                          const Screen& scr, const Rotation& R, SpinBlock& out, const OpRotator* rotator = nullptr);
// The same complementary operators built in B's own (composite) basis before
// the rotation is known -- they depend only on B1, the dot and the integrals,
// so the sweep builds them on a host thread while the GPU runs the Davidson
// solve -- and then rotated into out as one batch.
struct CompOpsAhead;
std::shared_ptr<CompOpsAhead> build_comp_ops(const Integrals& I, const SpinBlock& B, const SpinBlock& B1,
                                             const SpinBlock& dot, const Screen& scr);
void add_rotated_comp_ops(const CompOpsAhead& ahead, const Rotation& R, SpinBlock& out,
                          const OpRotator* rotator = nullptr);
// the operators built ahead (the ones add_rotated_comp_ops rotates)
std::vector<OpPtr> comp_ops_list(const CompOpsAhead& ahead);
// Another place to build composite blocks (the GPU): the same operators as
// build_sum_block / build_comp_ops, possibly device-only.
struct BlockBuilder {
  virtual ~BlockBuilder() = default;
  virtual BlockPtr sum_block(const Integrals& I, const SpinBlock& B1, const SpinBlock& dot, const Quantum& target,
                             int nrem, const BuildSpec& spec, const Screen& scr) const = 0;
  virtual std::shared_ptr<CompOpsAhead> comp_ops(const Integrals& I, const SpinBlock& B, const SpinBlock& B1,
                                                 const SpinBlock& dot, const Screen& scr) const = 0;
};
// seconds spent building complementary operators before a batch rotation (accumulated)
double comp_build_seconds();

}  // namespace blk
