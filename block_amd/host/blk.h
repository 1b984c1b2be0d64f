// Host-side data model of the spin-adapted two-site DMRG path that feeds the
// MI355X sigma kernels.  It restates, with its own storage, the objects Block
// builds around multiply_h:
//   SpinQuantum / StateInfo          (SpinQuantum.h, StateInfo.h:79-206)
//   SparseMatrix / Transposeview     (BaseOperator.h:75-243)
//   Wavefunction                     (wavefunction.h:16-56)
// All spins are stored doubled (2S) as in the reference.
#pragma once
#include <cstdint>
#include <functional>
#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace blk {

// ---------------------------------------------------------------- quantum numbers
// Sector label (N, 2S, irrep).  Abelian point groups (C1, Ci, Cs, C2, C2v, C2h,
// D2, D2h) are handled with irreps numbered 0..7 and the XOR product table
// (Symmetry.C:90-104); spatial 9j factors are then 1 for allowed couplings
// (Symmetry.C:534-538).
struct Quantum {
  int n = 0, s2 = 0, irrep = 0;
  bool operator==(const Quantum& o) const { return n == o.n && s2 == o.s2 && irrep == o.irrep; }
  bool operator!=(const Quantum& o) const { return !(*this == o); }
  bool operator<(const Quantum& o) const {
    if (n != o.n) return n < o.n;
    if (s2 != o.s2) return s2 < o.s2;
    return irrep < o.irrep;
  }
};
inline int irrep_mul(int a, int b) { return a ^ b; }
inline bool triangle(int a2, int b2, int c2) {
  return c2 >= std::abs(a2 - b2) && c2 <= a2 + b2 && ((a2 + b2 + c2) & 1) == 0;
}

// ---------------------------------------------------------------- SU(2) algebra
// Clebsch-Gordan, 6j and 9j from the Racah closed forms, all arguments doubled.
double cg(int j1, int m1, int j2, int m2, int J, int M);
double sixj(int a, int b, int c, int d, int e, int f);
double ninej(int a, int b, int c, int d, int e, int f, int g, int h, int i);
// Recoupling weight of a tensor product [A^{ka} (x) B^{kb}]^K between coupled
// states, i.e. Block's normalised Ninej (couplingCoeffs.h:88-95):
//   sqrt((jl+1)(jr+1)(J'+1)(K+1)) * {jl' jr' J'; ka kb K; jl jr J}
// arguments in the reference's order (ket-left, ket-right, ket-total, ka, kb,
// K, bra-left, bra-right, bra-total).  Cached.
double wninej(int jlk, int jrk, int jk, int ka, int kb, int K, int jlb, int jrb, int jb);
// Scaling of the reduced element of the adjoint tensor T^dag (components
// (-1)^{k-q} (T_{-q})^T) relative to the transposed reduced element of T:
//   R_{T^dag}(J, J') = dagger_phase(J, J', k) * R_T(J', J).
// This is the per-sector scaling Block's Transposeview::get_scaling computes
// numerically (BaseOperator.C:56-91).  Cached.
double dagger_phase(int J2, int Jp2, int k2);

// ---------------------------------------------------------------- state info
// Sector list of a block.  For a composite block (sys (x) dot) it also keeps
// the collected-quanta map (StateInfo.C:123-340): every collected sector is a
// stack of sub-blocks (left sector, right sector) of size dl*dr, ordered
// lexicographically by (left, right); inside a sub-block the state index is
// il*dr + ir.
struct StateInfo {
  std::vector<Quantum> q;
  std::vector<int> dim;
  std::vector<long> off;  // unblocked offset of each sector
  long total = 0;

  struct Sub { int l, r, off; };
  std::shared_ptr<const StateInfo> left, right;  // set for composite blocks
  std::vector<std::vector<Sub>> subs;            // per collected sector
  // (l*nr + r) -> list of (sector, offset-in-sector); one entry per total spin
  std::vector<std::vector<std::pair<int, int>>> prod_to_sector;

  int nsec() const { return (int)q.size(); }
  int find(const Quantum& x) const;  // -1 if absent
  void finalize();                   // recompute off / total
};
using SIPtr = std::shared_ptr<const StateInfo>;

// Product of two blocks' states, collected by total quanta and pruned by the
// keep predicate (sectors unable to reach the target are dropped, like the
// constraint passed to BuildSumBlock, spinblock.C).
SIPtr make_product(const SIPtr& l, const SIPtr& r, const std::function<bool(const Quantum&)>& keep);

// ---------------------------------------------------------------- block-sparse operators
// Allowed-block pattern shared by every operator with the same delta quantum on
// the same StateInfo.  Blocks are dense row-major d_row x d_col and packed in
// (row sector, col sector) order; an operator is the pattern plus one flat
// array, which is also its layout in HBM.
struct OpPattern {
  SIPtr si;
  int dn = 0, k2 = 0, dirrep = 0;  // delta N, 2*rank, irrep
  struct Blk { int r, c, rows, cols; long off; };
  std::vector<Blk> blocks;
  std::vector<int> row_start;  // CSR over row sectors
  long size = 0;
  int find(int r, int c) const;
  bool allowed(const Quantum& row, const Quantum& col) const {
    return row.n == col.n + dn && triangle(col.s2, k2, row.s2) && row.irrep == irrep_mul(col.irrep, dirrep);
  }
};
using PatPtr = std::shared_ptr<const OpPattern>;
PatPtr get_pattern(const SIPtr& si, int dn, int k2, int dirrep);

// An operator's elements live in host memory (`data`), in device memory
// (`dev`, kept alive by `dev_owner`), or both.  Operators built or rotated on
// the GPU are device-only (host_valid false, `data` empty) until something on
// the host asks for them (materialise_host).
struct SparseOp : std::enable_shared_from_this<SparseOp> {
  PatPtr pat;
  bool fermion = false;
  std::vector<double> data;
  double* dev = nullptr;
  std::shared_ptr<void> dev_owner;
  bool host_valid = true;
  SparseOp() = default;
  SparseOp(PatPtr p, bool f);
  struct DeviceOnly {};
  SparseOp(PatPtr p, bool f, DeviceOnly) : pat(std::move(p)), fermion(f), host_valid(false) {}
  double* blk(int id) { return data.data() + pat->blocks[id].off; }
  const double* blk(int id) const { return data.data() + pat->blocks[id].off; }
  int dn() const { return pat->dn; }
  int k2() const { return pat->k2; }
  int dirrep() const { return pat->dirrep; }
  long size() const { return pat->size; }
  void axpy(double a, const SparseOp& x);
};
using OpPtr = std::shared_ptr<SparseOp>;

// Download of device-only operators (set by the GPU layer); materialise_host
// makes `data` valid for host readers (a no-op for host operators).
using HostMaterialiser = void (*)(const std::vector<SparseOp*>& ops);
void set_host_materialiser(HostMaterialiser f);
void materialise_host(const std::vector<SparseOp*>& ops);

// ---------------------------------------------------------------- build recording
// While a BuildSession is open, tensor_product and SparseOp::axpy on the
// session's threads do not compute: they append the scaled block copies they
// stand for, and operators created meanwhile get no host storage.  The GPU
// layer then executes the items (composite operators built in HBM from
// device-resident block operators).  Item: dst[r dst_ld + c] += (alpha *
// (trans ? src[c src_ld + r] : src[r src_ld + c])) * beta for r < rows,
// c < cols (the host build's product order), offsets in doubles from the
// operators' first element; level 0 items (linear combinations of whole
// operators) run before level 1 items.
struct BuildItem {
  const SparseOp* dst;
  const SparseOp* src;
  int64_t dst_off, src_off;
  int32_t dst_ld, src_ld, rows, cols, trans, level;
  double alpha, beta;
};
struct BuildSession {
  BuildSession();
  ~BuildSession();
  BuildSession(const BuildSession&) = delete;
  BuildSession& operator=(const BuildSession&) = delete;
  // all threads' items in place (call after the recorded work returned;
  // valid while the session lives)
  std::vector<const BuildItem*> item_ptrs() const;
  // every operator an item writes, kept alive until the session ends
  std::vector<OpPtr> written() const;
  struct Impl;
  std::unique_ptr<Impl> impl;
};
bool build_recording();
// The calling thread's open session (null if none), and entering one on a
// worker thread (OpenMP workers of recorded work inherit their parent's).
BuildSession::Impl* current_build_session();
void enter_build_session(BuildSession::Impl* s);

// Read-only view: op itself (dag=false) or its adjoint tensor (dag=true),
// times coef.  The adjoint's block (r,c) is dagger_phase(S_r,S_c,k) *
// op.block(c,r)^T and its delta quantum has N negated (irreps are self-inverse
// for the Abelian groups handled here).  Mirrors Transposeview
// (BaseOperator.h:208-243).
struct OpView {
  const SparseOp* op = nullptr;
  bool dag = false;
  double coef = 1.0;
  int dn() const { return dag ? -op->dn() : op->dn(); }
  int k2() const { return op->k2(); }
  int dirrep() const { return op->dirrep(); }
  bool fermion() const { return op->fermion; }
  const StateInfo& si() const { return *op->pat->si; }
  // block id in op storage for view block (r,c), and whether it is stored transposed
  int find(int r, int c) const { return dag ? op->pat->find(c, r) : op->pat->find(r, c); }
  double scale(int r, int c) const;  // coef * (dag ? dagger_phase : 1)
};
inline OpView V(const SparseOp& o, double c = 1.0) { return OpView{&o, false, c}; }
inline OpView Vd(const SparseOp& o, double c = 1.0) { return OpView{&o, true, c}; }

// c(B1 (x) B2) += scale * [a(B1) (x) b(B2)]^{K}, K = c's rank.  B1 is the
// first (left) factor of c's StateInfo.  Restates TensorProductElement
// (operatorfunctions.C:167-254) including the fermion sign of b moving past
// the left ket state.
void tensor_product(const OpView& a, const OpView& b, SparseOp& c, double scale);

// Dense helpers (row-major).  C = alpha * op(A) * op(B) + beta * C.
void dgemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
           double beta, double* C, int ldc);
// Symmetric eigendecomposition (Householder tridiagonalisation + implicit QL,
// the EISPACK tred2/tql2 method); eigenvalues ascending, eigenvectors in
// columns of z (row-major n x n).
void syev(int n, double* a_to_z, double* w);

// ---------------------------------------------------------------- wavefunction layout
// Allowed (lQ, rQ) blocks of a wavefunction with target quanta on the
// superblock L (x) R (wavefunction.C:91-127).  Element order is sectors
// row-major then block row-major, the order Linear::precondition walks
// (linear.C:31-41).
struct WaveLayout {
  SIPtr L, R;
  Quantum target;
  struct Blk { int l, r, rows, cols; long off; };
  std::vector<Blk> blocks;
  std::vector<int> index;  // l*nR + r -> block or -1
  long size = 0;
  int find(int l, int r) const { return index[(size_t)l * R->nsec() + r]; }
};
std::shared_ptr<const WaveLayout> make_wave_layout(const SIPtr& L, const SIPtr& R, const Quantum& target);

}  // namespace blk
