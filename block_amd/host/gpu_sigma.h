// Work-list compiler and device residency for one superblock: the
// MI355X-side multiply_h.  Constructing a GpuSigma corresponds to constructing
// multiply_h(big, onedot) (davidson.h:22-29) for one Davidson solve: every
// operator block the term list touches is packed into one HBM arena and the
// term list is flattened into the two-phase GEMM-pair plan of
// include/block_sigma.h.  apply() is multiply_h::operator() (v += H c).
#pragma once
#include <memory>
#include <unordered_map>

#include "../../include/block_sigma.h"
#include "davidson.h"
#include "superblock.h"

namespace blk {

// fail loudly: every C-ABI status is checked
void bsg_check(int rc, const char* what);

// Device context plus buffers reused by successive superblocks of a sweep:
// the operator arena (leased by one GpuSigma at a time) and the page-locked
// staging it is uploaded through, both grown on demand.
struct GpuContext {
  int device = 0;
  bsg_context* ctx = nullptr;
  double* arena = nullptr;
  size_t arena_cap = 0;  // doubles
  bool arena_leased = false;
  double* staging = nullptr;
  size_t staging_cap = 0;  // doubles
  double* work = nullptr;  // scratch of the operator rotations
  size_t work_cap = 0;     // doubles
  // operators of the last arena upload and their offsets: valid until the
  // next upload (the sweep rotates the superblock's operators right after
  // its Davidson solve, while they are still in the arena)
  std::unordered_map<const SparseOp*, int64_t> resident;
  // Operators uploaded ahead of use by another host thread, on a second
  // context (its own stream, staging and buffers) so the copies run beside
  // the sigma kernels: slot 0 for the next rotation, slot 1 for the next
  // sigma arena.  An upload (upload_ops) that meets a preloaded operator
  // copies it on the device instead.  Preloaded operators are kept alive
  // until used, so no other operator can take one's address.
  struct Preload {
    std::vector<OpPtr> held;
    std::unordered_map<const SparseOp*, int64_t> at;  // offsets in buf
    double* buf = nullptr;
    size_t cap = 0;  // doubles
  };
  Preload pre[2];
  std::unique_ptr<GpuContext> side;  // created by the first preload
  explicit GpuContext(int device = 0);
  // grow-only buffers (contents not kept when they grow)
  double* ensure_staging(size_t n);
  double* ensure_work(size_t n);
  ~GpuContext();
  GpuContext(const GpuContext&) = delete;
  GpuContext& operator=(const GpuContext&) = delete;
};

// Upload ops to slot `slot` of g's preload store through g's side context
// (replaces what the slot held).  Called from a host thread other than the
// one using g, never while an upload of g runs.
void preload_ops(GpuContext& g, int slot, const std::vector<OpPtr>& ops);
// every operator of a block
std::vector<OpPtr> block_ops(const SpinBlock& b);

struct WorkList {
  std::vector<bsg_vblock_t> vb;    // V blocks with their pair ranges
  std::vector<bsg_pair_t> pairs;   // GEMM pairs grouped by V block, in term order
  std::vector<bsg_diag_t> diag;
  std::unordered_map<const SparseOp*, int64_t> arena_off;
  int64_t arena_size = 0;
  // per pair (same order): term index and the sectors l, l', r, r' it joins
  // (V block (l, r), operator blocks a(l, l') and b(r, r'), C block (l', r'))
  struct Prov {
    int term, l, lp, r, rp;
  };
  std::vector<Prov> prov;
};
// Flatten the term list into GEMM pairs grouped by the V block they update.
WorkList compile_worklist(const SuperBlock& sb);

// Seconds spent constructing GpuSigma objects, by phase (accumulated).
struct SetupTimes {
  double compile = 0, upload = 0, plan = 0;
  // GpuRotator::rotate: host pair lists, operator and matrix copies in, plan
  // creation, the rotation kernels (synchronised), the results out
  double rot_pairs = 0, rot_in = 0, rot_plan = 0, rot_run = 0, rot_out = 0;
};
const SetupTimes& gpu_setup_times();

class GpuSigma {
 public:
  GpuSigma(GpuContext& g, const SuperBlock& sb, int max_vec = 1);
  ~GpuSigma();
  GpuSigma(const GpuSigma&) = delete;
  GpuSigma& operator=(const GpuSigma&) = delete;

  long wave_size() const { return wave_size_; }
  double flops() const { return flops_; }
  size_t n_pairs() const { return n_pairs_; }
  size_t arena_bytes() const { return (size_t)arena_size_ * 8; }
  // arithmetic of the following applications: BSG_FP64 or BSG_SPLIT_BF16
  // (the BF16 slices of the arena are built on first use)
  void set_precision(int mode);
  int precision() const { return precision_; }
  // share of the flops the split mode runs on BF16 slices (bsg_plan_split_share)
  double split_share() const { return bsg_plan_split_share(plan_); }
  // host vectors: v = H c for nvec <= max_vec vectors (overwrites v)
  void apply(const double* c_host, double* v_host, int nvec = 1);
  // roots per launch of apply_device (bsg_plan_set_batch)
  void set_batch(int nroot);
  // the plan built again from the work list (after bsg_set_tile_table)
  void replan();
  int max_vec() const { return max_vec_; }
  // device vectors: v += H c
  void apply_device(const double* dc, double* dv, int nvec = 1);
  // apply_device through a captured HIP graph (one launch per application;
  // the input and output pass through the handle's own buffers)
  void set_graph(bool on);
  bool graph() const { return use_graph_; }
  std::vector<double> diagonal();
  bsg_context* ctx() const { return g_.ctx; }
  double* dev_c() const { return d_c_; }
  double* dev_v() const { return d_v_; }
  const double* dev_arena() const { return d_arena_; }
  bsg_plan* plan() const { return plan_; }

 private:
  GpuContext& g_;
  long wave_size_ = 0;
  double flops_ = 0.0;
  size_t n_pairs_ = 0;
  int64_t arena_size_ = 0;
  double* d_arena_ = nullptr;
  bool owns_arena_ = false;
  double* d_c_ = nullptr;
  double* d_v_ = nullptr;
  int max_vec_ = 1;
  int batch_ = 1;
  bsg_plan* plan_ = nullptr;
  int precision_ = BSG_FP64;
  bool split_ready_ = false;
  bool use_graph_ = false;
  int graph_precision_ = -1, graph_nvec_ = 0;
  std::vector<bsg_diag_t> diag_terms_;
  std::vector<bsg_vblock_t> vb_;  // the work list the plan is built from
  std::vector<bsg_pair_t> pairs_;
};

// Operator rotations O' = R_r^T O R_c on the GPU (transform_operators,
// save_load_block.C:267-319, through MatrixRotate, MatrixBLAS.C:553-572):
// every new operator block is one GEMM pair of the sigma kernels, with the
// rotation matrices as the operator arena and the old operator blocks as the
// wavefunction.  Operators are uploaded through page-locked staging and the
// rotated ones copied back.
// preload() uploads operators ahead of their rotation into slot 0 of the
// context's preload store, so the copy runs beside the sigma kernels of the
// solve; rotate() then copies them on the device (the sweep joins the
// preloading thread before it rotates).
// device_out: the rotated operators stay in HBM (device-only operators, one
// buffer per batch) instead of coming back to host memory.
class GpuRotator : public OpRotator {
 public:
  explicit GpuRotator(GpuContext& g, bool device_out = false) : g_(g), device_out_(device_out) {}
  std::vector<OpPtr> rotate(const std::vector<const SparseOp*>& ops, const Rotation& R) const override;
  void preload(const std::vector<OpPtr>& ops) const override {
    if (!device_out_) preload_ops(g_, 0, ops);
  }

 private:
  GpuContext& g_;
  bool device_out_;
};

// A device allocation of n doubles freed (through g) when the last owner
// goes; g must outlive it.
std::shared_ptr<void> device_buffer(GpuContext& g, size_t n, double** ptr);

// Composite blocks built in HBM (the GPU side of BuildSumBlock and of the
// complementary operators of the renormalisation): the host records the
// scaled block copies of blk::build_sum_block / build_comp_ops (BuildSession)
// and the device executes them from the block operators' device copies
// (bsg_build_apply); the new operators are device-only.  Same operators,
// bitwise, as the host build.
class GpuBlockBuilder : public BlockBuilder {
 public:
  explicit GpuBlockBuilder(GpuContext& g) : g_(g) {}
  BlockPtr sum_block(const Integrals& I, const SpinBlock& B1, const SpinBlock& dot, const Quantum& target, int nrem,
                     const BuildSpec& spec, const Screen& scr) const override;
  std::shared_ptr<CompOpsAhead> comp_ops(const Integrals& I, const SpinBlock& B, const SpinBlock& B1,
                                         const SpinBlock& dot, const Screen& scr) const override;

 private:
  GpuContext& g_;
};
// Execute a recorded build: every operator of `targets` and every operator an
// item writes gets device storage (targets in one buffer they own, the
// other written ones in scratch), sources are read from their device copies
// (host-only sources are uploaded first).
void execute_build(GpuContext& g, const BuildSession& session, const std::vector<SparseOp*>& targets);
// The grouped form bsg_build_apply takes (level 0 then level 1), with the
// operators' addresses from addr; run_build_host is the same arithmetic on
// the host (k_build's loop), execute_build_host the whole recorded build in
// host memory (the targets become host operators).  They check the
// recording and the grouping on the CPU.
// allocator whose resize leaves trivial elements uninitialised (the plan
// writes every element afterwards, in parallel)
template <class T>
struct UninitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = UninitAlloc<U>;
  };
  UninitAlloc() = default;
  template <class U>
  UninitAlloc(const UninitAlloc<U>&) {}
  template <class U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
struct BuildPlan {
  std::vector<bsg_build_group_t, UninitAlloc<bsg_build_group_t>> groups[2];
  std::vector<bsg_build_src_t, UninitAlloc<bsg_build_src_t>> srcs[2];
  double bytes = 0.0;  // algorithmic bytes (sources read, destinations read and written)
};
BuildPlan plan_build(const std::vector<const BuildItem*>& items,
                     const std::function<const double*(const SparseOp*)>& addr);
void run_build_host(const BuildPlan& plan);
void execute_build_host(const BuildSession& session, const std::vector<SparseOp*>& targets);
// Copy operators to dst + offset (doubles): device-resident ones (preloaded
// or device-only) on the device, host ones through g's page-locked staging.
// n: doubles spanned by the offsets.
void upload_ops(GpuContext& g, const std::vector<std::pair<const SparseOp*, int64_t>>& ops, double* dst, size_t n);
// Move host operators into one device buffer (they become device-only);
// operators already on the device are left as they are.
void make_device_resident(GpuContext& g, const std::vector<OpPtr>& ops);
// Make g the context that downloads device-only operators for host readers
// (materialise_host); null: none.
void set_materialiser_context(GpuContext* g);
// Seconds spent recording and executing device builds (accumulated).
struct BuildTimes {
  double record = 0, execute = 0;
  double plan = 0, apply = 0;  // of execute: host grouping (plan_build), device copies and bsg_build_apply
  double bytes = 0;  // algorithmic bytes of the executed builds (sources read, destinations written)
  double items = 0;  // recorded block copies executed
  double download = 0, download_bytes = 0;  // device-only operators copied to the host for host readers
};
const BuildTimes& gpu_build_times();

// Largest Davidson subspace (deflation_max + 1 vectors) the device solve takes.
// min_split_share: the ladder's split phase runs only when h.split_share()
// reaches it.
constexpr int kDeviceDavidsonMaxVectors = 128;
// Default smallest share of a superblock's flops on BF16 slices
// (split_share) for which a sweep's device Davidson runs the split phase of
// its precision ladder (below it the split sigma is mostly the FP64
// small-block kernel, and the restart at the switch costs more than it saves).
constexpr double kLadderMinSplitShare = 0.5;
// Linear::block_davidson on device vectors with the sigma of h (same control
// flow and outputs as blk::block_davidson; b, h_diag host vectors in/out).
void block_davidson_device(GpuSigma& h, std::vector<std::vector<double>>& b, std::vector<double>& h_diag,
                           double normtol, bool useprecond, int deflation_min, int deflation_max, DavidsonStats* st,
                           int max_iter, double ladder, const std::vector<std::vector<double>>* lower = nullptr,
                           double min_split_share = 0.0);

}  // namespace blk
