// FMM3d: the MI355X counterpart of the reference's FMM3d_MPI (fmm3d_mpi.hpp:36-249)
// on one rank.  Same configuration accessors (knl, np, ctr, rootLevel, ptsMax,
// maxLevel), same setup()/evaluate() call order and error convention (int
// return, 0 = success).  setup() builds the octree, lists and operator tables
// on the host and leaves everything device-resident; evaluate() runs the whole
// pass on one HIP stream.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "let3d.hpp"
#include "matmgnt3d.hpp"

namespace kifmm {

enum Precision { PREC_FP64 = 0, PREC_MIXED = 1 };

// Phase names follow the reference's PETSc events (manage_petsc_events.cpp:77-143).
enum Phase { PH_INI = 0, PH_UPW, PH_VLIST, PH_XLIST, PH_DOWN, PH_NEAR, PH_FIN, PH_COUNT };
// One more timed interval inside PH_VLIST: the Hadamard kernel alone.
constexpr int PH_HADAMARD = PH_COUNT, PH_TIMES = PH_COUNT + 1;
extern const char* const kPhaseNames[PH_TIMES];

class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  ~DeviceBuffer();
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  hipError_t alloc(size_t bytes);
  hipError_t upload(const void* host, size_t bytes);
  void release();
  template <typename T> T* as() const { return static_cast<T*>(ptr_); }
  size_t bytes() const { return bytes_; }
 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

// V-list interactions as sibling groups: the 8 children of one parent and
// their parent's 6^3 halo of colleague children.  Per group: halo slot ->
// source slot (-1 empty), target child -> target slot, V membership bits, and
// the halo slots in node order as pair lists for the ordered FP64 kernel.
struct VGroups {
  std::vector<int> src_node, trg_node, halo, hord, tgt, psrc, pslot, pcnt;
  std::vector<uint64_t> mask;
  double pairs = 0;
  DeviceBuffer src_node_d, trg_node_d, halo_d, hord_d, tgt_d, psrc_d, pslot_d, pcnt_d, mask_d;
  int ngrp() const { return int(tgt.size() / 8); }
};

// x as three BF16 parts (bit patterns), each the rounding of what the previous
// ones leave: x = p0 + p1 + p2 to about 2^-24 |x|.
void split_bf16x3(double x, uint16_t* p);

// FP32 copy of FP64 spectra rows (each row H complex values in the r2c layout
// (ka n + kb) h + kc, re and im interleaved) in the packed half-spectrum
// layout of the mixed path (spec_pack): rows x spec_packed_len(np) complex.
std::vector<float> pack_spectra32(const std::vector<double>& full, int np);

// Whether the 3 x 3 DOF block of FP64 translation spectra rows ([slot][sd*td][H]
// complex, DOF pair (j, i) at j * td + i) is symmetric, T(j, i) = T(i, j), to
// 1e-12 of the table's largest value.
bool table_sym3(const std::vector<double>& t, int sd, int td, int H);

// Split K of a mixed-path level GEMM launch that would have too few workgroups.
void split_k(GemmArgs& g);
// The split of a launch over nrhs right-hand sides is the one-RHS launch's
// (same K chunks per column), so every column sums in the one-RHS order.
void split_k_rhs(GemmArgs& g, int nrhs);

// Launch shapes the autotuner chooses (FMM3d::autotune), per handle.
struct Tuning {
  int near_mode = 1;   // mixed near field: 0 one workgroup per chunk in leaf order, 1 the same in cost
                       // order (largest first), 2 a persistent grid over a global work queue in cost order
  int near_stage = 1;  // staging weight of the chunk-split cost model: FMM3d::kStageWeights[near_stage]
  int trans = 0;       // mixed translations: 0 FP64 matrix cores (fused M2M), 6 split-BF16x6
  int vec_hadamard = 0;  // mixed Hadamard of vector kernels: 0 per pair, 1 / 2 halo sweep, 1 / 2 groups a tile
};
constexpr int kNumTuning = 4;

class FMM3d {
 public:
  static const double kStageWeights[3];
  FMM3d() = default;
  ~FMM3d();

  // --- configuration (fmm3d_mpi.hpp:170-193, knlmat3d_mpi.hpp:50-69) ---
  int& kernelType() { return kt_; }
  std::vector<double>& coefs() { return coefs_; }
  int& np() { return np_; }
  int& ptsMax() { return pts_max_; }
  int& maxLevel() { return max_level_; }
  double* ctr() { return ctr_; }
  int& rootLevel() { return root_level_; }
  int& precision() { return precision_; }
  int srcDOF() const;
  int trgDOF() const;

  // Points are copied (the caller keeps ownership); normals may be null for
  // kernels that do not use them.
  int setPoints(int64_t ns, const double* spos, const double* snor, int64_t nt, const double* tpos);
  // Many independent problems in one handle (a forest): problem p has ns[p]
  // sources and nt[p] targets; the arrays hold the problems one after another.
  // setup() builds one tree per problem and numbers them as one forest, so
  // every pass of evaluate() runs one launch over all problems' boxes.
  int setPointsBatch(int nprob, const int64_t* ns, const double* spos, const double* snor, const int64_t* nt,
                     const double* tpos);
  int numProblems() const { return prob_ns_.empty() ? 1 : int(prob_ns_.size()); }

  // Host half of setup(): octree, lists, operator tables and task tables, no
  // device calls (usable on a machine without a GPU).
  int setupHost();
  // setup() builds the octree and lists on the GPU (default) or on the host;
  // both give the same tree.
  int setup();
  void setTreeOnDevice(bool on) { tree_on_device_ = on; }
  // Mixed path translations (M2M, L2L): 0 = FP64 matrix cores with M2M fused
  // into the upward pseudo-inverse (default), 3 = split-BF16x3, 6 = split-BF16x6
  // products on the BF16 matrix cores.  Takes effect at the next setup().
  int setTranslationTerms(int terms);
  int translationTerms() const { return trans_terms_; }
  // Autotuner: times the mixed path's launch-shape choices (Tuning) on this
  // handle's tree with the caller's densities (nrhs right-hand sides), one
  // choice at a time, and keeps the fastest; d_trg_val receives the values of
  // the last trial.  No-op for the FP64 path.
  int autotune(const double* d_src_den, double* d_trg_val, int nrhs, hipStream_t stream);
  const Tuning& tuning() const { return tune_; }
  // Applies given choices (e.g. a previous autotune's); re-uploads what changes.
  int apply_tuning(const Tuning& t);
  struct TuneTrial { int knob, value; double ms; };
  const std::vector<TuneTrial>& tuneReport() const { return tune_report_; }
  // Wall times of the last setup, milliseconds: tree (total), of which device
  // build and copy-back (device builder only), operators, task tables, upload.
  enum SetupPart { SU_TREE = 0, SU_TREE_DEVICE, SU_TREE_COPY, SU_OPS, SU_TASKS, SU_UPLOAD, SU_COUNT };
  const std::array<double, SU_COUNT>& setupMs() const { return setup_ms_; }
  // Device pointers, caller's point order (srcDOF / trgDOF interleaved).
  int evaluate(const double* d_src_den, double* d_trg_val, hipStream_t stream);
  // nrhs right-hand sides in one pass: d_src_den holds nrhs density vectors and
  // d_trg_val receives nrhs value vectors, one after another.  Every pass runs
  // one launch for all of them (the point kernels stage each source once for
  // all right-hand sides, the GEMMs take nrhs times the columns, the V list
  // nrhs times the boxes); the right-hand-side buffers grow on the first call
  // with a larger nrhs.
  int evaluateMulti(const double* d_src_den, double* d_trg_val, int nrhs, hipStream_t stream);
  // Transposed pass: d_src_out (srcDOF per source) = A^T d_trg_in (trgDOF per
  // target), A the evaluate operator.  Needs a symmetric equivalent kernel
  // (111 or 511); the adjoint tables are built on the first call.
  int evaluateAdjoint(const double* d_trg_in, double* d_src_out, hipStream_t stream);
  // Gradient of the evaluated potential at the targets (3 values per target,
  // caller's order); Laplace single layer only.
  int field(const double* d_src_den, double* d_trg_grad, hipStream_t stream);
  // Upward equivalent densities as a checkpoint: evaluate() that also copies UE
  // (checkpointSize() doubles, node-major) to d_ue_out once the up pass is
  // done, and field() that takes UE from such a checkpoint instead of running
  // the up pass again (the densities must be those of the checkpoint).
  int64_t checkpointSize() const;
  int evaluateCheckpoint(const double* d_src_den, double* d_trg_val, double* d_ue_out, hipStream_t stream);
  int fieldFromCheckpoint(const double* d_src_den, const double* d_ue_in, double* d_trg_grad, hipStream_t stream);
  // FP64 direct sum at the given targets (device pointers, caller's order);
  // the partial-sum workspace is the handle's, grown on demand.
  int direct(const double* d_src_den, int nchk, const double* d_chk_pos, double* d_chk_val,
             hipStream_t stream);
  // Same with a caller-planned chunking and a caller-provided workspace of
  // d.work_doubles doubles (allocates nothing).
  int direct_with(const double* d_src_den, int nchk, const double* d_chk_pos, double* d_chk_val,
                  const DirectPlan& d, double* work, hipStream_t stream);
  // direct() without any allocation: the partial sums go to the library's
  // fixed per-GPU scratch (kScratchDoubles doubles); nchk * trgDOF beyond it is
  // rejected.  Calls that overlap on one GPU share the scratch and race.
  int directNoAlloc(const double* d_src_den, int nchk, const double* d_chk_pos, double* d_chk_val,
                    hipStream_t stream);

  const Tree& tree() const { return tree_; }
  const Operators* ops() const { return ops_.get(); }
  const std::string& lastError() const { return err_; }
  void setError(const std::string& e) { err_ = e; }
  bool isSetup() const { return setup_done_; }
  bool isHostReady() const { return host_ready_; }
  int64_t numSrc() const { return ns_; }
  int64_t numTrg() const { return nt_; }
  const std::vector<double>& targetPositions() const { return tpos_; }
  void enableTiming(bool on) { timing_ = on; }
  // Replay evaluate() from a captured hipGraph when it is called again with the
  // same device buffers on the same non-default stream (timing off).
  void enableGraph(bool on) { use_graph_ = on; if (!on) drop_graph(); }
  std::array<float, PH_TIMES> phaseMs() const { return phase_ms_; }
  // Work counts for the roofline (pairs per phase), filled at setup.
  struct Work {
    double near_pairs = 0, near_exact_pairs = 0, s2m_pairs = 0, s2l_pairs = 0;
    double vlist_pairs = 0, n_vsrc = 0, n_vtrg = 0, n_vgrp = 0;
    double up_cols = 0, m2m_cols = 0, dn_cols = 0, l2l_cols = 0;
  };
  const Work& work() const { return work_; }

 private:
  int fail(int code, const std::string& msg) { err_ = msg; return code; }
  int setup_prepare(bool device_tree);
  int check(hipError_t e, const char* what);
  // ue_in: UE from a checkpoint (the up pass is skipped); ue_out: UE copied
  // there after the up pass
  int evaluate_impl(const double* d_src_den, double* d_trg_val, hipStream_t s, bool grad, int nrhs = 1,
                    const double* ue_in = nullptr, double* ue_out = nullptr);
  int run_graph(const double* d_src_den, double* d_trg_val, hipStream_t s, int nrhs);
  int reserve_rhs(int nrhs);
  int prepare_rhs_steps(int nrhs);
  void drop_graph();
  int build_tasks();
  int retask();  // task tables and device state again (after a change of translations)
  NearChunk make_near_chunk(int B, int c, int k, int ibeg, int iend) const;
  void build_mixed_chunks();
  int upload_state();
  int build_vgroups(const std::vector<int>& order, const std::vector<int64_t>& off, const std::vector<int>& lst,
                    bool need_evtr, VGroups& vg);
  int upload_vgroups(VGroups& vg);
  void vlist_args(VlistArgs& v, const VGroups& vg) const;
  // m2l32: FP32 translation table of the mixed path (null: the forward one)
  int run_vlist(VlistArgs& v, hipStream_t s, bool timed, const float2* m2l32 = nullptr);
  int build_adjoint();

  int kt_ = 111, np_ = 6, pts_max_ = 64, max_level_ = 30, root_level_ = 1, precision_ = PREC_FP64;
  std::vector<double> coefs_{1.0, 0.25};
  double ctr_[3] = {0.5, 0.5, 0.5};
  int64_t ns_ = 0, nt_ = 0;
  std::vector<double> spos_, snor_, tpos_;
  std::vector<int64_t> prob_ns_, prob_nt_;  // forest: points per problem (empty: one problem)
  bool setup_done_ = false, host_ready_ = false, timing_ = false, tree_on_device_ = true;
  int trans_terms_ = 0;
  Tuning tune_;
  std::vector<TuneTrial> tune_report_;
  DeviceBuffer m2m_bf_d_, l2l_bf_d_;  // split-BF16 translation tables [3][nz][Mp][Kp]
  int m2m_mp_ = 0, m2m_kp_ = 0, l2l_mp_ = 0, l2l_kp_ = 0;
  std::array<double, SU_COUNT> setup_ms_{};
  std::string err_;
  Tree tree_;
  std::shared_ptr<const Operators> ops_;
  Work work_;
  std::array<float, PH_TIMES> phase_ms_{};

  // host task tables
  // One level-batched GEMM launch of the up or down pass.  mat: 0 = M2M
  // (the 8 child matrices side by side), 1 = UC->UE pinv, 2 = DC->DE pinv,
  // 3 = L2L (the 8 child matrices, one per z slice); transposed tables of the
  // adjoint: 4 = M2M^T (stacked, 8*UE x UC), 5 = pinv_up^T, 6 = pinv_dn^T,
  // 7 = L2L^T (side by side, DE x 8*DC); mixed up pass: 8 = P_up [M_0 ... M_7]
  // (UE of the children -> UE of the parent).  src/dst: 0 = UE, 1 = UC, 2 = DE, 3 = DC.
  struct GemmStep {
    int mat, src, dst, ncol, accumulate;
    int64_t idx_off;  // bidx at idx_off, cidx at idx_off + ncol
    int ks_off, rs_off;
    int64_t zoff_off = -1;  // 9 column offsets of the 8 L2L matrices, in lvl_idx_h_
    int max_zcol = 0;
  };
  std::vector<GemmStep> up_steps_, dn_steps_;
  // the same steps over nrhs right-hand sides (columns of right-hand side r
  // address node rows r * n_nodes + node), built for rhs_steps_n_
  std::vector<GemmStep> up_steps_r_, dn_steps_r_;
  std::vector<int> lvl_idx_r_h_;
  DeviceBuffer lvl_idx_r_d_;
  int rhs_steps_n_ = 1, rhs_cap_ = 1;
  DeviceBuffer splitk_d_;  // mixed path: split-K partial products (ordered sum, no atomics)
  int reserve_splitk(size_t doubles);
  double* splitk_for(const GemmArgs& g);
  std::vector<int> lvl_idx_h_;
  std::vector<double> scales_h_;
  std::vector<NearChunk> near_chunks_h_;
  std::vector<NearChunk> near_chunks_mx_h_;  // mixed near field: leaves split for full lanes
  std::vector<int> near_order_mx_h_;         // the mixed chunks by estimated cost, largest first
  std::vector<int> near_leaves_;             // evaluator leaves, node order
  std::vector<int64_t> near_ioff_;           // their item ranges in near_items_h_
  std::vector<double> sbox_;                 // per node: source bounding box (lo xyz, hi xyz)
  DeviceBuffer near_order_mx_d_, queue_d_;
  std::vector<int> near_items_h_;
  std::vector<SurfTask> s2m_tasks_h_, s2l_tasks_h_;
  std::vector<int4> s2m_leaf_h_;  // S2M tasks as (leaf, first source, source count)
  std::vector<int> surf_items_h_;
  VGroups vfwd_, vadj_;
  // adjoint tables (built on the first evaluateAdjoint)
  bool adj_ready_ = false;
  std::vector<NearChunk> nt_chunks_h_;
  std::vector<int> nt_items_h_;
  std::vector<SurfTask> ue_t_tasks_h_, de_t_tasks_h_;
  std::vector<int> surf_t_items_h_;
  std::vector<GemmStep> up_t_steps_, dn_t_steps_;

  // device state
  DeviceBuffer spos_d_, snor_d_, tpos_d_, sperm_d_, tperm_d_, sden_d_, tval_d_;
  DeviceBuffer sinv_d_;  // inverse of the source permutation: caller index -> leaf order
  DeviceBuffer geom_d_, depth_d_, sbeg_d_, snum_d_;
  DeviceBuffer ue_d_, uc_d_, dc_d_, de_d_;
  DeviceBuffer near_chunks_mx_d_;
  DeviceBuffer near_chunks_d_, near_items_d_, s2m_tasks_d_, s2l_tasks_d_, surf_items_d_, s2m_leaf_d_;
  std::vector<int> ue_zero_h_, dc_zero_h_;  // mixed evaluate: UE / DC rows zeroed at its start
  DeviceBuffer ue_zero_d_, dc_zero_d_;
  DeviceBuffer lvl_idx_d_;
  DeviceBuffer sam_d_[4], shell_d_, twiddle_d_, deg_d_, ones_d_, scales_d_;
  DeviceBuffer pup_d_, pdn_d_, m2m_d_, l2l_d_, m2l_d_;
  DeviceBuffer dspec_d_, espec_d_, tgrad_d_;
  DeviceBuffer nt_chunks_d_, nt_items_d_, ue_t_tasks_d_, de_t_tasks_d_, surf_t_items_d_, tbeg_d_, tnum_d_;
  DeviceBuffer m2mT_d_, pupT_d_, pdnT_d_, l2lT_d_;
  DeviceBuffer m2lT_d_, m2lT32_d_, deg0_d_;  // adjoint V list of the Stokes equivalent kernel
  DeviceBuffer wup_d_;  // mixed path: P_up [M_0 ... M_7]
  DeviceBuffer wdn_d_;  // mixed path: [P_dn | P_dn L2L_z], z = 0..7
  DeviceBuffer direct_work_d_;  // direct sum: partial sums of the source chunks
  DeviceBuffer m2l32_d_, dspec32_d_, espec32_d_;  // mixed path: FP32 spectra
  bool m2l_sym3_ = false, m2lT_sym3_ = false;     // table_sym3 of the forward / transposed table
  bool m2l_mirror_ = false;  // the FP32 table is the wrap-free one and table_mirror holds
  std::vector<hipEvent_t> events_;
  bool use_graph_ = false;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  const double* graph_src_ = nullptr;
  double* graph_dst_ = nullptr;
  hipStream_t graph_stream_ = nullptr;
  int graph_nrhs_ = 0;
};

}  // namespace kifmm
