// FMM3d setup and evaluate (see fmm3d.hpp).  Citations are to the reference's
// fmm3d_mpi/ sources.
#include "fmm3d.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>

#include "kernel3d.hpp"
#include "tree_device.hpp"

namespace kifmm {

const char* const kPhaseNames[PH_TIMES] = {"EvlIni", "EvlUpwComp", "EvlVList", "EvlXList",
                                           "EvlCombine", "EvlUList", "EvlFinalize", "VListHadamard"};

DeviceBuffer::~DeviceBuffer() { release(); }

hipError_t DeviceBuffer::alloc(size_t bytes) {
  release();
  if (bytes == 0) return hipSuccess;
  hipError_t e = hipMalloc(&ptr_, bytes);
  if (e == hipSuccess) bytes_ = bytes;
  else ptr_ = nullptr;
  return e;
}

hipError_t DeviceBuffer::upload(const void* host, size_t bytes) {
  hipError_t e = alloc(bytes);
  if (e != hipSuccess || bytes == 0) return e;
  return hipMemcpy(ptr_, host, bytes, hipMemcpyHostToDevice);
}

void DeviceBuffer::release() {
  if (ptr_) (void)hipFree(ptr_);
  ptr_ = nullptr;
  bytes_ = 0;
}

FMM3d::~FMM3d() {
  drop_graph();
  for (auto ev : events_) (void)hipEventDestroy(ev);
}

// Rows and depth of the forward passes' level GEMM steps, by matrix kind.
static void step_shape(const Operators& op, int mat, int& M, int& K) {
  switch (mat) {
    case 0: M = op.dat(UC); K = 8 * op.dat(UE); break;             // M2M
    case 1: M = op.dat(UE); K = op.dat(UC); break;                 // upward pseudo-inverse
    case 8: M = op.dat(UE); K = 8 * op.dat(UE); break;             // P_up [M_0 ... M_7]
    case 2: M = op.dat(DE); K = op.dat(DC); break;                 // downward pseudo-inverse
    case 9: M = op.dat(DE); K = op.dat(DC) + op.dat(DE); break;    // [P_dn | P_dn L2L_z]
    default: M = op.dat(DC); K = op.dat(DE);                       // L2L
  }
}

void FMM3d::drop_graph() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
  graph_src_ = nullptr;
  graph_dst_ = nullptr;
  graph_stream_ = nullptr;
  graph_nrhs_ = 0;
}

int FMM3d::srcDOF() const { return src_dof(kt_); }
int FMM3d::trgDOF() const { return trg_dof(kt_); }

int FMM3d::check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  err_ = std::string(what) + ": " + hipGetErrorString(e);
  return -5;
}

int FMM3d::setPoints(int64_t ns, const double* spos, const double* snor, int64_t nt, const double* tpos) {
  if (ns < 0 || nt < 0) return fail(-2, "negative point count");
  if ((ns > 0 && !spos) || (nt > 0 && !tpos)) return fail(-2, "null point array");
  ns_ = ns; nt_ = nt;
  spos_.assign(spos, spos + 3 * ns);
  if (snor) snor_.assign(snor, snor + 3 * ns);
  else snor_.assign(3 * ns, 0.0);
  tpos_.assign(tpos, tpos + 3 * nt);
  prob_ns_.clear();
  prob_nt_.clear();
  setup_done_ = false;
  host_ready_ = false;
  return 0;
}

int FMM3d::setPointsBatch(int nprob, const int64_t* ns, const double* spos, const double* snor, const int64_t* nt,
                          const double* tpos) {
  if (nprob < 1 || !ns || !nt) return fail(-2, "a batch needs at least one problem and its point counts");
  int64_t S = 0, T = 0;
  for (int p = 0; p < nprob; p++) {
    if (ns[p] < 0 || nt[p] < 0) return fail(-2, "negative point count in problem " + std::to_string(p));
    S += ns[p];
    T += nt[p];
  }
  if (S > INT32_MAX || T > INT32_MAX) return fail(-2, "too many points for 32-bit indices");
  int rc = setPoints(S, spos, snor, T, tpos);
  if (rc) return rc;
  if (nprob > 1) {
    prob_ns_.assign(ns, ns + nprob);
    prob_nt_.assign(nt, nt + nprob);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// setup: fmm3d_setup_mpi.cpp:41-565 (single rank)
// ---------------------------------------------------------------------------
int FMM3d::setup() {
  drop_graph();
  int rc = setup_prepare(tree_on_device_);
  if (rc) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  rc = upload_state();
  if (rc) return rc;
  setup_ms_[SU_UPLOAD] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  setup_done_ = true;
  return 0;
}

int FMM3d::setupHost() { return setup_prepare(false); }

int FMM3d::setTranslationTerms(int terms) {
  if (terms != 0 && terms != 3 && terms != 6) return fail(-2, "translation terms must be 0 (FP64), 3 or 6 (split BF16)");
  if (terms != trans_terms_) {
    trans_terms_ = terms;
    tune_.trans = terms == 6 ? 6 : 0;
    drop_graph();
    setup_done_ = false;  // the step tables differ: setup() again
    host_ready_ = false;
  }
  return 0;
}

// x = p0 + p1 + p2 + O(2^-24 |x|): each part the BF16 (round to nearest even)
// of what the previous parts leave, via FP32.
void split_bf16x3(double x, uint16_t* p) {
  double r = x;
  for (int q = 0; q < 3; q++) {
    const float f = float(r);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) { p[q] = uint16_t(u >> 16); r = 0.0; continue; }  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    p[q] = uint16_t(u >> 16);
    const uint32_t back = uint32_t(p[q]) << 16;
    float b;
    std::memcpy(&b, &back, 4);
    r -= double(b);
  }
}

int FMM3d::setup_prepare(bool device_tree) {
  using clk = std::chrono::steady_clock;
  auto ms_since = [](clk::time_point a) { return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
  setup_ms_.fill(0.0);
  setup_done_ = false;
  host_ready_ = false;
  if (!kernel_known(kt_)) return fail(-2, "unknown kernel type " + std::to_string(kt_));
  const int km = mm_kernel(kt_);
  if (km < 0) return fail(-2, "kernel " + std::to_string(kt_) + " has no equivalent-density kernel");
  if ((kt_ == 511 || kt_ == 521) && coefs_.size() < 2) return fail(-2, "Navier kernels need (mu, nu)");
  if (coefs_.empty()) return fail(-2, "kernel coefficients missing");
  if (np_ < 3 || np_ > 8) return fail(-2, "np must lie in [3,8]");
  if (precision_ != PREC_FP64 && precision_ != PREC_MIXED) return fail(-2, "unknown precision mode");
  int rc = 0;
  auto t0 = clk::now();
  if (device_tree) {
    // the same tree, built on the GPU (one launch sequence for all problems)
    const int P = numProblems();
    const int64_t one_s = ns_, one_t = nt_;
    const int64_t* pns = prob_ns_.empty() ? &one_s : prob_ns_.data();
    const int64_t* pnt = prob_ns_.empty() ? &one_t : prob_nt_.data();
    double tms[2] = {0, 0};
    rc = build_tree_device(P, pns, spos_.data(), pnt, tpos_.data(), ctr_, root_level_, pts_max_, max_level_, tree_,
                           &err_, nullptr, tms);
    if (rc) return rc;
    setup_ms_[SU_TREE_DEVICE] = tms[0];
    setup_ms_[SU_TREE_COPY] = tms[1];
  } else if (prob_ns_.empty()) {
    rc = build_tree(spos_.data(), ns_, tpos_.data(), nt_, ctr_, root_level_, pts_max_, max_level_, tree_, &err_);
    if (rc) return rc;
  } else {
    // one tree per problem (problems in parallel, each builder serial inside),
    // then one numbering for the forest
    const int P = int(prob_ns_.size());
    std::vector<int64_t> so(P, 0), to(P, 0);
    for (int p = 1; p < P; p++) { so[p] = so[p - 1] + prob_ns_[p - 1]; to[p] = to[p - 1] + prob_nt_[p - 1]; }
    std::vector<Tree> parts(P);
    std::vector<int> prc(P, 0);
    std::vector<std::string> perr(P);
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < P; p++)
      prc[p] = build_tree(spos_.data() + 3 * so[p], prob_ns_[p], tpos_.data() + 3 * to[p], prob_nt_[p], ctr_,
                          root_level_, pts_max_, max_level_, parts[p], &perr[p]);
    for (int p = 0; p < P; p++)
      if (prc[p]) return fail(prc[p], "problem " + std::to_string(p) + ": " + perr[p]);
    merge_trees(parts, so, to, tree_);
  }
  setup_ms_[SU_TREE] = ms_since(t0);
  t0 = clk::now();
  std::string e;
  std::vector<double> mmc = coefs_;
  if (kt_ == 312 || kt_ == 322) mmc.clear();  // fmm3d_setup_mpi.cpp:119,121
  if (mmc.empty()) mmc = {1.0, 0.25};
  ops_ = get_operators(km, mmc, np_, &e);
  if (!ops_) return fail(-2, e);
  setup_ms_[SU_OPS] = ms_since(t0);
  t0 = clk::now();
  rc = build_tasks();
  if (rc) return rc;
  setup_ms_[SU_TASKS] = ms_since(t0);
  rhs_steps_n_ = 1;
  up_steps_r_.clear(); dn_steps_r_.clear(); lvl_idx_r_h_.clear();
  host_ready_ = true;
  return 0;
}

// A near-field chunk of leaf B: targets [c, c + k) of the leaf, items
// [ibeg, iend); own / unchk from the bounding boxes (see NearChunk).
NearChunk FMM3d::make_near_chunk(int B, int c, int k, int ibeg, int iend) const {
  const Tree& t = tree_;
  NearChunk ch{B, t.trg_beg[B] + c, k, ibeg, iend};
  double tb[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int j = 0; j < ch.tnum; j++) {
    const int64_t o = t.trg_perm[size_t(ch.tbeg) + j];
    for (int d = 0; d < 3; d++) {
      tb[d] = std::min(tb[d], tpos_[3 * o + d]);
      tb[3 + d] = std::max(tb[3 + d], tpos_[3 * o + d]);
    }
  }
  bool sep = true;
  for (int i = ibeg; i < iend && sep; i++) {
    const int code = near_items_h_[size_t(i)], A = code >> 2;
    if ((code & 3) != ITEM_EXACT) continue;  // surfaces of non-adjacent nodes or the DE surface
    if (A == B && ch.own < 0) { ch.own = i - ibeg; continue; }
    const double* bx = &sbox_[size_t(A) * 6];
    double d2 = 0;
    for (int d = 0; d < 3; d++) {
      const double g = std::max(0.0, std::max(bx[d] - tb[3 + d], tb[d] - bx[3 + d]));
      d2 += g * g;
    }
    if (d2 < 1e-14) sep = false;  // within 1e-7: keep the cut on every pair
  }
  ch.unchk = sep ? 1 : 0;
  return ch;
}

const double FMM3d::kStageWeights[3] = {0.04, 0.15, 0.30};

// Mixed path: the near kernels put S = min(16, 128 / k) copies of a chunk's k
// targets on the 128 lanes, so a leaf of 70 targets leaves 58 lanes idle.
// Leaves are cut into chunks of sizes that fill the lanes: cost of a chunk =
// (active waves) / S sweeps of its sources, plus the staging weight per wave
// for staging them again; the cut minimises the sum (dynamic programming over
// sizes).  The chunks are also ordered by that cost, largest first (the order
// of the work queue).
void FMM3d::build_mixed_chunks() {
  const Tree& t = tree_;
  const double wstage = kStageWeights[tune_.near_stage];
  int max_tn = 0;
  for (int B : near_leaves_) max_tn = std::max(max_tn, t.trg_num[B]);
  std::vector<double> split_cost(size_t(max_tn) + 1, 0.0);
  std::vector<int> split_pick(size_t(max_tn) + 1, 0);
  for (int n = 1; n <= max_tn; n++) {
    split_cost[n] = 1e300;
    for (int k = 1; k <= std::min(n, kNearChunk); k++) {
      const int S = std::min(16, kNearChunk / k);
      const double c = double((S * k + 63) / 64) / S + 2 * wstage + split_cost[n - k];
      if (c < split_cost[n] - 1e-12) { split_cost[n] = c; split_pick[n] = k; }
    }
  }
  const int NL = int(near_leaves_.size());
  std::vector<int64_t> moff(size_t(NL) + 1, 0);
  for (int l = 0; l < NL; l++) {
    const int tn = t.trg_num[near_leaves_[l]];
    int nm = 0;
    for (int c = 0; c < tn; c += split_pick[tn - c]) nm++;
    moff[size_t(l) + 1] = moff[size_t(l)] + nm;
  }
  near_chunks_mx_h_.assign(size_t(moff[size_t(NL)]), NearChunk{});
  std::vector<double> cost(near_chunks_mx_h_.size(), 0.0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int l = 0; l < NL; l++) {
    const int B = near_leaves_[l], tn = t.trg_num[B];
    const int ibeg = int(near_ioff_[size_t(l)]), iend = int(near_ioff_[size_t(l) + 1]);
    double src = 0;  // source points over the leaf's items
    for (int i = ibeg; i < iend; i++) {
      const int code = near_items_h_[size_t(i)], A = code >> 2, kind = code & 3;
      src += kind == ITEM_EXACT ? t.src_num[A] : ops_->n_sam[kind == ITEM_UE ? UE : DE];
    }
    int64_t mi = moff[size_t(l)];
    for (int c = 0; c < tn;) {
      const int k = split_pick[tn - c];
      const int S = std::min(16, kNearChunk / k);
      cost[size_t(mi)] = double((S * k + 63) / 64) / S * src;
      near_chunks_mx_h_[size_t(mi++)] = make_near_chunk(B, c, k, ibeg, iend);
      c += k;
    }
  }
  near_order_mx_h_.resize(near_chunks_mx_h_.size());
  for (size_t i = 0; i < near_order_mx_h_.size(); i++) near_order_mx_h_[i] = int(i);
  std::stable_sort(near_order_mx_h_.begin(), near_order_mx_h_.end(),
                   [&](int a, int b) { return cost[size_t(a)] > cost[size_t(b)]; });
}

int FMM3d::build_tasks() {
  const Tree& t = tree_;
  const Operators& op = *ops_;
  const int N = t.n_nodes, SD = srcDOF(), TD = trgDOF(), sd = op.sd;
  const int szUE = op.dat(UE), szDC = op.dat(DC);
  work_ = Work();

  // --- near field: U, W (S2T or M2T), X (S2T fallback), L2T ---
  near_chunks_h_.clear(); near_chunks_mx_h_.clear(); near_items_h_.clear();
  // bounding boxes of every leaf's sources (the near kernel of the mixed path
  // applies the 1e-8 pair cut only where a source can come that close)
  sbox_.assign(size_t(N) * 6, 0.0);
  std::vector<double>& sbox = sbox_;
#pragma omp parallel for schedule(dynamic, 256)
  for (int A = 0; A < N; A++) {
    double* bx = &sbox[size_t(A) * 6];
    for (int d = 0; d < 3; d++) { bx[d] = 1e300; bx[3 + d] = -1e300; }
    if (t.src_beg[A] < 0) continue;
    for (int k = 0; k < t.src_num[A]; k++) {
      const int64_t o = t.src_perm[size_t(t.src_beg[A]) + k];
      for (int d = 0; d < 3; d++) {
        bx[d] = std::min(bx[d], spos_[3 * o + d]);
        bx[3 + d] = std::max(bx[3 + d], spos_[3 * o + d]);
      }
    }
  }
  // Evaluator leaves in node order; their items and chunks are counted first,
  // then filled in parallel at their prefix offsets (the same order as a
  // sequential pass over the leaves).
  std::vector<int>& near_leaves = near_leaves_;
  near_leaves.clear();
  for (int B = 0; B < N; B++)
    if ((t.tag[B] & TAG_EVTR) && t.leaf(B)) near_leaves.push_back(B);
  const int NL = int(near_leaves.size());
  near_ioff_.assign(size_t(NL) + 1, 0);
  std::vector<int64_t>& ioff = near_ioff_;
  std::vector<int64_t> coff(size_t(NL) + 1, 0);
  auto x_fallback = [&](int B) { return t.trg_num[B] * TD < szDC; };  // :1249
#pragma omp parallel for schedule(dynamic, 256)
  for (int l = 0; l < NL; l++) {
    const int B = near_leaves[l], tn = t.trg_num[B];
    int64_t ni = (t.off[LIST_U][B + 1] - t.off[LIST_U][B]) + (t.off[LIST_W][B + 1] - t.off[LIST_W][B]);
    if (x_fallback(B)) ni += t.off[LIST_X][B + 1] - t.off[LIST_X][B];
    if (t.depth[B] >= 2) ni++;
    ioff[size_t(l) + 1] = ni;
    coff[size_t(l) + 1] = (tn + kNearChunk - 1) / kNearChunk;
  }
  for (int l = 0; l < NL; l++) {
    ioff[size_t(l) + 1] += ioff[size_t(l)];
    coff[size_t(l) + 1] += coff[size_t(l)];
  }
  if (ioff[size_t(NL)] > INT32_MAX) return fail(-2, "too many near-field items for 32-bit offsets");
  near_items_h_.assign(size_t(ioff[size_t(NL)]), 0);
  near_chunks_h_.assign(size_t(coff[size_t(NL)]), NearChunk{});
  double near_exact = 0, near_all = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : near_exact, near_all)
  for (int l = 0; l < NL; l++) {
    const int B = near_leaves[l];
    const int ibeg = int(ioff[size_t(l)]);
    int* it = near_items_h_.data();
    int ni = ibeg;
    const int tn = t.trg_num[B];
    double exact = 0, surf = 0;
    for (int64_t q = t.off[LIST_U][B]; q < t.off[LIST_U][B + 1]; q++) {
      const int A = t.lst[LIST_U][q];
      it[ni++] = A << 2 | ITEM_EXACT;
      exact += t.src_num[A];
    }
    for (int64_t q = t.off[LIST_W][B]; q < t.off[LIST_W][B + 1]; q++) {  // :1146-1165
      const int A = t.lst[LIST_W][q];
      if (t.src_beg[A] >= 0 && t.src_num[A] * SD < szUE) {
        it[ni++] = A << 2 | ITEM_EXACT;
        exact += t.src_num[A];
      } else {
        it[ni++] = A << 2 | ITEM_UE;
        surf += op.n_sam[UE];
      }
    }
    if (x_fallback(B))
      for (int64_t q = t.off[LIST_X][B]; q < t.off[LIST_X][B + 1]; q++) {
        const int A = t.lst[LIST_X][q];
        it[ni++] = A << 2 | ITEM_EXACT;
        exact += t.src_num[A];
      }
    if (t.depth[B] >= 2) {  // DE is zero above depth 2 (:1378)
      it[ni++] = B << 2 | ITEM_DE;
      surf += op.n_sam[DE];
    }
    int64_t ci = coff[size_t(l)];
    for (int c = 0; c < tn; c += kNearChunk)
      near_chunks_h_[size_t(ci++)] = make_near_chunk(B, c, std::min(kNearChunk, tn - c), ibeg, ni);
    near_exact += double(tn) * exact;
    near_all += double(tn) * (exact + surf);
  }
  work_.near_exact_pairs = near_exact;
  work_.near_pairs = near_all;
  build_mixed_chunks();

  // --- S2M (:557-577) and X-list S2L (:1241-1257) ---
  s2m_tasks_h_.clear(); s2l_tasks_h_.clear(); surf_items_h_.clear(); s2m_leaf_h_.clear();
  for (int B = 1; B < N; B++)
    if ((t.tag[B] & TAG_CBTR) && t.leaf(B)) {
      int ib = int(surf_items_h_.size());
      surf_items_h_.push_back(B);
      s2m_tasks_h_.push_back(SurfTask{B, ib, ib + 1});
      s2m_leaf_h_.push_back(make_int4(B, t.src_beg[B], t.src_num[B], 0));
      work_.s2m_pairs += double(t.src_num[B]) * op.n_sam[UC];
    }
  for (int B = 0; B < N; B++) {
    if (!(t.tag[B] & TAG_EVTR)) continue;
    if (t.off[LIST_X][B] == t.off[LIST_X][B + 1]) continue;
    if (t.leaf(B) && t.trg_num[B] * TD < szDC) continue;
    int ib = int(surf_items_h_.size());
    for (int64_t q = t.off[LIST_X][B]; q < t.off[LIST_X][B + 1]; q++) {
      surf_items_h_.push_back(t.lst[LIST_X][q]);
      work_.s2l_pairs += double(t.src_num[t.lst[LIST_X][q]]) * op.n_sam[DC];
    }
    s2l_tasks_h_.push_back(SurfTask{B, ib, int(surf_items_h_.size())});
  }

  // --- level-batched GEMM steps ---
  up_steps_.clear(); dn_steps_.clear(); lvl_idx_h_.clear(); scales_h_.clear();
  auto scale_vec = [&](int lvl, int sign) {
    int off = int(scales_h_.size());
    for (int s = 0; s < sd; s++) scales_h_.push_back(std::ldexp(1.0, sign * int(lvl * op.deg[s])));
    return off;
  };
  auto add_step = [&](std::vector<GemmStep>& v, int mat, int src, int dst, const std::vector<int>& b,
                      const std::vector<int>& c, int ks, int rs, int acc) {
    if (b.empty()) return;
    GemmStep st{mat, src, dst, int(b.size()), acc, int64_t(lvl_idx_h_.size()), ks, rs};
    lvl_idx_h_.insert(lvl_idx_h_.end(), b.begin(), b.end());
    lvl_idx_h_.insert(lvl_idx_h_.end(), c.begin(), c.end());
    v.push_back(st);
  };
  const int rl = t.root_level;
  std::vector<std::vector<int>> by_depth(t.max_depth + 1);
  for (int i = 0; i < N; i++) by_depth[t.depth[i]].push_back(i);
  const bool fuse_up = precision_ == PREC_MIXED && trans_terms_ == 0, fuse_dn = fuse_up;
  for (int d = t.max_depth; d >= 1; d--) {  // :557-595, children before parents
    std::vector<int> mb, mc, pb, lb;
    for (int i : by_depth[d]) {
      if (!(t.tag[i] & TAG_CBTR)) continue;
      pb.push_back(i);
      if (!t.leaf(i)) { mb.push_back(t.child[i]); mc.push_back(i); }
      else lb.push_back(i);
    }
    work_.m2m_cols += double(mc.size());
    work_.up_cols += double(pb.size());
    if (fuse_up) {
      // mixed path: a parent's UE straight from its children's UE through
      // W = P_up [M_0 ... M_7] (UC is formed for the leaves only)
      add_step(up_steps_, 8, 0, 0, mb, mc, scale_vec(d + 1 + rl, +1), scale_vec(d + rl, -1), 0);
      add_step(up_steps_, 1, 1, 0, lb, lb, -1, scale_vec(d + rl, -1), 0);
    } else {
      add_step(up_steps_, 0, 0, 1, mb, mc, scale_vec(d + 1 + rl, +1), -1, 0);
      add_step(up_steps_, 1, 1, 0, pb, pb, -1, scale_vec(d + rl, -1), 0);
    }
  }
  for (int d = 2; d <= t.max_depth; d++) {  // :1368-1382, parents before children
    std::vector<int> pb;
    std::vector<int> lb[8], lc[8];
    for (int i : by_depth[d]) {
      if (!(t.tag[i] & TAG_EVTR)) continue;
      pb.push_back(i);
      if (d >= 3) {
        int p = t.parent[i], c = i - t.child[p];
        lb[c].push_back(p); lc[c].push_back(i);
      }
    }
    if (d >= 3 && fuse_dn) {
      // mixed path: DE straight from DC and the parent's DE through the 8
      // matrices [P_dn | P_dn L2L_z], one launch a level (no DC round trip)
      const int ks = scale_vec(d - 1 + rl, +1);
      std::vector<int> b, c;
      int zoff[9] = {0}, mz = 0;
      for (int q = 0; q < 8; q++) {
        b.insert(b.end(), lb[q].begin(), lb[q].end());
        c.insert(c.end(), lc[q].begin(), lc[q].end());
        zoff[q + 1] = int(b.size());
        mz = std::max(mz, int(lb[q].size()));
      }
      work_.l2l_cols += double(c.size());
      work_.dn_cols += double(pb.size());
      if (!b.empty()) {
        add_step(dn_steps_, 9, 2, 2, b, c, ks, scale_vec(d + rl, -1), 0);
        dn_steps_.back().zoff_off = int64_t(lvl_idx_h_.size());
        dn_steps_.back().max_zcol = mz;
        lvl_idx_h_.insert(lvl_idx_h_.end(), zoff, zoff + 9);
      }
      continue;
    }
    if (d >= 3) {  // the 8 child matrices of the level in one launch
      int ks = scale_vec(d - 1 + rl, +1);
      std::vector<int> b, c;
      int zoff[9] = {0}, mz = 0;
      for (int q = 0; q < 8; q++) {
        b.insert(b.end(), lb[q].begin(), lb[q].end());
        c.insert(c.end(), lc[q].begin(), lc[q].end());
        zoff[q + 1] = int(b.size());
        mz = std::max(mz, int(lb[q].size()));
      }
      work_.l2l_cols += double(c.size());
      if (!b.empty()) {
        add_step(dn_steps_, 3, 2, 3, b, c, ks, -1, 1);
        dn_steps_.back().zoff_off = int64_t(lvl_idx_h_.size());
        dn_steps_.back().max_zcol = mz;
        lvl_idx_h_.insert(lvl_idx_h_.end(), zoff, zoff + 9);
      }
    }
    work_.dn_cols += double(pb.size());
    add_step(dn_steps_, 2, 3, 2, pb, pb, -1, scale_vec(d + rl, -1), 0);
  }

  // --- V list as sibling groups ---
  std::vector<int> by_depth_order;
  by_depth_order.reserve(N);
  for (const auto& lv : by_depth) by_depth_order.insert(by_depth_order.end(), lv.begin(), lv.end());
  int rc = build_vgroups(by_depth_order, t.off[LIST_V], t.lst[LIST_V], true, vfwd_);
  if (rc) return rc;
  work_.vlist_pairs = vfwd_.pairs;
  work_.n_vsrc = double(vfwd_.src_node.size());
  work_.n_vtrg = double(vfwd_.trg_node.size());
  work_.n_vgrp = double(vfwd_.ngrp());
  // rows the mixed evaluate reads without writing them first, zeroed at its
  // start instead of the whole buffers: UE of the nodes without sources (the
  // fused M2M reads all 8 children of a parent), DC of the evaluator nodes of
  // depth >= 2 that are not V-list targets (the V-list inverse stores the others;
  // S2L adds after it and the downward step reads them)
  ue_zero_h_.clear(); dc_zero_h_.clear();
  {
    std::vector<char> vt(N, 0);
    for (int B : vfwd_.trg_node) vt[B] = 1;
    for (int B = 0; B < N; B++) {
      if (t.depth[B] >= 1 && !(t.tag[B] & TAG_CBTR)) ue_zero_h_.push_back(B);
      if (t.depth[B] >= 2 && (t.tag[B] & TAG_EVTR) && !vt[B]) dc_zero_h_.push_back(B);
    }
  }
  adj_ready_ = false;
  return 0;
}

// Groups the V lists (CSR over nodes) of the nodes in `order` by parent; the
// sources get slots in order of first use.  need_evtr: only evaluator nodes
// are targets (the forward lists); the adjoint passes the inverted lists.
int FMM3d::build_vgroups(const std::vector<int>& order, const std::vector<int64_t>& off,
                         const std::vector<int>& lst, bool need_evtr, VGroups& vg) {
  const Tree& t = tree_;
  const int N = t.n_nodes;
  vg.src_node.clear(); vg.trg_node.clear(); vg.halo.clear(); vg.hord.clear(); vg.tgt.clear();
  vg.psrc.clear(); vg.pslot.clear(); vg.pcnt.clear(); vg.mask.clear();
  vg.pairs = 0;
  std::vector<int> src_slot(N, -1);
  std::map<int, int> grp;
  // level-major order (depth, then node index): consecutive groups, and the
  // sources they first touch, are neighbours on one level
  for (int B : order) {
    if (need_evtr && !(t.tag[B] & TAG_EVTR)) continue;
    const int64_t b0 = off[B], b1 = off[B + 1];
    if (b0 == b1) continue;
    const int P = t.parent[B], tau = B - t.child[P];
    auto it = grp.find(P);
    int g;
    if (it == grp.end()) {
      g = int(grp.size());
      grp[P] = g;
      vg.halo.resize(size_t(g + 1) * 216, -1);
      vg.tgt.resize(size_t(g + 1) * 8, -1);
      vg.mask.resize(size_t(g + 1) * 32, 0);
    } else g = it->second;
    vg.tgt[size_t(g) * 8 + tau] = int(vg.trg_node.size());
    vg.trg_node.push_back(B);
    for (int64_t q = b0; q < b1; q++) {
      const int A = lst[q];
      if (src_slot[A] < 0) { src_slot[A] = int(vg.src_node.size()); vg.src_node.push_back(A); }
      int r = 0;
      for (int d = 0; d < 3; d++) {
        int rel = t.path[3 * A + d] - 2 * t.path[3 * P + d] + 2;
        if (rel < 0 || rel > 5) return fail(-6, "V-list member outside the parent's colleague halo");
        r = r * 6 + rel;
      }
      int& h = vg.halo[size_t(g) * 216 + r];
      if (h >= 0 && h != src_slot[A]) return fail(-6, "inconsistent V-list halo");
      h = src_slot[A];
      vg.mask[size_t(g) * 32 + tau * 4 + (r >> 6)] |= (uint64_t(1) << (r & 63));
      vg.pairs += 1;
    }
  }
  const int ngrp = int(grp.size());
  // The halo-sweep Hadamard relies on V(tau) = {occupied halo slots not adjacent
  // to tau} for every target child tau of a group; check it on the lists.
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int g = 0; g < ngrp; g++)
    for (int tau = 0; tau < 8; tau++) {
      if (vg.tgt[size_t(g) * 8 + tau] < 0) continue;
      const int tc[3] = {tau >> 2, (tau >> 1) & 1, tau & 1};
      for (int r = 0; r < 216; r++) {
        const int rc[3] = {r / 36 - 2, (r / 6) % 6 - 2, r % 6 - 2};
        bool adj = true;
        for (int d = 0; d < 3; d++) adj = adj && std::abs(rc[d] - tc[d]) <= 1;
        const bool in_v = (vg.mask[size_t(g) * 32 + tau * 4 + (r >> 6)] >> (r & 63)) & 1;
        if (in_v != (!adj && vg.halo[size_t(g) * 216 + r] >= 0)) bad = 1;
      }
    }
  if (bad) return fail(-6, "V list differs from the non-adjacent halo of its sibling group");
  // per group: halo slots ordered by node index (the V lists' order)
  vg.hord.assign(size_t(ngrp) * 216, -1);
#pragma omp parallel for schedule(static)
  for (int g = 0; g < ngrp; g++) {
    std::vector<std::pair<int, int>> nodes;
    for (int r = 0; r < 216; r++) {
      int sl = vg.halo[size_t(g) * 216 + r];
      if (sl >= 0) nodes.push_back({vg.src_node[sl], r});
    }
    std::sort(nodes.begin(), nodes.end());
    for (size_t k = 0; k < nodes.size(); k++) vg.hord[size_t(g) * 216 + k] = nodes[k].second;
  }
  vg.psrc.assign(size_t(ngrp) * kPairStride, -1);
  vg.pslot.assign(size_t(ngrp) * kPairStride, 0);
  vg.pcnt.assign(ngrp, 0);
#pragma omp parallel for schedule(static)
  for (int g = 0; g < ngrp; g++) {
    int k = 0;
    for (; k < 216 && vg.hord[size_t(g) * 216 + k] >= 0; k++) {
      const int r = vg.hord[size_t(g) * 216 + k];
      const int rx = r / 36, ry = (r / 6) % 6, rz = r % 6;  // halo coordinate + 2
      const int slot = (rx + 1) * 49 + (ry + 1) * 7 + (rz + 1);  // offset (r - 2) + 3
      int tm = 0;  // targets of the group whose V list holds this source
      for (int tt = 0; tt < 8; tt++)
        if ((vg.mask[size_t(g) * 32 + tt * 4 + (r >> 6)] >> (r & 63)) & 1) tm |= 1 << tt;
      vg.psrc[size_t(g) * kPairStride + k] = vg.halo[size_t(g) * 216 + r];
      vg.pslot[size_t(g) * kPairStride + k] = r | slot << 8 | tm << 20;
    }
    vg.pcnt[g] = (k + 7) / 8 * 8;
  }
  return 0;
}

int FMM3d::upload_state() {
  const Tree& t = tree_;
  const Operators& op = *ops_;
  const int N = t.n_nodes, SD = srcDOF(), TD = trgDOF();
  int rc;
#define KIFMM_UP(buf, vec, what) \
  if ((rc = check((buf).upload((vec).data(), sizeof((vec)[0]) * (vec).size()), what))) return rc;
  // packed points (leaf order)
  std::vector<double> sp(3 * ns_), sn(3 * ns_), tp(3 * nt_);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < ns_; k++) {
    int64_t o = t.src_perm[k];
    for (int d = 0; d < 3; d++) { sp[3 * k + d] = spos_[3 * o + d]; sn[3 * k + d] = snor_[3 * o + d]; }
  }
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nt_; k++) {
    int64_t o = t.trg_perm[k];
    for (int d = 0; d < 3; d++) tp[3 * k + d] = tpos_[3 * o + d];
  }
  KIFMM_UP(spos_d_, sp, "upload sources");
  KIFMM_UP(snor_d_, sn, "upload normals");
  KIFMM_UP(tpos_d_, tp, "upload targets");
  KIFMM_UP(sperm_d_, t.src_perm, "upload source permutation");
  {
    // the densities are scattered into leaf order: the caller's array is read in
    // order and each value is written once (random reads of 8 bytes cost a
    // whole memory burst each)
    std::vector<int> inv(size_t(std::max<int64_t>(1, ns_)), -1);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < ns_; k++) inv[size_t(t.src_perm[size_t(k)])] = int(k);
    int64_t holes = 0;
#pragma omp parallel for schedule(static) reduction(+ : holes)
    for (int64_t k = 0; k < ns_; k++) holes += inv[size_t(k)] < 0;
    if (holes) return fail(-6, "source permutation is not a permutation");
    KIFMM_UP(sinv_d_, inv, "upload source permutation");
  }
  KIFMM_UP(tperm_d_, t.trg_perm, "upload target permutation");
  if ((rc = check(sden_d_.alloc(sizeof(double) * std::max<int64_t>(1, ns_ * SD)), "alloc densities"))) return rc;
  if ((rc = check(tval_d_.alloc(sizeof(double) * std::max<int64_t>(1, nt_ * TD)), "alloc values"))) return rc;
  // node geometry
  std::vector<double4> geom(N);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double c[3];
    t.center(i, c);
    geom[i] = make_double4(c[0], c[1], c[2], t.radius(i));
  }
  KIFMM_UP(geom_d_, geom, "upload geometry");
  KIFMM_UP(depth_d_, t.depth, "upload depths");
  KIFMM_UP(sbeg_d_, t.src_beg, "upload ranges");
  KIFMM_UP(snum_d_, t.src_num, "upload ranges");
  rhs_cap_ = 1;
  // per-node equivalent / check buffers
  const size_t n = size_t(N);
  if ((rc = check(ue_d_.alloc(sizeof(double) * n * op.dat(UE)), "alloc UE"))) return rc;
  if ((rc = check(uc_d_.alloc(sizeof(double) * n * op.dat(UC)), "alloc UC"))) return rc;
  if ((rc = check(de_d_.alloc(sizeof(double) * n * op.dat(DE)), "alloc DE"))) return rc;
  if ((rc = check(dc_d_.alloc(sizeof(double) * n * op.dat(DC)), "alloc DC"))) return rc;
  // tasks
  KIFMM_UP(near_chunks_d_, near_chunks_h_, "upload near chunks");
  KIFMM_UP(near_chunks_mx_d_, near_chunks_mx_h_, "upload near chunks");
  KIFMM_UP(near_order_mx_d_, near_order_mx_h_, "upload near chunk order");
  if (!queue_d_.bytes() && (rc = check(queue_d_.alloc(256), "alloc work-queue counter"))) return rc;
  KIFMM_UP(near_items_d_, near_items_h_, "upload near items");
  KIFMM_UP(s2m_tasks_d_, s2m_tasks_h_, "upload S2M tasks");
  KIFMM_UP(s2m_leaf_d_, s2m_leaf_h_, "upload S2M tasks");
  KIFMM_UP(ue_zero_d_, ue_zero_h_, "upload zero rows");
  KIFMM_UP(dc_zero_d_, dc_zero_h_, "upload zero rows");
  KIFMM_UP(s2l_tasks_d_, s2l_tasks_h_, "upload S2L tasks");
  KIFMM_UP(surf_items_d_, surf_items_h_, "upload surface items");
  KIFMM_UP(lvl_idx_d_, lvl_idx_h_, "upload level columns");
  KIFMM_UP(scales_d_, scales_h_, "upload level scales");
  std::vector<double> ones(8, 1.0);
  KIFMM_UP(ones_d_, ones, "upload ones");
  // operators
  for (int s = 0; s < 4; s++) KIFMM_UP(sam_d_[s], op.sam[s], "upload surfaces");
  {
    std::vector<int> shell_blk(op.shell.size());
    const int np = op.np, nn = op.n, o = np / 2;
    for (size_t q = 0; q < op.shell.size(); q++) {
      int g = op.shell[q];
      int c = g % nn - o, b = (g / nn) % nn - o, a = g / (nn * nn) - o;
      shell_blk[q] = (a * np + b) * np + c;
    }
    KIFMM_UP(shell_d_, shell_blk, "upload shell");
    std::vector<double2> tw(nn);
    for (int m = 0; m < nn; m++) tw[m] = make_double2(std::cos(2.0 * M_PI * m / nn), -std::sin(2.0 * M_PI * m / nn));
    KIFMM_UP(twiddle_d_, tw, "upload twiddles");
    std::vector<double> deg(op.deg, op.deg + op.sd);
    KIFMM_UP(deg_d_, deg, "upload degrees");
  }
  KIFMM_UP(pup_d_, op.pinv_up, "upload pinv up");
  KIFMM_UP(pdn_d_, op.pinv_dn, "upload pinv down");
  KIFMM_UP(m2m_d_, op.m2m, "upload M2M");
  if (precision_ == PREC_MIXED) {
    // W = P_up M (nUE x 8 nUE, column-major), formed once in FP64
    const int ue = op.dat(UE), uc = op.dat(UC);
    std::vector<double> w(size_t(ue) * 8 * ue, 0.0);
#pragma omp parallel for
    for (int k = 0; k < 8 * ue; k++)
      for (int j = 0; j < uc; j++) {
        const double mjk = op.m2m[size_t(k) * uc + j];
        for (int m = 0; m < ue; m++) w[size_t(k) * ue + m] += op.pinv_up[size_t(j) * ue + m] * mjk;
      }
    KIFMM_UP(wup_d_, w, "upload P_up M2M");
    // W_z = [P_dn | P_dn L2L_z] (nDE x (nDC + nDE), column-major), z = 0..7
    const int de = op.dat(DE), dc = op.dat(DC), kw = dc + de;
    std::vector<double> wd(size_t(8) * de * kw, 0.0);
#pragma omp parallel for
    for (int zk = 0; zk < 8 * kw; zk++) {
      const int z = zk / kw, k = zk % kw;
      double* col = wd.data() + (size_t(z) * kw + k) * de;
      if (k < dc) {
        for (int m = 0; m < de; m++) col[m] = op.pinv_dn[size_t(k) * de + m];
      } else {
        const double* l = op.l2l.data() + size_t(z) * dc * de + size_t(k - dc) * dc;
        for (int j = 0; j < dc; j++)
          for (int m = 0; m < de; m++) col[m] += op.pinv_dn[size_t(j) * de + m] * l[j];
      }
    }
    KIFMM_UP(wdn_d_, wd, "upload P_dn L2L");
  } else {
    wup_d_.release();
    wdn_d_.release();
  }
  KIFMM_UP(l2l_d_, op.l2l, "upload L2L");
  if (precision_ == PREC_MIXED && trans_terms_ > 0) {
    // the translation tables as three BF16 parts each, row-major and padded
    auto split_table = [&](const std::vector<double>& a, int nz, int M, int K, int& Mp, int& Kp,
                           DeviceBuffer& out) -> int {
      Mp = (M + 15) / 16 * 16;
      Kp = (K + 31) / 32 * 32;
      const size_t per = size_t(nz) * Mp * Kp;
      std::vector<uint16_t> parts(3 * per, 0);
#pragma omp parallel for
      for (int zm = 0; zm < nz * Mp; zm++) {
        const int zz = zm / Mp, m = zm % Mp;
        for (int k = 0; k < Kp; k++) {
          const double v = m < M && k < K ? a[size_t(zz) * M * K + size_t(k) * M + m] : 0.0;
          uint16_t p[3];
          split_bf16x3(v, p);
          for (int q = 0; q < 3; q++) parts[size_t(q) * per + size_t(zm) * Kp + k] = p[q];
        }
      }
      return check(out.upload(parts.data(), parts.size() * sizeof(uint16_t)), "upload BF16 translation table");
    };
    if ((rc = split_table(op.m2m, 1, op.dat(UC), 8 * op.dat(UE), m2m_mp_, m2m_kp_, m2m_bf_d_))) return rc;
    if ((rc = split_table(op.l2l, 8, op.dat(DC), op.dat(DE), l2l_mp_, l2l_kp_, l2l_bf_d_))) return rc;
  } else {
    m2m_bf_d_.release();
    l2l_bf_d_.release();
  }
  if (precision_ == PREC_MIXED) {
    // split-K workspace: the same split rule as the evaluate over every step
    size_t need = 1;
    for (const auto* steps : {&up_steps_, &dn_steps_})
      for (const GemmStep& st : *steps) {
        GemmArgs g{};
        step_shape(op, st.mat, g.M, g.K);
        g.ncol = st.ncol; g.max_zcol = st.ncol;
        if (st.zoff_off >= 0) { g.zoff = &st.ncol; g.nz = 8; g.max_zcol = st.max_zcol; }  // only tested for null
        split_k(g);
        if (g.ksplit > 1) need = std::max(need, size_t(g.ksplit) * size_t(g.ncol) * size_t(g.M));
      }
    if ((rc = reserve_splitk(need))) return rc;
  }
  KIFMM_UP(m2l_d_, op.m2l, "upload M2L");
  // V list
  if ((rc = upload_vgroups(vfwd_))) return rc;
  {
    const size_t ns = vfwd_.src_node.size() * op.sd, nt = vfwd_.trg_node.size() * op.td;
    if (precision_ == PREC_FP64) {
      if ((rc = check(dspec_d_.alloc(sizeof(double2) * ns * op.H), "alloc V spectra"))) return rc;
      if ((rc = check(espec_d_.alloc(sizeof(double2) * nt * op.H), "alloc V spectra"))) return rc;
      m2l32_d_.release(); dspec32_d_.release(); espec32_d_.release();
    } else {
      // the FP32 table without the kernel grid's wrap entries (equal in every
      // product an evaluate forms) when its opposite offsets are then
      // conjugates (even kernels: 111, 511; not 301, whose fourth column is odd)
      const std::vector<double> nowrap = m2l_table_nowrap(op);
      m2l_mirror_ = table_mirror(nowrap, op.sd * op.td, op.H);
      const std::vector<float> m2l32 = pack_spectra32(m2l_mirror_ ? nowrap : op.m2l, op.np);
      KIFMM_UP(m2l32_d_, m2l32, "upload M2L (FP32)");
      m2l_sym3_ = table_sym3(op.m2l, op.sd, op.td, op.H);
      const size_t hc = size_t(spec_packed_len(op.np));
      if ((rc = check(dspec32_d_.alloc(sizeof(float2) * ns * hc), "alloc V spectra"))) return rc;
      if ((rc = check(espec32_d_.alloc(sizeof(float2) * nt * hc), "alloc V spectra"))) return rc;
      dspec_d_.release(); espec_d_.release();
    }
  }
#undef KIFMM_UP
  if (events_.empty()) {
    events_.resize(PH_COUNT + 3);  // phase boundaries, then the Hadamard's start and end
    for (auto& ev : events_)
      if ((rc = check(hipEventCreate(&ev), "create events"))) return rc;
  }
  return 0;
}

int FMM3d::upload_vgroups(VGroups& vg) {
  int rc;
#define KIFMM_UPV(buf, vec, what) \
  if ((rc = check((buf).upload((vec).data(), sizeof((vec)[0]) * (vec).size()), what))) return rc;
  KIFMM_UPV(vg.src_node_d, vg.src_node, "upload V sources");
  KIFMM_UPV(vg.trg_node_d, vg.trg_node, "upload V targets");
  KIFMM_UPV(vg.halo_d, vg.halo, "upload V halo");
  KIFMM_UPV(vg.hord_d, vg.hord, "upload V order");
  KIFMM_UPV(vg.psrc_d, vg.psrc, "upload V pairs");
  KIFMM_UPV(vg.pslot_d, vg.pslot, "upload V pairs");
  KIFMM_UPV(vg.pcnt_d, vg.pcnt, "upload V pairs");
  KIFMM_UPV(vg.tgt_d, vg.tgt, "upload V groups");
  KIFMM_UPV(vg.mask_d, vg.mask, "upload V masks");
#undef KIFMM_UPV
  return 0;
}

// Group tables and operators of a V pass; the caller sets the input
// (UE, sz_ue: equivalent densities of the sources) and the output (DC, sz_dc:
// check values of the targets, added to).
void FMM3d::vlist_args(VlistArgs& v, const VGroups& vg) const {
  const Operators& op = *ops_;
  v.np = op.np; v.n = op.n; v.sd = op.sd; v.td = op.td;
  v.H = precision_ == PREC_MIXED ? spec_packed_len(op.np) : op.H;  // FP32 spectra are packed
  v.nsrc = int(vg.src_node.size()); v.ntrg = int(vg.trg_node.size()); v.ngrp = vg.ngrp();
  v.src_node = vg.src_node_d.as<int>(); v.trg_node = vg.trg_node_d.as<int>();
  v.depth = depth_d_.as<int>(); v.root_level = tree_.root_level;
  v.deg = deg_d_.as<double>(); v.shell = shell_d_.as<int>(); v.twiddle = twiddle_d_.as<double2>();
  v.halo = vg.halo_d.as<int>(); v.hord = vg.hord_d.as<int>();
  v.psrc = vg.psrc_d.as<int>(); v.pslot = vg.pslot_d.as<int>(); v.pcnt = vg.pcnt_d.as<int>();
  v.mask = vg.mask_d.as<uint64_t>(); v.tgt = vg.tgt_d.as<int>();
  v.m2l = m2l_d_.as<double2>(); v.dspec = dspec_d_.as<double2>(); v.espec = espec_d_.as<double2>();
  v.had_form = tune_.vec_hadamard;
  v.sym3 = m2l_sym3_ ? 1 : 0;
  v.mirror = m2l_mirror_ ? 1 : 0;
}

// timed: record the Hadamard's start and end events (forward pass with timing on).
int FMM3d::run_vlist(VlistArgs& v, hipStream_t s, bool timed, const float2* m2l32) {
  int rc;
  if (v.ngrp <= 0) return 0;
  hipEvent_t e0 = timed && timing_ ? events_[PH_COUNT + 1] : nullptr;
  hipEvent_t e1 = timed && timing_ ? events_[PH_COUNT + 2] : nullptr;
  if (precision_ != PREC_MIXED) return check(launch_vlist(v, s, e0, e1), "V list");
  if ((rc = check(launch_fft_fwd32(v, dspec32_d_.as<float2>(), s), "V list forward"))) return rc;
  if (e0) (void)hipEventRecord(e0, s);
  if ((rc = check(launch_hadamard_mixed(v, dspec32_d_.as<float2>(), m2l32 ? m2l32 : m2l32_d_.as<float2>(),
                                        espec32_d_.as<float2>(), s),
                  "V list Hadamard")))
    return rc;
  if (e1) (void)hipEventRecord(e1, s);
  return check(launch_fft_inv32(v, espec32_d_.as<float2>(), s), "V list inverse");
}

// ---------------------------------------------------------------------------
// evaluate: fmm3d_eval_mpi.cpp:396-1515
// ---------------------------------------------------------------------------
int FMM3d::reserve_splitk(size_t doubles) {
  if (doubles * sizeof(double) <= splitk_d_.bytes()) return 0;
  drop_graph();  // a captured pass holds the old workspace address
  return check(splitk_d_.alloc(doubles * sizeof(double)), "alloc split-K workspace");
}

// The workspace for a split launch, or null if it does not fit (the launch is
// then rejected; setup reserves enough for every step it plans).
double* FMM3d::splitk_for(const GemmArgs& g) {
  if (g.ksplit <= 1) return nullptr;
  const size_t need = size_t(g.ksplit) * size_t(g.ncol) * size_t(g.M) * sizeof(double);
  return need <= splitk_d_.bytes() ? splitk_d_.as<double>() : nullptr;
}

bool table_sym3(const std::vector<double>& t, int sd, int td, int H) {
  if (sd < 3 || td < 3) return false;
  const size_t per = size_t(sd) * td * H * 2, nslot = t.size() / per;
  double big = 0.0, diff = 0.0;
  for (double x : t) big = std::max(big, std::fabs(x));
  for (size_t sl = 0; sl < nslot; sl++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < j; i++) {
        const double* a = t.data() + sl * per + size_t(j * td + i) * H * 2;
        const double* b = t.data() + sl * per + size_t(i * td + j) * H * 2;
        for (int f = 0; f < 2 * H; f++) diff = std::max(diff, std::fabs(a[f] - b[f]));
      }
  return diff <= 1e-12 * big;
}

std::vector<float> pack_spectra32(const std::vector<double>& full, int np) {
  const int n = 2 * np, hh = np + 1, h = n * n * hh, hc = spec_packed_len(np);  // padding stays zero
  const size_t rows = full.size() / (2 * size_t(h));
  std::vector<float> out(rows * size_t(hc) * 2);
  for (size_t r = 0; r < rows; r++)
    for (int ka = 0; ka < n; ka++)
      for (int kb = 0; kb < n; kb++)
        for (int kc = 0; kc < hh; kc++) {
          const int code = spec_pack(np, ka, kb, kc);
          if (code & 1) continue;  // the conjugate of a stored value
          const size_t src = (r * h + size_t((ka * n + kb) * hh + kc)) * 2, dst = (r * hc + size_t(code >> 1)) * 2;
          out[dst] = float(full[src]);
          out[dst + 1] = float(full[src + 1]);
        }
  return out;
}

void split_k(GemmArgs& g) {
  // A launch's time is its tile count over the CUs times the serial K loop
  // (~2-3 us per 16-deep slab when the launch is small: the slab's loads are
  // not hidden).  Splitting K pays where that loop is long and the tiles few:
  // the top levels (<= 16 tiles, any K: up to 64 chunks, down to one slab
  // each) and the fused or plain M2M with K = 8 nUE >= 1024 up to two tiles per
  // CU (<= 512 tiles; the two deepest parent levels of the 8M-point workload
  // have ~165 tiles and ran one wave per SIMD unsplit).  Elsewhere the
  // workspace round trip and the summing launch cost more than the split saves.
  const int64_t base = int64_t((g.M + 159) / 160) * ((g.max_zcol + 63) / 64) * (g.zoff ? g.nz : 1);
  int64_t want = 0, cap = 64;
  if (base <= 16) want = 128;
  else if (g.K >= 1024 && base <= 512) { want = 512; cap = 8; }
  if (want == 0 || g.K <= 16) return;
  int ks = int(std::min<int64_t>({cap, (want + base - 1) / base, (g.K + 15) / 16}));
  if (ks <= 1) return;
  int chunk = ((g.K + ks - 1) / ks + 15) / 16 * 16;
  g.kchunk = chunk;
  g.ksplit = (g.K + chunk - 1) / chunk;
  if (g.ksplit <= 1) { g.ksplit = 0; g.kchunk = 0; }
}

void split_k_rhs(GemmArgs& g, int nrhs) {
  GemmArgs g1 = g;
  g1.ncol = g.ncol / nrhs;
  g1.max_zcol = g.max_zcol / nrhs;
  split_k(g1);
  g.ksplit = g1.ksplit;
  g.kchunk = g1.kchunk;
}

int FMM3d::evaluate(const double* d_src_den, double* d_trg_val, hipStream_t s) {
  if (!setup_done_) return fail(-3, "evaluate called before setup");
  if (!use_graph_ || !s || timing_) return evaluate_impl(d_src_den, d_trg_val, s, false);
  return run_graph(d_src_den, d_trg_val, s, 1);
}

int FMM3d::evaluateMulti(const double* d_src_den, double* d_trg_val, int nrhs, hipStream_t s) {
  if (!setup_done_) return fail(-3, "evaluate called before setup");
  if (nrhs < 0) return fail(-2, "negative nrhs");
  if (nrhs == 0) return 0;
  if (nrhs == 1) return evaluate(d_src_den, d_trg_val, s);
  int rc;
  if ((rc = reserve_rhs(nrhs))) return rc;
  if ((rc = prepare_rhs_steps(nrhs))) return rc;
  if (!use_graph_ || !s || timing_) return evaluate_impl(d_src_den, d_trg_val, s, false, nrhs);
  return run_graph(d_src_den, d_trg_val, s, nrhs);
}

// Replays the captured pass when (buffers, stream, nrhs) match the capture,
// else captures anew.
int FMM3d::run_graph(const double* d_src_den, double* d_trg_val, hipStream_t s, int nrhs) {
  if (!(graph_exec_ && graph_src_ == d_src_den && graph_dst_ == d_trg_val && graph_stream_ == s &&
        graph_nrhs_ == nrhs)) {
    drop_graph();
    int rc = check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin graph capture");
    if (rc) return rc;
    rc = evaluate_impl(d_src_den, d_trg_val, s, false, nrhs);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if ((rc = check(e, "end graph capture"))) return rc;
    graph_ = g;
    if ((rc = check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "instantiate graph"))) {
      drop_graph();
      return rc;
    }
    graph_src_ = d_src_den;
    graph_dst_ = d_trg_val;
    graph_stream_ = s;
    graph_nrhs_ = nrhs;
  }
  return check(hipGraphLaunch(graph_exec_, s), "graph launch");
}

// Right-hand-side buffers for nrhs: node buffers, densities, values and
// spectra hold nrhs blocks (allocated once per larger nrhs; a captured graph
// holds the old addresses, so it is dropped).
int FMM3d::reserve_rhs(int nrhs) {
  if (nrhs <= rhs_cap_) return 0;
  drop_graph();
  const Operators& op = *ops_;
  const size_t N = size_t(tree_.n_nodes), R = size_t(nrhs);
  const int SD = srcDOF(), TD = trgDOF();
  int rc;
  auto grow = [&](DeviceBuffer& b, size_t bytes, const char* what) {
    return b.bytes() >= bytes ? 0 : check(b.alloc(bytes), what);
  };
  if ((rc = grow(ue_d_, sizeof(double) * R * N * op.dat(UE), "alloc UE"))) return rc;
  if ((rc = grow(uc_d_, sizeof(double) * R * N * op.dat(UC), "alloc UC"))) return rc;
  if ((rc = grow(de_d_, sizeof(double) * R * N * op.dat(DE), "alloc DE"))) return rc;
  if ((rc = grow(dc_d_, sizeof(double) * R * N * op.dat(DC), "alloc DC"))) return rc;
  if ((rc = grow(sden_d_, sizeof(double) * R * std::max<int64_t>(1, ns_ * SD), "alloc densities"))) return rc;
  if ((rc = grow(tval_d_, sizeof(double) * R * std::max<int64_t>(1, nt_ * TD), "alloc values"))) return rc;
  const size_t nsv = vfwd_.src_node.size() * op.sd * R, ntv = vfwd_.trg_node.size() * op.td * R;
  if (precision_ == PREC_FP64) {
    if ((rc = grow(dspec_d_, sizeof(double2) * nsv * op.H, "alloc V spectra"))) return rc;
    if ((rc = grow(espec_d_, sizeof(double2) * ntv * op.H, "alloc V spectra"))) return rc;
  } else {
    const size_t hc = size_t(spec_packed_len(op.np));
    if ((rc = grow(dspec32_d_, sizeof(float2) * nsv * hc, "alloc V spectra"))) return rc;
    if ((rc = grow(espec32_d_, sizeof(float2) * ntv * hc, "alloc V spectra"))) return rc;
  }
  rhs_cap_ = nrhs;
  return 0;
}

// The level GEMM steps over nrhs right-hand sides: the columns of right-hand
// side r are the single-RHS columns shifted to node rows r * n_nodes; the 8
// L2L matrices keep their column blocks contiguous (matrix z: all right-hand
// sides' columns of z).
int FMM3d::prepare_rhs_steps(int nrhs) {
  if (nrhs == rhs_steps_n_ && !up_steps_r_.empty()) return 0;
  drop_graph();  // the column tables are uploaded anew
  const int N = tree_.n_nodes;
  lvl_idx_r_h_.clear();
  auto expand = [&](const std::vector<GemmStep>& in, std::vector<GemmStep>& out) {
    out.clear();
    for (const GemmStep& st : in) {
      GemmStep o = st;
      const int* b = lvl_idx_h_.data() + st.idx_off;
      const int* c = b + st.ncol;
      o.ncol = st.ncol * nrhs;
      o.idx_off = int64_t(lvl_idx_r_h_.size());
      if (st.zoff_off < 0) {
        for (const int* arr : {b, c})
          for (int r = 0; r < nrhs; r++)
            for (int k = 0; k < st.ncol; k++) lvl_idx_r_h_.push_back(arr[k] + r * N);
        o.zoff_off = -1;
        o.max_zcol = 0;
      } else {
        const int* zo = lvl_idx_h_.data() + st.zoff_off;
        for (const int* arr : {b, c})
          for (int z = 0; z < 8; z++)
            for (int r = 0; r < nrhs; r++)
              for (int k = zo[z]; k < zo[z + 1]; k++) lvl_idx_r_h_.push_back(arr[k] + r * N);
        o.zoff_off = int64_t(lvl_idx_r_h_.size());
        for (int z = 0; z < 9; z++) lvl_idx_r_h_.push_back(zo[z] * nrhs);
        o.max_zcol = st.max_zcol * nrhs;
      }
      out.push_back(o);
    }
  };
  expand(up_steps_, up_steps_r_);
  expand(dn_steps_, dn_steps_r_);
  int rc;
  if ((rc = check(lvl_idx_r_d_.upload(lvl_idx_r_h_.data(), sizeof(int) * lvl_idx_r_h_.size()), "upload level columns")))
    return rc;
  if (precision_ == PREC_MIXED) {
    const Operators& op = *ops_;
    size_t need = 1;
    for (const auto* steps : {&up_steps_r_, &dn_steps_r_})
      for (const GemmStep& st : *steps) {
        GemmArgs g{};
        step_shape(op, st.mat, g.M, g.K);
        g.ncol = st.ncol; g.max_zcol = st.ncol;
        if (st.zoff_off >= 0) { g.zoff = &st.ncol; g.nz = 8; g.max_zcol = st.max_zcol; }
        split_k_rhs(g, nrhs);
        if (g.ksplit > 1) need = std::max(need, size_t(g.ksplit) * size_t(g.ncol) * size_t(g.M));
      }
    if ((rc = reserve_splitk(need))) return rc;
  }
  rhs_steps_n_ = nrhs;
  return 0;
}

int FMM3d::field(const double* d_src_den, double* d_trg_grad, hipStream_t s) {
  if (!setup_done_) return fail(-3, "field called before setup");
  if (kt_ != 111) return fail(-2, "the field pass supports the Laplace single layer (kernel 111) only");
  if (tgrad_d_.bytes() < sizeof(double) * 3 * size_t(std::max<int64_t>(1, nt_))) {
    int rc = check(tgrad_d_.alloc(sizeof(double) * 3 * size_t(std::max<int64_t>(1, nt_))), "alloc field");
    if (rc) return rc;
  }
  return evaluate_impl(d_src_den, d_trg_grad, s, true);
}

// grad: the same far-field pass, then the gradient of the near-field and local
// expansion kernels at the targets (3 values per target) instead of the values.
int64_t FMM3d::checkpointSize() const { return ops_ ? int64_t(tree_.n_nodes) * ops_->dat(UE) : 0; }

int FMM3d::evaluateCheckpoint(const double* d_src_den, double* d_trg_val, double* d_ue_out, hipStream_t s) {
  if (!setup_done_) return fail(-3, "evaluate called before setup");
  if (!d_ue_out) return fail(-2, "null checkpoint pointer");
  return evaluate_impl(d_src_den, d_trg_val, s, false, 1, nullptr, d_ue_out);
}

int FMM3d::fieldFromCheckpoint(const double* d_src_den, const double* d_ue_in, double* d_trg_grad, hipStream_t s) {
  if (!setup_done_) return fail(-3, "field called before setup");
  if (kt_ != 111) return fail(-2, "the field pass supports the Laplace single layer (kernel 111) only");
  if (!d_ue_in) return fail(-2, "null checkpoint pointer");
  if (tgrad_d_.bytes() < sizeof(double) * 3 * size_t(std::max<int64_t>(1, nt_))) {
    int rc = check(tgrad_d_.alloc(sizeof(double) * 3 * size_t(std::max<int64_t>(1, nt_))), "alloc field");
    if (rc) return rc;
  }
  return evaluate_impl(d_src_den, d_trg_grad, s, true, 1, d_ue_in, nullptr);
}

int FMM3d::evaluate_impl(const double* d_src_den, double* d_trg_val, hipStream_t s, bool grad, int nrhs,
                         const double* ue_in, double* ue_out) {
  if ((ns_ > 0 && !d_src_den) || (nt_ > 0 && !d_trg_val)) return fail(-2, "null density or value pointer");
  const int R = nrhs;
  if (R < 1 || R > rhs_cap_ || (R > 1 && (grad || rhs_steps_n_ != R))) return fail(-6, "right-hand sides not prepared");
  const Tree& t = tree_;
  const Operators& op = *ops_;
  const int N = t.n_nodes, SD = srcDOF(), TD = trgDOF();
  const KParams<double> pu = make_params(kt_, coefs_.data());
  const KParams<double> pm = make_params(op.kt_mm, op.coefs.data());
  int rc;
  auto mark = [&](int ph) { if (timing_) (void)hipEventRecord(events_[ph], s); };
  const bool mixed = precision_ == PREC_MIXED;
  const bool fused_out = mixed && kt_ == 111 && !grad;

  mark(PH_INI);
  // zero the work vectors (:425-438)
  const int64_t NR = int64_t(N) * R;  // node rows over the right-hand sides
  // mixed path: only the rows read before any write (the rest is stored first)
  if (!ue_in && (rc = check(mixed ? launch_zero_rows(ue_d_.as<double>(), ue_zero_d_.as<int>(), int(ue_zero_h_.size()),
                                                     op.dat(UE), R, N, s)
                                  : launch_fill(ue_d_.as<double>(), NR * op.dat(UE), 0.0, s),
                            "zero UE")))
    return rc;
  // mixed path: UC is read only for the contributor leaves, whose S2M task
  // stores it whole (the fused up pass forms no UC for parents)
  if (!mixed && (rc = check(launch_fill(uc_d_.as<double>(), NR * op.dat(UC), 0.0, s), "zero UC")))
    return rc;
  // DE: the downward pseudo-inverse stores it for every evaluator node at depth
  // >= 2, and only those rows are read (L2L, L2T); the mixed path skips the fill
  if (!mixed && (rc = check(launch_fill(de_d_.as<double>(), NR * op.dat(DE), 0.0, s), "zero DE")))
    return rc;
  if ((rc = check(mixed ? launch_zero_rows(dc_d_.as<double>(), dc_zero_d_.as<int>(), int(dc_zero_h_.size()), op.dat(DC),
                                           R, N, s)
                        : launch_fill(dc_d_.as<double>(), NR * op.dat(DC), 0.0, s),
                  "zero DC")))
    return rc;
  if ((rc = check(launch_load_rhs(d_src_den, sinv_d_.as<int>(), sden_d_.as<double>(), ns_, SD, R, s),
                  "load densities")))
    return rc;

  SurfArgs sa{};
  sa.items = surf_items_d_.as<int>();
  sa.spos = spos_d_.as<double>(); sa.snor = snor_d_.as<double>(); sa.sden = sden_d_.as<double>();
  sa.src_beg = sbeg_d_.as<int>(); sa.src_num = snum_d_.as<int>();
  sa.geom = geom_d_.as<double4>();
  sa.nrhs = R;
  sa.nstride = N;

  double* bufs[4] = {ue_d_.as<double>(), uc_d_.as<double>(), de_d_.as<double>(), dc_d_.as<double>()};
  const int szs[4] = {op.dat(UE), op.dat(UC), op.dat(DE), op.dat(DC)};
  const int* lvl_idx = R > 1 ? lvl_idx_r_d_.as<int>() : lvl_idx_d_.as<int>();
  auto run_steps = [&](const std::vector<GemmStep>& steps) -> int {
    for (const GemmStep& st : steps) {
      GemmArgs g{};
      step_shape(op, st.mat, g.M, g.K);
      switch (st.mat) {
        case 0: g.A = m2m_d_.as<double>(); break;
        case 1: g.A = pup_d_.as<double>(); break;
        case 8: g.A = wup_d_.as<double>(); break;
        case 2: g.A = pdn_d_.as<double>(); break;
        case 9:
          // DE = [P_dn | P_dn L2L_z] [DC; DE of the parent]: the first K part
          // is the output node's DC row
          g.A = wdn_d_.as<double>();
          g.B1 = bufs[3]; g.ldb1 = szs[3]; g.k1 = op.dat(DC);
          break;
        default: g.A = l2l_d_.as<double>();
      }
      g.max_zcol = st.ncol;
      if (st.zoff_off >= 0) {
        g.zoff = lvl_idx + st.zoff_off;
        g.nz = 8;
        g.a_zstride = int64_t(g.M) * g.K;
        g.max_zcol = st.max_zcol;
      }
      g.B = bufs[st.src]; g.ldb = szs[st.src];
      g.C = bufs[st.dst]; g.ldc = szs[st.dst];
      g.bidx = lvl_idx + st.idx_off;
      g.cidx = lvl_idx + st.idx_off + st.ncol;
      if (st.ks_off >= 0) { g.kscale = scales_d_.as<double>() + st.ks_off; g.ksd = op.sd; }
      else { g.kscale = ones_d_.as<double>(); g.ksd = 1; }
      if (st.rs_off >= 0) { g.rscale = scales_d_.as<double>() + st.rs_off; g.rsd = op.sd; }
      else { g.rscale = ones_d_.as<double>(); g.rsd = 1; }
      g.ncol = st.ncol; g.accumulate = st.accumulate;
      if (mixed && trans_terms_ > 0 && (st.mat == 0 || st.mat == 3)) {
        // translations as split-BF16 products (the pseudo-inverses stay FP64)
        const bool m2m = st.mat == 0;
        int r = check(launch_gemm_bf16s(g, m2m ? m2m_bf_d_.as<void>() : l2l_bf_d_.as<void>(), m2m ? m2m_mp_ : l2l_mp_,
                                        m2m ? m2m_kp_ : l2l_kp_, trans_terms_, s),
                      "BF16 translation GEMM");
        if (r) return r;
        continue;
      }
      // mixed path: launches with very few workgroups split K into chunks of
      // whole 16-deep slabs; the chunks' partial products go to the workspace
      // and are summed in chunk order into the output
      if (mixed) {
        split_k_rhs(g, R);
        g.work = splitk_for(g);
        if (g.ksplit > 1 && !g.work) return fail(-6, "split-K workspace smaller than the launch needs");
      }
      int r = check(mixed ? launch_gemm_fma(g, s) : launch_gemm(g, s), "level GEMM");
      if (r) return r;
    }
    return 0;
  };

  // up pass (:480-619)
  mark(PH_UPW);
  sa.tasks = s2m_tasks_d_.as<SurfTask>(); sa.ntasks = int(s2m_tasks_h_.size());
  sa.leaf = mixed ? s2m_leaf_d_.as<int4>() : nullptr;
  sa.sam = sam_d_[UC].as<double>(); sa.n_sam = op.n_sam[UC];
  sa.out = uc_d_.as<double>(); sa.sz_out = op.dat(UC);
  if (!ue_in) {
    if ((rc = check(mixed ? launch_surface_mixed(kt_, sa, pu, s) : launch_surface(kt_, sa, pu, s), "S2M")))
      return rc;
    if ((rc = run_steps(R > 1 ? up_steps_r_ : up_steps_))) return rc;
  }
  if (ue_out &&
      (rc = check(hipMemcpyAsync(ue_out, ue_d_.as<double>(), sizeof(double) * size_t(N) * op.dat(UE),
                                 hipMemcpyDeviceToDevice, s), "checkpoint UE")))
    return rc;
  const double* ue_src = ue_in ? ue_in : ue_d_.as<double>();

  // V list (:1007-1059)
  mark(PH_VLIST);
  VlistArgs v{};
  vlist_args(v, vfwd_);
  v.UE = ue_src; v.sz_ue = op.dat(UE);
  v.DC = dc_d_.as<double>(); v.sz_dc = op.dat(DC);
  v.dc_store = 1;  // DC was zeroed above and the V list writes it first (S2L adds after)
  v.nrhs = R;
  v.nstride = N;
  if ((rc = run_vlist(v, s, true))) return rc;

  // X list S2L (:1233-1265)
  mark(PH_XLIST);
  sa.tasks = s2l_tasks_d_.as<SurfTask>(); sa.ntasks = int(s2l_tasks_h_.size());
  sa.leaf = nullptr;
  sa.sam = sam_d_[DC].as<double>(); sa.n_sam = op.n_sam[DC];
  sa.out = dc_d_.as<double>(); sa.sz_out = op.dat(DC);
  if ((rc = check(mixed ? launch_surface_mixed(kt_, sa, pu, s) : launch_surface(kt_, sa, pu, s), "S2L"))) return rc;

  // down pass (:1267-1382)
  mark(PH_DOWN);
  if ((rc = run_steps(R > 1 ? dn_steps_r_ : dn_steps_))) return rc;

  // U, W, X-fallback and L2T at the targets (:653-794, :1062-1231, :1383-1435)
  mark(PH_NEAR);
  NearArgs na{};
  na.chunks = near_chunks_d_.as<NearChunk>(); na.nchunks = int(near_chunks_h_.size());
  if (mixed && !grad) {
    na.chunks = near_chunks_mx_d_.as<NearChunk>();
    na.nchunks = int(near_chunks_mx_h_.size());
    if (tune_.near_mode >= 1) na.order = near_order_mx_d_.as<int>();
    if (tune_.near_mode == 2) na.queue = queue_d_.as<int>();
  }
  na.items = near_items_d_.as<int>();
  na.spos = spos_d_.as<double>(); na.snor = snor_d_.as<double>(); na.sden = sden_d_.as<double>();
  na.src_beg = sbeg_d_.as<int>(); na.src_num = snum_d_.as<int>();
  na.geom = geom_d_.as<double4>();
  na.sam_ue = sam_d_[UE].as<double>(); na.n_ue = op.n_sam[UE];
  na.sam_de = sam_d_[DE].as<double>(); na.n_de = op.n_sam[DE];
  na.UE = ue_src; na.sz_ue = op.dat(UE);
  na.DE = de_d_.as<double>(); na.sz_de = op.dat(DE);
  na.tpos = tpos_d_.as<double>(); na.tval = tval_d_.as<double>();
  na.nrhs = R;
  na.nstride = N;
  na.tstride = nt_ * TD;
  if (grad) {
    na.tval = tgrad_d_.as<double>();
    if ((rc = check(launch_near_grad(kt_, na, pu, pm, s), "near field gradient"))) return rc;
  } else {
    // mixed Laplace: the near kernel writes the final values in the caller's order
    if (fused_out) { na.tval = d_trg_val; na.tperm = tperm_d_.as<int>(); na.tstride = nt_; }
    if ((rc = check(mixed ? launch_near_mixed(kt_, na, pu, pm, s) : launch_near(kt_, na, pu, pm, s), "near field")))
      return rc;
  }

  // write back (:1490-1510)
  mark(PH_FIN);
  if (!fused_out &&
      (rc = check(launch_store_rhs(grad ? tgrad_d_.as<double>() : tval_d_.as<double>(), tperm_d_.as<int>(), d_trg_val,
                                   nt_, grad ? 3 : TD, R, s), "scatter")))
    return rc;
  mark(PH_COUNT);
  if (timing_) {
    if ((rc = check(hipEventSynchronize(events_[PH_COUNT]), "sync"))) return rc;
    for (int p = 0; p < PH_COUNT; p++) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, events_[p], events_[p + 1]);
      phase_ms_[p] = ms;
    }
    float ms = 0;
    if (vfwd_.ngrp() > 0) (void)hipEventElapsedTime(&ms, events_[PH_COUNT + 1], events_[PH_COUNT + 2]);
    phase_ms_[PH_HADAMARD] = ms;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Autotuner: each launch-shape choice of the mixed path is timed on this
// handle's own tree and inputs (phase times of warm evaluates, the best of
// two), one choice at a time, and the fastest kept.
// ---------------------------------------------------------------------------
int FMM3d::retask() {
  drop_graph();
  int rc = build_tasks();
  if (rc) return rc;
  rhs_steps_n_ = 1;
  up_steps_r_.clear(); dn_steps_r_.clear(); lvl_idx_r_h_.clear();
  rhs_cap_ = 1;
  return upload_state();
}

int FMM3d::apply_tuning(const Tuning& t) {
  if (t.near_mode < 0 || t.near_mode > 2 || t.near_stage < 0 || t.near_stage > 2 || (t.trans != 0 && t.trans != 6) ||
      t.vec_hadamard < 0 || t.vec_hadamard > 2)
    return fail(-2, "tuning values out of range");
  if (setup_done_ && t.vec_hadamard != tune_.vec_hadamard) drop_graph();
  if (!setup_done_) { tune_ = t; trans_terms_ = t.trans; return 0; }
  const bool new_tasks = t.trans != trans_terms_;
  const bool new_chunks = t.near_stage != tune_.near_stage;
  tune_ = t;
  trans_terms_ = t.trans;
  int rc = 0;
  if (new_tasks) {
    if ((rc = retask())) { setup_done_ = false; return rc; }
  } else if (new_chunks) {
    drop_graph();
    build_mixed_chunks();
    if ((rc = check(near_chunks_mx_d_.upload(near_chunks_mx_h_.data(), sizeof(NearChunk) * near_chunks_mx_h_.size()),
                    "upload near chunks")) ||
        (rc = check(near_order_mx_d_.upload(near_order_mx_h_.data(), sizeof(int) * near_order_mx_h_.size()),
                    "upload near chunk order")))
      return rc;
  }
  return 0;
}

int FMM3d::autotune(const double* d_src_den, double* d_trg_val, int nrhs, hipStream_t s) {
  if (!setup_done_) return fail(-3, "autotune called before setup");
  if (nrhs < 1) return fail(-2, "nrhs must be >= 1");
  tune_report_.clear();
  if (precision_ != PREC_MIXED) return 0;  // the FP64 path keeps its fixed summation order
  const bool was_timing = timing_;
  timing_ = true;
  auto run = [&](double* ms_out) -> int {
    double best[PH_COUNT];
    for (int p = 0; p < PH_COUNT; p++) best[p] = 1e300;
    for (int rep = 0; rep < 3; rep++) {  // one warm-up, then the best of two
      const int rc = nrhs == 1 ? evaluate(d_src_den, d_trg_val, s) : evaluateMulti(d_src_den, d_trg_val, nrhs, s);
      if (rc) return rc;
      if (rep == 0) continue;
      for (int p = 0; p < PH_COUNT; p++) best[p] = std::min(best[p], double(phase_ms_[size_t(p)]));
    }
    for (int p = 0; p < PH_COUNT; p++) ms_out[p] = best[p];
    return 0;
  };
  auto total = [](const double* ms) {
    double t = 0;
    for (int p = 0; p < PH_COUNT; p++) t += ms[p];
    return t;
  };
  struct Knob { int id; std::vector<int> values; };
  // vector kernels: the Hadamard form first (the V list dominates their
  // evaluate, so the other trials run with the chosen form)
  std::vector<Knob> knobs;
  if (ops_->sd > 1 || ops_->td > 1) knobs.push_back({3, {0, 1, 2}});
  knobs.insert(knobs.end(), {{2, {0, 6}}, {1, {0, 1, 2}}, {0, {0, 1, 2}}});
  int rc = 0;
  for (const Knob& k : knobs) {
    Tuning base = tune_;
    int best_v = -1;
    double best_t = 1e300;
    for (int v : k.values) {
      Tuning t = base;
      if (k.id == 0) t.near_mode = v;
      else if (k.id == 1) t.near_stage = v;
      else if (k.id == 2) t.trans = v;
      else t.vec_hadamard = v;
      if ((rc = apply_tuning(t))) break;
      double ms[PH_COUNT];
      if ((rc = run(ms))) break;
      // the translations touch the up and down passes, the near-field choices
      // the near-field phase, the Hadamard form the V-list phase
      const double tm = k.id == 2 ? total(ms) : k.id == 3 ? ms[PH_VLIST] : ms[PH_NEAR];
      tune_report_.push_back({k.id, v, tm});
      if (tm < best_t) { best_t = tm; best_v = v; }
    }
    if (rc) break;
    Tuning t = base;
    if (k.id == 0) t.near_mode = best_v;
    else if (k.id == 1) t.near_stage = best_v;
    else if (k.id == 2) t.trans = best_v;
    else t.vec_hadamard = best_v;
    if ((rc = apply_tuning(t))) break;
  }
  timing_ = was_timing;
  return rc;
}

int FMM3d::direct(const double* d_src_den, int nchk, const double* d_chk_pos, double* d_chk_val,
                  hipStream_t s) {
  if (!setup_done_) return fail(-3, "direct called before setup");
  if (nchk < 0) return fail(-2, "negative check-target count");
  if (nchk == 0) return 0;
  if ((ns_ > 0 && !d_src_den) || !d_chk_pos || !d_chk_val) return fail(-2, "null density, position or value pointer");
  const int TD = trgDOF();
  // the workspace grows to what the plan asks for (at most 64 chunks' worth
  // beyond one chunk of partial sums per target)
  const int64_t cap = std::max<int64_t>(int64_t(nchk) * TD * 64, int64_t(1) << 20);
  DirectPlan d = plan_direct(ns_, nchk, TD, cap);
  if (d.nchunk <= 0) return fail(-6, "direct sum plan failed");
  int rc;
  if (size_t(d.work_doubles) * sizeof(double) > direct_work_d_.bytes())
    if ((rc = check(direct_work_d_.alloc(size_t(d.work_doubles) * sizeof(double)), "alloc direct-sum workspace")))
      return rc;
  return direct_with(d_src_den, nchk, d_chk_pos, d_chk_val, d, direct_work_d_.as<double>(), s);
}

int FMM3d::directNoAlloc(const double* d_src_den, int nchk, const double* d_chk_pos, double* d_chk_val,
                         hipStream_t s) {
  if (!setup_done_) return fail(-3, "direct called before setup");
  if (nchk < 0) return fail(-2, "negative check-target count");
  if (nchk == 0) return 0;
  if ((ns_ > 0 && !d_src_den) || !d_chk_pos || !d_chk_val) return fail(-2, "null density, position or value pointer");
  const int TD = trgDOF();
  if (int64_t(nchk) * TD > kScratchDoubles)
    return fail(-2, "nchk * trgDOF = " + std::to_string(int64_t(nchk) * TD) + " exceeds the scratch capacity of " +
                        std::to_string(kScratchDoubles) + " values");
  double* work = nullptr;
  int rc = check(scratch_address(&work), "scratch address");
  if (rc) return rc;
  DirectPlan d = plan_direct(ns_, nchk, TD, kScratchDoubles);
  if (d.nchunk <= 0) return fail(-6, "direct sum plan failed");
  return direct_with(d_src_den, nchk, d_chk_pos, d_chk_val, d, work, s);
}

int FMM3d::direct_with(const double* d_src_den, int nchk, const double* d_chk_pos, double* d_chk_val,
                       const DirectPlan& d, double* work, hipStream_t s) {
  const KParams<double> pu = make_params(kt_, coefs_.data());
  return check(launch_direct(kt_, pu, ns_, spos_d_.as<double>(), snor_d_.as<double>(), sperm_d_.as<int>(), d_src_den,
                             nchk, d_chk_pos, d, work, d_chk_val, s),
               "direct sum");
}

}  // namespace kifmm
