// Adjoint (transposed) evaluate pass: sigma_bar = A^T w for the operator A that
// FMM3d::evaluate applies, built as the exact transpose of every forward step
// of fmm3d_eval_mpi.cpp:396-1515 taken in reverse order:
//
//   forward                                   adjoint
//   S2M      UC(B) = K(UC_B, src B) sigma     sigma_bar += K(UC_B, .)^T UCbar      (near^T, ADJ_UC)
//   pinv_up  UE = rs * P_up UC                UCbar += P_up^T (rs * UEbar)
//   M2M      UC(P) = sum_c M_c (ks * UE(c))   UEbar(c) += ks * M_c^T UCbar(P)
//   V list   DC(B) += M2L(B<-A) UE(A)         UEbar(A) += M2L(B<-A)^T DCbar(B)
//   X list   DC(B) += K(DC_B, src A) sigma    sigma_bar += K(DC_B, .)^T DCbar      (near^T, ADJ_DC)
//   L2L      DC(c) += L_c (ks * DE(P))        DEbar(P) += ks * sum_c L_c^T DCbar(c)
//   pinv_dn  DE = rs * P_dn DC                DCbar += P_dn^T (rs * DEbar)
//   near     phi += K sigma + K_mm UE + K_mm DE   sigma_bar, UEbar, DEbar += transposes
//
// The V-list step uses the forward machinery with the roles of the two boxes
// swapped: the transpose of the translation B <- A is a translation A <- B on
// the same sample grid with the kernel K'(r) = K_mm(-r)^T.  For the even,
// symmetric equivalent kernels (111, 511) that is the forward table at the
// negated offset, and the homogeneous level scale is a scalar shared by A and
// B (same level).  For the Stokes equivalent kernel 301 (sd 4, td 3; the
// velocity part even, the pressure-doublet column odd) it is the forward table
// with the DOFs transposed and the odd column negated (3 -> 4 DOFs), and the
// level scale 2^(l deg), whose degrees differ per DOF (1, 1, 1, 2), moves from
// the input to the output of the pass.
#include <algorithm>
#include <cmath>

#include "fmm3d.hpp"
#include "kernel3d.hpp"

namespace kifmm {

int FMM3d::build_adjoint() {
  // the level tables and spectra buffers below are re-allocated: a captured
  // evaluate graph would replay against the old ones
  drop_graph();
  const Tree& t = tree_;
  const Operators& op = *ops_;
  const int N = t.n_nodes, sd = op.sd;
  if (op.kt_mm != 111 && op.kt_mm != 511 && op.kt_mm != 301)
    return fail(-2, "no adjoint pass for kernel " + std::to_string(kt_) + " (equivalent kernel " +
                        std::to_string(op.kt_mm) + ")");

  // --- source-centric near field: per source leaf, the target leaves whose
  // near lists hold it as exact sources, the S2L boxes that read it, and its
  // own S2M check surface ---
  std::vector<std::vector<int>> items(N);
  std::vector<std::vector<int>> ue_from(N);
  std::vector<int> de_leaves;
  for (size_t c = 0; c < near_chunks_h_.size(); c++) {
    const NearChunk& ch = near_chunks_h_[c];
    if (ch.tbeg != t.trg_beg[ch.node]) continue;  // one pass per target leaf
    for (int it = ch.ibeg; it < ch.iend; it++) {
      const int code = near_items_h_[it], A = code >> 2, kind = code & 3;
      if (kind == ITEM_EXACT) items[A].push_back(ch.node << 2 | ADJ_TRG);
      else if (kind == ITEM_UE) ue_from[A].push_back(ch.node);
      else de_leaves.push_back(ch.node);
    }
  }
  for (const SurfTask& tk : s2l_tasks_h_)
    for (int it = tk.ibeg; it < tk.iend; it++) items[surf_items_h_[it]].push_back(tk.node << 2 | ADJ_DC);
  for (const SurfTask& tk : s2m_tasks_h_) items[tk.node].push_back(tk.node << 2 | ADJ_UC);
  nt_chunks_h_.clear(); nt_items_h_.clear();
  for (int A = 0; A < N; A++) {
    if (!t.leaf(A) || t.src_beg[A] < 0 || t.src_num[A] <= 0) continue;
    const int ibeg = int(nt_items_h_.size());
    nt_items_h_.insert(nt_items_h_.end(), items[A].begin(), items[A].end());
    const int iend = int(nt_items_h_.size()), ns = t.src_num[A];
    for (int c = 0; c < ns; c += kNearChunk)
      nt_chunks_h_.push_back(NearChunk{A, t.src_beg[A] + c, std::min(kNearChunk, ns - c), ibeg, iend});
  }
  // --- equivalent surfaces <- targets (W-list M2T and L2T transposed) ---
  ue_t_tasks_h_.clear(); de_t_tasks_h_.clear(); surf_t_items_h_.clear();
  for (int A = 0; A < N; A++) {
    if (ue_from[A].empty()) continue;
    const int ib = int(surf_t_items_h_.size());
    surf_t_items_h_.insert(surf_t_items_h_.end(), ue_from[A].begin(), ue_from[A].end());
    ue_t_tasks_h_.push_back(SurfTask{A, ib, int(surf_t_items_h_.size())});
  }
  for (int B : de_leaves) {
    const int ib = int(surf_t_items_h_.size());
    surf_t_items_h_.push_back(B);
    de_t_tasks_h_.push_back(SurfTask{B, ib, ib + 1});
  }

  // --- transposed level steps (scales as in the forward steps) ---
  up_t_steps_.clear(); dn_t_steps_.clear();
  auto scale_vec = [&](int lvl, int sign) {
    int off = int(scales_h_.size());
    for (int s = 0; s < sd; s++) scales_h_.push_back(std::ldexp(1.0, sign * int(lvl * op.deg[s])));
    return off;
  };
  auto add_step = [&](std::vector<GemmStep>& v, int mat, int src, int dst, const std::vector<int>& b,
                      const std::vector<int>& c, int ks, int rs) {
    if (b.empty()) return;
    GemmStep st{mat, src, dst, int(b.size()), 1, int64_t(lvl_idx_h_.size()), ks, rs};
    lvl_idx_h_.insert(lvl_idx_h_.end(), b.begin(), b.end());
    lvl_idx_h_.insert(lvl_idx_h_.end(), c.begin(), c.end());
    v.push_back(st);
  };
  const int rl = t.root_level;
  std::vector<std::vector<int>> by_depth(t.max_depth + 1);
  for (int i = 0; i < N; i++) by_depth[t.depth[i]].push_back(i);
  // down pass, deepest level first: pinv_dn^T at depth d, then L2L^T into the
  // parents at depth d-1 (children at depth >= 3 only, :1371-1376)
  for (int d = t.max_depth; d >= 2; d--) {
    std::vector<int> pb, par, ch0;
    for (int i : by_depth[d])
      if (t.tag[i] & TAG_EVTR) pb.push_back(i);
    add_step(dn_t_steps_, 6, 2, 3, pb, pb, scale_vec(d + rl, -1), -1);
    if (d >= 3)
      for (int p : by_depth[d - 1]) {
        if (t.leaf(p)) continue;
        bool any = false;
        for (int c = 0; c < 8; c++) any = any || (t.tag[t.child[p] + c] & TAG_EVTR);
        if (any) { par.push_back(p); ch0.push_back(t.child[p]); }
      }
    if (d >= 3) add_step(dn_t_steps_, 7, 3, 2, ch0, par, -1, scale_vec(d - 1 + rl, +1));
  }
  // up pass, root side first: pinv_up^T at depth d, then M2M^T into depth d+1
  for (int d = 1; d <= t.max_depth; d++) {
    std::vector<int> pb, par, ch0;
    for (int i : by_depth[d]) {
      if (!(t.tag[i] & TAG_CBTR)) continue;
      pb.push_back(i);
      if (!t.leaf(i)) { par.push_back(i); ch0.push_back(t.child[i]); }
    }
    add_step(up_t_steps_, 5, 0, 1, pb, pb, scale_vec(d + rl, -1), -1);
    add_step(up_t_steps_, 4, 1, 0, par, ch0, -1, scale_vec(d + 1 + rl, +1));
  }

  // --- V list with the roles swapped: the inverted lists, level-major ---
  std::vector<int64_t> off(N + 1, 0);
  for (int B = 0; B < N; B++)
    if (t.tag[B] & TAG_EVTR)
      for (int64_t q = t.off[LIST_V][B]; q < t.off[LIST_V][B + 1]; q++) off[t.lst[LIST_V][q] + 1]++;
  for (int i = 0; i < N; i++) off[i + 1] += off[i];
  std::vector<int> lst(off[N]);
  std::vector<int64_t> fill(off.begin(), off.end() - 1);
  for (int B = 0; B < N; B++)
    if (t.tag[B] & TAG_EVTR)
      for (int64_t q = t.off[LIST_V][B]; q < t.off[LIST_V][B + 1]; q++) lst[fill[t.lst[LIST_V][q]]++] = B;
  std::vector<int> order;
  order.reserve(N);
  for (const auto& lv : by_depth) order.insert(order.end(), lv.begin(), lv.end());
  int rc = build_vgroups(order, off, lst, false, vadj_);
  if (rc) return rc;

  // --- device tables ---
#define KIFMM_UP(buf, vec, what) \
  if ((rc = check((buf).upload((vec).data(), sizeof((vec)[0]) * (vec).size()), what))) return rc;
  KIFMM_UP(nt_chunks_d_, nt_chunks_h_, "upload adjoint near chunks");
  KIFMM_UP(nt_items_d_, nt_items_h_, "upload adjoint near items");
  KIFMM_UP(ue_t_tasks_d_, ue_t_tasks_h_, "upload adjoint surface tasks");
  KIFMM_UP(de_t_tasks_d_, de_t_tasks_h_, "upload adjoint surface tasks");
  KIFMM_UP(surf_t_items_d_, surf_t_items_h_, "upload adjoint surface items");
  KIFMM_UP(tbeg_d_, t.trg_beg, "upload target ranges");
  KIFMM_UP(tnum_d_, t.trg_num, "upload target ranges");
  KIFMM_UP(lvl_idx_d_, lvl_idx_h_, "upload level columns");
  KIFMM_UP(scales_d_, scales_h_, "upload level scales");
  {
    const int ue = op.dat(UE), uc = op.dat(UC), de = op.dat(DE), dc = op.dat(DC);
    std::vector<double> m(size_t(8) * ue * uc);  // (8 UE) x UC: transpose of the UC x (8 UE) M2M
    for (int k = 0; k < 8 * ue; k++)
      for (int i = 0; i < uc; i++) m[size_t(i) * 8 * ue + k] = op.m2m[size_t(k) * uc + i];
    KIFMM_UP(m2mT_d_, m, "upload M2M^T");
    std::vector<double> pu(size_t(uc) * ue);  // UC x UE
    for (int i = 0; i < uc; i++)
      for (int j = 0; j < ue; j++) pu[size_t(j) * uc + i] = op.pinv_up[size_t(i) * ue + j];
    KIFMM_UP(pupT_d_, pu, "upload pinv_up^T");
    std::vector<double> pd(size_t(dc) * de);  // DC x DE
    for (int i = 0; i < dc; i++)
      for (int j = 0; j < de; j++) pd[size_t(j) * dc + i] = op.pinv_dn[size_t(i) * de + j];
    KIFMM_UP(pdnT_d_, pd, "upload pinv_dn^T");
    std::vector<double> l(size_t(de) * 8 * dc);  // DE x (8 DC): L_c^T side by side
    for (int c = 0; c < 8; c++)
      for (int i = 0; i < dc; i++)
        for (int j = 0; j < de; j++)
          l[(size_t(c) * dc + i) * de + j] = op.l2l[size_t(c) * dc * de + size_t(j) * dc + i];
    KIFMM_UP(l2lT_d_, l, "upload L2L^T");
  }
  if ((rc = upload_vgroups(vadj_))) return rc;
  if (op.kt_mm == 301) {
    // transposed translation spectra: [slot][i * sd + j] = s_j T[slot][j * td + i]
    // (s = +1, +1, +1, -1: the odd pressure-doublet column), FP64 and FP32
    const int sd4 = op.sd, td3 = op.td, H = op.H;
    std::vector<double> t(op.m2l.size());
    for (int slot = 0; slot < 343; slot++)
      for (int j = 0; j < sd4; j++)
        for (int i = 0; i < td3; i++) {
          const double sg = j == 3 ? -1.0 : 1.0;
          const double* src = op.m2l.data() + ((size_t(slot) * sd4 * td3 + size_t(j) * td3 + i) * H) * 2;
          double* dst = t.data() + ((size_t(slot) * sd4 * td3 + size_t(i) * sd4 + j) * H) * 2;
          for (int f = 0; f < 2 * H; f++) dst[f] = sg * src[f];
        }
    KIFMM_UP(m2lT_d_, t, "upload transposed M2L");
    if (precision_ == PREC_MIXED) {
      const std::vector<float> t32 = pack_spectra32(t, op.np);
      KIFMM_UP(m2lT32_d_, t32, "upload transposed M2L (FP32)");
      m2lT_sym3_ = table_sym3(t, td3, sd4, H);  // transposed: td DOFs in, sd out
    }
    std::vector<double> z(size_t(op.td), 0.0);
    KIFMM_UP(deg0_d_, z, "upload degrees");
  }
#undef KIFMM_UP
  // spectra buffers hold the larger of the two passes (a captured forward pass
  // holds the old addresses)
  drop_graph();
  {
    const size_t dmax = size_t(std::max(op.sd, op.td));
    const size_t nsrc = std::max(vfwd_.src_node.size(), vadj_.src_node.size()) * dmax;
    const size_t ntrg = std::max(vfwd_.trg_node.size(), vadj_.trg_node.size()) * dmax;
    if (precision_ == PREC_FP64) {
      if (dspec_d_.bytes() < sizeof(double2) * nsrc * op.H &&
          (rc = check(dspec_d_.alloc(sizeof(double2) * nsrc * op.H), "alloc V spectra")))
        return rc;
      if (espec_d_.bytes() < sizeof(double2) * ntrg * op.H &&
          (rc = check(espec_d_.alloc(sizeof(double2) * ntrg * op.H), "alloc V spectra")))
        return rc;
    } else {
      const size_t hc = size_t(spec_packed_len(op.np));
      if (dspec32_d_.bytes() < sizeof(float2) * nsrc * hc &&
          (rc = check(dspec32_d_.alloc(sizeof(float2) * nsrc * hc), "alloc V spectra")))
        return rc;
      if (espec32_d_.bytes() < sizeof(float2) * ntrg * hc &&
          (rc = check(espec32_d_.alloc(sizeof(float2) * ntrg * hc), "alloc V spectra")))
        return rc;
    }
  }
  if (precision_ == PREC_MIXED) {
    // split-K workspace for the transposed steps (the same rule as the pass)
    size_t need = 1;
    for (const auto* steps : {&dn_t_steps_, &up_t_steps_})
      for (const GemmStep& st : *steps) {
        GemmArgs g{};
        switch (st.mat) {
          case 4: g.M = 8 * op.dat(UE); g.K = op.dat(UC); break;
          case 5: g.M = op.dat(UC); g.K = op.dat(UE); break;
          case 6: g.M = op.dat(DC); g.K = op.dat(DE); break;
          default: g.M = op.dat(DE); g.K = 8 * op.dat(DC);
        }
        g.ncol = st.ncol; g.max_zcol = st.ncol;
        split_k(g);
        if (g.ksplit > 1) need = std::max(need, size_t(g.ksplit) * size_t(g.ncol) * size_t(g.M));
      }
    if ((rc = reserve_splitk(need))) return rc;
  }
  adj_ready_ = true;
  return 0;
}

int FMM3d::evaluateAdjoint(const double* d_trg_in, double* d_src_out, hipStream_t s) {
  if (!setup_done_) return fail(-3, "evaluateAdjoint called before setup");
  if ((nt_ > 0 && !d_trg_in) || (ns_ > 0 && !d_src_out)) return fail(-2, "null input or output pointer");
  int rc;
  if (!adj_ready_) {
    // the adjoint tables are built and uploaded before any adjoint kernel runs
    if ((rc = check(hipStreamSynchronize(s), "sync"))) return rc;
    if ((rc = build_adjoint())) return rc;
  }
  const Tree& t = tree_;
  const Operators& op = *ops_;
  const int N = t.n_nodes, SD = srcDOF(), TD = trgDOF();
  const KParams<double> pu = make_params(kt_, coefs_.data());
  const KParams<double> pm = make_params(op.kt_mm, op.coefs.data());
  double* bufs[4] = {ue_d_.as<double>(), uc_d_.as<double>(), de_d_.as<double>(), dc_d_.as<double>()};
  const int szs[4] = {op.dat(UE), op.dat(UC), op.dat(DE), op.dat(DC)};

  for (int b = 0; b < 4; b++)
    if ((rc = check(launch_fill(bufs[b], int64_t(N) * szs[b], 0.0, s), "zero work vectors"))) return rc;
  if ((rc = check(launch_gather(d_trg_in, tperm_d_.as<int>(), tval_d_.as<double>(), nt_, TD, s), "gather"))) return rc;

  // L2T^T and M2T^T: targets -> DE / UE surfaces
  SurfTArgs st{};
  st.items = surf_t_items_d_.as<int>();
  st.tpos = tpos_d_.as<double>(); st.tw = tval_d_.as<double>();
  st.trg_beg = tbeg_d_.as<int>(); st.trg_num = tnum_d_.as<int>();
  st.geom = geom_d_.as<double4>();
  st.tasks = de_t_tasks_d_.as<SurfTask>(); st.ntasks = int(de_t_tasks_h_.size());
  st.sam = sam_d_[DE].as<double>(); st.n_sam = op.n_sam[DE];
  st.out = de_d_.as<double>(); st.sz_out = op.dat(DE);
  if ((rc = check(launch_surface_adjoint(op.kt_mm, st, pm, s), "L2T adjoint"))) return rc;
  st.tasks = ue_t_tasks_d_.as<SurfTask>(); st.ntasks = int(ue_t_tasks_h_.size());
  st.sam = sam_d_[UE].as<double>(); st.n_sam = op.n_sam[UE];
  st.out = ue_d_.as<double>(); st.sz_out = op.dat(UE);
  if ((rc = check(launch_surface_adjoint(op.kt_mm, st, pm, s), "M2T adjoint"))) return rc;

  auto run_steps = [&](const std::vector<GemmStep>& steps) -> int {
    for (const GemmStep& stp : steps) {
      GemmArgs g{};
      switch (stp.mat) {
        case 4: g.M = 8 * op.dat(UE); g.K = op.dat(UC); g.A = m2mT_d_.as<double>(); break;
        case 5: g.M = op.dat(UC); g.K = op.dat(UE); g.A = pupT_d_.as<double>(); break;
        case 6: g.M = op.dat(DC); g.K = op.dat(DE); g.A = pdnT_d_.as<double>(); break;
        default:
          g.M = op.dat(DE); g.K = 8 * op.dat(DC); g.A = l2lT_d_.as<double>();
      }
      g.max_zcol = stp.ncol;
      g.B = bufs[stp.src]; g.ldb = szs[stp.src];
      g.C = bufs[stp.dst]; g.ldc = szs[stp.dst];
      g.bidx = lvl_idx_d_.as<int>() + stp.idx_off;
      g.cidx = lvl_idx_d_.as<int>() + stp.idx_off + stp.ncol;
      if (stp.ks_off >= 0) { g.kscale = scales_d_.as<double>() + stp.ks_off; g.ksd = op.sd; }
      else { g.kscale = ones_d_.as<double>(); g.ksd = 1; }
      if (stp.rs_off >= 0) { g.rscale = scales_d_.as<double>() + stp.rs_off; g.rsd = op.sd; }
      else { g.rscale = ones_d_.as<double>(); g.rsd = 1; }
      g.ncol = stp.ncol; g.accumulate = 1;
      // mixed mode: few-tile launches split K, the chunks' partial products are
      // summed in chunk order (deterministic, as the FP64 mode's one workgroup
      // per output tile)
      if (precision_ == PREC_MIXED) {
        split_k(g);
        g.work = splitk_for(g);
        if (g.ksplit > 1 && !g.work) return fail(-6, "split-K workspace smaller than the launch needs");
      }
      int r = check(launch_gemm_fma(g, s), "adjoint level GEMM");
      if (r) return r;
    }
    return 0;
  };
  if ((rc = run_steps(dn_t_steps_))) return rc;

  // V list transposed: DCbar of the forward targets -> UEbar of the forward sources
  VlistArgs v{};
  vlist_args(v, vadj_);
  v.UE = dc_d_.as<double>(); v.sz_ue = op.dat(DC);
  v.DC = ue_d_.as<double>(); v.sz_dc = op.dat(UE);
  const float2* m2l32 = nullptr;
  if (op.kt_mm == 301) {  // 3 DOFs in, 4 out; the level scale on the output
    v.sd = op.td;
    v.td = op.sd;
    v.m2l = m2lT_d_.as<double2>();
    m2l32 = m2lT32_d_.as<float2>();
    v.sym3 = m2lT_sym3_ ? 1 : 0;
    v.mirror = 0;  // the transposed table keeps the wrap entries
    v.deg = deg0_d_.as<double>();
    v.deg_out = deg_d_.as<double>();
  }
  if ((rc = run_vlist(v, s, false, m2l32))) return rc;

  if ((rc = run_steps(up_t_steps_))) return rc;

  // U, W, X point pairs, S2L^T and S2M^T at the sources
  NearTArgs na{};
  na.chunks = nt_chunks_d_.as<NearChunk>(); na.nchunks = int(nt_chunks_h_.size());
  na.items = nt_items_d_.as<int>();
  na.spos = spos_d_.as<double>(); na.snor = snor_d_.as<double>();
  na.tpos = tpos_d_.as<double>(); na.tw = tval_d_.as<double>();
  na.trg_beg = tbeg_d_.as<int>(); na.trg_num = tnum_d_.as<int>();
  na.geom = geom_d_.as<double4>();
  na.sam_dc = sam_d_[DC].as<double>(); na.n_dc = op.n_sam[DC]; na.DC = dc_d_.as<double>(); na.sz_dc = op.dat(DC);
  na.sam_uc = sam_d_[UC].as<double>(); na.n_uc = op.n_sam[UC]; na.UC = uc_d_.as<double>(); na.sz_uc = op.dat(UC);
  na.sout = sden_d_.as<double>();
  if ((rc = check(launch_fill(sden_d_.as<double>(), std::max<int64_t>(1, ns_ * SD), 0.0, s), "zero output"))) return rc;
  if (precision_ == PREC_MIXED && kt_ == 111) {
    // Laplace is symmetric, so the transposed pairs are the forward mixed near
    // kernel's with the roles swapped: the chunk's points are the sources, the
    // exact items the target leaves (target points weighted by the adjoint
    // input), and the two surface kinds the S2L check surfaces (DCbar) and the
    // leaf's own S2M check surface (UCbar).  The item codes ADJ_TRG, ADJ_DC and
    // ADJ_UC are ITEM_EXACT, ITEM_UE and ITEM_DE.
    static_assert(int(ADJ_TRG) == int(ITEM_EXACT) && int(ADJ_DC) == int(ITEM_UE) && int(ADJ_UC) == int(ITEM_DE),
                  "adjoint item kinds must match the near-kernel item kinds");
    NearArgs a{};
    a.chunks = na.chunks; a.nchunks = na.nchunks; a.items = na.items;
    a.spos = na.tpos; a.snor = nullptr; a.sden = na.tw;
    a.src_beg = na.trg_beg; a.src_num = na.trg_num;
    a.geom = na.geom;
    a.sam_ue = na.sam_dc; a.n_ue = na.n_dc; a.UE = na.DC; a.sz_ue = na.sz_dc;
    a.sam_de = na.sam_uc; a.n_de = na.n_uc; a.DE = na.UC; a.sz_de = na.sz_uc;
    a.tpos = na.spos; a.tval = na.sout; a.tperm = nullptr;
    a.nrhs = 1; a.r0 = 0; a.nr = 1; a.nstride = N; a.tstride = 0;
    if ((rc = check(launch_near_mixed(kt_, a, pu, pu, s), "near field adjoint"))) return rc;
  } else if ((rc = check(launch_near_adjoint(kt_, na, pu, s), "near field adjoint"))) {
    return rc;
  }
  return check(launch_scatter(sden_d_.as<double>(), sperm_d_.as<int>(), d_src_out, ns_, SD, s), "scatter");
}

}  // namespace kifmm
