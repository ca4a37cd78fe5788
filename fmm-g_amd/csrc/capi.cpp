// extern "C" boundary (include/kifmm.h) over kifmm::FMM3d.
#include "../../include/kifmm.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fmm3d.hpp"
#include "kernel3d.hpp"
#include "options.hpp"

struct kifmm_fmm {
  kifmm::FMM3d fmm;
  // staging of the host-pointer entry points: kept and grown, so repeated calls
  // allocate nothing
  kifmm::DeviceBuffer in, out, aux;
};

namespace {
thread_local std::string g_create_error;

int set_err(kifmm_handle h, int code, const std::string& msg) {
  if (h) h->fmm.setError(msg);
  else g_create_error = msg;
  return code;
}

// Device staging for host arrays: grows b to at least `bytes` (contents not kept).
hipError_t grow(kifmm::DeviceBuffer& b, size_t bytes) {
  return b.bytes() >= bytes ? hipSuccess : b.alloc(bytes);
}

// Copies `bytes` of host data into the (grown) staging buffer b.
hipError_t stage_in(kifmm::DeviceBuffer& b, const void* host, size_t bytes) {
  hipError_t e = grow(b, bytes);
  if (e != hipSuccess || bytes == 0) return e;
  return hipMemcpy(b.as<void>(), host, bytes, hipMemcpyHostToDevice);
}

// SplitMix64: sample indices for kifmm_check.
uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
}  // namespace

extern "C" {

int kifmm_create(int kernel_type, const double* coefs, int ncoefs, int np, int pts_max, int max_level,
                 const double ctr[3], int root_level, int precision, kifmm_handle* out) {
  if (!out) return set_err(nullptr, KIFMM_ERR_ARG, "null output handle");
  *out = nullptr;
  if (!kifmm::kernel_known(kernel_type))
    return set_err(nullptr, KIFMM_ERR_ARG, "unknown kernel type " + std::to_string(kernel_type));
  if (kifmm::mm_kernel(kernel_type) < 0)
    return set_err(nullptr, KIFMM_ERR_ARG, "kernel " + std::to_string(kernel_type) + " has no FMM equivalent kernel");
  if (np < 3 || np > 8) return set_err(nullptr, KIFMM_ERR_ARG, "np must lie in [3,8]");
  if (pts_max < 1) return set_err(nullptr, KIFMM_ERR_ARG, "ptsMax must be >= 1");
  if (max_level < 1 || max_level > 30) return set_err(nullptr, KIFMM_ERR_ARG, "maxLevel must lie in [1,30]");
  if (precision != KIFMM_PREC_FP64 && precision != KIFMM_PREC_MIXED)
    return set_err(nullptr, KIFMM_ERR_ARG, "unknown precision mode");
  if (ncoefs < 0 || (ncoefs > 0 && !coefs)) return set_err(nullptr, KIFMM_ERR_ARG, "bad coefficient array");
  auto h = std::make_unique<kifmm_fmm>();
  auto& f = h->fmm;
  f.kernelType() = kernel_type;
  if (ncoefs > 0) f.coefs().assign(coefs, coefs + ncoefs);
  f.np() = np;
  f.ptsMax() = pts_max;
  f.maxLevel() = max_level;
  if (ctr) for (int d = 0; d < 3; d++) f.ctr()[d] = ctr[d];
  f.rootLevel() = root_level;
  f.precision() = precision;
  *out = h.release();
  return KIFMM_OK;
}

int kifmm_create_from_options(const char* path, const char* prefix, kifmm_handle* out) {
  if (!out || !path) return set_err(nullptr, KIFMM_ERR_ARG, "null argument");
  *out = nullptr;
  kifmm::Options o;
  std::string err;
  if (o.load(path, &err)) return set_err(nullptr, KIFMM_ERR_OPTIONS, err);
  const std::string pf = prefix ? prefix : "fmm3d_";
  long kt, np, pm, ml;
  if (!o.get_int("-kt", &kt)) return set_err(nullptr, KIFMM_ERR_OPTIONS, "option -kt missing");
  if (!o.get_int("-" + pf + "np", &np)) return set_err(nullptr, KIFMM_ERR_OPTIONS, "option -" + pf + "np missing");
  if (!o.get_int("-" + pf + "let3d_ptsMax", &pm))
    return set_err(nullptr, KIFMM_ERR_OPTIONS, "option -" + pf + "let3d_ptsMax missing");
  if (!o.get_int("-" + pf + "let3d_maxLevel", &ml))
    return set_err(nullptr, KIFMM_ERR_OPTIONS, "option -" + pf + "let3d_maxLevel missing");
  double coefs[2] = {1.0, 0.25};
  o.get_double("-kifmm_mu", &coefs[0]);
  o.get_double("-kifmm_nu", &coefs[1]);
  int prec = KIFMM_PREC_FP64;
  std::string ps;
  if (o.get("-kifmm_precision", &ps)) {
    if (ps == "mixed") prec = KIFMM_PREC_MIXED;
    else if (ps != "fp64") return set_err(nullptr, KIFMM_ERR_OPTIONS, "-kifmm_precision must be fp64 or mixed");
  }
  const double ctr[3] = {0.5, 0.5, 0.5};
  return kifmm_create(int(kt), coefs, 2, int(np), int(pm), int(ml), ctr, 1, prec, out);
}

int kifmm_destroy(kifmm_handle h) {
  delete h;
  return KIFMM_OK;
}

int kifmm_set_points(kifmm_handle h, int64_t n_src, const double* src_pos, const double* src_nor, int64_t n_trg,
                     const double* trg_pos) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  if (kifmm::needs_normals(h->fmm.kernelType()) && n_src > 0 && !src_nor)
    return set_err(h, KIFMM_ERR_ARG, "kernel needs source normals");
  return h->fmm.setPoints(n_src, src_pos, src_nor, n_trg, trg_pos);
}

int kifmm_set_points_batch(kifmm_handle h, int nprob, const int64_t* n_src, const double* src_pos,
                           const double* src_nor, const int64_t* n_trg, const double* trg_pos) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  if (nprob < 1 || !n_src) return set_err(h, KIFMM_ERR_ARG, "a batch needs at least one problem and its point counts");
  int64_t S = 0;
  for (int p = 0; p < nprob; p++) S += n_src[p] > 0 ? n_src[p] : 0;
  if (kifmm::needs_normals(h->fmm.kernelType()) && S > 0 && !src_nor)
    return set_err(h, KIFMM_ERR_ARG, "kernel needs source normals");
  if (!trg_pos) return h->fmm.setPointsBatch(nprob, n_src, src_pos, src_nor, n_src, src_pos);
  if (!n_trg) return set_err(h, KIFMM_ERR_ARG, "target counts missing");
  return h->fmm.setPointsBatch(nprob, n_src, src_pos, src_nor, n_trg, trg_pos);
}

int kifmm_setup(kifmm_handle h) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.setup();
}

int kifmm_setup_host(kifmm_handle h) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.setupHost();
}

int kifmm_src_dof(kifmm_handle h) { return h ? h->fmm.srcDOF() : KIFMM_ERR_ARG; }
int kifmm_trg_dof(kifmm_handle h) { return h ? h->fmm.trgDOF() : KIFMM_ERR_ARG; }

int kifmm_evaluate_device(kifmm_handle h, const double* d_src_den, double* d_trg_val, void* stream) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.evaluate(d_src_den, d_trg_val, static_cast<hipStream_t>(stream));
}

int kifmm_evaluate(kifmm_handle h, const double* src_den, double* trg_val) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  auto& f = h->fmm;
  if (!f.isSetup()) return set_err(h, KIFMM_ERR_STATE, "evaluate called before setup");
  if ((f.numSrc() > 0 && !src_den) || (f.numTrg() > 0 && !trg_val)) return set_err(h, KIFMM_ERR_ARG, "null array");
  const size_t nd = size_t(f.numSrc()) * f.srcDOF(), nv = size_t(f.numTrg()) * f.trgDOF();
  kifmm::DeviceBuffer &d = h->in, &v = h->out;
  hipError_t e;
  if ((e = stage_in(d, src_den, sizeof(double) * nd)) != hipSuccess) return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  if ((e = grow(v, sizeof(double) * nv)) != hipSuccess) return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  int rc = f.evaluate(d.as<double>(), v.as<double>(), nullptr);
  if (rc) return rc;
  if ((e = hipMemcpy(trg_val, v.as<double>(), sizeof(double) * nv, hipMemcpyDeviceToHost)) != hipSuccess)
    return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  return KIFMM_OK;
}

int kifmm_evaluate_multi_device(kifmm_handle h, const double* d_src_den, double* d_trg_val, int nrhs,
                                void* stream) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  if (nrhs < 0) return set_err(h, KIFMM_ERR_ARG, "negative nrhs");
  return h->fmm.evaluateMulti(d_src_den, d_trg_val, nrhs, static_cast<hipStream_t>(stream));
}

int kifmm_evaluate_multi(kifmm_handle h, const double* src_den, double* trg_val, int nrhs) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  auto& f = h->fmm;
  if (!f.isSetup()) return set_err(h, KIFMM_ERR_STATE, "evaluate called before setup");
  if (nrhs < 0) return set_err(h, KIFMM_ERR_ARG, "negative nrhs");
  if (nrhs == 0) return KIFMM_OK;
  if ((f.numSrc() > 0 && !src_den) || (f.numTrg() > 0 && !trg_val)) return set_err(h, KIFMM_ERR_ARG, "null array");
  const size_t nd = size_t(f.numSrc()) * f.srcDOF() * nrhs, nv = size_t(f.numTrg()) * f.trgDOF() * nrhs;
  kifmm::DeviceBuffer &d = h->in, &v = h->out;
  hipError_t e;
  if ((e = stage_in(d, src_den, sizeof(double) * nd)) != hipSuccess) return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  if ((e = grow(v, sizeof(double) * nv)) != hipSuccess) return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  int rc = kifmm_evaluate_multi_device(h, d.as<double>(), v.as<double>(), nrhs, nullptr);
  if (rc) return rc;
  if ((e = hipMemcpy(trg_val, v.as<double>(), sizeof(double) * nv, hipMemcpyDeviceToHost)) != hipSuccess)
    return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  return KIFMM_OK;
}

int kifmm_evaluate_adjoint_device(kifmm_handle h, const double* d_trg_in, double* d_src_out, void* stream) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.evaluateAdjoint(d_trg_in, d_src_out, static_cast<hipStream_t>(stream));
}

int kifmm_evaluate_adjoint(kifmm_handle h, const double* trg_in, double* src_out) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  auto& f = h->fmm;
  if (!f.isSetup()) return set_err(h, KIFMM_ERR_STATE, "evaluate_adjoint called before setup");
  if ((f.numTrg() > 0 && !trg_in) || (f.numSrc() > 0 && !src_out)) return set_err(h, KIFMM_ERR_ARG, "null array");
  const size_t nw = size_t(f.numTrg()) * f.trgDOF(), ns = size_t(f.numSrc()) * f.srcDOF();
  kifmm::DeviceBuffer &w = h->in, &o = h->out;
  hipError_t e;
  if ((e = stage_in(w, trg_in, sizeof(double) * nw)) != hipSuccess) return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  if ((e = grow(o, sizeof(double) * ns)) != hipSuccess) return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  int rc = f.evaluateAdjoint(w.as<double>(), o.as<double>(), nullptr);
  if (rc) return rc;
  if ((e = hipMemcpy(src_out, o.as<double>(), sizeof(double) * ns, hipMemcpyDeviceToHost)) != hipSuccess)
    return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  return KIFMM_OK;
}

int kifmm_field_device(kifmm_handle h, const double* d_src_den, double* d_trg_grad, void* stream) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.field(d_src_den, d_trg_grad, static_cast<hipStream_t>(stream));
}

int kifmm_field(kifmm_handle h, const double* src_den, double* trg_grad) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  auto& f = h->fmm;
  if (!f.isSetup()) return set_err(h, KIFMM_ERR_STATE, "field called before setup");
  if ((f.numSrc() > 0 && !src_den) || (f.numTrg() > 0 && !trg_grad)) return set_err(h, KIFMM_ERR_ARG, "null array");
  const size_t nd = size_t(f.numSrc()) * f.srcDOF(), ng = size_t(f.numTrg()) * 3;
  kifmm::DeviceBuffer &d = h->in, &g = h->out;
  hipError_t e;
  if ((e = stage_in(d, src_den, sizeof(double) * nd)) != hipSuccess) return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  if ((e = grow(g, sizeof(double) * ng)) != hipSuccess) return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  int rc = f.field(d.as<double>(), g.as<double>(), nullptr);
  if (rc) return rc;
  if ((e = hipMemcpy(trg_grad, g.as<double>(), sizeof(double) * ng, hipMemcpyDeviceToHost)) != hipSuccess)
    return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  return KIFMM_OK;
}

int kifmm_checkpoint_size(kifmm_handle h, int64_t* ndoubles) {
  if (!h || !ndoubles) return set_err(h, KIFMM_ERR_ARG, "null argument");
  if (!h->fmm.isHostReady()) return set_err(h, KIFMM_ERR_STATE, "checkpoint size requested before setup");
  *ndoubles = h->fmm.checkpointSize();
  return KIFMM_OK;
}

int kifmm_evaluate_checkpoint_device(kifmm_handle h, const double* d_src_den, double* d_trg_val, double* d_ue_out,
                                     void* stream) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.evaluateCheckpoint(d_src_den, d_trg_val, d_ue_out, static_cast<hipStream_t>(stream));
}

int kifmm_field_checkpoint_device(kifmm_handle h, const double* d_src_den, const double* d_ue_in, double* d_trg_grad,
                                  void* stream) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.fieldFromCheckpoint(d_src_den, d_ue_in, d_trg_grad, static_cast<hipStream_t>(stream));
}

int kifmm_direct_device(kifmm_handle h, const double* d_src_den, int nchk, const double* d_chk_pos, double* d_chk_val,
                        void* stream) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.direct(d_src_den, nchk, d_chk_pos, d_chk_val, static_cast<hipStream_t>(stream));
}

int kifmm_direct_device_noalloc(kifmm_handle h, const double* d_src_den, int nchk, const double* d_chk_pos,
                                double* d_chk_val, void* stream) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.directNoAlloc(d_src_den, nchk, d_chk_pos, d_chk_val, static_cast<hipStream_t>(stream));
}

int kifmm_check(kifmm_handle h, const double* src_den, const double* trg_val, int numchk, uint64_t seed,
                double* rerr) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  auto& f = h->fmm;
  if (!f.isSetup()) return set_err(h, KIFMM_ERR_STATE, "check called before setup");
  if (!src_den || !trg_val || !rerr || numchk < 1) return set_err(h, KIFMM_ERR_ARG, "bad check arguments");
  const int64_t nt = f.numTrg();
  const int TD = f.trgDOF();
  std::vector<int64_t> idx;
  if (numchk >= nt) {  // fmm3d_check_mpi.cpp:46-52: all targets
    for (int64_t k = 0; k < nt; k++) idx.push_back(k);
  } else {             // :53-59: random sample
    uint64_t s = seed;
    for (int k = 0; k < numchk; k++) idx.push_back(int64_t(splitmix(s) % uint64_t(nt)));
  }
  const auto& tp = f.targetPositions();
  std::vector<double> chk_pos(3 * idx.size()), chk_val(TD * idx.size());
  for (size_t k = 0; k < idx.size(); k++)
    for (int d = 0; d < 3; d++) chk_pos[3 * k + d] = tp[3 * idx[k] + d];
  kifmm::DeviceBuffer &dd = h->in, &dp = h->aux, &dv = h->out;
  hipError_t e;
  const size_t nd = size_t(f.numSrc()) * f.srcDOF();
  if ((e = stage_in(dd, src_den, sizeof(double) * nd)) != hipSuccess) return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  if ((e = stage_in(dp, chk_pos.data(), sizeof(double) * chk_pos.size())) != hipSuccess)
    return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  if ((e = grow(dv, sizeof(double) * chk_val.size())) != hipSuccess) return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  int rc = f.direct(dd.as<double>(), int(idx.size()), dp.as<double>(), dv.as<double>(), nullptr);
  if (rc) return rc;
  if ((e = hipMemcpy(chk_val.data(), dv.as<double>(), sizeof(double) * chk_val.size(), hipMemcpyDeviceToHost)) !=
      hipSuccess)
    return set_err(h, KIFMM_ERR_HIP, hipGetErrorString(e));
  double nerr = 0, npot = 0;  // :159-165
  for (size_t k = 0; k < idx.size(); k++)
    for (int d = 0; d < TD; d++) {
      double ref = chk_val[TD * k + d], dif = trg_val[TD * idx[k] + d] - ref;
      nerr += dif * dif;
      npot += ref * ref;
    }
  *rerr = std::sqrt(nerr) / std::sqrt(npot);
  return KIFMM_OK;
}

int kifmm_tree_info(kifmm_handle h, int64_t info[8]) {
  if (!h || !info) return set_err(h, KIFMM_ERR_ARG, "null argument");
  if (!h->fmm.isHostReady()) return set_err(h, KIFMM_ERR_STATE, "tree requested before setup");
  const auto& t = h->fmm.tree();
  info[0] = t.n_nodes; info[1] = t.max_depth;
  for (int k = 0; k < 4; k++) info[2 + k] = int64_t(t.lst[k].size());
  info[6] = h->fmm.numSrc(); info[7] = h->fmm.numTrg();
  return KIFMM_OK;
}

int kifmm_get_tree(kifmm_handle h, int* parent, int* child, int* depth, int* path, int* tag, int* src_beg,
                   int* src_num, int* trg_beg, int* trg_num, int* src_perm, int* trg_perm) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  if (!h->fmm.isHostReady()) return set_err(h, KIFMM_ERR_STATE, "tree requested before setup");
  const auto& t = h->fmm.tree();
  auto cp = [](int* dst, const std::vector<int>& src) { if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(int) * src.size()); };
  cp(parent, t.parent); cp(child, t.child); cp(depth, t.depth); cp(path, t.path); cp(tag, t.tag);
  cp(src_beg, t.src_beg); cp(src_num, t.src_num); cp(trg_beg, t.trg_beg); cp(trg_num, t.trg_num);
  cp(src_perm, t.src_perm); cp(trg_perm, t.trg_perm);
  return KIFMM_OK;
}

int kifmm_get_list(kifmm_handle h, int kind, int64_t* off, int* lst) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  if (kind < 0 || kind > 3) return set_err(h, KIFMM_ERR_ARG, "list kind must be 0..3");
  if (!h->fmm.isHostReady()) return set_err(h, KIFMM_ERR_STATE, "lists requested before setup");
  const auto& t = h->fmm.tree();
  if (off) std::memcpy(off, t.off[kind].data(), sizeof(int64_t) * t.off[kind].size());
  if (lst && !t.lst[kind].empty()) std::memcpy(lst, t.lst[kind].data(), sizeof(int) * t.lst[kind].size());
  return KIFMM_OK;
}

int kifmm_get_operator(kifmm_handle h, int which, double* out, int64_t* count) {
  if (!h || !count) return set_err(h, KIFMM_ERR_ARG, "null argument");
  const kifmm::Operators* op = h->fmm.ops();
  if (!h->fmm.isHostReady() || !op) return set_err(h, KIFMM_ERR_STATE, "operators requested before setup");
  const std::vector<double>* v = nullptr;
  switch (which) {
    case 0: v = &op->pinv_up; break;
    case 1: v = &op->pinv_dn; break;
    case 2: v = &op->m2m; break;
    case 3: v = &op->l2l; break;
    case 4: v = &op->m2l; break;
    default: return set_err(h, KIFMM_ERR_ARG, "operator id must be 0..4");
  }
  *count = int64_t(v->size());
  if (out) std::memcpy(out, v->data(), sizeof(double) * v->size());
  return KIFMM_OK;
}

int kifmm_set_timing(kifmm_handle h, int on) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  h->fmm.enableTiming(on != 0);
  return KIFMM_OK;
}

int kifmm_set_graph(kifmm_handle h, int on) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  h->fmm.enableGraph(on != 0);
  return KIFMM_OK;
}

int kifmm_phase_times(kifmm_handle h, float* ms, int n) {
  if (!h || !ms) return set_err(h, KIFMM_ERR_ARG, "null argument");
  auto p = h->fmm.phaseMs();
  for (int i = 0; i < n && i < kifmm::PH_TIMES; i++) ms[i] = p[i];
  return KIFMM_OK;
}

const char* kifmm_phase_name(int phase) {
  return (phase >= 0 && phase < kifmm::PH_TIMES) ? kifmm::kPhaseNames[phase] : "";
}

int kifmm_set_tree_builder(kifmm_handle h, int on_device) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  h->fmm.setTreeOnDevice(on_device != 0);
  return KIFMM_OK;
}

int kifmm_set_translation(kifmm_handle h, int terms) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.setTranslationTerms(terms);
}

int kifmm_autotune(kifmm_handle h, const double* d_src_den, double* d_trg_val, int nrhs, void* stream) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  return h->fmm.autotune(d_src_den, d_trg_val, nrhs, static_cast<hipStream_t>(stream));
}

int kifmm_get_tuning(kifmm_handle h, int* knobs, int n) {
  if (!h || !knobs) return set_err(h, KIFMM_ERR_ARG, "null argument");
  const kifmm::Tuning& t = h->fmm.tuning();
  const int v[KIFMM_NTUNE] = {t.near_mode, t.near_stage, t.trans, t.vec_hadamard};
  for (int i = 0; i < n && i < KIFMM_NTUNE; i++) knobs[i] = v[i];
  return KIFMM_OK;
}

int kifmm_set_tuning(kifmm_handle h, const int* knobs, int n) {
  if (!h || !knobs || n < KIFMM_NTUNE) return set_err(h, KIFMM_ERR_ARG, "need KIFMM_NTUNE knob values");
  kifmm::Tuning t;
  t.near_mode = knobs[0];
  t.near_stage = knobs[1];
  t.trans = knobs[2];
  t.vec_hadamard = knobs[3];
  return h->fmm.apply_tuning(t);
}

int kifmm_tuning_report(kifmm_handle h, double* rows, int maxrows) {
  if (!h) return set_err(nullptr, KIFMM_ERR_ARG, "null handle");
  const auto& r = h->fmm.tuneReport();
  for (size_t i = 0; rows && i < r.size() && int(i) < maxrows; i++) {
    rows[3 * i] = r[i].knob;
    rows[3 * i + 1] = r[i].value;
    rows[3 * i + 2] = r[i].ms;
  }
  return int(r.size());
}

int kifmm_setup_times(kifmm_handle h, double* ms, int n) {
  if (!h || !ms) return set_err(h, KIFMM_ERR_ARG, "null argument");
  const auto& s = h->fmm.setupMs();
  for (int i = 0; i < n && i < KIFMM_NSETUP; i++) ms[i] = s[size_t(i)];
  return KIFMM_OK;
}

int kifmm_work_counts(kifmm_handle h, double* out, int n) {
  if (!h || !out) return set_err(h, KIFMM_ERR_ARG, "null argument");
  if (!h->fmm.isHostReady()) return set_err(h, KIFMM_ERR_STATE, "work counts requested before setup");
  const auto& w = h->fmm.work();
  const double v[KIFMM_NWORK] = {w.near_pairs, w.near_exact_pairs, w.s2m_pairs, w.s2l_pairs, w.vlist_pairs,
                                 w.n_vsrc, w.n_vtrg, w.n_vgrp, w.up_cols, w.m2m_cols, w.dn_cols, w.l2l_cols};
  for (int i = 0; i < n && i < KIFMM_NWORK; i++) out[i] = v[i];
  return KIFMM_OK;
}

int kifmm_options_lookup(const char* path, const char* name, char* value, int len) {
  if (!path || !name) return set_err(nullptr, KIFMM_ERR_ARG, "null argument");
  kifmm::Options o;
  std::string err, v;
  if (o.load(path, &err)) return set_err(nullptr, KIFMM_ERR_OPTIONS, err);
  if (!o.get(name, &v)) return 0;
  if (value && len > 0) {
    std::strncpy(value, v.c_str(), size_t(len) - 1);
    value[len - 1] = 0;
  }
  return 1;
}

const char* kifmm_last_error(kifmm_handle h) {
  return h ? h->fmm.lastError().c_str() : g_create_error.c_str();
}

}  // extern "C"
