// Kernels of the adjoint (transposed) evaluate pass and of the field pass, FP64.
// In the adjoint every block is the forward kernel block K(target - source;
// source normal) contracted over the target DOF instead of the source DOF, so
// the pass is the exact transpose of the forward operator whatever the
// kernel's symmetry.  No rounding-order contract applies (FMA contraction on).
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace kifmm {

// ---------------------------------------------------------------------------
// Source-centric near field: one workgroup per chunk of <= kNearChunk sources
// of a source leaf; the points of each item (targets of a target leaf with
// their adjoint inputs, or the DC / UC check samples of a node with the
// adjoint check values) are staged through LDS, and the lanes are split into
// S groups over the staged points as in the forward near field.
// sigma_bar(s) = sum_t K(t - s; n_s)^T w(t)   (U, W and X point pairs)
//              + sum_q K(x_q - s; n_s)^T DCbar_q / UCbar_q   (S2L, S2M)
// ---------------------------------------------------------------------------
template <int KT>
__global__ __launch_bounds__(kNearChunk) void k_near_t(NearTArgs a, KParams<double> pu) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  constexpr bool NOR = needs_normals_ct(KT);
  __shared__ double sp[kNearChunk * 3];
  __shared__ double sv[kNearChunk * TD];
  __shared__ double red[kNearChunk * SD];
  const NearChunk ch = a.chunks[blockIdx.x];
  const int lane = threadIdx.x;
  const int tn = max(ch.tnum, 1);
  int S = kNearChunk / tn;
  S = S > 16 ? 16 : S;
  const int si = lane % tn, sub = lane / tn;
  const bool active = sub < S && ch.tnum > 0;
  double sx = 0, sy = 0, sz = 0, nx = 0, ny = 0, nz = 0;
  if (active) {
    const int64_t q = int64_t(ch.tbeg + si);
    sx = a.spos[3 * q]; sy = a.spos[3 * q + 1]; sz = a.spos[3 * q + 2];
    if constexpr (NOR) { nx = a.snor[3 * q]; ny = a.snor[3 * q + 1]; nz = a.snor[3 * q + 2]; }
  }
  double acc[SD];
#pragma unroll
  for (int s = 0; s < SD; s++) acc[s] = 0.0;
  for (int it = ch.ibeg; it < ch.iend; it++) {
    const int code = a.items[it];
    const int node = code >> 2, kind = code & 3;
    int np_, b = 0;
    double4 g = make_double4(0, 0, 0, 0);
    const double* sam = nullptr;
    const double* val = nullptr;
    if (kind == ADJ_TRG) {
      b = a.trg_beg[node]; np_ = a.trg_num[node];
    } else {
      g = a.geom[node];
      const bool dc = kind == ADJ_DC;
      sam = dc ? a.sam_dc : a.sam_uc;
      np_ = dc ? a.n_dc : a.n_uc;
      val = dc ? a.DC + int64_t(node) * a.sz_dc : a.UC + int64_t(node) * a.sz_uc;
    }
    for (int base = 0; base < np_; base += kNearChunk) {
      const int m = min(kNearChunk, np_ - base);
      __syncthreads();
      if (lane < m) {
        if (kind == ADJ_TRG) {
          const int64_t q = int64_t(b + base + lane);
          for (int d = 0; d < 3; d++) sp[3 * lane + d] = a.tpos[3 * q + d];
          for (int t = 0; t < TD; t++) sv[lane * TD + t] = a.tw[q * TD + t];
        } else {
          const int q = base + lane;
          sp[3 * lane] = g.x + g.w * sam[3 * q];
          sp[3 * lane + 1] = g.y + g.w * sam[3 * q + 1];
          sp[3 * lane + 2] = g.z + g.w * sam[3 * q + 2];
          for (int t = 0; t < TD; t++) sv[lane * TD + t] = val[q * TD + t];
        }
      }
      __syncthreads();
      if (active) {
        for (int j = sub; j < m; j += S) {
          double blk[SD * TD];
          kblock<KT, double>(pu, sp[3 * j] - sx, sp[3 * j + 1] - sy, sp[3 * j + 2] - sz, nx, ny, nz, blk);
#pragma unroll
          for (int t = 0; t < TD; t++)
#pragma unroll
            for (int s = 0; s < SD; s++) acc[s] += blk[t * SD + s] * sv[j * TD + t];
        }
      }
    }
  }
  __syncthreads();
  if (active && sub > 0)
    for (int s = 0; s < SD; s++) red[(lane - tn) * SD + s] = acc[s];
  __syncthreads();
  if (lane < ch.tnum) {
    for (int u = 1; u < S; u++)
      for (int s = 0; s < SD; s++) acc[s] += red[((u - 1) * tn + lane) * SD + s];
    double* o = a.sout + int64_t(ch.tbeg + lane) * SD;
#pragma unroll
    for (int s = 0; s < SD; s++) o[s] = acc[s];
  }
}

// ---------------------------------------------------------------------------
// Equivalent surface <- targets (adjoint of M2T into UE and of L2T into DE):
// out(node)[q] += sum_t K_mm(t - y_q)^T w(t); one workgroup per task, each lane
// owns up to two surface samples.
// ---------------------------------------------------------------------------
constexpr int kSurfT = 256;

template <int KM>
__global__ __launch_bounds__(kSurfT) void k_surf_t(SurfTArgs a, KParams<double> pm) {
  constexpr int sd = src_dof(KM), td = trg_dof(KM);
  __shared__ double sp[kSurfT * 3];
  __shared__ double sv[kSurfT * td];
  const SurfTask tk = a.tasks[blockIdx.x];
  const double4 g = a.geom[tk.node];
  const int lane = threadIdx.x;
  double yp[2][3];
  double acc[2][sd];
  bool act[2];
  for (int u = 0; u < 2; u++) {
    const int q = lane + u * kSurfT;
    act[u] = q < a.n_sam;
    const int qq = act[u] ? q : 0;
    yp[u][0] = g.x + g.w * a.sam[3 * qq];
    yp[u][1] = g.y + g.w * a.sam[3 * qq + 1];
    yp[u][2] = g.z + g.w * a.sam[3 * qq + 2];
    for (int s = 0; s < sd; s++) acc[u][s] = 0.0;
  }
  for (int it = tk.ibeg; it < tk.iend; it++) {
    const int leaf = a.items[it];
    const int b = a.trg_beg[leaf], nt = a.trg_num[leaf];
    for (int base = 0; base < nt; base += kSurfT) {
      const int m = min(kSurfT, nt - base);
      __syncthreads();
      if (lane < m) {
        const int64_t q = int64_t(b + base + lane);
        for (int d = 0; d < 3; d++) sp[3 * lane + d] = a.tpos[3 * q + d];
        for (int t = 0; t < td; t++) sv[lane * td + t] = a.tw[q * td + t];
      }
      __syncthreads();
      for (int j = 0; j < m; j++) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
          if (!act[u]) continue;
          double blk[sd * td];
          kblock<KM, double>(pm, sp[3 * j] - yp[u][0], sp[3 * j + 1] - yp[u][1], sp[3 * j + 2] - yp[u][2], 0.0,
                             0.0, 0.0, blk);
#pragma unroll
          for (int t = 0; t < td; t++)
#pragma unroll
            for (int s = 0; s < sd; s++) acc[u][s] += blk[t * sd + s] * sv[j * td + t];
        }
      }
    }
  }
  double* o = a.out + int64_t(tk.node) * a.sz_out;
  for (int u = 0; u < 2; u++) {
    if (!act[u]) continue;
    const int q = lane + u * kSurfT;
    for (int s = 0; s < sd; s++) o[q * sd + s] += acc[u][s];
  }
}

// ---------------------------------------------------------------------------
// Field pass, Laplace single layer: grad_t of a/|t - s| = -a (t - s)/|t - s|^3
// over the same items as the value near field (exact sources with the user
// kernel, UE / DE surfaces with the equivalent kernel; pairs closer than the
// reference's 1e-8 cut contribute nothing).  3 values per target.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNearChunk) void k_near_grad_lap(NearArgs a, double pa_u, double pa_m) {
  __shared__ double sp[kNearChunk * 3];
  __shared__ double sv[kNearChunk];
  __shared__ double red[kNearChunk * 3];
  const NearChunk ch = a.chunks[blockIdx.x];
  const int lane = threadIdx.x;
  const int tn = max(ch.tnum, 1);
  int S = kNearChunk / tn;
  S = S > 16 ? 16 : S;
  const int ti = lane % tn, sub = lane / tn;
  const bool active = sub < S;
  double tx = 0, ty = 0, tz = 0;
  if (active) {
    const double* p = a.tpos + 3 * int64_t(ch.tbeg + ti);
    tx = p[0]; ty = p[1]; tz = p[2];
  }
  double gx = 0, gy = 0, gz = 0;
  for (int it = ch.ibeg; it < ch.iend; it++) {
    const int code = a.items[it];
    const int node = code >> 2, kind = code & 3;
    const bool exact = kind == ITEM_EXACT;
    int ns, b = 0;
    double4 g = make_double4(0, 0, 0, 0);
    const double* sam = nullptr;
    const double* den = nullptr;
    if (exact) {
      b = a.src_beg[node]; ns = a.src_num[node];
    } else {
      g = a.geom[node];
      const bool ue = kind == ITEM_UE;
      sam = ue ? a.sam_ue : a.sam_de;
      ns = ue ? a.n_ue : a.n_de;
      den = ue ? a.UE + int64_t(node) * a.sz_ue : a.DE + int64_t(node) * a.sz_de;
    }
    for (int base = 0; base < ns; base += kNearChunk) {
      const int m = min(kNearChunk, ns - base);
      __syncthreads();
      if (lane < m) {
        if (exact) {
          const int64_t q = int64_t(b + base + lane);
          for (int d = 0; d < 3; d++) sp[3 * lane + d] = a.spos[3 * q + d];
          sv[lane] = a.sden[q] * pa_u;
        } else {
          const int q = base + lane;
          sp[3 * lane] = g.x + g.w * sam[3 * q];
          sp[3 * lane + 1] = g.y + g.w * sam[3 * q + 1];
          sp[3 * lane + 2] = g.z + g.w * sam[3 * q + 2];
          sv[lane] = den[q] * pa_m;
        }
      }
      __syncthreads();
      if (active) {
        for (int j = sub; j < m; j += S) {
          const double dx = tx - sp[3 * j], dy = ty - sp[3 * j + 1], dz = tz - sp[3 * j + 2];
          const double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < kMinDif * kMinDif) continue;
          const double inv = rsqrt(r2);
          const double c = -sv[j] * inv * inv * inv;
          gx += c * dx; gy += c * dy; gz += c * dz;
        }
      }
    }
  }
  __syncthreads();
  if (active && sub > 0) {
    red[(lane - tn) * 3] = gx; red[(lane - tn) * 3 + 1] = gy; red[(lane - tn) * 3 + 2] = gz;
  }
  __syncthreads();
  if (lane < ch.tnum) {
    for (int u = 1; u < S; u++) {
      gx += red[((u - 1) * tn + lane) * 3];
      gy += red[((u - 1) * tn + lane) * 3 + 1];
      gz += red[((u - 1) * tn + lane) * 3 + 2];
    }
    double* o = a.tval + int64_t(ch.tbeg + lane) * 3;
    o[0] = gx; o[1] = gy; o[2] = gz;
  }
}

hipError_t launch_near_grad(int kt, const NearArgs& a, const KParams<double>& pu, const KParams<double>& pm,
                            hipStream_t s) {
  if (kt != 111) return hipErrorInvalidValue;
  if (a.nchunks <= 0) return hipSuccess;
  k_near_grad_lap<<<dim3(unsigned(a.nchunks)), dim3(kNearChunk), 0, s>>>(a, pu.a, pm.a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
#define KIFMM_ADJ_KERNELS(X) X(111) X(121) X(311) X(312) X(321) X(322) X(511) X(521)

hipError_t launch_near_adjoint(int kt, const NearTArgs& a, const KParams<double>& pu, hipStream_t s) {
  if (a.nchunks <= 0) return hipSuccess;
  switch (kt) {
#define KIFMM_CASE(K) case K: k_near_t<K><<<dim3(unsigned(a.nchunks)), dim3(kNearChunk), 0, s>>>(a, pu); break;
    KIFMM_ADJ_KERNELS(KIFMM_CASE)
#undef KIFMM_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_surface_adjoint(int km, const SurfTArgs& a, const KParams<double>& pm, hipStream_t s) {
  if (a.ntasks <= 0) return hipSuccess;
  if (a.n_sam > 2 * kSurfT) return hipErrorInvalidValue;
  switch (km) {
    case 111: k_surf_t<111><<<dim3(unsigned(a.ntasks)), dim3(kSurfT), 0, s>>>(a, pm); break;
    case 301: k_surf_t<301><<<dim3(unsigned(a.ntasks)), dim3(kSurfT), 0, s>>>(a, pm); break;
    case 511: k_surf_t<511><<<dim3(unsigned(a.ntasks)), dim3(kSurfT), 0, s>>>(a, pm); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace kifmm
