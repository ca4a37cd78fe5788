// Mixed-precision kernels of the evaluate pass (gfx950).  Built with FP
// contraction: no rounding-order contract applies here.
//
//   near field (U, W, X fallbacks, M2T, L2T)  FP32 kernel evaluation in
//       coordinates relative to the target leaf centre, FP32 partial sums per
//       staged source block, FP64 running sums; near and far (L2T) parts summed
//       in FP64 at the end
//   S2M / S2L                                 same scheme, FP64 check values
//   V-list Hadamard                           FP32 spectra, FP32 FMA accumulation
//   translations and pseudo-inverse applies   FP64 GEMM with FMA
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "gemm.hpp"

namespace kifmm {

static KParams<float> to_f32(const KParams<double>& p) { return KParams<float>{float(p.a), float(p.b), float(p.c)}; }

// ---------------------------------------------------------------------------
template <int KT>
__device__ inline void pair_f32(const KParams<float>& p, float dx, float dy, float dz, float nx, float ny,
                                float nz, const float* den, float* acc) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  if constexpr (KT == 111) {
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float g = r2 < float(kMinDif * kMinDif) ? 0.0f : p.a * rsqrtf(r2);
    acc[0] = fmaf(g, den[0], acc[0]);
  } else {
    float blk[SD * TD];
    kblock<KT, float>(p, dx, dy, dz, nx, ny, nz, blk);
#pragma unroll
    for (int t = 0; t < TD; t++)
#pragma unroll
      for (int s = 0; s < SD; s++) acc[t] = fmaf(blk[t * SD + s], den[s], acc[t]);
  }
}

// Double-layer kernels: r and r.n from FP64 positions (r.n cancels for close
// pairs on a smooth surface), the rest of the kernel in FP32.
template <int KT>
__device__ inline void pair_f32_dl(const KParams<float>& p, double dx, double dy, double dz, float nx, float ny,
                                   float nz, const float* den, float* acc) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  const float rn = float(dx * double(nx) + dy * double(ny) + dz * double(nz));
  float blk[SD * TD];
  kblock_rn<KT, float>(p, float(dx), float(dy), float(dz), nx, ny, nz, rn, blk);
#pragma unroll
  for (int t = 0; t < TD; t++)
#pragma unroll
    for (int s = 0; s < SD; s++) acc[t] = fmaf(blk[t * SD + s], den[s], acc[t]);
}

template <int KT>
__global__ __launch_bounds__(kNearChunk) void k_near_mx(NearArgs a, KParams<float> pu, KParams<float> pm) {
  constexpr int KM = mm_kernel(KT);
  constexpr int SD = src_dof(KT), TD = trg_dof(KT), sd = src_dof(KM);
  constexpr int MAXD = SD > sd ? SD : sd;
  constexpr bool NOR = needs_normals_ct(KT);
  __shared__ float sp[kNearChunk * 3];
  __shared__ float sn[NOR ? kNearChunk * 3 : 1];
  __shared__ double sp64[NOR ? kNearChunk * 3 : 1];
  __shared__ float sv[kNearChunk * MAXD];
  __shared__ double red[kNearChunk * TD * 2];
  const NearChunk ch = a.chunks[blockIdx.x];
  const double4 cb = a.geom[ch.node];
  const int lane = threadIdx.x;
  int S = kNearChunk / max(ch.tnum, 1);
  S = S > 16 ? 16 : S;
  const int ti = lane % ch.tnum, sub = lane / ch.tnum;
  const bool active = sub < S;
  float tx = 0, ty = 0, tz = 0;
  double tx64 = 0, ty64 = 0, tz64 = 0;
  if (active) {
    const double* p = a.tpos + 3 * int64_t(ch.tbeg + ti);
    tx64 = p[0]; ty64 = p[1]; tz64 = p[2];
    tx = float(p[0] - cb.x); ty = float(p[1] - cb.y); tz = float(p[2] - cb.z);
  }
  double near[TD], far[TD];
#pragma unroll
  for (int t = 0; t < TD; t++) near[t] = far[t] = 0.0;

  for (int it = ch.ibeg; it < ch.iend; it++) {
    const int code = a.items[it];
    const int node = code >> 2, kind = code & 3;
    float acc[TD];
#pragma unroll
    for (int t = 0; t < TD; t++) acc[t] = 0.0f;
    if (kind == ITEM_EXACT) {
      const int b = a.src_beg[node], ns = a.src_num[node];
      for (int base = 0; base < ns; base += kNearChunk) {
        const int m = min(kNearChunk, ns - base);
        __syncthreads();
        if (lane < m) {
          const int64_t g = int64_t(b + base + lane);
          sp[3 * lane] = float(a.spos[3 * g] - cb.x);
          sp[3 * lane + 1] = float(a.spos[3 * g + 1] - cb.y);
          sp[3 * lane + 2] = float(a.spos[3 * g + 2] - cb.z);
          if constexpr (NOR)
            for (int d = 0; d < 3; d++) { sn[3 * lane + d] = float(a.snor[3 * g + d]); sp64[3 * lane + d] = a.spos[3 * g + d]; }
          for (int s = 0; s < SD; s++) sv[lane * SD + s] = float(a.sden[g * SD + s]);
        }
        __syncthreads();
        if (active) {
          for (int j = sub; j < m; j += S) {
            if constexpr (NOR) {
              pair_f32_dl<KT>(pu, tx64 - sp64[3 * j], ty64 - sp64[3 * j + 1], tz64 - sp64[3 * j + 2], sn[3 * j],
                              sn[3 * j + 1], sn[3 * j + 2], &sv[j * SD], acc);
            } else {
              pair_f32<KT>(pu, tx - sp[3 * j], ty - sp[3 * j + 1], tz - sp[3 * j + 2], 0.f, 0.f, 0.f, &sv[j * SD], acc);
            }
          }
        }
      }
#pragma unroll
      for (int t = 0; t < TD; t++) near[t] += double(acc[t]);
    } else {
      const double4 g = a.geom[node];
      const bool ue = kind == ITEM_UE;
      const double* sam = ue ? a.sam_ue : a.sam_de;
      const int ns = ue ? a.n_ue : a.n_de;
      const double* den = ue ? a.UE + int64_t(node) * a.sz_ue : a.DE + int64_t(node) * a.sz_de;
      for (int base = 0; base < ns; base += kNearChunk) {
        const int m = min(kNearChunk, ns - base);
        __syncthreads();
        if (lane < m) {
          const int q = base + lane;
          sp[3 * lane] = float((g.x - cb.x) + g.w * sam[3 * q]);
          sp[3 * lane + 1] = float((g.y - cb.y) + g.w * sam[3 * q + 1]);
          sp[3 * lane + 2] = float((g.z - cb.z) + g.w * sam[3 * q + 2]);
          for (int s = 0; s < sd; s++) sv[lane * sd + s] = float(den[q * sd + s]);
        }
        __syncthreads();
        if (active) {
          for (int j = sub; j < m; j += S) {
            const float x = sp[3 * j], y = sp[3 * j + 1], z = sp[3 * j + 2];
            pair_f32<KM>(pm, tx - x, ty - y, tz - z, x, y, z, &sv[j * sd], acc);
          }
        }
      }
      if (ue) {
#pragma unroll
        for (int t = 0; t < TD; t++) near[t] += double(acc[t]);
      } else {
#pragma unroll
        for (int t = 0; t < TD; t++) far[t] += double(acc[t]);
      }
    }
  }
  __syncthreads();
  if (active && sub > 0)
    for (int t = 0; t < TD; t++) {
      red[((lane - ch.tnum) * TD + t) * 2] = near[t];
      red[((lane - ch.tnum) * TD + t) * 2 + 1] = far[t];
    }
  __syncthreads();
  if (lane < ch.tnum) {
    for (int u = 1; u < S; u++)
      for (int t = 0; t < TD; t++) {
        near[t] += red[(((u - 1) * ch.tnum + lane) * TD + t) * 2];
        far[t] += red[(((u - 1) * ch.tnum + lane) * TD + t) * 2 + 1];
      }
    double* o = a.tval + int64_t(ch.tbeg + lane) * TD;
#pragma unroll
    for (int t = 0; t < TD; t++) o[t] = near[t] + far[t];
  }
}

// ---------------------------------------------------------------------------
constexpr int kSurfMx = 256;

template <int KT>
__global__ __launch_bounds__(kSurfMx) void k_surface_mx(SurfArgs a, KParams<float> pu) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  constexpr bool NOR = needs_normals_ct(KT);
  __shared__ float sp[kSurfMx * 3];
  __shared__ float sn[NOR ? kSurfMx * 3 : 1];
  __shared__ float sv[kSurfMx * SD];
  const SurfTask tk = a.tasks[blockIdx.x];
  const double4 g = a.geom[tk.node];
  const int lane = threadIdx.x;
  float tp[2][3];
  double acc[2][TD];
  bool act[2];
  for (int u = 0; u < 2; u++) {
    const int q = lane + u * kSurfMx;
    act[u] = q < a.n_sam;
    const int qq = act[u] ? q : 0;
    for (int d = 0; d < 3; d++) tp[u][d] = float(g.w * a.sam[3 * qq + d]);
    for (int t = 0; t < TD; t++) acc[u][t] = 0.0;
  }
  for (int it = tk.ibeg; it < tk.iend; it++) {
    const int node = a.items[it];
    const int b = a.src_beg[node], ns = a.src_num[node];
    for (int base = 0; base < ns; base += kSurfMx) {
      const int m = min(kSurfMx, ns - base);
      __syncthreads();
      if (lane < m) {
        const int64_t gg = int64_t(b + base + lane);
        sp[3 * lane] = float(a.spos[3 * gg] - g.x);
        sp[3 * lane + 1] = float(a.spos[3 * gg + 1] - g.y);
        sp[3 * lane + 2] = float(a.spos[3 * gg + 2] - g.z);
        if constexpr (NOR) for (int d = 0; d < 3; d++) sn[3 * lane + d] = float(a.snor[3 * gg + d]);
        for (int s = 0; s < SD; s++) sv[lane * SD + s] = float(a.sden[gg * SD + s]);
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 2; u++) {
        if (!act[u]) continue;
        float part[TD];
        for (int t = 0; t < TD; t++) part[t] = 0.0f;
        for (int j = 0; j < m; j++) {
          float nx = 0, ny = 0, nz = 0;
          if constexpr (NOR) { nx = sn[3 * j]; ny = sn[3 * j + 1]; nz = sn[3 * j + 2]; }
          pair_f32<KT>(pu, tp[u][0] - sp[3 * j], tp[u][1] - sp[3 * j + 1], tp[u][2] - sp[3 * j + 2], nx, ny, nz,
                       &sv[j * SD], part);
        }
        for (int t = 0; t < TD; t++) acc[u][t] += double(part[t]);
      }
    }
  }
  double* o = a.out + int64_t(tk.node) * a.sz_out;
  for (int u = 0; u < 2; u++) {
    if (!act[u]) continue;
    const int q = lane + u * kSurfMx;
    for (int t = 0; t < TD; t++) o[q * TD + t] += acc[u][t];
  }
}

// ---------------------------------------------------------------------------
// V-list Hadamard with FP32 spectra: one wave per sibling group and 64
// frequencies, lane = frequency, the 8 targets' accumulators per lane.
template <int SD, int TD>
__global__ __launch_bounds__(64) void k_hadamard_mx(VlistArgs v, const float2* __restrict__ dspec,
                                                   const float2* __restrict__ m2l, float2* __restrict__ espec,
                                                   float nrm) {
  const int g = blockIdx.x;
  const int f = blockIdx.y * 64 + threadIdx.x;
  const int H = v.H;
  const bool ok = f < H;
  const int ff = ok ? f : 0;
  float2 acc[8][TD];
#pragma unroll
  for (int t = 0; t < 8; t++)
#pragma unroll
    for (int i = 0; i < TD; i++) acc[t][i] = make_float2(0.f, 0.f);
  const int cnt = v.pcnt[g];
  const int* ps = v.psrc + int64_t(g) * kPairStride;
  const int* pr = v.pslot + int64_t(g) * kPairStride;
  for (int k = 0; k < cnt; k++) {
    const int s = ps[k], word = pr[k];
    if (s < 0) continue;
    float2 d[SD];
#pragma unroll
    for (int j = 0; j < SD; j++) d[j] = dspec[(int64_t(s) * SD + j) * H + ff];
    const int rs = (word >> 8) & 0x3ff, tm = word >> 20;
#pragma unroll
    for (int tau = 0; tau < 8; tau++) {
      if (!((tm >> tau) & 1)) continue;
      const int slot = rs - ((tau >> 2) * 49 + ((tau >> 1) & 1) * 7 + (tau & 1));
      const float2* T = m2l + int64_t(slot) * SD * TD * H;
#pragma unroll
      for (int i = 0; i < TD; i++)
#pragma unroll
        for (int j = 0; j < SD; j++) {
          const float2 t = T[int64_t(j * TD + i) * H + ff];
          acc[tau][i].x = fmaf(t.x, d[j].x, fmaf(-t.y, d[j].y, acc[tau][i].x));
          acc[tau][i].y = fmaf(t.x, d[j].y, fmaf(t.y, d[j].x, acc[tau][i].y));
        }
    }
  }
  if (!ok) return;
#pragma unroll
  for (int tau = 0; tau < 8; tau++) {
    const int t = v.tgt[g * 8 + tau];
    if (t < 0) continue;
#pragma unroll
    for (int i = 0; i < TD; i++)
      espec[(int64_t(t) * TD + i) * H + f] = make_float2(acc[tau][i].x * nrm, acc[tau][i].y * nrm);
  }
}

// ---------------------------------------------------------------------------
// Scalar-kernel Hadamard (sd = td = 1): the FP32 translation spectra of a
// 16-frequency slice for all 343 offset slots sit in LDS (47 KB), so a spectrum
// value is read from LDS instead of L2 for every pair.  Lane = (frequency,
// sibling group): 16 frequencies x 4 groups per wave; each lane accumulates the
// 8 target children of its group, loads each halo source spectrum once, and
// runs 4 sources ahead to keep loads in flight.
constexpr int kHxF = 16, kHxRow = kHxF + 1, kHxGroupsPerWave = 64 / kHxF;
constexpr int kHxWaves = 4, kHxBatch = 8;

__global__ __launch_bounds__(64 * kHxWaves) void k_hadamard_mx_s(VlistArgs v, const float2* __restrict__ dspec,
                                                              const float2* __restrict__ m2l,
                                                              float2* __restrict__ espec, float nrm) {
  __shared__ float2 Ts[343 * kHxRow];
  const int H = v.H;
  const int f0 = blockIdx.x * kHxF;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int fl = lane % kHxF, gq = lane / kHxF;
  const int f = f0 + fl;
  const bool fok = f < H;
  const int fs = fok ? f : 0;
  for (int e = tid; e < 343 * kHxF; e += 64 * kHxWaves) {
    const int sl = e / kHxF, ff = f0 + e % kHxF;
    Ts[sl * kHxRow + e % kHxF] = ff < H ? m2l[int64_t(sl) * H + ff] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  constexpr int per_pass = kHxWaves * kHxGroupsPerWave;
  const int gbase = blockIdx.y * kHxBatch * per_pass;
  for (int pass = 0; pass < kHxBatch; pass++) {
    const int g = gbase + pass * per_pass + wave * kHxGroupsPerWave + gq;
    const bool gok = g < v.ngrp;
    const int cnt = gok ? v.pcnt[g] : 0;
    const int* ps = v.psrc + int64_t(gok ? g : 0) * kPairStride;
    const int* pr = v.pslot + int64_t(gok ? g : 0) * kPairStride;
    float2 acc[8];
#pragma unroll
    for (int t = 0; t < 8; t++) acc[t] = make_float2(0.f, 0.f);
    for (int k = 0; k < cnt; k += 4) {  // cnt is a multiple of 8
      int sr[4], wd[4];
      float2 d[4];
#pragma unroll
      for (int u = 0; u < 4; u++) { sr[u] = ps[k + u]; wd[u] = pr[k + u]; }
#pragma unroll
      for (int u = 0; u < 4; u++) d[u] = dspec[int64_t(sr[u] < 0 ? 0 : sr[u]) * H + fs];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (sr[u] < 0) continue;
        const int rs = (wd[u] >> 8) & 0x3ff, tm = wd[u] >> 20;
#pragma unroll
        for (int t = 0; t < 8; t++) {
          if (!((tm >> t) & 1)) continue;
          const float2 tt = Ts[(rs - ((t >> 2) * 49 + ((t >> 1) & 1) * 7 + (t & 1))) * kHxRow + fl];
          acc[t].x = fmaf(tt.x, d[u].x, fmaf(-tt.y, d[u].y, acc[t].x));
          acc[t].y = fmaf(tt.x, d[u].y, fmaf(tt.y, d[u].x, acc[t].y));
        }
      }
    }
    if (gok && fok) {
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const int tg = v.tgt[g * 8 + t];
        if (tg >= 0) espec[int64_t(tg) * H + f] = make_float2(acc[t].x * nrm, acc[t].y * nrm);
      }
    }
  }
}

// ---------------------------------------------------------------------------
#define KIFMM_USER_KERNELS_MX(X) X(111) X(121) X(311) X(312) X(321) X(322) X(511) X(521)

hipError_t launch_near_mixed(int kt, const NearArgs& a, const KParams<double>& pu, const KParams<double>& pm,
                             hipStream_t s) {
  if (a.nchunks <= 0) return hipSuccess;
  const KParams<float> fu = to_f32(pu), fm = to_f32(pm);
  switch (kt) {
#define KIFMM_CASE(K) case K: k_near_mx<K><<<dim3(unsigned(a.nchunks)), dim3(kNearChunk), 0, s>>>(a, fu, fm); break;
    KIFMM_USER_KERNELS_MX(KIFMM_CASE)
#undef KIFMM_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_surface_mixed(int kt, const SurfArgs& a, const KParams<double>& pu, hipStream_t s) {
  if (a.ntasks <= 0) return hipSuccess;
  if (a.n_sam > 2 * kSurfMx) return hipErrorInvalidValue;
  const KParams<float> fu = to_f32(pu);
  switch (kt) {
#define KIFMM_CASE(K) case K: k_surface_mx<K><<<dim3(unsigned(a.ntasks)), dim3(kSurfMx), 0, s>>>(a, fu); break;
    KIFMM_USER_KERNELS_MX(KIFMM_CASE)
#undef KIFMM_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_hadamard_mixed(const VlistArgs& v, const float2* dspec, const float2* m2l, float2* espec,
                                 hipStream_t s) {
  if (v.ngrp <= 0) return hipSuccess;
  const float nrm = float(1.0 / double(v.n * v.n * v.n));
  dim3 grid(unsigned(v.ngrp), unsigned((v.H + 63) / 64));
  if (v.sd == 1 && v.td == 1) {
    constexpr int per_wg = kHxBatch * kHxWaves * kHxGroupsPerWave;
    k_hadamard_mx_s<<<dim3(unsigned((v.H + kHxF - 1) / kHxF), unsigned((v.ngrp + per_wg - 1) / per_wg)),
                      dim3(64 * kHxWaves), 0, s>>>(v, dspec, m2l, espec, nrm);
  }
  else if (v.sd == 4 && v.td == 3) k_hadamard_mx<4, 3><<<grid, dim3(64), 0, s>>>(v, dspec, m2l, espec, nrm);
  else if (v.sd == 3 && v.td == 3) k_hadamard_mx<3, 3><<<grid, dim3(64), 0, s>>>(v, dspec, m2l, espec, nrm);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_gemm_fma(const GemmArgs& g, hipStream_t s) { return launch_gemm_t<true>(g, s); }

}  // namespace kifmm
