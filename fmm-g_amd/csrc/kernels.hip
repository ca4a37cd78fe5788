// HIP kernels of the FP64 evaluate pass for gfx950 (CDNA4, wave64).
//
// Phase -> kernel:
//   load densities (fmm3d_eval_mpi.cpp:443-461)           k_gather
//   S2M (:577) / X-list S2L (:1253)                       k_surface
//   M2M, UC->UE, L2L, DC->DE (:583-592, :1371-1382)       k_gemm (level-batched, indexed columns)
//   V-list (:1007-1059)                                   k_fft_fwd, k_hadamard, k_fft_inv
//   U-list S2T, W-list S2T/M2T, X-list S2T, L2T
//     (:773-785, :1146-1165, :1249, :1383-1435)            k_near (one pass per target chunk)
//   write back (:1490-1510)                               k_scatter
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "device.hpp"
#include "gemm.hpp"

namespace kifmm {

// ---------------------------------------------------------------------------
__global__ void k_gather(const double* __restrict__ in, const int* __restrict__ perm,
                         double* __restrict__ out, int64_t n, int dof) {
  int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int64_t o = perm[k];
  for (int d = 0; d < dof; d++) out[k * dof + d] = in[o * dof + d];
}

__global__ void k_scatter(const double* __restrict__ in, const int* __restrict__ perm,
                          double* __restrict__ out, int64_t n, int dof) {
  int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int64_t o = perm[k];
  for (int d = 0; d < dof; d++) out[o * dof + d] = in[k * dof + d];
}

// Densities of nrhs right-hand sides, caller's layout [r][k][dof] -> leaf order
// [sinv[k]][r][dof] (the right-hand sides of a source side by side).
__global__ void k_load_rhs(const double* __restrict__ in, const int* __restrict__ sinv, double* __restrict__ out,
                           int64_t n, int dof, int nrhs) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * nrhs) return;
  const int64_t r = i / n, k = i % n, o = sinv[k];
  for (int d = 0; d < dof; d++) out[(o * nrhs + r) * dof + d] = in[i * dof + d];
}

// Values of nrhs right-hand sides, leaf order [r][k][dof] -> caller's [r][perm[k]][dof].
__global__ void k_store_rhs(const double* __restrict__ in, const int* __restrict__ perm, double* __restrict__ out,
                            int64_t n, int dof, int nrhs) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * nrhs) return;
  const int64_t r = i / n, k = i % n, o = r * n + perm[k];
  for (int d = 0; d < dof; d++) out[o * dof + d] = in[i * dof + d];
}

__global__ void k_fill(double* p, int64_t n, double v) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = v;
}

hipError_t launch_gather(const double* in, const int* perm, double* out, int64_t n, int dof, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_gather<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, s>>>(in, perm, out, n, dof);
  return hipGetLastError();
}
hipError_t launch_scatter(const double* in, const int* perm, double* out, int64_t n, int dof, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_scatter<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, s>>>(in, perm, out, n, dof);
  return hipGetLastError();
}
hipError_t launch_load_rhs(const double* in, const int* sinv, double* out, int64_t n, int dof, int nrhs,
                           hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  k_load_rhs<<<dim3(unsigned((n * nrhs + 255) / 256)), dim3(256), 0, s>>>(in, sinv, out, n, dof, nrhs);
  return hipGetLastError();
}
hipError_t launch_store_rhs(const double* in, const int* perm, double* out, int64_t n, int dof, int nrhs,
                            hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  k_store_rhs<<<dim3(unsigned((n * nrhs + 255) / 256)), dim3(256), 0, s>>>(in, perm, out, n, dof, nrhs);
  return hipGetLastError();
}
// One thread per (row, right-hand side, column); 16-byte stores when the row
// length is even (rows of 8-byte-aligned length: 152, 296, ...).
__global__ void k_zero_rows(double* base, const int* rows, int nrows, int len, int nrhs, int64_t nstride) {
  const int half = len / 2;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(nrows) * nrhs * half) return;
  const int col = int(i % half);
  const int64_t kr = i / half;
  const int k = int(kr % nrows), r = int(kr / nrows);
  double* row = base + (int64_t(rows[k]) + r * nstride) * len;
  reinterpret_cast<double2*>(row)[col] = make_double2(0.0, 0.0);
}
__global__ void k_zero_rows_odd(double* base, const int* rows, int nrows, int len, int nrhs, int64_t nstride) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(nrows) * nrhs * len) return;
  const int col = int(i % len);
  const int64_t kr = i / len;
  const int k = int(kr % nrows), r = int(kr / nrows);
  base[(int64_t(rows[k]) + r * nstride) * len + col] = 0.0;
}
hipError_t launch_zero_rows(double* base, const int* rows, int nrows, int len, int nrhs, int64_t nstride,
                            hipStream_t s) {
  if (nrows <= 0 || nrhs <= 0 || len <= 0) return hipSuccess;
  if (len % 2 == 0) {
    const int64_t n = int64_t(nrows) * nrhs * (len / 2);
    k_zero_rows<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, s>>>(base, rows, nrows, len, nrhs, nstride);
  } else {
    const int64_t n = int64_t(nrows) * nrhs * len;
    k_zero_rows_odd<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, s>>>(base, rows, nrows, len, nrhs, nstride);
  }
  return hipGetLastError();
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_fill<<<dim3(unsigned(blocks)), dim3(256), 0, s>>>(p, n, v);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Near field: one workgroup per chunk of <= 128 targets of an evaluator leaf.
// Source items are staged through LDS 128 at a time.  Leaves are small (a few
// tens of targets), so the 128 lanes are split into S groups that take every
// S-th source of a stage (S = 128 / targets, at most 16); the S partial sums of
// a target are reduced through LDS at the end.  The near field feeds no
// pseudo-inverse, so its summation order is free; for the Laplace kernel 1/r
// comes from rsqrt.
// ---------------------------------------------------------------------------
template <int KT, int NR>
__device__ inline void near_pair(const KParams<double>& p, double dx, double dy, double dz, double nx,
                                 double ny, double nz, const double* den, double (*acc)[trg_dof(KT)]) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  if constexpr (KT == 111) {
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double g = r2 < kMinDif * kMinDif ? 0.0 : p.a * rsqrt(r2);
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r][0] += g * den[r];
  } else {
    double blk[SD * TD];
    kblock<KT, double>(p, dx, dy, dz, nx, ny, nz, blk);
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int t = 0; t < TD; t++)
#pragma unroll
        for (int s = 0; s < SD; s++) acc[r][t] += blk[t * SD + s] * den[r * SD + s];
  }
}

// NR: right-hand sides per launch (the kernel value of a pair is formed once and
// applied to all of them); each right-hand side's sums run in the single-RHS
// order, so its values are bitwise those of a one-RHS launch.
template <int KT, int NR>
__global__ __launch_bounds__(kNearChunk) void k_near(NearArgs a, KParams<double> pu, KParams<double> pm) {
  constexpr int KM = mm_kernel(KT);
  constexpr int SD = src_dof(KT), TD = trg_dof(KT), sd = src_dof(KM);
  constexpr int MAXD = SD > sd ? SD : sd;
  constexpr bool NOR = needs_normals_ct(KT);
  static_assert(trg_dof(KM) == TD, "user and equivalent kernels share the target DOF");
  __shared__ double sp[kNearChunk * 3];
  __shared__ double sn[NOR ? kNearChunk * 3 : 1];
  __shared__ double sv[kNearChunk * MAXD * NR];
  __shared__ double red[kNearChunk * TD];
  const NearChunk ch = a.chunks[blockIdx.x];
  const int lane = threadIdx.x;
  int S = kNearChunk / max(ch.tnum, 1);
  S = S > 16 ? 16 : S;
  const int ti = lane % ch.tnum, sub = lane / ch.tnum;
  const bool active = sub < S;
  double tx = 0, ty = 0, tz = 0;
  if (active) {
    const double* p = a.tpos + 3 * int64_t(ch.tbeg + ti);
    tx = p[0]; ty = p[1]; tz = p[2];
  }
  double acc[NR][TD];
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int t = 0; t < TD; t++) acc[r][t] = 0.0;

  for (int it = ch.ibeg; it < ch.iend; it++) {
    const int code = a.items[it];
    const int node = code >> 2, kind = code & 3;
    if (kind == ITEM_EXACT) {
      const int b = a.src_beg[node], ns = a.src_num[node];
      for (int base = 0; base < ns; base += kNearChunk) {
        const int m = min(kNearChunk, ns - base);
        __syncthreads();
        if (lane < m) {
          const int64_t g = int64_t(b + base + lane);
          for (int d = 0; d < 3; d++) sp[3 * lane + d] = a.spos[3 * g + d];
          if constexpr (NOR) for (int d = 0; d < 3; d++) sn[3 * lane + d] = a.snor[3 * g + d];
#pragma unroll
          for (int r = 0; r < NR; r++)
            for (int s = 0; s < SD; s++)
              sv[(lane * NR + r) * SD + s] = r < a.nr ? a.sden[(g * a.nrhs + a.r0 + r) * SD + s] : 0.0;
        }
        __syncthreads();
        if (active) {
          for (int j = sub; j < m; j += S) {
            double nx = 0, ny = 0, nz = 0;
            if constexpr (NOR) { nx = sn[3 * j]; ny = sn[3 * j + 1]; nz = sn[3 * j + 2]; }
            near_pair<KT, NR>(pu, tx - sp[3 * j], ty - sp[3 * j + 1], tz - sp[3 * j + 2], nx, ny, nz,
                              &sv[j * NR * SD], acc);
          }
        }
      }
    } else {
      const double4 g = a.geom[node];
      const bool ue = kind == ITEM_UE;
      const double* sam = ue ? a.sam_ue : a.sam_de;
      const int ns = ue ? a.n_ue : a.n_de;
      const double* den = ue ? a.UE + int64_t(node) * a.sz_ue : a.DE + int64_t(node) * a.sz_de;
      const int64_t rstep = a.nstride * (ue ? a.sz_ue : a.sz_de);
      for (int base = 0; base < ns; base += kNearChunk) {
        const int m = min(kNearChunk, ns - base);
        __syncthreads();
        if (lane < m) {
          const int q = base + lane;
          sp[3 * lane] = g.x + g.w * sam[3 * q];
          sp[3 * lane + 1] = g.y + g.w * sam[3 * q + 1];
          sp[3 * lane + 2] = g.z + g.w * sam[3 * q + 2];
#pragma unroll
          for (int r = 0; r < NR; r++)
            for (int s = 0; s < sd; s++)
              sv[(lane * NR + r) * sd + s] = r < a.nr ? den[(a.r0 + r) * rstep + q * sd + s] : 0.0;
        }
        __syncthreads();
        if (active) {
          for (int j = sub; j < m; j += S) {
            const double x = sp[3 * j], y = sp[3 * j + 1], z = sp[3 * j + 2];
            near_pair<KM, NR>(pm, tx - x, ty - y, tz - z, x, y, z, &sv[j * NR * sd], acc);
          }
        }
      }
    }
  }
  // reduce the S partial sums of each target, one right-hand side at a time
#pragma unroll
  for (int r = 0; r < NR; r++) {
    __syncthreads();
    if (active && sub > 0)
      for (int t = 0; t < TD; t++) red[(lane - ch.tnum) * TD + t] = acc[r][t];
    __syncthreads();
    if (lane < ch.tnum && r < a.nr) {
      double v[TD];
#pragma unroll
      for (int t = 0; t < TD; t++) v[t] = acc[r][t];
      for (int u = 1; u < S; u++)
        for (int t = 0; t < TD; t++) v[t] += red[((u - 1) * ch.tnum + lane) * TD + t];
      double* o = a.tval + int64_t(a.r0 + r) * a.tstride + int64_t(ch.tbeg + lane) * TD;
#pragma unroll
      for (int t = 0; t < TD; t++) o[t] = v[t];
    }
  }
}

// ---------------------------------------------------------------------------
// Exact sources -> check surface of a node (S2M into UC, S2L into DC).
// One workgroup per task; each lane owns up to two surface points.
// ---------------------------------------------------------------------------
constexpr int kSurfThreads = 256;

template <int KT, int NR>
__global__ __launch_bounds__(kSurfThreads) void k_surface(SurfArgs a, KParams<double> pu) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  __shared__ double sp[kSurfThreads * 3];
  __shared__ double sn[kSurfThreads * 3];
  __shared__ double sv[kSurfThreads * SD * NR];
  const SurfTask tk = a.tasks[blockIdx.x];
  const double4 g = a.geom[tk.node];
  const int lane = threadIdx.x;
  const int64_t rstep = a.nstride * a.sz_out;
  double tp[2][3];
  double acc[2][NR][TD];
  bool act[2];
  for (int u = 0; u < 2; u++) {
    const int q = lane + u * kSurfThreads;
    act[u] = q < a.n_sam;
    const int qq = act[u] ? q : 0;
    tp[u][0] = g.x + g.w * a.sam[3 * qq];
    tp[u][1] = g.y + g.w * a.sam[3 * qq + 1];
    tp[u][2] = g.z + g.w * a.sam[3 * qq + 2];
    // start from the current check value and add term by term (dgemv order)
#pragma unroll
    for (int r = 0; r < NR; r++)
      for (int t = 0; t < TD; t++)
        acc[u][r][t] = act[u] && r < a.nr
                           ? a.out[(a.r0 + r) * rstep + int64_t(tk.node) * a.sz_out + q * TD + t] : 0.0;
  }
  for (int it = tk.ibeg; it < tk.iend; it++) {
    const int node = a.items[it];
    const int b = a.src_beg[node], ns = a.src_num[node];
    for (int base = 0; base < ns; base += kSurfThreads) {
      const int m = min(kSurfThreads, ns - base);
      __syncthreads();
      if (lane < m) {
        const int64_t gg = int64_t(b + base + lane);
        for (int d = 0; d < 3; d++) { sp[3 * lane + d] = a.spos[3 * gg + d]; sn[3 * lane + d] = a.snor[3 * gg + d]; }
#pragma unroll
        for (int r = 0; r < NR; r++)
          for (int s = 0; s < SD; s++)
            sv[(lane * NR + r) * SD + s] = r < a.nr ? a.sden[(gg * a.nrhs + a.r0 + r) * SD + s] : 0.0;
      }
      __syncthreads();
      for (int j = 0; j < m; j++) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
          if (!act[u]) continue;
          double blk[SD * TD];
          kblock<KT, double>(pu, tp[u][0] - sp[3 * j], tp[u][1] - sp[3 * j + 1], tp[u][2] - sp[3 * j + 2],
                             sn[3 * j], sn[3 * j + 1], sn[3 * j + 2], blk);
#pragma unroll
          for (int r = 0; r < NR; r++)
#pragma unroll
            for (int t = 0; t < TD; t++)
#pragma unroll
              for (int s = 0; s < SD; s++) acc[u][r][t] += blk[t * SD + s] * sv[(j * NR + r) * SD + s];
        }
      }
    }
  }
  for (int u = 0; u < 2; u++) {
    if (!act[u]) continue;
    const int q = lane + u * kSurfThreads;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      if (r >= a.nr) continue;
      double* o = a.out + (a.r0 + r) * rstep + int64_t(tk.node) * a.sz_out;
      for (int t = 0; t < TD; t++) o[q * TD + t] = acc[u][r][t];
    }
  }
}

// ---------------------------------------------------------------------------
// Level-batched translations and pseudo-inverse applies: see gemm.hpp.
hipError_t launch_gemm(const GemmArgs& g, hipStream_t s) { return launch_gemm_t<false>(g, s); }

// ---------------------------------------------------------------------------
// V list.  Forward: UE samples (scaled 2^(l*deg)) scattered onto the (2np)^3
// grid and transformed (r2c, FFTW layout).  Only the np^3 block at offset np/2
// is non-zero, so the three DFT passes run over that block.  One wave per
// node; each lane transforms whole lines, and since the twiddle exponent of
// input x and output k is ((x + np/2) k) mod n with x and k unrolled at compile
// time (NP is a template parameter), the n twiddles sit in registers.  Every
// output is summed over its inputs in ascending order, as the separable DFT of
// the reference layout does.
// ---------------------------------------------------------------------------
constexpr int kMaxNp = 8;

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// 3- and 6-point DFTs, X[k] = sum_n x[n] exp(SGN 2 pi i n k / N) (SGN = -1:
// forward, +1: inverse).  Three points: t1 = b + c, t2 = a - t1/2,
// r = SGN i (sqrt(3)/2)(b - c); X = (a + t1, t2 + r, t2 - r).  Six points: radix 2
// over two 3-point transforms of the even and odd samples with twiddles
// w6^k = exp(SGN 2 pi i k / 6), X[k] = E[k] + w6^k D[k], X[k+3] = E[k] - w6^k D[k].
template <int SGN>
__device__ inline void dft3(double2 a, double2 b, double2 c, double2& X0, double2& X1, double2& X2) {
  constexpr double kS = 0.86602540378443864676;
  const double2 t1 = cadd(b, c);
  const double2 t2 = make_double2(a.x - 0.5 * t1.x, a.y - 0.5 * t1.y);
  const double2 d = csub(b, c);
  const double2 r = make_double2(-SGN * kS * d.y, SGN * kS * d.x);
  X0 = cadd(a, t1);
  X1 = cadd(t2, r);
  X2 = csub(t2, r);
}
template <int SGN>
__device__ inline void dft6(const double2 (&x)[6], double2 (&X)[6]) {
  constexpr double kS = 0.86602540378443864676;
  double2 E[3], D[3];
  dft3<SGN>(x[0], x[2], x[4], E[0], E[1], E[2]);
  dft3<SGN>(x[1], x[3], x[5], D[0], D[1], D[2]);
  const double2 t1 = cmul(make_double2(0.5, SGN * kS), D[1]);
  const double2 t2 = cmul(make_double2(-0.5, SGN * kS), D[2]);
  X[0] = cadd(E[0], D[0]); X[3] = csub(E[0], D[0]);
  X[1] = cadd(E[1], t1);   X[4] = csub(E[1], t1);
  X[2] = cadd(E[2], t2);   X[5] = csub(E[2], t2);
}
// Mixed path, np = 6.  Forward line transform from the 6 block samples at grid
// positions b + 3 to the 12 frequencies: out[k] = sum_b in[b] w^((b+3) k); the
// even outputs are the 6-point DFT of the samples placed at (b + 3) mod 6, the
// odd outputs the same of in[b] w^(b+3).  w: the 12 forward twiddles.
__device__ inline void pad12_fwd(const double2 (&in)[6], const double2 (&w)[12], double2 (&out)[12]) {
  double2 e[6], o[6], E[6], D[6];
#pragma unroll
  for (int b = 0; b < 6; b++) {
    const int p = (b + 3) % 6;
    e[p] = in[b];
    o[p] = cmul(in[b], w[(b + 3) % 12]);
  }
  dft6<-1>(e, E);
  dft6<-1>(o, D);
#pragma unroll
  for (int m = 0; m < 6; m++) { out[2 * m] = E[m]; out[2 * m + 1] = D[m]; }
}
// Inverse line transform from 12 frequencies to the 6 block positions a + 3:
// out[a] = sum_k in[k] wi^((a+3) k) = E[j] + wi^(a+3) D[j], j = (a + 3) mod 6,
// with E, D the 6-point inverse DFTs of the even and odd frequencies.
__device__ inline void pad12_inv(const double2 (&in)[12], const double2 (&wi)[12], double2 (&out)[6]) {
  double2 e[6], o[6], E[6], D[6];
#pragma unroll
  for (int m = 0; m < 6; m++) { e[m] = in[2 * m]; o[m] = in[2 * m + 1]; }
  dft6<1>(e, E);
  dft6<1>(o, D);
#pragma unroll
  for (int a = 0; a < 6; a++) out[a] = cadd(E[(a + 3) % 6], cmul(wi[(a + 3) % 12], D[(a + 3) % 6]));
}

__device__ inline void store_c(double2* p, double2 v) { *p = v; }
__device__ inline void store_c(float2* p, double2 v) { *p = make_float2(float(v.x), float(v.y)); }
__device__ inline double2 load_c(const double2* p) { return *p; }
__device__ inline double2 load_c(const float2* p) { const float2 v = *p; return make_double2(v.x, v.y); }

// C2 = storage type of the spectra (double2, or float2 for the mixed path).
template <int NP, typename C2>
__global__ __launch_bounds__(64) void k_fft_fwd(VlistArgs v, C2* dspec) {
  constexpr int N = 2 * NP, HH = NP + 1, O = NP / 2;
  // mixed path at np = 6: the y and z line transforms by two 6-point DFTs
  // (the FP64 path keeps the ascending-order sums of the reference layout)
  constexpr bool kFast = NP == 6 && std::is_same<C2, float2>::value;
  // mixed path: packed half spectrum (spec_pack), the FP64 path the full r2c
  // layout (ka n + kb) h + kc
  constexpr bool kPack = std::is_same<C2, float2>::value;
  constexpr int HS = kPack ? spec_packed_len(NP) : N * N * HH;
  constexpr int NSAM = NP * NP * NP - (NP - 2) * (NP - 2) * (NP - 2);
  // mixed path: the line results between passes are kept in LDS as FP32 (the
  // spectra are stored in FP32 anyway), which halves the LDS of a workgroup
  using LT = typename std::conditional<std::is_same<C2, float2>::value, float2, double2>::type;
  __shared__ double xb[NP * NP * NP];
  __shared__ LT t1[NP * NP * HH];
  __shared__ LT t2[NP * N * HH];
  const int slot = blockIdx.x;  // right-hand side slot / nsrc, source slot % nsrc
  const int node = v.src_node[slot % v.nsrc];
  const int64_t urow = int64_t(node) + int64_t(slot / v.nsrc) * v.nstride;
  const int lane = threadIdx.x;
  double2 w[N];  // exp(-2 pi i m / n)
#pragma unroll
  for (int m = 0; m < N; m++) w[m] = v.twiddle[m];
  const int l = v.depth[node] + v.root_level;
  constexpr int R = (NSAM + 63) / 64;  // lane rounds over the surface samples
  for (int s = 0; s < v.sd; s++) {
    const double scale = ldexp(1.0, int(l * v.deg[s]));
    // all rounds' samples are loaded before the grid is cleared and filled
    double uv[R];
    int sh[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      // unconditional loads (a lane past the samples reads sample 0), so the
      // rounds' loads are issued together
      const int q = lane + 64 * r;
      const bool ok = q < NSAM;
      const int qq = ok ? q : 0;
      sh[r] = v.shell[qq];
      const double u = v.UE[urow * v.sz_ue + qq * v.sd + s];
      uv[r] = ok ? u * scale : 0.0;
    }
    __syncthreads();
    for (int e = lane; e < NP * NP * NP; e += 64) xb[e] = 0.0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++)
      if (lane + 64 * r < NSAM) xb[sh[r]] = uv[r];
    __syncthreads();
    // along x (c): real -> half spectrum; line (a, b)
    for (int ab = lane; ab < NP * NP; ab += 64) {
      double xv[NP];
#pragma unroll
      for (int c = 0; c < NP; c++) xv[c] = xb[ab * NP + c];
#pragma unroll
      for (int kc = 0; kc < HH; kc++) {
        double2 acc = make_double2(0, 0);
#pragma unroll
        for (int c = 0; c < NP; c++) {
          const double2 tw = w[((c + O) * kc) % N];
          acc.x += xv[c] * tw.x; acc.y += xv[c] * tw.y;
        }
        store_c(&t1[ab * HH + kc], acc);
      }
    }
    __syncthreads();
    // along y (b); line (a, kc)
    for (int e = lane; e < NP * HH; e += 64) {
      const int a = e / HH, kc = e % HH;
      double2 in[NP];
#pragma unroll
      for (int b = 0; b < NP; b++) in[b] = load_c(&t1[(a * NP + b) * HH + kc]);
      if constexpr (kFast) {
        double2 out[N];
        pad12_fwd(in, w, out);
#pragma unroll
        for (int kb = 0; kb < N; kb++) store_c(&t2[(a * N + kb) * HH + kc], out[kb]);
      } else {
#pragma unroll
        for (int kb = 0; kb < N; kb++) {
          double2 acc = make_double2(0, 0);
#pragma unroll
          for (int b = 0; b < NP; b++) acc = cadd(acc, cmul(in[b], w[((b + O) * kb) % N]));
          store_c(&t2[(a * N + kb) * HH + kc], acc);
        }
      }
    }
    __syncthreads();
    // along z (a) -> global, frequency (ka, kb, kc); line (kb, kc)
    C2* out = dspec + (int64_t(slot) * v.sd + s) * HS;
    auto zline = [&](int e, double2 (&o)[N]) {  // the n outputs of line e = kb h + kc
      double2 in[NP];
#pragma unroll
      for (int a = 0; a < NP; a++) in[a] = load_c(&t2[a * N * HH + e]);
      if constexpr (kFast) {
        pad12_fwd(in, w, o);
      } else {
#pragma unroll
        for (int ka = 0; ka < N; ka++) {
          double2 acc = make_double2(0, 0);
#pragma unroll
          for (int a = 0; a < NP; a++) acc = cadd(acc, cmul(in[a], w[((a + O) * ka) % N]));
          o[ka] = acc;
        }
      }
    };
    if constexpr (kPack) {
      // the lines of the planes kc in [1, np-1] first: line li = kb (np-1) +
      // kc - 1 stores frequency ka at li + ka NI; then the lines p = 2 kb +
      // (kc == np) of the planes kc = 0, np, which store ka in [0, np] without
      // the mirrors.  Every lane round is of one kind.
      constexpr int NI = N * (NP - 1);
      for (int li = lane; li < NI; li += 64) {
        double2 o[N];
        zline((li / (NP - 1)) * HH + li % (NP - 1) + 1, o);
#pragma unroll
        for (int ka = 0; ka < N; ka++) store_c(out + li + ka * NI, o[ka]);
      }
      for (int p = lane; p < 2 * N; p += 64) {
        const int e = (p >> 1) * HH + (p & 1) * NP;
        const SpecLine<NP> L(e);
        double2 o[N];
        zline(e, o);
#pragma unroll
        for (int ka = 0; ka <= NP; ka++)
          if (!L.plane_conj(ka)) store_c(out + L.plane_index(ka), o[ka]);
      }
      // the row's padding (the Hadamard sweeps it with the other frequencies)
      constexpr int NC = spec_packed_count(NP);
      for (int q = NC + lane; q < HS; q += 64) store_c(out + q, make_double2(0, 0));
    } else {
      for (int e = lane; e < N * HH; e += 64) {
        double2 o[N];
        zline(e, o);
#pragma unroll
        for (int ka = 0; ka < N; ka++) store_c(out + ka * N * HH + e, o[ka]);
      }
    }
  }
}

// Hadamard accumulation: one workgroup (one wave) per sibling group (the 8
// children of a parent) and 64 frequencies; lane = frequency, each lane holds
// the 8 targets' accumulators.  The host lists the group's sources (children
// of the parent's colleagues) in V-list (node) order with a packed word (halo
// index, offset slot of the halo index, 8-bit target membership), so every
// target sees its terms in V-list order, with the 1/G factor applied per term
// (UpwEqu2DwnChk_dgemv + cptwvv).
template <int SD, int TD>
__global__ __launch_bounds__(64) void k_hadamard(VlistArgs v, double nrm) {
  const int g = blockIdx.x / v.nrhs, rh = blockIdx.x % v.nrhs;  // virtual group g * nrhs + rh
  const int soff = rh * v.nsrc, toff = rh * v.ntrg;
  const int f = blockIdx.y * 64 + threadIdx.x;
  const int H = v.H;
  const bool ok = f < H;
  const int ff = ok ? f : 0;
  double2 acc[8][TD];
#pragma unroll
  for (int t = 0; t < 8; t++)
#pragma unroll
    for (int i = 0; i < TD; i++) acc[t][i] = make_double2(0, 0);
  const int cnt = v.pcnt[g];
  const int* ps = v.psrc + int64_t(g) * kPairStride;
  const int* pr = v.pslot + int64_t(g) * kPairStride;
  for (int k = 0; k < cnt; k++) {
    const int s = ps[k], word = pr[k];
    if (s < 0) continue;
    double2 d[SD];
#pragma unroll
    for (int j = 0; j < SD; j++) d[j] = v.dspec[(int64_t(s + soff) * SD + j) * H + ff];
    const int rs = (word >> 8) & 0x3ff, tm = word >> 20;
#pragma unroll
    for (int tau = 0; tau < 8; tau++) {
      if (!((tm >> tau) & 1)) continue;
      const int slot = rs - ((tau >> 2) * 49 + ((tau >> 1) & 1) * 7 + (tau & 1));
      const double2* T = v.m2l + int64_t(slot) * SD * TD * H;
#pragma unroll
      for (int i = 0; i < TD; i++)
#pragma unroll
        for (int j = 0; j < SD; j++) {
          const double2 p = cmul(T[int64_t(j * TD + i) * H + ff], d[j]);
          acc[tau][i] = make_double2(acc[tau][i].x + nrm * p.x, acc[tau][i].y + nrm * p.y);
        }
    }
  }
  if (!ok) return;
#pragma unroll
  for (int tau = 0; tau < 8; tau++) {
    const int t = v.tgt[g * 8 + tau];
    if (t < 0) continue;
#pragma unroll
    for (int i = 0; i < TD; i++) v.espec[(int64_t(t + toff) * TD + i) * H + f] = acc[tau][i];
  }
}

// Inverse: unnormalised c2r of each target spectrum, evaluated only on the
// surface samples, added into DC (effVal2PlnVal + regVal2SamVal).  Same line
// scheme as the forward transform; the first pass reads the spectrum straight
// from global memory (lines are contiguous across lanes).
template <int NP, typename C2>
__global__ __launch_bounds__(64) void k_fft_inv(VlistArgs v, const C2* espec) {
  constexpr int N = 2 * NP, HH = NP + 1, O = NP / 2;
  constexpr bool kFast = NP == 6 && std::is_same<C2, float2>::value;
  constexpr bool kPack = std::is_same<C2, float2>::value;
  constexpr int HS = kPack ? spec_packed_len(NP) : N * N * HH;
  constexpr int NSAM = NP * NP * NP - (NP - 2) * (NP - 2) * (NP - 2);
  using LT = typename std::conditional<std::is_same<C2, float2>::value, float2, double2>::type;
  __shared__ LT y1[NP * N * HH];
  __shared__ LT y2[NP * NP * HH];
  __shared__ int qof[NP * NP * NP];  // block index -> surface sample, -1 inside
  const int slot = blockIdx.x;  // right-hand side slot / ntrg, target slot % ntrg
  const int node = v.trg_node[slot % v.ntrg];
  const int64_t drow = int64_t(node) + int64_t(slot / v.ntrg) * v.nstride;
  const int lane = threadIdx.x;
  double2 w[N];  // exp(+2 pi i m / n)
#pragma unroll
  for (int m = 0; m < N; m++) { const double2 t = v.twiddle[m]; w[m] = make_double2(t.x, -t.y); }
  for (int e = lane; e < NP * NP * NP; e += 64) qof[e] = -1;
  __syncthreads();
  for (int q = lane; q < NSAM; q += 64) qof[v.shell[q]] = q;
  double* dc = v.DC + drow * v.sz_dc;
  const int lvl = v.depth[node] + v.root_level;
  for (int i = 0; i < v.td; i++) {
    const double oscale = v.deg_out ? ldexp(1.0, int(lvl * v.deg_out[i])) : 1.0;
    const C2* in = espec + (int64_t(slot) * v.td + i) * HS;
    __syncthreads();
    // along z (a), only block rows a' in [0,np); line (kb, kc)
    auto zline = [&](const double2 (&x)[N], int e) {  // line e = kb h + kc from its n inputs
      if constexpr (kFast) {
        double2 o6[NP];
        pad12_inv(x, w, o6);
#pragma unroll
        for (int a = 0; a < NP; a++) store_c(&y1[a * N * HH + e], o6[a]);
      } else {
#pragma unroll
        for (int a = 0; a < NP; a++) {
          double2 acc = make_double2(0, 0);
#pragma unroll
          for (int ka = 0; ka < N; ka++) acc = cadd(acc, cmul(x[ka], w[((a + O) * ka) % N]));
          store_c(&y1[a * N * HH + e], acc);
        }
      }
    };
    if constexpr (kPack) {
      // packed spectrum: the lines of the planes kc in [1, np-1] (li = kb (np-1)
      // + kc - 1 reads frequency ka at li + ka NI), then those of the planes
      // kc = 0, np (p = 2 kb + (kc == np), a value that is not stored is the
      // conjugate of its mirror), so each lane round is of one kind.  With one
      // round of each (np <= 6) both rounds' inputs are loaded before either is
      // transformed.
      constexpr int NI = N * (NP - 1);
      auto eline = [](int li) { return (li / (NP - 1)) * HH + li % (NP - 1) + 1; };
      auto eplane = [](int p) { return (p >> 1) * HH + (p & 1) * NP; };
      auto ldp = [&](const SpecLine<NP>& L, int ka) -> C2 {
        C2 t = in[L.plane_index(ka)];
        if (L.plane_conj(ka)) t.y = -t.y;
        return t;
      };
      if constexpr (NI <= 64 && 2 * N <= 64) {
        const int li = lane < NI ? lane : 0, p = lane < 2 * N ? lane : 0;
        const SpecLine<NP> L(eplane(p));
        C2 ri[N], rp[N];
#pragma unroll
        for (int ka = 0; ka < N; ka++) ri[ka] = in[li + ka * NI];
#pragma unroll
        for (int ka = 0; ka < N; ka++) rp[ka] = ldp(L, ka);
        double2 x[N];
        if (lane < NI) {
#pragma unroll
          for (int ka = 0; ka < N; ka++) x[ka] = load_c(&ri[ka]);
          zline(x, eline(li));
        }
        if (lane < 2 * N) {
#pragma unroll
          for (int ka = 0; ka < N; ka++) x[ka] = load_c(&rp[ka]);
          zline(x, eplane(p));
        }
      } else {
        for (int li = lane; li < NI; li += 64) {
          double2 x[N];
#pragma unroll
          for (int ka = 0; ka < N; ka++) x[ka] = load_c(in + li + ka * NI);
          zline(x, eline(li));
        }
        for (int p = lane; p < 2 * N; p += 64) {
          const SpecLine<NP> L(eplane(p));
          double2 x[N];
#pragma unroll
          for (int ka = 0; ka < N; ka++) {
            const C2 t = ldp(L, ka);
            x[ka] = load_c(&t);
          }
          zline(x, eplane(p));
        }
      }
    } else if constexpr (N * HH > 128) {
      for (int e = lane; e < N * HH; e += 64) {
        double2 x[N];
#pragma unroll
        for (int ka = 0; ka < N; ka++) x[ka] = load_c(in + ka * N * HH + e);
        zline(x, e);
      }
    } else {
      // the n h lines fit two rounds of the 64 lanes (np <= 7): both rounds'
      // inputs are loaded before either is transformed, so the load latency is
      // paid once per node
      const int e0 = lane, e1 = lane + 64;
      const bool has1 = e1 < N * HH;
      C2 r0[N], r1[N];
#pragma unroll
      for (int ka = 0; ka < N; ka++) r0[ka] = in[ka * N * HH + (e0 < N * HH ? e0 : 0)];
#pragma unroll
      for (int ka = 0; ka < N; ka++) r1[ka] = in[ka * N * HH + (has1 ? e1 : 0)];
#pragma unroll
      for (int half = 0; half < 2; half++) {
        const int e = half ? e1 : e0;
        if (e >= N * HH) continue;
        double2 x[N];
#pragma unroll
        for (int ka = 0; ka < N; ka++) x[ka] = load_c(half ? &r1[ka] : &r0[ka]);
        zline(x, e);
      }
    }
    __syncthreads();
    // along y (b), only b' in [0,np); line (a, kc)
    for (int e = lane; e < NP * HH; e += 64) {
      const int a = e / HH, kc = e % HH;
      double2 x[N];
#pragma unroll
      for (int kb = 0; kb < N; kb++) x[kb] = load_c(&y1[(a * N + kb) * HH + kc]);
      if constexpr (kFast) {
        double2 o6[NP];
        pad12_inv(x, w, o6);
#pragma unroll
        for (int b = 0; b < NP; b++) store_c(&y2[(a * NP + b) * HH + kc], o6[b]);
      } else {
#pragma unroll
        for (int b = 0; b < NP; b++) {
          double2 acc = make_double2(0, 0);
#pragma unroll
          for (int kb = 0; kb < N; kb++) acc = cadd(acc, cmul(x[kb], w[((b + O) * kb) % N]));
          store_c(&y2[(a * NP + b) * HH + kc], acc);
        }
      }
    }
    __syncthreads();
    // along x (c), c2r at the surface samples; line (a, b)
    for (int ab = lane; ab < NP * NP; ab += 64) {
      double2 yy[HH];
#pragma unroll
      for (int kc = 0; kc < HH; kc++) yy[kc] = load_c(&y2[ab * HH + kc]);
      // the line's check values are loaded together (absent samples read
      // element 0 and are not stored), so the loads are not serialised
      int qc[NP];
      double old[NP];
#pragma unroll
      for (int c = 0; c < NP; c++) qc[c] = qof[ab * NP + c];
      if (v.dc_store) {  // 0 + x: the same sums as reading the zeroed values
#pragma unroll
        for (int c = 0; c < NP; c++) old[c] = 0.0;
      } else {
#pragma unroll
        for (int c = 0; c < NP; c++) old[c] = dc[(qc[c] < 0 ? 0 : qc[c]) * v.td + i];
      }
#pragma unroll
      for (int c = 0; c < NP; c++) {
        double val = yy[0].x + yy[N / 2].x * (((c + O) & 1) ? -1.0 : 1.0);
#pragma unroll
        for (int kc = 1; kc < N / 2; kc++) {
          const double2 tw = w[((c + O) * kc) % N];
          val += 2.0 * (yy[kc].x * tw.x - yy[kc].y * tw.y);
        }
        if (qc[c] >= 0) dc[qc[c] * v.td + i] = old[c] + (v.deg_out ? oscale * val : val);
      }
    }
  }
}

template <typename C2>
static hipError_t launch_fft_fwd_any(const VlistArgs& v, C2* dspec, hipStream_t s) {
  const dim3 g(unsigned(int64_t(v.nsrc) * v.nrhs)), b(64);
  switch (v.np) {
    case 3: k_fft_fwd<3, C2><<<g, b, 0, s>>>(v, dspec); break;
    case 4: k_fft_fwd<4, C2><<<g, b, 0, s>>>(v, dspec); break;
    case 5: k_fft_fwd<5, C2><<<g, b, 0, s>>>(v, dspec); break;
    case 6: k_fft_fwd<6, C2><<<g, b, 0, s>>>(v, dspec); break;
    case 7: k_fft_fwd<7, C2><<<g, b, 0, s>>>(v, dspec); break;
    case 8: k_fft_fwd<8, C2><<<g, b, 0, s>>>(v, dspec); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename C2>
static hipError_t launch_fft_inv_any(const VlistArgs& v, const C2* espec, hipStream_t s) {
  const dim3 g(unsigned(int64_t(v.ntrg) * v.nrhs)), b(64);
  switch (v.np) {
    case 3: k_fft_inv<3, C2><<<g, b, 0, s>>>(v, espec); break;
    case 4: k_fft_inv<4, C2><<<g, b, 0, s>>>(v, espec); break;
    case 5: k_fft_inv<5, C2><<<g, b, 0, s>>>(v, espec); break;
    case 6: k_fft_inv<6, C2><<<g, b, 0, s>>>(v, espec); break;
    case 7: k_fft_inv<7, C2><<<g, b, 0, s>>>(v, espec); break;
    case 8: k_fft_inv<8, C2><<<g, b, 0, s>>>(v, espec); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_vlist(const VlistArgs& v, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  if (v.ngrp <= 0) return hipSuccess;
  if (v.np > kMaxNp || v.nrhs < 1) return hipErrorInvalidValue;
  hipError_t e = launch_fft_fwd_any<double2>(v, v.dspec, s);
  if (e != hipSuccess) return e;
  if (e0) (void)hipEventRecord(e0, s);
  const double nrm = 1.0 / double(v.n * v.n * v.n);
  dim3 grid(unsigned(int64_t(v.ngrp) * v.nrhs), unsigned((v.H + 63) / 64));
  if (v.sd == 1 && v.td == 1) k_hadamard<1, 1><<<grid, dim3(64), 0, s>>>(v, nrm);
  else if (v.sd == 4 && v.td == 3) k_hadamard<4, 3><<<grid, dim3(64), 0, s>>>(v, nrm);
  else if (v.sd == 3 && v.td == 3) k_hadamard<3, 3><<<grid, dim3(64), 0, s>>>(v, nrm);
  else if (v.sd == 3 && v.td == 4) k_hadamard<3, 4><<<grid, dim3(64), 0, s>>>(v, nrm);
  else return hipErrorInvalidValue;
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (e1) (void)hipEventRecord(e1, s);
  return launch_fft_inv_any<double2>(v, v.espec, s);
}

hipError_t launch_fft_fwd32(const VlistArgs& v, float2* dspec, hipStream_t s) {
  if (v.nsrc <= 0) return hipSuccess;
  if (v.np > kMaxNp || v.nrhs < 1) return hipErrorInvalidValue;
  return launch_fft_fwd_any<float2>(v, dspec, s);
}

hipError_t launch_fft_inv32(const VlistArgs& v, const float2* espec, hipStream_t s) {
  if (v.ntrg <= 0) return hipSuccess;
  if (v.np > kMaxNp || v.nrhs < 1) return hipErrorInvalidValue;
  return launch_fft_inv_any<float2>(v, espec, s);
}

// ---------------------------------------------------------------------------
// FP64 direct sum (Dense3d_MPI::evaluate, dense3d_mpi.cpp:75-141) at nt
// check targets over all sources.  Workgroup (x, y) takes 256 targets (one per
// thread) and source chunk y; the chunk's sources (leaf order, density read
// through the source permutation) are staged through LDS 256 at a time and
// every thread sweeps them with the exact kernel block (r < 1e-8 excluded).
// Each chunk writes its partial sums to part[y][target][TD]; k_direct_sum adds
// the chunks in chunk order (deterministic for a given chunk count).
template <int KT>
__global__ __launch_bounds__(256) void k_direct_part(KParams<double> p, int64_t ns, int64_t chunk,
                                                     const double* __restrict__ spos, const double* __restrict__ snor,
                                                     const int* __restrict__ sperm, const double* __restrict__ sden,
                                                     int nt, const double* __restrict__ tpos,
                                                     double* __restrict__ part) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  constexpr bool NOR = needs_normals_ct(KT);
  __shared__ double sx[256], sy[256], sz[256], snx[NOR ? 256 : 1], sny[NOR ? 256 : 1], snz[NOR ? 256 : 1];
  __shared__ double sdn[256 * SD];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * 256 + tid;
  const int ic = i < nt ? i : nt - 1;
  const double tx = tpos[3 * int64_t(ic)], ty = tpos[3 * int64_t(ic) + 1], tz = tpos[3 * int64_t(ic) + 2];
  const int64_t j0 = int64_t(blockIdx.y) * chunk, j1 = j0 + chunk < ns ? j0 + chunk : ns;
  double acc[TD];
#pragma unroll
  for (int t = 0; t < TD; t++) acc[t] = 0.0;
  for (int64_t b = j0; b < j1; b += 256) {
    const int64_t j = b + tid;
    __syncthreads();
    if (j < j1) {
      sx[tid] = spos[3 * j]; sy[tid] = spos[3 * j + 1]; sz[tid] = spos[3 * j + 2];
      if constexpr (NOR) { snx[tid] = snor[3 * j]; sny[tid] = snor[3 * j + 1]; snz[tid] = snor[3 * j + 2]; }
      const int64_t o = sperm[j];
#pragma unroll
      for (int q = 0; q < SD; q++) sdn[tid * SD + q] = sden[o * SD + q];
    } else {  // padding: far away, zero density
      sx[tid] = sy[tid] = sz[tid] = 1e18;
      if constexpr (NOR) { snx[tid] = sny[tid] = snz[tid] = 0.0; }
#pragma unroll
      for (int q = 0; q < SD; q++) sdn[tid * SD + q] = 0.0;
    }
    __syncthreads();
    const int cnt = j1 - b < 256 ? int(j1 - b) : 256;
    for (int k = 0; k < cnt; k++) {
      double blk[SD * TD], nx = 0.0, ny = 0.0, nz = 0.0;
      if constexpr (NOR) { nx = snx[k]; ny = sny[k]; nz = snz[k]; }
      kblock<KT, double>(p, tx - sx[k], ty - sy[k], tz - sz[k], nx, ny, nz, blk);
#pragma unroll
      for (int t = 0; t < TD; t++)
#pragma unroll
        for (int q = 0; q < SD; q++) acc[t] += blk[t * SD + q] * sdn[k * SD + q];
    }
  }
  if (i < nt) {
#pragma unroll
    for (int t = 0; t < TD; t++) part[(int64_t(blockIdx.y) * nt + i) * TD + t] = acc[t];
  }
}

__global__ void k_direct_sum(const double* __restrict__ part, int nchunk, int64_t n, double* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  for (int c = 0; c < nchunk; c++) v += part[int64_t(c) * n + i];
  out[i] = v;
}

// ---------------------------------------------------------------------------
#define KIFMM_USER_KERNELS(X) X(111) X(121) X(311) X(312) X(321) X(322) X(511) X(521)

hipError_t launch_near(int kt, const NearArgs& a0, const KParams<double>& pu, const KParams<double>& pm,
                       hipStream_t s) {
  if (a0.nchunks <= 0) return hipSuccess;
  if (a0.nrhs < 1) return hipErrorInvalidValue;
  for (int r0 = 0; r0 < a0.nrhs; r0 += kMaxNr) {
    NearArgs a = a0;
    a.r0 = r0;
    a.nr = std::min(kMaxNr, a0.nrhs - r0);
    const dim3 g(unsigned(a.nchunks)), b(kNearChunk);
    switch (kt * 16 + rhs_group(a.nr)) {
#define KIFMM_CASE_NR(K, R) case K * 16 + R: k_near<K, R><<<g, b, 0, s>>>(a, pu, pm); break;
#define KIFMM_CASE(K) KIFMM_CASE_NR(K, 1) KIFMM_CASE_NR(K, 2) KIFMM_CASE_NR(K, 4) KIFMM_CASE_NR(K, 8)
      KIFMM_USER_KERNELS(KIFMM_CASE)
#undef KIFMM_CASE
#undef KIFMM_CASE_NR
      default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_surface(int kt, const SurfArgs& a0, const KParams<double>& pu, hipStream_t s) {
  if (a0.ntasks <= 0) return hipSuccess;
  if (a0.n_sam > 2 * kSurfThreads || a0.nrhs < 1) return hipErrorInvalidValue;
  for (int r0 = 0; r0 < a0.nrhs; r0 += kMaxNr) {
    SurfArgs a = a0;
    a.r0 = r0;
    a.nr = std::min(kMaxNr, a0.nrhs - r0);
    const dim3 g(unsigned(a.ntasks)), b(kSurfThreads);
    switch (kt * 16 + rhs_group(a.nr)) {
#define KIFMM_CASE_NR(K, R) case K * 16 + R: k_surface<K, R><<<g, b, 0, s>>>(a, pu); break;
#define KIFMM_CASE(K) KIFMM_CASE_NR(K, 1) KIFMM_CASE_NR(K, 2) KIFMM_CASE_NR(K, 4) KIFMM_CASE_NR(K, 8)
      KIFMM_USER_KERNELS(KIFMM_CASE)
#undef KIFMM_CASE
#undef KIFMM_CASE_NR
      default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

DirectPlan plan_direct(int64_t ns, int nt, int td, int64_t cap_doubles) {
  DirectPlan d{};
  if (nt <= 0) return d;
  const int64_t per_chunk = int64_t(nt) * td;
  if (per_chunk > cap_doubles) return d;  // nchunk 0: the workspace cannot hold one chunk
  const int64_t tiles = (nt + 255) / 256;
  const int64_t blocks = std::max<int64_t>(1, (ns + 255) / 256);
  // about 8 workgroups per CU over 256 CUs, within the workspace
  int64_t nchunk = std::min<int64_t>({blocks, std::max<int64_t>(1, (2048 + tiles - 1) / tiles),
                                      cap_doubles / per_chunk, 65535});
  nchunk = std::max<int64_t>(1, nchunk);
  d.chunk = ((ns + nchunk - 1) / nchunk + 255) / 256 * 256;
  if (d.chunk == 0) d.chunk = 256;
  d.nchunk = int(std::max<int64_t>(1, (ns + d.chunk - 1) / d.chunk));
  d.work_doubles = int64_t(d.nchunk) * per_chunk;
  return d;
}

hipError_t launch_direct(int kt, const KParams<double>& p, int64_t ns, const double* spos, const double* snor,
                         const int* sperm, const double* sden, int nt, const double* tpos, const DirectPlan& d,
                         double* work, double* tval, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (d.nchunk <= 0 || !work || d.chunk <= 0 || int64_t(d.nchunk) * d.chunk < ns) return hipErrorInvalidValue;
  const dim3 grid(unsigned((nt + 255) / 256), unsigned(d.nchunk));
  switch (kt) {
#define KIFMM_CASE(K) \
    case K: k_direct_part<K><<<grid, dim3(256), 0, s>>>(p, ns, d.chunk, spos, snor, sperm, sden, nt, tpos, work); break;
    KIFMM_USER_KERNELS(KIFMM_CASE)
#undef KIFMM_CASE
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t n = int64_t(nt) * trg_dof(kt);
  k_direct_sum<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, s>>>(work, d.nchunk, n, tval);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fixed scratch for callers that cannot allocate: the direct sum's partial
// sums live in this array, one copy per GPU, held until the process ends.
// Every slot a call reads is written earlier in the same call, so nothing is
// cleared; calls that overlap on one GPU share it and race.
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t scratch_address(double** out) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    // the array's device address on the current GPU (the host-side name is
    // only a stand-in and must never reach a kernel)
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace kifmm
