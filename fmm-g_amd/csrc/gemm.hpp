// Indexed batched GEMM shared by the FP64 path (exact dgemv order: separate
// multiply and add, k ascending, accumulate steps seeded from C) and the mixed
// path (FMA).  C[:, cidx[j]] = (acc ? C : 0) + rs (.) (A_z * (ks (.) B[:, bidx[j]])).
// A_z is M x K column-major; z = blockIdx.z selects one of several matrices and
// its column range [zoff[z], zoff[z+1]) (the 8 L2L child matrices of a level run
// in one launch).  64x64 tile per workgroup, 4x4 outputs per lane, K staged
// through LDS in slabs of 16 with the next slab prefetched into registers.
#pragma once
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace kifmm {

constexpr int GT_M = 64, GT_N = 64, GT_K = 16;

template <bool FMA>
__global__ __launch_bounds__(256) void k_gemm_t(GemmArgs g) {
  __shared__ double As[GT_K][GT_M + 1];
  __shared__ double Bs[GT_K][GT_N + 1];
  __shared__ int64_t bcol[GT_N];
  const int z = blockIdx.z;
  const int cbeg = g.zoff ? g.zoff[z] : 0, cend = g.zoff ? g.zoff[z + 1] : g.ncol;
  const int n0 = cbeg + int(blockIdx.y) * GT_N;
  if (n0 >= cend) return;
  const double* A = g.A + int64_t(z) * g.a_zstride;
  const int m0 = blockIdx.x * GT_M;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  if (tid < GT_N) {
    const int j = n0 + tid;
    bcol[tid] = j < cend ? int64_t(g.bidx[j]) * g.ldb : int64_t(-1);
  }
  double c[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int n = n0 + tx + 16 * j;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int m = m0 + ty + 16 * i;
      c[i][j] = (!FMA && g.accumulate && n < cend && m < g.M) ? g.C[int64_t(g.cidx[n]) * g.ldc + m] : 0.0;
    }
  }
  __syncthreads();
  // each lane stages 4 A and 4 B elements per slab
  double ra[4], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = tid + 256 * u;
      const int mm = e % GT_M, kk = e / GT_M;
      const int m = m0 + mm, k = k0 + kk;
      ra[u] = (m < g.M && k < g.K) ? A[int64_t(k) * g.M + m] : 0.0;
      const int kb = e % GT_K, nn = e / GT_K;
      const int kq = k0 + kb;
      const int64_t col = bcol[nn];
      rb[u] = (col >= 0 && kq < g.K) ? g.B[col + kq] * g.kscale[kq % g.ksd] : 0.0;
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < g.K; k0 += GT_K) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = tid + 256 * u;
      As[e / GT_M][e % GT_M] = ra[u];
      Bs[e % GT_K][e / GT_K] = rb[u];
    }
    __syncthreads();
    if (k0 + GT_K < g.K) fetch(k0 + GT_K);
#pragma unroll
    for (int kk = 0; kk < GT_K; kk++) {
      double av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; i++) av[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if constexpr (FMA) c[i][j] = fma(av[i], bv[j], c[i][j]);
          else c[i][j] = c[i][j] + av[i] * bv[j];
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int n = n0 + tx + 16 * j;
    if (n >= cend) continue;
    double* col = g.C + int64_t(g.cidx[n]) * g.ldc;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int m = m0 + ty + 16 * i;
      if (m >= g.M) continue;
      if constexpr (FMA) {
        const double v = g.rscale[m % g.rsd] * c[i][j];
        col[m] = g.accumulate ? col[m] + v : v;
      } else {
        col[m] = g.accumulate ? c[i][j] : g.rscale[m % g.rsd] * c[i][j];
      }
    }
  }
}

template <bool FMA>
inline hipError_t launch_gemm_t(const GemmArgs& g, hipStream_t s) {
  if (g.ncol <= 0) return hipSuccess;
  const int nz = g.zoff ? g.nz : 1;
  dim3 grid(unsigned((g.M + GT_M - 1) / GT_M), unsigned((g.max_zcol + GT_N - 1) / GT_N), unsigned(nz));
  k_gemm_t<FMA><<<grid, dim3(256), 0, s>>>(g);
  return hipGetLastError();
}

}  // namespace kifmm
