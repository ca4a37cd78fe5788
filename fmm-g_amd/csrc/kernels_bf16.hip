// Translations (M2M, L2L) of the mixed path as split-BF16 products on the
// matrix cores, FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// C[:, cidx[n]] (+)= rscale . (A (kscale . B[:, bidx[n]])) for the gathered
// columns of a level, as k_gemm_mfma computes it in FP64.  Both operands are
// split into three BF16 parts, x = x0 + x1 + x2 (each part the BF16 rounding of
// what the previous ones leave, so the three carry about 24 bits): A once on the
// host (its parts are tables in HBM), B while it is loaded (the level scale
// 2^(l deg) of the equivalent densities is applied in FP64 first, so every part
// is a scaled O(1) value).  The product keeps the terms A_i x_j with i + j <= t:
//   BF16x3 (t = 1): A0 x0 + A0 x1 + A1 x0, about 16 bits (relative 2^-16);
//   BF16x6 (t = 2): also A0 x2 + A1 x1 + A2 x0, about FP32 (2^-24).
// The smallest terms are accumulated first.
//
// Each 32-deep K slab is accumulated in FP32 by the instruction chain and then
// added to FP64 accumulators: the equivalent densities of the vector kernels
// cancel strongly in these products, and FP32 accumulation over the whole K
// (up to 3648 terms) cost a factor of 4-5 in accuracy (CPU emulation on
// physical densities: BF16x6 Navier np=6 2.0e-7 -> 4.8e-8, Laplace 8.7e-8 ->
// 1.9e-8 against FP64).
//
// Tile: a workgroup of 4 waves covers 160 rows and 64 columns; wave w owns the
// five 16-row blocks of row half w / 2 and two 16-column blocks (columns
// 32 (w % 2) .. + 32), so each A operand serves two products.  No LDS: a lane
// loads its 8 consecutive K values of a column (contiguous in HBM) and splits
// them in registers, and its A operands (8 consecutive K of one row, 16 bytes
// a part) straight from the tables, which L2 holds.  Operand layout of the
// 16x16x32 instruction: lane l supplies A[l%16][8(l/16) .. +7] and
// B[8(l/16) .. +7][l%16]; result register r holds D[4(l/16) + r][l%16].
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace kifmm {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int BM = 160, BMB = 5, BWAVES = 4, BCOL = 64;

__device__ inline void split3(double x, __bf16& a, __bf16& b, __bf16& c) {
  a = (__bf16)float(x);
  const double r1 = x - double(float(a));
  b = (__bf16)float(r1);
  const double r2 = r1 - double(float(b));
  c = (__bf16)float(r2);
}

template <int T>
__global__ __launch_bounds__(64 * BWAVES, 2) void k_gemm_bf16s(GemmArgs g, const bf8* __restrict__ A3, int Mp,
                                                                int Kp) {
  const int z = int(blockIdx.z);
  const int cbeg = g.zoff ? g.zoff[z] : 0, cend = g.zoff ? g.zoff[z + 1] : g.ncol;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int nw0 = cbeg + int(blockIdx.y) * BCOL + 32 * (wave & 1);  // first column of this wave
  if (nw0 >= cend) return;
  const int m0 = int(blockIdx.x) * BM + 80 * (wave >> 1);  // first row of this wave
  if (m0 >= Mp) return;
  const int mb = min(BMB, (Mp - m0) / 16);  // row blocks of this wave (Mp is a multiple of 16)
  // this lane's B columns, one per column block
  int64_t bc[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int n = nw0 + 16 * h + li;
    bc[h] = n < cend ? int64_t(g.bidx[n]) * g.ldb : int64_t(-1);
  }
  // A part p of matrix z: rows of 8-element K groups, parts [3][nz][Mp][Kp]
  const int64_t kq = Kp / 8;
  const bf8* Az = A3 + int64_t(z) * Mp * kq;
  const int64_t pstride = int64_t(g.zoff ? g.nz : 1) * Mp * kq;
  double acc[2][BMB][4];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int mi = 0; mi < BMB; mi++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[h][mi][r] = 0.0;
  for (int k0 = 0; k0 < Kp; k0 += 32) {
    bf8 x[2][3];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int k = k0 + 8 * lk + e;
        const double v = bc[h] >= 0 && k < g.K ? g.B[bc[h] + k] * g.kscale[k % g.ksd] : 0.0;
        __bf16 p0, p1, p2;
        split3(v, p0, p1, p2);
        x[h][0][e] = p0;
        x[h][1][e] = p1;
        x[h][2][e] = p2;
      }
    const int64_t kk = k0 / 8 + lk;
#pragma unroll
    for (int mi = 0; mi < BMB; mi++) {
      if (mi >= mb) continue;  // (uniform; continue keeps the loop unrolled)
      const int64_t row = int64_t(m0 + 16 * mi + li) * kq + kk;
      const bf8 a0 = Az[row], a1 = Az[pstride + row];
      f4 sl[2] = {f4{0.f, 0.f, 0.f, 0.f}, f4{0.f, 0.f, 0.f, 0.f}};
      if constexpr (T == 2) {
        const bf8 a2 = Az[2 * pstride + row];
#pragma unroll
        for (int h = 0; h < 2; h++) {
          sl[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, x[h][2], sl[h], 0, 0, 0);
          sl[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, x[h][1], sl[h], 0, 0, 0);
          sl[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, x[h][0], sl[h], 0, 0, 0);
        }
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        sl[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, x[h][1], sl[h], 0, 0, 0);
        sl[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, x[h][0], sl[h], 0, 0, 0);
        sl[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, x[h][0], sl[h], 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; r++) acc[h][mi][r] += double(sl[h][r]);
      }
    }
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int n = nw0 + 16 * h + li;
    if (n >= cend) continue;
    double* col = g.C + int64_t(g.cidx[n]) * g.ldc;
#pragma unroll
    for (int mi = 0; mi < BMB; mi++) {
      if (mi >= mb) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + 16 * mi + 4 * lk + r;
        if (m >= g.M) continue;
        const double v = g.rscale[m % g.rsd] * acc[h][mi][r];
        col[m] = g.accumulate ? col[m] + v : v;
      }
    }
  }
}

}  // namespace

hipError_t launch_gemm_bf16s(const GemmArgs& g, const void* a3, int Mp, int Kp, int terms, hipStream_t s) {
  if (g.ncol <= 0) return hipSuccess;
  if (!a3 || Mp % 16 || Kp % 32 || Mp < g.M || Kp < g.K || g.ksplit > 1) return hipErrorInvalidValue;
  const int nz = g.zoff ? g.nz : 1;
  const dim3 grid(unsigned((g.M + BM - 1) / BM), unsigned((g.max_zcol + BCOL - 1) / BCOL), unsigned(nz));
  const bf8* A = static_cast<const bf8*>(a3);
  if (terms == 3) k_gemm_bf16s<1><<<grid, dim3(64 * BWAVES), 0, s>>>(g, A, Mp, Kp);
  else if (terms == 6) k_gemm_bf16s<2><<<grid, dim3(64 * BWAVES), 0, s>>>(g, A, Mp, Kp);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace kifmm
