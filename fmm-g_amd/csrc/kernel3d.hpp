// Kernel library: the reference's Kernel3d_MPI kernel ids, DOF tables,
// homogeneity and closed forms (kernel3d_mpi.hpp:37-78, kernel3d_mpi.cpp:40-466),
// written once as __host__ __device__ templates so the host operator builder and
// the HIP interaction kernels evaluate the same expressions.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace kifmm {

enum KernelType {
  KNL_LAP_S_U = 111, KNL_LAP_D_U = 121, KNL_LAP_I = 191,
  KNL_STK_F_U = 301, KNL_STK_S_U = 311, KNL_STK_S_P = 312, KNL_STK_D_U = 321,
  KNL_STK_D_P = 322, KNL_STK_R_U = 331, KNL_STK_R_P = 332, KNL_STK_I = 391, KNL_STK_E = 392,
  KNL_NAV_S_U = 511, KNL_NAV_D_U = 521, KNL_NAV_R_U = 531, KNL_NAV_I = 591, KNL_NAV_E = 592,
  KNL_ERR = -1
};

// Pairs closer than this contribute nothing (kernel3d_mpi.cpp:35).
constexpr double kMinDif = 1e-8;

__host__ __device__ constexpr int src_dof(int kt) {
  return (kt == 111 || kt == 121 || kt == 191) ? 1
       : (kt == 301) ? 4
       : (kt == 311 || kt == 312 || kt == 321 || kt == 322 || kt == 331 || kt == 332 ||
          kt == 391 || kt == 392 || kt == 511 || kt == 521 || kt == 531 || kt == 591 ||
          kt == 592) ? 3 : 0;
}

__host__ __device__ constexpr int trg_dof(int kt) {
  return (kt == 111 || kt == 121 || kt == 191 || kt == 312 || kt == 322 || kt == 332) ? 1
       : (kt == 301 || kt == 311 || kt == 321 || kt == 331 || kt == 391 || kt == 392 ||
          kt == 511 || kt == 521 || kt == 531 || kt == 591 || kt == 592) ? 3 : 0;
}

// Kernel used for equivalent densities (fmm3d_setup_mpi.cpp:113-126); -1 = unsupported.
constexpr int mm_kernel(int kt) {
  return (kt == 111 || kt == 121 || kt == 312 || kt == 322) ? 111
       : (kt == 311 || kt == 321) ? 301
       : (kt == 511 || kt == 521) ? 511 : -1;
}

// Homogeneity degrees per source DOF (kernel3d_mpi.cpp:106-131); returns count.
inline int homogeneous_deg(int kt, double* deg) {
  switch (kt) {
    case 111: deg[0] = 1; return 1;
    case 301: deg[0] = deg[1] = deg[2] = 1; deg[3] = 2; return 4;
    case 511: deg[0] = deg[1] = deg[2] = 1; return 3;
    default: return 0;
  }
}

inline bool needs_normals(int kt) { return kt == 121 || kt == 321 || kt == 322 || kt == 521; }
__host__ __device__ constexpr bool needs_normals_ct(int kt) {
  return kt == 121 || kt == 321 || kt == 322 || kt == 521;
}

// Per-kernel constants, derived on the host in double from the coefficient
// vector (mu, nu) exactly as the reference derives them.
template <typename T>
struct KParams { T a, b, c; };

inline KParams<double> make_params(int kt, const double* coefs) {
  const double pi = M_PI;
  double mu = coefs ? coefs[0] : 1.0, ve = coefs ? coefs[1] : 0.25;
  KParams<double> p{0, 0, 0};
  switch (kt) {
    case 111: case 121: p.a = 1.0 / (4.0 * pi); break;
    case 301: { double oomu = 1.0 / mu; p.a = oomu * (1.0 / (8.0 * pi)); p.b = 1.0 / (4.0 * pi); break; }
    case 311: p.a = 1.0 / (8.0 * pi); p.b = 1.0 / mu; break;
    case 312: p.a = 1.0 / (4.0 * pi); break;
    case 321: p.a = 6.0 / (8.0 * pi); break;
    case 322: p.a = mu * (1.0 / (2.0 * pi)); break;
    case 331: case 531: p.a = 1.0 / (8.0 * pi * mu); break;
    case 511:
      p.a = (3.0 - 4.0 * ve) / (16.0 * pi * (1.0 - ve));
      p.b = 1.0 / (16.0 * pi * (1.0 - ve));
      p.c = 1.0 / mu;
      break;
    case 521:
      p.a = -(1 - 2.0 * ve) / (8.0 * pi * (1.0 - ve));
      p.b = (1 - 2.0 * ve) / (8.0 * pi * (1.0 - ve));
      p.c = -3.0 / (8.0 * pi * (1.0 - ve));
      break;
    default: break;
  }
  return p;
}

template <typename T> __host__ __device__ inline T ksqrt(T v) { return std::sqrt(v); }

// One (target, source) block, row-major TD x SD: blk[t*SD+s]; (x,y,z) = target - source;
// rn = (x,y,z).n, used by the double-layer kernels only.
template <int KT, typename T>
__host__ __device__ inline void kblock_rn(const KParams<T>& p, T x, T y, T z, T nx, T ny, T nz, T rn,
                                          T* blk) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  T r2 = x * x + y * y + z * z;
  T r = ksqrt(r2);
  bool sing = r < T(kMinDif);
  if constexpr (KT == 111) {
    blk[0] = sing ? T(0) : p.a / r;
  } else if constexpr (KT == 121) {
    T r3 = r2 * r;
    blk[0] = sing ? T(0) : -p.a / r3 * rn;
  } else if constexpr (KT == 191) {
    blk[0] = T(1);
  } else if constexpr (KT == 301) {
    T r3 = r2 * r;
    T G = p.a / r, H = p.a / r3, A = p.b / r3;
    if (sing) { G = H = A = T(0); }
    blk[0] = G + H * x * x; blk[1] = H * x * y;     blk[2] = H * x * z;      blk[3] = A * x;
    blk[4] = H * y * x;     blk[5] = G + H * y * y; blk[6] = H * y * z;      blk[7] = A * y;
    blk[8] = H * z * x;     blk[9] = H * z * y;     blk[10] = G + H * z * z; blk[11] = A * z;
  } else if constexpr (KT == 311) {
    T r3 = r2 * r;
    T G = p.a / r, H = p.a / r3, m = p.b;
    if (sing) { G = H = m = T(0); }
    blk[0] = m * (G + H * x * x); blk[1] = m * (H * x * y);     blk[2] = m * (H * x * z);
    blk[3] = m * (H * y * x);     blk[4] = m * (G + H * y * y); blk[5] = m * (H * y * z);
    blk[6] = m * (H * z * x);     blk[7] = m * (H * z * y);     blk[8] = m * (G + H * z * z);
  } else if constexpr (KT == 312) {
    T r3 = r2 * r;
    T f = sing ? T(0) : T(1);
    blk[0] = f * (p.a * x / r3); blk[1] = f * (p.a * y / r3); blk[2] = f * (p.a * z / r3);
    if (sing) { blk[0] = blk[1] = blk[2] = T(0); }
  } else if constexpr (KT == 321) {
    T r5 = r2 * r2 * r;
    T C = sing ? T(0) : -p.a / r5;
    blk[0] = C * rn * x * x; blk[1] = C * rn * x * y; blk[2] = C * rn * x * z;
    blk[3] = C * rn * y * x; blk[4] = C * rn * y * y; blk[5] = C * rn * y * z;
    blk[6] = C * rn * z * x; blk[7] = C * rn * z * y; blk[8] = C * rn * z * z;
  } else if constexpr (KT == 322) {
    T r3 = r2 * r, r5 = r3 * r2;
    blk[0] = p.a * (nx / r3 - 3 * rn * x / r5);
    blk[1] = p.a * (ny / r3 - 3 * rn * y / r5);
    blk[2] = p.a * (nz / r3 - 3 * rn * z / r5);
    if (sing) { blk[0] = blk[1] = blk[2] = T(0); }
  } else if constexpr (KT == 331 || KT == 531) {
    T r3 = r2 * r;
    T c = sing ? T(0) : p.a / r3;
    blk[0] = c * 0;    blk[1] = c * z;    blk[2] = c * (-y);
    blk[3] = c * (-z); blk[4] = c * 0;    blk[5] = c * x;
    blk[6] = c * y;    blk[7] = c * (-x); blk[8] = c * 0;
  } else if constexpr (KT == 332) {
    blk[0] = blk[1] = blk[2] = T(0);
  } else if constexpr (KT == 391 || KT == 591) {
    for (int i = 0; i < 9; i++) blk[i] = T(0);
    blk[0] = blk[4] = blk[8] = T(1);
  } else if constexpr (KT == 392 || KT == 592) {
    blk[0] = 0;  blk[1] = z;  blk[2] = -y;
    blk[3] = -z; blk[4] = 0;  blk[5] = x;
    blk[6] = y;  blk[7] = -x; blk[8] = 0;
  } else if constexpr (KT == 511) {
    T r3 = r2 * r;
    T G = p.a / r, H = p.b / r3, m = p.c;
    if (sing) { G = H = m = T(0); }
    blk[0] = m * (G + H * x * x); blk[1] = m * (H * x * y);     blk[2] = m * (H * x * z);
    blk[3] = m * (H * y * x);     blk[4] = m * (G + H * y * y); blk[5] = m * (H * y * z);
    blk[6] = m * (H * z * x);     blk[7] = m * (H * z * y);     blk[8] = m * (G + H * z * z);
  } else if constexpr (KT == 521) {
    T r3 = r2 * r, r5 = r3 * r2;
    T A = p.a / r3, B = p.b / r3, C = p.c / r5;
    if (sing) { A = B = C = T(0); rn = T(0); }
    blk[0] = A * (rn + nx * x) + B * (x * nx) + C * rn * x * x;
    blk[1] = A * (nx * y) + B * (x * ny) + C * rn * x * y;
    blk[2] = A * (nx * z) + B * (x * nz) + C * rn * x * z;
    blk[3] = A * (ny * x) + B * (y * nx) + C * rn * y * x;
    blk[4] = A * (rn + ny * y) + B * (y * ny) + C * rn * y * y;
    blk[5] = A * (ny * z) + B * (y * nz) + C * rn * y * z;
    blk[6] = A * (nz * x) + B * (z * nx) + C * rn * z * x;
    blk[7] = A * (nz * y) + B * (z * ny) + C * rn * z * y;
    blk[8] = A * (rn + nz * z) + B * (z * nz) + C * rn * z * z;
  } else {
    for (int i = 0; i < SD * TD; i++) blk[i] = T(0);
  }
}

template <int KT, typename T>
__host__ __device__ inline void kblock(const KParams<T>& p, T x, T y, T z, T nx, T ny, T nz, T* blk) {
  kblock_rn<KT, T>(p, x, y, z, nx, ny, nz, x * nx + y * ny + z * nz, blk);
}

// Runtime dispatch to the compile-time kernel id.
#define KIFMM_FOR_EACH_KERNEL(X) \
  X(111) X(121) X(191) X(301) X(311) X(312) X(321) X(322) X(331) X(332) X(391) X(392) \
  X(511) X(521) X(531) X(591) X(592)

inline bool kernel_known(int kt) { return src_dof(kt) > 0; }

// Dense (nt*TD) x (ns*SD) interaction matrix, column-major (kernel3d_mpi.cpp:142-177).
// Host only; used to build the translation tables.
int build_kernel_matrix(int kt, const double* coefs, int ns, const double* spos,
                        const double* snor, int nt, const double* tpos, double* mat);

}  // namespace kifmm
