// Translation operators of the kernel-independent FMM (the reference's
// MatMgnt3d_MPI, matmgnt3d_mpi.hpp/.cpp): surface and grid geometry, the
// truncated-SVD check-to-equivalent pseudo-inverses, the 8 M2M and 8 L2L
// matrices and the 316 FFT M2L spectra.  Homogeneous mm kernels only
// (Laplace 111, Stokes 301, Navier 511), so every table is built once at R = 1
// and scaled per level by powers of two.  Built in FP64 on the host, cached per
// (kernel, coefficients, np), uploaded once per handle.
#pragma once
#include <memory>
#include <vector>

namespace kifmm {

enum Surface { UE = 0, UC = 1, DE = 2, DC = 3 };

struct Operators {
  int kt_mm = 0, np = 0, sd = 0, td = 0;
  std::vector<double> coefs;
  double deg[4] = {0, 0, 0, 0};
  int n_sam[4] = {0, 0, 0, 0};          // points per surface
  int n = 0, G = 0, H = 0;              // grid edge 2np, grid size n^3, half spectrum (np+1)n^2
  std::vector<double> sam[4];           // surface samples at the reference radii (xyz interleaved)
  std::vector<int> shell;               // surface sample q -> regular-grid linear index
  std::vector<double> pinv_up;          // (n_ue*sd) x (n_uc*td), column-major
  std::vector<double> pinv_dn;          // (n_de*sd) x (n_dc*td)
  std::vector<double> m2m;              // 8 x (n_uc*td) x (n_ue*sd), child 4a+2b+c
  std::vector<double> l2l;              // 8 x (n_dc*td) x (n_de*sd)
  std::vector<double> m2l;              // 343 slots x (sd*td) x H complex (re,im), frequency fastest
  int rank_up = 0, rank_dn = 0;         // kept singular values
  int dat(int s) const { return n_sam[s] * ((s == UE || s == DE) ? sd : td); }
  static int slot(int i0, int i1, int i2) { return (i0 + 3) * 49 + (i1 + 3) * 7 + (i2 + 3); }
};

// Cached, thread-safe.  Returns nullptr (and sets err) for unsupported kernels.
std::shared_ptr<const Operators> get_operators(int kt_mm, const std::vector<double>& coefs, int np,
                                               std::string* err);

// The M2L spectra of op recomputed with the kernel grid's wrap entries (a
// coordinate index np: displacement -np, which no target of the block
// convolution reads) set to zero; same layout as Operators::m2l, and equal to it
// in what any evaluate computes.  For an even kernel the spectra of opposite
// offsets are then conjugates: table_mirror checks that, T(-s) = conj(T(s)) to
// 1e-12 of the table's largest value.
std::vector<double> m2l_table_nowrap(const Operators& op);
bool table_mirror(const std::vector<double>& t, int sdtd, int H);

// matmgnt3d_mpi.cpp:451-499
int sample_count(int np);
void sample_positions(int np, double R, double* pos);
void regular_positions(int np, double R, double* pos);

// Truncated pseudo-inverse: singular values below eps*sigma_0 dropped
// (common/vecmatop.cpp:129-164, common/svdrep.cpp:38-80).  A is m x n
// column-major, P is n x m column-major.  Returns the kept rank.
int pinv_truncated(int m, int n, const double* A, double eps, double* P);

// 3-D real DFT with FFTW's r2c layout on an n^3 grid, `howmany` interleaved
// components; output complex (re,im) at ((a*n+b)*(n/2+1)+c)*howmany+h.
void host_rfft3(int n, int howmany, const double* in, double* out);

}  // namespace kifmm
