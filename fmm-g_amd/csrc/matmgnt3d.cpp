// Translation-operator tables (see matmgnt3d.hpp).  Citations are to the
// reference's fmm3d_mpi/matmgnt3d_mpi.cpp unless stated otherwise.
#include "matmgnt3d.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>

#include "kernel3d.hpp"

namespace kifmm {

int sample_count(int np) { return np * np * np - (np - 2) * (np - 2) * (np - 2); }

// samPosCal :451-475 -- boundary points of an np^3 lattice on [-R,R]^3, i (x) outer.
void sample_positions(int np, double R, double* pos) {
  const double step = 2.0 / (np - 1);
  int q = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++) {
        bool edge = i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1;
        if (!edge) continue;
        double x = -1.0 + i * step, y = -1.0 + j * step, z = -1.0 + k * step;
        pos[3 * q] = R * x; pos[3 * q + 1] = R * y; pos[3 * q + 2] = R * z;
        q++;
      }
}

// regPosCal :478-499 -- (2np)^3 grid with wrap-around coordinates, x fastest.
void regular_positions(int np, double R, double* pos) {
  const int n = 2 * np;
  const double step = 2.0 / (np - 1);
  int q = 0;
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++, q++) {
        int g[3] = {i < np ? i : i - n, j < np ? j : j - n, k < np ? k : k - n};
        for (int d = 0; d < 3; d++) pos[3 * q + d] = R * g[d] * step;
      }
}

// ---------------------------------------------------------------------------
// One-sided (Hestenes) Jacobi SVD with a parallel round-robin pair schedule.
// ---------------------------------------------------------------------------
int pinv_truncated(int m, int n, const double* A, double eps, double* P) {
  const bool tr = m < n;          // work on the tall orientation
  const int p = tr ? n : m, q = tr ? m : n;
  std::vector<double> B(size_t(p) * q), V(size_t(q) * q, 0.0);
  for (int j = 0; j < q; j++)
    for (int i = 0; i < p; i++)
      B[size_t(j) * p + i] = tr ? A[size_t(i) * m + j] : A[size_t(j) * m + i];
  for (int j = 0; j < q; j++) V[size_t(j) * q + j] = 1.0;

  const int qq = q + (q & 1);
  const double tol = std::sqrt(double(p)) * DBL_EPSILON;
  for (int sweep = 0; sweep < 100; sweep++) {
    int rotated = 0;
    for (int r = 0; r < qq - 1; r++) {
#pragma omp parallel for schedule(static) reduction(+ : rotated)
      for (int k = 0; k < qq / 2; k++) {
        auto player = [&](int s) { return s == 0 ? 0 : ((s - 1 + r) % (qq - 1)) + 1; };
        int a = player(k), b = player(qq - 1 - k);
        if (a >= q || b >= q) continue;
        double* ba = &B[size_t(a) * p];
        double* bb = &B[size_t(b) * p];
        double al = 0, be = 0, ga = 0;
        for (int i = 0; i < p; i++) { al += ba[i] * ba[i]; be += bb[i] * bb[i]; ga += ba[i] * bb[i]; }
        if (al == 0.0 || be == 0.0 || std::fabs(ga) <= tol * std::sqrt(al * be)) continue;
        rotated++;
        double zeta = (be - al) / (2.0 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < p; i++) {
          double x = ba[i], y = bb[i];
          ba[i] = c * x - s * y; bb[i] = s * x + c * y;
        }
        double* va = &V[size_t(a) * q];
        double* vb = &V[size_t(b) * q];
        for (int i = 0; i < q; i++) {
          double x = va[i], y = vb[i];
          va[i] = c * x - s * y; vb[i] = s * x + c * y;
        }
      }
    }
    if (rotated == 0) break;
  }
  std::vector<double> sig(q);
  for (int j = 0; j < q; j++) {
    double s = 0;
    for (int i = 0; i < p; i++) s += B[size_t(j) * p + i] * B[size_t(j) * p + i];
    sig[j] = std::sqrt(s);
  }
  std::vector<int> ord(q);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return sig[x] > sig[y]; });
  const double cutoff = eps * sig[ord[0]];
  int rank = 0;
  while (rank < q && sig[ord[rank]] >= cutoff) rank++;  // svdrep.cpp:58-63
  // left singular vectors of the tall matrix: columns of B / sigma
  // pinv(A) (n x m):  not transposed -> V S^-1 Ub^T ;  transposed -> Ub S^-1 V^T
  std::memset(P, 0, sizeof(double) * size_t(n) * m);
#pragma omp parallel for schedule(static)
  for (int col = 0; col < m; col++) {
    for (int kk = 0; kk < rank; kk++) {
      int k = ord[kk];
      double inv = 1.0 / sig[k];
      const double* ub = &B[size_t(k) * p];
      const double* vk = &V[size_t(k) * q];
      if (!tr) {
        double f = ub[col] * inv * inv;  // ub = sigma*u
        for (int row = 0; row < n; row++) P[size_t(col) * n + row] += vk[row] * f;
      } else {
        double f = vk[col] * inv * inv;
        for (int row = 0; row < n; row++) P[size_t(col) * n + row] += ub[row] * f;
      }
    }
  }
  return rank;
}

// ---------------------------------------------------------------------------
// Host 3-D real DFT (table setup only): three passes of length-n complex DFTs
// through a precomputed twiddle matrix.
// ---------------------------------------------------------------------------
void host_rfft3(int n, int howmany, const double* in, double* out) {
  using cd = std::complex<double>;
  const int h = n / 2 + 1;
  std::vector<cd> w(n);
  for (int m = 0; m < n; m++) w[m] = std::polar(1.0, -2.0 * M_PI * m / n);
  std::vector<cd> a1(size_t(n) * n * h), a2(size_t(n) * n * h);
  for (int hm = 0; hm < howmany; hm++) {
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int kx = 0; kx < h; kx++) {
          cd s = 0;
          for (int x = 0; x < n; x++) s += in[((size_t(z) * n + y) * n + x) * howmany + hm] * w[(x * kx) % n];
          a1[(size_t(z) * n + y) * h + kx] = s;
        }
    for (int z = 0; z < n; z++)
      for (int ky = 0; ky < n; ky++)
        for (int kx = 0; kx < h; kx++) {
          cd s = 0;
          for (int y = 0; y < n; y++) s += a1[(size_t(z) * n + y) * h + kx] * w[(y * ky) % n];
          a2[(size_t(z) * n + ky) * h + kx] = s;
        }
    for (int kz = 0; kz < n; kz++)
      for (int ky = 0; ky < n; ky++)
        for (int kx = 0; kx < h; kx++) {
          cd s = 0;
          for (int z = 0; z < n; z++) s += a2[(size_t(z) * n + ky) * h + kx] * w[(z * kz) % n];
          size_t o = ((size_t(kz) * n + ky) * h + kx) * howmany + hm;
          out[2 * o] = s.real(); out[2 * o + 1] = s.imag();
        }
  }
}

// M2L spectra: for each offset slot s (|s|_inf >= 2) the r2c spectrum of
// K(grid point - source at 2 s) over the n^3 regular grid, [slot][sd*td][H]
// complex.  nowrap: the grid points with a coordinate index np (displacement
// -np, the only one without its mirror; no target of the block convolution
// reads it) hold zero, so that for an even kernel the spectra of opposite
// offsets are conjugates.
static void m2l_spectra(int kt, const double* cf, int np, int sd, int td, bool nowrap, std::vector<double>& out) {
  const int n = 2 * np, G = n * n * n, H = (np + 1) * n * n, sdtd = sd * td;
  std::vector<double> grid(3 * size_t(G));
  regular_positions(np, 1.0, grid.data());
  out.assign(size_t(343) * sdtd * H * 2, 0.0);
  std::vector<int> offs;
  for (int i0 = -3; i0 <= 3; i0++)
    for (int i1 = -3; i1 <= 3; i1++)
      for (int i2 = -3; i2 <= 3; i2++)
        if (std::max({std::abs(i0), std::abs(i1), std::abs(i2)}) > 1) offs.push_back(Operators::slot(i0, i1, i2));
#pragma omp parallel for schedule(dynamic)
  for (size_t u = 0; u < offs.size(); u++) {
    int s = offs[u];
    int i0 = s / 49 - 3, i1 = (s / 7) % 7 - 3, i2 = s % 7 - 3;
    double src[3] = {i0 * 2.0, i1 * 2.0, i2 * 2.0};
    std::vector<double> tt(size_t(G) * td * sd), tmp(size_t(G) * sdtd), spec(size_t(H) * sdtd * 2);
    build_kernel_matrix(kt, cf, 1, src, src, G, grid.data(), tt.data());
    for (int g = 0; g < G; g++) {
      const bool wrap = nowrap && (g % n == np || (g / n) % n == np || g / (n * n) == np);
      for (int i = 0; i < td; i++)
        for (int j = 0; j < sd; j++)
          tmp[size_t(g) * sdtd + j * td + i] = wrap ? 0.0 : tt[(size_t(g) * td + i) + size_t(j) * G * td];
    }
    host_rfft3(n, sdtd, tmp.data(), spec.data());
    double* dst = &out[size_t(s) * sdtd * H * 2];
    for (int f = 0; f < H; f++)
      for (int c = 0; c < sdtd; c++) {
        dst[(size_t(c) * H + f) * 2] = spec[(size_t(f) * sdtd + c) * 2];
        dst[(size_t(c) * H + f) * 2 + 1] = spec[(size_t(f) * sdtd + c) * 2 + 1];
      }
  }
}

std::vector<double> m2l_table_nowrap(const Operators& op) {
  std::vector<double> t;
  m2l_spectra(op.kt_mm, op.coefs.data(), op.np, op.sd, op.td, true, t);
  return t;
}

bool table_mirror(const std::vector<double>& t, int sdtd, int H) {
  const size_t per = size_t(sdtd) * H * 2;
  double big = 0.0, diff = 0.0;
  for (double x : t) big = std::max(big, std::fabs(x));
  for (int i0 = -3; i0 <= 3; i0++)
    for (int i1 = -3; i1 <= 3; i1++)
      for (int i2 = -3; i2 <= 3; i2++) {
        if (std::max({std::abs(i0), std::abs(i1), std::abs(i2)}) <= 1) continue;
        const double* a = t.data() + size_t(Operators::slot(i0, i1, i2)) * per;
        const double* b = t.data() + size_t(Operators::slot(-i0, -i1, -i2)) * per;
        for (size_t f = 0; f < per; f += 2) {
          diff = std::max(diff, std::fabs(a[f] - b[f]));
          diff = std::max(diff, std::fabs(a[f + 1] + b[f + 1]));
        }
      }
  return diff <= 1e-12 * big;
}

// ---------------------------------------------------------------------------
static std::shared_ptr<Operators> build_operators(int kt, const std::vector<double>& coefs, int np) {
  auto op = std::make_shared<Operators>();
  op->kt_mm = kt; op->np = np; op->coefs = coefs;
  op->sd = src_dof(kt); op->td = trg_dof(kt);
  const int sd = op->sd, td = op->td;
  homogeneous_deg(kt, op->deg);
  op->n_sam[UE] = op->n_sam[DE] = op->n_sam[DC] = sample_count(np);
  op->n_sam[UC] = sample_count(np + 2);
  // setup :63-68 -- UE (np, R=1), UC (np+2, R=3), DE (np, R=3), DC (np, R=1)
  const int npp[4] = {np, np + 2, np, np};
  const double rad[4] = {1.0, 3.0, 3.0, 1.0};
  for (int s = 0; s < 4; s++) {
    op->sam[s].resize(3 * op->n_sam[s]);
    sample_positions(npp[s], rad[s], op->sam[s].data());
  }
  op->n = 2 * np; op->G = op->n * op->n * op->n; op->H = (np + 1) * op->n * op->n;
  const int n = op->n;
  // samDen2RegDen :309-331 offsets
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++)
        if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1)
          op->shell.push_back((k + np / 2) * n * n + (j + np / 2) * n + (i + np / 2));

  const double* cf = coefs.data();
  const double eps = std::pow(0.1, np + 1);  // alt() :37-40
  const int nue = op->n_sam[UE], nuc = op->n_sam[UC], nde = op->n_sam[DE], ndc = op->n_sam[DC];
  {  // UpwChk2UpwEqu :107-126
    std::vector<double> M(size_t(nuc) * td * nue * sd);
    build_kernel_matrix(kt, cf, nue, op->sam[UE].data(), op->sam[UE].data(), nuc, op->sam[UC].data(), M.data());
    op->pinv_up.resize(M.size());
    op->rank_up = pinv_truncated(nuc * td, nue * sd, M.data(), eps, op->pinv_up.data());
  }
  {  // DwnChk2DwnEqu :196-210
    std::vector<double> M(size_t(ndc) * td * nde * sd);
    build_kernel_matrix(kt, cf, nde, op->sam[DE].data(), op->sam[DE].data(), ndc, op->sam[DC].data(), M.data());
    op->pinv_dn.resize(M.size());
    op->rank_dn = pinv_truncated(ndc * td, nde * sd, M.data(), eps, op->pinv_dn.data());
  }
  {  // UpwEqu2UpwChk :151-173 and DwnEqu2DwnChk :231-254, R = 1
    size_t mm = size_t(nuc) * td * nue * sd, ll = size_t(ndc) * td * nde * sd;
    op->m2m.resize(8 * mm);
    op->l2l.resize(8 * ll);
    std::vector<double> chk(3 * std::max(nuc, ndc)), den(3 * std::max(nue, nde));
    for (int c = 0; c < 8; c++) {
      int idx[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
      for (int q = 0; q < nuc; q++)
        for (int d = 0; d < 3; d++) chk[3 * q + d] = 2.0 * op->sam[UC][3 * q + d];
      for (int q = 0; q < nue; q++)
        for (int d = 0; d < 3; d++) den[3 * q + d] = op->sam[UE][3 * q + d] + (2 * idx[d] - 1);
      build_kernel_matrix(kt, cf, nue, den.data(), den.data(), nuc, chk.data(), &op->m2m[c * mm]);
      for (int q = 0; q < nde; q++)
        for (int d = 0; d < 3; d++) den[3 * q + d] = op->sam[DE][3 * q + d];
      for (int q = 0; q < ndc; q++)
        for (int d = 0; d < 3; d++) chk[3 * q + d] = 0.5 * op->sam[DC][3 * q + d] + (double(idx[d]) - 0.5);
      build_kernel_matrix(kt, cf, nde, den.data(), den.data(), ndc, chk.data(), &op->l2l[c * ll]);
    }
  }
  m2l_spectra(kt, cf, np, sd, td, false, op->m2l);  // UpwEqu2DwnChk table :383-410
  return op;
}

std::shared_ptr<const Operators> get_operators(int kt_mm, const std::vector<double>& coefs, int np,
                                               std::string* err) {
  static std::mutex mu;
  static std::vector<std::shared_ptr<Operators>> cache;  // getmmptr :526-543
  double deg[4];
  if (homogeneous_deg(kt_mm, deg) != src_dof(kt_mm) || src_dof(kt_mm) == 0) {
    if (err) *err = "kernel " + std::to_string(kt_mm) + " is not a supported equivalent-density kernel";
    return nullptr;
  }
  if (np < 3 || np > 16) {
    if (err) *err = "np must lie in [3,16]";
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(mu);
  for (auto& c : cache)
    if (c->kt_mm == kt_mm && c->np == np && c->coefs == coefs) return c;
  auto op = build_operators(kt_mm, coefs, np);
  cache.push_back(op);
  return op;
}

}  // namespace kifmm
