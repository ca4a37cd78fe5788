/*
 * KIFMM3d CPU oracle -- TEST INFRASTRUCTURE ONLY (see kifmm_oracle.h).
 *
 * Plain C99 restatement of the reference CPU evaluate path on one rank.  Each
 * function cites the reference file:line it follows (paths relative to the
 * reference root).  It is deliberately written like the reference: every box
 * interaction materialises the dense kernel matrix and applies it as a GEMV;
 * M2L uses the FFT-on-regular-grid scheme with a scalar Hadamard product.
 *
 * Parity status of each part (what pins it) is documented in the tests that
 * exercise it: kernels against closed-form identities, the FFT against numpy,
 * the FFT-M2L against the dense M2L product, the tree/lists against brute-force
 * invariants, and the whole evaluate against the direct sum.
 */
#include "kifmm_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* kernel3d_mpi.cpp:35 -- pairs closer than this contribute nothing */
static const double ORC_MINDIF = 1e-8;

/* ======================================================================
 * Kernels: kernel3d_mpi.cpp:40-131 (DOF / homogeneity tables)
 * ====================================================================== */
int orc_src_dof(int kt) {
  switch (kt) {
    case 111: case 121: case 191: return 1;
    case 301: return 4;
    case 311: case 312: case 321: case 322: case 331: case 332: case 391: case 392: return 3;
    case 511: case 521: case 531: case 591: case 592: return 3;
    default: return 0;
  }
}

int orc_trg_dof(int kt) {
  switch (kt) {
    case 111: case 121: case 191: return 1;
    case 301: case 311: case 321: case 331: case 391: case 392: return 3;
    case 312: case 322: case 332: return 1;
    case 511: case 521: case 531: case 591: case 592: return 3;
    default: return 0;
  }
}

/* fmm3d_setup_mpi.cpp:113-126: kernel used for equivalent densities */
int orc_mm_kernel(int kt) {
  switch (kt) {
    case 111: case 121: return 111;
    case 311: case 321: return 301;
    case 312: case 322: return 111;
    case 511: case 521: return 511;
    default: return -1;
  }
}

/* kernel3d_mpi.cpp:122-131 */
int orc_hom_deg(int kt, double *deg) {
  switch (kt) {
    case 111: deg[0] = 1; return 1;
    case 301: deg[0] = 1; deg[1] = 1; deg[2] = 1; deg[3] = 2; return 4;
    case 511: deg[0] = 1; deg[1] = 1; deg[2] = 1; return 3;
    default: return 0;
  }
}

/* One (target, source) block, row-major td x sd: blk[t*sd+s].
 * (x,y,z) = target - source.  Formulas restate kernel3d_mpi.cpp:142-466. */
static void kblock(int kt, const double *c, double x, double y, double z,
                   const double *n, double *blk) {
  double r2 = x * x + y * y + z * z;
  double r = sqrt(r2);
  int sd = orc_src_dof(kt), td = orc_trg_dof(kt);
  int sing = (r < ORC_MINDIF);
  switch (kt) {
    case 111: { /* :152-167 */
      double OOFP = 1.0 / (4.0 * M_PI);
      blk[0] = sing ? 0.0 : OOFP / r;
      return;
    }
    case 121: { /* :169-186 */
      double OOFP = 1.0 / (4.0 * M_PI);
      if (sing) { blk[0] = 0.0; return; }
      double rn = x * n[0] + y * n[1] + z * n[2];
      double r3 = r2 * r;
      blk[0] = -OOFP / r3 * rn;
      return;
    }
    case 191: blk[0] = 1; return; /* :188-193 */
    case 301: { /* :195-219 */
      double mu = c[0];
      double OOFP = 1.0 / (4.0 * M_PI), OOEP = 1.0 / (8.0 * M_PI), oomu = 1.0 / mu;
      if (sing) { memset(blk, 0, sizeof(double) * 12); return; }
      double r3 = r2 * r;
      double G = oomu * OOEP / r, H = oomu * OOEP / r3, A = OOFP / r3;
      blk[0] = G + H * x * x; blk[1] = H * x * y;     blk[2] = H * x * z;     blk[3] = A * x;
      blk[4] = H * y * x;     blk[5] = G + H * y * y; blk[6] = H * y * z;     blk[7] = A * y;
      blk[8] = H * z * x;     blk[9] = H * z * y;     blk[10] = G + H * z * z; blk[11] = A * z;
      return;
    }
    case 311: { /* :221-243 */
      double mu = c[0];
      double OOEP = 1.0 / (8.0 * M_PI), oomu = 1.0 / mu;
      if (sing) { memset(blk, 0, sizeof(double) * 9); return; }
      double r3 = r2 * r;
      double G = OOEP / r, H = OOEP / r3;
      blk[0] = oomu * (G + H * x * x); blk[1] = oomu * (H * x * y);     blk[2] = oomu * (H * x * z);
      blk[3] = oomu * (H * y * x);     blk[4] = oomu * (G + H * y * y); blk[5] = oomu * (H * y * z);
      blk[6] = oomu * (H * z * x);     blk[7] = oomu * (H * z * y);     blk[8] = oomu * (G + H * z * z);
      return;
    }
    case 312: { /* :245-261 */
      double OOFP = 1.0 / (4.0 * M_PI);
      if (sing) { blk[0] = blk[1] = blk[2] = 0.0; return; }
      double r3 = r2 * r;
      blk[0] = OOFP * x / r3; blk[1] = OOFP * y / r3; blk[2] = OOFP * z / r3;
      return;
    }
    case 321: { /* :263-283 */
      double SOEP = 6.0 / (8.0 * M_PI);
      if (sing) { memset(blk, 0, sizeof(double) * 9); return; }
      double rn = x * n[0] + y * n[1] + z * n[2];
      double r5 = r2 * r2 * r;
      double C = -SOEP / r5;
      blk[0] = C * rn * x * x; blk[1] = C * rn * x * y; blk[2] = C * rn * x * z;
      blk[3] = C * rn * y * x; blk[4] = C * rn * y * y; blk[5] = C * rn * y * z;
      blk[6] = C * rn * z * x; blk[7] = C * rn * z * y; blk[8] = C * rn * z * z;
      return;
    }
    case 322: { /* :285-307 */
      double mu = c[0];
      double coef = mu * (1.0 / (2.0 * M_PI));
      if (sing) { blk[0] = blk[1] = blk[2] = 0.0; return; }
      double rn = x * n[0] + y * n[1] + z * n[2];
      double r3 = r2 * r, r5 = r3 * r2;
      blk[0] = coef * (n[0] / r3 - 3 * rn * x / r5);
      blk[1] = coef * (n[1] / r3 - 3 * rn * y / r5);
      blk[2] = coef * (n[2] / r3 - 3 * rn * z / r5);
      return;
    }
    case 331: case 531: { /* :308-324, :438-454 */
      double mu = c[0];
      if (sing) { memset(blk, 0, sizeof(double) * 9); return; }
      double r3 = r2 * r;
      double coef = 1.0 / (8.0 * M_PI * mu) / r3;
      blk[0] = coef * 0;    blk[1] = coef * z;    blk[2] = coef * (-y);
      blk[3] = coef * (-z); blk[4] = coef * 0;    blk[5] = coef * x;
      blk[6] = coef * y;    blk[7] = coef * (-x); blk[8] = coef * 0;
      return;
    }
    case 332: blk[0] = blk[1] = blk[2] = 0; return; /* :325-334 */
    case 391: case 591: /* :335-345 */
      memset(blk, 0, sizeof(double) * 9); blk[0] = blk[4] = blk[8] = 1; return;
    case 392: case 592: /* :347-358 */
      blk[0] = 0;  blk[1] = z;  blk[2] = -y;
      blk[3] = -z; blk[4] = 0;  blk[5] = x;
      blk[6] = y;  blk[7] = -x; blk[8] = 0;
      return;
    case 511: { /* :360-385 */
      double mu = c[0], ve = c[1];
      double sc1 = (3.0 - 4.0 * ve) / (16.0 * M_PI * (1.0 - ve));
      double sc2 = 1.0 / (16.0 * M_PI * (1.0 - ve));
      double oomu = 1.0 / mu;
      if (sing) { memset(blk, 0, sizeof(double) * 9); return; }
      double r3 = r2 * r;
      double G = sc1 / r, H = sc2 / r3;
      blk[0] = oomu * (G + H * x * x); blk[1] = oomu * (H * x * y);     blk[2] = oomu * (H * x * z);
      blk[3] = oomu * (H * y * x);     blk[4] = oomu * (G + H * y * y); blk[5] = oomu * (H * y * z);
      blk[6] = oomu * (H * z * x);     blk[7] = oomu * (H * z * y);     blk[8] = oomu * (G + H * z * z);
      return;
    }
    case 521: { /* :387-436 */
      double ve = c[1];
      double dc1 = -(1 - 2.0 * ve) / (8.0 * M_PI * (1.0 - ve));
      double dc2 = (1 - 2.0 * ve) / (8.0 * M_PI * (1.0 - ve));
      double dc3 = -3.0 / (8.0 * M_PI * (1.0 - ve));
      if (sing) { memset(blk, 0, sizeof(double) * 9); return; }
      double nx = n[0], ny = n[1], nz = n[2];
      double rn = x * nx + y * ny + z * nz;
      double r3 = r2 * r, r5 = r3 * r2;
      double A = dc1 / r3, B = dc2 / r3, C = dc3 / r5;
      blk[0] = A * (rn + nx * x) + B * (x * nx) + C * rn * x * x;
      blk[1] = A * (nx * y) + B * (x * ny) + C * rn * x * y;
      blk[2] = A * (nx * z) + B * (x * nz) + C * rn * x * z;
      blk[3] = A * (ny * x) + B * (y * nx) + C * rn * y * x;
      blk[4] = A * (rn + ny * y) + B * (y * ny) + C * rn * y * y;
      blk[5] = A * (ny * z) + B * (y * nz) + C * rn * y * z;
      blk[6] = A * (nz * x) + B * (z * nx) + C * rn * z * x;
      blk[7] = A * (nz * y) + B * (z * ny) + C * rn * z * y;
      blk[8] = A * (rn + nz * z) + B * (z * nz) + C * rn * z * z;
      return;
    }
    default:
      for (int i = 0; i < sd * td; i++) blk[i] = 0.0;
      return;
  }
}

/* kernel3d_mpi.cpp:142-177 buildKnlIntCtx: dense (nt*td) x (ns*sd), column-major */
int orc_kernel_matrix(int kt, const double *coefs, int ns, const double *spos,
                      const double *snor, int nt, const double *tpos, double *mat) {
  int sd = orc_src_dof(kt), td = orc_trg_dof(kt);
  if (sd == 0) return -1;
  int64_t M = (int64_t)nt * td;
  double blk[16];
  for (int j = 0; j < ns; j++)
    for (int i = 0; i < nt; i++) {
      kblock(kt, coefs, tpos[3 * i] - spos[3 * j], tpos[3 * i + 1] - spos[3 * j + 1],
             tpos[3 * i + 2] - spos[3 * j + 2], snor + 3 * j, blk);
      for (int t = 0; t < td; t++)
        for (int s = 0; s < sd; s++)
          mat[(int64_t)(i * td + t) + (int64_t)(j * sd + s) * M] = blk[t * sd + s];
    }
  return 0;
}

/* y += A x, A column-major m x n (reference BLAS dgemv 'N' column order) */
static void gemv_acc(int m, int n, const double *A, const double *x, double *y) {
  for (int j = 0; j < n; j++) {
    double xj = x[j];
    const double *a = A + (int64_t)j * m;
    for (int i = 0; i < m; i++) y[i] += a[i] * xj;
  }
}

/* ======================================================================
 * Geometry: matmgnt3d_mpi.cpp:451-499
 * ====================================================================== */
int orc_sample_count(int np) { return np * np * np - (np - 2) * (np - 2) * (np - 2); }

void orc_sample_pos(int np, double R, double *pos) { /* samPosCal :451-475 */
  double step = 2.0 / (np - 1), init = -1.0;
  int cnt = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++)
        if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1) {
          double x = init + i * step, y = init + j * step, z = init + k * step;
          pos[3 * cnt] = R * x; pos[3 * cnt + 1] = R * y; pos[3 * cnt + 2] = R * z;
          cnt++;
        }
}

void orc_regular_pos(int np, double R, double *pos) { /* regPosCal :478-499 */
  int n = 2 * np;
  double step = 2.0 / (np - 1);
  int cnt = 0;
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        int gi = (i < np) ? i : i - n, gj = (j < np) ? j : j - n, gk = (k < np) ? k : k - n;
        pos[3 * cnt] = R * gi * step; pos[3 * cnt + 1] = R * gj * step; pos[3 * cnt + 2] = R * gk * step;
        cnt++;
      }
}

/* locPos :434-448 */
static void loc_pos(int n, const double *bas, const double ctr[3], double rad, double *pos) {
  for (int j = 0; j < n; j++)
    for (int d = 0; d < 3; d++) pos[3 * j + d] = ctr[d] + rad * bas[3 * j + d];
}

/* ======================================================================
 * 3-D real DFT with FFTW's r2c / c2r conventions (plan_many_dft_r2c/_c2r with
 * dims {n,n,n}, istride=howmany, idist=1; matmgnt3d_mpi.cpp:302,341).  Index of
 * real sample (a,b,c) is (a*n+b)*n+c with c fastest; the half spectrum runs along
 * c.  The c2r is unnormalised.  Written as separable direct DFTs.
 * ====================================================================== */
static void twiddles(int n, double *cr, double *ci) {
  for (int m = 0; m < n; m++) { cr[m] = cos(2.0 * M_PI * m / n); ci[m] = -sin(2.0 * M_PI * m / n); }
}

void orc_rfft3(int n, int howmany, const double *in, double *out) {
  int h = n / 2 + 1;
  double cr[64], ci[64];
  twiddles(n, cr, ci);
  size_t nc = (size_t)n * n * h;
  double *t1 = (double *)calloc(2 * nc, sizeof(double));
  double *t2 = (double *)calloc(2 * nc, sizeof(double));
  for (int hm = 0; hm < howmany; hm++) {
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++)
        for (int kc = 0; kc < h; kc++) {
          double sr = 0, si = 0;
          for (int c = 0; c < n; c++) {
            double v = in[(((size_t)a * n + b) * n + c) * howmany + hm];
            int m = (c * kc) % n;
            sr += v * cr[m]; si += v * ci[m];
          }
          size_t o = ((size_t)a * n + b) * h + kc;
          t1[2 * o] = sr; t1[2 * o + 1] = si;
        }
    for (int a = 0; a < n; a++)
      for (int kb = 0; kb < n; kb++)
        for (int kc = 0; kc < h; kc++) {
          double sr = 0, si = 0;
          for (int b = 0; b < n; b++) {
            size_t o = ((size_t)a * n + b) * h + kc;
            int m = (b * kb) % n;
            sr += t1[2 * o] * cr[m] - t1[2 * o + 1] * ci[m];
            si += t1[2 * o] * ci[m] + t1[2 * o + 1] * cr[m];
          }
          size_t o = ((size_t)a * n + kb) * h + kc;
          t2[2 * o] = sr; t2[2 * o + 1] = si;
        }
    for (int ka = 0; ka < n; ka++)
      for (int kb = 0; kb < n; kb++)
        for (int kc = 0; kc < h; kc++) {
          double sr = 0, si = 0;
          for (int a = 0; a < n; a++) {
            size_t o = ((size_t)a * n + kb) * h + kc;
            int m = (a * ka) % n;
            sr += t2[2 * o] * cr[m] - t2[2 * o + 1] * ci[m];
            si += t2[2 * o] * ci[m] + t2[2 * o + 1] * cr[m];
          }
          size_t o = (((size_t)ka * n + kb) * h + kc) * howmany + hm;
          out[2 * o] = sr; out[2 * o + 1] = si;
        }
  }
  free(t1); free(t2);
}

void orc_irfft3(int n, int howmany, const double *in, double *out) {
  int h = n / 2 + 1;
  double cr[64], ci[64];
  twiddles(n, cr, ci); /* inverse uses conj: e^{+i} = cr - i*ci */
  size_t nc = (size_t)n * n * h;
  double *t1 = (double *)calloc(2 * nc, sizeof(double));
  double *t2 = (double *)calloc(2 * nc, sizeof(double));
  for (int hm = 0; hm < howmany; hm++) {
    for (int a = 0; a < n; a++)
      for (int kb = 0; kb < n; kb++)
        for (int kc = 0; kc < h; kc++) {
          double sr = 0, si = 0;
          for (int ka = 0; ka < n; ka++) {
            size_t o = (((size_t)ka * n + kb) * h + kc) * howmany + hm;
            int m = (a * ka) % n;
            double xr = in[2 * o], xi = in[2 * o + 1];
            sr += xr * cr[m] + xi * ci[m];
            si += xi * cr[m] - xr * ci[m];
          }
          size_t o = ((size_t)a * n + kb) * h + kc;
          t1[2 * o] = sr; t1[2 * o + 1] = si;
        }
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++)
        for (int kc = 0; kc < h; kc++) {
          double sr = 0, si = 0;
          for (int kb = 0; kb < n; kb++) {
            size_t o = ((size_t)a * n + kb) * h + kc;
            int m = (b * kb) % n;
            double xr = t1[2 * o], xi = t1[2 * o + 1];
            sr += xr * cr[m] + xi * ci[m];
            si += xi * cr[m] - xr * ci[m];
          }
          size_t o = ((size_t)a * n + b) * h + kc;
          t2[2 * o] = sr; t2[2 * o + 1] = si;
        }
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++)
        for (int c = 0; c < n; c++) {
          size_t base = ((size_t)a * n + b) * h;
          double s = t2[2 * base];                                   /* kc = 0 */
          s += t2[2 * (base + n / 2)] * ((c & 1) ? -1.0 : 1.0);       /* kc = n/2 */
          for (int kc = 1; kc < n / 2; kc++) {
            int m = (c * kc) % n;
            double yr = t2[2 * (base + kc)], yi = t2[2 * (base + kc) + 1];
            s += 2.0 * (yr * cr[m] + yi * ci[m]);
          }
          out[(((size_t)a * n + b) * n + c) * howmany + hm] = s;
        }
  }
  free(t1); free(t2);
}

/* samDen2RegDen :309-331 (adds into reg) and regVal2SamVal :349-371 (adds into sam) */
static void sam2reg(int np, int dof, const double *sam, double *reg) {
  int n = 2 * np, cnt = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++)
        if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1) {
          int off = (k + np / 2) * n * n + (j + np / 2) * n + (i + np / 2);
          for (int f = 0; f < dof; f++) reg[dof * off + f] += sam[dof * cnt + f];
          cnt++;
        }
}
static void reg2sam(int np, int dof, const double *reg, double *sam) {
  int n = 2 * np, cnt = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++)
        if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1) {
          int off = (k + np / 2) * n * n + (j + np / 2) * n + (i + np / 2);
          for (int f = 0; f < dof; f++) sam[dof * cnt + f] += reg[dof * off + f];
          cnt++;
        }
}

/* ======================================================================
 * Operators: matmgnt3d_mpi.cpp:52-71 (surfaces), :107-274 (pinv, M2M, L2L),
 * :373-431 (M2L spectra).  Homogeneous mm kernels only (R = 1 tables).
 * ====================================================================== */
static double *sample_tab(int np, int which) { /* setup :63-68 */
  int npp = (which == ORC_UC) ? np + 2 : np;
  double R = (which == ORC_UE || which == ORC_DC) ? 1.0 : 3.0;
  double *p = (double *)malloc(sizeof(double) * 3 * orc_sample_count(npp));
  orc_sample_pos(npp, R, p);
  return p;
}

/* the matrix that pinv inverts: which=0 -> UC<-UE (:114-126), which=1 -> DC<-DE (:201-210) */
int orc_check_matrix(int kt_mm, const double *coefs, int ncoefs, int np, int which, double *mat) {
  (void)ncoefs;
  double *chk = sample_tab(np, which == 0 ? ORC_UC : ORC_DC);
  double *den = sample_tab(np, which == 0 ? ORC_UE : ORC_DE);
  int nchk = orc_sample_count(which == 0 ? np + 2 : np), nden = orc_sample_count(np);
  /* R = 1: daxpy(R, samPos, 0) leaves the samples unchanged */
  int rc = orc_kernel_matrix(kt_mm, coefs, nden, den, den, nchk, chk, mat);
  free(chk); free(den);
  return rc;
}

int orc_ops_init(orc_ops *op, int kt_mm, const double *coefs, int ncoefs, int np,
                 const double *pinv_up, const double *pinv_dn) {
  memset(op, 0, sizeof(*op));
  op->kt_mm = kt_mm; op->np = np;
  op->sd = orc_src_dof(kt_mm); op->td = orc_trg_dof(kt_mm);
  if (op->sd == 0 || orc_hom_deg(kt_mm, op->deg) != op->sd) return -1;
  op->ncoefs = ncoefs < 4 ? ncoefs : 4;
  for (int i = 0; i < op->ncoefs; i++) op->coefs[i] = coefs[i];
  int sd = op->sd, td = op->td;
  op->n_ue = op->n_de = op->n_dc = orc_sample_count(np);
  op->n_uc = orc_sample_count(np + 2);
  size_t up = (size_t)op->n_ue * sd * op->n_uc * td, dn = (size_t)op->n_de * sd * op->n_dc * td;
  op->pinv_up = (double *)malloc(sizeof(double) * up);
  op->pinv_dn = (double *)malloc(sizeof(double) * dn);
  memcpy(op->pinv_up, pinv_up, sizeof(double) * up);
  memcpy(op->pinv_dn, pinv_dn, sizeof(double) * dn);

  double *sUE = sample_tab(np, ORC_UE), *sUC = sample_tab(np, ORC_UC);
  double *sDE = sample_tab(np, ORC_DE), *sDC = sample_tab(np, ORC_DC);
  double R = 1.0;
  /* M2M :151-173 -- chk = 2R*UC, den = R*UE + (2*idx-1)*R */
  size_t m2m_sz = (size_t)op->n_uc * td * op->n_ue * sd;
  op->m2m = (double *)malloc(sizeof(double) * 8 * m2m_sz);
  double *chk = (double *)malloc(sizeof(double) * 3 * op->n_uc);
  double *den = (double *)malloc(sizeof(double) * 3 * op->n_uc);
  for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++) {
    int idx[3] = {a, b, c};
    for (int j = 0; j < op->n_uc; j++)
      for (int d = 0; d < 3; d++) chk[3 * j + d] = 0.0 + 2.0 * R * sUC[3 * j + d];
    for (int j = 0; j < op->n_ue; j++)
      for (int d = 0; d < 3; d++) {
        double v = 0.0 + R * sUE[3 * j + d];
        den[3 * j + d] = v + (2 * idx[d] - 1) * R;
      }
    orc_kernel_matrix(kt_mm, op->coefs, op->n_ue, den, den, op->n_uc, chk,
                      op->m2m + (size_t)(4 * a + 2 * b + c) * m2m_sz);
  }
  /* L2L :231-254 -- den = R*DE, chk = 0.5R*DC + (idx-0.5)*R */
  size_t l2l_sz = (size_t)op->n_dc * td * op->n_de * sd;
  op->l2l = (double *)malloc(sizeof(double) * 8 * l2l_sz);
  for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++) {
    int idx[3] = {a, b, c};
    for (int j = 0; j < op->n_de; j++)
      for (int d = 0; d < 3; d++) den[3 * j + d] = 0.0 + R * sDE[3 * j + d];
    for (int j = 0; j < op->n_dc; j++)
      for (int d = 0; d < 3; d++) {
        double v = 0.0 + 0.5 * R * sDC[3 * j + d];
        chk[3 * j + d] = v + ((double)idx[d] - 0.5) * R;
      }
    orc_kernel_matrix(kt_mm, op->coefs, op->n_de, den, den, op->n_dc, chk,
                      op->l2l + (size_t)(4 * a + 2 * b + c) * l2l_sz);
  }
  free(chk); free(den);
  /* M2L spectra :373-410 */
  int n = 2 * np, G = n * n * n, H = (np + 1) * n * n, sdtd = sd * td;
  double *reg = (double *)malloc(sizeof(double) * 3 * G);
  orc_regular_pos(np, 1.0, reg);
  double *rs = (double *)malloc(sizeof(double) * 3 * G);
  for (int j = 0; j < 3 * G; j++) rs[j] = 0.0 + R * reg[j];
  double *tt = (double *)malloc(sizeof(double) * (size_t)G * td * sd);
  double *tmp = (double *)malloc(sizeof(double) * (size_t)G * td * sd);
  op->m2l = (double *)calloc((size_t)343 * sdtd * H * 2, sizeof(double));
  for (int i0 = -3; i0 <= 3; i0++) for (int i1 = -3; i1 <= 3; i1++) for (int i2 = -3; i2 <= 3; i2++) {
    int lin = abs(i0); if (abs(i1) > lin) lin = abs(i1); if (abs(i2) > lin) lin = abs(i2);
    if (lin <= 1) continue;
    double dp[3] = {i0 * 2.0 * R, i1 * 2.0 * R, i2 * 2.0 * R};
    orc_kernel_matrix(kt_mm, op->coefs, 1, dp, dp, G, rs, tt); /* (G*td) x sd */
    for (int k = 0; k < G; k++)
      for (int i = 0; i < td; i++)
        for (int j = 0; j < sd; j++)
          tmp[i + td * (j + k * sd)] = tt[(i + k * td) + (size_t)j * G * td];
    int slot = (i0 + 3) * 49 + (i1 + 3) * 7 + (i2 + 3);
    orc_rfft3(n, sdtd, tmp, op->m2l + (size_t)slot * sdtd * H * 2);
  }
  free(reg); free(rs); free(tt); free(tmp);
  free(sUE); free(sUC); free(sDE); free(sDC);
  return 0;
}

void orc_ops_free(orc_ops *op) {
  free(op->pinv_up); free(op->pinv_dn); free(op->m2m); free(op->l2l); free(op->m2l);
  memset(op, 0, sizeof(*op));
}

/* ======================================================================
 * Tree and lists.
 *
 * Leaves: restates Dendro 1.0 ot::points2Octree as called at let3d_mpi.cpp:268
 * (external, not in the reference tree; rule chosen and documented): points map
 * to their depth-maxLevel cell; a node splits iff it holds more than ptsMax
 * distinct cells and its depth is below maxLevel; the root always splits
 * (let3d_mpi.cpp:283-288).  Node numbering follows CreateAncestors
 * (let3d_mpi.cpp:611-726): root is 0, its children 1..8, and child blocks are
 * appended when the first leaf beneath them is reached in Dendro's Morton order
 * (z most significant), i.e. a pre-order walk in that order.  Child index is
 * 4x+2y+z (let3d_mpi.cpp:1800-1806).
 * ====================================================================== */
typedef unsigned __int128 u128;

typedef struct {
  int n, cap;
  int *parent, *child, *depth, *path;
} nodevec;

static int nv_push(nodevec *v, int par, int d, int px, int py, int pz) {
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 1024;
    v->parent = (int *)realloc(v->parent, sizeof(int) * v->cap);
    v->child = (int *)realloc(v->child, sizeof(int) * v->cap);
    v->depth = (int *)realloc(v->depth, sizeof(int) * v->cap);
    v->path = (int *)realloc(v->path, sizeof(int) * 3 * v->cap);
  }
  int i = v->n++;
  v->parent[i] = par; v->child[i] = -1; v->depth[i] = d;
  v->path[3 * i] = px; v->path[3 * i + 1] = py; v->path[3 * i + 2] = pz;
  return i;
}

static void nv_split(nodevec *v, int p) {
  int first = v->n;
  v->child[p] = first;
  for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++)
    nv_push(v, p, v->depth[p] + 1, 2 * v->path[3 * p] + a, 2 * v->path[3 * p + 1] + b,
            2 * v->path[3 * p + 2] + c);
}

static int64_t lower_u128(const u128 *k, int64_t lo, int64_t hi, u128 key) {
  while (lo < hi) { int64_t m = lo + (hi - lo) / 2; if (k[m] < key) lo = m + 1; else hi = m; }
  return lo;
}

typedef struct { nodevec *v; const u128 *keys; int L, pts_max; } refine_ctx;

static void refine(refine_ctx *cx, int p, u128 base, int64_t lo, int64_t hi) {
  int d = cx->v->depth[p] + 1;
  u128 span = ((u128)1) << (3 * (cx->L - d));
  for (int k = 0; k < 8; k++) { /* Dendro Morton order: bits (z,y,x) */
    int x = k & 1, y = (k >> 1) & 1, z = (k >> 2) & 1;
    int c = cx->v->child[p] + 4 * x + 2 * y + z;
    u128 cb = base + (u128)k * span;
    int64_t clo = lower_u128(cx->keys, lo, hi, cb);
    int64_t chi = lower_u128(cx->keys, clo, hi, cb + span);
    if (d < cx->L && chi - clo > cx->pts_max) {
      nv_split(cx->v, c);
      refine(cx, c, cb, clo, chi);
    }
  }
}

static int cmp_u128(const void *a, const void *b) {
  u128 x = *(const u128 *)a, y = *(const u128 *)b;
  return (x < y) ? -1 : (x > y);
}

static u128 morton(uint32_t x, uint32_t y, uint32_t z, int L) {
  u128 k = 0;
  for (int b = L - 1; b >= 0; b--)
    k = (k << 3) | (u128)((((z >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((x >> b) & 1u));
  return k;
}

static int is_leaf(const orc_tree *t, int i) { return t->child[i] < 0; }

/* findGlbNode :1833-1853 */
static int find_node(const orc_tree *t, int want_depth, const int *wpath) {
  int cur = 0;
  int tp[3] = {wpath[0], wpath[1], wpath[2]};
  while (t->depth[cur] < want_depth && !is_leaf(t, cur)) {
    int dif = want_depth - t->depth[cur];
    int tmp = 1 << (dif - 1);
    int ch[3];
    for (int d = 0; d < 3; d++) { ch[d] = tp[d] / tmp; tp[d] -= ch[d] * tmp; }
    cur = t->child[cur] + ch[0] * 4 + ch[1] * 2 + ch[2];
  }
  return cur;
}

/* adjacent :1856-1876 */
static int adjacent(const orc_tree *t, int me, int you) {
  int dm = t->depth[me], dy = t->depth[you];
  int md = dm > dy ? dm : dy;
  int64_t mr = (int64_t)1 << (md - dm), yr = (int64_t)1 << (md - dy);
  int64_t rad = mr + yr, mx = 0;
  int ok = 1;
  for (int d = 0; d < 3; d++) {
    int64_t mc = (2 * (int64_t)t->path[3 * me + d] + 1) * mr;
    int64_t yc = (2 * (int64_t)t->path[3 * you + d] + 1) * yr;
    int64_t df = mc > yc ? mc - yc : yc - mc;
    if (df > rad) ok = 0;
    if (df > mx) mx = df;
  }
  return ok && mx == rad;
}

typedef struct { int *a; int n, cap; } ivec;
static void iv_push(ivec *v, int x) {
  if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 64; v->a = (int *)realloc(v->a, sizeof(int) * v->cap); }
  v->a[v->n++] = x;
}
static int cmp_int(const void *a, const void *b) { int x = *(const int *)a, y = *(const int *)b; return (x > y) - (x < y); }
static void iv_sort_unique(ivec *v) {
  if (v->n < 2) return;
  qsort(v->a, v->n, sizeof(int), cmp_int);
  int m = 1;
  for (int i = 1; i < v->n; i++) if (v->a[i] != v->a[m - 1]) v->a[m++] = v->a[i];
  v->n = m;
}

/* calGlbNodeLists :1637-1768 */
static void node_lists(const orc_tree *t, int B, ivec *U, ivec *V, ivec *W, ivec *X, ivec *queue) {
  U->n = V->n = W->n = X->n = 0;
  if (B != 0) {
    int P = t->parent[B], D = t->depth[B];
    int maxIdx = 1 << D;
    for (int i = -2; i < 4; i++) for (int j = -2; j < 4; j++) for (int k = -2; k < 4; k++) {
      int tp[3] = {2 * t->path[3 * P] + i, 2 * t->path[3 * P + 1] + j, 2 * t->path[3 * P + 2] + k};
      int inside = 1, same = 1;
      for (int d = 0; d < 3; d++) {
        if (tp[d] < 0 || tp[d] >= maxIdx) inside = 0;
        if (tp[d] != t->path[3 * B + d]) same = 0;
      }
      if (!inside || same) continue;
      int R = find_node(t, D, tp);
      int adj = adjacent(t, R, B);
      if (t->depth[R] < D) {
        if (adj) { if (is_leaf(t, B)) iv_push(U, R); }
        else iv_push(X, R);
      }
      if (t->depth[R] == D) {
        if (!adj) iv_push(V, R);
        else if (is_leaf(t, B)) {
          queue->n = 0; iv_push(queue, R);
          for (int q = 0; q < queue->n; q++) {
            int F = queue->a[q];
            if (!adjacent(t, F, B)) iv_push(W, F);
            else if (is_leaf(t, F)) iv_push(U, F);
            else for (int c = 0; c < 8; c++) iv_push(queue, t->child[F] + c);
          }
        }
      }
    }
  }
  if (is_leaf(t, B)) iv_push(U, B);
  iv_sort_unique(U); iv_sort_unique(V); iv_sort_unique(W); iv_sort_unique(X);
  /* filters :1748-1765 */
  int m;
  m = 0; for (int i = 0; i < U->n; i++) if (t->src_num[U->a[i]] > 0) U->a[m++] = U->a[i]; U->n = m;
  m = 0; for (int i = 0; i < V->n; i++) if (t->tag[V->a[i]] & ORC_SRCE) V->a[m++] = V->a[i]; V->n = m;
  m = 0; for (int i = 0; i < W->n; i++) if (t->tag[W->a[i]] & ORC_SRCE) W->a[m++] = W->a[i]; W->n = m;
  m = 0; for (int i = 0; i < X->n; i++) if (t->src_num[X->a[i]] > 0) X->a[m++] = X->a[i]; X->n = m;
}

/* descend by x >= centre comparisons (srcData :1255-1267, center :1808-1818) */
static int locate(const orc_tree *t, const double *p, const double ll[3], double rad) {
  int node = 0;
  while (!is_leaf(t, node)) {
    int tmp = 1 << t->depth[node];
    int idx[3];
    for (int d = 0; d < 3; d++) {
      double c = ll[d] + (2 * rad) * (t->path[3 * node + d] + 0.5) / (double)tmp;
      idx[d] = (p[d] >= c);
    }
    node = t->child[node] + idx[0] * 4 + idx[1] * 2 + idx[2];
  }
  return node;
}

int orc_tree_build(orc_tree *t, int n_src, const double *spos, int n_trg, const double *tpos,
                   const double ctr[3], int root_level, int pts_max, int max_level) {
  memset(t, 0, sizeof(*t));
  if (max_level < 1 || max_level > 30 || pts_max < 1) return -1;
  int L = max_level;
  double rad = pow(2.0, -root_level);
  double ll[3] = {ctr[0] - rad, ctr[1] - rad, ctr[2] - rad};
  /* depth-L cells, deduplicated */
  u128 *keys = (u128 *)malloc(sizeof(u128) * (n_src > 0 ? n_src : 1));
  double scale = ldexp(1.0, L - 1 + root_level); /* cells per unit length */
  uint32_t cmax = (1u << L) - 1u;
  for (int i = 0; i < n_src; i++) {
    uint32_t c[3];
    for (int d = 0; d < 3; d++) {
      double v = (spos[3 * i + d] - ll[d]) * scale;
      c[d] = (v <= 0.0) ? 0u : (v >= (double)cmax ? cmax : (uint32_t)v);
    }
    keys[i] = morton(c[0], c[1], c[2], L);
  }
  qsort(keys, n_src, sizeof(u128), cmp_u128);
  int64_t nu = 0;
  for (int i = 0; i < n_src; i++) if (nu == 0 || keys[i] != keys[nu - 1]) keys[nu++] = keys[i];

  nodevec v; memset(&v, 0, sizeof(v));
  nv_push(&v, -1, 0, 0, 0, 0);
  nv_split(&v, 0);
  refine_ctx cx = {&v, keys, L, pts_max};
  refine(&cx, 0, (u128)0, 0, nu);
  free(keys);

  int N = v.n;
  t->n_nodes = N; t->max_level = L; t->n_src = n_src; t->n_trg = n_trg;
  t->parent = v.parent; t->child = v.child; t->depth = v.depth; t->path = v.path;
  t->tag = (int *)calloc(N, sizeof(int));
  t->src_cnt = (int *)calloc(N, sizeof(int));
  t->trg_cnt = (int *)calloc(N, sizeof(int));
  t->src_beg = (int *)malloc(sizeof(int) * N); t->src_num = (int *)malloc(sizeof(int) * N);
  t->trg_beg = (int *)malloc(sizeof(int) * N); t->trg_num = (int *)malloc(sizeof(int) * N);

  /* srcData :1209-1465 (single rank) */
  int *leaf_of = (int *)malloc(sizeof(int) * (n_src > n_trg ? (n_src > 0 ? n_src : 1) : (n_trg > 0 ? n_trg : 1)));
  for (int i = 0; i < n_src; i++) {
    int node = locate(t, spos + 3 * i, ll, rad);
    leaf_of[i] = node;
    for (int q = node; q >= 0; q = t->parent[q]) t->src_cnt[q]++;
  }
  int off = 0;
  for (int i = 0; i < N; i++) {
    t->src_beg[i] = -1; t->src_num[i] = -1;
    if (i >= 1 && (!is_leaf(t, i) || t->src_cnt[i] > 0)) {
      t->tag[i] |= ORC_SRCE | ORC_OWNR;
      if (is_leaf(t, i)) { t->src_beg[i] = off; t->src_num[i] = t->src_cnt[i]; off += t->src_cnt[i]; }
    }
    if (i >= 1 && t->src_cnt[i] > 0) t->tag[i] |= ORC_CBTR;
  }
  t->src_perm = (int *)malloc(sizeof(int) * (n_src > 0 ? n_src : 1));
  int *fill = (int *)calloc(N, sizeof(int));
  for (int i = 0; i < n_src; i++) { int l = leaf_of[i]; t->src_perm[t->src_beg[l] + fill[l]++] = i; }

  /* trgData :1470-1631 */
  for (int i = 0; i < n_trg; i++) {
    int node = locate(t, tpos + 3 * i, ll, rad);
    leaf_of[i] = node;
    for (int q = node; q >= 0; q = t->parent[q]) t->trg_cnt[q]++;
  }
  off = 0;
  for (int i = 0; i < N; i++) {
    t->trg_beg[i] = -1; t->trg_num[i] = -1;
    if (t->trg_cnt[i] > 0) {
      t->tag[i] |= ORC_EVTR;
      if (is_leaf(t, i)) { t->trg_beg[i] = off; t->trg_num[i] = t->trg_cnt[i]; off += t->trg_cnt[i]; }
    }
  }
  t->trg_perm = (int *)malloc(sizeof(int) * (n_trg > 0 ? n_trg : 1));
  memset(fill, 0, sizeof(int) * N);
  for (int i = 0; i < n_trg; i++) { int l = leaf_of[i]; t->trg_perm[t->trg_beg[l] + fill[l]++] = i; }
  free(fill); free(leaf_of);

  /* lists for every node holding targets */
  ivec U = {0}, V = {0}, W = {0}, X = {0}, Q = {0};
  ivec Ua = {0}, Va = {0}, Wa = {0}, Xa = {0};
  t->u_off = (int64_t *)calloc(N + 1, sizeof(int64_t));
  t->v_off = (int64_t *)calloc(N + 1, sizeof(int64_t));
  t->w_off = (int64_t *)calloc(N + 1, sizeof(int64_t));
  t->x_off = (int64_t *)calloc(N + 1, sizeof(int64_t));
  for (int i = 0; i < N; i++) {
    if (t->trg_cnt[i] > 0) {
      node_lists(t, i, &U, &V, &W, &X, &Q);
      for (int k = 0; k < U.n; k++) { iv_push(&Ua, U.a[k]); t->tag[U.a[k]] |= ORC_USER; }
      for (int k = 0; k < V.n; k++) { iv_push(&Va, V.a[k]); t->tag[V.a[k]] |= ORC_USER; }
      for (int k = 0; k < W.n; k++) { iv_push(&Wa, W.a[k]); t->tag[W.a[k]] |= ORC_USER; }
      for (int k = 0; k < X.n; k++) { iv_push(&Xa, X.a[k]); t->tag[X.a[k]] |= ORC_USER; }
    }
    t->u_off[i + 1] = Ua.n; t->v_off[i + 1] = Va.n; t->w_off[i + 1] = Wa.n; t->x_off[i + 1] = Xa.n;
  }
  t->u_lst = Ua.a; t->v_lst = Va.a; t->w_lst = Wa.a; t->x_lst = Xa.a;
  free(U.a); free(V.a); free(W.a); free(X.a); free(Q.a);
  return 0;
}

void orc_tree_free(orc_tree *t) {
  free(t->parent); free(t->child); free(t->depth); free(t->path); free(t->tag);
  free(t->src_cnt); free(t->trg_cnt); free(t->src_beg); free(t->src_num);
  free(t->trg_beg); free(t->trg_num); free(t->src_perm); free(t->trg_perm);
  free(t->u_off); free(t->v_off); free(t->w_off); free(t->x_off);
  free(t->u_lst); free(t->v_lst); free(t->w_lst); free(t->x_lst);
  memset(t, 0, sizeof(*t));
}

/* ======================================================================
 * M2L through the regular grid (matmgnt3d_mpi.cpp:276-431).
 * ====================================================================== */
/* plnDen2EffDen :276-307: scale by 2^(l*deg), scatter onto the grid, r2c */
void orc_eff_den(const orc_ops *op, int level, const double *ue, double *eff) {
  int np = op->np, sd = op->sd, n = 2 * np, G = n * n * n;
  int nue = op->n_ue;
  double *tmp = (double *)malloc(sizeof(double) * nue * sd);
  double *reg = (double *)calloc((size_t)G * sd, sizeof(double));
  for (int q = 0; q < nue; q++)
    for (int s = 0; s < sd; s++) tmp[q * sd + s] = ue[q * sd + s] * pow(2.0, level * op->deg[s]);
  sam2reg(np, sd, tmp, reg);
  orc_rfft3(n, sd, reg, eff);
  free(tmp); free(reg);
}

/* UpwEqu2DwnChk_dgemv :411-430 with cptwvv :502-512; idx = offset of the source
 * box from the target box in box widths */
void orc_m2l_hadamard(const orc_ops *op, const int idx[3], const double *eff_den, double *eff_val) {
  int np = op->np, sd = op->sd, td = op->td, n = 2 * np, H = (np + 1) * n * n;
  double nrmfc = 1.0 / (double)(n * n * n);
  int slot = (idx[0] + 3) * 49 + (idx[1] + 3) * 7 + (idx[2] + 3);
  const double *mat = op->m2l + (size_t)slot * sd * td * H * 2;
  for (int it = 0; it < td; it++)
    for (int js = 0; js < sd; js++) {
      int moff = js * td + it;
      for (int f = 0; f < H; f++) {
        const double *x = mat + 2 * ((size_t)f * sd * td + moff);
        const double *y = eff_den + 2 * ((size_t)f * sd + js);
        double *z = eff_val + 2 * ((size_t)f * td + it);
        z[0] += nrmfc * (x[0] * y[0] - x[1] * y[1]);
        z[1] += nrmfc * (x[0] * y[1] + x[1] * y[0]);
      }
    }
}

/* effVal2PlnVal :333-371: c2r then gather the surface samples, added into dc */
void orc_eff_val(const orc_ops *op, const double *eff_val, double *dc) {
  int np = op->np, td = op->td, n = 2 * np, G = n * n * n;
  double *regv = (double *)malloc(sizeof(double) * (size_t)G * td);
  orc_irfft3(n, td, eff_val, regv);
  reg2sam(np, td, regv, dc);
  free(regv);
}

/* ======================================================================
 * Evaluate: fmm3d_eval_mpi.cpp:396-1515, CPU branches, one rank.
 * ====================================================================== */
static void node_center(const orc_tree *t, int i, const double ctr[3], double rad, double *c) {
  int tmp = 1 << t->depth[i];
  for (int d = 0; d < 3; d++) c[d] = (ctr[d] - rad) + (2 * rad) * (t->path[3 * i + d] + 0.5) / (double)tmp;
}

/* fmm3d_mpi.cpp:81-119 (and :124-249 with generated positions): dense block + GEMV,
 * targets in chunks of TMAX = 1024 */
static void dense_apply(int kt, const double *coefs, int ns, const double *sp, const double *sn,
                        const double *den, int nt, const double *tp, double *val, double *work) {
  int sd = orc_src_dof(kt), td = orc_trg_dof(kt);
  const int TMAX = 1024;
  for (int s = 0; s < nt; s += TMAX) {
    int num = nt - s < TMAX ? nt - s : TMAX;
    orc_kernel_matrix(kt, coefs, ns, sp, sn, num, tp + 3 * s, work);
    gemv_acc(num * td, ns * sd, work, den, val + (size_t)s * td);
  }
}

int orc_evaluate(const orc_tree *t, const orc_ops *op, int kt_user, const double *coefs,
                 int root_level, const double ctr[3], const double *spos, const double *snor,
                 const double *tpos, const double *src_den, double *trg_val, double *ue_out,
                 double *de_out) {
  int SD = orc_src_dof(kt_user), TD = orc_trg_dof(kt_user);
  int km = op->kt_mm, sd = op->sd, td = op->td, np = op->np;
  if (orc_mm_kernel(kt_user) != km) return -1;
  int N = t->n_nodes;
  int nUE = op->n_ue, nUC = op->n_uc, nDE = op->n_de, nDC = op->n_dc;
  int szUE = nUE * sd, szUC = nUC * td, szDE = nDE * sd, szDC = nDC * td;
  double rad0 = pow(2.0, -root_level);
  double *UE = (double *)calloc((size_t)N * szUE, sizeof(double));
  double *UC = (double *)calloc((size_t)N * szUC, sizeof(double));
  double *DE = (double *)calloc((size_t)N * szDE, sizeof(double));
  double *DC = (double *)calloc((size_t)N * szDC, sizeof(double));
  int ns = t->n_src, nt = t->n_trg;
  double *SP = (double *)malloc(sizeof(double) * 3 * (ns + 1));
  double *SN = (double *)malloc(sizeof(double) * 3 * (ns + 1));
  double *SDN = (double *)malloc(sizeof(double) * SD * (ns + 1));
  double *TP = (double *)malloc(sizeof(double) * 3 * (nt + 1));
  double *TV = (double *)calloc((size_t)TD * (nt + 1), sizeof(double));
  for (int k = 0; k < ns; k++) { /* :443-461 load exact densities into leaf order */
    int o = t->src_perm[k];
    for (int d = 0; d < 3; d++) { SP[3 * k + d] = spos[3 * o + d]; SN[3 * k + d] = snor[3 * o + d]; }
    for (int d = 0; d < SD; d++) SDN[SD * k + d] = src_den[SD * o + d];
  }
  for (int k = 0; k < nt; k++) {
    int o = t->trg_perm[k];
    for (int d = 0; d < 3; d++) TP[3 * k + d] = tpos[3 * o + d];
  }
  double *sUE = sample_tab(np, ORC_UE), *sUC = sample_tab(np, ORC_UC);
  double *sDE = sample_tab(np, ORC_DE), *sDC = sample_tab(np, ORC_DC);
  int maxw = 1024 * (TD > td ? TD : td);
  int maxs = 1;
  for (int i = 0; i < N; i++) if (t->src_num[i] > maxs) maxs = t->src_num[i];
  int cols = maxs * SD; if (szUE > cols) cols = szUE; if (szDE > cols) cols = szDE;
  int rows = maxw; if (szUC > rows) rows = szUC; if (szDC > rows) rows = szDC;
  double *work = (double *)malloc(sizeof(double) * (size_t)rows * cols);
  double *pts = (double *)malloc(sizeof(double) * 3 * nUC);
  double *tmpd = (double *)malloc(sizeof(double) * (szUC > szDC ? szUC : szDC));
  double c[3];

  /* up pass :557-595, children first (descending index) */
  for (int i = N - 1; i >= 0; i--) {
    if (!(t->tag[i] & ORC_CBTR)) continue;
    double r = rad0 / (double)(1 << t->depth[i]);
    double *uc = UC + (size_t)i * szUC;
    if (is_leaf(t, i)) { /* S2M fmm3d_mpi.cpp:124-137 */
      node_center(t, i, ctr, rad0, c);
      loc_pos(nUC, sUC, c, r, pts);
      int b = t->src_beg[i];
      dense_apply(kt_user, coefs, t->src_num[i], SP + 3 * b, SN + 3 * b, SDN + SD * b, nUC, pts, uc, work);
    } else { /* M2M matmgnt3d_mpi.cpp:151-193 */
      for (int a = 0; a < 8; a++) {
        int ch = t->child[i] + a;
        if (!(t->tag[ch] & ORC_CBTR)) continue;
        int l = t->depth[ch] + root_level;
        for (int q = 0; q < nUE; q++)
          for (int s = 0; s < sd; s++)
            tmpd[q * sd + s] = UE[(size_t)ch * szUE + q * sd + s] * pow(2.0, l * op->deg[s]);
        gemv_acc(szUC, szUE, op->m2m + (size_t)a * szUC * szUE, tmpd, uc);
      }
    }
    /* UC -> UE matmgnt3d_mpi.cpp:107-149 */
    int l = t->depth[i] + root_level;
    memset(tmpd, 0, sizeof(double) * szUE);
    gemv_acc(szUE, szUC, op->pinv_up, uc, tmpd);
    for (int q = 0; q < nUE; q++)
      for (int s = 0; s < sd; s++)
        UE[(size_t)i * szUE + q * sd + s] += tmpd[q * sd + s] * pow(2.0, -l * op->deg[s]);
  }

  /* U list :764-786 */
  for (int i = N - 1; i >= 0; i--) {
    if (!(t->tag[i] & ORC_EVTR) || !is_leaf(t, i)) continue;
    int tb = t->trg_beg[i], tn = t->trg_num[i];
    for (int64_t q = t->u_off[i]; q < t->u_off[i + 1]; q++) {
      int A = t->u_lst[q], b = t->src_beg[A];
      dense_apply(kt_user, coefs, t->src_num[A], SP + 3 * b, SN + 3 * b, SDN + SD * b, tn,
                  TP + 3 * tb, TV + (size_t)TD * tb, work);
    }
  }

  /* V list :1007-1059 */
  {
    int n = 2 * np, H = (np + 1) * n * n;
    double **eff_den = (double **)calloc(N, sizeof(double *));
    double *effv = (double *)malloc(sizeof(double) * 2 * (size_t)H * td);
    for (int i = N - 1; i >= 0; i--) {
      if (!(t->tag[i] & ORC_EVTR)) continue;
      if (t->v_off[i + 1] == t->v_off[i]) continue;
      double ci[3]; node_center(t, i, ctr, rad0, ci);
      double D = 2.0 * (rad0 / (double)(1 << t->depth[i]));
      int l = t->depth[i] + root_level;
      memset(effv, 0, sizeof(double) * 2 * (size_t)H * td);
      for (int64_t q = t->v_off[i]; q < t->v_off[i + 1]; q++) {
        int A = t->v_lst[q];
        double ca[3]; node_center(t, A, ctr, rad0, ca);
        int idx[3];
        for (int d = 0; d < 3; d++) idx[d] = (int)floor((ca[d] - ci[d]) / D + 0.5);
        if (!eff_den[A]) {
          eff_den[A] = (double *)malloc(sizeof(double) * 2 * (size_t)H * sd);
          orc_eff_den(op, l, UE + (size_t)A * szUE, eff_den[A]);
        }
        orc_m2l_hadamard(op, idx, eff_den[A], effv);
      }
      orc_eff_val(op, effv, DC + (size_t)i * szDC);
    }
    for (int i = 0; i < N; i++) free(eff_den[i]);
    free(eff_den); free(effv);
  }

  /* W list :1062-1231 (CPU part) */
  for (int i = N - 1; i >= 0; i--) {
    if (!(t->tag[i] & ORC_EVTR) || !is_leaf(t, i)) continue;
    int tb = t->trg_beg[i], tn = t->trg_num[i];
    for (int64_t q = t->w_off[i]; q < t->w_off[i + 1]; q++) {
      int A = t->w_lst[q];
      if (t->src_beg[A] >= 0 && t->src_num[A] * SD < szUE) {
        int b = t->src_beg[A];
        dense_apply(kt_user, coefs, t->src_num[A], SP + 3 * b, SN + 3 * b, SDN + SD * b, tn,
                    TP + 3 * tb, TV + (size_t)TD * tb, work);
      } else { /* M2T fmm3d_mpi.cpp:208-249 */
        double ca[3]; node_center(t, A, ctr, rad0, ca);
        loc_pos(nUE, sUE, ca, rad0 / (double)(1 << t->depth[A]), pts);
        dense_apply(km, op->coefs, nUE, pts, pts, UE + (size_t)A * szUE, tn, TP + 3 * tb,
                    TV + (size_t)TD * tb, work);
      }
    }
  }

  /* X list :1233-1265 */
  for (int i = N - 1; i >= 0; i--) {
    if (!(t->tag[i] & ORC_EVTR)) continue;
    for (int64_t q = t->x_off[i]; q < t->x_off[i + 1]; q++) {
      int A = t->x_lst[q], b = t->src_beg[A];
      if (is_leaf(t, i) && t->trg_num[i] * TD < szDC) {
        int tb = t->trg_beg[i];
        dense_apply(kt_user, coefs, t->src_num[A], SP + 3 * b, SN + 3 * b, SDN + SD * b,
                    t->trg_num[i], TP + 3 * tb, TV + (size_t)TD * tb, work);
      } else { /* S2L fmm3d_mpi.cpp:142-154 */
        double cb[3]; node_center(t, i, ctr, rad0, cb);
        loc_pos(nDC, sDC, cb, rad0 / (double)(1 << t->depth[i]), pts);
        dense_apply(kt_user, coefs, t->src_num[A], SP + 3 * b, SN + 3 * b, SDN + SD * b, nDC,
                    pts, DC + (size_t)i * szDC, work);
      }
    }
  }

  /* down pass :1368-1438, parents first (ascending index) */
  for (int i = 0; i < N; i++) {
    if (!(t->tag[i] & ORC_EVTR)) continue;
    double *dc = DC + (size_t)i * szDC, *de = DE + (size_t)i * szDE;
    if (t->depth[i] >= 3) { /* L2L matmgnt3d_mpi.cpp:231-274 */
      int p = t->parent[i];
      int a = t->path[3 * i] - 2 * t->path[3 * p], b = t->path[3 * i + 1] - 2 * t->path[3 * p + 1];
      int cc = t->path[3 * i + 2] - 2 * t->path[3 * p + 2];
      int l = t->depth[p] + root_level;
      for (int q = 0; q < nDE; q++)
        for (int s = 0; s < sd; s++)
          tmpd[q * sd + s] = DE[(size_t)p * szDE + q * sd + s] * pow(2.0, l * op->deg[s]);
      gemv_acc(szDC, szDE, op->l2l + (size_t)(4 * a + 2 * b + cc) * szDC * szDE, tmpd, dc);
    }
    if (t->depth[i] >= 2) { /* DC -> DE :196-229 */
      int l = t->depth[i] + root_level;
      double *tmp2 = (double *)calloc(szDE, sizeof(double));
      gemv_acc(szDE, szDC, op->pinv_dn, dc, tmp2);
      for (int q = 0; q < nDE; q++)
        for (int s = 0; s < sd; s++) de[q * sd + s] += tmp2[q * sd + s] * pow(2.0, -l * op->deg[s]);
      free(tmp2);
    }
    if (is_leaf(t, i)) { /* L2T fmm3d_mpi.cpp:159-202 */
      node_center(t, i, ctr, rad0, c);
      loc_pos(nDE, sDE, c, rad0 / (double)(1 << t->depth[i]), pts);
      int tb = t->trg_beg[i];
      dense_apply(km, op->coefs, nDE, pts, pts, de, t->trg_num[i], TP + 3 * tb,
                  TV + (size_t)TD * tb, work);
    }
  }

  /* write back :1490-1510 */
  for (int k = 0; k < nt; k++)
    for (int d = 0; d < TD; d++) trg_val[(size_t)TD * t->trg_perm[k] + d] = TV[(size_t)TD * k + d];
  if (ue_out) memcpy(ue_out, UE, sizeof(double) * (size_t)N * szUE);
  if (de_out) memcpy(de_out, DE, sizeof(double) * (size_t)N * szDE);
  free(UE); free(UC); free(DE); free(DC); free(SP); free(SN); free(SDN); free(TP); free(TV);
  free(sUE); free(sUC); free(sDE); free(sDC); free(work); free(pts); free(tmpd);
  return 0;
}

/* dense3d_mpi.cpp:122-129: direct sum with the user kernel (r < 1e-8 excluded) */
int orc_direct(int kt, const double *coefs, int ns, const double *spos, const double *snor,
               const double *sden, int nt, const double *tpos, double *tval, int nthreads) {
  int sd = orc_src_dof(kt), td = orc_trg_dof(kt);
  if (sd == 0) return -1;
  (void)nthreads;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int i = 0; i < nt; i++) {
    double acc[3] = {0, 0, 0}, blk[16];
    for (int j = 0; j < ns; j++) {
      kblock(kt, coefs, tpos[3 * i] - spos[3 * j], tpos[3 * i + 1] - spos[3 * j + 1],
             tpos[3 * i + 2] - spos[3 * j + 2], snor + 3 * j, blk);
      for (int a = 0; a < td; a++)
        for (int s = 0; s < sd; s++) acc[a] += blk[a * sd + s] * sden[sd * j + s];
    }
    for (int a = 0; a < td; a++) tval[(size_t)td * i + a] = acc[a];
  }
  return 0;
}
