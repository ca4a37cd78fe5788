// Host-side launchers for the HIP kernels of the evaluate pass.  All pointers are
// device pointers; every launcher enqueues on `stream` and returns the launch
// status.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernel3d.hpp"

namespace kifmm {

// A chunk of at most kNearChunk targets of one leaf and its source items.
// own: position (from ibeg) of the leaf's own exact item, -1 if absent.
// unchk: 1 if no other item has a source within 1e-7 of the chunk's targets
// (bounding boxes), so that only the own item's sources can fall under the
// 1e-8 pair cut; 0 (the default) keeps the cut on every pair.
struct NearChunk {
  int node, tbeg, tnum, ibeg, iend;
  int own = -1, unchk = 0;
};
constexpr int kNearChunk = 128;
// Right-hand sides per launch of the point kernels: groups of 1, 2, 4 or 8
// (compile-time accumulator arrays); more go in several launches.
constexpr int kMaxNr = 8;
__host__ __device__ constexpr int rhs_group(int nr) { return nr <= 1 ? 1 : nr <= 2 ? 2 : nr <= 4 ? 4 : 8; }
constexpr int kPairStride = 224;  // 216 halo slots rounded up to a multiple of 8

// Packed half spectrum of the mixed path.  The r2c layout of an n^3 real grid
// (n = 2 np) holds (np + 1) n^2 values, frequency (ka, kb, kc), kc in [0, np].
// In the planes kc = 0 and kc = np the value at (-ka, -kb) is the conjugate of
// the one at (ka, kb), and so is every product of two such spectra, so the
// V-list sum there is needed on one of each pair only: the packed form keeps
// n^2 np + 4 values (868 of 1,008 at np = 6).
//   kc in [1, np-1]: (ka n + kb)(np - 1) + kc - 1;
//   kc in {0, np}:   base n^2 (np-1) (+ n^2/2 + 2 for kc = np), then ka = 0 with
//                    kb in [0, np], ka in [1, np-1] with every kb, ka = np with
//                    kb in [0, np].
// spec_pack returns 2 * packed index + 1 when the value at (ka, kb, kc) is the
// conjugate of the stored one (its mirror is the representative).  A row is
// padded with zeros to a multiple of 16 values (880 at np = 6), so that every
// 64-frequency chunk of a row starts on a 128-byte line.
__host__ __device__ constexpr int spec_packed_count(int np) { return 4 * np * np * np + 4; }
__host__ __device__ constexpr int spec_packed_len(int np) { return (spec_packed_count(np) + 15) / 16 * 16; }
__host__ __device__ constexpr int spec_pack(int np, int ka, int kb, int kc) {
  const int n = 2 * np;
  if (kc > 0 && kc < np) return ((ka * n + kb) * (np - 1) + kc - 1) << 1;
  int a = ka, b = kb, cj = 0;
  if (a > np || ((a == 0 || a == np) && b > np)) {
    a = (n - a) % n;
    b = (n - b) % n;
    cj = 1;
  }
  const int base = n * n * (np - 1) + (kc ? n * n / 2 + 2 : 0);
  const int p = a == 0 ? b : a == np ? np + 1 + (np - 1) * n + b : np + 1 + (a - 1) * n + b;
  return ((base + p) << 1) | cj;
}
// spec_pack along one line e = kb (np + 1) + kc of the transforms, with the
// line's part computed once: in the planes kc in [1, np-1] frequency ka sits at
// lin + ka n (np - 1); in the planes kc = 0, np at pb + ka n for ka in [1, np-1],
// mirrored at pm + (n - ka) n for ka in [np+1, n-1], and at z0, zh for ka = 0, np
// (mirrored when kb > np).  With ka a constant each index is one add or select.
template <int NP>
struct SpecLine {
  static constexpr int N = 2 * NP;
  int inner, hi, lin, pb, pm, z0, zh;
  __host__ __device__ explicit SpecLine(int e) {
    const int kb = e / (NP + 1), kc = e % (NP + 1);
    inner = kc > 0 && kc < NP;
    hi = kb > NP;
    lin = kb * (NP - 1) + kc - 1;
    const int base = N * N * (NP - 1) + (kc ? N * N / 2 + 2 : 0);
    const int bm = kb ? N - kb : 0;  // the mirror's kb
    pb = base + NP + 1 - N + kb;
    pm = base + NP + 1 - N + bm;
    const int bz = hi ? bm : kb;
    z0 = base + bz;
    zh = base + NP + 1 + (NP - 1) * N + bz;
  }
  // a line of the planes kc = 0, np
  __host__ __device__ int plane_index(int ka) const {
    return ka == 0 ? z0 : ka == NP ? zh : ka < NP ? pb + ka * N : pm + (N - ka) * N;
  }
  __host__ __device__ bool plane_conj(int ka) const { return ka == 0 || ka == NP ? hi != 0 : ka > NP; }
  __host__ __device__ int index(int ka) const { return inner ? lin + ka * N * (NP - 1) : plane_index(ka); }
  // the value at ka is the conjugate of the stored one (not stored by this line)
  __host__ __device__ bool conj(int ka) const { return !inner && plane_conj(ka); }
};
// Source item code: (node << 2) | kind
enum ItemKind { ITEM_EXACT = 0, ITEM_UE = 1, ITEM_DE = 2 };

// A to-surface task: check surface of `node` receives the field of the exact
// sources of the leaves items[ibeg..iend).
struct SurfTask {
  int node, ibeg, iend;
};

struct NearArgs {
  const NearChunk* chunks; int nchunks;
  const int* items;
  const double* spos; const double* snor; const double* sden;
  const int* src_beg; const int* src_num;
  const double4* geom;
  const double* sam_ue; int n_ue;
  const double* sam_de; int n_de;
  const double* UE; int sz_ue;
  const double* DE; int sz_de;
  const double* tpos; double* tval;
  // mixed Laplace near field: if set, value i (leaf order) goes to
  // tval[tperm[i]], the caller's order (no separate scatter)
  const int* tperm;
  // Several right-hand sides: sden holds nrhs densities per source
  // (sden[(k * nrhs + r) * SD + s]); right-hand side r's equivalent densities
  // are rows r * nstride + node of UE and DE, and its values start at
  // tval + r * tstride (tstride = n_trg * TD).  One launch covers the nr
  // right-hand sides r0 .. r0 + nr - 1.
  int nrhs = 1, r0 = 0, nr = 1;
  int64_t nstride = 0, tstride = 0;
  // Mixed near kernels: chunk c of the launch is chunks[order[c]] (order null:
  // chunks[c]); with queue set, a persistent grid of workgroups takes chunk
  // indices from the global counter *queue (zeroed by the launcher) instead
  // of one workgroup per chunk.
  const int* order = nullptr;
  int* queue = nullptr;
};

struct SurfArgs {
  const SurfTask* tasks; int ntasks;
  const int* items;
  const double* spos; const double* snor; const double* sden;
  const int* src_beg; const int* src_num;
  const double4* geom;
  const double* sam; int n_sam;
  double* out; int sz_out;
  // tasks of one leaf each, as (leaf, first source, source count): saves the
  // task -> item -> range lookups (mixed S2M; null elsewhere)
  const int4* leaf;
  // several right-hand sides (as NearArgs): nrhs densities per source, right-
  // hand side r writes rows r * nstride + node; this launch: r0 .. r0 + nr - 1
  int nrhs = 1, r0 = 0, nr = 1;
  int64_t nstride = 0;
};

struct GemmArgs {
  int M, K; const double* A;
  const double* B; const int* bidx; int64_t ldb; const double* kscale; int ksd;
  double* C; const int* cidx; int64_t ldc; const double* rscale; int rsd;
  int ncol; int accumulate;
  // two-part B (mixed path, k1 > 0): rows [0, k1) of column j are
  // B1[cidx[j] * ldb1 + k], unscaled; rows [k1, K) are B[bidx[j] * ldb + k - k1]
  // scaled by kscale[(k - k1) % ksd].  k1 = 0: B alone, as above.
  const double* B1; int64_t ldb1; int k1;
  // several matrices in one launch: matrix z is A + z*a_zstride and owns the
  // columns [zoff[z], zoff[z+1]); zoff == nullptr means one matrix, all columns
  const int* zoff; int nz; int64_t a_zstride; int max_zcol;
  // split-K (mixed path only): blockIdx.z = matrix * ksplit + c sums k in
  // [c*kchunk, (c+1)*kchunk).  The partial products go to work (ksplit * ncol *
  // M doubles, required) and a second kernel sums them in chunk order into C
  // (deterministic).
  int ksplit; int kchunk;
  double* work;
};

struct VlistArgs {
  int np, n, H, sd, td;
  int nsrc, ntrg, ngrp;
  const int* src_node; const int* trg_node; const int* depth; int root_level;
  const double* deg;           // per source DOF, scale 2^(l deg) applied to the input samples
  // per target DOF, scale 2^(l deg_out) applied to the output samples (null:
  // none); the transposed pass of a kernel whose DOFs have different degrees
  // scales its output, as the forward pass scales its input
  const double* deg_out = nullptr;
  const int* shell;            // surface sample -> block-relative (k*np+j)*np+i
  const double2* twiddle;      // n entries exp(-2 pi i m/n)
  const double* UE; int sz_ue;
  double* DC; int sz_dc;
  const int* halo; const int* hord; const uint64_t* mask; const int* tgt;
  // per group, the halo sources in V-list (node) order, padded to a multiple of 8
  // with -1: source slot, and (halo index | offset slot of the halo index << 8)
  const int* psrc; const int* pslot; const int* pcnt;
  const double2* m2l;          // [343][sd*td][H]
  double2* dspec;              // [nrhs][nsrc][sd][H]
  double2* espec;              // [nrhs][ntrg][td][H]
  // Several right-hand sides: right-hand side r of source slot s is spectrum
  // row r * nsrc + s, read from UE row r * nstride + node (likewise targets
  // and DC); the groups run as ngrp * nrhs virtual groups g * nrhs + r, so
  // consecutive virtual groups share their translation spectra.
  int nrhs = 1;
  int64_t nstride = 0;
  // mixed Hadamard of the vector kernels (sd, td > 1): 0 one wave per group
  // reading the translation spectra per pair, 1 and 2 the halo sweep over tiles
  // of 1 or 2 groups
  int had_form = 0;
  // the translation spectra of DOFs (j, i) and (i, j), i, j < 3, are equal (the
  // 3 x 3 block G delta_ij + H x_i x_j of the Stokes and Navier equivalent
  // kernels): the mixed per-pair Hadamard loads one of each pair
  int sym3 = 0;
  // the FP32 translation spectra of opposite offsets are conjugates, T(-s) =
  // conj(T(s)) (scalar even kernel, table without the kernel grid's wrap
  // entries): the mixed halo sweep serves offsets q and -q from one load
  int mirror = 0;
  // the targets' check values are zero when the inverse transform runs (and
  // no other writer comes first): it stores instead of reading and adding
  int dc_store = 0;
};

// Adjoint near field (source-centric): chunk = <= kNearChunk sources of a
// source leaf (node, tbeg = first source in leaf order, tnum); items are
// (node << 2) | kind over the points that feed them.
enum AdjItemKind { ADJ_TRG = 0, ADJ_DC = 1, ADJ_UC = 2 };
struct NearTArgs {
  const NearChunk* chunks; int nchunks;
  const int* items;
  const double* spos; const double* snor;            // sources, leaf order
  const double* tpos; const double* tw;              // targets, leaf order, adjoint input (TD each)
  const int* trg_beg; const int* trg_num;
  const double4* geom;
  const double* sam_dc; int n_dc; const double* DC; int sz_dc;
  const double* sam_uc; int n_uc; const double* UC; int sz_uc;
  double* sout;                                      // SD per source, leaf order
};

// Adjoint of M2T / L2T: the equivalent surface of `node` receives from the
// targets of the target leaves items[ibeg..iend).
struct SurfTArgs {
  const SurfTask* tasks; int ntasks;
  const int* items;
  const double* tpos; const double* tw; const int* trg_beg; const int* trg_num;
  const double4* geom;
  const double* sam; int n_sam;
  double* out; int sz_out;
};

// Gradient (with respect to the target) of the near-field and local-expansion
// terms, 3 values per target written to a.tval; Laplace single layer.
hipError_t launch_near_grad(int kt, const NearArgs& a, const KParams<double>& pu, const KParams<double>& pm,
                            hipStream_t s);
hipError_t launch_near_adjoint(int kt, const NearTArgs& a, const KParams<double>& pu, hipStream_t s);
hipError_t launch_surface_adjoint(int km, const SurfTArgs& a, const KParams<double>& pm, hipStream_t s);

hipError_t launch_gather(const double* in, const int* perm, double* out, int64_t n, int dof, hipStream_t s);
hipError_t launch_scatter(const double* in, const int* perm, double* out, int64_t n, int dof, hipStream_t s);
hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s);
// Zero rows[k] + r * nstride (k < nrows, r < nrhs) of a row-major array of rows
// of len doubles.
hipError_t launch_zero_rows(double* base, const int* rows, int nrows, int len, int nrhs, int64_t nstride,
                            hipStream_t s);
// nrhs density vectors [r][k][dof] (caller's order) -> leaf order [sinv[k]][r][dof]
hipError_t launch_load_rhs(const double* in, const int* sinv, double* out, int64_t n, int dof, int nrhs,
                           hipStream_t s);
// nrhs value vectors in leaf order [r][k][dof] -> caller's order [r][perm[k]][dof]
hipError_t launch_store_rhs(const double* in, const int* perm, double* out, int64_t n, int dof, int nrhs,
                            hipStream_t s);
hipError_t launch_near(int kt, const NearArgs& a, const KParams<double>& pu, const KParams<double>& pm,
                       hipStream_t s);
hipError_t launch_surface(int kt, const SurfArgs& a, const KParams<double>& pu, hipStream_t s);
hipError_t launch_gemm(const GemmArgs& g, hipStream_t s);
// FFT, Hadamard, inverse FFT; e0/e1 (may be null) are recorded around the Hadamard.
hipError_t launch_vlist(const VlistArgs& v, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Mixed-precision path (kernels_mixed.hip) and FP32-storage FFTs.
hipError_t launch_near_mixed(int kt, const NearArgs& a, const KParams<double>& pu, const KParams<double>& pm,
                             hipStream_t s);
hipError_t launch_surface_mixed(int kt, const SurfArgs& a, const KParams<double>& pu, hipStream_t s);
hipError_t launch_hadamard_mixed(const VlistArgs& v, const float2* dspec, const float2* m2l, float2* espec,
                                 hipStream_t s);
hipError_t launch_gemm_fma(const GemmArgs& g, hipStream_t s);
// Translation GEMM as split-BF16 products on the matrix cores (kernels_bf16.hip):
// a3 holds A's three BF16 parts [3][nz][Mp][Kp] (row-major, Mp a multiple of
// 16 and Kp of 32, zero padded), terms 3 (BF16x3) or 6 (BF16x6); no split K.
hipError_t launch_gemm_bf16s(const GemmArgs& g, const void* a3, int Mp, int Kp, int terms, hipStream_t s);
hipError_t launch_fft_fwd32(const VlistArgs& v, float2* dspec, hipStream_t s);
hipError_t launch_fft_inv32(const VlistArgs& v, const float2* espec, hipStream_t s);

// FP64 direct sum at nt targets: sources in leaf order (spos, snor), densities
// in the caller's order read through sperm (leaf position -> caller index).
// The sources are cut into nchunk chunks of `chunk` (a multiple of 256); the
// partial sums need work_doubles = nchunk * nt * TD doubles of workspace.
struct DirectPlan {
  int nchunk; int64_t chunk; int64_t work_doubles;
};
// nchunk == 0 when the workspace capacity cannot hold one chunk's partial sums.
DirectPlan plan_direct(int64_t ns, int nt, int td, int64_t cap_doubles);
hipError_t launch_direct(int kt, const KParams<double>& p, int64_t ns, const double* spos, const double* snor,
                         const int* sperm, const double* sden, int nt, const double* tpos, const DirectPlan& d,
                         double* work, double* tval, hipStream_t s);

// Library-held scratch of kScratchDoubles doubles per GPU (a static device
// array, no allocation): its device address on the current GPU, looked up on
// the first call for that GPU and kept.
constexpr int64_t kScratchDoubles = int64_t(1) << 21;  // 16 MiB
hipError_t scratch_address(double** out);

}  // namespace kifmm
