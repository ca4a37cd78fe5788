/*
 * kifmm.h -- C ABI of the MI355X KIFMM3d evaluate path (libkifmm.so).
 *
 * Drop-in for the evaluate path of the reference's FMM3d_MPI on one MPI rank.
 * Every entry point below replaces the reference interface cited beside it
 * (paths relative to the reference root).  Plain pointers and sizes only; every
 * call returns an int status (0 = success, negative = error class) and never
 * aborts; kifmm_last_error() gives the message.  Handles are independent and
 * thread-compatible (one host thread per handle at a time).
 *
 * Layouts (the reference's, tt0.cpp:472-476, fmm3d_eval_mpi.cpp:455,1503):
 *   positions / normals: xyz interleaved, 3 doubles per point;
 *   densities: srcDOF doubles per point; values: trgDOF doubles per point.
 */
#ifndef KIFMM_H
#define KIFMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KIFMM_OK 0
#define KIFMM_ERR_ARG (-2)      /* bad argument or unsupported configuration */
#define KIFMM_ERR_STATE (-3)    /* call out of order (e.g. evaluate before setup) */
#define KIFMM_ERR_OPTIONS (-4)  /* options file missing or a required option absent */
#define KIFMM_ERR_HIP (-5)      /* HIP runtime error (no device, allocation, launch) */
#define KIFMM_ERR_INTERNAL (-6)

#define KIFMM_PREC_FP64 0
#define KIFMM_PREC_MIXED 1

typedef struct kifmm_fmm *kifmm_handle;

/* Constructor + configuration: FMM3d_MPI(prefix) and its setters knl(), np(),
 * ctr(), rootLevel() (fmm3d_mpi.hpp:170-193, tt0.cpp:429-437) plus the
 * -fmm3d_let3d_ptsMax / -maxLevel options (let3d_mpi.cpp:231-232). */
int kifmm_create(int kernel_type, const double *coefs, int ncoefs, int np, int pts_max,
                 int max_level, const double ctr[3], int root_level, int precision,
                 kifmm_handle *out);

/* Same, configured from a PETSc-style options file as tt0 reads it
 * (tt0.cpp:150-183, fmm3d_setup_mpi.cpp:48, let3d_mpi.cpp:231-232):
 * -kt, -<prefix>np, -<prefix>let3d_ptsMax, -<prefix>let3d_maxLevel; optional
 * -kifmm_mu, -kifmm_nu (default 1, 0.25 as tt0.cpp:179-183) and
 * -kifmm_precision fp64|mixed.  prefix NULL means "fmm3d_". */
int kifmm_create_from_options(const char *options_path, const char *prefix, kifmm_handle *out);

/* ~FMM3d_MPI (fmm3d_mpi.cpp:41-76). */
int kifmm_destroy(kifmm_handle h);

/* srcPos() / srcNor() / trgPos() setters (knlmat3d_mpi.hpp:50-56).  Copies the
 * arrays; src_nor may be NULL for kernels without normals. */
int kifmm_set_points(kifmm_handle h, int64_t n_src, const double *src_pos, const double *src_nor,
                     int64_t n_trg, const double *trg_pos);

/* Many independent problems in one handle, the batched form of the point
 * setters above (the reference builds one FMM3d_MPI per problem, tt0.cpp:429-518;
 * no batched counterpart): problem p has n_src[p] sources and n_trg[p]
 * targets, and the arrays hold the problems one after another (positions and
 * normals 3 per point).  trg_pos NULL (n_trg ignored) means targets = sources.
 * kifmm_setup builds one octree per problem and numbers them as one forest;
 * each pass of kifmm_evaluate* then runs one launch over every problem's
 * boxes.  Densities and values are the problems' vectors one after another
 * (per right-hand side), in each problem's point order. */
int kifmm_set_points_batch(kifmm_handle h, int nprob, const int64_t *n_src, const double *src_pos,
                           const double *src_nor, const int64_t *n_trg, const double *trg_pos);

/* KnlMat3d_MPI::setup() (knlmat3d_mpi.hpp:59; fmm3d_setup_mpi.cpp:41). */
int kifmm_setup(kifmm_handle h);

/* Host half of setup: octree, lists, operator tables and task tables, without
 * touching the GPU (the tree, list, operator and work-count exports below need
 * only this).  kifmm_setup() performs it too.  No reference counterpart: the
 * reference's setup has no device half. */
int kifmm_setup_host(kifmm_handle h);

/* srcDOF() / trgDOF() (knlmat3d_mpi.hpp:67-69); negative on error. */
int kifmm_src_dof(kifmm_handle h);
int kifmm_trg_dof(kifmm_handle h);

/* KnlMat3d_MPI::evaluate(Vec srcDen, Vec trgVal) (knlmat3d_mpi.hpp:62;
 * fmm3d_eval_mpi.cpp:396).  Host pointers; copies in, evaluates on the GPU,
 * copies out, synchronous.  trg_val is overwritten. */
int kifmm_evaluate(kifmm_handle h, const double *src_den, double *trg_val);

/* Same with device pointers, enqueued on `stream` (a hipStream_t; NULL = the
 * default stream); returns once enqueued. */
int kifmm_evaluate_device(kifmm_handle h, const double *d_src_den, double *d_trg_val, void *stream);

/* nrhs right-hand sides in one call (the nrhs argument proposed for
 * KnlMat3d_MPI::evaluate, knlmat3d_mpi.hpp:62, which takes one Vec per call):
 * src_den holds nrhs density vectors one after another (nrhs x n_src x
 * srcDOF), trg_val receives nrhs value vectors (nrhs x n_trg x trgDOF),
 * overwritten.  One pass serves all right-hand sides: each phase is one
 * launch over every right-hand side (groups of 8 for the point kernels), and
 * each right-hand side's values equal those of a one-RHS evaluate to
 * round-off (bitwise in FP64).  The handle's work buffers grow to nrhs blocks
 * on the first call with a larger nrhs.  Host pointers, synchronous. */
int kifmm_evaluate_multi(kifmm_handle h, const double *src_den, double *trg_val, int nrhs);

/* Same with device pointers, enqueued on `stream`. */
int kifmm_evaluate_multi_device(kifmm_handle h, const double *d_src_den, double *d_trg_val, int nrhs,
                                void *stream);

/* Transposed evaluate: src_out = A^T trg_in, where A is the operator that
 * kifmm_evaluate applies (the transposed FMM the reference lacks: its
 * evaluate, fmm3d_eval_mpi.cpp:396-1515, has no adjoint; every forward step is
 * transposed in reverse order).  trg_in: trgDOF per target, src_out: srcDOF
 * per source, caller's order, src_out overwritten.  All eight FMM kernels
 * (111, 121, 311, 312, 321, 322, 511, 521).  Host pointers, synchronous. */
int kifmm_evaluate_adjoint(kifmm_handle h, const double *trg_in, double *src_out);

/* Same with device pointers, enqueued on `stream` (the first call on a handle
 * builds the adjoint tables and synchronises `stream` first). */
int kifmm_evaluate_adjoint_device(kifmm_handle h, const double *d_trg_in, double *d_src_out, void *stream);

/* Field pass: the gradient of the evaluated potential at every target,
 * 3 values per target in the caller's order (the reference computes values only;
 * this is the position-gradient building block of gradient-based callers).
 * Laplace single layer (kernel 111) only; KIFMM_ERR_ARG otherwise.  Host
 * pointers, synchronous. */
int kifmm_field(kifmm_handle h, const double *src_den, double *trg_grad);

/* Same with device pointers, enqueued on `stream`. */
int kifmm_field_device(kifmm_handle h, const double *d_src_den, double *d_trg_grad, void *stream);

/* Checkpoint of the upward equivalent densities (the UE vectors of
 * fmm3d_eval_mpi.cpp:480-619, kept for reuse by a gradient pass; no
 * reference counterpart).  kifmm_checkpoint_size gives the doubles a
 * checkpoint holds.  kifmm_evaluate_checkpoint_device is
 * kifmm_evaluate_device that also copies UE to d_ue_out once the up pass is
 * done; kifmm_field_checkpoint_device is kifmm_field_device that reads UE from
 * such a checkpoint instead of running the up pass again (d_src_den must be
 * the checkpointed densities).  Device pointers, enqueued on `stream`. */
int kifmm_checkpoint_size(kifmm_handle h, int64_t *ndoubles);
int kifmm_evaluate_checkpoint_device(kifmm_handle h, const double *d_src_den, double *d_trg_val, double *d_ue_out,
                                     void *stream);
int kifmm_field_checkpoint_device(kifmm_handle h, const double *d_src_den, const double *d_ue_in, double *d_trg_grad,
                                  void *stream);

/* FP64 direct sum at nchk device-resident targets with the user kernel
 * (Dense3d_MPI::evaluate, dense3d_mpi.cpp:75-141). */
int kifmm_direct_device(kifmm_handle h, const double *d_src_den, int nchk, const double *d_chk_pos,
                        double *d_chk_val, void *stream);

/* kifmm_direct_device that allocates nothing, for callers that cannot call
 * hipMalloc (code under graph capture, serving loops, harnesses that forbid
 * it).  It has no counterpart in the reference.  The partial sums go to a
 * fixed scratch array the library holds on each GPU (2^21 doubles, 16 MiB,
 * kept until the process ends); nchk * trgDOF above 2^21 is rejected with
 * KIFMM_ERR_ARG.  Calls that overlap on one GPU share that scratch and race. */
int kifmm_direct_device_noalloc(kifmm_handle h, const double *d_src_den, int nchk, const double *d_chk_pos,
                                double *d_chk_val, void *stream);

/* FMM3d_MPI::check (fmm3d_check_mpi.cpp:34-173): relative 2-norm error of
 * trg_val against the direct sum at numchk targets sampled with `seed`
 * (all targets when numchk >= n_trg).  Host pointers. */
int kifmm_check(kifmm_handle h, const double *src_den, const double *trg_val, int numchk,
                uint64_t seed, double *rerr);

/* Tree / list export for parity tests (Let3d_MPI accessors, let3d_mpi.hpp).
 * info: n_nodes, max_depth, |U|, |V|, |W|, |X| (total entries), n_src, n_trg. */
int kifmm_tree_info(kifmm_handle h, int64_t info[8]);
int kifmm_get_tree(kifmm_handle h, int *parent, int *child, int *depth, int *path, int *tag,
                   int *src_beg, int *src_num, int *trg_beg, int *trg_num, int *src_perm,
                   int *trg_perm);
/* kind: 0 U, 1 V, 2 W, 3 X (Node::Unodes()..Xnodes()); off has n_nodes+1 entries. */
int kifmm_get_list(kifmm_handle h, int kind, int64_t *off, int *lst);

/* Where kifmm_setup builds the octree and lists (the tree part of
 * Let3d_MPI::setup, let3d_mpi.cpp:611-726, 1209-1876): on_device != 0 (the
 * default) on the GPU, 0 on the host.  Both give the same tree, bit for bit;
 * kifmm_setup_host always uses the host. */
int kifmm_set_tree_builder(kifmm_handle h, int on_device);
/* Wall times of the last setup in milliseconds (FMM3d_MPI::setup,
 * fmm3d_setup_mpi.cpp:41-565, has PETSc events instead): tree and lists
 * (total), of which the device build and its copy back to the host (device
 * builder), operator tables, task tables, upload.  ms has KIFMM_NSETUP slots. */
#define KIFMM_NSETUP 6
int kifmm_setup_times(kifmm_handle h, double *ms, int n);

/* Mixed-precision translations (M2M, matmgnt3d_mpi.cpp:151-193, and L2L,
 * :231-274; the reference applies them as FP64 GEMVs): terms = 0 (default)
 * runs them on the FP64 matrix cores with M2M fused into the upward
 * pseudo-inverse; 3 or 6 runs them as split-BF16 products (BF16x3, BF16x6) on
 * the BF16 matrix cores, accumulated in FP32, with the equivalent densities
 * scaled by their level's 2^(l deg) before the split.  The pseudo-inverse
 * applications stay FP64 either way.  No effect on the FP64 path.  Takes
 * effect at the next kifmm_setup (the handle needs setup again). */
int kifmm_set_translation(kifmm_handle h, int terms);

/* Autotuner of the mixed path's launch shapes (no reference counterpart; the
 * reference has no launch shapes to choose).  Times each choice on this
 * handle's own tree with the caller's device densities (nrhs right-hand sides,
 * as kifmm_evaluate_multi_device; d_trg_val receives values) and keeps the
 * fastest, one choice at a time, in the order 3 (vector kernels only), 2, 1, 0:
 *   knob 2: translations, 0 (FP64 matrix cores, fused M2M) or 6 (split BF16x6),
 *           timed over the whole evaluate;
 *   knob 1: staging weight of the near-field chunk split, index 0..2 of
 *           {0.04, 0.15, 0.30}, timed on the near-field phase;
 *   knob 0: near-field launch, 0 one workgroup per chunk in leaf order, 1 the
 *           same largest chunk first, 2 a persistent grid over a global work
 *           queue, largest chunk first; timed on the near-field phase;
 *   knob 3: V-list Hadamard of the vector kernels (Stokes, Navier; not tuned
 *           for Laplace), 0 translation spectra read per pair, 1 and 2 the halo
 *           sweep over tiles of 1 or 2 sibling groups; timed on the V list.
 * Synchronous.  No-op (success) for the FP64 path, whose summation order is
 * fixed.  kifmm_get_tuning / kifmm_set_tuning read and apply the knobs
 * (KIFMM_NTUNE ints, index = knob); kifmm_tuning_report fills rows of
 * (knob, value, ms) for every trial of the last autotune and returns their
 * number (rows may be NULL to count only). */
#define KIFMM_NTUNE 4
int kifmm_autotune(kifmm_handle h, const double *d_src_den, double *d_trg_val, int nrhs, void *stream);
int kifmm_get_tuning(kifmm_handle h, int *knobs, int n);
int kifmm_set_tuning(kifmm_handle h, const int *knobs, int n);
int kifmm_tuning_report(kifmm_handle h, double *rows, int maxrows);

/* Operator tables of MatMgnt3d_MPI (matmgnt3d_mpi.cpp:107-431), column-major
 * FP64: which 0 = UC->UE pinv, 1 = DC->DE pinv, 2 = 8 M2M, 3 = 8 L2L,
 * 4 = M2L spectra [343][sd*td][H] complex.  With out == NULL only *count
 * (number of doubles) is returned. */
int kifmm_get_operator(kifmm_handle h, int which, double *out, int64_t *count);

/* Per-phase device time of the last evaluate (PETSc event names of
 * manage_petsc_events.cpp:77-143) when timing is on; ms has KIFMM_NPHASE slots.
 * The last slot, "VListHadamard", is the V-list Hadamard kernel alone, an
 * interval inside "EvlVList". */
#define KIFMM_NPHASE 8
int kifmm_set_timing(kifmm_handle h, int on);
/* Graph replay of kifmm_evaluate_device (no reference counterpart; for callers
 * that evaluate repeatedly from the same device buffers): with on != 0, the
 * first call with a given (d_src_den, d_trg_val, stream) on a non-default stream
 * captures the whole pass into a hipGraph and later calls with the same triple
 * replay it.  Other calls, the default stream and timing mode run eagerly. */
int kifmm_set_graph(kifmm_handle h, int on);
int kifmm_phase_times(kifmm_handle h, float *ms, int n);
const char *kifmm_phase_name(int phase);
/* Work counts from the built tree, used for roofline figures: near pairs, exact near
 * pairs, S2M pairs, S2L pairs, V pairs, #V sources, #V targets, #V groups,
 * UC->UE columns, M2M columns, DC->DE columns, L2L columns. */
#define KIFMM_NWORK 12
int kifmm_work_counts(kifmm_handle h, double *out, int n);

/* PetscOptionsGet* over an options file (%/# comments, case-insensitive
 * names): 1 if found (value copied, "" for a flag), 0 if absent, <0 on error. */
int kifmm_options_lookup(const char *options_path, const char *name, char *value, int len);

/* Last error message of the handle; with h == NULL, of the last failed create
 * on this thread. */
const char *kifmm_last_error(kifmm_handle h);

#ifdef __cplusplus
}
#endif
#endif
