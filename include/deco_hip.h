/*
 * deco_hip.h -- C ABI of libdeco_hip.so, the MI355X (gfx950) implementation of deco's
 * CPR-preconditioned Krylov solve path.
 *
 * Conventions
 *   - Every entry point returns an int status: DECO_OK (0) or a negative DECO_ERR_* code.
 *     The message of the last failure on the calling thread is deco_last_error().
 *     Nothing in the library calls exit() or throws across this boundary (the reference
 *     exits on numerical failure, e.g. precond/CPR.hpp:416-419, solver/rgmres.tpp:96-99;
 *     the C++ facade under deco_amd/cpp maps the codes back to that behaviour).
 *   - Arguments named *_dev are device pointers (hipMalloc'd, or torch tensor data_ptr()).
 *     All other pointer arguments are host pointers.
 *   - Index arrays are int32 CSR with column indices sorted ascending in each row, as the
 *     reference assumes (core/CSR.hpp:3-13, precond/CPR.hpp:321,371).
 *   - All device work of the library is ordered on one HIP stream per process
 *     (deco_stream()); calls return after the work they queued has completed unless the
 *     function says otherwise.
 *   - A handle is not re-entrant; distinct handles may be used from distinct threads only
 *     if the caller serialises the library stream.
 */
#ifndef DECO_HIP_H
#define DECO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes -------- */
#define DECO_OK               0
#define DECO_ERR_ARG         -1   /* invalid argument / shape (e.g. n % bs != 0, CPR.hpp:69-72)   */
#define DECO_ERR_DEVICE      -2   /* HIP runtime failure                                           */
#define DECO_ERR_SINGULAR    -3   /* zero determinant / zero pivot (CPR.hpp:416-419,643-646)       */
#define DECO_ERR_BREAKDOWN   -4   /* Krylov breakdown / zero pivot in back-substitution            */
#define DECO_ERR_UNSUPPORTED -5   /* parameter value the library does not implement                */
#define DECO_ERR_NOMEM       -6   /* device allocation failed                                      */
#define DECO_ERR_STATE       -7   /* handle used in a state that does not allow the call           */

/* ---------------------------------------------------------------- opaque handles ------ */
typedef struct deco_csr  deco_csr;   /* device CSR matrix                         */
typedef struct deco_prec deco_prec;  /* any preconditioner: CPR, FSAI, Jacobi, AMG */
typedef struct deco_gen  deco_gen;   /* host-side synthetic reservoir system       */

/* ---------------------------------------------------------------- parameters --------- */
/* Weight types of CPR ("weights" key, precond/CPR.hpp:61-87). */
#define DECO_W_NONE           0
#define DECO_W_QUASIIMPES     1
#define DECO_W_QUASIIMPES_GE  2

/* Global-stage types ("prec.prec_glo.type", cpr.cpp:64-69). */
#define DECO_GLO_FSAI   0
#define DECO_GLO_JACOBI 1

/* Precision of the preconditioner (FP32 preconditioner inside FP64 Krylov iterations; the
 * Krylov part is always FP64). */
#define DECO_PREC_FP64  0
#define DECO_PREC_FP32  1

/* Storage the preconditioner's products read (FSAI factors, the decoupled A and its pressure
 * rows; the AMG levels are scalar and not affected), in either precision: cell-block shadows
 * (default) or the CSR arrays (the reference's format and summation order). No reference
 * counterpart (a layout choice). */
#define DECO_LAYOUT_BLOCK 0
#define DECO_LAYOUT_CSR   1

/* Solver types ("sol.type", solver/SOL.hpp:25-38). */
#define DECO_SOL_RGMRES    0
#define DECO_SOL_LGMRES    1
#define DECO_SOL_BICGSTAB  2
#define DECO_SOL_LBICGSTAB 3
#define DECO_SOL_PCG       4

/* Every key the reference reads under "prec" (precond/CPR.hpp:49-65,118,130),
 * "prec.prec_glo" (precond/FSAI/FSAI.hpp:19-20,34; precond/Jacobi/Jacobi.hpp:18) and
 * "prec.prec_loc" (precond/hypre_AMG/hypre_AMG.tpp:48-66). Fields after the first blank
 * line of each group are extensions with documented defaults. */
typedef struct deco_cpr_params {
  int    block_size;        /* prec.block_size                                  */
  int    pressure_id;       /* prec.pressure_id (must be 0: the strided products take
                               rows bs*i, as alg.hpp:55,62 and CPR.hpp:241 do)     */
  int    seq;               /* prec.seq: 0..3                                   */
  int    weights;           /* DECO_W_*                                         */
  int    use_drs;           /* prec.use_drs                                     */
  double drs_e_dd;          /* prec.drs.e_dd                                    */
  double drs_e_ps;          /* prec.drs.e_ps                                    */
  int    cpd;               /* prec.CPD                                         */
  double tol_zero;          /* prec.tol_zero                                    */

  int    glo_type;          /* DECO_GLO_*                                       */
  double glo_tol_rel;       /* prec.prec_glo.tol_rel                            */
  double glo_tol_abs;       /* prec.prec_glo.tol_abs                            */
  int    glo_power;         /* prec.prec_glo.power: 1 or 2                      */
  int    glo_block_size;    /* prec.prec_glo.block_size (Jacobi only)           */

  double loc_tol_rel;       /* prec.prec_loc.tol_rel (filter_proper of Ap)      */
  int    amg_coarse_type;   /* prec.prec_loc.amg_coarse_type: 8 PMIS, 10 HMIS   */
  int    amg_relax_type;    /* prec.prec_loc.amg_relax_type. As in the reference only a
                               positive value names a type (hypre_AMG.tpp:50-51): -1 and 0 mean
                               hypre's default, Gauss-Seidel forward down / backward up, run as
                               two-stage Gauss-Seidel (one inner step); 11/13 forward, 12 forward
                               with two inner steps, 14 backward; 7 Jacobi (weight
                               amg_relax_weight); 17 F-C-F Jacobi; 18 l1-Jacobi */
  int    amg_agg_levs;      /* prec.prec_loc.amg_agg_levs (only 0 supported)    */
  int    amg_agg_path;      /* prec.prec_loc.amg_agg_path (unused when 0 levels)*/
  double amg_A_drop_tol;    /* prec.prec_loc.amg_A_drop_tol                     */
  double amg_prol_trunc_fac;/* prec.prec_loc.amg_prol_trunc_fac                 */
  double amg_SoC_alpha;     /* prec.prec_loc.amg_SoC_alpha                      */

  int    amg_max_coarse;    /* default 300 (hypre_AMG.tpp:64)                   */
  int    amg_max_levels;    /* default 25                                       */
  int    amg_P_max_elmts;   /* default 4                                        */
  double amg_max_row_sum;   /* default 0.9                                      */
  int    precision;         /* DECO_PREC_FP64 / DECO_PREC_FP32                  */
  double amg_relax_weight;  /* weight of the Jacobi smoother (type 7); 0 = 1             */
  int    layout;            /* DECO_LAYOUT_*; 0 (BLOCK) in a zero-initialised struct */
} deco_cpr_params;

/* "sol" keys (solver/rgmres.tpp:13-16, bicgstab.tpp:16-22). */
typedef struct deco_sol_params {
  int    type;              /* DECO_SOL_*          */
  double reltol;            /* sol.reltol          */
  int    restart;           /* sol.restart (GMRES) */
  int    itmax;             /* sol.itmax           */
  int    verbosity;         /* sol.verbosity       */
  int    use_x0;            /* 1: x_dev holds the initial guess (rgmres.tpp:42-47) */
  int    ortho;             /* GMRES Gram-Schmidt: DECO_ORTHO_DCGS2 (default) or DECO_ORTHO_MGS, the
                               reference's modified Gram-Schmidt order (rgmres.tpp:70-78); no
                               counterpart key in the reference's "sol" tree */
  int    flexible;          /* right GMRES with the default Gram-Schmidt: 1 keeps the directions
                               M^{-1} u_j (flexible GMRES) also with an FP64 preconditioner, so no
                               final M^{-1}(V y) application; 0 (default) keeps the form of
                               rgmres.tpp:106-118 there. An FP32 preconditioner always runs
                               flexible. No counterpart key in the reference's "sol" tree */
} deco_sol_params;

#define DECO_ORTHO_DCGS2 0   /* one reduction per step, delayed reorthogonalisation */
#define DECO_ORTHO_MGS   1

/* Results of a solve: the members SOL keeps (solver/SOL.hpp:10-13) plus timings. */
typedef struct deco_sol_info {
  int    iter;              /* SOL::iter  (BiCGSTAB: (int) of the half-step count, bicgstab.tpp:81) */
  double res_abs;           /* SOL::res_abs */
  double res_rel;           /* SOL::res_rel */
  int    flag;              /* SOL::flag: 0 converged, 1 not converged */
  double bnorm;             /* norm the tolerance is relative to */
  double time_ms;           /* device-timed solve (hipEvents on the library stream) */
  int    n_prec;            /* preconditioner applications */
  int    n_spmv;            /* outer SpMVs with A */
} deco_sol_info;

/* ---------------------------------------------------------------- runtime ------------- */
int         deco_init(int device);                 /* select the device, create the stream */
const char* deco_last_error(void);
int         deco_sync(void);                       /* wait for the library stream          */
void*       deco_stream(void);                     /* the hipStream_t used by the library   */
int         deco_malloc(size_t bytes, void** out_dev);
int         deco_free(void* dev);
int         deco_memcpy_h2d(void* dst_dev, const void* src, size_t bytes);
int         deco_memcpy_d2h(void* dst, const void* src_dev, size_t bytes);
int         deco_memcpy_d2d(void* dst_dev, const void* src_dev, size_t bytes);
int         deco_memset(void* dst_dev, int value, size_t bytes);

/* ---------------------------------------------------------------- device CSR ---------- */
/* CSR<R,C,V>(n,m,nz,r,c,v) (core/CSR.hpp:57-59): copies host arrays to the device. */
int deco_csr_create(int n, int m, int nnz, const int* r, const int* c, const double* v,
                    deco_csr** out);
/* Same, from device arrays (copied). */
int deco_csr_create_dev(int n, int m, int nnz, const int* r_dev, const int* c_dev,
                        const double* v_dev, deco_csr** out);
int deco_csr_destroy(deco_csr* A);                 /* ~CSR (core/CSR.hpp:214-221)          */
int deco_csr_info(const deco_csr* A, int* n, int* m, int* nnz);
int deco_csr_download(const deco_csr* A, int* r, int* c, double* v);   /* any may be NULL */
int deco_csr_device_ptrs(const deco_csr* A, int** r_dev, int** c_dev, double** v_dev);
int deco_csr_update_values(deco_csr* A, const double* v);              /* host values      */
int deco_csr_copy(const deco_csr* A, deco_csr** out);                  /* CSR::copy :138   */
/* CSR::filter(tol) (core/CSR.hpp:423-473): keep |a| > tol or the diagonal; in place. */
int deco_csr_filter(deco_csr* A, double tol);
/* CSR::filter_proper(tol) (core/CSR.hpp:476-525): keep |a| > |a_ii*tol| or the diagonal. */
int deco_csr_filter_proper(deco_csr* A, double tol);
/* CSR::local_matrix(bs,p,w) (core/CSR.hpp:338-420): rows bs*i+p, columns c%bs==p -> c/bs;
 * w_dev (length n) divides the kept values when not NULL. */
int deco_csr_local_matrix(const deco_csr* A, int bs, int p, const double* w_dev, deco_csr** out);
/* In-cell reordering of the unknowns, the step before the solve (AD::block_reorder,
 * adapter/csr2coo.cpp:19-88, called at cpr.cpp:41). row_order / col_order are the
 * "prec.row_order" / "prec.col_order" digit strings (1-based, e.g. "213"): in every cell of bs
 * unknowns, position i receives the unknown numbered by the i-th digit. Rows and columns of A
 * are permuted, rows stay sorted by column, and b_dev (length n, may be NULL) is permuted like
 * the rows. In place. bs 1, and bs 2 with both orders "12", leave A unchanged as the reference
 * does. */
int deco_block_reorder(deco_csr* A, int bs, const char* row_order, const char* col_order, double* b_dev);
/* CSR::trans_mat_full (core/CSR.hpp:151-211) as an out-of-place transpose (rows sorted). */
int deco_csr_transpose(const deco_csr* A, deco_csr** out);
/* CSR::apply(x,b,bs) -> alg::spmv (core/alg.hpp:51-64): y = A x for bs == 1; for bs > 1
 * y[i] = A[bs*i,:] x for i < n/bs (the strided pressure-row product, CPR.hpp:241). */
int deco_spmv(const deco_csr* A, const double* x_dev, double* y_dev, int bs);
/* y = z - A x (the residual loops of solver/rgmres.tpp:114-118 and CPR.hpp:258-262). */
int deco_spmv_residual(const deco_csr* A, const double* x_dev, const double* z_dev, double* y_dev);
/* y = A^T x through a cached device transpose (adjoint path; no reference counterpart). */
int deco_spmv_transpose(deco_csr* A, const double* x_dev, double* y_dev);
/* FP32-value shadow of A (created on first use) times an FP32 vector. */
int deco_spmv_f32(deco_csr* A, const float* x_dev, float* y_dev);
/* The same product through the FP32 block shadow of A with R x C blocks (R, C in 1..3; rows
 * and columns of A in groups of R and C), the form the FP32 preconditioner uses for the
 * cell-structured operators of CPR.hpp (decoupled A, its pressure rows, the FSAI factors).
 * No reference counterpart (the reference's FP64 CSR product is alg::spmv). */
int deco_spmv_block_f32(deco_csr* A, int R, int C, const float* x_dev, float* y_dev);
/* y = A x through the same block shadow with FP64 values: the form of the solvers' product with
 * the decoupled CPR matrix (CSR::apply, core/CSR.hpp:132-136, via alg::spmv, core/alg.hpp:52-63).
 * Each row's sum runs over its blocks in column order, two products per block. */
int deco_spmv_block(deco_csr* A, int R, int C, const double* x_dev, double* y_dev);
/* Read bandwidth of the current GPU in GB/s: a streaming read of `bytes` (at least 1 MiB) with
 * 16-byte non-temporal loads, 64 workgroups per CU and one partial sum per workgroup, `iters`
 * back-to-back launches timed with hipEvents (reads well past the 256 MB cache measure HBM). The reference rate for the products' read-dominated traffic; no reference
 * counterpart. */
int deco_read_bandwidth(size_t bytes, int iters, double* gbps);
/* Layout of the block shadows made from now on: the sliced one (64 block rows per slice, one
 * lane per block row) when it pads the blocks by at most max_fraction (default 0.10), the
 * chunked block CSR otherwise; a negative value selects the chunked layout always. No reference
 * counterpart (a layout choice of this library). */
int deco_set_sell_padding(double max_fraction);
/* Whether sliced block shadows made from now on store their block columns as 16-bit offsets
 * from the block row (default 1; taken when every offset fits and the block rows do not
 * outnumber the block columns, else 32-bit columns), and whether the stream products of the
 * AMG levels read their columns as 16-bit offsets from each row chunk's smallest column (taken
 * when every chunk spans at most 65535 columns). No reference counterpart. */
int deco_set_sell_offsets16(int enable);
/* Whether block shadows made from now on take the uniform-offset form of the sliced layout when it
 * applies (default 1): at most 16 distinct block-column offsets (block column - block row, as on
 * stencil grids) and zero blocks within the padding threshold; entry k of every block row is the
 * block at its offset k, so no column is stored. No reference counterpart. */
int deco_set_sell_dia(int enable);
/* Whether the stream products of matrices without a block shadow (the AMG levels below the
 * finest, their restriction and interpolation operators) made from now on use the row-piece
 * sliced form (default 1): rows cut into pieces of 1, 2, 4 or 8 entries, 64 pieces per
 * wavefront, a row's pieces summed in a segmented shuffle tree; taken for matrices of at most
 * 2.8M entries (latency-bound products) when its slots stay within 1.3 times the entries, else
 * the CSR-stream kernel. No reference counterpart. */
int deco_set_stream_pieces(int enable);
/* FP32 products on uniform-offset sliced block shadows with two 64-row slices per wavefront (the
 * loads of both slices' batches issued together): 0 never, 1 when a batch holds at most 8
 * values per lane (default: the scalar rows of the pressure level), 2 always. The sums are the same bit for bit. No
 * reference counterpart. */
int deco_set_sell_pairs(int mode);
/* When the CSR-stream products (those of the matrices above, and the plain CSR products with a
 * matrix's own values, such as deco_spmv) read the column-sorted chunk form: each stream
 * chunk's entries stored by ascending column with the slot of their row-order position, so that
 * the x gathers of one wavefront fall on a few cache lines and the products still sum in row order
 * (the same results bit for bit). The form applies when every chunk's columns span fewer than 2^20
 * and no row exceeds half the chunk tile. mode 1 (default): wherever the row-order form would
 * store 32-bit columns (the same bytes), and in place of 16-bit columns for products with FP32
 * values whose row-order gathers touch at least 9 distinct 64-byte segments of x per wavefront on
 * average; mode 2: wherever it applies; mode 0: never. The row-piece form goes first where it applies. No reference
 * counterpart. */
int deco_set_stream_sorted(int mode);
/* The R x C block shadow of A (created if needed): its layout (0 chunked, 1 sliced, 2 sliced with
 * 16-bit block-column offsets, 3 sliced with uniform offsets), its number of blocks, and the
 * blocks it stores (padding included). No reference counterpart. */
int deco_block_info(deco_csr* A, int R, int C, int* sliced, int* nnzb, long long* entries);
/* deco_spmv_time for the block-shadow product (precision as there); nnzb (may be NULL)
 * receives the number of R x C blocks. */
int deco_spmv_block_time(deco_csr* A, int R, int C, int precision, int iters, double* avg_ms, int* nnzb);
/* Times `iters` back-to-back SpMVs on the library stream with hipEvents and returns the
 * average kernel duration in milliseconds. precision: DECO_PREC_FP64 (f64 values/vectors)
 * or DECO_PREC_FP32 (f32 values/vectors). */
int deco_spmv_time(deco_csr* A, int precision, int iters, double* avg_ms);

/* ---------------------------------------------------------------- SpGEMM -------------- */
/* MxM (core/MxM.cpp:113-146): C = A*B, structural pattern, columns sorted ascending. */
int deco_spgemm(const deco_csr* A, const deco_csr* B, deco_csr** C);
/* FSAI power-2 symbolic pattern (precond/FSAI/symbolic.tpp:149-172): rows of
 * tril(pattern(A*A)), values zero. */
int deco_spgemm_tril_pattern(const deco_csr* A, deco_csr** F);

/* ---------------------------------------------------------------- BLAS-1 -------------- */
/* Deterministic fixed-grid reductions (alg::dot core/alg.hpp:278-307, alg::l2norm :334-373). */
int deco_dot(int n, const double* x_dev, const double* y_dev, double* result);
int deco_nrm2(int n, const double* x_dev, double* result);
/* deco_dot without allocating: no reference counterpart; for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The same kernels, grid and result as deco_dot, with the block partials and the
 * value in a file-scope device array of 8448 bytes per GPU (any n: the reduction grid is capped
 * at 1024 partials), looked up once per device; the value reaches the host through the pinned
 * buffer deco_init allocates. deco_init must have run (DECO_ERR_STATE otherwise). Calls that
 * overlap on one GPU share the array and race. */
int deco_dot_noalloc(int n, const double* x_dev, const double* y_dev, double* result);
/* y = b*y + a*x (alg::axpby core/alg.hpp:185-190). */
int deco_axpby(int n, double a, const double* x_dev, double b, double* y_dev);

/* ---------------------------------------------------------------- preconditioners ----- */
/* CPR<R,C,V,FSAI|Jacobi,AMG>(param, A, b) (precond/CPR.hpp:42-172). As in the reference the
 * setup decouples A and b IN PLACE (b_dev may be NULL). The AMG hierarchy of the pressure
 * matrix is built by this library on the host (hypre is not available) and uploaded once;
 * the Galerkin products run on the device. */
int deco_cpr_create(const deco_cpr_params* p, deco_csr* A, double* b_dev, deco_prec** out);
/* FSAI(param, A) (precond/FSAI/FSAI.hpp:14-37). */
int deco_fsai_create(const deco_cpr_params* p, const deco_csr* A, deco_prec** out);
/* Jacobi(param, A) (precond/Jacobi/Jacobi.hpp:15-25); block_size from glo_block_size. */
int deco_jacobi_create(const deco_cpr_params* p, const deco_csr* A, deco_prec** out);
/* hypre_AMG(param, Ap) (precond/hypre_AMG/hypre_AMG.hpp:16-21): BoomerAMG-like setup
 * (strength, HMIS/PMIS, extended+i interpolation, truncation, device RAP); apply = one
 * V-cycle from a zero initial guess (hypre_AMG.hpp:25-37). */
int deco_amg_create(const deco_cpr_params* p, const deco_csr* Ap, deco_prec** out);
/* Identity (the reference's dummy(), precond/CPR.hpp:301-305). */
int deco_identity_create(int n, deco_prec** out);
int deco_prec_destroy(deco_prec* M);
/* apply(r, x): x = M^{-1} r (CPRBase::apply, precond/CPR.hpp:3-8). */
int deco_prec_apply(deco_prec* M, const double* r_dev, double* x_dev);
/* x = M^{-T} r: the transposed operator (adjoint path; no reference counterpart). */
int deco_prec_apply_transpose(deco_prec* M, const double* r_dev, double* x_dev);
int deco_prec_size(const deco_prec* M, int* n);
int deco_prec_set_precision(deco_prec* M, int precision);
/* Storage of the preconditioner's products from now on (DECO_LAYOUT_*); block shadows are
 * built on first use. No reference counterpart. */
int deco_prec_set_layout(deco_prec* M, int layout);

/* Introspection used by the parity tests. */
int deco_cpr_pressure_matrix(const deco_prec* M, const deco_csr** Ap);   /* CPR::Ap         */
int deco_cpr_fsai(const deco_prec* M, const deco_prec** F);              /* CPR::Fglo       */
int deco_cpr_amg(const deco_prec* M, const deco_prec** G);               /* CPR::Floc       */
/* Per-cell left transform W (bs*bs row-major per cell, length n*bs) such that the decoupled
 * system is (W A) x = W b. Used by the adjoint: lambda = W^T mu. */
int deco_cpr_left_transform(const deco_prec* M, double* W);
/* FSAI factors: FL (lower CSR) and FUt (upper CSR, the transposed FU, FSAI.hpp:43-44). */
int deco_fsai_factors(const deco_prec* M, const deco_csr** FL, const deco_csr** FUt);
/* AMG hierarchy export, level by level (A_l, P_l, smoother diagonal). */
int deco_amg_num_levels(const deco_prec* M, int* levels);
int deco_amg_level(const deco_prec* M, int level, const deco_csr** A, const deco_csr** P,
                   const deco_csr** R, const double** smoother_dev, const int** cf_dev);
/* Dense FP64 inverse of the coarsest matrix (row-major nc*nc) and its size. */
int deco_amg_coarse_inverse(const deco_prec* M, int* nc, const double** inv_dev);
/* Rows of the AMG's dense tail operator (0 when the hierarchy has none): the last level's
 * sub-cycle formed as one dense matrix at setup. No reference counterpart. */
int deco_amg_dense_tail(const deco_prec* M, int* rows);
/* AMG hierarchies set up from now on form the whole sub-cycle of their highest level of at most
 * max_rows rows (its smoothing, residual, restriction, the levels below it with the coarsest solve,
 * prolongation and smoothing) as one dense operator, applied by one product per cycle (default
 * 4608 rows, at most 8192; 0 keeps the level-by-level cycle; the F-C-F smoother 17 always does).
 * Forming it costs O(n^2 n_c) device flops and 24 n^2 bytes of device memory (FP64 and FP32,
 * forward and transposed). Equal to the level-by-level sub-cycle up to rounding. No reference
 * counterpart (hypre's cycle runs level by level). */
int deco_set_amg_dense_tail(int max_rows);
/* Host-only: one level of the AMG setup (strength, C/F splitting, extended+i interpolation and
 * truncation) on host CSR arrays, as deco_amg_create runs it per level. cf[i] = 1 (C) / -1 (F);
 * P (n x nc) is returned in malloc'd arrays released with deco_host_free. No device is used. */
int deco_amg_interp_host(const deco_cpr_params* p, int n, const int* r, const int* c, const double* v,
                         int* cf, int* nc, int* p_nnz, int** P_r, int** P_c, double** P_v);
void deco_host_free(void* ptr);

/* ---------------------------------------------------------------- solvers ------------- */
/* SOL(param, A, precond, b, x) (solver/SOL.hpp:19-39): runs rgmres/lgmres/bicgstab/
 * lbicgstab/pcg on the device. x_dev receives the solution (zero initial guess unless
 * use_x0). M may be NULL (identity). */
int deco_solve(const deco_sol_params* sp, const deco_csr* A, deco_prec* M,
               const double* b_dev, double* x_dev, deco_sol_info* info);
/* Phase timer tree in the reference's format (PROFILE, core/profiler.hpp:1-178): off by default.
 * When enabled, the library times its phases under the reference's labels ("decoupling",
 * "quasiimpes weights", "IMPES", "CPD decoup", "CPD apply", "filter A", "get local matrix",
 * "Global precond" with "symbolic" / "numeric", "Local  precond", "FSAI apply", "AMG apply",
 * "CPR_rest", ...) and waits for the device at every boundary, so profiled runs are slower.
 * Callers add their own phases with start/stop (cpr.cpp:28,57-58,75). report writes the tree as
 * PROFILE::report prints it (needed = bytes including the terminator). */
int deco_profile_enable(int on);
int deco_profile_reset(void);
int deco_profile_start(const char* name);
int deco_profile_stop(void);
int deco_profile_time(const char* name, double* seconds);
int deco_profile_report(char* buf, size_t len, size_t* needed);
/* Algorithmic byte count of the device work queued from now on (on != 0; 0 stops counting; both
 * reset the count): every product adds the bytes it must move at least once in the storage it
 * reads -- stored values and column indices (16-bit where the product reads 16-bit offsets), row
 * pointers or slice descriptors, one pass over x and over every vector its epilogue reads or
 * writes -- and every vector kernel, copy and fill the vectors it reads and writes once. The
 * strided products of deco_spmv (bs > 1) are not counted. Divided by a solve's time it gives the
 * solve's achieved algorithmic bandwidth. No reference counterpart. */
int deco_count_bytes(int on);
int deco_counted_bytes(double* bytes, long long* launches);
/* Event timing of the solvers' operator products with the system matrix, the products that form
 * the Krylov directions (on != 0 starts, 0 stops; both reset; the residual products b - A x are
 * not timed): a pair of events on the library stream around each product, so the average
 * in-solve duration of the roofline kernel can be read beside its back-to-back time. total_ms
 * sums the recorded products (waiting for the last one), count is their number. No reference
 * counterpart. */
int deco_time_outer_products(int on);
int deco_outer_product_times(double* total_ms, long long* count);
/* The solvers keep their work vectors (Krylov basis, flexible directions, partial sums) between
 * calls and grow them on demand; this frees them. No reference counterpart (the reference
 * allocates per solve, solver/rgmres.tpp:18-29). */
int deco_workspace_release(void);
/* Solves A^T y = g with the transposed preconditioner M^T (adjoint path). */
int deco_solve_transpose(const deco_sol_params* sp, deco_csr* A, deco_prec* M,
                         const double* g_dev, double* y_dev, deco_sol_info* info);
/* Sampled product on A's pattern: dA[k] = -lambda[row(k)] * x[col(k)] (dL/dA_ij). */
int deco_grad_A_pattern(const deco_csr* A, const double* lambda_dev, const double* x_dev,
                        double* dA_dev);
/* lambda = W^T mu with the per-cell transform of deco_cpr_left_transform. */
int deco_cpr_transform_transpose(const deco_prec* M, const double* mu_dev, double* lambda_dev);

/* ---------------------------------------------------------------- synthetic systems --- */
/* Seeded synthetic reservoir Jacobians (configs documented in deco_amd/csrc/gen.cpp).
 * Host only: no device is touched. cfg: 1 (32x32x16), 11 (32x32x32), 2, 3, 4 (16x16x8),
 * 5. seed 0 selects the default seed 1000+cfg. */
int deco_gen_create(int cfg, uint64_t seed, deco_gen** out);
int deco_gen_custom(int nx, int ny, int nz, int bs, uint64_t seed, deco_gen** out);
/* The physics of configuration cfg on an nx x ny x nz grid (tests at reduced size). */
int deco_gen_scaled(int cfg, int nx, int ny, int nz, uint64_t seed, deco_gen** out);
int deco_gen_info(const deco_gen* g, int* n, int* nnz, int* bs, int* ncells);
/* Borrowed host pointers valid until deco_gen_destroy. */
int deco_gen_arrays(const deco_gen* g, const int** r, const int** c, const double** v,
                    const double** b, const double** xstar);
int deco_gen_destroy(deco_gen* g);

#ifdef __cplusplus
}
#endif
#endif /* DECO_HIP_H */
