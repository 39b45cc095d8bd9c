/*
 * CPU oracle of the deco CPR-Krylov solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * A clean-room C++/OpenMP restatement of the reference algorithms (each function cites the
 * reference file:line it follows). Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so, and only as the checker / the timed CPU baseline;
 * the product library (deco_amd/libdeco_hip.so) never calls it.
 *
 * Memory: every orc_csr / array returned by the oracle is malloc'd; release with
 * orc_csr_free / orc_free.
 */
#ifndef DECO_ORACLE_H
#define DECO_ORACLE_H
#include "deco_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_csr {
  int n, m, nnz;
  int* r;
  int* c;
  double* v;
} orc_csr;

typedef struct orc_prec orc_prec;

void orc_free(void* p);
void orc_csr_free(orc_csr* A);            /* frees the arrays (not the struct)            */
int  orc_csr_clone(const orc_csr* A, orc_csr* out);
int  orc_num_threads(void);
void orc_set_threads(int t);

/* core/alg.hpp */
int    orc_spmv(const orc_csr* A, const double* x, double* y, int bs);
double orc_dot(int n, const double* x, const double* y);
double orc_nrm2(int n, const double* x);

/* core/CSR.hpp */
int orc_block_reorder(orc_csr* A, double* b, int bs, const char* row_order, const char* col_order);
int orc_filter(orc_csr* A, double tol);
int orc_filter_proper(orc_csr* A, double tol);
int orc_local_matrix(const orc_csr* A, int bs, int p, const double* w, orc_csr* out);
int orc_trans_mat_full(const orc_csr* A, orc_csr* out);

/* core/MxM.cpp and precond/FSAI/symbolic.tpp */
int orc_spgemm(const orc_csr* A, const orc_csr* B, orc_csr* C);
int orc_fsai_pattern(const orc_csr* A, int power, orc_csr* F);

/* precond/CPR.hpp pieces (in place on A / b); weights outputs have length n */
int orc_fix_diag(orc_csr* A, double* b, int bs, unsigned char* check_diag);
int orc_quasi_impes_ge(const orc_csr* A, int bs, double* w);
int orc_quasi_impes(const orc_csr* A, int bs, double* w);
int orc_drs_weights(const orc_csr* A, int bs, double e_dd, double e_ps, unsigned char* w_drs);

/* Preconditioners */
int orc_fsai_create(const deco_cpr_params* p, const orc_csr* A, orc_prec** out);
int orc_jacobi_create(int block_size, const orc_csr* A, orc_prec** out);
int orc_dense_create(const orc_csr* A, orc_prec** out);         /* exact LU solve (tiny A) */
int orc_identity_create(int n, orc_prec** out);
/* AMG pressure stage (role of precond/hypre_AMG): from the finest matrix A0 and the imported
 * interpolation operators P_l (n_l x n_{l+1}) and C/F markers cf_l (1 = C) of levels
 * 0..nlev-1, the oracle forms R_l = P_l^T, A_{l+1} = drop(R_l (A_l P_l), drop_tol), the smoother
 * diagonals of relax_type (relax_weight for type 7) and the coarsest inverse itself. Apply = one
 * V-cycle from zero. */
int orc_amg_build(const orc_csr* A0, int nlev, const orc_csr* P, const int* const* cf, int relax_type,
                  double relax_weight, double drop_tol, orc_prec** out);
int orc_amg_num_levels(const orc_prec* M, int* nlev, int* nc);
/* copies of level `level`'s A, R and smoother diagonal (level nlev: the coarsest A only) */
int orc_amg_level(const orc_prec* M, int level, orc_csr* A, orc_csr* R, double* d);
int orc_amg_coarse_inverse(const orc_prec* M, double* inv);   /* row-major nc x nc */
/* CPR setup: decouples A and b in place (precond/CPR.hpp:42-172); the local (pressure)
 * stage is attached afterwards from the extracted pressure matrix. */
int orc_cpr_create(const deco_cpr_params* p, orc_csr* A, double* b, orc_prec** out);
int orc_cpr_pressure_matrix(const orc_prec* M, orc_csr* Ap);      /* copy of CPR::Ap */
int orc_cpr_set_local(orc_prec* M, orc_prec* local);              /* takes ownership */
int orc_cpr_left_transform(const orc_prec* M, double* W);         /* n*bs per-cell W */
int orc_fsai_factors(const orc_prec* M, orc_csr* FL, orc_csr* FUt);
int orc_prec_apply(orc_prec* M, const double* r, double* x);
int orc_prec_apply_transpose(orc_prec* M, const double* r, double* x);
void orc_prec_destroy(orc_prec* M);

/* solver/ *.tpp files; resvec (length itmax+1, may be NULL) receives the residual history. */
int orc_solve(const deco_sol_params* sp, const orc_csr* A, orc_prec* M, const double* b,
              double* x, deco_sol_info* info, double* resvec);
int orc_solve_transpose(const deco_sol_params* sp, const orc_csr* A, orc_prec* M,
                        const double* g, double* y, deco_sol_info* info);

/* LU with partial pivoting (LAPACK dgetf2/dgetrs semantics) on a column-major k x k matrix. */
int orc_lu_factor(int k, double* a, int* ipiv);
int orc_lu_solve(int k, const double* a, const int* ipiv, int trans, double* b);

#ifdef __cplusplus
}
#endif
#endif
