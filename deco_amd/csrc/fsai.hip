// Nonsymmetric static-pattern FSAI on the device (precond/FSAI/*).
//
// Setup (FSAI.hpp:14-37, FSAI.tpp:11-32, numeric.tpp:60-133):
//   1. filtered copy of A: filter_proper(tol_rel), filter(tol_abs) (FSAI.hpp:22-28)
//   2. pattern: tril(A_f) (power 1) or tril(pattern(A_f^2)) (power 2, device SpGEMM)
//   3. per row i with P = pattern minus i: one group of G lanes solves, in LDS,
//        A[P,P] fU = -A[P,i]   and   A[P,P]^T fL = -A[i,P]^T
//      by factoring M = A[P,P]^T with partial pivoting (the matrix LAPACK sees in
//      LU_main.cpp:41-43) and solving M^T fU = bU ('T') and M fL = bL ('N'). Lane t owns column
//      t of M (= row P[t] of A, gathered with a merge walk). Every update is done in the same
//      order as dgetf2/dgetrs, so the factors match the CPU restatement.
//   4. unit diagonal, d_i = (FL A FU^T)_ii, FL *= sign(d)/sqrt|d|, FU *= 1/sqrt|d|
//      (diagonal.cpp:7-73), FUt = FU^T (device transpose, numeric.tpp:114).
// Apply (FSAI.hpp:43-44): temp = FL x ; y = FUt temp.
#include "prec.hpp"
#include <climits>

namespace deco {

deco_csr* spgemm_tril_pattern(const deco_csr* A);
void csr_filter(deco_csr* A, double tol);
void csr_filter_proper(deco_csr* A, double tol);

__device__ __forceinline__ void group_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------- tril pattern ----------
__global__ void k_tril_count(int n, const int* __restrict__ r, const int* __restrict__ c, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k = 0;
  for (int j = r[i]; j < r[i + 1]; ++j) {
    if (c[j] <= i) ++k; else break;   // symbolic.tpp:204-207
  }
  cnt[i] = k;
}
__global__ void k_tril_fill(int n, const int* __restrict__ r, const int* __restrict__ c, const int* __restrict__ rf,
                            int* __restrict__ cf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int o = rf[i];
  for (int j = r[i]; j < r[i + 1]; ++j) {
    if (c[j] <= i) cf[o++] = c[j]; else break;
  }
}

// ---------------------------------------------------------------- binning ---------------
__global__ void k_fsai_bin(int n, const int* __restrict__ rf, int* __restrict__ cnt, int* __restrict__ lists,
                           int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = (i < n) ? rf[i + 1] - rf[i] - 1 : 0;
  int b = -1;
  if (k <= 0) b = -1;
  else if (k <= 8) b = 0;
  else if (k <= 16) b = 1;
  else if (k <= 32) b = 2;
  else if (k <= 64) b = 3;
  else atomicMax(bad, k);
  wave_bin_append<4>(b, i, cnt, lists, (size_t)n);
}

// ---------------------------------------------------------------- batched LU ------------
// G lanes per row system (G in 8..64), GPB groups per block; LDS per block <= 32 KiB.
template <int G>
constexpr int fsai_gpb() { return G == 64 ? 1 : G == 32 ? 4 : G == 16 ? 16 : 32; }

template <int G>
__global__ __launch_bounds__(G * fsai_gpb<G>()) void k_fsai_numeric(const int* __restrict__ list, int nrows,
                                                     const int* __restrict__ ra, const int* __restrict__ ca,
                                                     const double* __restrict__ va, const int* __restrict__ rf,
                                                     const int* __restrict__ cf, double* __restrict__ vl,
                                                     double* __restrict__ vu, int* __restrict__ singular) {
  constexpr int GPB = fsai_gpb<G>();
  __shared__ double sM[GPB][G * G];     // column-major: M(a, b) = sM[b*G + a]
  __shared__ int sP[GPB][G];
  __shared__ int sPiv[GPB][G];
  __shared__ double sX[GPB][G];
  const int grp = threadIdx.x / G, t = threadIdx.x % G;
  const int idx = blockIdx.x * GPB + grp;
  const bool active = idx < nrows;
  const int i = active ? list[idx] : 0;
  const int iaF = active ? rf[i] : 0;
  const int k = active ? rf[i + 1] - 1 - iaF : 0;
  double* M = sM[grp];
  int* P = sP[grp];
  int* piv = sPiv[grp];
  double* X = sX[grp];
  if (t < k) P[t] = cf[iaF + t];
  for (int q = t; q < G * G; q += G) M[q] = 0.0;
  group_sync();
  // gather: lane t walks row P[t] of A; column t of M holds A[P[t], P[b]] at row b
  double bU = 0.0, bL = 0.0;
  if (t < k) {
    const int row = P[t];
    int b = 0;
    for (int s = ra[row]; s < ra[row + 1]; ++s) {
      const int col = ca[s];
      if (col == i) bU = -va[s];                                       // rhsU (gather.cpp:201-204)
      while (b < k && P[b] < col) ++b;
      if (b < k && P[b] == col) M[t * G + b] = va[s];
    }
  }
  // rhsL[b] = -A[i, P[b]] (gather.cpp:210-217): lanes split row i of A
  if (t < k) X[t] = 0.0;
  group_sync();
  if (active)
    for (int s = ra[i] + t; s < ra[i + 1]; s += G) {
      const int col = ca[s];
      int lo = 0, hi = k;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (P[mid] < col) lo = mid + 1; else hi = mid;
      }
      if (lo < k && P[lo] == col) X[lo] = -va[s];
    }
  group_sync();
  if (t < k) bL = X[t];
  group_sync();
  // ---- dgetf2 on M (k x k, column-major, lane t owns column t)
  int info = 0;
  for (int j = 0; j < k; ++j) {
    if (t == j) {
      int p = j;
      double mx = fabs(M[j * G + j]);
      for (int a = j + 1; a < k; ++a) {
        const double v = fabs(M[j * G + a]);
        if (v > mx) { mx = v; p = a; }
      }
      piv[j] = p;
    }
    group_sync();
    const int p = piv[j];
    const double pv = M[j * G + p];
    if (pv != 0.0) {
      if (p != j && t < k) {        // swap rows j and p of every column
        const double tmp = M[t * G + j];
        M[t * G + j] = M[t * G + p];
        M[t * G + p] = tmp;
      }
      group_sync();
      if (t == j) {
        const double pivot = M[j * G + j];
        if (fabs(pivot) >= 2.2250738585072014e-308) {
          const double rp = 1.0 / pivot;
          for (int a = j + 1; a < k; ++a) M[j * G + a] *= rp;
        } else {
          for (int a = j + 1; a < k; ++a) M[j * G + a] /= pivot;
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
    group_sync();
    if (t > j && t < k) {           // rank-1 update of column t
      const double tt = M[t * G + j];
      if (tt != 0.0)
        for (int a = j + 1; a < k; ++a) M[t * G + a] -= M[j * G + a] * tt;
    }
    group_sync();
  }
  if (active && t == 0 && info != 0) atomicMin(singular, i);
  // ---- dgetrs 'T' : M^T fU = bU   (U^T y = b, then L^T x = y, then the pivots backwards)
  if (t < k) X[t] = bU;
  group_sync();
  for (int j = 0; j < k; ++j) {
    if (t == j) {
      double s = X[j];
      for (int a = 0; a < j; ++a) s -= M[j * G + a] * X[a];
      X[j] = s / M[j * G + j];
    }
    group_sync();
  }
  for (int j = k - 1; j >= 0; --j) {
    if (t == j) {
      double s = X[j];
      for (int a = j + 1; a < k; ++a) s -= M[j * G + a] * X[a];
      X[j] = s;
    }
    group_sync();
  }
  if (t == 0)
    for (int j = k - 1; j >= 0; --j)
      if (piv[j] != j) { const double tmp = X[j]; X[j] = X[piv[j]]; X[piv[j]] = tmp; }
  group_sync();
  if (t < k) vu[iaF + t] = X[t];
  group_sync();
  // ---- dgetrs 'N' : M fL = bL   (pivots, L y = b, U x = y), lane a keeps b[a]
  if (t < k) X[t] = bL;
  group_sync();
  if (t == 0)
    for (int j = 0; j < k; ++j)
      if (piv[j] != j) { const double tmp = X[j]; X[j] = X[piv[j]]; X[piv[j]] = tmp; }
  group_sync();
  for (int j = 0; j < k; ++j) {
    const double bj = X[j];
    if (bj != 0.0 && t > j && t < k) X[t] -= bj * M[j * G + t];
    group_sync();
  }
  for (int j = k - 1; j >= 0; --j) {
    if (t == j && X[j] != 0.0) X[j] /= M[j * G + j];
    group_sync();
    const double xj = X[j];
    if (xj != 0.0 && t < j) X[t] -= xj * M[j * G + t];
    group_sync();
  }
  if (t < k) vl[iaF + t] = X[t];
}

// ---------------------------------------------------------------- diagonal scaling ------
// d_i = (FL A FU^T)_ii with unit diagonals (diagonal.cpp:7-52), then Jacobi_scale (:57-73)
__global__ void k_fsai_unit_diag(int n, const int* __restrict__ rf, double* __restrict__ vl, double* __restrict__ vu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vl[rf[i + 1] - 1] = 1.0;
  vu[rf[i + 1] - 1] = 1.0;
}

__global__ void k_fsai_diag_scale(int n, const int* __restrict__ ra, const int* __restrict__ ca,
                                  const double* __restrict__ va, const int* __restrict__ rf,
                                  const int* __restrict__ cf, double* __restrict__ vl, double* __restrict__ vu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int aFL = rf[i], bFL = rf[i + 1];
  double s = 0.0;
  for (int j = aFL; j < bFL; ++j) {
    const double vfl = vl[j];
    if (vfl != 0.0) {
      const int col = cf[j];
      int aA = ra[col];
      const int bA = ra[col + 1];
      int aFU = aFL;
      double sAFU = 0.0;
      while (aA < bA && aFU < bFL) {
        const int x = ca[aA], y = cf[aFU];
        if (x == y) { sAFU += va[aA] * vu[aFU]; ++aA; ++aFU; }
        else if (x < y) ++aA;
        else ++aFU;
      }
      s += vfl * sAFU;
    }
  }
  const double dd = (s < 0.0) ? (-1.0 / sqrt(fabs(s))) : (1.0 / sqrt(s));
  const double da = fabs(dd);
  for (int j = aFL; j < bFL; ++j) {
    vl[j] = vl[j] * dd;
    vu[j] = vu[j] * da;
  }
}

// ---------------------------------------------------------------- FSAI object -----------
struct FSAI : deco_prec {
  deco_csr* FL = nullptr;
  deco_csr* FU = nullptr;    // lower storage of FU (kept for the transposed apply)
  deco_csr* FUt = nullptr;   // upper CSR used by apply
  deco_csr* FLt = nullptr;   // transposed FL (transposed apply), built on demand
  DBuf<double> temp;
  DBuf<float> temp32;
  int cb = 1;                // cell size of the unknowns: the factors' block shadows use cb x cb blocks

  FSAI(const deco_cpr_params& p, const deco_csr* A_in) {
    kind = FSAI_KIND;
    n = A_in->n;
    if (p.layout != DECO_LAYOUT_BLOCK && p.layout != DECO_LAYOUT_CSR) fail(DECO_ERR_ARG, "bad layout %d", p.layout);
    layout = p.layout;
    if ((p.block_size == 2 || p.block_size == 3) && n % p.block_size == 0) cb = p.block_size;
    deco_csr* A = nullptr;
    const deco_csr* Au = A_in;
    if (p.glo_tol_rel != 0.0 || p.glo_tol_abs != 0.0) {
      ProfPhase ph("filter A");                                          // FSAI.hpp:23-27
      A = csr_clone(A_in);
      if (p.glo_tol_rel != 0.0) csr_filter_proper(A, p.glo_tol_rel);
      if (p.glo_tol_abs != 0.0) csr_filter(A, p.glo_tol_abs);
      Au = A;
    }
    try {
      build(p, Au);
    } catch (...) {
      delete A;
      cleanup();
      throw;
    }
    delete A;
    temp.alloc(n);
    temp32.alloc(n);
  }
  static constexpr int FSAI_KIND = deco_prec::FSAI;

  void cleanup() {
    delete FL; delete FU; delete FUt; delete FLt;
    FL = FU = FUt = FLt = nullptr;
  }
  ~FSAI() override { cleanup(); }

  void build(const deco_cpr_params& p, const deco_csr* A) {
    ProfPhase ph_sym("symbolic");                                        // FSAI.tpp:13-15
    if (p.glo_power == 1) {
      FL = csr_new(n, n);
      DBuf<int> cnt(n + 1);
      hipLaunchKernelGGL(k_tril_count, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, A->r.p, A->c.p, cnt.p);
      DECO_CHECK_LAUNCH();
      build_row_ptr(FL, cnt.p, n);
      hipLaunchKernelGGL(k_tril_fill, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, A->r.p, A->c.p,
                         (const int*)FL->r.p, FL->c.p);
      DECO_CHECK_LAUNCH();
    } else if (p.glo_power == 2) {
      FL = spgemm_tril_pattern(A);
    } else {
      fail(DECO_ERR_ARG, "k > 2 is not implemented, k = %d", p.glo_power);
    }
    ph_sym.end();
    ProfPhase ph_num("numeric");                                         // FSAI.tpp:28-30
    // FU shares FL's pattern (FSAI.tpp:21-23)
    FU = csr_new(n, n);
    FU->nnz = FL->nnz;
    FU->r.alloc(n + 1);
    FU->c.alloc(FL->nnz);
    FU->v.alloc(FL->nnz);
    copy_d2d(FU->r.p, FL->r.p, n + 1);
    copy_d2d(FU->c.p, FL->c.p, FL->nnz);
    if (FL->nnz) {
      DECO_HIP(hipMemsetAsync(FL->v.p, 0, sizeof(double) * FL->nnz, stream()));
      DECO_HIP(hipMemsetAsync(FU->v.p, 0, sizeof(double) * FL->nnz, stream()));
    }
    // batched local systems, binned by size
    ProfPhase ph_bin("sort");
    DBuf<int> cnt(4), lists((size_t)4 * n), flags(2);
    cnt.zero();
    int init[2] = {0, INT_MAX};
    copy_h2d(flags.p, init, 2);
    hipLaunchKernelGGL(k_fsai_bin, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, (const int*)FL->r.p, cnt.p,
                       lists.p, flags.p);
    DECO_CHECK_LAUNCH();
    int hc[4], hf[2];
    copy_d2h(hc, cnt.p, 4);
    copy_d2h(hf, flags.p, 1);
    if (hf[0] > 0) fail(DECO_ERR_UNSUPPORTED, "FSAI local system of size %d exceeds 64", hf[0]);
    ph_bin.end();
    ProfPhase ph_ls("linsolve");                                         // gather + LU + scatter
    const int* L = lists.p;
    if (hc[0]) launch<8>(L, hc[0], A, flags.p + 1);
    if (hc[1]) launch<16>(L + n, hc[1], A, flags.p + 1);
    if (hc[2]) launch<32>(L + 2 * (size_t)n, hc[2], A, flags.p + 1);
    if (hc[3]) launch<64>(L + 3 * (size_t)n, hc[3], A, flags.p + 1);
    copy_d2h(hf + 1, flags.p + 1, 1);
    if (hf[1] != INT_MAX) fail(DECO_ERR_SINGULAR, "singular FSAI local system at row %d", hf[1]);
    ph_ls.end();
    ProfPhase ph_dg("diagonal");
    hipLaunchKernelGGL(k_fsai_unit_diag, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, (const int*)FL->r.p,
                       FL->v.p, FU->v.p);
    DECO_CHECK_LAUNCH();
    ph_dg.end();
    {
      ProfPhase ph("Jacobi_scale");
      hipLaunchKernelGGL(k_fsai_diag_scale, dim3(grid_for(n, 128)), dim3(128), 0, stream(), n, A->r.p, A->c.p,
                         A->v.p, (const int*)FL->r.p, (const int*)FL->c.p, FL->v.p, FU->v.p);
      DECO_CHECK_LAUNCH();
    }
    FL->invalidate();
    FU->invalidate();
    ProfPhase ph_t("Transpose");
    FUt = csr_transpose(FU);
  }

  template <int G>
  void launch(const int* list, int cnt, const deco_csr* A, int* singular) {
    constexpr int GPB = fsai_gpb<G>();
    hipLaunchKernelGGL(k_fsai_numeric<G>, dim3(grid_for(cnt, GPB)), dim3(G * GPB), 0, stream(), list, cnt, A->r.p, A->c.p,
                       A->v.p, (const int*)FL->r.p, (const int*)FL->c.p, FL->v.p, FU->v.p, singular);
    DECO_CHECK_LAUNCH();
  }

  void apply(const double* r, double* x) override {
    ProfPhase ph("FSAI apply");                                          // FSAI.hpp:42-45
    if (precision == DECO_PREC_FP32) {   // FP32 factors, FP64 vectors in and out
      f32_product(FL, r, EpiStd<EPI_SET, float, float>{temp32.p, nullptr});
      f32_product(FUt, (const float*)temp32.p, EpiStd<EPI_SET, double, float>{x, nullptr});
      return;
    }
    f64_product(FL, r, EpiStd<EPI_SET, double, double>{temp.p, nullptr});
    f64_product(FUt, (const double*)temp.p, EpiStd<EPI_SET, double, double>{x, nullptr});
  }
  // x += M r
  void apply_add(const double* r, double* x) {
    ProfPhase ph("FSAI apply");
    f64_product(FL, r, EpiStd<EPI_SET, double, double>{temp.p, nullptr});
    f64_product(FUt, (const double*)temp.p, EpiStd<EPI_ACC, double, double>{x, nullptr});
  }
  void ensure_t() {
    if (!FLt) FLt = csr_transpose(FL);
  }
  void apply_t(const double* r, double* x) override {  // (FUt FL)^T = FLt FU
    ensure_t();
    if (precision == DECO_PREC_FP32) {
      f32_product(FU, r, EpiStd<EPI_SET, float, float>{temp32.p, nullptr});
      f32_product(FLt, (const float*)temp32.p, EpiStd<EPI_SET, double, float>{x, nullptr});
      return;
    }
    f64_product(FU, r, EpiStd<EPI_SET, double, double>{temp.p, nullptr});
    f64_product(FLt, (const double*)temp.p, EpiStd<EPI_SET, double, double>{x, nullptr});
  }
  void apply_t_add(const double* r, double* x) {
    ensure_t();
    f64_product(FU, r, EpiStd<EPI_SET, double, double>{temp.p, nullptr});
    f64_product(FLt, (const double*)temp.p, EpiStd<EPI_ACC, double, double>{x, nullptr});
  }
  // y = F x with FP64 factor values: cell blocks when selected and the unknowns come in cells,
  // else the CSR arrays
  template <class E>
  void f64_product(const deco_csr* F, const double* x, E e) {
    const deco_bsr* B = cb > 1 && blocks64() ? F->block64(cb, cb) : nullptr;
    if (B) bsr_stream_d(B, x, e);
    else spmv_stream_f(F, F->v.p, x, e);
  }
  // y = F x with FP32 factor values (cell blocks when the unknowns come in cells), epilogue e
  template <class XT, class E>
  void f32_product(const deco_csr* F, const XT* x, E e) {
    const deco_bsr* B = cb > 1 && blocks32() ? F->block32(cb, cb) : nullptr;
    if (B) bsr_stream_f(B, x, e);
    else spmv_stream_f(F, const_cast<deco_csr*>(F)->values_f32(), x, e);
  }
  // FP32 factors; input of any precision, output f32 (or f64 with EPI_ADDZ from the caller)
  template <class TI>
  void apply32_in(const TI* r, float* x) {
    ProfPhase ph("FSAI apply");
    f32_product(FL, r, EpiStd<EPI_SET, float, float>{temp32.p, nullptr});
    f32_product(FUt, (const float*)temp32.p, EpiStd<EPI_SET, float, float>{x, nullptr});
  }
  // xo = z (f32) + M r, output f64 or f32
  template <class TI, class TO>
  void apply32_addz(const TI* r, const float* z, TO* xo) {
    ProfPhase ph("FSAI apply");
    f32_product(FL, r, EpiStd<EPI_SET, float, float>{temp32.p, nullptr});
    f32_product(FUt, (const float*)temp32.p, EpiStd<EPI_ADDZ, TO, float>{xo, z});
  }
  template <class TI>
  void apply32_t_in(const TI* r, float* x) {
    ProfPhase ph("FSAI apply");
    ensure_t();
    f32_product(FU, r, EpiStd<EPI_SET, float, float>{temp32.p, nullptr});
    f32_product(FLt, (const float*)temp32.p, EpiStd<EPI_SET, float, float>{x, nullptr});
  }
  template <class TI, class TO>
  void apply32_t_addz(const TI* r, const float* z, TO* xo) {
    ProfPhase ph("FSAI apply");
    ensure_t();
    f32_product(FU, r, EpiStd<EPI_SET, float, float>{temp32.p, nullptr});
    f32_product(FLt, (const float*)temp32.p, EpiStd<EPI_ADDZ, TO, float>{xo, z});
  }
  void apply32(const float* r, float* x) override { apply32_in(r, x); }
  void apply32_t(const float* r, float* x) override { apply32_t_in(r, x); }
  void apply_f32out(const double* r, float* x) override {
    if (precision != DECO_PREC_FP32) { deco_prec::apply_f32out(r, x); return; }
    apply32_in(r, x);
  }
  void apply_t_f32out(const double* r, float* x) override {
    if (precision != DECO_PREC_FP32) { deco_prec::apply_t_f32out(r, x); return; }
    apply32_t_in(r, x);
  }
};

deco_prec* make_fsai(const deco_cpr_params& p, const deco_csr* A) { return new FSAI(p, A); }

FSAI* as_fsai(deco_prec* M) { return (M && M->kind == deco_prec::FSAI) ? static_cast<FSAI*>(M) : nullptr; }
void fsai_apply_add(FSAI* F, const double* r, double* x) { F->apply_add(r, x); }
void fsai_apply_t_add(FSAI* F, const double* r, double* x) { F->apply_t_add(r, x); }
void fsai_apply32(FSAI* F, const double* r, float* x) { F->apply32_in<double>(r, x); }
void fsai_apply32_f(FSAI* F, const float* r, float* x) { F->apply32_in<float>(r, x); }
void fsai_apply32_addz(FSAI* F, const float* r, const float* z, double* x) { F->apply32_addz<float, double>(r, z, x); }
void fsai_apply32_addz(FSAI* F, const float* r, const float* z, float* x) { F->apply32_addz<float, float>(r, z, x); }
void fsai_apply32_t(FSAI* F, const double* r, float* x) { F->apply32_t_in<double>(r, x); }
void fsai_apply32_t_addz(FSAI* F, const float* r, const float* z, double* x) {
  F->apply32_t_addz<float, double>(r, z, x);
}
void fsai_apply32_t_addz(FSAI* F, const float* r, const float* z, float* x) {
  F->apply32_t_addz<float, float>(r, z, x);
}

}  // namespace deco

using namespace deco;

extern "C" {

int deco_fsai_factors(const deco_prec* M, const deco_csr** FL, const deco_csr** FUt) {
  return guard([&] {
    FSAI* F = as_fsai(const_cast<deco_prec*>(M));
    if (!F && M && M->kind == deco_prec::CPR) {
      const deco_prec* G = nullptr;
      if (deco_cpr_fsai(M, &G) != DECO_OK) fail(DECO_ERR_ARG, "not an FSAI preconditioner");
      F = as_fsai(const_cast<deco_prec*>(G));
    }
    if (!F) fail(DECO_ERR_ARG, "not an FSAI preconditioner");
    if (FL) *FL = F->FL;
    if (FUt) *FUt = F->FUt;
  });
}

int deco_fsai_create(const deco_cpr_params* p, const deco_csr* A, deco_prec** out) {
  return guard([&] {
    if (!p || !A || !out) fail(DECO_ERR_ARG, "null argument");
    if (A->n != A->m) fail(DECO_ERR_ARG, "matrix is not square");
    *out = make_fsai(*p, A);
    sync();
  });
}

}  // extern "C"
