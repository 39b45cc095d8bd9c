// CPR preconditioner on the device (precond/CPR.hpp).
//
// Setup (CPR.hpp:42-172), all on the device, decoupling A and b in place:
//   fix_diag -> quasi-IMPES(-GE) weights -> row scaling -> DRS weights + unscale -> IMPES ->
//   CPD build/apply -> tol_zero filter -> pressure extraction Ap (+ filter_proper) ->
//   global stage (FSAI or Jacobi) -> local stage (AMG on Ap) -> undo fix_diag.
// One thread per cell; every arithmetic step follows the reference's operation order, so the
// decoupled matrix is identical to the CPU restatement's.
//
// Apply (CPR.hpp:176-299): sequences 0..3. In FP32 mode the global stage, the two in-CPR
// products with A (f32 shadow of the decoupled A) and the AMG run in f32 with f32 work vectors;
// the input r and the output x stay FP64.
#include "prec.hpp"

namespace deco {

deco_prec* make_fsai(const deco_cpr_params& p, const deco_csr* A);
deco_prec* make_amg(const deco_cpr_params& p, const deco_csr* Ap);
deco_csr* local_matrix(const deco_csr* A, int bs, int p, const double* w);
void csr_filter(deco_csr* A, double tol);
void csr_filter_proper(deco_csr* A, double tol);

// FSAI pieces used by CPR (defined in fsai.hip)
struct FSAI;
FSAI* as_fsai(deco_prec* M);
void fsai_apply_add(FSAI* F, const double* r, double* x);
void fsai_apply_t_add(FSAI* F, const double* r, double* x);
void fsai_apply32(FSAI* F, const double* r, float* x);
void fsai_apply32_f(FSAI* F, const float* r, float* x);
void fsai_apply32_addz(FSAI* F, const float* r, const float* z, double* x);
void fsai_apply32_addz(FSAI* F, const float* r, const float* z, float* x);
void fsai_apply32_t(FSAI* F, const double* r, float* x);
void fsai_apply32_t_addz(FSAI* F, const float* r, const float* z, double* x);
void fsai_apply32_t_addz(FSAI* F, const float* r, const float* z, float* x);

enum CprErr { ERR_DET = 1, ERR_ROWLEN = 2, ERR_NODIAG = 4 };

// every row holds its diagonal entry (the decoupling reads it by binary search in the row)
__global__ void k_diag_present(int n, const int* __restrict__ r, const int* __restrict__ c, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = r[i], hi = r[i + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (c[mid] < i) lo = mid + 1; else hi = mid;
  }
  if (lo == r[i + 1] || c[lo] != i) atomicOr(err, ERR_NODIAG);
}

__device__ __forceinline__ int bsearch_dev(const int* __restrict__ c, int lower, int upper, int target) {
  int mn = lower, mx = upper, mid = (lower + upper) >> 1;  // core::binarysearch (core/alg.hpp:16-25)
  while (mx - mn > 1) {
    if (c[mid] > target) mx = mid; else mn = mid;
    mid = (mn + mx) >> 1;
  }
  return mid;
}

// fix_diag (CPR.hpp:753-840) with small = big = 1e-20
__global__ void k_fix_diag(int sn, int bs, const int* __restrict__ r, const int* __restrict__ c, double* __restrict__ v,
                           double* __restrict__ b, unsigned char* __restrict__ flag, int* __restrict__ swapped,
                           int* __restrict__ nzero) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sn) return;
  const double small_tol = 1e-20, big_tol = 1e-20;
  const int i = s * bs;
  const int j0 = bsearch_dev(c, r[i], r[i + 1], i);
  if (fabs(v[j0]) < small_tol) {
    v[j0] = (v[j0] > 0) ? small_tol : -small_tol;
    flag[i] = 1;
    atomicAdd(nzero, 1);
  }
  if (bs == 3) {
    const int j1 = bsearch_dev(c, r[i + 1], r[i + 2], i + 1);
    const int j2 = bsearch_dev(c, r[i + 2], r[i + 3], i + 2);
    if (fabs(v[j1]) < small_tol || fabs(v[j2]) < small_tol) {
      if (fabs(v[j1 + 1]) > small_tol || fabs(v[j2 - 1]) > small_tol) {
        const int ia1 = r[i + 1], ib1 = r[i + 2];
        for (int k = 0; k < ib1 - ia1; ++k) {     // swap the values of rows i+1 and i+2
          const double t = v[ia1 + k];
          v[ia1 + k] = v[ib1 + k];
          v[ib1 + k] = t;
        }
        if (b) { const double t = b[i + 1]; b[i + 1] = b[i + 2]; b[i + 2] = t; }
        swapped[s] = 1;
      } else {
        if (fabs(v[j1]) < small_tol) { v[j1] = (v[j1] > 0) ? big_tol : -big_tol; flag[i + 1] = 1; }
        if (fabs(v[j2]) < small_tol) { v[j2] = (v[j2] > 0) ? big_tol : -big_tol; flag[i + 2] = 1; }
      }
    }
  } else if (bs == 2) {
    const int j1 = bsearch_dev(c, r[i + 1], r[i + 2], i + 1);
    if (fabs(v[j1]) < small_tol) { v[j1] = (v[j1] > 0) ? big_tol : -big_tol; flag[i + 1] = 1; }
  }
}

// quasi_IMPES_GE (CPR.hpp:372-447)
__global__ void k_qimpes_ge(int sn, int bs, const int* __restrict__ r, const int* __restrict__ c,
                            const double* __restrict__ v, double* __restrict__ w, int* __restrict__ err) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sn) return;
  const int bs1 = bs - 1;
  double Ap[4] = {0, 0, 0, 0}, bp[2] = {0, 0};
  w[s * bs] = 1.0;
  for (int p = 0; p < bs; ++p) {
    const int i = s * bs + p;
    const int j = bsearch_dev(c, r[i], r[i + 1], i);
    for (int k = 1; k < bs; ++k) {
      if (p == 0) bp[k - 1] = -v[j + k - p];
      else Ap[(p - 1) * bs1 + c[j + k - p] % bs - 1] = v[j + k - p];
    }
  }
  if (bs1 == 2) {
    const double det = Ap[0] * Ap[3] - Ap[1] * Ap[2];
    if (det == 0.0) { atomicOr(err, ERR_DET); return; }
    const double invdet = 1.0 / det;
    const double i0 = Ap[3] * invdet, i1 = -Ap[1] * invdet, i2 = -Ap[2] * invdet, i3 = Ap[0] * invdet;
    w[s * bs + 1] = i0 * bp[0] + i2 * bp[1];
    w[s * bs + 2] = i1 * bp[0] + i3 * bp[1];
  } else if (bs1 == 1) {
    if (Ap[0] == 0.0) { atomicOr(err, ERR_DET); return; }
    w[s * bs + 1] = bp[0] / Ap[0];
  }
}

// quasi_IMPES (CPR.hpp:322-369): B^T w = e_0 per cell by LU with partial pivoting (dgesv)
__global__ void k_qimpes(int sn, int bs, const int* __restrict__ r, const int* __restrict__ c,
                         const double* __restrict__ v, double* __restrict__ w) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sn) return;
  double M[9];
  int piv[3];
  for (int q = 0; q < bs * bs; ++q) M[q] = 0.0;
  for (int p = 0; p < bs; ++p) {
    const int i = s * bs + p;
    const int ia = r[i], ib = r[i + 1];
    const int j = bsearch_dev(c, ia, ib, i);
    if (c[j] == i)
      for (int k = 0; k < bs; ++k) {
        const int jj = j + k - p;
        if (jj < ib && jj >= ia && c[jj] / bs == s) M[p * bs + c[jj] % bs] = v[jj];  // read column-major: B^T
      }
  }
  double* x = w + (size_t)s * bs;
  for (int q = 0; q < bs; ++q) x[q] = (q == 0) ? 1.0 : 0.0;
  for (int j = 0; j < bs; ++j) {  // dgetf2
    int p = j;
    double mx = fabs(M[j + j * bs]);
    for (int i = j + 1; i < bs; ++i) if (fabs(M[i + j * bs]) > mx) { mx = fabs(M[i + j * bs]); p = i; }
    piv[j] = p;
    if (M[p + j * bs] != 0.0) {
      if (p != j) for (int q = 0; q < bs; ++q) { const double t = M[j + q * bs]; M[j + q * bs] = M[p + q * bs]; M[p + q * bs] = t; }
      const double pv = M[j + j * bs];
      if (fabs(pv) >= 2.2250738585072014e-308) { const double rp = 1.0 / pv; for (int i = j + 1; i < bs; ++i) M[i + j * bs] *= rp; }
      else for (int i = j + 1; i < bs; ++i) M[i + j * bs] /= pv;
    }
    for (int q = j + 1; q < bs; ++q) {
      const double t = M[j + q * bs];
      if (t != 0.0) for (int i = j + 1; i < bs; ++i) M[i + q * bs] -= M[i + j * bs] * t;
    }
  }
  for (int j = 0; j < bs; ++j) if (piv[j] != j) { const double t = x[j]; x[j] = x[piv[j]]; x[piv[j]] = t; }
  for (int j = 0; j < bs; ++j) { const double t = x[j]; if (t != 0.0) for (int i = j + 1; i < bs; ++i) x[i] -= t * M[i + j * bs]; }
  for (int j = bs - 1; j >= 0; --j)
    if (x[j] != 0.0) { x[j] /= M[j + j * bs]; const double t = x[j]; for (int i = 0; i < j; ++i) x[i] -= t * M[i + j * bs]; }
}

// alg::diag_scale (core/alg.hpp:153-177) for the matrix and the vector
__global__ void k_row_scale(int n, const int* __restrict__ r, double* __restrict__ v, const double* __restrict__ w,
                            double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = w[i];
  for (int j = r[i]; j < r[i + 1]; ++j) v[j] *= d;
  if (b) b[i] *= d;
}

// drs_weights (CPR.hpp:518-584), one thread per cell
__global__ void k_drs(int sn, int bs, const int* __restrict__ r, const int* __restrict__ c,
                      const double* __restrict__ v, double e_dd, double e_ps, unsigned char* __restrict__ wd) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sn) return;
  for (int p = 0; p < bs; ++p) {
    const int i = s * bs + p;
    const int ia = r[i], ib = r[i + 1];
    const double diag = v[bsearch_dev(c, ia, ib, i - p)];
    double off = -fabs(diag);
    for (int j = ia; j < ib; ++j) off += (c[j] % bs == 0) ? fabs(v[j]) : 0.;
    wd[i] = (diag < e_dd * off) ? 0 : 1;
  }
  if (e_ps != 0.0) {
    const int i = s * bs;
    const int ia = r[i], ib = r[i + 1];
    const int jd = bsearch_dev(c, ia, ib, i);
    const double diag = fabs(v[jd]);
    double off[8];
    for (int p = 0; p < bs; ++p) off[p] = 0.0;
    for (int j = ia; j < ib; ++j) {
      if (j == jd) continue;
      off[c[j] % bs] += fabs(v[j]);
    }
    for (int p = 0; p < bs; ++p) {
      const bool ok = !(off[p] < e_ps * diag);
      wd[i + p] = (unsigned char)(wd[i + p] && ok);
    }
  }
  wd[s * bs] = 1;
}

// unscale_drs_weights (CPR.hpp:586-600)
__global__ void k_drs_unscale(int n, const int* __restrict__ r, double* __restrict__ v, double* __restrict__ b,
                              const unsigned char* __restrict__ wd, double* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || wd[i]) return;
  const double d = w[i];
  for (int j = r[i]; j < r[i + 1]; ++j) v[j] /= d;
  if (b) b[i] /= d;
  w[i] = 1.;
}

// IMPES (CPR.hpp:475-516 matrix, :450-471 vector): pressure row += sum_k wd_k row_k, positional
__global__ void k_impes(int sn, int bs, const int* __restrict__ r, double* __restrict__ v, double* __restrict__ b,
                        const unsigned char* __restrict__ wd, int* __restrict__ err) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sn) return;
  const int ip = s * bs;
  const int len = r[ip + 1] - r[ip];
  double* table = v + r[ip];
  for (int k = 1; k < bs; ++k) {
    const int i = s * bs + k;
    const int ia = r[i], ib = r[i + 1];
    if (ib - ia != len) atomicOr(err, ERR_ROWLEN);
    const double wk = wd ? (double)wd[i] : 1.0;
    for (int j = ia; j < ib && j - ia < len; ++j) {
      if (wd) table[j - ia] += wk * v[j];
      else table[j - ia] += v[j];
    }
  }
  if (b) {
    double sum = 0.0;
    for (int k = 0; k < bs; ++k) sum += wd ? (double)wd[s * bs + k] * b[s * bs + k] : b[s * bs + k];
    b[s * bs] = sum;
  }
}

// CPD_build (CPR.hpp:603-685)
__global__ void k_cpd_build(int sn, int bs, const int* __restrict__ r, const int* __restrict__ c,
                            const double* __restrict__ v, double* __restrict__ J, int* __restrict__ err) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sn) return;
  double a[9], Dp[3];
  for (int q = 0; q < bs; ++q) {
    const int i = s * bs + q;
    const int j = bsearch_dev(c, r[i], r[i + 1], i);
    Dp[q] = v[j];
    for (int k = 0; k < bs; ++k) a[q * bs + c[j + k - q] % bs] = v[j + k - q];
  }
  double* vj = J + (size_t)s * bs * bs;
  if (bs == 3) {
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[3] * (a[1] * a[8] - a[2] * a[7]) +
                       a[6] * (a[1] * a[5] - a[2] * a[4]);
    if (det == 0.0) { atomicOr(err, ERR_DET); return; }
    const double invdet = 1.0 / det;
    vj[0] = (a[4] * a[8] - a[5] * a[7]) * invdet * Dp[0];
    vj[1] = -(a[1] * a[8] - a[2] * a[7]) * invdet * Dp[0];
    vj[2] = (a[1] * a[5] - a[2] * a[4]) * invdet * Dp[0];
    vj[3] = -(a[3] * a[8] - a[5] * a[6]) * invdet * Dp[1];
    vj[4] = (a[0] * a[8] - a[2] * a[6]) * invdet * Dp[1];
    vj[5] = -(a[0] * a[5] - a[2] * a[3]) * invdet * Dp[1];
    vj[6] = (a[3] * a[7] - a[4] * a[6]) * invdet * Dp[2];
    vj[7] = -(a[0] * a[7] - a[1] * a[6]) * invdet * Dp[2];
    vj[8] = (a[0] * a[4] - a[1] * a[3]) * invdet * Dp[2];
  } else {
    const double det = a[0] * a[3] - a[1] * a[2];
    if (det == 0.0) { atomicOr(err, ERR_DET); return; }
    const double invdet = 1.0 / det;
    vj[0] = a[3] * invdet * Dp[0];
    vj[1] = -a[1] * invdet * Dp[0];
    vj[2] = -a[2] * invdet * Dp[1];
    vj[3] = a[0] * invdet * Dp[1];
  }
}

// CPD_apply (CPR.hpp:688-750): A <- blockdiag(J) A positionally, b <- blockdiag(J) b
__global__ void k_cpd_apply(int sn, int bs, const int* __restrict__ r, double* __restrict__ v, double* __restrict__ b,
                            const double* __restrict__ J, int* __restrict__ err) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sn) return;
  const double* vj = J + (size_t)s * bs * bs;
  const int len = r[s * bs + 1] - r[s * bs];
  for (int q = 1; q < bs; ++q)
    if (r[s * bs + q + 1] - r[s * bs + q] != len) { atomicOr(err, ERR_ROWLEN); return; }
  for (int t = 0; t < len; ++t) {
    double o[3], nv[3];
    for (int k = 0; k < bs; ++k) o[k] = v[r[s * bs + k] + t];
    for (int q = 0; q < bs; ++q) {
      double acc = o[q] * vj[q * bs + q];
      for (int k = 0; k < bs; ++k) {
        if (k == q) continue;
        acc += o[k] * vj[q * bs + k];
      }
      nv[q] = acc;
    }
    for (int q = 0; q < bs; ++q) v[r[s * bs + q] + t] = nv[q];
  }
  if (b) {
    double sum[3];
    for (int q = 0; q < bs; ++q) {
      sum[q] = 0.0;
      for (int k = 0; k < bs; ++k) sum[q] += vj[q * bs + k] * b[s * bs + k];
    }
    for (int q = 0; q < bs; ++q) b[s * bs + q] = sum[q];
  }
}

// per-cell left transform W_s = J_s I_s diag(w_s) Pi_s (row-major), applied as in the CPU code
__global__ void k_left_transform(int sn, int bs, const int* __restrict__ swapped, const double* __restrict__ w,
                                 const unsigned char* __restrict__ wd, const double* __restrict__ J, int use_w,
                                 int use_j, double* __restrict__ W) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sn) return;
  double Ws[9], T[9], o[9];
  for (int q = 0; q < bs * bs; ++q) Ws[q] = 0.0;
  for (int q = 0; q < bs; ++q) Ws[q * bs + q] = 1.0;
  if (swapped[s])
    for (int q = 0; q < bs; ++q) { const double t = Ws[1 * bs + q]; Ws[1 * bs + q] = Ws[2 * bs + q]; Ws[2 * bs + q] = t; }
  auto left_mul = [&](const double* TT) {
    for (int a = 0; a < bs; ++a)
      for (int cc = 0; cc < bs; ++cc) {
        double t = 0.0;
        for (int k = 0; k < bs; ++k) t += TT[a * bs + k] * Ws[k * bs + cc];
        o[a * bs + cc] = t;
      }
    for (int q = 0; q < bs * bs; ++q) Ws[q] = o[q];
  };
  if (use_w) {
    for (int q = 0; q < bs * bs; ++q) T[q] = 0.0;
    for (int q = 0; q < bs; ++q) T[q * bs + q] = w[s * bs + q];
    left_mul(T);
  }
  for (int q = 0; q < bs * bs; ++q) T[q] = 0.0;
  for (int q = 0; q < bs; ++q) T[q * bs + q] = 1.0;
  for (int k = 0; k < bs; ++k) T[k] = wd ? (double)wd[s * bs + k] : 1.0;
  left_mul(T);
  if (use_j) left_mul(J + (size_t)s * bs * bs);
  for (int q = 0; q < bs * bs; ++q) W[(size_t)s * bs * bs + q] = Ws[q];
}

// lambda_s = W_s^T mu_s
__global__ void k_transform_t(int sn, int bs, const double* __restrict__ W, const double* __restrict__ mu,
                              double* __restrict__ lam) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sn) return;
  const double* Ws = W + (size_t)s * bs * bs;
  for (int cc = 0; cc < bs; ++cc) {
    double t = 0.0;
    for (int a = 0; a < bs; ++a) t += Ws[a * bs + cc] * mu[s * bs + a];
    lam[s * bs + cc] = t;
  }
}

__global__ void k_undo_fix(int n, const int* __restrict__ r, const int* __restrict__ c, double* __restrict__ v,
                           const unsigned char* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  v[bsearch_dev(c, r[i], r[i + 1], i)] = 0.;
}

// ---------------------------------------------------------------- apply kernels ---------
template <class TX, class TP>
__global__ void k_add_pressure(int sn, int bs, TX* __restrict__ x, const TP* __restrict__ xp) {  // x[bs i] += xp[i]
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < sn) x[(size_t)i * bs] += (TX)xp[i];
}
template <class TX, class TP>
__global__ void k_set_pressure(int sn, int bs, TX* __restrict__ x, const TP* __restrict__ xp) {   // x[bs i] = xp[i]
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < sn) x[(size_t)i * bs] = (TX)xp[i];
}
template <class TI, class TO>
__global__ void k_gather_pressure(int sn, int bs, const TI* __restrict__ r, TO* __restrict__ rt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < sn) rt[i] = (TO)r[(size_t)i * bs];
}
template <class TA, class TB, class TO = double>
__global__ void k_addz(int n, const TA* __restrict__ z, const TB* __restrict__ a, TO* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // x = z + a (z, a of any precision)
  if (i < n) x[i] = (TO)((double)z[i] + (double)a[i]);
}
__global__ void k_add_f(int n, const float* __restrict__ a, float* __restrict__ e) {   // e += a
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) e[i] = a[i] + e[i];
}
__global__ void k_add3(int n, const double* __restrict__ a, const double* __restrict__ e, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // x += a + e
  if (i < n) x[i] += a[i] + e[i];
}
// e = a + P ep: ep (one value per cell) at the pressure unknowns bs*i, 0 elsewhere
__global__ void k_add_p(int n, int bs, const float* __restrict__ a, const float* __restrict__ ep, float* __restrict__ e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) e[i] = a[i] + (i % bs == 0 ? ep[i / bs] : 0.f);
}
// x += a + P ep
__global__ void k_add3_p(int n, int bs, const double* __restrict__ a, const double* __restrict__ ep,
                         double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += a[i] + (i % bs == 0 ? ep[i / bs] : 0.0);
}

struct CPR : deco_prec {
  deco_cpr_params p{};
  deco_csr* A = nullptr;                 // caller's matrix, decoupled in place
  deco_csr* Ap = nullptr;
  deco_csr* Apr = nullptr;               // pressure rows (bs*i) of the decoupled A, copied at setup
  // pressure columns (bs*j) of the decoupled A^T, n x sn, made on the first transposed application
  // and remade when A changes (atp_gen = A->gen when made): A^T P e for a pressure correction e
  // reads only them
  deco_csr* Atp = nullptr;
  unsigned long long atp_gen = 0;
  deco_prec* glo = nullptr;
  deco_prec* loc = nullptr;
  deco_prec* jac = nullptr;
  int bs = 1, sn = 0;
  DBuf<double> W, rt, xp, tmp, tmp2, x2;
  DBuf<float> x32, rt32, xp32, t32, s32, e32;

  CPR(const deco_cpr_params& pp, deco_csr* A_, double* b) : p(pp), A(A_) {
    kind = deco_prec::CPR;
    n = A->n;
    bs = p.block_size;
    if (bs <= 0 || n % bs != 0) fail(DECO_ERR_ARG, "n is not divisible by bs, n = %d, bs = %d", n, bs);
    if (bs != 2 && bs != 3) fail(DECO_ERR_UNSUPPORTED, "unknown block size CPD decoupling (bs = %d)", bs);
    if (p.pressure_id != 0) fail(DECO_ERR_UNSUPPORTED, "pressure_id must be 0 (strided products assume it)");
    if (p.seq < 0 || p.seq > 3) fail(DECO_ERR_ARG, "unknown sequence, seq = %d", p.seq);
    if (p.layout != DECO_LAYOUT_BLOCK && p.layout != DECO_LAYOUT_CSR) fail(DECO_ERR_ARG, "bad layout %d", p.layout);
    if (p.weights != DECO_W_NONE && p.weights != DECO_W_QUASIIMPES && p.weights != DECO_W_QUASIIMPES_GE)
      fail(DECO_ERR_ARG, "Unknown weights type %d", p.weights);
    precision = p.precision;
    layout = p.layout;
    sn = n / bs;
    try {
      setup(b);
    } catch (...) {
      release();
      throw;
    }
  }

  void release() {
    delete glo; delete loc; delete jac; delete Ap; delete Apr; delete Atp;
    glo = loc = jac = nullptr;
    Ap = Apr = Atp = nullptr;
  }
  ~CPR() override { release(); }

  void check_err(const int* err_dev) {
    int e = 0;
    copy_d2h(&e, err_dev, 1);
    if (e & ERR_DET) fail(DECO_ERR_SINGULAR, "determinant is zero in CPD decoupling");
    if (e & ERR_ROWLEN) fail(DECO_ERR_ARG, "rows of a cell do not share one column pattern");
  }

  void setup(double* b) {
    const unsigned gc = grid_for(sn, 128), gr = grid_for(n, 256);
    DBuf<unsigned char> flag(n), wd(n);
    DBuf<int> swapped(sn), err(2);
    flag.zero();
    swapped.zero();
    err.zero();
    ProfPhase ph_dec("decoupling");                                    // CPR.hpp:47-138
    hipLaunchKernelGGL(k_diag_present, dim3(gr), dim3(256), 0, stream(), n, A->r.p, A->c.p, err.p);
    DECO_CHECK_LAUNCH();
    {
      int e = 0;
      copy_d2h(&e, err.p, 1);
      if (e & ERR_NODIAG) fail(DECO_ERR_ARG, "a row of A has no diagonal entry (CPR needs every diagonal)");
    }
    hipLaunchKernelGGL(k_fix_diag, dim3(gc), dim3(128), 0, stream(), sn, bs, A->r.p, A->c.p, A->v.p, b, flag.p,
                       swapped.p, err.p + 1);
    DECO_CHECK_LAUNCH();
    DBuf<double> w;
    ProfPhase ph_w("quasiimpes weights");
    if (p.weights != DECO_W_NONE) {
      w.alloc(n);
      if (p.weights == DECO_W_QUASIIMPES_GE)
        hipLaunchKernelGGL(k_qimpes_ge, dim3(gc), dim3(128), 0, stream(), sn, bs, A->r.p, A->c.p, A->v.p, w.p, err.p);
      else
        hipLaunchKernelGGL(k_qimpes, dim3(gc), dim3(128), 0, stream(), sn, bs, A->r.p, A->c.p, A->v.p, w.p);
      DECO_CHECK_LAUNCH();
      check_err(err.p);
      hipLaunchKernelGGL(k_row_scale, dim3(gr), dim3(256), 0, stream(), n, A->r.p, A->v.p, (const double*)w.p, b);
      DECO_CHECK_LAUNCH();
    }
    ph_w.end();
    if (p.use_drs) {
      ProfPhase ph("DRS weights");
      hipLaunchKernelGGL(k_drs, dim3(gc), dim3(128), 0, stream(), sn, bs, A->r.p, A->c.p, A->v.p, p.drs_e_dd,
                         p.drs_e_ps, wd.p);
      DECO_CHECK_LAUNCH();
      if (p.weights != DECO_W_NONE) {
        hipLaunchKernelGGL(k_drs_unscale, dim3(gr), dim3(256), 0, stream(), n, A->r.p, A->v.p, b,
                           (const unsigned char*)wd.p, w.p);
        DECO_CHECK_LAUNCH();
      }
    }
    const unsigned char* wdp = p.use_drs ? wd.p : nullptr;
    ProfPhase ph_impes("IMPES");
    hipLaunchKernelGGL(k_impes, dim3(gc), dim3(128), 0, stream(), sn, bs, A->r.p, A->v.p, b, wdp, err.p);
    DECO_CHECK_LAUNCH();
    check_err(err.p);
    ph_impes.end();
    DBuf<double> J;
    if (p.cpd) {
      J.alloc((size_t)sn * bs * bs);
      ProfPhase ph_cd("CPD decoup");
      hipLaunchKernelGGL(k_cpd_build, dim3(gc), dim3(128), 0, stream(), sn, bs, A->r.p, A->c.p, A->v.p, J.p, err.p);
      DECO_CHECK_LAUNCH();
      check_err(err.p);
      ph_cd.end();
      ProfPhase ph_ca("CPD apply");
      hipLaunchKernelGGL(k_cpd_apply, dim3(gc), dim3(128), 0, stream(), sn, bs, A->r.p, A->v.p, b,
                         (const double*)J.p, err.p);
      DECO_CHECK_LAUNCH();
      check_err(err.p);
    }
    W.alloc((size_t)sn * bs * bs);
    hipLaunchKernelGGL(k_left_transform, dim3(gc), dim3(128), 0, stream(), sn, bs, (const int*)swapped.p,
                       (const double*)w.p, wdp, (const double*)J.p, p.weights != DECO_W_NONE ? 1 : 0, p.cpd ? 1 : 0,
                       W.p);
    DECO_CHECK_LAUNCH();
    A->invalidate();
    if (A->At) { delete A->At; A->At = nullptr; }
    if (p.tol_zero != 0.0) {
      ProfPhase ph("filter A");
      csr_filter(A, p.tol_zero);
    }
    {
      ProfPhase ph("get local matrix");
      Ap = local_matrix(A, bs, 0, nullptr);
    }
    if (p.loc_tol_rel > 1e-20) {
      ProfPhase ph("filter Ap");
      csr_filter_proper(Ap, p.loc_tol_rel);
    }
    ph_dec.end();
    {
      ProfPhase ph("Global precond");                                 // CPR.hpp:141-150
      if (p.glo_type == DECO_GLO_FSAI) glo = make_fsai(p, A);
      else glo = new Jacobi(A, p.glo_block_size);
    }
    {
      ProfPhase ph("Local  precond");
      loc = make_amg(p, Ap);
    }
    ProfPhase ph_dec2("decoupling");                                  // CPR.hpp:152-170
    if (p.seq == 2) {
      ProfPhase ph("Jacobi precond");
      jac = new Jacobi(A, p.glo_block_size);
    }
    hipLaunchKernelGGL(k_undo_fix, dim3(gr), dim3(256), 0, stream(), n, A->r.p, A->c.p, A->v.p,
                       (const unsigned char*)flag.p);
    DECO_CHECK_LAUNCH();
    A->invalidate();
    // the Krylov product with the decoupled A runs on its cell-block shadow (FP64 values)
    A->cell = bs;
    A->block64(bs, bs);
    // the strided pressure residual of CPR.hpp:241-246 runs as a contiguous stream SpMV over
    // a copy of the pressure rows (same entries, same per-row summation order)
    if (p.seq != 1) Apr = csr_row_subset(A, bs, 0);
    rt.alloc(n); xp.alloc(sn); tmp.alloc(n); tmp2.alloc(n); x2.alloc(n);
    x32.alloc(n); rt32.alloc(n); xp32.alloc(sn); t32.alloc(n);
    sync();
  }

  unsigned gn() const { return grid_for(n, 256); }
  unsigned gs() const { return grid_for(sn, 256); }

  // x = M_glo r (FP64), or x += M_glo r
  void glo_apply(const double* r, double* x) { glo->apply(r, x); }
  void glo_apply_add(const double* r, double* x) {
    if (deco::FSAI* F = as_fsai(glo)) fsai_apply_add(F, r, x);
    else {
      glo->apply(r, tmp.p);
      hipLaunchKernelGGL((k_addz<double, double>), dim3(gn()), dim3(256), 0, stream(), n, (const double*)x,
                         (const double*)tmp.p, x);
      count_bytes(3.0 * n * 8);
      DECO_CHECK_LAUNCH();
    }
  }

  // Products with the pressure rows and the decoupled A: on their cell-block shadows (1 x bs and
  // bs x bs blocks) or the CSR arrays, as the layout selects.
  template <class E>
  void apr_product(const double* x, E e) {
    if (const deco_bsr* B = blocks64() ? Apr->block64(1, bs) : nullptr) bsr_stream_d(B, x, e);
    else spmv_stream_f(Apr, Apr->v.p, x, e);
  }
  template <class E>
  void apr_product(const float* x, E e) {
    if (const deco_bsr* B = blocks32() ? Apr->block32(1, bs) : nullptr) bsr_stream_f(B, x, e);
    else spmv_stream_f(Apr, Apr->values_f32(), x, e);
  }
  template <class E>
  void a_product(const deco_csr* M, const double* x, E e) {
    if (const deco_bsr* B = blocks64() ? M->block64(bs, bs) : nullptr) bsr_stream_d(B, x, e);
    else spmv_stream_f(M, M->v.p, x, e);
  }
  template <class E>
  void a_product(const deco_csr* M, const float* x, E e) {
    if (const deco_bsr* B = blocks32() ? M->block32(bs, bs) : nullptr) bsr_stream_f(B, x, e);
    else spmv_stream_f(M, const_cast<deco_csr*>(M)->values_f32(), x, e);
  }

  // A^T (P ep) for a vector ep on the cells, through the pressure columns of A^T (bs x 1 blocks)
  const deco_csr* pressure_cols_t() {
    if (!A->At) A->At = csr_transpose(A);
    if (!Atp || atp_gen != A->gen) {
      delete Atp;
      Atp = nullptr;
      Atp = csr_col_subset(A->At, bs, 0);
      atp_gen = A->gen;
    }
    return Atp;
  }
  template <class E>
  void atp_product(const double* ep, E e) {
    const deco_csr* M = pressure_cols_t();
    if (const deco_bsr* B = blocks64() ? M->block64(bs, 1) : nullptr) bsr_stream_d(B, ep, e);
    else spmv_stream_f(M, M->v.p, ep, e);
  }
  template <class E>
  void atp_product(const float* ep, E e) {
    const deco_csr* M = pressure_cols_t();
    if (const deco_bsr* B = blocks32() ? M->block32(bs, 1) : nullptr) bsr_stream_f(B, ep, e);
    else spmv_stream_f(M, const_cast<deco_csr*>(M)->values_f32(), ep, e);
  }

  // strided pressure residual rt[i] = r[bs i] - (A x)[bs i]   (CPR.hpp:241-246)
  template <class T>
  void pressure_residual(const T* x, const double* r, T* out) {
    apr_product(x, EpiStrideRes<T, double>{out, r, bs});
  }
  // the same, also writing the AMG's first zero-guess sweep x0 = rt/d0
  template <class T>
  void pressure_residual_j0(const T* x, const double* r, T* out, const T* d0, T* x0) {
    apr_product(x, EpiStrideResJ0<T, double>{out, r, bs, d0, x0});
  }
  // out = r - A x with FP64 values of the decoupled A
  void residual64(const double* x, const double* r, double* out) {
    a_product(A, x, EpiStd<EPI_RES, double, double>{out, r});
  }
  // out = r - A x with FP32 values of the decoupled A
  void residual32(const float* x, const double* r, float* out) {
    a_product(A, x, EpiStd<EPI_RES, float, double>{out, r});
  }

  void apply(const double* r, double* x) override {
    if (precision == DECO_PREC_FP32) { apply32_seq(r, x); return; }
    const int seq = p.seq;
    if (seq == 1) {                                                   // CPR.hpp:176-200
      hipLaunchKernelGGL((k_gather_pressure<double, double>), dim3(gs()), dim3(256), 0, stream(), sn, bs, r, rt.p);
      count_bytes((double)sn * 12);
      loc->apply(rt.p, xp.p);
      zero_async(x, sizeof(double) * n);
      hipLaunchKernelGGL((k_set_pressure<double, double>), dim3(gs()), dim3(256), 0, stream(), sn, bs, x,
                         (const double*)xp.p);
      count_bytes((double)sn * 12);
      residual64(x, r, rt.p);
      glo->apply(rt.p, x);
      hipLaunchKernelGGL((k_add_pressure<double, double>), dim3(gs()), dim3(256), 0, stream(), sn, bs, x,
                         (const double*)xp.p);
      count_bytes((double)sn * 20);
      DECO_CHECK_LAUNCH();
      return;
    }
    if (seq == 2) jac->apply(r, x); else glo_apply(r, x);
    // With an AMG of Jacobi-type smoothers the residual epilogue also performs the finest level's
    // zero-guess sweep, and the cycle's last sweep adds xp into x (same arithmetic as the
    // separate k_relax_zero and k_add_pressure).
    const double* d0 = nullptr;
    const bool fused = loc->stage_diag(&d0);
    {
      ProfPhase ph("CPR_rest");                                          // CPR.hpp:240-296
      if (fused) pressure_residual_j0((const double*)x, r, rt.p, d0, xp.p);
      else pressure_residual((const double*)x, r, rt.p);
    }
    if (fused) loc->stage(rt.p, xp.p, x, bs);
    else loc->apply(rt.p, xp.p);
    ProfPhase ph_rest("CPR_rest");
    if (!fused) {
      hipLaunchKernelGGL((k_add_pressure<double, double>), dim3(gs()), dim3(256), 0, stream(), sn, bs, x,
                         (const double*)xp.p);
      count_bytes((double)sn * 20);
      DECO_CHECK_LAUNCH();
    }
    if (seq == 0) return;
    residual64(x, r, rt.p);
    ph_rest.end();
    glo_apply_add(rt.p, x);
  }

  // FP32 internals, FP64 in, FP64 or FP32 out
  template <class TO>
  void apply32_seq(const double* r, TO* xo) {
    deco::FSAI* F = as_fsai(glo);
    auto glo32_from64 = [&](const double* in, float* out) {
      if (F) fsai_apply32(F, in, out);
      else {
        convert(in, t32.p, n);
        glo->apply32(t32.p, out);
      }
    };
    const int seq = p.seq;
    if (seq == 1) {
      hipLaunchKernelGGL((k_gather_pressure<double, float>), dim3(gs()), dim3(256), 0, stream(), sn, bs, r, rt32.p);
      count_bytes((double)sn * 12);
      loc->apply32(rt32.p, xp32.p);
      zero_async(x32.p, sizeof(float) * n);
      hipLaunchKernelGGL((k_set_pressure<float, float>), dim3(gs()), dim3(256), 0, stream(), sn, bs, x32.p,
                         (const float*)xp32.p);
      count_bytes((double)sn * 12);
      residual32(x32.p, r, rt32.p);
      if (F) fsai_apply32_f(F, rt32.p, x32.p);
      else glo->apply32(rt32.p, x32.p);
      hipLaunchKernelGGL((k_add_pressure<float, float>), dim3(gs()), dim3(256), 0, stream(), sn, bs, x32.p,
                         (const float*)xp32.p);
      count_bytes((double)sn * 20);
      convert((const float*)x32.p, xo, n);
      return;
    }
    if (seq == 2) { convert(r, t32.p, n); jac->apply32(t32.p, x32.p); }
    else glo32_from64(r, x32.p);
    const float* d0 = nullptr;
    const bool fused = loc->stage_diag32(&d0);
    {
      ProfPhase ph("CPR_rest");
      if (fused) pressure_residual_j0((const float*)x32.p, r, rt32.p, d0, xp32.p);
      else pressure_residual((const float*)x32.p, r, rt32.p);
    }
    if (fused) loc->stage32(rt32.p, xp32.p, x32.p, bs);
    else loc->apply32(rt32.p, xp32.p);
    ProfPhase ph_rest("CPR_rest");
    if (!fused) {
      hipLaunchKernelGGL((k_add_pressure<float, float>), dim3(gs()), dim3(256), 0, stream(), sn, bs, x32.p,
                         (const float*)xp32.p);
      count_bytes((double)sn * 20);
      DECO_CHECK_LAUNCH();
    }
    if (seq == 0) { convert((const float*)x32.p, xo, n); return; }
    residual32(x32.p, r, rt32.p);
    ph_rest.end();
    if (F) fsai_apply32_addz(F, rt32.p, x32.p, xo);
    else {
      glo->apply32(rt32.p, t32.p);
      hipLaunchKernelGGL((k_addz<float, float, TO>), dim3(gn()), dim3(256), 0, stream(), n, (const float*)x32.p,
                         (const float*)t32.p, xo);
      count_bytes(3.0 * n * 8);
      DECO_CHECK_LAUNCH();
    }
  }

  // out = r - A^T x with FP32 values of the decoupled A^T
  template <class TR>
  void residual32_t(const float* x, const TR* r, float* out) {
    a_product(A->At, x, EpiStd<EPI_RES, float, TR>{out, r});
  }

  // M^T u for seq 3 with FP32 internals, FP64 in, FP64 or FP32 out (the sequence of apply_t)
  template <class TO>
  void apply32_t_seq(const double* u, TO* xo) {
    deco::FSAI* F = as_fsai(glo);
    if (!F) fail(DECO_ERR_UNSUPPORTED, "the FP32 transposed CPR needs the FSAI global stage");
    s32.ensure(n);
    e32.ensure(n);
    pressure_cols_t();
    float* a = t32.p;
    fsai_apply32_t(F, u, a);                                           // a = G^T u
    {
      ProfPhase ph("CPR_rest");
      // s = u - A^T a, and its pressure entries rt = s[bs i] from the same pass
      a_product(A->At, (const float*)a, EpiResP<float, double>{s32.p, u, rt32.p, bs});
    }
    loc->apply32_t(rt32.p, xp32.p);                                    // e = P ep, ep = L^T R s
    ProfPhase ph_rest("CPR_rest");
    atp_product((const float*)xp32.p, EpiStd<EPI_RES, float, float>{x32.p, s32.p});   // s - A^T e
    hipLaunchKernelGGL(k_add_p, dim3(gn()), dim3(256), 0, stream(), n, bs, (const float*)a, (const float*)xp32.p,
                       e32.p);
    count_bytes(8.0 * n + 4.0 * sn);   // a + e
    DECO_CHECK_LAUNCH();
    ph_rest.end();
    fsai_apply32_t_addz(F, x32.p, e32.p, xo);                          // out = a + e + G^T (s - A^T e)
  }

  // M^T u for seq 3 (FP64): a = G^T u, s = u - A^T a, e = P L^T R s, out = a + e + G^T (s - A^T e)
  void apply_t(const double* u, double* x) override {
    if (p.seq != 3) fail(DECO_ERR_UNSUPPORTED, "transposed CPR is implemented for seq 3");
    if (!A->At) A->At = csr_transpose(A);
    if (precision == DECO_PREC_FP32) { apply32_t_seq(u, x); return; }
    deco_csr* At = A->At;
    pressure_cols_t();
    double* a = tmp.p;
    double* s = tmp2.p;
    glo->apply_t(u, a);
    a_product(At, (const double*)a, EpiResP<double, double>{s, u, rt.p, bs});   // s = u - A^T a, rt = s[bs i]
    loc->apply_t(rt.p, xp.p);                                          // e = P ep
    atp_product((const double*)xp.p, EpiStd<EPI_RES, double, double>{rt.p, s});   // rt (length n) = s - A^T e
    glo->apply_t(rt.p, x);
    hipLaunchKernelGGL(k_add3_p, dim3(gn()), dim3(256), 0, stream(), n, bs, (const double*)a, (const double*)xp.p, x);
    count_bytes(24.0 * n + 8.0 * sn);
    DECO_CHECK_LAUNCH();
  }

  void apply_f32out(const double* r, float* x) override {
    if (precision != DECO_PREC_FP32) { deco_prec::apply_f32out(r, x); return; }
    apply32_seq(r, x);
  }
  void apply_t_f32out(const double* u, float* x) override {
    if (precision != DECO_PREC_FP32) { deco_prec::apply_t_f32out(u, x); return; }
    if (p.seq != 3) fail(DECO_ERR_UNSUPPORTED, "transposed CPR is implemented for seq 3");
    if (!A->At) A->At = csr_transpose(A);
    apply32_t_seq(u, x);
  }
  void set_precision(int pr) override { precision = pr; }
  void set_layout(int l) override {
    layout = l;
    if (glo) glo->set_layout(l);
  }
};

deco_prec* make_cpr(const deco_cpr_params& p, deco_csr* A, double* b) { return new CPR(p, A, b); }

static CPR* as_cpr(const deco_prec* M) {
  if (!M || M->kind != deco_prec::CPR) fail(DECO_ERR_ARG, "not a CPR preconditioner");
  return static_cast<CPR*>(const_cast<deco_prec*>(M));
}

}  // namespace deco

using namespace deco;

extern "C" {

int deco_cpr_create(const deco_cpr_params* p, deco_csr* A, double* b, deco_prec** out) {
  return guard([&] {
    if (!p || !A || !out) fail(DECO_ERR_ARG, "null argument");
    if (A->n != A->m) fail(DECO_ERR_ARG, "matrix is not square");
    *out = make_cpr(*p, A, b);
  });
}

int deco_cpr_pressure_matrix(const deco_prec* M, const deco_csr** Ap) {
  return guard([&] {
    if (!Ap) fail(DECO_ERR_ARG, "null argument");
    *Ap = as_cpr(M)->Ap;
  });
}

int deco_cpr_fsai(const deco_prec* M, const deco_prec** F) {
  return guard([&] {
    if (!F) fail(DECO_ERR_ARG, "null argument");
    *F = as_cpr(M)->glo;
  });
}

int deco_cpr_amg(const deco_prec* M, const deco_prec** G) {
  return guard([&] {
    if (!G) fail(DECO_ERR_ARG, "null argument");
    *G = as_cpr(M)->loc;
  });
}

int deco_cpr_left_transform(const deco_prec* M, double* W) {
  return guard([&] {
    if (!W) fail(DECO_ERR_ARG, "null argument");
    CPR* c = as_cpr(M);
    c->W.download(W, c->W.n);
  });
}

int deco_cpr_transform_transpose(const deco_prec* M, const double* mu, double* lam) {
  return guard([&] {
    CPR* c = as_cpr(M);
    need(mu, c->n, "mu");
    need(lam, c->n, "lambda");
    hipLaunchKernelGGL(k_transform_t, dim3(grid_for(c->sn, 128)), dim3(128), 0, stream(), c->sn, c->bs,
                       (const double*)c->W.p, mu, lam);
    count_bytes((double)c->sn * c->bs * (c->bs + 2) * 8);
    DECO_CHECK_LAUNCH();
    sync();
  });
}

}  // extern "C"
