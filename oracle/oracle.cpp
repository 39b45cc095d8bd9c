// CPU oracle of the deco CPR-Krylov solve path -- TEST INFRASTRUCTURE ONLY.
//
// Clean-room restatement, in C++/OpenMP, of the reference algorithms; every function cites the
// reference file:line it follows. Built by oracle/Makefile into oracle/liboracle.so. Only the
// tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product library
// never does.
//
// Parity status of each part (tests/test_oracle_*.py hold the checks that pin them):
//   spmv/BLAS, filters, local_matrix, transpose, SpGEMM, FSAI pattern     pinned (closed forms,
//       scipy.sparse, brute force)
//   CPR decoupling (fix_diag, weights, DRS, IMPES, CPD)                   pinned (invariants)
//   FSAI numeric                                                          pinned (invariants:
//       (FL A)[i,P_i] = 0, (A FU^T)[P_i,i] = 0, unit diagonal, closed forms)
//   rgmres/lgmres/bicgstab/lbicgstab/pcg                                  pinned (exact on tiny
//       systems, Arnoldi relation, scipy solutions)
//   AMG: R = P^T, Galerkin RAP, drop rule, smoother diagonals, coarsest   pinned (scipy RAP,
//       inverse and the V-cycle with every smoother (l1-Jacobi, Jacobi,    numpy inverse, dense
//       F-C-F, two-stage Gauss-Seidel) on imported P_l and C/F splittings    multi-level closed
//                                                                          forms)
//       The interpolation and coarsening follow no reference code (hypre is absent): parity of
//       the hierarchy with hypre BoomerAMG is unpinned.
#include "oracle.h"
#include <omp.h>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <memory>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <functional>

namespace orc {

// ------------------------------------------------------------------ CSR ----------------
struct CSR {
  int n = 0, m = 0;
  std::vector<int> r, c;
  std::vector<double> v;
  int nnz() const { return (int)c.size(); }
};

static CSR from_c(const orc_csr* A) {
  CSR B;
  B.n = A->n;
  B.m = A->m;
  B.r.assign(A->r, A->r + A->n + 1);
  B.c.assign(A->c, A->c + A->nnz);
  B.v.assign(A->v, A->v + A->nnz);
  return B;
}

static void to_c(const CSR& B, orc_csr* out) {
  out->n = B.n;
  out->m = B.m;
  out->nnz = B.nnz();
  out->r = (int*)malloc(sizeof(int) * (B.n + 1));
  out->c = (int*)malloc(sizeof(int) * std::max(1, B.nnz()));
  out->v = (double*)malloc(sizeof(double) * std::max(1, B.nnz()));
  std::memcpy(out->r, B.r.data(), sizeof(int) * (B.n + 1));
  if (B.nnz()) {
    std::memcpy(out->c, B.c.data(), sizeof(int) * B.nnz());
    std::memcpy(out->v, B.v.data(), sizeof(double) * B.nnz());
  }
}

// core::binarysearch (core/alg.hpp:16-25): largest index in [lower, upper) whose key <= target
static inline int binarysearch(const int* arr, int lower, int upper, int target) {
  int mn = lower, mx = upper, mid = (lower + upper) >> 1;
  while (mx - mn > 1) {
    if (arr[mid] > target) mx = mid; else mn = mid;
    mid = (mn + mx) >> 1;
  }
  return mid;
}

// ------------------------------------------------------------------ alg ----------------
// alg::spmv (core/alg.hpp:51-64): rows 0, bs, 2bs, ... ; b[i/bs] = sum in column order
static void spmv(const CSR& A, const double* x, double* y, int bs = 1) {
  const int n = A.n;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i += bs) {
    double s = 0.0;
    for (int j = A.r[i]; j < A.r[i + 1]; ++j) s += A.v[j] * x[A.c[j]];
    y[i / bs] = s;
  }
}

// alg::dot (core/alg.hpp:278-291) with per-thread partials combined in thread order
static double dot(int n, const double* a, const double* b) {
  const int T = omp_get_max_threads();
  std::vector<double> part(T, 0.0);
#pragma omp parallel
  {
    double s = 0.0;
#pragma omp for schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    part[omp_get_thread_num()] = s;
  }
  double s = 0.0;
  for (double p : part) s += p;
  return s;
}
// alg::l2norm (core/alg.hpp:334-352)
static double nrm2(int n, const double* a) { return std::sqrt(dot(n, a, a)); }

// alg::diag_scale (core/alg.hpp:153-177)
static void diag_scale(CSR& A, const double* D) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < A.n; ++i)
    for (int j = A.r[i]; j < A.r[i + 1]; ++j) A.v[j] *= D[i];
}

// ------------------------------------------------------------------ CSR methods --------
// CSR::filter (core/CSR.hpp:423-473) / filter_proper (:476-525)
static void filter(CSR& A, double tol, bool proper) {
  const int n = A.n;
  std::vector<int> rp(n + 1, 0);
  auto keep_tol = [&](int i) {
    if (!proper) return tol;
    int j = binarysearch(A.c.data(), A.r[i], A.r[i + 1], i);
    return std::fabs(A.v[j] * tol);
  };
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const double D = keep_tol(i);
    int k = 0;
    for (int j = A.r[i]; j < A.r[i + 1]; ++j)
      if (std::fabs(A.v[j]) > D || A.c[j] == i) ++k;
    rp[i + 1] = k;
  }
  for (int i = 1; i <= n; ++i) rp[i] += rp[i - 1];
  std::vector<int> cp(rp[n]);
  std::vector<double> vp(rp[n]);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const double D = keep_tol(i);
    int o = rp[i];
    for (int j = A.r[i]; j < A.r[i + 1]; ++j)
      if (std::fabs(A.v[j]) > D || A.c[j] == i) { cp[o] = A.c[j]; vp[o] = A.v[j]; ++o; }
  }
  A.r.swap(rp);
  A.c.swap(cp);
  A.v.swap(vp);
}

// CSR::local_matrix (core/CSR.hpp:338-420)
static CSR local_matrix(const CSR& A, int bs, int p, const double* w) {
  if (A.n % bs != 0) throw std::invalid_argument("n is not divisible by bs");
  CSR L;
  L.n = A.n / bs;
  L.m = A.m / bs;
  L.r.assign(L.n + 1, 0);
  for (int i = 0; i < L.n; ++i) {
    int k = 0;
    for (int j = A.r[bs * i + p]; j < A.r[bs * i + p + 1]; ++j)
      if (A.c[j] % bs == p) ++k;
    L.r[i + 1] = L.r[i] + k;
  }
  L.c.resize(L.r[L.n]);
  L.v.resize(L.r[L.n]);
  for (int i = 0; i < L.n; ++i) {
    int o = L.r[i];
    const double scale = w ? w[bs * i + p] : 1.0;
    for (int j = A.r[bs * i + p]; j < A.r[bs * i + p + 1]; ++j)
      if (A.c[j] % bs == p) {
        L.c[o] = A.c[j] / bs;
        L.v[o] = w ? A.v[j] / scale : A.v[j];
        ++o;
      }
  }
  return L;
}

// CSR::trans_mat_full (core/CSR.hpp:151-211): counting sort by column, rows of the result in
// increasing order of the original row index.
static CSR transpose(const CSR& A) {
  CSR T;
  T.n = A.m;
  T.m = A.n;
  T.r.assign(A.m + 1, 0);
  for (int j : A.c) T.r[j + 1]++;
  for (int i = 0; i < A.m; ++i) T.r[i + 1] += T.r[i];
  T.c.resize(A.nnz());
  T.v.resize(A.nnz());
  std::vector<int> cur(T.r.begin(), T.r.end() - 1);
  for (int i = 0; i < A.n; ++i)
    for (int j = A.r[i]; j < A.r[i + 1]; ++j) {
      int k = cur[A.c[j]]++;
      T.c[k] = i;
      T.v[k] = A.v[j];
    }
  return T;
}

// ------------------------------------------------------------------ SpGEMM -------------
// MxM (core/MxM.cpp:113-146): structural product, column indices sorted ascending. The
// reference accumulates each row in a hash table in (j,k) loop order; the values here are
// accumulated per column in the same (j,k) order with a dense marker.
static CSR spgemm(const CSR& A, const CSR& B) {
  CSR C;
  C.n = A.n;
  C.m = B.m;
  C.r.assign(A.n + 1, 0);
  std::vector<std::vector<int>> cols(A.n);
  std::vector<std::vector<double>> vals(A.n);
#pragma omp parallel
  {
    std::vector<int> pos(B.m, -1);
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < A.n; ++i) {
      std::vector<int>& cc = cols[i];
      std::vector<double>& vv = vals[i];
      for (int j = A.r[i]; j < A.r[i + 1]; ++j) {
        const int ca = A.c[j];
        const double va = A.v[j];
        for (int k = B.r[ca]; k < B.r[ca + 1]; ++k) {
          const int key = B.c[k];
          const double val = B.v[k] * va;
          if (pos[key] < 0) { pos[key] = (int)cc.size(); cc.push_back(key); vv.push_back(val); }
          else vv[pos[key]] += val;
        }
      }
      for (int key : cc) pos[key] = -1;
      std::vector<int> idx(cc.size());
      for (size_t q = 0; q < idx.size(); ++q) idx[q] = (int)q;
      std::sort(idx.begin(), idx.end(), [&](int a, int b) { return cc[a] < cc[b]; });
      std::vector<int> c2(cc.size());
      std::vector<double> v2(cc.size());
      for (size_t q = 0; q < idx.size(); ++q) { c2[q] = cc[idx[q]]; v2[q] = vv[idx[q]]; }
      cc.swap(c2);
      vv.swap(v2);
    }
  }
  for (int i = 0; i < A.n; ++i) C.r[i + 1] = C.r[i] + (int)cols[i].size();
  C.c.resize(C.r[A.n]);
  C.v.resize(C.r[A.n]);
  for (int i = 0; i < A.n; ++i) {
    std::copy(cols[i].begin(), cols[i].end(), C.c.begin() + C.r[i]);
    std::copy(vals[i].begin(), vals[i].end(), C.v.begin() + C.r[i]);
  }
  return C;
}

// FSAI symbolic (precond/FSAI/symbolic.tpp:102-177).
//   power 1: the lower triangle of A including the diagonal (:104-148), relying on sorted rows;
//   power 2: tril(pattern(A*A)): for each j in row i, the keys of row j up to the first key > i
//            (early break, :136,162), union, sorted (:176).
static CSR fsai_pattern(const CSR& A, int power) {
  CSR F;
  F.n = A.n;
  F.m = A.m;
  F.r.assign(A.n + 1, 0);
  if (power == 1) {
    for (int i = 0; i < A.n; ++i) {
      int k = 0;
      for (int j = A.r[i]; j < A.r[i + 1]; ++j) {
        if (A.c[j] <= i) ++k; else break;
      }
      F.r[i + 1] = F.r[i] + k;
    }
    F.c.resize(F.r[A.n]);
    for (int i = 0; i < A.n; ++i) {
      int o = F.r[i];
      for (int j = A.r[i]; j < A.r[i + 1]; ++j) {
        if (A.c[j] <= i) F.c[o++] = A.c[j]; else break;
      }
    }
  } else if (power == 2) {
    std::vector<std::vector<int>> rows(A.n);
#pragma omp parallel
    {
      std::vector<char> mark(A.n, 0);
#pragma omp for schedule(dynamic, 64)
      for (int i = 0; i < A.n; ++i) {
        std::vector<int>& out = rows[i];
        for (int j = A.r[i]; j < A.r[i + 1]; ++j) {
          const int ca = A.c[j];
          for (int k = A.r[ca]; k < A.r[ca + 1]; ++k) {
            const int key = A.c[k];
            if (key > i) break;
            if (!mark[key]) { mark[key] = 1; out.push_back(key); }
          }
        }
        for (int key : out) mark[key] = 0;
        std::sort(out.begin(), out.end());
      }
    }
    for (int i = 0; i < A.n; ++i) F.r[i + 1] = F.r[i] + (int)rows[i].size();
    F.c.resize(F.r[A.n]);
    for (int i = 0; i < A.n; ++i) std::copy(rows[i].begin(), rows[i].end(), F.c.begin() + F.r[i]);
  } else {
    throw std::invalid_argument("FSAI power must be 1 or 2");
  }
  F.v.assign(F.c.size(), 0.0);
  return F;
}

// ------------------------------------------------------------------ dense LU -----------
// LAPACK dgetf2 semantics on a column-major k x k matrix: pivot = first index of the largest
// |a| in the column (idamax), row swap, scale by the reciprocal of the pivot, rank-1 update.
// Returns 0 or the 1-based index of the first zero pivot (the matrix is then singular).
static int lu_factor(int k, double* a, int* ipiv) {
  int info = 0;
  for (int j = 0; j < k; ++j) {
    int p = j;
    double mx = std::fabs(a[j + j * k]);
    for (int i = j + 1; i < k; ++i) {
      double t = std::fabs(a[i + j * k]);
      if (t > mx) { mx = t; p = i; }
    }
    ipiv[j] = p;
    if (a[p + j * k] != 0.0) {
      if (p != j)
        for (int q = 0; q < k; ++q) std::swap(a[j + q * k], a[p + q * k]);
      const double piv = a[j + j * k];
      if (std::fabs(piv) >= 2.2250738585072014e-308) {
        const double rp = 1.0 / piv;
        for (int i = j + 1; i < k; ++i) a[i + j * k] *= rp;
      } else {
        for (int i = j + 1; i < k; ++i) a[i + j * k] /= piv;
      }
    } else if (info == 0) {
      info = j + 1;
    }
    for (int q = j + 1; q < k; ++q) {
      const double t = a[j + q * k];
      if (t != 0.0)
        for (int i = j + 1; i < k; ++i) a[i + q * k] -= a[i + j * k] * t;
    }
  }
  return info;
}

// LAPACK dgetrs: trans = 0 solves A x = b, trans = 1 solves A^T x = b (A factored by lu_factor)
static void lu_solve(int k, const double* a, const int* ipiv, int trans, double* b) {
  if (!trans) {
    for (int j = 0; j < k; ++j)
      if (ipiv[j] != j) std::swap(b[j], b[ipiv[j]]);
    for (int j = 0; j < k; ++j) {            // L y = b (unit lower)
      const double t = b[j];
      if (t != 0.0)
        for (int i = j + 1; i < k; ++i) b[i] -= t * a[i + j * k];
    }
    for (int j = k - 1; j >= 0; --j) {       // U x = y
      if (b[j] != 0.0) {
        b[j] /= a[j + j * k];
        const double t = b[j];
        for (int i = 0; i < j; ++i) b[i] -= t * a[i + j * k];
      }
    }
  } else {
    for (int j = 0; j < k; ++j) {            // U^T y = b
      double t = b[j];
      for (int i = 0; i < j; ++i) t -= a[i + j * k] * b[i];
      b[j] = t / a[j + j * k];
    }
    for (int j = k - 1; j >= 0; --j) {       // L^T x = y
      double t = b[j];
      for (int i = j + 1; i < k; ++i) t -= a[i + j * k] * b[i];
      b[j] = t;
    }
    for (int j = k - 1; j >= 0; --j)
      if (ipiv[j] != j) std::swap(b[j], b[ipiv[j]]);
  }
}

// ------------------------------------------------------------------ preconditioners ----
struct Prec {
  int n = 0;
  virtual ~Prec() = default;
  virtual void apply(const double* r, double* x) = 0;
  virtual void apply_t(const double* r, double* x) { (void)r; (void)x; throw std::runtime_error("transpose not available"); }
};

struct Identity : Prec {
  explicit Identity(int n_) { n = n_; }
  void apply(const double* r, double* x) override { std::memcpy(x, r, sizeof(double) * n); }
  void apply_t(const double* r, double* x) override { apply(r, x); }
};

// FSAI (precond/FSAI/FSAI.hpp:14-47, FSAI.tpp:11-32, numeric.tpp:60-133)
struct FSAI : Prec {
  CSR FL, FU, FUt, FLt;
  std::vector<double> temp;
  FSAI(const deco_cpr_params& p, const CSR& A_in) {
    n = A_in.n;
    CSR A = A_in;                                      // filtered copy (FSAI.hpp:22-28)
    if (p.glo_tol_rel != 0.0) filter(A, p.glo_tol_rel, true);
    if (p.glo_tol_abs != 0.0) filter(A, p.glo_tol_abs, false);
    FL = fsai_pattern(A, p.glo_power);                 // symbolic (FSAI.tpp:13)
    FU = FL;                                           // one pattern (FSAI.tpp:21-23)
    numeric(A);
    temp.assign(n, 0.0);
  }
  // numeric.tpp:60-133 with gather.cpp:15-88 and LU_main.cpp:15-48
  void numeric(const CSR& A) {
    const int nr = n;
#pragma omp parallel
    {
      std::vector<double> M, bU, bL;
      std::vector<int> ipiv, pos(A.m, -1);
#pragma omp for schedule(dynamic, 256)
      for (int i = 0; i < nr; ++i) {
        const int iaF = FL.r[i], ibF = FL.r[i + 1] - 1;   // ibF: the diagonal position
        const int k = ibF - iaF;
        if (k <= 0) continue;
        M.assign((size_t)k * k, 0.0);
        bU.assign(k, 0.0);
        bL.assign(k, 0.0);
        ipiv.assign(k, 0);
        for (int q = 0; q < k; ++q) pos[FL.c[iaF + q]] = q;
        // full_A row-major [rowid][colid] = A[P[rowid], P[colid]] (gather.cpp:188-199);
        // LAPACK reads it column-major, i.e. M = A[P,P]^T: M(col=rowid, row=colid).
        for (int q = 0; q < k; ++q) {
          const int row = FL.c[iaF + q];
          for (int s = A.r[row]; s < A.r[row + 1]; ++s) {
            const int cc = A.c[s];
            if (cc < (int)pos.size() && pos[cc] >= 0) M[(size_t)q * k + pos[cc]] = A.v[s];
          }
          // rhsU[rowid] = -A[P[rowid], i] (gather.cpp:201-204)
          const int jd = binarysearch(A.c.data(), A.r[row], A.r[row + 1], i);
          if (A.c[jd] == i) bU[q] = -A.v[jd];
        }
        // rhsL[colid] = -A[i, P[colid]] (gather.cpp:210-217)
        for (int s = A.r[i]; s < A.r[i + 1]; ++s) {
          const int cc = A.c[s];
          if (pos[cc] >= 0) bL[pos[cc]] = -A.v[s];
        }
        for (int q = 0; q < k; ++q) pos[FL.c[iaF + q]] = -1;
        lu_factor(k, M.data(), ipiv.data());           // dgetrf (LU_main.cpp:41)
        lu_solve(k, M.data(), ipiv.data(), 1, bU.data());  // dgetrs 'T' -> FU (:42)
        lu_solve(k, M.data(), ipiv.data(), 0, bL.data());  // dgetrs 'N' -> FL (:43)
        for (int q = 0; q < k; ++q) {                  // scatter (numeric.tpp:34-54)
          FL.v[iaF + q] = bL[q];
          FU.v[iaF + q] = bU[q];
        }
      }
    }
    // MxM_diag (diagonal.cpp:7-52)
    std::vector<double> D(nr);
    for (int i = 0; i < nr; ++i) {
      FL.v[FL.r[i + 1] - 1] = 1.0;
      FU.v[FL.r[i + 1] - 1] = 1.0;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nr; ++i) {
      const int aFL = FL.r[i], bFL = FL.r[i + 1];
      double s = 0.0;
      for (int j = aFL; j < bFL; ++j) {
        const double vfl = FL.v[j];
        if (vfl != 0.0) {
          const int col = FL.c[j];
          int aA = A.r[col];
          const int bA = A.r[col + 1];
          int aFU = aFL;
          double sAFU = 0.0;
          while (aA < bA && aFU < bFL) {
            if (A.c[aA] == FL.c[aFU]) { sAFU += A.v[aA] * FU.v[aFU]; ++aA; ++aFU; }
            else if (A.c[aA] < FL.c[aFU]) ++aA;
            else ++aFU;
          }
          s += vfl * sAFU;
        }
      }
      D[i] = (s < 0.0) ? (-1.0 / std::sqrt(std::fabs(s))) : (1.0 / std::sqrt(s));
    }
    // Jacobi_scale (diagonal.cpp:57-73): FL signed, FU unsigned
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nr; ++i) {
      const double dd = D[i], da = std::fabs(dd);
      for (int j = FL.r[i]; j < FL.r[i + 1]; ++j) {
        FL.v[j] = FL.v[j] * dd;
        FU.v[j] = FU.v[j] * da;
      }
    }
    FUt = transpose(FU);                                // trans_mat_full (numeric.tpp:114)
    FLt = transpose(FL);                                // for the transposed apply
  }
  void apply(const double* r, double* x) override {    // FSAI.hpp:43-44
    spmv(FL, r, temp.data());
    spmv(FUt, temp.data(), x);
  }
  void apply_t(const double* r, double* x) override {  // (FUt FL)^T = FLt FU
    spmv(FU, r, temp.data());
    spmv(FLt, temp.data(), x);
  }
};

// Jacobi (precond/Jacobi/Jacobi.hpp:15-68, Jacobi.tpp:3-69)
struct Jacobi : Prec {
  int bs = 1;
  std::vector<double> J;  // bs==1: 1/a_ii; else per cell bs*bs row-major inverse
  Jacobi(int bs_, const CSR& A) {
    n = A.n;
    bs = bs_;
    if (bs == 1) {
      J.resize(n);
      for (int i = 0; i < n; ++i) J[i] = 1.0 / A.v[binarysearch(A.c.data(), A.r[i], A.r[i + 1], i)];
    } else {
      const int sn = n / bs;
      J.assign((size_t)sn * bs * bs, 0.0);
      std::vector<double> B(bs * bs), col(bs);
      std::vector<int> ipiv(bs);
      for (int s = 0; s < sn; ++s) {
        std::fill(B.begin(), B.end(), 0.0);
        for (int p = 0; p < bs; ++p) {
          const int i = s * bs + p;
          const int ia = A.r[i], ib = A.r[i + 1];
          const int j = binarysearch(A.c.data(), ia, ib, i);
          for (int k = 0; k < bs; ++k) {
            const int jj = j + k - p;
            if (jj < ib && jj >= ia && A.c[jj] / bs == s) B[p + (A.c[jj] % bs) * bs] = A.v[jj];  // col-major B
          }
        }
        lu_factor(bs, B.data(), ipiv.data());
        for (int k = 0; k < bs; ++k) {   // inverse column by column: J[s] row-major = inv(B)
          std::fill(col.begin(), col.end(), 0.0);
          col[k] = 1.0;
          lu_solve(bs, B.data(), ipiv.data(), 0, col.data());
          for (int p = 0; p < bs; ++p) J[(size_t)s * bs * bs + p * bs + k] = col[p];
        }
      }
    }
  }
  void apply_impl(const double* x, double* b, bool tr) const {
    if (bs == 1) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) b[i] = J[i] * x[i];
      return;
    }
    const int sn = n / bs;
#pragma omp parallel for schedule(static)
    for (int s = 0; s < sn; ++s) {
      const double* Ap = J.data() + (size_t)s * bs * bs;
      for (int p = 0; p < bs; ++p) {
        double sum = (tr ? Ap[p] : Ap[p * bs]) * x[s * bs];
        for (int k = 1; k < bs; ++k) sum += (tr ? Ap[k * bs + p] : Ap[p * bs + k]) * x[s * bs + k];
        b[s * bs + p] = sum;
      }
    }
  }
  void apply(const double* r, double* x) override { apply_impl(r, x, false); }
  void apply_t(const double* r, double* x) override { apply_impl(r, x, true); }
};

// Exact dense solve (tiny matrices): the "perfect" local stage used by CPU-only tests.
struct Dense : Prec {
  std::vector<double> LU;
  std::vector<int> ipiv;
  explicit Dense(const CSR& A) {
    n = A.n;
    LU.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = A.r[i]; j < A.r[i + 1]; ++j) LU[i + (size_t)A.c[j] * n] = A.v[j];
    ipiv.resize(n);
    if (lu_factor(n, LU.data(), ipiv.data()) != 0) throw std::runtime_error("singular dense matrix");
  }
  void apply(const double* r, double* x) override {
    std::memcpy(x, r, sizeof(double) * n);
    lu_solve(n, LU.data(), ipiv.data(), 0, x);
  }
  void apply_t(const double* r, double* x) override {
    std::memcpy(x, r, sizeof(double) * n);
    lu_solve(n, LU.data(), ipiv.data(), 1, x);
  }
};

// AMG pressure stage (the role of hypre_AMG, precond/hypre_AMG/hypre_AMG.hpp:16-37 and
// hypre_AMG.tpp:44-67). hypre (v2.31.0) is absent, so the C/F splittings and interpolation
// operators P_l are imported (they are the product's host setup, which follows no reference
// code); everything else of the hierarchy is formed here from the finest matrix: R_l = P_l^T,
// the Galerkin product A_{l+1} = R_l (A_l P_l) with MxM, the drop rule of amg_A_drop_tol, the
// smoother diagonals and the coarsest inverse (dgetf2/dgetrs on unit columns).
//
// Apply: one V-cycle from x = 0 (hypre_AMG.hpp:28-33, hypre_AMG.tpp:67), one sweep down and up.
// Smoothers by amg_relax_type (the reference passes a type to hypre only when it is positive,
// hypre_AMG.tpp:50-51, so -1 and 0 both mean hypre's default, Gauss-Seidel down and up):
//   -1, 0  two-stage Gauss-Seidel: forward (strict lower part) down, backward (strict upper) up
//   11, 13 two-stage Gauss-Seidel, forward both ways;  14: backward both ways
//   12     two-stage Gauss-Seidel with two inner iterations, forward both ways
//   7      Jacobi with weight amg_relax_weight (1 when unset)
//   17     F-C-F Jacobi half-sweeps;  18: l1-Jacobi
// A two-stage sweep with s inner iterations on x (Neumann series of (D + T)^{-1}, T the strict
// lower or upper part): rho = b - A x, g_0 = rho / d, g_k = (rho - T g_{k-1}) / d, x += g_s.
struct AMG : Prec {
  struct Level {
    CSR A, At, P, R;
    CSR Lo, Up, Lot, Upt;      // strict lower / upper parts of A and of A^T (two-stage sweeps)
    std::vector<double> d;
    std::vector<int> cf;
    std::vector<double> x, b, r, t, g;
  };
  enum Kind { L1JAC, JAC, FCF, TWOSTAGE };
  std::vector<Level> L;
  CSR Ac;                      // coarsest matrix
  Kind kind = L1JAC;
  double weight = 1.0;
  int dir_down = 1, dir_up = -1, inner = 1;   // two-stage: +1 forward, -1 backward
  int nc = 0;
  std::vector<double> Cinv, xc, bc;
  bool have_t = false;

  void set_relax(int relax, double w) {
    weight = 1.0;
    switch (relax) {
      case 18: kind = L1JAC; break;
      case 7: kind = JAC; weight = w > 0.0 ? w : 1.0; break;
      case 17: kind = FCF; break;
      case -1: case 0: kind = TWOSTAGE; dir_down = 1; dir_up = -1; inner = 1; break;
      case 11: case 13: kind = TWOSTAGE; dir_down = 1; dir_up = 1; inner = 1; break;
      case 12: kind = TWOSTAGE; dir_down = 1; dir_up = 1; inner = 2; break;
      case 14: kind = TWOSTAGE; dir_down = -1; dir_up = -1; inner = 1; break;
      default: throw std::runtime_error("amg_relax_type " + std::to_string(relax) + " not supported");
    }
  }

  // keep a_ij if j == i or |a_ij| >= tol * max_k |a_ik| (hypre's A-drop with the row's max norm)
  static CSR drop_small(const CSR& A, double tol) {
    CSR B;
    B.n = A.n;
    B.m = A.m;
    B.r.assign(A.n + 1, 0);
    for (int i = 0; i < A.n; ++i) {
      double mx = 0.0;
      for (int j = A.r[i]; j < A.r[i + 1]; ++j) mx = std::max(mx, std::fabs(A.v[j]));
      const double thr = tol * mx;
      for (int j = A.r[i]; j < A.r[i + 1]; ++j)
        if (A.c[j] == i || std::fabs(A.v[j]) >= thr) {
          B.c.push_back(A.c[j]);
          B.v.push_back(A.v[j]);
        }
      B.r[i + 1] = B.nnz();
    }
    return B;
  }
  // strict lower (lower = true) or strict upper part, rows kept sorted
  static CSR strict_part(const CSR& A, bool lower) {
    CSR B;
    B.n = A.n;
    B.m = A.m;
    B.r.assign(A.n + 1, 0);
    for (int i = 0; i < A.n; ++i) {
      for (int j = A.r[i]; j < A.r[i + 1]; ++j)
        if (lower ? A.c[j] < i : A.c[j] > i) {
          B.c.push_back(A.c[j]);
          B.v.push_back(A.v[j]);
        }
      B.r[i + 1] = B.nnz();
    }
    return B;
  }
  void smoother_diag(Level& lv) const {
    const CSR& A = lv.A;
    lv.d.assign(A.n, 0.0);
    for (int i = 0; i < A.n; ++i) {
      double s = 0.0;
      for (int j = A.r[i]; j < A.r[i + 1]; ++j) {
        if (kind == L1JAC) s += std::fabs(A.v[j]);
        else if (A.c[j] == i) s += A.v[j];
      }
      lv.d[i] = kind == L1JAC ? s : s / weight;
    }
  }

  // hierarchy from the finest matrix and the imported P_l, cf_l
  void build(const CSR& A0, int nlev, const orc_csr* P, const int* const* cf, double drop_tol) {
    n = A0.n;
    CSR A = A0;
    L.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
      Level& lv = L[l];
      lv.P = from_c(&P[l]);
      if (lv.P.n != A.n) throw std::runtime_error("P_" + std::to_string(l) + " does not match A_l's rows");
      lv.cf.assign(cf[l], cf[l] + A.n);
      lv.R = transpose(lv.P);
      CSR AP = spgemm(A, lv.P);
      CSR RAP = spgemm(lv.R, AP);
      lv.A = std::move(A);
      A = drop_tol > 0.0 ? drop_small(RAP, drop_tol) : std::move(RAP);
    }
    Ac = std::move(A);
    nc = Ac.n;
    for (auto& lv : L) {
      smoother_diag(lv);
      if (kind == TWOSTAGE) {
        lv.Lo = strict_part(lv.A, true);
        lv.Up = strict_part(lv.A, false);
      }
      const int nl = lv.A.n;
      lv.x.assign(nl, 0.0);
      lv.b.assign(nl, 0.0);
      lv.r.assign(nl, 0.0);
      lv.t.assign(nl, 0.0);
      lv.g.assign(nl, 0.0);
    }
    // coarsest inverse, row-major: column k solves Ac y = e_k
    std::vector<double> lu((size_t)nc * nc, 0.0);
    for (int i = 0; i < nc; ++i)
      for (int j = Ac.r[i]; j < Ac.r[i + 1]; ++j) lu[i + (size_t)Ac.c[j] * nc] = Ac.v[j];
    std::vector<int> ipiv(nc);
    if (nc && lu_factor(nc, lu.data(), ipiv.data()) != 0) throw std::runtime_error("singular coarsest matrix");
    Cinv.assign((size_t)nc * nc, 0.0);
    std::vector<double> col(nc);
    for (int k = 0; k < nc; ++k) {
      std::fill(col.begin(), col.end(), 0.0);
      col[k] = 1.0;
      lu_solve(nc, lu.data(), ipiv.data(), 0, col.data());
      for (int i = 0; i < nc; ++i) Cinv[(size_t)i * nc + k] = col[i];
    }
    xc.assign(nc, 0.0);
    bc.assign(nc, 0.0);
  }

  // One relaxation sweep x <- x + D^{-1}(b - A x) on the rows of `subset`
  // (-1 all, 1 C points, 0 F points); out of place through t.
  void half_sweep(Level& lv, const CSR& A, int subset) {
    const int nn = A.n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nn; ++i) {
      const bool on = (subset < 0) || (subset == 1 ? lv.cf[i] == 1 : lv.cf[i] != 1);
      if (!on) { lv.t[i] = lv.x[i]; continue; }
      double s = 0.0;
      for (int j = A.r[i]; j < A.r[i + 1]; ++j) s += A.v[j] * lv.x[A.c[j]];
      lv.t[i] = lv.x[i] + (lv.b[i] - s) / lv.d[i];
    }
    lv.x.swap(lv.t);
  }
  // two-stage Gauss-Seidel sweep on x with T (strict lower or upper part of A)
  void two_stage(Level& lv, const CSR& A, const CSR& T) {
    const int nn = A.n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nn; ++i) {
      double s = 0.0;
      for (int j = A.r[i]; j < A.r[i + 1]; ++j) s += A.v[j] * lv.x[A.c[j]];
      lv.r[i] = lv.b[i] - s;                 // rho
      lv.g[i] = lv.r[i] / lv.d[i];           // g_0
    }
    for (int k = 0; k < inner; ++k) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < nn; ++i) {
        double s = 0.0;
        for (int j = T.r[i]; j < T.r[i + 1]; ++j) s += T.v[j] * lv.g[T.c[j]];
        lv.t[i] = (lv.r[i] - s) / lv.d[i];   // g_{k+1}
      }
      lv.g.swap(lv.t);
    }
    for (int i = 0; i < nn; ++i) lv.x[i] += lv.g[i];
  }
  // pre (down = true) or post smoother; the transposed cycle is the cycle on A^T with the
  // sweep directions of the other side reversed (its pre-smoother is the transpose of the
  // forward cycle's post-smoother)
  void smooth(Level& lv, bool tr, bool down) {
    const CSR& A = tr ? lv.At : lv.A;
    if (kind == FCF) { half_sweep(lv, A, 0); half_sweep(lv, A, 1); half_sweep(lv, A, 0); return; }
    if (kind != TWOSTAGE) { half_sweep(lv, A, -1); return; }
    int dir = down ? dir_down : dir_up;
    if (tr) dir = -(down ? dir_up : dir_down);
    const CSR& T = dir > 0 ? (tr ? lv.Lot : lv.Lo) : (tr ? lv.Upt : lv.Up);
    two_stage(lv, A, T);
  }
  void cycle(size_t l, bool tr) {
    if (l == L.size()) {   // coarsest: dense inverse, row-major
      for (int i = 0; i < nc; ++i) {
        double s = 0.0;
        for (int k = 0; k < nc; ++k) s += (tr ? Cinv[(size_t)k * nc + i] : Cinv[(size_t)i * nc + k]) * bc[k];
        xc[i] = s;
      }
      return;
    }
    Level& lv = L[l];
    const CSR& A = tr ? lv.At : lv.A;
    std::fill(lv.x.begin(), lv.x.end(), 0.0);
    smooth(lv, tr, true);
    // r = b - A x
#pragma omp parallel for schedule(static)
    for (int i = 0; i < A.n; ++i) {
      double s = 0.0;
      for (int j = A.r[i]; j < A.r[i + 1]; ++j) s += A.v[j] * lv.x[A.c[j]];
      lv.r[i] = lv.b[i] - s;
    }
    double* bnext = (l + 1 == L.size()) ? bc.data() : L[l + 1].b.data();
    spmv(lv.R, lv.r.data(), bnext);
    cycle(l + 1, tr);
    const double* xnext = (l + 1 == L.size()) ? xc.data() : L[l + 1].x.data();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < lv.P.n; ++i) {
      double s = 0.0;
      for (int j = lv.P.r[i]; j < lv.P.r[i + 1]; ++j) s += lv.P.v[j] * xnext[lv.P.c[j]];
      lv.x[i] += s;
    }
    smooth(lv, tr, false);
  }
  void run(const double* r, double* x, bool tr) {
    if (L.empty()) {
      std::memcpy(bc.data(), r, sizeof(double) * nc);
      cycle(0, tr);
      std::memcpy(x, xc.data(), sizeof(double) * nc);
      return;
    }
    if (tr && !have_t) {
      for (auto& lv : L) {
        lv.At = transpose(lv.A);
        if (kind == TWOSTAGE) {
          lv.Lot = strict_part(lv.At, true);
          lv.Upt = strict_part(lv.At, false);
        }
      }
      have_t = true;
    }
    std::memcpy(L[0].b.data(), r, sizeof(double) * n);
    cycle(0, tr);
    std::memcpy(x, L[0].x.data(), sizeof(double) * n);
  }
  void apply(const double* r, double* x) override { run(r, x, false); }
  void apply_t(const double* r, double* x) override { run(r, x, true); }
};

// CPR (precond/CPR.hpp)
struct CPR : Prec {
  deco_cpr_params p{};
  CSR* A = nullptr;          // the caller's matrix, decoupled in place (CPR.hpp:91-114)
  CSR Ap;
  int bs = 1, sn = 0, pid = 0;
  std::unique_ptr<Prec> glo, loc, jac;
  std::vector<double> rt, xp, W, x2;
  std::vector<double> tmp1, tmp2;
  CSR At;                    // transpose of the decoupled A (adjoint)
  bool have_t = false;

  void apply(const double* r, double* x) override;
  void apply_t(const double* r, double* x) override;
};

// fix_diag (precond/CPR.hpp:753-840) with small = big = 1e-20 as called at :76
static void fix_diag(CSR& A, double* b, int bs, std::vector<char>& flag, std::vector<int>& swapped) {
  const double small_tol = 1e-20, big_tol = 1e-20;
  const int sn = A.n / bs;
  flag.assign(A.n, 0);
  swapped.assign(sn, 0);
  for (int s = 0; s < sn; ++s) {
    const int i = s * bs;
    const int j0 = binarysearch(A.c.data(), A.r[i], A.r[i + 1], i);
    if (std::fabs(A.v[j0]) < small_tol) {
      A.v[j0] = (A.v[j0] > 0) ? small_tol : -small_tol;
      flag[i] = 1;
    }
    if (bs == 3) {
      const int j1 = binarysearch(A.c.data(), A.r[i + 1], A.r[i + 2], i + 1);
      const int j2 = binarysearch(A.c.data(), A.r[i + 2], A.r[i + 3], i + 2);
      if (std::fabs(A.v[j1]) < small_tol || std::fabs(A.v[j2]) < small_tol) {
        if (std::fabs(A.v[j1 + 1]) > small_tol || std::fabs(A.v[j2 - 1]) > small_tol) {
          const int ia1 = A.r[i + 1], ib1 = A.r[i + 2], ib2 = A.r[i + 3];
          std::vector<double> scratch(A.v.begin() + ia1, A.v.begin() + ib1);
          for (int j = ib1, k = ia1; k < ib1; ++j, ++k) A.v[k] = A.v[j];
          for (int j = ib1, k = 0; j < ib2; ++j, ++k) A.v[j] = scratch[k];
          if (b) std::swap(b[i + 1], b[i + 2]);
          swapped[s] = 1;
        } else {
          if (std::fabs(A.v[j1]) < small_tol) { A.v[j1] = (A.v[j1] > 0) ? big_tol : -big_tol; flag[i + 1] = 1; }
          if (std::fabs(A.v[j2]) < small_tol) { A.v[j2] = (A.v[j2] > 0) ? big_tol : -big_tol; flag[i + 2] = 1; }
        }
      }
    } else if (bs == 2) {
      const int j1 = binarysearch(A.c.data(), A.r[i + 1], A.r[i + 2], i + 1);
      if (std::fabs(A.v[j1]) < small_tol) { A.v[j1] = (A.v[j1] > 0) ? big_tol : -big_tol; flag[i + 1] = 1; }
    }
  }
}

// quasi_IMPES_GE (precond/CPR.hpp:372-447)
static std::vector<double> quasi_impes_ge(const CSR& A, int bs) {
  const int sn = A.n / bs, bs1 = bs - 1;
  std::vector<double> rhs(A.n, 0.0);
  for (int s = 0; s < sn; ++s) rhs[s * bs] = 1.0;
  for (int s = 0; s < sn; ++s) {
    double Ap[4] = {0, 0, 0, 0}, bp[2] = {0, 0};
    for (int p = 0; p < bs; ++p) {
      const int i = s * bs + p;
      const int j = binarysearch(A.c.data(), A.r[i], A.r[i + 1], i);
      for (int k = 1; k < bs; ++k) {
        if (p == 0) bp[k - 1] = -A.v[j + k - p];
        else Ap[(p - 1) * bs1 + A.c[j + k - p] % bs - 1] = A.v[j + k - p];
      }
    }
    if (bs1 == 2) {
      const double det = Ap[0] * Ap[3] - Ap[1] * Ap[2];
      if (det == 0.0) throw std::domain_error("determinant is zero in CPD decoupling");
      const double invdet = 1.0 / det;
      const double i0 = Ap[3] * invdet, i1 = -Ap[1] * invdet, i2 = -Ap[2] * invdet, i3 = Ap[0] * invdet;
      rhs[s * bs + 1] = i0 * bp[0] + i2 * bp[1];
      rhs[s * bs + 2] = i1 * bp[0] + i3 * bp[1];
    } else if (bs1 == 1) {
      if (Ap[0] == 0.0) throw std::domain_error("determinant is zero in CPD decoupling");
      rhs[s * bs + 1] = bp[0] / Ap[0];
    } else {
      throw std::invalid_argument("unknown block size CPD decoupling");
    }
  }
  return rhs;
}

// quasi_IMPES (precond/CPR.hpp:322-369): per cell B^T w = e_p with LU (dgesv)
static std::vector<double> quasi_impes(const CSR& A, int bs) {
  const int sn = A.n / bs;
  std::vector<double> rhs(A.n, 0.0);
  for (int s = 0; s < sn; ++s) rhs[s * bs] = 1.0;
  std::vector<double> M(bs * bs);
  std::vector<int> ipiv(bs);
  for (int s = 0; s < sn; ++s) {
    std::fill(M.begin(), M.end(), 0.0);
    for (int p = 0; p < bs; ++p) {
      const int i = s * bs + p;
      const int ia = A.r[i], ib = A.r[i + 1];
      const int j = binarysearch(A.c.data(), ia, ib, i);
      if (A.c[j] == i)
        for (int k = 0; k < bs; ++k) {
          const int jj = j + k - p;
          if (jj < ib && jj >= ia && A.c[jj] / bs == s) M[p * bs + A.c[jj] % bs] = A.v[jj];  // row-major B
        }
    }
    // LAPACK reads M column-major (= B^T): solve B^T w = e_p
    lu_factor(bs, M.data(), ipiv.data());
    lu_solve(bs, M.data(), ipiv.data(), 0, rhs.data() + (size_t)s * bs);
  }
  return rhs;
}

// drs_weights (precond/CPR.hpp:518-584); the diagonal is compared signed (CPR.hpp:545)
static std::vector<char> drs_weights(const CSR& A, int bs, double e_dd, double e_ps) {
  const int sn = A.n / bs;
  std::vector<char> w(A.n, 1);
  for (int s = 0; s < sn; ++s)
    for (int p = 0; p < bs; ++p) {
      const int i = s * bs + p;
      const int ia = A.r[i], ib = A.r[i + 1];
      const double diag = A.v[binarysearch(A.c.data(), ia, ib, i - p)];
      double off = -std::fabs(diag);
      for (int j = ia; j < ib; ++j) off += (A.c[j] % bs == 0) ? std::fabs(A.v[j]) : 0.0;
      w[i] = (diag < e_dd * off) ? 0 : 1;
    }
  if (e_ps != 0.0) {
    std::vector<double> off(bs);
    for (int s = 0; s < sn; ++s) {
      const int i = s * bs;
      const int ia = A.r[i], ib = A.r[i + 1];
      const int jd = binarysearch(A.c.data(), ia, ib, i);
      const double diag = std::fabs(A.v[jd]);
      std::fill(off.begin(), off.end(), 0.0);
      for (int j = ia; j < ib; ++j) {
        if (j == jd) continue;
        off[A.c[j] % bs] += std::fabs(A.v[j]);
      }
      for (int p = 0; p < bs; ++p) {
        const bool ok = !(off[p] < e_ps * diag);
        w[i + p] = (char)(w[i + p] && ok);
      }
    }
  }
  for (int s = 0; s < sn; ++s) w[s * bs] = 1;
  return w;
}

static CPR* cpr_create(const deco_cpr_params& p, CSR& A, double* b) {
  if (p.pressure_id != 0) throw std::invalid_argument("pressure_id must be 0");
  auto* M = new CPR();
  M->p = p;
  M->A = &A;
  M->bs = p.block_size;
  M->pid = p.pressure_id;
  M->n = A.n;
  if (M->bs <= 0 || A.n % M->bs != 0) { delete M; throw std::invalid_argument("n is not divisible by bs"); }
  const int bs = M->bs, n = A.n, sn = n / bs;
  M->sn = sn;
  M->xp.assign(sn, 0.0);
  M->rt.assign(n, 0.0);
  M->x2.assign(n, 0.0);
  M->tmp1.assign(n, 0.0);
  M->tmp2.assign(n, 0.0);

  // per-cell left transform, starts as the permutation of fix_diag
  std::vector<char> flag;
  std::vector<int> swapped;
  fix_diag(A, b, bs, flag, swapped);                         // :76
  std::vector<double> W((size_t)sn * bs * bs, 0.0);
  for (int s = 0; s < sn; ++s)
    for (int q = 0; q < bs; ++q) W[(size_t)s * bs * bs + q * bs + q] = 1.0;
  for (int s = 0; s < sn; ++s)
    if (swapped[s]) {
      double* Ws = W.data() + (size_t)s * bs * bs;
      for (int q = 0; q < bs; ++q) std::swap(Ws[1 * bs + q], Ws[2 * bs + q]);
    }
  auto left_mul = [&](int s, const double* T) {               // W_s <- T W_s
    double* Ws = W.data() + (size_t)s * bs * bs;
    std::vector<double> o(bs * bs, 0.0);
    for (int a = 0; a < bs; ++a)
      for (int c = 0; c < bs; ++c) {
        double t = 0.0;
        for (int k = 0; k < bs; ++k) t += T[a * bs + k] * Ws[k * bs + c];
        o[a * bs + c] = t;
      }
    std::copy(o.begin(), o.end(), Ws);
  };

  std::vector<double> w;
  if (p.weights == DECO_W_QUASIIMPES) w = quasi_impes(A, bs);           // :79-80
  else if (p.weights == DECO_W_QUASIIMPES_GE) w = quasi_impes_ge(A, bs);  // :81-82
  else if (p.weights != DECO_W_NONE) { delete M; throw std::invalid_argument("Unknown weights type"); }
  if (p.weights != DECO_W_NONE) {                              // :89-93
    diag_scale(A, w.data());
    if (b) for (int i = 0; i < n; ++i) b[i] *= w[i];
  }
  std::vector<char> wd;
  if (p.use_drs) {                                             // :96-101
    wd = drs_weights(A, bs, p.drs_e_dd, p.drs_e_ps);
    if (p.weights != DECO_W_NONE)                              // unscale_drs_weights :586-600
      for (int i = 0; i < n; ++i)
        if (!wd[i]) {
          const double d = w[i];
          for (int j = A.r[i]; j < A.r[i + 1]; ++j) A.v[j] /= d;
          if (b) b[i] /= d;
          w[i] = 1.0;
        }
  }
  if (p.weights != DECO_W_NONE)
    for (int s = 0; s < sn; ++s) {
      std::vector<double> T(bs * bs, 0.0);
      for (int q = 0; q < bs; ++q) T[q * bs + q] = w[s * bs + q];
      left_mul(s, T.data());
    }
  // IMPES (:104-105 -> :475-516, :450-471): pressure row = sum_k wd_k row_k (positional)
  for (int s = 0; s < sn; ++s) {
    const int ip = s * bs;
    double* table = A.v.data() + A.r[ip];
    for (int k = 0; k < bs; ++k) {
      if (k == 0) continue;
      const int i = s * bs + k;
      const double wk = wd.empty() ? 1.0 : (double)wd[i];
      for (int j = A.r[i]; j < A.r[i + 1]; ++j) {
        if (wd.empty()) table[j - A.r[i]] += A.v[j];
        else table[j - A.r[i]] += wk * A.v[j];
      }
    }
    if (b) {
      double sum = 0.0;
      for (int k = 0; k < bs; ++k) sum += (wd.empty() ? b[s * bs + k] : (double)wd[s * bs + k] * b[s * bs + k]);
      b[s * bs] = sum;
    }
    std::vector<double> T(bs * bs, 0.0);
    for (int q = 0; q < bs; ++q) T[q * bs + q] = 1.0;
    for (int k = 0; k < bs; ++k) T[k] = wd.empty() ? 1.0 : (double)wd[s * bs + k];
    left_mul(s, T.data());
  }
  if (p.cpd) {                                                 // :108-116
    std::vector<double> J((size_t)sn * bs * bs);
    for (int s = 0; s < sn; ++s) {                             // CPD_build :603-685
      double Bm[9], Dp[3];
      for (int q = 0; q < bs; ++q) {
        const int i = s * bs + q;
        const int j = binarysearch(A.c.data(), A.r[i], A.r[i + 1], i);
        Dp[q] = A.v[j];
        for (int k = 0; k < bs; ++k) Bm[q * bs + A.c[j + k - q] % bs] = A.v[j + k - q];
      }
      double* vj = J.data() + (size_t)s * bs * bs;
      if (bs == 3) {
        const double* a = Bm;
        const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[3] * (a[1] * a[8] - a[2] * a[7]) +
                           a[6] * (a[1] * a[5] - a[2] * a[4]);
        if (det == 0.0) { delete M; throw std::domain_error("determinant is zero in CPD decoupling"); }
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
      } else if (bs == 2) {
        const double det = Bm[0] * Bm[3] - Bm[1] * Bm[2];
        if (det == 0.0) { delete M; throw std::domain_error("determinant is zero in CPD decoupling"); }
        const double invdet = 1.0 / det;
        vj[0] = Bm[3] * invdet * Dp[0];
        vj[1] = -Bm[1] * invdet * Dp[0];
        vj[2] = -Bm[2] * invdet * Dp[1];
        vj[3] = Bm[0] * invdet * Dp[1];
      } else {
        delete M;
        throw std::invalid_argument("unknown block size CPD decoupling");
      }
    }
    const std::vector<double> scr = A.v;                       // CPD_apply (mat) :688-726
    for (int s = 0; s < sn; ++s)
      for (int q = 0; q < bs; ++q) {
        const int i = s * bs + q;
        const double* vj = J.data() + (size_t)s * bs * bs + q * bs;
        const int ia = A.r[i], ib = A.r[i + 1];
        for (int j = ia; j < ib; ++j) A.v[j] *= vj[q];
        for (int k = 0; k < bs; ++k) {
          if (k == q) continue;
          const int i2 = s * bs + k;
          const int ia2 = A.r[i2], ib2 = A.r[i2 + 1];
          for (int j = ia2; j < ib2; ++j) A.v[ia + (j - ia2)] += scr[j] * vj[k];
        }
      }
    if (b)                                                     // CPD_apply (vec) :730-750
      for (int s = 0; s < sn; ++s) {
        double sum[3];
        for (int q = 0; q < bs; ++q) {
          const double* vj = J.data() + (size_t)(s * bs + q) * bs;
          sum[q] = 0.0;
          for (int k = 0; k < bs; ++k) sum[q] += vj[k] * b[s * bs + k];
        }
        for (int q = 0; q < bs; ++q) b[s * bs + q] = sum[q];
      }
    for (int s = 0; s < sn; ++s) left_mul(s, J.data() + (size_t)s * bs * bs);
  }
  M->W.swap(W);
  if (p.tol_zero != 0.0) filter(A, p.tol_zero, false);         // :118-123
  M->Ap = local_matrix(A, bs, 0, nullptr);                     // :127
  if (p.loc_tol_rel > 1e-20) filter(M->Ap, p.loc_tol_rel, true);  // :130-136
  if (p.glo_type == DECO_GLO_FSAI) M->glo.reset(new FSAI(p, A));  // :142
  else M->glo.reset(new Jacobi(p.glo_block_size, A));
  if (p.seq == 2) {                                            // :154-163
    if (p.glo_type == DECO_GLO_JACOBI) M->jac.reset(new Jacobi(p.glo_block_size, A));
    else M->jac.reset(new Jacobi(p.glo_block_size, A));
  }
  for (int i = 0; i < n; ++i)                                  // undo fix_diag :169, :842-855
    if (flag[i]) A.v[binarysearch(A.c.data(), A.r[i], A.r[i + 1], i)] = 0.0;
  return M;
}

// apply_seq0..3 (precond/CPR.hpp:176-299)
void CPR::apply(const double* r, double* x) {
  if (!loc) throw std::runtime_error("CPR local stage not attached");
  const int seq = p.seq;
  if (seq == 1) {                                              // :176-200
    for (int i = 0; i < sn; ++i) rt[i] = r[i * bs + pid];
    loc->apply(rt.data(), xp.data());
    std::fill(x, x + n, 0.0);
    for (int i = 0; i < sn; ++i) x[i * bs + pid] = xp[i];
    spmv(*A, x, rt.data());
    for (int i = 0; i < n; ++i) rt[i] = r[i] - rt[i];
    glo->apply(rt.data(), x);
    for (int i = 0; i < sn; ++i) x[i * bs + pid] += xp[i];
    return;
  }
  if (seq == 2) jac->apply(r, x); else glo->apply(r, x);      // :205 / :238 / :279
  spmv(*A, x, rt.data(), bs);                                  // rows bs*i: pressure rows as pid = 0
  for (int i = 0; i < sn; ++i) rt[i] = r[i * bs + pid] - rt[i];
  loc->apply(rt.data(), xp.data());
  for (int i = 0; i < sn; ++i) x[i * bs + pid] += xp[i];
  if (seq == 0) return;
  spmv(*A, x, rt.data());
  for (int i = 0; i < n; ++i) rt[i] = r[i] - rt[i];
  glo->apply(rt.data(), tmp1.data());   // the reference aliases in and out here (CPR.hpp:265)
  for (int i = 0; i < n; ++i) x[i] += tmp1[i];
}

// Transpose of seq3: M = F + (I - F A)(F + P G R (I - A F)) gives
// M^T u = a + e + F^T (s - A^T e), a = F^T u, s = u - A^T a, e = P G^T R s.
void CPR::apply_t(const double* u, double* x) {
  if (p.seq != 3) throw std::runtime_error("transposed CPR implemented for seq 3");
  if (!have_t) { At = transpose(*A); have_t = true; }
  std::vector<double>& a = tmp1;
  std::vector<double>& s = tmp2;
  glo->apply_t(u, a.data());
  spmv(At, a.data(), s.data());
  for (int i = 0; i < n; ++i) s[i] = u[i] - s[i];
  for (int i = 0; i < sn; ++i) rt[i] = s[i * bs + pid];
  loc->apply_t(rt.data(), xp.data());
  std::fill(x2.begin(), x2.end(), 0.0);
  for (int i = 0; i < sn; ++i) x2[i * bs + pid] = xp[i];
  std::vector<double> ate(n);
  spmv(At, x2.data(), ate.data());
  for (int i = 0; i < n; ++i) ate[i] = s[i] - ate[i];
  glo->apply_t(ate.data(), x);
  for (int i = 0; i < n; ++i) x[i] += a[i] + x2[i];
}

// ------------------------------------------------------------------ solvers ------------
static void apply_givens(double c, double s, double& dx, double& dy) {   // rgmres.tpp:156-160
  const double t = c * dx + s * dy;
  dy = -s * dx + c * dy;
  dx = t;
}
static void comp_givens(double x, double y, double& c, double& s) {      // rgmres.tpp:141-154
  const double EPS = 2.220446049250312e-16;
  if (std::fabs(y) < EPS) { c = 1.0; s = 0.0; }
  else if (std::fabs(y) > std::fabs(x)) { const double nu = x / y; s = 1.0 / std::sqrt(1.0 + nu * nu); c = nu * s; }
  else { const double nu = y / x; c = 1.0 / std::sqrt(1.0 + nu * nu); s = nu * c; }
}

using ApplyA = std::function<void(const double*, double*)>;
using ApplyM = std::function<void(const double*, double*)>;

// rgmres (solver/rgmres.tpp:9-139) and lgmres (solver/lgmres.tpp:2-134)
static void gmres(bool left, const deco_sol_params& sp, int n, const ApplyA& Aop, const ApplyM& M,
                  const double* b, double* x, deco_sol_info& info, double* resvec) {
  const double tol = sp.reltol;
  const int m = sp.restart, itmax = sp.itmax;
  std::vector<double> r(n), z(n), w(n), V((size_t)n * (m + 1)), H((size_t)(m + 1) * (m + 3), 0.0);
  double* g = H.data() + (size_t)(m + 1) * m;
  double* cs = g + (m + 1);
  double* sn = cs + (m + 1);
  auto Hm = [&](int i, int j) -> double& { return H[i + (size_t)j * (m + 1)]; };
  int flag = 1, iter = 0;
  double bnorm, absTol, rnorm = 0.0;
  if (!left) {
    bnorm = nrm2(n, b);
    absTol = tol * bnorm;
    if (!sp.use_x0) { std::fill(x, x + n, 0.0); std::memcpy(r.data(), b, sizeof(double) * n); }
    else { Aop(x, r.data()); for (int i = 0; i < n; ++i) r[i] = b[i] - r[i]; }
  } else {
    if (!sp.use_x0) { std::fill(x, x + n, 0.0); M(b, r.data()); }
    else { Aop(x, z.data()); for (int i = 0; i < n; ++i) z[i] = b[i] - z[i]; M(z.data(), r.data()); }
    bnorm = nrm2(n, r.data());
    absTol = tol * bnorm;
  }
  int napp = 0, nspmv = 0;
  while (iter < itmax) {
    std::fill(H.begin(), H.end(), 0.0);
    g[0] = nrm2(n, r.data());
    const double inv0 = 1.0 / g[0];
    for (int i = 0; i < n; ++i) V[i] = inv0 * r[i];
    int j = 0;
    while (j < m && iter < itmax) {
      rnorm = std::fabs(g[j]);
      if (resvec) resvec[iter] = rnorm / bnorm;
      if (rnorm < absTol) { flag = 0; break; }
      double* vj = V.data() + (size_t)j * n;
      if (!left) { M(vj, z.data()); Aop(z.data(), w.data()); }
      else { Aop(vj, z.data()); M(z.data(), w.data()); }
      ++napp; ++nspmv;
      for (int i = 0; i <= j; ++i) {
        const double* vi = V.data() + (size_t)i * n;
        Hm(i, j) = dot(n, vi, w.data());
        const double h = -Hm(i, j);
        for (int k = 0; k < n; ++k) w[k] = 1.0 * w[k] + h * vi[k];
      }
      Hm(j + 1, j) = nrm2(n, w.data());
      const double inv = 1.0 / Hm(j + 1, j);
      double* vn = V.data() + (size_t)(j + 1) * n;
      for (int k = 0; k < n; ++k) vn[k] = inv * w[k];
      for (int i = 0; i <= j - 1; ++i) apply_givens(cs[i], sn[i], Hm(i, j), Hm(i + 1, j));
      comp_givens(Hm(j, j), Hm(j + 1, j), cs[j], sn[j]);
      apply_givens(cs[j], sn[j], Hm(j, j), Hm(j + 1, j));
      apply_givens(cs[j], sn[j], g[j], g[j + 1]);
      ++j;
      ++iter;
    }
    for (int i = j - 1; i >= 0; --i) {
      if (Hm(i, i) == 0.0) throw std::domain_error("Zero diagonal element in upper triangular matrix");
      g[i] /= Hm(i, i);
      for (int k = 0; k < i; ++k) g[k] -= Hm(k, i) * g[i];
    }
    if (j > 0) {
      for (int k = 0; k < n; ++k) w[k] = g[0] * V[k];
      for (int i = 1; i < j; ++i) {
        const double* vi = V.data() + (size_t)i * n;
        for (int k = 0; k < n; ++k) w[k] = 1.0 * w[k] + g[i] * vi[k];
      }
    } else {
      for (int k = 0; k < n; ++k) w[k] = g[0] * V[k];
    }
    if (!left) {
      M(w.data(), z.data()); ++napp;
      for (int k = 0; k < n; ++k) x[k] += z[k];
      Aop(x, r.data()); ++nspmv;
      for (int k = 0; k < n; ++k) r[k] = b[k] - r[k];
    } else {
      for (int k = 0; k < n; ++k) x[k] += w[k];
      Aop(x, z.data()); ++nspmv;
      for (int k = 0; k < n; ++k) z[k] = b[k] - z[k];
      M(z.data(), r.data()); ++napp;
    }
    rnorm = nrm2(n, r.data());
    if (rnorm < absTol) { flag = 0; break; }
  }
  info.iter = iter;
  info.res_abs = rnorm;
  info.res_rel = rnorm / bnorm;
  info.flag = flag;
  info.bnorm = bnorm;
  info.n_prec = napp;
  info.n_spmv = nspmv;
}

// bicgstab (solver/bicgstab.tpp:2-99) and lbicgstab (solver/lbicgstab.tpp:2-95)
static void bicgstab(bool left, const deco_sol_params& sp, int n, const ApplyA& Aop, const ApplyM& M,
                     const double* b, double* x, deco_sol_info& info, double* resvec) {
  std::vector<double> r(n), rt(n), p(n), y(n), v(n), t(n);
  double rho_old, rho = 1, alpha = 1, beta, omega = 1;
  const double reltol = sp.reltol;
  const int itmax = sp.itmax;
  float it = 0.0f;
  int flag = 1, napp = 0, nspmv = 0;
  double norm_b, tol, res_norm;
  if (!left) {
    norm_b = nrm2(n, b);
    tol = norm_b * reltol;
    if (!sp.use_x0) { std::fill(x, x + n, 0.0); std::memcpy(r.data(), b, sizeof(double) * n); }
    else { Aop(x, r.data()); for (int i = 0; i < n; ++i) r[i] = b[i] - r[i]; }
    rt = r;
    res_norm = nrm2(n, r.data());
  } else {
    if (!sp.use_x0) { std::fill(x, x + n, 0.0); M(b, r.data()); }
    else { Aop(x, y.data()); for (int i = 0; i < n; ++i) y[i] = b[i] - y[i]; M(y.data(), r.data()); }
    rt = r;
    res_norm = nrm2(n, r.data());
    norm_b = res_norm;
    tol = norm_b * reltol;
  }
  int hist = 0;
  for (it = 0.5f; it < itmax; it += 0.5f) {
    rho_old = rho;
    rho = dot(n, r.data(), rt.data());
    if (it < 1) p = r;
    else {
      beta = (std::fabs(res_norm) < 1e-40) ? 0.0 : (rho / rho_old) * (alpha / omega);
      // axpby(n, 1.0, r, -beta*omega, v, beta, p): p = beta*p + (-beta*omega)*v + 1.0*r
      const double a1 = -beta * omega, d1 = 1.0;
      for (int i = 0; i < n; ++i) p[i] = beta * p[i] + a1 * v[i] + d1 * r[i];
    }
    if (!left) { M(p.data(), y.data()); Aop(y.data(), v.data()); }
    else { Aop(p.data(), y.data()); M(y.data(), v.data()); }
    ++napp; ++nspmv;
    alpha = (std::fabs(res_norm) < 1e-40) ? 0.0 : rho / dot(n, v.data(), rt.data());
    for (int i = 0; i < n; ++i) r[i] = 1.0 * r[i] + (-alpha) * v[i];
    if (!left) for (int i = 0; i < n; ++i) x[i] = 1.0 * x[i] + alpha * y[i];
    res_norm = nrm2(n, r.data());
    if (resvec) resvec[hist++] = res_norm / norm_b;
    if (res_norm <= tol) {
      // lbicgstab.tpp:193 updates x with y = A p (as written in the reference)
      if (left) for (int i = 0; i < n; ++i) x[i] = 1.0 * x[i] + alpha * y[i];
      flag = 0;
      break;
    }
    it += 0.5f;
    if (!left) { M(r.data(), y.data()); Aop(y.data(), t.data()); }
    else { Aop(r.data(), y.data()); M(y.data(), t.data()); }
    ++napp; ++nspmv;
    omega = (std::fabs(res_norm) < 1e-40) ? 0.0 : dot(n, t.data(), r.data()) / dot(n, t.data(), t.data());
    if (!left) {
      for (int i = 0; i < n; ++i) r[i] = 1.0 * r[i] + (-omega) * t[i];
      for (int i = 0; i < n; ++i) x[i] = 1.0 * x[i] + omega * y[i];
    } else {
      // lbicgstab.tpp:202: x = x + (alpha*omega) * y with y = A r (as written in the reference)
      for (int i = 0; i < n; ++i) x[i] = 1.0 * x[i] + (alpha * omega) * y[i];
      for (int i = 0; i < n; ++i) r[i] = 1.0 * r[i] + (-omega) * t[i];
    }
    res_norm = nrm2(n, r.data());
    if (resvec) resvec[hist++] = res_norm / norm_b;
    if (res_norm <= tol) { flag = 0; break; }
  }
  info.iter = (int)it;
  info.res_abs = res_norm;
  info.res_rel = res_norm / norm_b;
  info.flag = flag;
  info.bnorm = norm_b;
  info.n_prec = napp;
  info.n_spmv = nspmv;
}

// pcg (solver/pcg.tpp:3-62)
static void pcg(const deco_sol_params& sp, int n, const ApplyA& Aop, const ApplyM& M, const double* b, double* x,
                deco_sol_info& info, double* resvec) {
  std::vector<double> r(n), p(n), z(n), g(n);
  const double norm_b = nrm2(n, b), tol = norm_b * sp.reltol;
  int it = 0, napp = 0, nspmv = 0;
  if (!sp.use_x0) { std::fill(x, x + n, 0.0); std::memcpy(r.data(), b, sizeof(double) * n); }
  else { Aop(x, r.data()); for (int i = 0; i < n; ++i) r[i] = b[i] - r[i]; }
  M(r.data(), p.data()); ++napp;
  double rho1 = dot(n, r.data(), p.data());
  double res_norm = nrm2(n, r.data());
  while (res_norm > tol && it < sp.itmax) {
    Aop(p.data(), z.data()); ++nspmv;
    const double alpha = rho1 / dot(n, z.data(), p.data());
    for (int i = 0; i < n; ++i) x[i] = 1.0 * x[i] + alpha * p[i];
    for (int i = 0; i < n; ++i) r[i] = 1.0 * r[i] + (-alpha) * z[i];
    M(r.data(), g.data()); ++napp;
    const double rho2 = dot(n, r.data(), g.data());
    const double beta = rho2 / rho1;
    rho1 = rho2;
    for (int i = 0; i < n; ++i) p[i] = beta * p[i] + 1.0 * g[i];
    res_norm = nrm2(n, r.data());
    if (resvec) resvec[it] = res_norm / norm_b;
    ++it;
  }
  info.iter = it;
  info.res_abs = res_norm;
  info.res_rel = res_norm / norm_b;
  info.flag = res_norm > tol ? 1 : 0;
  info.bnorm = norm_b;
  info.n_prec = napp;
  info.n_spmv = nspmv;
}

static void solve(const deco_sol_params& sp, int n, const ApplyA& Aop, const ApplyM& M, const double* b, double* x,
                  deco_sol_info& info, double* resvec) {
  switch (sp.type) {
    case DECO_SOL_RGMRES: gmres(false, sp, n, Aop, M, b, x, info, resvec); break;
    case DECO_SOL_LGMRES: gmres(true, sp, n, Aop, M, b, x, info, resvec); break;
    case DECO_SOL_BICGSTAB: bicgstab(false, sp, n, Aop, M, b, x, info, resvec); break;
    case DECO_SOL_LBICGSTAB: bicgstab(true, sp, n, Aop, M, b, x, info, resvec); break;
    case DECO_SOL_PCG: pcg(sp, n, Aop, M, b, x, info, resvec); break;
    default: throw std::invalid_argument("Invalid solver type");
  }
}

}  // namespace orc

struct orc_prec {
  std::unique_ptr<orc::Prec> impl;
};

using namespace orc;

static thread_local std::string t_err;

template <class F>
static int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::domain_error& e) {
    t_err = e.what();
    return DECO_ERR_SINGULAR;
  } catch (const std::exception& e) {
    t_err = e.what();
    return DECO_ERR_ARG;
  }
}

extern "C" {

const char* orc_last_error(void) { return t_err.c_str(); }
void orc_free(void* p) { free(p); }
void orc_csr_free(orc_csr* A) {
  if (!A) return;
  free(A->r); free(A->c); free(A->v);
  A->r = nullptr; A->c = nullptr; A->v = nullptr;
}
int orc_csr_clone(const orc_csr* A, orc_csr* out) { return guard([&] { to_c(from_c(A), out); }); }
int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int t) { if (t > 0) omp_set_num_threads(t); }

int orc_spmv(const orc_csr* A, const double* x, double* y, int bs) {
  // alg::spmv (core/alg.hpp:51-64) directly on the caller's arrays
  const int n = A->n;
  const int* r = A->r;
  const int* c = A->c;
  const double* v = A->v;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i += bs) {
    double s = 0.0;
    for (int j = r[i]; j < r[i + 1]; ++j) s += v[j] * x[c[j]];
    y[i / bs] = s;
  }
  return 0;
}
double orc_dot(int n, const double* x, const double* y) { return dot(n, x, y); }
double orc_nrm2(int n, const double* x) { return nrm2(n, x); }

// in place: arrays keep their allocation, A->nnz shrinks
// AD::block_reorder (adapter/csr2coo.cpp:19-88): COO row indices, rows and columns mapped
// inside each cell (:54-66), one sort of the entries by (row, column) (:69-72), row pointers from
// the sorted rows (:75), b permuted per cell (:78-86). Returns false when A is left unchanged.
static bool block_reorder(CSR& A, double* b, int bs, const char* row_ord, const char* col_ord) {
  if (bs < 1 || bs > 9) throw std::invalid_argument("block size out of range");
  if (A.n % bs || A.m % bs) throw std::invalid_argument("n not divisible by bs");
  std::vector<int> reord_row(bs), reord_col(bs), ord_row(bs), seen_r(bs, 0), seen_c(bs, 0);
  if (bs == 1) return false;
  for (int i = 0; i < bs; ++i) {
    const int row = row_ord[i] - '1', col = col_ord[i] - '1';
    if (row < 0 || row >= bs || col < 0 || col >= bs || seen_r[row] || seen_c[col])
      throw std::invalid_argument("order is not a permutation");
    seen_r[row] = seen_c[col] = 1;
    reord_row[row] = i;
    reord_col[col] = i;
    ord_row[i] = row;
  }
  if (bs == 2 && reord_row[0] == 0 && reord_row[1] == 1 && reord_col[0] == 0 && reord_col[1] == 1) return false;
  const int nz = A.nnz();
  std::vector<int> l(nz);
  for (int i = 0; i < A.n; ++i)
    for (int k = A.r[i]; k < A.r[i + 1]; ++k) l[k] = i;
  for (int k = 0; k < nz; ++k) {
    const int rr = l[k] % bs;
    l[k] = l[k] - rr + reord_row[rr];
    const int rc = A.c[k] % bs;
    A.c[k] = A.c[k] - rc + reord_col[rc];
  }
  std::vector<int> idx(nz);
  for (int k = 0; k < nz; ++k) idx[k] = k;
  std::sort(idx.begin(), idx.end(), [&](int a, int b2) {
    if (l[a] != l[b2]) return l[a] < l[b2];
    return A.c[a] < A.c[b2];
  });
  std::vector<int> c2(nz);
  std::vector<double> v2(nz);
  std::fill(A.r.begin(), A.r.end(), 0);
  for (int k = 0; k < nz; ++k) {
    c2[k] = A.c[idx[k]];
    v2[k] = A.v[idx[k]];
    ++A.r[l[idx[k]] + 1];
  }
  for (int i = 0; i < A.n; ++i) A.r[i + 1] += A.r[i];
  A.c.swap(c2);
  A.v.swap(v2);
  if (b) {
    std::vector<double> scratch(bs);
    for (int s0 = 0; s0 < A.n / bs; ++s0) {
      for (int q = 0; q < bs; ++q) scratch[q] = b[s0 * bs + q];
      for (int q = 0; q < bs; ++q) b[s0 * bs + q] = scratch[ord_row[q]];
    }
  }
  return true;
}

static void write_back(const CSR& B, orc_csr* A) {
  A->nnz = B.nnz();
  std::memcpy(A->r, B.r.data(), sizeof(int) * (B.n + 1));
  if (B.nnz()) {
    std::memcpy(A->c, B.c.data(), sizeof(int) * B.nnz());
    std::memcpy(A->v, B.v.data(), sizeof(double) * B.nnz());
  }
}
int orc_block_reorder(orc_csr* A, double* b, int bs, const char* row_order, const char* col_order) {
  return guard([&] { CSR B = from_c(A); if (block_reorder(B, b, bs, row_order, col_order)) write_back(B, A); });
}
int orc_filter(orc_csr* A, double tol) {
  return guard([&] { CSR B = from_c(A); filter(B, tol, false); write_back(B, A); });
}
int orc_filter_proper(orc_csr* A, double tol) {
  return guard([&] { CSR B = from_c(A); filter(B, tol, true); write_back(B, A); });
}
int orc_local_matrix(const orc_csr* A, int bs, int p, const double* w, orc_csr* out) {
  return guard([&] { to_c(local_matrix(from_c(A), bs, p, w), out); });
}
int orc_trans_mat_full(const orc_csr* A, orc_csr* out) {
  return guard([&] { to_c(transpose(from_c(A)), out); });
}
int orc_spgemm(const orc_csr* A, const orc_csr* B, orc_csr* C) {
  return guard([&] { to_c(spgemm(from_c(A), from_c(B)), C); });
}
int orc_fsai_pattern(const orc_csr* A, int power, orc_csr* F) {
  return guard([&] { to_c(fsai_pattern(from_c(A), power), F); });
}

int orc_fix_diag(orc_csr* A, double* b, int bs, unsigned char* check_diag) {
  return guard([&] {
    CSR B = from_c(A);
    std::vector<char> flag;
    std::vector<int> sw;
    fix_diag(B, b, bs, flag, sw);
    write_back(B, A);
    if (check_diag) for (int i = 0; i < B.n; ++i) check_diag[i] = (unsigned char)flag[i];
  });
}
int orc_quasi_impes_ge(const orc_csr* A, int bs, double* w) {
  return guard([&] { auto v = quasi_impes_ge(from_c(A), bs); std::copy(v.begin(), v.end(), w); });
}
int orc_quasi_impes(const orc_csr* A, int bs, double* w) {
  return guard([&] { auto v = quasi_impes(from_c(A), bs); std::copy(v.begin(), v.end(), w); });
}
int orc_drs_weights(const orc_csr* A, int bs, double e_dd, double e_ps, unsigned char* w) {
  return guard([&] {
    auto v = drs_weights(from_c(A), bs, e_dd, e_ps);
    for (size_t i = 0; i < v.size(); ++i) w[i] = (unsigned char)v[i];
  });
}

int orc_fsai_create(const deco_cpr_params* p, const orc_csr* A, orc_prec** out) {
  return guard([&] { *out = new orc_prec{std::unique_ptr<Prec>(new FSAI(*p, from_c(A)))}; });
}
int orc_jacobi_create(int block_size, const orc_csr* A, orc_prec** out) {
  return guard([&] { *out = new orc_prec{std::unique_ptr<Prec>(new Jacobi(block_size, from_c(A)))}; });
}
int orc_dense_create(const orc_csr* A, orc_prec** out) {
  return guard([&] { *out = new orc_prec{std::unique_ptr<Prec>(new Dense(from_c(A)))}; });
}
int orc_identity_create(int n, orc_prec** out) {
  return guard([&] { *out = new orc_prec{std::unique_ptr<Prec>(new Identity(n))}; });
}
int orc_amg_build(const orc_csr* A0, int nlev, const orc_csr* P, const int* const* cf, int relax_type,
                  double relax_weight, double drop_tol, orc_prec** out) {
  return guard([&] {
    if (!A0 || nlev < 0 || (nlev > 0 && (!P || !cf)) || !out) throw std::invalid_argument("null argument");
    auto* G = new AMG();
    std::unique_ptr<Prec> own(G);
    G->set_relax(relax_type, relax_weight);
    G->build(from_c(A0), nlev, P, cf, drop_tol);
    *out = new orc_prec{std::move(own)};
  });
}
static AMG* as_amg(const orc_prec* M) {
  auto* G = dynamic_cast<AMG*>(M->impl.get());
  if (!G) throw std::invalid_argument("not an AMG preconditioner");
  return G;
}
int orc_amg_num_levels(const orc_prec* M, int* nlev, int* nc) {
  return guard([&] {
    AMG* G = as_amg(M);
    if (nlev) *nlev = (int)G->L.size();
    if (nc) *nc = G->nc;
  });
}
int orc_amg_level(const orc_prec* M, int level, orc_csr* A, orc_csr* R, double* d) {
  return guard([&] {
    AMG* G = as_amg(M);
    if (level < 0 || level > (int)G->L.size()) throw std::invalid_argument("level out of range");
    if (level == (int)G->L.size()) {
      if (A) to_c(G->Ac, A);
      return;
    }
    const auto& lv = G->L[level];
    if (A) to_c(lv.A, A);
    if (R) to_c(lv.R, R);
    if (d) std::memcpy(d, lv.d.data(), sizeof(double) * lv.d.size());
  });
}
int orc_amg_coarse_inverse(const orc_prec* M, double* inv) {
  return guard([&] {
    AMG* G = as_amg(M);
    if (inv && G->nc) std::memcpy(inv, G->Cinv.data(), sizeof(double) * G->Cinv.size());
  });
}

// The orc_csr passed here is decoupled in place and must outlive the CPR handle; its arrays
// are resized through malloc/realloc (filters may shrink nnz).
struct CprHolder : CPR {
  CSR own;
  orc_csr* ext = nullptr;
};

int orc_cpr_create(const deco_cpr_params* p, orc_csr* A, double* b, orc_prec** out) {
  return guard([&] {
    auto* H = new CprHolder();
    H->own = from_c(A);
    CPR* M;
    try {
      M = cpr_create(*p, H->own, b);
    } catch (...) {
      delete H;
      throw;
    }
    // move the built state into the holder
    H->p = M->p; H->A = &H->own; H->Ap = std::move(M->Ap); H->bs = M->bs; H->sn = M->sn; H->pid = M->pid;
    H->glo = std::move(M->glo); H->jac = std::move(M->jac); H->rt = M->rt; H->xp = M->xp; H->W = M->W;
    H->x2 = M->x2; H->tmp1 = M->tmp1; H->tmp2 = M->tmp2; H->n = M->n;
    delete M;
    write_back(H->own, A);
    H->ext = A;
    *out = new orc_prec{std::unique_ptr<Prec>(H)};
  });
}

static CPR* as_cpr(const orc_prec* M) {
  auto* c = dynamic_cast<CPR*>(M->impl.get());
  if (!c) throw std::invalid_argument("not a CPR preconditioner");
  return c;
}
static FSAI* as_fsai(const orc_prec* M) {
  auto* f = dynamic_cast<FSAI*>(M->impl.get());
  if (!f) {
    auto* c = dynamic_cast<CPR*>(M->impl.get());
    if (c) f = dynamic_cast<FSAI*>(c->glo.get());
  }
  if (!f) throw std::invalid_argument("not an FSAI preconditioner");
  return f;
}

int orc_cpr_pressure_matrix(const orc_prec* M, orc_csr* Ap) { return guard([&] { to_c(as_cpr(M)->Ap, Ap); }); }
int orc_cpr_set_local(orc_prec* M, orc_prec* local) {
  return guard([&] {
    as_cpr(M)->loc = std::move(local->impl);
    delete local;
  });
}
int orc_cpr_left_transform(const orc_prec* M, double* W) {
  return guard([&] { auto* c = as_cpr(M); std::copy(c->W.begin(), c->W.end(), W); });
}
int orc_fsai_factors(const orc_prec* M, orc_csr* FL, orc_csr* FUt) {
  return guard([&] { auto* f = as_fsai(M); to_c(f->FL, FL); to_c(f->FUt, FUt); });
}
int orc_prec_apply(orc_prec* M, const double* r, double* x) { return guard([&] { M->impl->apply(r, x); }); }
int orc_prec_apply_transpose(orc_prec* M, const double* r, double* x) {
  return guard([&] { M->impl->apply_t(r, x); });
}
void orc_prec_destroy(orc_prec* M) { delete M; }

int orc_solve(const deco_sol_params* sp, const orc_csr* A, orc_prec* M, const double* b, double* x,
              deco_sol_info* info, double* resvec) {
  return guard([&] {
    CSR B = from_c(A);
    ApplyA Aop = [&](const double* xx, double* yy) { spmv(B, xx, yy); };
    ApplyM Mop = [&](const double* rr, double* zz) {
      if (M) M->impl->apply(rr, zz); else std::memcpy(zz, rr, sizeof(double) * B.n);
    };
    deco_sol_info inf{};
    const double t0 = omp_get_wtime();
    solve(*sp, B.n, Aop, Mop, b, x, inf, resvec);
    inf.time_ms = (omp_get_wtime() - t0) * 1e3;
    if (info) *info = inf;
  });
}

int orc_solve_transpose(const deco_sol_params* sp, const orc_csr* A, orc_prec* M, const double* g, double* y,
                        deco_sol_info* info) {
  return guard([&] {
    CSR B = transpose(from_c(A));
    ApplyA Aop = [&](const double* xx, double* yy) { spmv(B, xx, yy); };
    ApplyM Mop = [&](const double* rr, double* zz) {
      if (M) M->impl->apply_t(rr, zz); else std::memcpy(zz, rr, sizeof(double) * B.n);
    };
    deco_sol_info inf{};
    const double t0 = omp_get_wtime();
    solve(*sp, B.n, Aop, Mop, g, y, inf, nullptr);
    inf.time_ms = (omp_get_wtime() - t0) * 1e3;
    if (info) *info = inf;
  });
}

int orc_lu_factor(int k, double* a, int* ipiv) { return lu_factor(k, a, ipiv); }
int orc_lu_solve(int k, const double* a, const int* ipiv, int trans, double* b) {
  lu_solve(k, a, ipiv, trans, b);
  return 0;
}

}  // extern "C"
