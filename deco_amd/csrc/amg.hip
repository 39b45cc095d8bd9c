// AMG preconditioner for the CPR pressure stage (the role of hypre_AMG,
// precond/hypre_AMG/hypre_AMG.hpp:16-37, hypre_AMG.tpp:3-80).
//
// Setup, level by level: the host computes strength, C/F splitting and the extended+i
// interpolation P (amg_setup.cpp); the device forms R = P^T, the Galerkin product
// A_c = R (A P) with the device SpGEMM, drops entries below amg_A_drop_tol times the row
// maximum, and computes the smoother diagonal. The coarsest matrix (<= amg_max_coarse rows) is
// inverted densely in FP64 on the host and uploaded.
//
// Apply: one V-cycle from a zero initial guess (hypre_AMG.hpp:28-33), one sweep down and up.
// Smoothers by amg_relax_type. The reference passes a type to hypre only when it is positive
// (hypre_AMG.tpp:50-51), so -1 and 0 both mean hypre's default: Gauss-Seidel forward on the way
// down and backward on the way up (hypre's l1 hybrid types 13/14). A sequential sweep has no
// wavefront form, so the Gauss-Seidel types run as two-stage Gauss-Seidel (the form hypre itself
// uses on GPUs, its types 11/12): x += g_s with rho = b - A x, g_0 = rho/d,
// g_k = (rho - T g_{k-1})/d, T the strict lower (forward) or upper (backward) part, d = a_ii.
//   -1, 0   forward down, backward up, s = 1      11, 13   forward both ways, s = 1
//   12      forward both ways, s = 2              14       backward both ways, s = 1
//   7       Jacobi with weight amg_relax_weight (1 when unset: weighted Jacobi is this key)
//   17      F-C-F Jacobi half-sweeps              18       l1-Jacobi (diagonal sum_j |a_ij|)
// Relaxation is fused into the stream SpMV epilogue. FP32 mode runs the levels on f32 shadows and
// the coarsest solve in FP64.
#include "prec.hpp"
#include "amg_setup.hpp"
#include <vector>
#include <cstdlib>
#include <cstring>

namespace deco {

deco_csr* spgemm(const deco_csr* A, const deco_csr* B);

// keep a_ij if i == j or |a_ij| >= tol * max_k |a_ik|
__global__ void k_drop_count(int n, const int* __restrict__ r, const int* __restrict__ c, const double* __restrict__ v,
                             double tol, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double mx = 0.0;
  for (int j = r[i]; j < r[i + 1]; ++j) mx = fmax(mx, fabs(v[j]));
  const double thr = tol * mx;
  int k = 0;
  for (int j = r[i]; j < r[i + 1]; ++j)
    if (c[j] == i || fabs(v[j]) >= thr) ++k;
  cnt[i] = k;
}
__global__ void k_drop_fill(int n, const int* __restrict__ r, const int* __restrict__ c, const double* __restrict__ v,
                            double tol, const int* __restrict__ rn, int* __restrict__ cn, double* __restrict__ vn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double mx = 0.0;
  for (int j = r[i]; j < r[i + 1]; ++j) mx = fmax(mx, fabs(v[j]));
  const double thr = tol * mx;
  int o = rn[i];
  for (int j = r[i]; j < r[i + 1]; ++j)
    if (c[j] == i || fabs(v[j]) >= thr) { cn[o] = c[j]; vn[o] = v[j]; ++o; }
}

static deco_csr* drop_small(const deco_csr* A, double tol) {
  deco_csr* B = csr_new(A->n, A->m);
  DBuf<int> cnt(A->n + 1);
  if (A->n) {
    hipLaunchKernelGGL(k_drop_count, dim3(grid_for(A->n, 256)), dim3(256), 0, stream(), A->n, A->r.p, A->c.p, A->v.p,
                       tol, cnt.p);
    DECO_CHECK_LAUNCH();
  }
  build_row_ptr(B, cnt.p, A->n);
  if (A->n) {
    hipLaunchKernelGGL(k_drop_fill, dim3(grid_for(A->n, 256)), dim3(256), 0, stream(), A->n, A->r.p, A->c.p, A->v.p,
                       tol, (const int*)B->r.p, B->c.p, B->v.p);
    DECO_CHECK_LAUNCH();
  }
  return B;
}

// strict lower (lower != 0) or strict upper part of A, rows kept sorted
__global__ void k_strict_count(int n, const int* __restrict__ r, const int* __restrict__ c, int lower,
                               int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k = 0;
  for (int j = r[i]; j < r[i + 1]; ++j) k += lower ? (c[j] < i) : (c[j] > i);
  cnt[i] = k;
}
__global__ void k_strict_fill(int n, const int* __restrict__ r, const int* __restrict__ c,
                              const double* __restrict__ v, int lower, const int* __restrict__ rn,
                              int* __restrict__ cn, double* __restrict__ vn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int o = rn[i];
  for (int j = r[i]; j < r[i + 1]; ++j)
    if (lower ? c[j] < i : c[j] > i) { cn[o] = c[j]; vn[o] = v[j]; ++o; }
}

static deco_csr* strict_part(const deco_csr* A, bool lower) {
  deco_csr* B = csr_new(A->n, A->m);
  DBuf<int> cnt(A->n + 1);
  if (A->n) {
    hipLaunchKernelGGL(k_strict_count, dim3(grid_for(A->n, 256)), dim3(256), 0, stream(), A->n, A->r.p, A->c.p,
                       lower ? 1 : 0, cnt.p);
    DECO_CHECK_LAUNCH();
  }
  build_row_ptr(B, cnt.p, A->n);
  if (A->n) {
    hipLaunchKernelGGL(k_strict_fill, dim3(grid_for(A->n, 256)), dim3(256), 0, stream(), A->n, A->r.p, A->c.p,
                       A->v.p, lower ? 1 : 0, (const int*)B->r.p, B->c.p, B->v.p);
    DECO_CHECK_LAUNCH();
  }
  return B;
}

// Residual of a two-stage Gauss-Seidel sweep: rho[o] = b[o] - (A x)[o] and g0[o] = rho[o]/d[o]
template <class T>
struct EpiResG {
  const T* __restrict__ b;
  const T* __restrict__ d;
  T* __restrict__ rho;
  T* __restrict__ g0;
  struct Pre { T b, d; };
  __device__ __forceinline__ Pre pre(int o) const { return Pre{b[o], d[o]}; }
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc, const Pre& p) const {
    const T r = (T)((AT)p.b - acc);
    rho[o] = r;
    g0[o] = r / p.d;
  }
  size_t row_bytes() const { return 4 * sizeof(T); }
};

// One inner step of a two-stage Gauss-Seidel sweep, the product being (T g)[o]:
// y = x[o] + (rho[o] - (T g)[o])/d[o] (x NULL: 0 + ...); stored in out, or with add_to set added to
// add_to[o*stride] (the finest level's last sweep of the fused CPR stage)
template <class T>
struct EpiTwoStage {
  const T* __restrict__ x;
  const T* __restrict__ rho;
  const T* __restrict__ d;
  T* __restrict__ out;
  T* __restrict__ add_to;
  int stride;
  struct Pre { T x, rho, d, a; };
  __device__ __forceinline__ Pre pre(int o) const {
    return Pre{x ? x[o] : (T)0, rho[o], d[o], add_to ? add_to[(size_t)o * stride] : (T)0};
  }
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc, const Pre& p) const {
    const T y = p.x + (T)((AT)p.rho - acc) / p.d;
    if (add_to) add_to[(size_t)o * stride] = p.a + y;
    else out[o] = y;
  }
  size_t row_bytes() const { return sizeof(T) * ((x ? 1 : 0) + 2 + (add_to ? 2 : 1)); }
};

// smoother diagonal: l1 (sum |a_ij|) or the diagonal entry divided by the Jacobi weight
__global__ void k_smoother_diag(int n, const int* __restrict__ r, const int* __restrict__ c,
                                const double* __restrict__ v, int l1, double weight, double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = r[i]; j < r[i + 1]; ++j) {
    if (l1) s += fabs(v[j]);
    else if (c[j] == i) s += v[j];
  }
  d[i] = l1 ? s : s / weight;
}

// x = b/d on the rows of the subset (others 0): first half-sweep from a zero guess
template <class T>
__global__ void k_relax_zero(int n, const T* __restrict__ b, const T* __restrict__ d, const int* __restrict__ cf,
                             int subset, T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool on = subset < 0 || (subset == 1 ? cf[i] == 1 : cf[i] != 1);
  x[i] = on ? (T)0 + (b[i] - (T)0) / d[i] : (T)0;
}

// y = x + (b - A x)/d on the subset rows, y = x elsewhere; with add_to set, also
// add_to[o*stride] += y (k_add_pressure's update, from the finest level's last sweep)
template <class T>
struct EpiRelax {
  const T* __restrict__ x;
  const T* __restrict__ b;
  const T* __restrict__ d;
  const int* __restrict__ cf;
  int subset;
  T* __restrict__ y;
  T* __restrict__ add_to = nullptr;
  int stride = 0;
  struct Pre { T x, b, d, a; bool on; };
  __device__ __forceinline__ Pre pre(int o) const {
    return Pre{x[o], b[o], d[o], add_to ? add_to[(size_t)o * stride] : (T)0,
               subset < 0 || (subset == 1 ? cf[o] == 1 : cf[o] != 1)};
  }
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc, const Pre& p) const {
    const T yo = p.on ? (T)(p.x + (p.b - (T)acc) / p.d) : p.x;
    y[o] = yo;
    if (add_to) add_to[(size_t)o * stride] = p.a + yo;
  }
  size_t row_bytes() const { return sizeof(T) * (4 + (add_to ? 2 : 0)) + (subset >= 0 ? sizeof(int) : 0); }
};

// Restriction b_c = R r fused with the next level's zero-guess Jacobi-type sweep: each coarse
// row stores b_c and x_c = (0 + (b_c - 0)/d_c) (k_relax_zero's arithmetic).
template <class T>
struct EpiRestrictJacobi0 {
  T* __restrict__ bn;
  const T* __restrict__ dn;
  T* __restrict__ xn;
  struct Pre { T d; };
  __device__ __forceinline__ Pre pre(int o) const { return Pre{dn[o]}; }
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc, const Pre& p) const {
    const T bo = (T)acc;
    bn[o] = bo;
    xn[o] = (T)0 + (bo - (T)0) / p.d;
  }
  size_t row_bytes() const { return 3 * sizeof(T); }
};

// Up-sweep of a level with a Jacobi-type smoother as one product: with t = x + P e and the
// residual r = b - A x kept from the down-sweep, t + (b - A t)/d = x + r/d + S e for
// S = P - D^{-1} (A P). Writes x, or with add_to set adds it to add_to[o*stride] instead.
template <class T>
struct EpiUp {
  T* __restrict__ x;
  const T* __restrict__ r;
  const T* __restrict__ d;
  T* __restrict__ add_to;
  int stride;
  struct Pre { T x, r, d, a; };
  __device__ __forceinline__ Pre pre(int o) const {
    return Pre{x[o], r[o], d[o], add_to ? add_to[(size_t)o * stride] : (T)0};
  }
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc, const Pre& p) const {
    const T y = (T)(p.x + p.r / p.d) + (T)acc;
    if (add_to) add_to[(size_t)o * stride] = p.a + y;   // the caller's iterate takes the result
    else x[o] = y;
  }
  size_t row_bytes() const { return sizeof(T) * (3 + (add_to ? 2 : 1)); }
};

// S = P - D^{-1} Q (Q = A P) on the union of the two row patterns (both sorted by column)
__global__ void k_upop_count(int n, const int* __restrict__ pr, const int* __restrict__ pc, const int* __restrict__ qr,
                             const int* __restrict__ qc, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int a = pr[i], b = qr[i], k = 0;
  const int ae = pr[i + 1], be = qr[i + 1];
  while (a < ae || b < be) {
    if (b >= be || (a < ae && pc[a] < qc[b])) ++a;
    else if (a >= ae || qc[b] < pc[a]) ++b;
    else { ++a; ++b; }
    ++k;
  }
  cnt[i] = k;
}
__global__ void k_upop_fill(int n, const int* __restrict__ pr, const int* __restrict__ pc, const double* __restrict__ pv,
                            const int* __restrict__ qr, const int* __restrict__ qc, const double* __restrict__ qv,
                            const double* __restrict__ d, const int* __restrict__ sr, int* __restrict__ sc,
                            double* __restrict__ sv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int a = pr[i], b = qr[i], o = sr[i];
  const int ae = pr[i + 1], be = qr[i + 1];
  const double di = d[i];
  while (a < ae || b < be) {
    if (b >= be || (a < ae && pc[a] < qc[b])) { sc[o] = pc[a]; sv[o] = pv[a]; ++a; }
    else if (a >= ae || qc[b] < pc[a]) { sc[o] = qc[b]; sv[o] = -(qv[b] / di); ++b; }
    else { sc[o] = pc[a]; sv[o] = pv[a] - qv[b] / di; ++a; ++b; }
    ++o;
  }
}

static deco_csr* up_operator(const deco_csr* P, const deco_csr* Q, const double* d) {
  deco_csr* S = csr_new(P->n, P->m);
  DBuf<int> cnt(P->n + 1);
  if (P->n) {
    hipLaunchKernelGGL(k_upop_count, dim3(grid_for(P->n, 256)), dim3(256), 0, stream(), P->n, P->r.p, P->c.p, Q->r.p,
                       Q->c.p, cnt.p);
    DECO_CHECK_LAUNCH();
  }
  build_row_ptr(S, cnt.p, P->n);
  if (P->n) {
    hipLaunchKernelGGL(k_upop_fill, dim3(grid_for(P->n, 256)), dim3(256), 0, stream(), P->n, P->r.p, P->c.p, P->v.p,
                       Q->r.p, Q->c.p, Q->v.p, d, (const int*)S->r.p, S->c.p, S->v.p);
    DECO_CHECK_LAUNCH();
  }
  return S;
}

// K = [R, -R A] with interleaved columns (column 2j: R on b_j, column 2j+1: -(R A) on t_j), so
// that b_c = R (b - A t) = K [b; t] is one product: the residual and the restriction of a
// two-stage level's down-sweep in one launch. Rows merge the two sorted patterns.
__global__ void k_rcat_count(int n, const int* __restrict__ rr, const int* __restrict__ qr, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cnt[i] = (rr[i + 1] - rr[i]) + (qr[i + 1] - qr[i]);
}
__global__ void k_rcat_fill(int n, const int* __restrict__ rr, const int* __restrict__ rc, const double* __restrict__ rv,
                            const int* __restrict__ qr, const int* __restrict__ qc, const double* __restrict__ qv,
                            const int* __restrict__ kr, int* __restrict__ kc, double* __restrict__ kv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int a = rr[i], b = qr[i], o = kr[i];
  const int ae = rr[i + 1], be = qr[i + 1];
  while (a < ae || b < be) {   // column j of R (2j) precedes column j of R A (2j + 1)
    if (b >= be || (a < ae && rc[a] <= qc[b])) { kc[o] = 2 * rc[a]; kv[o] = rv[a]; ++a; }
    else { kc[o] = 2 * qc[b] + 1; kv[o] = -qv[b]; ++b; }
    ++o;
  }
}

static deco_csr* restrict_operator(const deco_csr* R, const deco_csr* RA) {
  deco_csr* K = csr_new(R->n, 2 * R->m);
  DBuf<int> cnt(R->n + 1);
  if (R->n) {
    hipLaunchKernelGGL(k_rcat_count, dim3(grid_for(R->n, 256)), dim3(256), 0, stream(), R->n, R->r.p, RA->r.p, cnt.p);
    DECO_CHECK_LAUNCH();
  }
  build_row_ptr(K, cnt.p, R->n);
  if (R->n) {
    hipLaunchKernelGGL(k_rcat_fill, dim3(grid_for(R->n, 256)), dim3(256), 0, stream(), R->n, R->r.p, R->c.p, R->v.p,
                       RA->r.p, RA->c.p, RA->v.p, (const int*)K->r.p, K->c.p, K->v.p);
    DECO_CHECK_LAUNCH();
  }
  return K;
}

// x access of K's product: entry 2j is b_j, entry 2j+1 is t_j
template <class T>
struct CatGather {
  using type = T;
  const T* __restrict__ b;
  const T* __restrict__ t;
  __device__ __forceinline__ T operator()(int j) const {
    const T* p = (j & 1) ? t : b;   // one load per entry with a per-lane address
    return p[j >> 1];
  }
};

// x = C b with C row-major n x n in MT (the coarsest inverse for the forward cycle, its transpose
// for the transposed one; or a dense sub-cycle operator): one wavefront per row, lanes stride the
// columns (scalar loads: 16-byte vector loads with the b entries they pair with measured slower),
// fixed-order butterfly sum; input/output in T, FP64 arithmetic; with add_to set,
// add_to[i*stride] += x_i instead of storing x
template <class T, class MT>
__global__ __launch_bounds__(256) void k_dense(int n, const MT* __restrict__ C, const T* __restrict__ b,
                                               T* __restrict__ x, T* __restrict__ add_to, int stride) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const MT* cr = C + (size_t)row * n;
  double s = 0.0;
#pragma unroll 4
  for (int k = lane; k < n; k += 64) s += (double)cr[k] * (double)b[k];
  s = wave_sum(s);
  if (lane == 0) {
    const T y = (T)s;
    if (add_to) add_to[(size_t)row * stride] += y;
    else x[row] = y;
  }
}

// x = Z b for the dense sub-cycle operator Z, n rows of ld >= n entries (ld a multiple of 4, the
// columns past n zero, so every row starts on a 16-byte boundary): a workgroup takes two rows and
// each of its four wavefronts a quarter of their columns, so one round of loads covers a row pair
// (the whole operator, 76 MB in FP32 at 4365 rows, is in flight at once instead of in three rounds
// of two-row wavefronts); each lane loads 16-byte vectors of both rows and the b entries they pair
// with, kDenseU vectors per row in flight; FP64 arithmetic, fixed-order butterfly sums, the four
// quarters added in order; with add_to set, add_to[i*stride] += x_i instead of storing x
constexpr int kDenseU = 6;
template <class T, class MT>
__global__ __launch_bounds__(256) void k_dense_tail(int n, int ld, const MT* __restrict__ Z, const T* __restrict__ b,
                                                    T* __restrict__ x, T* __restrict__ add_to, int stride) {
  constexpr int VM = 16 / sizeof(MT);   // entries per 16-byte vector
  typedef MT mv __attribute__((ext_vector_type(VM)));
  __shared__ double part[4][2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row0 = blockIdx.x * 2;   // the grid covers the rows, so row0 < n
  const bool two = row0 + 1 < n;
  const mv* z0 = reinterpret_cast<const mv*>(Z + (size_t)row0 * ld);
  const mv* z1 = two ? z0 + ld / VM : z0;
  const bool bvec = ((uintptr_t)b & 15) == 0 && sizeof(T) == sizeof(MT);
  const int nv = ld / VM;
  const int q = (nv + 3) / 4;
  const int vb = w * q, ve = min(nv, vb + q);   // this wavefront's quarter of the vectors
  double s0 = 0.0, s1 = 0.0;
  for (int v0 = vb; v0 < ve; v0 += 64 * kDenseU) {
    mv a0[kDenseU], a1[kDenseU];
    T bb[kDenseU][VM];
#pragma unroll
    for (int u = 0; u < kDenseU; ++u) {
      const int v = v0 + lane + 64 * u;
      if (v < ve) {
        a0[u] = __builtin_nontemporal_load(z0 + v);
        a1[u] = __builtin_nontemporal_load(z1 + v);
      } else {
        a0[u] = mv(0);
        a1[u] = mv(0);
      }
    }
#pragma unroll
    for (int u = 0; u < kDenseU; ++u) {
      const int v = v0 + lane + 64 * u;
      const int e = v * VM;
      if (v < ve && bvec && e + VM <= n) {
        const mv t = *reinterpret_cast<const mv*>(b + e);
#pragma unroll
        for (int qq = 0; qq < VM; ++qq) bb[u][qq] = (T)t[qq];
      } else {
#pragma unroll
        for (int qq = 0; qq < VM; ++qq) bb[u][qq] = (v < ve && e + qq < n) ? b[e + qq] : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < kDenseU; ++u)
#pragma unroll
      for (int qq = 0; qq < VM; ++qq) {
        s0 += (double)a0[u][qq] * (double)bb[u][qq];
        s1 += (double)a1[u][qq] * (double)bb[u][qq];
      }
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if (lane == 0) {
    part[w][0] = s0;
    part[w][1] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t0 = ((part[0][0] + part[1][0]) + part[2][0]) + part[3][0];
    const double t1 = ((part[0][1] + part[1][1]) + part[2][1]) + part[3][1];
    if (add_to) {
      add_to[(size_t)row0 * stride] += (T)t0;
      if (two) add_to[(size_t)(row0 + 1) * stride] += (T)t1;
    } else {
      x[row0] = (T)t0;
      if (two) x[row0 + 1] = (T)t1;
    }
  }
}

// dst (n rows of ld entries) = src (n x n, row-major) with zero columns n .. ld-1
template <class S, class D>
__global__ __launch_bounds__(256) void k_pad_rows(int n, int ld, const S* __restrict__ src, D* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)n * ld) return;
  const int r = (int)(i / ld), c = (int)(i % ld);
  dst[i] = c < n ? (D)src[(size_t)r * n + c] : D(0);
}

// C (M x N) = A (M x K) B (K x N), all row-major FP64 (the dense tail's setup products): 64 x 64
// output tiles, 16-deep slices of A and B staged in LDS, 4 x 4 outputs per thread
constexpr int kGemmTile = 64, kGemmDepth = 16;
__global__ __launch_bounds__(256) void k_gemm64(int M, int N, int K, const double* __restrict__ A,
                                                const double* __restrict__ B, double* __restrict__ C) {
  __shared__ double As[kGemmDepth][kGemmTile + 1];
  __shared__ double Bs[kGemmDepth][kGemmTile + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int row0 = blockIdx.y * kGemmTile, col0 = blockIdx.x * kGemmTile;
  double acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += kGemmDepth) {
    for (int e = threadIdx.x; e < kGemmTile * kGemmDepth; e += 256) {
      const int ar = e / kGemmDepth, ak = e % kGemmDepth;          // A tile: 64 rows x 16
      const int gr = row0 + ar, gk = k0 + ak;
      As[ak][ar] = (gr < M && gk < K) ? A[(size_t)gr * K + gk] : 0.0;
      const int bk = e / kGemmTile, bc = e % kGemmTile;           // B tile: 16 x 64 columns
      const int hk = k0 + bk, hc = col0 + bc;
      Bs[bk][bc] = (hk < K && hc < N) ? B[(size_t)hk * N + hc] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kGemmDepth; ++k) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[k][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = row0 + ty * 4 + i;
    if (r >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = col0 + tx * 4 + j;
      if (c < N) C[(size_t)r * N + c] = acc[i][j];
    }
  }
}

// Z[i, c] += v for the entries of a CSR matrix (n rows, Z n x n row-major)
__global__ void k_add_csr_dense(int n, const int* __restrict__ r, const int* __restrict__ c,
                                const double* __restrict__ v, double* __restrict__ Z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int e = r[i]; e < r[i + 1]; ++e) Z[(size_t)i * n + c[e]] += v[e];
}

// Zt = Z^T for an n x n row-major matrix, through 32 x 32 LDS tiles
__global__ __launch_bounds__(256) void k_transpose_dense(int n, const double* __restrict__ Z, double* __restrict__ Zt) {
  __shared__ double t[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
  for (int k = ty; k < 32; k += 8) {
    const int r = by + k, c = bx + tx;
    t[k][tx] = (r < n && c < n) ? Z[(size_t)r * n + c] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int r = bx + k, c = by + tx;
    if (r < n && c < n) Zt[(size_t)r * n + c] = t[tx][k];
  }
}

static void gemm64(int M, int N, int K, const double* A, const double* B, double* C) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((N + kGemmTile - 1) / kGemmTile, (M + kGemmTile - 1) / kGemmTile);
  hipLaunchKernelGGL(k_gemm64, grid, dim3(256), 0, stream(), M, N, K, A, B, C);
  DECO_CHECK_LAUNCH();
}

// largest level (rows) whose whole sub-cycle, down to the coarsest solve, is formed as one dense
// operator at setup
static int g_dense_tail_rows = 4608;

struct AMG : deco_prec {
  struct Level {
    deco_csr* A = nullptr;
    deco_csr* At = nullptr;   // transposed level matrix (transposed apply)
    deco_csr* P = nullptr;
    deco_csr* R = nullptr;
    deco_csr* AP = nullptr;   // A P from the Galerkin product, until S is formed
    deco_csr* S = nullptr;    // up-sweep operator P - D^{-1} A P (Jacobi-type smoothers)
    deco_csr* St = nullptr;   // the same with A^T, for the transposed cycle
    deco_csr* Lo = nullptr;   // strict lower / upper parts of A (two-stage Gauss-Seidel)
    deco_csr* Up = nullptr;
    deco_csr* Lot = nullptr;  // the same of A^T
    deco_csr* Upt = nullptr;
    deco_csr* K = nullptr;    // [R, -R A] interleaved: residual and restriction of the down-sweep
    deco_csr* Kt = nullptr;   // [R, -R A^T] for the transposed cycle
    DBuf<double> d;
    DBuf<float> df;
    DBuf<int> cf;
    DBuf<double> x, b, t, r, u;
    DBuf<float> xf, bf, tf, rf, uf;
  };
  std::vector<Level> L;       // levels 0..nl-1 with interpolation; L.back() may be the coarsest
  deco_csr* Ac = nullptr;     // coarsest matrix
  int nc = 0;
  DBuf<double> CT, Cinv;      // coarse inverse: row-major (forward) and its transpose (transposed cycle)
  DBuf<double> xc, bc;
  DBuf<float> xcf, bcf;
  DBuf<float> r32, x32;       // FP32 copies of a standalone apply's input and output
  // the last level's sub-cycle (zero-guess sweep, residual, restriction, coarse solve, up-sweep)
  // as one dense operator Z, row-major, and its transpose for the transposed cycle (Jacobi-type
  // smoothers, last level of at most g_dense_tail_rows rows); tail = its rows, 0 when absent
  int tail = 0;
  int tail_lev = -1;          // the level whose sub-cycle the dense operator replaces (-1: none)
  int tail_ld = 0;            // row length of the stored operators: tail rounded up to a multiple of 4
  DBuf<double> Z64, Z64t;     // tail rows of tail_ld entries (zero past column tail), row-major
  DBuf<float> Z32, Z32t;
  int relax = 18;              // 18 l1-Jacobi, 0 Jacobi (weighted), 17 F-C-F, 11 two-stage Gauss-Seidel
  double weight = 1.0;
  int dir_down = 1, dir_up = -1, inner = 1;   // two-stage: +1 forward, -1 backward; inner steps
  bool have_t = false;
  bool jacobi_type() const { return relax == 18 || relax == 0; }

  AMG(const deco_cpr_params& p, const deco_csr* A0) {
    kind = deco_prec::AMG;
    n = A0->n;
    switch (p.amg_relax_type) {
      case 18: relax = 18; break;
      case 17: relax = 17; break;
      case 7: relax = 0; weight = p.amg_relax_weight > 0.0 ? p.amg_relax_weight : 1.0; break;
      case -1: case 0: relax = 11; dir_down = 1; dir_up = -1; inner = 1; break;
      case 11: case 13: relax = 11; dir_down = 1; dir_up = 1; inner = 1; break;
      case 12: relax = 11; dir_down = 1; dir_up = 1; inner = 2; break;
      case 14: relax = 11; dir_down = -1; dir_up = -1; inner = 1; break;
      default:
        fail(DECO_ERR_UNSUPPORTED, "amg_relax_type %d is not supported (use -1, 0, 7, 11-14, 17 or 18)",
             p.amg_relax_type);
    }
    if (p.amg_agg_levs > 0) fail(DECO_ERR_UNSUPPORTED, "aggressive coarsening (amg_agg_levs > 0) is not supported");
    if (A0->n != A0->m) fail(DECO_ERR_ARG, "AMG matrix is not square");
    const int max_coarse = p.amg_max_coarse > 0 ? p.amg_max_coarse : 300;
    const int max_levels = p.amg_max_levels > 0 ? p.amg_max_levels : 25;
    const double theta = p.amg_SoC_alpha;
    const double mrs = p.amg_max_row_sum > 0 ? p.amg_max_row_sum : 0.9;
    deco_csr* Al = csr_clone(A0);
    amg::HostCSR H;   // host copy of the current level, its arrays reused level to level
    try {
      for (int lev = 0;; ++lev) {
        if (Al->n <= max_coarse || lev >= max_levels - 1) break;
        H.n = Al->n;
        H.m = Al->m;
        H.r.clear();
        H.c.clear();
        H.v.clear();
        H.r.resize(Al->n + 1);
        H.c.resize(Al->nnz);
        H.v.resize(Al->nnz);
        {
          ProfPhase ph_w("queued device work");   // decoupling and FSAI kernels still running
          sync();
        }
        ProfPhase ph_dl("download");
        amg::prefault(H.r.data(), sizeof(int) * H.r.size());
        amg::prefault(H.c.data(), sizeof(int) * H.c.size());
        amg::prefault(H.v.data(), sizeof(double) * H.v.size());
        Al->r.download(H.r.data(), Al->n + 1);
        Al->c.download(H.c.data(), Al->nnz);
        Al->v.download(H.v.data(), Al->nnz);
        ph_dl.end();
        std::vector<char> strong;
        ProfPhase ph_st("strength");
        amg::strength(H, theta, mrs, strong);
        ph_st.end();
        std::vector<int> cf;
        ProfPhase ph_co("coarsening");
        amg::coarsen(H, strong, p.amg_coarse_type, 0x5eed0000ull + (uint64_t)lev, cf);
        ph_co.end();
        int ncl = 0;
        ProfPhase ph_in("interpolation");
        amg::HostCSR Ph = amg::interp_extpi(H, strong, cf, p.amg_prol_trunc_fac, p.amg_P_max_elmts, ncl);
        ph_in.end();
        if (ncl == 0 || ncl == Al->n) break;
        Level lv;
        lv.A = Al;
        ProfPhase ph_up("upload P");
        lv.cf.upload(cf.data(), cf.size());
        lv.P = csr_new(Ph.n, Ph.m);
        lv.P->nnz = (int)Ph.c.size();
        lv.P->r.upload(Ph.r.data(), Ph.r.size());
        lv.P->c.upload(Ph.c.data(), Ph.c.size());
        lv.P->v.upload(Ph.v.data(), Ph.v.size());
        ph_up.end();
        ProfPhase ph_t("Transpose");
        lv.R = csr_transpose(lv.P);
        ph_t.end();
        ProfPhase ph_rap("Galerkin RAP");
        deco_csr* AP = spgemm(Al, lv.P);
        lv.AP = AP;
        deco_csr* RAP = spgemm(lv.R, AP);
        if (p.amg_A_drop_tol > 0.0) {
          Al = drop_small(RAP, p.amg_A_drop_tol);
          delete RAP;
        } else {
          Al = RAP;
        }
        ph_rap.end();
        L.push_back(std::move(lv));
      }
      Ac = Al;
      Al = nullptr;
    } catch (...) {
      delete Al;
      release();
      throw;
    }
    // smoothers and work vectors
    const int l1 = (relax == 18) ? 1 : 0;
    for (auto& lv : L) {
      const int nl = lv.A->n;
      lv.d.alloc(nl);
      hipLaunchKernelGGL(k_smoother_diag, dim3(grid_for(nl, 256)), dim3(256), 0, stream(), nl, lv.A->r.p, lv.A->c.p,
                         lv.A->v.p, l1, relax == 0 ? weight : 1.0, lv.d.p);
      DECO_CHECK_LAUNCH();
      lv.df.alloc(nl);
      convert(lv.d.p, lv.df.p, nl);
      lv.x.alloc(nl); lv.b.alloc(nl); lv.t.alloc(nl); lv.r.alloc(nl);
      lv.xf.alloc(nl); lv.bf.alloc(nl); lv.tf.alloc(nl); lv.rf.alloc(nl);
      if (relax == 11) {
        lv.Lo = strict_part(lv.A, true);
        lv.Up = strict_part(lv.A, false);
        lv.u.alloc(nl);
        lv.uf.alloc(nl);
        if (&lv != &L[0]) lv.K = fused_restriction(lv, lv.A);
      }
      // the up-sweep's prolongation and Jacobi-type sweep as one operator
      if (jacobi_type()) lv.S = up_operator(lv.P, lv.AP, lv.d.p);
      delete lv.AP;
      lv.AP = nullptr;
    }
    // coarsest: dense FP64 inverse
    ProfPhase ph_ci("coarse inverse");
    nc = Ac->n;
    H.n = nc;
    H.m = nc;
    H.r.clear();
    H.c.clear();
    H.v.clear();
    H.r.resize(nc + 1);
    H.c.resize(Ac->nnz);
    H.v.resize(Ac->nnz);
    Ac->r.download(H.r.data(), nc + 1);
    Ac->c.download(H.c.data(), Ac->nnz);
    Ac->v.download(H.v.data(), Ac->nnz);
    std::vector<double> inv;
    if (nc > 0 && !amg::dense_inverse(H, inv)) {
      release();
      fail(DECO_ERR_SINGULAR, "coarsest AMG matrix (%d rows) is singular", nc);
    }
    std::vector<double> invT((size_t)nc * nc);
    for (int i = 0; i < nc; ++i)
      for (int k = 0; k < nc; ++k) invT[(size_t)k * nc + i] = inv[(size_t)i * nc + k];
    Cinv.upload(inv.data(), inv.size());
    CT.upload(invT.data(), invT.size());
    xc.alloc(nc); bc.alloc(nc); xcf.alloc(nc); bcf.alloc(nc);
    ph_ci.end();
    if (relax != 17 && !L.empty() && L.back().A->n <= g_dense_tail_rows) form_tail();
    sync();
  }

  static void download(const deco_csr* A, amg::HostCSR& H) {
    H.n = A->n;
    H.m = A->m;
    H.r.resize(A->n + 1);
    H.c.resize(A->nnz);
    H.v.resize(A->nnz);
    A->r.download(H.r.data(), A->n + 1);
    A->c.download(H.c.data(), A->nnz);
    A->v.download(H.v.data(), A->nnz);
  }

  // the smoother of the last level as a stationary step's matrix: D^{-1} (Jacobi-type), or the
  // two-stage Neumann series sum_k N^k D^{-1}, N = -D^{-1} T, T the strict lower (dir > 0) or
  // upper part, k = 0..inner
  amg::HostCSR smoother_matrix(const amg::HostCSR& A, const std::vector<double>& d, int dir) const {
    const int nl = A.n;
    amg::HostCSR X;
    X.n = X.m = nl;
    X.r.resize(nl + 1);
    for (int i = 0; i <= nl; ++i) X.r[i] = i;
    X.c.resize(nl);
    X.v.resize(nl);
    for (int i = 0; i < nl; ++i) { X.c[i] = i; X.v[i] = 1.0 / d[i]; }
    if (relax != 11) return X;
    amg::HostCSR N;
    N.n = N.m = nl;
    N.r.assign(nl + 1, 0);
    for (int i = 0; i < nl; ++i) {
      for (int e = A.r[i]; e < A.r[i + 1]; ++e)
        if (dir > 0 ? A.c[e] < i : A.c[e] > i) { N.c.push_back(A.c[e]); N.v.push_back(-A.v[e] / d[i]); }
      N.r[i + 1] = (int)N.c.size();
    }
    amg::HostCSR term = X;
    for (int k = 0; k < inner; ++k) {
      term = amg::host_spgemm(N, term);
      amg::HostCSR sum;                                // X + term, patterns merged
      sum.n = sum.m = nl;
      sum.r.assign(nl + 1, 0);
      for (int i = 0; i < nl; ++i) {
        int a = X.r[i], b = term.r[i];
        const int ae = X.r[i + 1], be = term.r[i + 1];
        while (a < ae || b < be) {
          if (b >= be || (a < ae && X.c[a] < term.c[b])) { sum.c.push_back(X.c[a]); sum.v.push_back(X.v[a]); ++a; }
          else if (a >= ae || term.c[b] < X.c[a]) { sum.c.push_back(term.c[b]); sum.v.push_back(term.v[b]); ++b; }
          else { sum.c.push_back(X.c[a]); sum.v.push_back(X.v[a] + term.v[b]); ++a; ++b; }
        }
        sum.r[i + 1] = (int)sum.c.size();
      }
      X = std::move(sum);
    }
    return X;
  }

  // The sub-cycle of level k (pre-smoother from zero, residual, restriction, the cycle below,
  // prolongation, post-smoother) is the fixed map Z_k = E + W C G with E = Xpre + Xpost - Xpost A Xpre,
  // W = P - Xpost A P, G = R - R A Xpre and C the operator of the cycle below (the coarsest inverse,
  // or Z_{k+1}). Starting from the last level with interpolation, Z is formed for every level of at
  // most g_dense_tail_rows rows: E, W and G on the host (sparse products), C G and W (C G) on the
  // device in FP64. The highest such level's Z is kept in FP64 and FP32 with its transpose (R = P^T
  // and the transposed cycle's smoothers are the transposes of the other side's, so the transposed
  // sub-cycle is Z^T).
  void form_tail() {
    ProfPhase ph("dense tail");
    int k = (int)L.size() - 1;
    while (k > 0 && L[k - 1].A->n <= g_dense_tail_rows) --k;
    DBuf<double> Cprev, Z, H, W, G;
    int nC = nc;
    const double* C = Cinv.p;
    for (int lev = (int)L.size() - 1; lev >= k; --lev) {
      Level& lv = L[lev];
      const int nl = lv.A->n;
      amg::HostCSR A, P, R;
      download(lv.A, A);
      download(lv.P, P);
      download(lv.R, R);
      std::vector<double> d(nl);
      lv.d.download(d.data(), nl);
      const amg::HostCSR Xpre = smoother_matrix(A, d, dir_down);
      const amg::HostCSR Xpost = smoother_matrix(A, d, dir_up);
      const amg::HostCSR AX = amg::host_spgemm(A, Xpre);
      const amg::HostCSR RAX = amg::host_spgemm(R, AX);
      const amg::HostCSR XAP = amg::host_spgemm(Xpost, amg::host_spgemm(A, P));
      const amg::HostCSR XAX = amg::host_spgemm(Xpost, AX);
      std::vector<double> Wh((size_t)nl * nC, 0.0), Gh((size_t)nC * nl, 0.0);
      for (int i = 0; i < nl; ++i) {
        for (int e = P.r[i]; e < P.r[i + 1]; ++e) Wh[(size_t)i * nC + P.c[e]] += P.v[e];
        for (int e = XAP.r[i]; e < XAP.r[i + 1]; ++e) Wh[(size_t)i * nC + XAP.c[e]] -= XAP.v[e];
      }
      for (int i = 0; i < nC; ++i) {
        for (int e = R.r[i]; e < R.r[i + 1]; ++e) Gh[(size_t)i * nl + R.c[e]] += R.v[e];
        for (int e = RAX.r[i]; e < RAX.r[i + 1]; ++e) Gh[(size_t)i * nl + RAX.c[e]] -= RAX.v[e];
      }
      // E = Xpre + Xpost - Xpost A Xpre as one CSR (duplicate columns within a row are summed by
      // the scatter-add)
      std::vector<int> er(nl + 1, 0), ec;
      std::vector<double> ev;
      for (int i = 0; i < nl; ++i) {
        for (int e = Xpre.r[i]; e < Xpre.r[i + 1]; ++e) { ec.push_back(Xpre.c[e]); ev.push_back(Xpre.v[e]); }
        for (int e = Xpost.r[i]; e < Xpost.r[i + 1]; ++e) { ec.push_back(Xpost.c[e]); ev.push_back(Xpost.v[e]); }
        for (int e = XAX.r[i]; e < XAX.r[i + 1]; ++e) { ec.push_back(XAX.c[e]); ev.push_back(-XAX.v[e]); }
        er[i + 1] = (int)ec.size();
      }
      W.upload(Wh.data(), Wh.size());
      G.upload(Gh.data(), Gh.size());
      DBuf<int> erd, ecd;
      DBuf<double> evd;
      erd.upload(er.data(), er.size());
      ecd.upload(ec.data(), ec.size());
      evd.upload(ev.data(), ev.size());
      H.alloc((size_t)nC * nl);
      gemm64(nC, nl, nC, C, G.p, H.p);                 // C G
      DBuf<double> Zl((size_t)nl * nl);
      gemm64(nl, nl, nC, W.p, H.p, Zl.p);              // W (C G)
      hipLaunchKernelGGL(k_add_csr_dense, dim3(grid_for(nl, 256)), dim3(256), 0, stream(), nl, (const int*)erd.p,
                         (const int*)ecd.p, (const double*)evd.p, Zl.p);
      DECO_CHECK_LAUNCH();
      sync();   // the host arrays go out of scope here
      Cprev = std::move(Zl);
      C = Cprev.p;
      nC = nl;
    }
    const int nl = nC;
    const int ld = (nl + 3) / 4 * 4;
    DBuf<double> Zt((size_t)nl * nl);
    hipLaunchKernelGGL(k_transpose_dense, dim3((nl + 31) / 32, (nl + 31) / 32), dim3(256), 0, stream(), nl,
                       (const double*)Cprev.p, Zt.p);
    DECO_CHECK_LAUNCH();
    const unsigned g = grid_for((int64_t)nl * ld, 256);
    Z64.alloc((size_t)nl * ld);
    Z64t.alloc((size_t)nl * ld);
    Z32.alloc((size_t)nl * ld);
    Z32t.alloc((size_t)nl * ld);
    hipLaunchKernelGGL((k_pad_rows<double, double>), dim3(g), dim3(256), 0, stream(), nl, ld, (const double*)Cprev.p,
                       Z64.p);
    hipLaunchKernelGGL((k_pad_rows<double, double>), dim3(g), dim3(256), 0, stream(), nl, ld, (const double*)Zt.p,
                       Z64t.p);
    hipLaunchKernelGGL((k_pad_rows<double, float>), dim3(g), dim3(256), 0, stream(), nl, ld, (const double*)Cprev.p,
                       Z32.p);
    hipLaunchKernelGGL((k_pad_rows<double, float>), dim3(g), dim3(256), 0, stream(), nl, ld, (const double*)Zt.p,
                       Z32t.p);
    DECO_CHECK_LAUNCH();
    sync();
    tail = nl;
    tail_ld = ld;
    tail_lev = k;
  }

  // x = C b (or add_to[stride*i] += (C b)_i) with the m x m row-major C
  template <class T, class MT>
  void dense(int m, const MT* C, const T* b, T* x, T* add_to, int stride) {
    hipLaunchKernelGGL((k_dense<T, MT>), dim3((m + 3) / 4), dim3(256), 0, stream(), m, C, b, x, add_to, stride);
    DECO_CHECK_LAUNCH();
    count_bytes((double)m * m * sizeof(MT) + (double)m * sizeof(T) * (add_to ? 3 : 2));
  }
  // x = Z b (or add_to[stride*i] += (Z b)_i) with the dense tail operator (tail rows of tail_ld)
  template <class T, class MT>
  void dense_tail(const MT* Zop, const T* b, T* x, T* add_to, int stride) {
    const int m = tail;
    hipLaunchKernelGGL((k_dense_tail<T, MT>), dim3((m + 1) / 2), dim3(256), 0, stream(), m, tail_ld, Zop, b, x, add_to,
                       stride);
    DECO_CHECK_LAUNCH();
    count_bytes((double)m * m * sizeof(MT) + (double)m * sizeof(T) * (add_to ? 3 : 2));
  }
  // the dense tail operator in the cycle's precision
  const double* tail_op(bool tr, const double*) const { return tr ? Z64t.p : Z64.p; }
  const float* tail_op(bool tr, const float*) const { return tr ? Z32t.p : Z32.p; }

  void release() {
    for (auto& lv : L) {
      delete lv.A; delete lv.At; delete lv.P; delete lv.R; delete lv.AP; delete lv.S; delete lv.St;
      delete lv.Lo; delete lv.Up; delete lv.Lot; delete lv.Upt; delete lv.K; delete lv.Kt;
      lv.A = lv.At = lv.P = lv.R = lv.AP = lv.S = lv.St = nullptr;
      lv.Lo = lv.Up = lv.Lot = lv.Upt = lv.K = lv.Kt = nullptr;
    }
    L.clear();
    delete Ac;
    Ac = nullptr;
  }
  ~AMG() override { release(); }

  // level data accessors by precision
  template <class T> static T* X(Level& l);
  template <class T> static T* B(Level& l);
  template <class T> static T* Tm(Level& l);
  template <class T> static T* Rr(Level& l);
  template <class T> static T* Uu(Level& l);
  template <class T> static const T* D(Level& l);
  template <class T> static const T* vals(deco_csr* A);
  template <class T> T* XC();
  template <class T> T* BC();

  template <class T>
  void relax_sweep(Level& lv, const deco_csr* A, const T* x, const T* b, T* y, int subset, T* add_to = nullptr,
                   int stride = 0) {
    spmv_auto<T>(A, x, EpiRelax<T>{x, b, D<T>(lv), lv.cf.p, subset, y, add_to, stride});
  }

  // one V-cycle on level l with right-hand side b, result in x (x != b); x_ready: x already
  // holds the zero-guess sweep; add_to: the last sweep adds the result to add_to[stride*i] (with
  // a Jacobi-type smoother instead of storing it in x, which then holds the zero-guess sweep)
  template <class T>
  void cycle(size_t l, const T* b, T* x, bool tr, bool x_ready = false, T* add_to = nullptr, int stride = 0) {
    if (l == L.size()) {
      if (nc == 0) return;
      dense(nc, (const double*)(tr ? CT.p : Cinv.p), b, x, (T*)nullptr, 0);
      return;
    }
    if (tail && (int)l == tail_lev) {   // the whole sub-cycle below as one product
      dense_tail(tail_op(tr, b), b, x, add_to, stride);
      return;
    }
    if (relax == 11) {
      two_stage_level<T>(l, b, x, tr, x_ready, add_to, stride);
      return;
    }
    Level& lv = L[l];
    const deco_csr* A = tr ? lv.At : lv.A;
    const int nl = A->n;
    T* t = Tm<T>(lv);
    T* r = Rr<T>(lv);
    // down: smoother from a zero guess
    if (relax == 17) {
      hipLaunchKernelGGL(k_relax_zero<T>, dim3(grid_for(nl, 256)), dim3(256), 0, stream(), nl, b, D<T>(lv), lv.cf.p, 0, x);
      DECO_CHECK_LAUNCH();
      count_bytes((double)nl * (3 * sizeof(T) + sizeof(int)));
      relax_sweep<T>(lv, A, x, b, t, 1);
      relax_sweep<T>(lv, A, t, b, x, 0);
    } else if (!x_ready) {
      hipLaunchKernelGGL(k_relax_zero<T>, dim3(grid_for(nl, 256)), dim3(256), 0, stream(), nl, b, D<T>(lv), lv.cf.p, -1, x);
      DECO_CHECK_LAUNCH();
      count_bytes((double)nl * (3 * sizeof(T)));
    }
    // residual and restriction (the restriction also does the next level's zero-guess sweep
    // for the Jacobi-type smoothers)
    spmv_auto<T>(A, x, EpiStd<EPI_RES, T, T>{r, b});
    T* xn = coarse_correction<T>(l, r, tr);
    if (deco_csr* S = tr ? lv.St : lv.S) {   // x <- x + r/d + S xn: prolongation and sweep in one product
      spmv_auto<T>(S, (const T*)xn, EpiUp<T>{x, (const T*)r, D<T>(lv), add_to, stride});
      return;
    }
    // prolongation t = x + P xn, then the up smoother back into x
    spmv_auto<T>(lv.P, (const T*)xn, EpiStd<EPI_ADDZ, T, T>{t, x});
    if (relax == 17) {
      relax_sweep<T>(lv, A, t, b, x, 0);
      relax_sweep<T>(lv, A, x, b, t, 1);
      relax_sweep<T>(lv, A, t, b, x, 0, add_to, stride);
    } else {
      relax_sweep<T>(lv, A, t, b, x, -1, add_to, stride);
    }
  }

  // restriction into level l+1 (restr(epi) forms the restricted residual with epilogue epi),
  // the cycle below, and the coarse solution's location; the restriction also writes the next
  // level's zero-guess sweep (g_0 = b/d) unless that level is the coarsest, the dense tail or an
  // F-C-F level
  template <class T, class F>
  T* coarse_correction_with(size_t l, bool tr, F restr) {
    const bool last = (l + 1 == L.size());
    T* bn = last ? BC<T>() : B<T>(L[l + 1]);
    T* xn = last ? XC<T>() : X<T>(L[l + 1]);
    const bool fuse = !last && relax != 17 && !(tail && (int)l + 1 == tail_lev);
    if (fuse) restr(EpiRestrictJacobi0<T>{bn, D<T>(L[l + 1]), xn});
    else restr(EpiStd<EPI_SET, T, T>{bn, nullptr});
    cycle<T>(l + 1, bn, xn, tr, fuse);
    return xn;
  }
  // the same with the residual r of level l restricted by R
  template <class T>
  T* coarse_correction(size_t l, const T* r, bool tr) {
    return coarse_correction_with<T>(l, tr, [&](auto epi) { spmv_auto<T>(L[l].R, r, epi); });
  }

  // K = [R, -R A] of level lv (A: its matrix or the transpose) for the fused down-sweep restriction
  static deco_csr* fused_restriction(const Level& lv, const deco_csr* A) {
    deco_csr* RA = spgemm(lv.R, A);
    deco_csr* K = nullptr;
    try {
      K = restrict_operator(lv.R, RA);
    } catch (...) {
      delete RA;
      throw;
    }
    delete RA;
    return K;
  }

  // strict part of the (transposed) level matrix for a two-stage sweep of direction dir
  const deco_csr* strict(const Level& lv, bool tr, int dir) const {
    return dir > 0 ? (tr ? lv.Lot : lv.Lo) : (tr ? lv.Upt : lv.Up);
  }

  // Level l with two-stage Gauss-Seidel sweeps. Down from a zero guess (rho = b): g_0 = b/d is in
  // x (x_ready, or written here), the inner steps leave the sweep's result in t; below level 0 the
  // residual b - A t and its restriction are one product, K [b; t] with K = [R, -R A]. Up: x = t + P e,
  // rho = b - A x and g_0 = rho/d (into r and u), the inner steps add g_s to x (or, with add_to
  // set, x + g_s to add_to[stride*i]). The transposed cycle runs on A^T with the other side's
  // directions reversed.
  template <class T>
  void two_stage_level(size_t l, const T* b, T* x, bool tr, bool x_ready, T* add_to, int stride) {
    Level& lv = L[l];
    const deco_csr* A = tr ? lv.At : lv.A;
    const int nl = A->n;
    T* t = Tm<T>(lv);
    T* r = Rr<T>(lv);
    T* u = Uu<T>(lv);
    const T* d = D<T>(lv);
    const int down = tr ? -dir_up : dir_down, up = tr ? -dir_down : dir_up;
    if (!x_ready) {
      hipLaunchKernelGGL(k_relax_zero<T>, dim3(grid_for(nl, 256)), dim3(256), 0, stream(), nl, b, d, lv.cf.p, -1, x);
      DECO_CHECK_LAUNCH();
      count_bytes((double)nl * (3 * sizeof(T)));
    }
    const deco_csr* Td = strict(lv, tr, down);
    if (inner == 1) {
      spmv_auto<T>(Td, (const T*)x, EpiTwoStage<T>{nullptr, b, d, t, nullptr, 0});
    } else {
      spmv_auto<T>(Td, (const T*)x, EpiTwoStage<T>{nullptr, b, d, u, nullptr, 0});
      spmv_auto<T>(Td, (const T*)u, EpiTwoStage<T>{nullptr, b, d, t, nullptr, 0});
    }
    const T* xn;
    if (deco_csr* K = tr ? lv.Kt : lv.K) {   // b_c = R (b - A t) as K [b; t]
      xn = coarse_correction_with<T>(l, tr, [&](auto epi) { spmv_stream_c16(K, vals<T>(K), CatGather<T>{b, t}, epi); });
    } else {
      spmv_auto<T>(A, (const T*)t, EpiStd<EPI_RES, T, T>{r, b});
      xn = coarse_correction<T>(l, r, tr);
    }
    spmv_auto<T>(lv.P, xn, EpiStd<EPI_ADDZ, T, T>{x, t});
    spmv_auto<T>(A, (const T*)x, EpiResG<T>{b, d, r, u});
    const deco_csr* Tu = strict(lv, tr, up);
    if (inner == 1) {
      spmv_auto<T>(Tu, (const T*)u, EpiTwoStage<T>{x, r, d, x, add_to, stride});
    } else {
      spmv_auto<T>(Tu, (const T*)u, EpiTwoStage<T>{nullptr, r, d, t, nullptr, 0});
      spmv_auto<T>(Tu, (const T*)t, EpiTwoStage<T>{x, r, d, x, add_to, stride});
    }
  }

  void ensure_t() {
    if (have_t) return;
    for (auto& lv : L) {
      lv.At = csr_transpose(lv.A);
      if (relax == 11) {
        lv.Lot = strict_part(lv.At, true);
        lv.Upt = strict_part(lv.At, false);
        if (&lv != &L[0]) lv.Kt = fused_restriction(lv, lv.At);
      }
      if (jacobi_type()) {
        deco_csr* Q = spgemm(lv.At, lv.P);
        lv.St = up_operator(lv.P, Q, lv.d.p);
        delete Q;
      }
    }
    have_t = true;
  }

  template <class T>
  void run(const T* r, T* x, bool tr) {
    if (tr) ensure_t();
    if (L.empty()) {
      dense(nc, (const double*)(tr ? CT.p : Cinv.p), r, x, (T*)nullptr, 0);
      return;
    }
    cycle<T>(0, r, x, tr);
  }
  // FP64 in and out; with FP32 precision the cycle runs on FP32 copies
  void run64(const double* r, double* x, bool tr) {
    if (precision != DECO_PREC_FP32) {
      run<double>(r, x, tr);
      return;
    }
    r32.ensure(n);
    x32.ensure(n);
    convert(r, r32.p, n);
    run<float>(r32.p, x32.p, tr);
    convert(x32.p, x, n);
  }
  void apply(const double* r, double* x) override {
    ProfPhase ph("AMG apply");                                           // hypre_AMG.hpp:27-35
    run64(r, x, false);
  }
  void apply_t(const double* r, double* x) override { run64(r, x, true); }
  void apply32(const float* r, float* x) override {
    ProfPhase ph("AMG apply");
    run<float>(r, x, false);
  }
  void apply32_t(const float* r, float* x) override { run<float>(r, x, true); }

  // fused CPR pressure stage: the caller's residual epilogue performs the finest level's
  // zero-guess Jacobi sweep (x = b/d with this diagonal), the last sweep adds the result to the
  // caller's iterate (Jacobi-type smoothers with at least one level below the finest)
  bool stage_diag(const double** d0) override {
    if (relax == 17 || L.empty()) return false;
    *d0 = L[0].d.p;
    return true;
  }
  bool stage_diag32(const float** d0) override {
    if (relax == 17 || L.empty()) return false;
    *d0 = L[0].df.p;
    return true;
  }
  void stage(const double* b, double* x, double* add_to, int stride) override {
    ProfPhase ph("AMG apply");
    cycle<double>(0, b, x, false, true, add_to, stride);
  }
  void stage32(const float* b, float* x, float* add_to, int stride) override {
    ProfPhase ph("AMG apply");
    cycle<float>(0, b, x, false, true, add_to, stride);
  }
};

template <> double* AMG::X<double>(Level& l) { return l.x.p; }
template <> float* AMG::X<float>(Level& l) { return l.xf.p; }
template <> double* AMG::B<double>(Level& l) { return l.b.p; }
template <> float* AMG::B<float>(Level& l) { return l.bf.p; }
template <> double* AMG::Tm<double>(Level& l) { return l.t.p; }
template <> float* AMG::Tm<float>(Level& l) { return l.tf.p; }
template <> double* AMG::Rr<double>(Level& l) { return l.r.p; }
template <> float* AMG::Rr<float>(Level& l) { return l.rf.p; }
template <> double* AMG::Uu<double>(Level& l) { return l.u.p; }
template <> float* AMG::Uu<float>(Level& l) { return l.uf.p; }
template <> const double* AMG::D<double>(Level& l) { return l.d.p; }
template <> const float* AMG::D<float>(Level& l) { return l.df.p; }
template <> const double* AMG::vals<double>(deco_csr* A) { return A->v.p; }
template <> const float* AMG::vals<float>(deco_csr* A) { return A->values_f32(); }
template <> double* AMG::XC<double>() { return xc.p; }
template <> float* AMG::XC<float>() { return xcf.p; }
template <> double* AMG::BC<double>() { return bc.p; }
template <> float* AMG::BC<float>() { return bcf.p; }

deco_prec* make_amg(const deco_cpr_params& p, const deco_csr* Ap) { return new AMG(p, Ap); }

static AMG* as_amg(const deco_prec* M) {
  if (!M || M->kind != deco_prec::AMG) fail(DECO_ERR_ARG, "not an AMG preconditioner");
  return static_cast<AMG*>(const_cast<deco_prec*>(M));
}

}  // namespace deco

using namespace deco;

extern "C" {

int deco_amg_create(const deco_cpr_params* p, const deco_csr* Ap, deco_prec** out) {
  return guard([&] {
    if (!p || !Ap || !out) fail(DECO_ERR_ARG, "null argument");
    *out = make_amg(*p, Ap);
  });
}

int deco_amg_num_levels(const deco_prec* M, int* levels) {
  return guard([&] {
    if (!levels) fail(DECO_ERR_ARG, "null argument");
    *levels = (int)as_amg(M)->L.size();
  });
}

int deco_amg_level(const deco_prec* M, int level, const deco_csr** A, const deco_csr** P, const deco_csr** R,
                   const double** d, const int** cf) {
  return guard([&] {
    AMG* G = as_amg(M);
    if (level < 0 || level > (int)G->L.size()) fail(DECO_ERR_ARG, "level %d out of range", level);
    if (level == (int)G->L.size()) {
      if (A) *A = G->Ac;
      if (P) *P = nullptr;
      if (R) *R = nullptr;
      if (d) *d = nullptr;
      if (cf) *cf = nullptr;
      return;
    }
    auto& lv = G->L[level];
    if (A) *A = lv.A;
    if (P) *P = lv.P;
    if (R) *R = lv.R;
    if (d) *d = lv.d.p;
    if (cf) *cf = lv.cf.p;
  });
}

int deco_set_amg_dense_tail(int max_rows) {
  return guard([&] {
    if (max_rows < 0 || max_rows > 8192) fail(DECO_ERR_ARG, "max_rows %d outside [0, 8192]", max_rows);
    g_dense_tail_rows = max_rows;
  });
}

int deco_amg_dense_tail(const deco_prec* M, int* rows) {
  return guard([&] {
    if (!rows) fail(DECO_ERR_ARG, "null argument");
    *rows = as_amg(M)->tail;
  });
}

int deco_amg_coarse_inverse(const deco_prec* M, int* nc, const double** inv) {
  return guard([&] {
    AMG* G = as_amg(M);
    if (nc) *nc = G->nc;
    if (inv) *inv = G->Cinv.p;
  });
}

int deco_amg_interp_host(const deco_cpr_params* p, int n, const int* r, const int* c, const double* v, int* cf,
                         int* nc, int* p_nnz, int** P_r, int** P_c, double** P_v) {
  return guard([&] {
    if (!p || n < 0 || !r || !cf || !nc || !p_nnz || !P_r || !P_c || !P_v) fail(DECO_ERR_ARG, "null argument");
    amg::HostCSR H;
    H.n = n;
    H.m = n;
    H.r.assign(r, r + n + 1);
    H.c.assign(c, c + r[n]);
    H.v.assign(v, v + r[n]);
    std::vector<char> strong;
    amg::strength(H, p->amg_SoC_alpha, p->amg_max_row_sum > 0 ? p->amg_max_row_sum : 0.9, strong);
    std::vector<int> cfv;
    amg::coarsen(H, strong, p->amg_coarse_type, 0x5eed0000ull, cfv);
    int ncl = 0;
    amg::HostCSR P = amg::interp_extpi(H, strong, cfv, p->amg_prol_trunc_fac, p->amg_P_max_elmts, ncl);
    std::memcpy(cf, cfv.data(), sizeof(int) * n);
    *nc = ncl;
    *p_nnz = (int)P.c.size();
    *P_r = (int*)malloc(sizeof(int) * (n + 1));
    *P_c = (int*)malloc(sizeof(int) * std::max<size_t>(1, P.c.size()));
    *P_v = (double*)malloc(sizeof(double) * std::max<size_t>(1, P.v.size()));
    std::memcpy(*P_r, P.r.data(), sizeof(int) * (n + 1));
    if (!P.c.empty()) {
      std::memcpy(*P_c, P.c.data(), sizeof(int) * P.c.size());
      std::memcpy(*P_v, P.v.data(), sizeof(double) * P.v.size());
    }
  });
}

void deco_host_free(void* ptr) { free(ptr); }

}  // extern "C"
