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
// Smoothers: relax 18 = l1-Jacobi (diagonal sum_j |a_ij|), 0/7 = Jacobi with amg_relax_weight,
// -1 = damped Jacobi (weight 2/3) in place of hypre's hybrid Gauss-Seidel, 17 = F-C-F
// Jacobi half-sweeps. Relaxation is fused into the stream SpMV epilogue. FP32 mode runs the
// levels on f32 shadows and the coarsest solve in FP64.
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

// x = C b with C row-major n x n in MT (the coarsest inverse for the forward cycle, its
// transpose for the transposed one; or the last level's dense sub-cycle operator): one wavefront
// per row, lanes stride the columns, fixed-order butterfly sum; input/output in T, FP64
// arithmetic; with add_to set, add_to[i*stride] += x_i instead of storing x
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

// largest last level (rows) whose sub-cycle is formed as one dense operator at setup
static int g_dense_tail_rows = 2048;

struct AMG : deco_prec {
  struct Level {
    deco_csr* A = nullptr;
    deco_csr* At = nullptr;   // transposed level matrix (transposed apply)
    deco_csr* P = nullptr;
    deco_csr* R = nullptr;
    deco_csr* AP = nullptr;   // A P from the Galerkin product, until S is formed
    deco_csr* S = nullptr;    // up-sweep operator P - D^{-1} A P (Jacobi-type smoothers)
    deco_csr* St = nullptr;   // the same with A^T, for the transposed cycle
    DBuf<double> d;
    DBuf<float> df;
    DBuf<int> cf;
    DBuf<double> x, b, t, r;
    DBuf<float> xf, bf, tf, rf;
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
  DBuf<double> Z64, Z64t;
  DBuf<float> Z32, Z32t;
  int relax = 18;
  double weight = 1.0;
  bool have_t = false;

  AMG(const deco_cpr_params& p, const deco_csr* A0) {
    kind = deco_prec::AMG;
    n = A0->n;
    relax = p.amg_relax_type;
    if (relax != -1 && relax != 0 && relax != 7 && relax != 17 && relax != 18)
      fail(DECO_ERR_UNSUPPORTED, "amg_relax_type %d is not supported (use -1, 0, 7, 17 or 18)", p.amg_relax_type);
    // -1 (hypre's CPU default, hybrid Gauss-Seidel) runs as damped Jacobi with weight 2/3
    weight = p.amg_relax_weight > 0.0 ? p.amg_relax_weight : (relax == -1 ? 2.0 / 3.0 : 1.0);
    if (relax == -1 || relax == 7) relax = 0;
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
                         lv.A->v.p, l1, relax == 17 ? 1.0 : weight, lv.d.p);
      DECO_CHECK_LAUNCH();
      lv.df.alloc(nl);
      convert(lv.d.p, lv.df.p, nl);
      lv.x.alloc(nl); lv.b.alloc(nl); lv.t.alloc(nl); lv.r.alloc(nl);
      lv.xf.alloc(nl); lv.bf.alloc(nl); lv.tf.alloc(nl); lv.rf.alloc(nl);
      // the up-sweep's prolongation and Jacobi-type sweep as one operator
      if (relax != 17) lv.S = up_operator(lv.P, lv.AP, lv.d.p);
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
    if (relax != 17 && !L.empty() && L.back().A->n <= g_dense_tail_rows) form_tail(inv);
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

  // Z = 2 D^{-1} - D^{-1} A D^{-1} + S C R (I - A D^{-1}) of the last level, formed on the host in
  // FP64 and stored in FP64 and FP32 with its transpose (R = P^T, so the transposed sub-cycle with
  // A^T, S^T's counterpart P - D^{-1} A^T P and C^T is Z^T)
  void form_tail(const std::vector<double>& inv) {
    ProfPhase ph("dense tail");
    Level& lv = L.back();
    const int nl = lv.A->n;
    amg::HostCSR A, R, S;
    download(lv.A, A);
    download(lv.R, R);
    download(lv.S, S);
    std::vector<double> d(nl), Z, Zt((size_t)nl * nl);
    lv.d.download(d.data(), nl);
    amg::dense_tail(A, d, R, S, inv, Z);
    for (int i = 0; i < nl; ++i)
      for (int k = 0; k < nl; ++k) Zt[(size_t)k * nl + i] = Z[(size_t)i * nl + k];
    std::vector<float> z32(Z.begin(), Z.end()), zt32(Zt.begin(), Zt.end());
    Z64.upload(Z.data(), Z.size());
    Z64t.upload(Zt.data(), Zt.size());
    Z32.upload(z32.data(), z32.size());
    Z32t.upload(zt32.data(), zt32.size());
    sync();   // the host arrays go out of scope here
    tail = nl;
  }

  template <class MT>
  void dense(int m, const MT* C, const float* b, float* x, float* add_to, int stride) {
    hipLaunchKernelGGL((k_dense<float, MT>), dim3((m + 3) / 4), dim3(256), 0, stream(), m, C, b, x, add_to, stride);
    DECO_CHECK_LAUNCH();
  }
  template <class MT>
  void dense(int m, const MT* C, const double* b, double* x, double* add_to, int stride) {
    hipLaunchKernelGGL((k_dense<double, MT>), dim3((m + 3) / 4), dim3(256), 0, stream(), m, C, b, x, add_to, stride);
    DECO_CHECK_LAUNCH();
  }
  // the dense tail operator in the cycle's precision
  const double* tail_op(bool tr, const double*) const { return tr ? Z64t.p : Z64.p; }
  const float* tail_op(bool tr, const float*) const { return tr ? Z32t.p : Z32.p; }

  void release() {
    for (auto& lv : L) {
      delete lv.A; delete lv.At; delete lv.P; delete lv.R; delete lv.AP; delete lv.S; delete lv.St;
      lv.A = lv.At = lv.P = lv.R = lv.AP = lv.S = lv.St = nullptr;
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
    if (tail && l + 1 == L.size()) {   // the whole sub-cycle of the last level as one product
      dense(tail, tail_op(tr, b), b, x, add_to, stride);
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
      relax_sweep<T>(lv, A, x, b, t, 1);
      relax_sweep<T>(lv, A, t, b, x, 0);
    } else if (!x_ready) {
      hipLaunchKernelGGL(k_relax_zero<T>, dim3(grid_for(nl, 256)), dim3(256), 0, stream(), nl, b, D<T>(lv), lv.cf.p, -1, x);
      DECO_CHECK_LAUNCH();
    }
    // residual and restriction (the restriction also does the next level's zero-guess sweep
    // for the Jacobi-type smoothers)
    spmv_auto<T>(A, x, EpiStd<EPI_RES, T, T>{r, b});
    const bool last = (l + 1 == L.size());
    T* bn = last ? BC<T>() : B<T>(L[l + 1]);
    T* xn = last ? XC<T>() : X<T>(L[l + 1]);
    const bool fuse = !last && relax != 17 && !(tail && l + 2 == L.size());
    if (fuse)
      spmv_auto<T>(lv.R, (const T*)r, EpiRestrictJacobi0<T>{bn, D<T>(L[l + 1]), xn});
    else
      spmv_auto<T>(lv.R, (const T*)r, EpiStd<EPI_SET, T, T>{bn, nullptr});
    cycle<T>(l + 1, bn, xn, tr, fuse);
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

  void ensure_t() {
    if (have_t) return;
    for (auto& lv : L) {
      lv.At = csr_transpose(lv.A);
      if (relax != 17) {
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
    if (max_rows < 0 || max_rows > 4096) fail(DECO_ERR_ARG, "max_rows %d outside [0, 4096]", max_rows);
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
