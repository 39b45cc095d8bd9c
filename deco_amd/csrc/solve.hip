// Krylov solvers on the device: rgmres, lgmres, bicgstab, lbicgstab, pcg (solver/*.tpp).
//
// The host drives the iteration; every vector operation is a device kernel on the library
// stream. Reductions are deterministic (fixed grid, fixed-order partial sums). The GMRES
// Hessenberg column comes back to the host each inner step (j+2 doubles) and the Givens
// rotations and back-substitution run on the host with the reference's arithmetic
// (solver/rgmres.tpp:80-104,141-160), so they are bit-identical given the same column.
#include "prec.hpp"
#include "blas.hpp"
#include <cmath>
#include <cstring>
#include <vector>
#include <chrono>

namespace deco {

// ------------------------------------------------------------------ kernels ------------
__global__ void k_scale(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}
__global__ void k_add(int n, const double* __restrict__ z, double* __restrict__ x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += z[i];
}
__global__ void k_sub_from(int n, const double* __restrict__ b, double* __restrict__ r) {  // r = b - r
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}

// One fused modified-Gram-Schmidt step. If vprev: h = reduce(part_prev) (every block gets the
// same value), w = 1.0*w + (-h)*vprev (rgmres.tpp:74) and block 0 stores h. Then the partial
// sums of vcur.w (rgmres.tpp:73), or of w.w when vcur is null (rgmres.tpp:76).
__global__ __launch_bounds__(kRedBlock) void k_mgs_step(int n, const double* __restrict__ part_prev, int np,
                                                       const double* __restrict__ vprev, double* __restrict__ w,
                                                       const double* __restrict__ vcur, double* __restrict__ part_out,
                                                       double* __restrict__ hout) {
  __shared__ double sm[kRedBlock / 64];
  double h = 0.0;
  if (vprev) {
    h = reduce_partials(part_prev, np);
    if (blockIdx.x == 0 && threadIdx.x == 0) *hout = h;
  }
  const double mh = -h;
  double s = 0.0;
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    double wi = w[i];
    if (vprev) {
      wi = 1.0 * wi + mh * vprev[i];
      w[i] = wi;
    }
    s += (vcur ? vcur[i] : wi) * wi;
  }
  s = block_sum<kRedBlock>(s, sm);
  if (threadIdx.x == 0) part_out[blockIdx.x] = s;
}

// nrm = sqrt(reduce(part)); v = (1/nrm) * w (rgmres.tpp:76-77); block 0 stores nrm.
__global__ __launch_bounds__(kRedBlock) void k_mgs_norm(int n, const double* __restrict__ part, int np,
                                                       const double* __restrict__ w, double* __restrict__ v,
                                                       double* __restrict__ hout) {
  const double nrm = sqrt(reduce_partials(part, np));
  if (blockIdx.x == 0 && threadIdx.x == 0) *hout = nrm;
  const double inv = 1.0 / nrm;
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) v[i] = inv * w[i];
}

// w = g0*V0 + sum_{i=1}^{j-1} g_i V_i in the order of rgmres.tpp:106-109; the coefficients come
// as kernel arguments (gc, j <= kCombineArgs) or from device memory (g)
constexpr int kCombineArgs = 64;
struct CombineCoef {
  double g[kCombineArgs];
};
template <class QT>
__global__ void k_combine(int n, int j, const QT* __restrict__ V, int64_t ldv, const double* __restrict__ g,
                          CombineCoef gc, double* __restrict__ w) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double* c = j <= kCombineArgs ? gc.g : g;
  double s = c[0] * (double)V[k];
  for (int i = 1; i < j; ++i) s = 1.0 * s + c[i] * (double)V[(int64_t)i * ldv + k];
  w[k] = s;
}

// ---------------- one-reduction Arnoldi step (delayed classical Gram-Schmidt, DCGS2) -------
// Each block owns kArnTile consecutive rows (kArnRpt per thread, strided by the block size so
// every load instruction is coalesced). Values of one step, in this order:
//   s_i = q_i.u (i < j), alpha = u.u, z_i = q_i.w (i < j), beta = u.w   (the last two only if w)
// part[v*nb + block] = fixed-order block partial of value v.
constexpr int kArnRpt = 8;
constexpr int kArnBlock = 256;
constexpr int kArnTile = kArnRpt * kArnBlock;
constexpr int kDcgs2MaxRestart = 1000;   // LDS of the dots kernel: (2 restart + 2) * 4 doubles
// the reduction's values and tags live in pinned host buffers of 4096 entries each
static_assert(2 * kDcgs2MaxRestart + 2 <= 4096, "DCGS2 reduction exceeds the pinned buffers");

__global__ __launch_bounds__(kArnBlock) void k_arn_dots(int n, int j, const double* __restrict__ Q,
                                                       const double* __restrict__ u, const double* __restrict__ w,
                                                       double* __restrict__ part, int nb) {
  extern __shared__ double red[];  // [value][wave]
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t base = (int64_t)blockIdx.x * kArnTile + tid;
  double uu[kArnRpt], ww[kArnRpt];
#pragma unroll
  for (int k = 0; k < kArnRpt; ++k) {
    const int64_t idx = base + k * kArnBlock;
    uu[k] = idx < n ? u[idx] : 0.0;
    ww[k] = (w && idx < n) ? w[idx] : 0.0;
  }
  const bool hw = (w != nullptr);
  for (int i = 0; i < j; ++i) {
    const double* q = Q + (size_t)i * n;
    double qq[kArnRpt];
#pragma unroll
    for (int k = 0; k < kArnRpt; ++k) {
      const int64_t idx = base + k * kArnBlock;
      qq[k] = idx < n ? __builtin_nontemporal_load(q + idx) : 0.0;
    }
    double ps = 0.0, pz = 0.0;
#pragma unroll
    for (int k = 0; k < kArnRpt; ++k) { ps += qq[k] * uu[k]; pz += qq[k] * ww[k]; }
    ps = wave_sum(ps);
    if (hw) pz = wave_sum(pz);
    if (lane == 0) {
      red[i * 4 + wid] = ps;
      if (hw) red[(j + 1 + i) * 4 + wid] = pz;
    }
  }
  double pa = 0.0, pb = 0.0;
#pragma unroll
  for (int k = 0; k < kArnRpt; ++k) { pa += uu[k] * uu[k]; pb += uu[k] * ww[k]; }
  pa = wave_sum(pa);
  if (hw) pb = wave_sum(pb);
  if (lane == 0) {
    red[j * 4 + wid] = pa;
    if (hw) red[(2 * j + 1) * 4 + wid] = pb;
  }
  __syncthreads();
  const int nv = hw ? 2 * j + 2 : j + 1;
  for (int v = tid; v < nv; v += kArnBlock)
    part[(size_t)v * nb + blockIdx.x] = ((red[v * 4] + red[v * 4 + 1]) + red[v * 4 + 2]) + red[v * 4 + 3];
}

// The tag that validates a reduced value published to the host: the value's bits XOR a bijective
// mix of the reduction's sequence number. The host accepts value v of reduction seq only when
// tag[v] ^ bits[v] == tag_mix(seq), so neither a value that reaches host memory after its tag
// nor a tag left from an earlier reduction passes (both have been observed: a value store to
// coherent host memory can become visible after a later, fenced tag store).
__host__ __device__ inline unsigned long long tag_mix(long long seq) {
  return (unsigned long long)seq * 0x9E3779B97F4A7C15ull;
}

// out[v] = sum over the nb block partials of value v: one workgroup per value, every lane a
// strided subset (loads issued together), then a fixed-order workgroup sum. The value's bits also
// go to host (coherent pinned host memory) with their tag, so the host reads them as soon as they
// validate (no event in the stream between this kernel and the update sweep behind it).
__global__ __launch_bounds__(256) void k_arn_reduce(int nb, const double* __restrict__ part, double* __restrict__ out,
                                                   unsigned long long* __restrict__ host,
                                                   unsigned long long* __restrict__ tag, long long seq) {
  __shared__ double sm[4];
  const double* p = part + (size_t)blockIdx.x * nb;
  double s = 0.0;
#pragma unroll 8
  for (int b = threadIdx.x; b < nb; b += 256) s += p[b];
  s = block_sum<256>(s, sm);
  if (threadIdx.x == 0) {
    out[blockIdx.x] = s;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(s);
    host[blockIdx.x] = bits;
    __threadfence_system();
    tag[blockIdx.x] = bits ^ tag_mix(seq);
  }
}

// Waits until the reduction with sequence number seq has published values [0, nv) to the pinned
// host buffer and copies each validated value into vals (host memory the solver then reads).
// While spinning it polls the stream now and then: a stream error is raised, and a drained
// stream without the values is an error rather than an endless wait.
static void wait_tags(int nv, long long seq, double* vals) {
  const unsigned long long* t = reinterpret_cast<const unsigned long long*>(ctx().tags);
  const unsigned long long* h = reinterpret_cast<const unsigned long long*>(ctx().pinned);
  const unsigned long long mix = tag_mix(seq);
  int done = 0;   // values [0, done) validated
  auto arrived = [&] {
    for (; done < nv; ++done) {
      const unsigned long long tv = __atomic_load_n(&t[done], __ATOMIC_ACQUIRE);
      const unsigned long long bv = __atomic_load_n(&h[done], __ATOMIC_ACQUIRE);
      if ((tv ^ bv) != mix) return false;
      std::memcpy(&vals[done], &bv, sizeof(double));
    }
    return true;
  };
  for (unsigned spin = 1;; ++spin) {
    if (arrived()) return;
    if ((spin & 1023) == 0) {
      const hipError_t e = hipStreamQuery(stream());
      if (e == hipSuccess) {
        if (arrived()) return;
        fail(DECO_ERR_DEVICE, "reduction values missing after the stream drained");
      }
      if (e != hipErrorNotReady) DECO_HIP(e);
    }
  }
}

// The step's scalars from the reduced values (same operation order on host and device):
// rho = sqrt(alpha - s.s) (norm of the reorthogonalised u), gamma = (beta - s.z) / rho^2.
__host__ __device__ inline void arn_scalars(int j, const double* red, double& rho, double& gam) {
  double ss = 0.0, sz = 0.0;
  for (int i = 0; i < j; ++i) { ss += red[i] * red[i]; sz += red[i] * red[j + 1 + i]; }
  const double d = red[j] - ss;
  rho = sqrt(d > 0.0 ? d : 0.0);
  gam = (red[2 * j + 1] - sz) / (rho * rho);
}

// v_j = (u - Q s) / rho  and  u_next = (w - Q (z - gamma s) - gamma u) / rho, one sweep over Q
__global__ __launch_bounds__(kArnBlock) void k_arn_update(int n, int j, const double* __restrict__ Q,
                                                         const double* __restrict__ u, const double* __restrict__ w,
                                                         const double* __restrict__ red, double* __restrict__ vout,
                                                         double* __restrict__ unext) {
  double rho, gam;
  arn_scalars(j, red, rho, gam);
  const double irho = 1.0 / rho;
  const int64_t base = (int64_t)blockIdx.x * kArnTile + threadIdx.x;
  double a1[kArnRpt], a2[kArnRpt];
#pragma unroll
  for (int k = 0; k < kArnRpt; ++k) { a1[k] = 0.0; a2[k] = 0.0; }
  for (int i = 0; i < j; ++i) {
    const double* q = Q + (size_t)i * n;
    const double si = red[i], ci = red[j + 1 + i] - gam * red[i];
#pragma unroll
    for (int k = 0; k < kArnRpt; ++k) {
      const int64_t idx = base + k * kArnBlock;
      if (idx < n) {
        const double qv = __builtin_nontemporal_load(q + idx);
        a1[k] += qv * si;
        a2[k] += qv * ci;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kArnRpt; ++k) {
    const int64_t idx = base + k * kArnBlock;
    if (idx < n) {
      const double uv = u[idx];
      vout[idx] = (uv - a1[k]) * irho;
      unext[idx] = ((w[idx] - a2[k]) - gam * uv) * irho;
    }
  }
}

// partial sums of up to three products over the grid: part[s*kRedMaxGrid + block]
template <int K>
__global__ __launch_bounds__(kRedBlock) void k_dots(int n, const double* __restrict__ a0, const double* __restrict__ b0,
                                                   const double* __restrict__ a1, const double* __restrict__ b1,
                                                   double* __restrict__ part) {
  __shared__ double sm[kRedBlock / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    s0 += a0[i] * b0[i];
    if (K > 1) s1 += a1[i] * b1[i];
  }
  s0 = block_sum<kRedBlock>(s0, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = s0;
  if (K > 1) {
    s1 = block_sum<kRedBlock>(s1, sm);
    if (threadIdx.x == 0) part[kRedMaxGrid + blockIdx.x] = s1;
  }
}

// BiCGSTAB: p = beta*p + (-beta*omega)*v + 1.0*r (bicgstab.tpp:254, axpby 3-term form)
__global__ void k_bicg_p(int n, double beta, double bo, const double* __restrict__ r, const double* __restrict__ v,
                         double* __restrict__ p) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = beta * p[i] + bo * v[i] + 1.0 * r[i];
}

// r = 1.0*r + (-a)*v ; [x = 1.0*x + a*y] ; partials of r.r (and r.rt)
template <bool UPDX, bool DOT_RT>
__global__ __launch_bounds__(kRedBlock) void k_bicg_update(int n, double a, const double* __restrict__ v,
                                                          double* __restrict__ r, double ax,
                                                          const double* __restrict__ y, double* __restrict__ x,
                                                          const double* __restrict__ rt, double* __restrict__ part) {
  __shared__ double sm[kRedBlock / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    if (UPDX) x[i] = 1.0 * x[i] + ax * y[i];
    const double ri = 1.0 * r[i] + (-a) * v[i];
    r[i] = ri;
    s0 += ri * ri;
    if (DOT_RT) s1 += ri * rt[i];
  }
  s0 = block_sum<kRedBlock>(s0, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = s0;
  if (DOT_RT) {
    s1 = block_sum<kRedBlock>(s1, sm);
    if (threadIdx.x == 0) part[kRedMaxGrid + blockIdx.x] = s1;
  }
}

// y = b*y + a*x
__global__ void k_axpby2(int n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = b * y[i] + a * x[i];
}

// ------------------------------------------------------------------ workspace ----------
// Work vectors of the Krylov solvers, kept between solves and grown on demand, so that a solve
// does not allocate and free its basis (1.6 GB at 4M unknowns and restart 50, half as much again
// for the FP32 flexible directions) every time. One solve at a time uses them (the library runs on one
// stream); deco_workspace_release() frees them. The array lives until the process ends and is
// never destroyed: a destructor at exit would free device memory after the HIP runtime (and any
// profiling tool attached to it) has shut down; the driver reclaims the memory with the process.
constexpr int kWsSlots = 12;
static DBuf<double>* const g_ws = new DBuf<double>[kWsSlots];
struct WsRef {
  double* p;
};
static WsRef ws(int slot, size_t count) {
  g_ws[slot].ensure(std::max<size_t>(count, 1));
  return WsRef{g_ws[slot].p};
}

// w = sum_i g_i V_i (j columns of V with stride ldv), the j host coefficients passed to the kernel
// as arguments when they fit, else uploaded to gdev first
template <class QT>
static void combine(int n, int j, const QT* V, int64_t ldv, const double* g, double* gdev, double* w) {
  CombineCoef gc{};
  if (j <= kCombineArgs) std::memcpy(gc.g, g, sizeof(double) * std::max(j, 0));
  else copy_h2d(gdev, g, j);
  hipLaunchKernelGGL(k_combine<QT>, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, j, V, ldv,
                     (const double*)gdev, gc, w);
  DECO_CHECK_LAUNCH();
  count_bytes((double)n * (j * sizeof(QT) + 8) + 8.0 * j);
}

// ------------------------------------------------------------------ host helpers -------
struct Solver {
  int n;
  int nb;                 // reduction grid
  RedScratch rs;
  std::vector<double> hpart;
  explicit Solver(int n_) : n(n_), nb(red_grid(n_)), hpart(2 * kRedMaxGrid) {}

  unsigned g256() const { return grid_for(n, 256); }
  // sum the first k partial slots on the host (fixed order)
  void host_sums(int k, double* out) {
    DECO_HIP(hipMemcpyAsync(ctx().pinned, rs.part.p, sizeof(double) * kRedMaxGrid * k, hipMemcpyDeviceToHost,
                            stream()));
    sync();
    for (int s = 0; s < k; ++s) {
      double t = 0.0;
      for (int b = 0; b < nb; ++b) t += ctx().pinned[s * kRedMaxGrid + b];
      out[s] = t;
    }
  }
  double dot(const double* a, const double* b) {
    hipLaunchKernelGGL((k_dots<1>), dim3(nb), dim3(kRedBlock), 0, stream(), n, a, b, nullptr, nullptr, rs.part.p);
    DECO_CHECK_LAUNCH();
    count_bytes(16.0 * n + 8.0 * nb);
    double s;
    host_sums(1, &s);
    return s;
  }
  double nrm2(const double* a) { return std::sqrt(dot(a, a)); }
  void dot2(const double* a0, const double* b0, const double* a1, const double* b1, double* out) {
    hipLaunchKernelGGL((k_dots<2>), dim3(nb), dim3(kRedBlock), 0, stream(), n, a0, b0, a1, b1, rs.part.p);
    DECO_CHECK_LAUNCH();
    count_bytes(32.0 * n + 16.0 * nb);
    host_sums(2, out);
  }
  void scale(double a, const double* x, double* y) {
    hipLaunchKernelGGL(k_scale, dim3(g256()), dim3(256), 0, stream(), n, a, x, y);
    DECO_CHECK_LAUNCH();
    count_bytes(16.0 * n);
  }
  void add(const double* z, double* x) {
    hipLaunchKernelGGL(k_add, dim3(g256()), dim3(256), 0, stream(), n, z, x);
    DECO_CHECK_LAUNCH();
    count_bytes(24.0 * n);
  }
  void axpby(double a, const double* x, double b, double* y) {
    hipLaunchKernelGGL(k_axpby2, dim3(g256()), dim3(256), 0, stream(), n, a, x, b, y);
    DECO_CHECK_LAUNCH();
    count_bytes(24.0 * n);
  }
};

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

// Event timing of the solvers' products with A (deco_time_outer_products): a pair of events
// around each product on the solver stream, created on first use and reused.
struct OuterTiming {
  bool on = false;
  std::vector<hipEvent_t> ev;
  size_t used = 0;   // events recorded since timing was switched on
};
static OuterTiming& outer_timing() {
  static OuterTiming* t = new OuterTiming();   // never destroyed: its events outlive the runtime's teardown
  return *t;
}
static void outer_mark() {
  OuterTiming& t = outer_timing();
  if (!t.on) return;
  if (t.used == t.ev.size()) {
    hipEvent_t e;
    DECO_HIP(hipEventCreate(&e));
    t.ev.push_back(e);
  }
  DECO_HIP(hipEventRecord(t.ev[t.used++], stream()));
}

// The solvers' operator: products with A (or A^T) on its FP64 cell-block shadow when the CPR
// setup recorded a cell size, else on the CSR arrays.
struct AOp {
  const deco_csr* A;
  bool transpose;
  template <class XT, class E>
  void product(const XT* x, E e) const {
    const deco_csr* M = transpose ? A->At : A;
    if (A->cell > 1)
      if (const deco_bsr* B = M->block64(A->cell, A->cell)) {
        bsr_stream_v<double>(B, B->bv64.p, x, e);
        return;
      }
    spmv_stream_f(M, M->v.p, x, e);
  }
  void apply(const double* x, double* y) const {
    outer_mark();
    product(x, EpiOuter{y});
    outer_mark();
  }
  // x held in FP32 (flexible GMRES directions), FP64 values and sums
  void apply(const float* x, double* y) const {
    outer_mark();
    product(x, EpiOuter{y});
    outer_mark();
  }
  void residual(const double* x, const double* b, double* r) const {  // r = b - A x
    product(x, EpiStd<EPI_RES, double, double>{r, b});
  }
};

struct MOp {
  deco_prec* M;
  bool transpose;
  int n;
  void apply(const double* r, double* x) const {
    if (!M) copy_d2d(x, r, n);
    else if (transpose) M->apply_t(r, x);
    else M->apply(r, x);
  }
  // the same written in FP32 (an FP32 preconditioner's flexible GMRES directions)
  void apply_f32(const double* r, float* x) const {
    if (transpose) M->apply_t_f32out(r, x);
    else M->apply_f32out(r, x);
  }
};

// Flexible GMRES directions Z = [M^{-1} v_j], n x m, stored in FP32: an FP32 preconditioner's
// results are FP32 values, and the method only needs the same stored z_j in w = A z_j and in
// the update x += Z y.
static float* flex_basis(int n, int m) {
  return reinterpret_cast<float*>(ws(6, ((size_t)n * m + 1) / 2).p);
}

// rgmres (solver/rgmres.tpp:9-139) and lgmres (solver/lgmres.tpp:2-134)
static void gmres(bool left, const deco_sol_params& sp, const AOp& A, const MOp& M, const double* b, double* x,
                  deco_sol_info& info) {
  const int n = A.transpose ? A.A->m : A.A->n;
  Solver S(n);
  const double tol = sp.reltol;
  const int m = sp.restart, itmax = sp.itmax;
  if (m <= 0) fail(DECO_ERR_ARG, "restart must be positive");
  // An FP32 preconditioner is not exactly linear (rounding), so M^{-1}(V y) differs from
  // sum_j M^{-1}(v_j) y_j by O(eps_f32) and the Arnoldi residual estimate stops describing the
  // update. For that case the right-preconditioned solve keeps Z = [M^{-1} v_j] (flexible GMRES)
  // and updates x += Z y; the FP64 preconditioner keeps the form of rgmres.tpp:106-118.
  const bool flex = !left && M.M && M.M->precision == DECO_PREC_FP32;
  WsRef r = ws(0, n), z = ws(1, n), w = ws(2, n), V = ws(3, (size_t)n * (m + 1)), gdev = ws(4, m + 1),
        hcol = ws(5, m + 2);
  float* Z = flex ? flex_basis(n, m) : nullptr;
  std::vector<double> H((size_t)(m + 1) * (m + 3), 0.0);
  double* g = H.data() + (size_t)(m + 1) * m;
  double* cs = g + (m + 1);
  double* sn = cs + (m + 1);
  auto Hm = [&](int i, int j) -> double& { return H[i + (size_t)j * (m + 1)]; };
  int flag = 1, iter = 0, napp = 0, nspmv = 0;
  double bnorm, absTol, rnorm = 0.0;
  if (!left) {
    bnorm = S.nrm2(b);
    absTol = tol * bnorm;
    if (!sp.use_x0) { zero_async(x, sizeof(double) * n); copy_d2d(r.p, b, n); }
    else { A.residual(x, b, r.p); ++nspmv; }
  } else {
    if (!sp.use_x0) { zero_async(x, sizeof(double) * n); M.apply(b, r.p); ++napp; }
    else { A.residual(x, b, z.p); ++nspmv; M.apply(z.p, r.p); ++napp; }
    bnorm = S.nrm2(r.p);
    absTol = tol * bnorm;
  }
  const int nb = S.nb;
  std::vector<double> hh(m + 2);
  while (iter < itmax) {
    std::fill(H.begin(), H.end(), 0.0);
    g[0] = S.nrm2(r.p);
    S.scale(1.0 / g[0], r.p, V.p);
    int j = 0;
    while (j < m && iter < itmax) {
      rnorm = std::fabs(g[j]);
      if (rnorm < absTol) { flag = 0; break; }
      double* vj = V.p + (size_t)j * n;
      if (!left && flex) {
        float* zj = Z + (size_t)j * n;
        M.apply_f32(vj, zj);
        A.apply(zj, w.p);
      } else if (!left) {
        M.apply(vj, z.p);
        A.apply(z.p, w.p);
      } else { A.apply(vj, z.p); M.apply(z.p, w.p); }
      ++napp; ++nspmv;
      // fused MGS: j+2 kernels, partial slots ping-pong
      double* slot[2] = {S.rs.slot(0), S.rs.slot(1)};
      hipLaunchKernelGGL(k_mgs_step, dim3(nb), dim3(kRedBlock), 0, stream(), n, (const double*)nullptr, 0,
                         (const double*)nullptr, w.p, (const double*)V.p, slot[0], (double*)nullptr);
      DECO_CHECK_LAUNCH();
      count_bytes(8.0 * n + 8.0 * nb);
      for (int i = 1; i <= j + 1; ++i) {
        const double* vprev = V.p + (size_t)(i - 1) * n;
        const double* vcur = (i <= j) ? V.p + (size_t)i * n : nullptr;
        hipLaunchKernelGGL(k_mgs_step, dim3(nb), dim3(kRedBlock), 0, stream(), n, (const double*)slot[(i - 1) & 1], nb,
                           vprev, w.p, vcur, slot[i & 1], hcol.p + (i - 1));
        DECO_CHECK_LAUNCH();
        count_bytes((vcur ? 32.0 : 24.0) * n + 16.0 * nb);
      }
      hipLaunchKernelGGL(k_mgs_norm, dim3(nb), dim3(kRedBlock), 0, stream(), n, (const double*)slot[(j + 1) & 1], nb,
                         (const double*)w.p, V.p + (size_t)(j + 1) * n, hcol.p + (j + 1));
      DECO_CHECK_LAUNCH();
      count_bytes(16.0 * n + 8.0 * nb);
      DECO_HIP(hipMemcpyAsync(hh.data(), hcol.p, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, stream()));
      sync();
      for (int i = 0; i <= j + 1; ++i) Hm(i, j) = hh[i];
      for (int i = 0; i <= j - 1; ++i) apply_givens(cs[i], sn[i], Hm(i, j), Hm(i + 1, j));
      comp_givens(Hm(j, j), Hm(j + 1, j), cs[j], sn[j]);
      apply_givens(cs[j], sn[j], Hm(j, j), Hm(j + 1, j));
      apply_givens(cs[j], sn[j], g[j], g[j + 1]);
      ++j;
      ++iter;
    }
    for (int i = j - 1; i >= 0; --i) {                      // rgmres.tpp:95-104
      if (Hm(i, i) == 0.0) fail(DECO_ERR_BREAKDOWN, "Zero diagonal element in upper triangular matrix");
      g[i] /= Hm(i, i);
      for (int k = 0; k < i; ++k) g[k] -= Hm(k, i) * g[i];
    }
    if (flex && j > 0) combine(n, j, (const float*)Z, (int64_t)n, g, gdev.p, w.p);
    else if (!flex) combine(n, std::max(j, 1), (const double*)V.p, (int64_t)n, g, gdev.p, w.p);
    if (flex) {
      if (j > 0) S.add(w.p, x);
      A.residual(x, b, r.p); ++nspmv;
    } else if (!left) {
      M.apply(w.p, z.p); ++napp;
      S.add(z.p, x);
      A.residual(x, b, r.p); ++nspmv;
    } else {
      S.add(w.p, x);
      A.residual(x, b, z.p); ++nspmv;
      M.apply(z.p, r.p); ++napp;
    }
    rnorm = S.nrm2(r.p);
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

// rgmres / lgmres with the one-reduction delayed-reorthogonalisation Arnoldi (DCGS2:
// Swirydowicz, Langou, Ananthan, Yang, Thomas, "Low synchronization Gram-Schmidt and GMRES
// algorithms", 2020; Bielich et al., Parallel Computing 2022). Step j applies the operator to
// u_j, the once-orthogonalised (unnormalised) candidate for v_j, and one sweep computes
// s = Q^T u, |u|^2, z = Q^T w, u.w. Then u_j = Q s + rho v_j (rho by Pythagoras), which
// finalises Hessenberg column j-1 (H(:,j-1) += s, H(j,j-1) = rho), and a second sweep writes
// v_j and the next candidate u_{j+1} = B v_j - Q_{j+1} c. Reading the basis twice per step
// instead of 2(j+1) times. The residual estimate of column j-1 is known at step j, so a solve
// applies the operator once more than it finalises columns. In exact arithmetic the iterates
// are those of solver/rgmres.tpp and lgmres.tpp (same Krylov space, same least-squares problem,
// the same Givens arithmetic, rgmres.tpp:80-104,141-160).
// With an FP32 preconditioner (right variant) Z~ = [M^{-1} u_j] is stored; the directions that
// span the Arnoldi relation are Z = Z~ T^{-1} with T(:,j) = [s; rho], so x += Z~ (T^{-1} y).
static void gmres_dcgs2(bool left, const deco_sol_params& sp, const AOp& A, const MOp& M, const double* b,
                        double* x, deco_sol_info& info) {
  const int n = A.transpose ? A.A->m : A.A->n;
  Solver S(n);
  const double tol = sp.reltol;
  const int m = sp.restart, itmax = sp.itmax;
  if (m <= 0) fail(DECO_ERR_ARG, "restart must be positive");
  // flexible directions: always with an FP32 preconditioner, and with an FP64 one when the
  // caller asks (sp.flexible), kept in FP64 then; the same Krylov space in exact arithmetic,
  // and no final M^{-1}(V y) application
  const bool flex32 = !left && M.M && M.M->precision == DECO_PREC_FP32;
  const bool flex64 = !left && M.M && !flex32 && sp.flexible;
  const bool flex = flex32 || flex64;
  const int nbk = (int)std::max<int64_t>(1, ((int64_t)n + kArnTile - 1) / kArnTile);
  WsRef r = ws(0, n), z = ws(1, n), w = ws(2, n), ua = ws(7, n), ub = ws(8, n), V = ws(3, (size_t)n * (m + 1)),
        gdev = ws(4, m + 1);
  WsRef part = ws(9, (size_t)(2 * m + 2) * nbk), redd = ws(10, 2 * m + 2);
  float* Z = flex32 ? flex_basis(n, m) : nullptr;
  double* Z64 = flex64 ? ws(11, (size_t)n * m).p : nullptr;
  const size_t ld = (size_t)m + 1;
  std::vector<double> Hr(ld * m), Rm(ld * m), T(flex ? ld * m : 0), g(m + 1), cs(m), sn(m), hs(m + 1);
  std::vector<double> red_h(2 * (size_t)m + 2);   // the reduced values of the current step
  double* red = red_h.data();
  int flag = 1, iter = 0, napp = 0, nspmv = 0;
  double bnorm = -1.0, absTol = 0.0, rnorm = 0.0;
  // Right variant from x = 0: r_0 = b, so step 0 reads b itself and |b| is the norm its reduction
  // returns (set at step 0); the first cycle's update writes x instead of adding to it. No host
  // round trip before the first operator application.
  const bool from_zero = !left && !sp.use_x0;
  if (!left) {
    if (sp.use_x0) {
      bnorm = S.nrm2(b);
      absTol = tol * bnorm;
      A.residual(x, b, r.p); ++nspmv;
    }
  } else {
    if (!sp.use_x0) { zero_async(x, sizeof(double) * n); M.apply(b, r.p); ++napp; }
    else { A.residual(x, b, z.p); ++nspmv; M.apply(z.p, r.p); ++napp; }
    bnorm = S.nrm2(r.p);
    absTol = tol * bnorm;
  }
  // w = B u_j for step j (B = A M^{-1} right, M^{-1} A left); z / Z column j hold M^{-1} u_j
  auto apply_op = [&](int jj, const double* uu) {
    if (flex32) {
      float* zj = Z + (size_t)jj * n;
      M.apply_f32(uu, zj);
      A.apply(zj, w.p);
    } else if (flex64) {
      double* zj = Z64 + (size_t)jj * n;
      M.apply(uu, zj);
      A.apply(zj, w.p);
    } else if (!left) {
      M.apply(uu, z.p);
      A.apply(z.p, w.p);
    } else { A.apply(uu, z.p); M.apply(z.p, w.p); }
    ++napp; ++nspmv;
  };
  if (from_zero && itmax <= 0) {   // no step: x = 0 and |b| here
    zero_async(x, sizeof(double) * n);
    bnorm = S.nrm2(b);
    absTol = tol * bnorm;
  }
  long long red_seq = 0;   // sequence number and value count of the last queued reduction
  int red_nv = 0;
  while (iter < itmax) {
    std::fill(Hr.begin(), Hr.end(), 0.0);
    std::fill(Rm.begin(), Rm.end(), 0.0);
    const int it0 = iter;
    double* bufs[2] = {ua.p, ub.p};
    const bool first_zero = from_zero && it0 == 0;   // x is still 0 (never written)
    const double* u = first_zero ? b : r.p;
    int j = 0, k = 0;
    bool have_w = false;   // step j's operator application is already queued
    double rn1 = -1.0, rn2 = -1.0, rn3 = -1.0;   // the last three residual estimates of this cycle
    // one reduction of step j: s = Q^T u, |u|^2 (and z = Q^T w, u.w with w); the update sweep
    // is queued behind it when w is there. The reduction writes the values to red (pinned host
    // memory) itself, each followed by its sequence tag, so the host reads them while the update
    // runs; the host has read red before it queues the next reduction.
    auto reduce_step = [&](const double* uu, const double* wp, double* un) {
      const int nv = wp ? 2 * j + 2 : j + 1;
      hipLaunchKernelGGL(k_arn_dots, dim3(nbk), dim3(kArnBlock), sizeof(double) * 4 * (2 * j + 2), stream(), n, j,
                         (const double*)V.p, uu, wp, part.p, nbk);
      DECO_CHECK_LAUNCH();
      count_bytes(8.0 * n * (j + (wp ? 2 : 1)) + 8.0 * nv * nbk);
      red_seq = ++ctx().tag_seq;
      red_nv = nv;
      hipLaunchKernelGGL(k_arn_reduce, dim3(nv), dim3(256), 0, stream(), nbk, (const double*)part.p, redd.p,
                         reinterpret_cast<unsigned long long*>(ctx().pinned_dev),
                         reinterpret_cast<unsigned long long*>(ctx().tags_dev), red_seq);
      DECO_CHECK_LAUNCH();
      count_bytes(8.0 * nv * (nbk + 2));
      if (wp) {   // queued before the host decision; unused if the cycle ends here
        hipLaunchKernelGGL(k_arn_update, dim3(nbk), dim3(kArnBlock), 0, stream(), n, j, (const double*)V.p, uu, wp,
                           (const double*)redd.p, V.p + (size_t)j * n, un);
        DECO_CHECK_LAUNCH();
        count_bytes(8.0 * n * (j + 4) + 8.0 * (2 * j + 2));
      }
    };
    // finalise column j-1 with u_j = Q s + rho v_j and rotate it (rgmres.tpp:80-93); true when
    // the cycle ends here
    auto finalize = [&](double rho, bool last) -> bool {
      for (int i = 0; i < j; ++i) Hr[i + (size_t)(j - 1) * ld] += red[i];
      Hr[j + (size_t)(j - 1) * ld] = rho;
      double* col = &Rm[(size_t)(j - 1) * ld];
      for (int i = 0; i <= j; ++i) col[i] = Hr[i + (size_t)(j - 1) * ld];
      for (int i = 0; i <= j - 2; ++i) apply_givens(cs[i], sn[i], col[i], col[i + 1]);
      comp_givens(col[j - 1], col[j], cs[j - 1], sn[j - 1]);
      apply_givens(cs[j - 1], sn[j - 1], col[j - 1], col[j]);
      g[j] = 0.0;
      apply_givens(cs[j - 1], sn[j - 1], g[j - 1], g[j]);
      ++iter;
      k = j;
      rnorm = std::fabs(g[j]);
      rn3 = rn2;
      rn2 = rn1;
      rn1 = rnorm;
      if (rnorm < absTol) { flag = 0; return true; }
      return last || !(rho > 0.0);
    };
    auto rho_of = [&](bool with_w, double& gam) {
      double rho;
      gam = 0.0;
      if (with_w) arn_scalars(j, red, rho, gam);
      else {
        double ss = 0.0;
        for (int i = 0; i < j; ++i) ss += red[i] * red[i];
        const double d = red[j] - ss;
        rho = std::sqrt(d > 0.0 ? d : 0.0);
      }
      return rho;
    };
    while (true) {
      const bool last = (j == m) || (it0 + j >= itmax);
      double* un = bufs[j & 1];
      double rho, gam;
      // When the last two estimates extrapolate below the tolerance at this step, column j-1 is
      // finalised from Q^T u_j and |u_j|^2 alone first, and the operator is applied to u_j only
      // if the solve goes on (the values of s and |u_j|^2 do not depend on w).
      bool finalized = false;
      if (!last && !have_w && j > 0 && rn1 > 0.0 && rn2 > 0.0 && rn1 * (rn1 / rn2) < absTol) {
        reduce_step(u, nullptr, un);
        wait_tags(red_nv, red_seq, red);
        rho = rho_of(false, gam);
        if (finalize(rho, false)) break;
        finalized = true;
      }
      const double* wp = nullptr;
      if (!last) {
        if (!have_w) apply_op(j, u);
        wp = w.p;
      }
      reduce_step(u, wp, un);
      // Queue step j+1's operator application on u_{j+1} before the host reads step j, so the
      // device works while the host rotates the Hessenberg column; if the solve stops at step j,
      // or at step j+1 (which finalises column j from Q^T u_{j+1} and |u_{j+1}|^2 alone), this
      // application is not used (stream order keeps every later use of w, z after it). Skipped
      // where the estimates extrapolate below the tolerance at column j, two columns past the
      // last finalised one, at the larger of the last two convergence factors (a wrong skip only
      // serialises step j+1's application behind the host; the iterates are unchanged).
      const bool next_last = (j + 1 == m) || (it0 + j + 1 >= itmax);
      bool likely_stop = false;
      if (rn1 > 0.0 && rn2 > 0.0) {
        const double q = rn3 > 0.0 ? std::max(rn1 / rn2, rn2 / rn3) : rn1 / rn2;
        likely_stop = rn1 * q * q < absTol;
      }
      have_w = wp && !next_last && !likely_stop;
      if (have_w) apply_op(j + 1, un);
      wait_tags(red_nv, red_seq, red);
      rho = rho_of(wp != nullptr, gam);
      if (j == 0) {
        if (bnorm < 0.0) {   // |b| from step 0 of a solve from x = 0
          bnorm = rho;
          absTol = tol * bnorm;
        }
        g[0] = rho;
        rnorm = std::fabs(g[0]);
        rn1 = rnorm;
        if (rnorm < absTol) { flag = 0; break; }
        if (last || !(rho > 0.0)) break;
      } else if (!finalized) {
        if (finalize(rho, last)) break;
      }
      // u_j = Q s + rho v_j: column j of T, provisional Hessenberg column j
      if (flex) {
        for (int i = 0; i < j; ++i) T[i + (size_t)j * ld] = red[i];
        T[j + (size_t)j * ld] = rho;
      }
      for (int i = 0; i <= j; ++i) {
        double t = 0.0;
        for (int q = 0; q < j; ++q) t += Hr[i + (size_t)q * ld] * red[q];
        hs[i] = t;
      }
      double sz = 0.0;
      for (int i = 0; i < j; ++i) sz += red[i] * red[j + 1 + i];
      for (int i = 0; i < j; ++i) Hr[i + (size_t)j * ld] = (red[j + 1 + i] - hs[i]) / rho;
      Hr[j + (size_t)j * ld] = ((red[2 * j + 1] - sz) / rho - hs[j]) / rho;
      u = un;
      ++j;
    }
    if (k > 0) {
      for (int i = k - 1; i >= 0; --i) {                      // rgmres.tpp:95-104
        const double d = Rm[i + (size_t)i * ld];
        if (d == 0.0) fail(DECO_ERR_BREAKDOWN, "Zero diagonal element in upper triangular matrix");
        g[i] /= d;
        for (int q = 0; q < i; ++q) g[q] -= Rm[q + (size_t)i * ld] * g[i];
      }
      if (flex) {
        for (int i = k - 1; i >= 0; --i) {
          g[i] /= T[i + (size_t)i * ld];
          for (int q = 0; q < i; ++q) g[q] -= T[q + (size_t)i * ld] * g[i];
        }
      }
      // x += (combination), or x = (combination) on the first cycle from x = 0
      double* wout = (first_zero && flex) ? x : w.p;
      if (flex32) combine(n, k, (const float*)Z, (int64_t)n, g.data(), gdev.p, wout);
      else if (flex64) combine(n, k, (const double*)Z64, (int64_t)n, g.data(), gdev.p, wout);
      else combine(n, k, (const double*)V.p, (int64_t)n, g.data(), gdev.p, w.p);
      if (left) S.add(w.p, x);
      else if (!flex) {
        if (first_zero) { M.apply(w.p, x); ++napp; }
        else { M.apply(w.p, z.p); ++napp; S.add(z.p, x); }
      } else if (!first_zero) S.add(w.p, x);
    } else if (first_zero) {
      zero_async(x, sizeof(double) * n);   // no direction: x stays 0
    }
    if (!left) { A.residual(x, b, r.p); ++nspmv; }
    else { A.residual(x, b, z.p); ++nspmv; M.apply(z.p, r.p); ++napp; }
    rnorm = S.nrm2(r.p);
    if (rnorm < absTol) { flag = 0; break; }
    if (k == 0) break;   // zero residual norm with a zero tolerance: no direction to extend
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
static void bicgstab(bool left, const deco_sol_params& sp, const AOp& A, const MOp& M, const double* b, double* x,
                     deco_sol_info& info) {
  const int n = A.transpose ? A.A->m : A.A->n;
  Solver S(n);
  WsRef r = ws(0, n), rt = ws(1, n), p = ws(2, n), y = ws(7, n), v = ws(8, n), t = ws(9, n);
  double rho_old, rho = 1, alpha = 1, beta, omega = 1;
  const double reltol = sp.reltol;
  const int itmax = sp.itmax;
  float it = 0.0f;
  int flag = 1, napp = 0, nspmv = 0;
  double norm_b, tol, res_norm;
  if (!sp.use_x0) zero_async(x, sizeof(double) * n);
  if (!left) {
    norm_b = S.nrm2(b);
    tol = norm_b * reltol;
    if (!sp.use_x0) copy_d2d(r.p, b, n);
    else { A.residual(x, b, r.p); ++nspmv; }
  } else {
    if (!sp.use_x0) { M.apply(b, r.p); ++napp; }
    else { A.residual(x, b, y.p); ++nspmv; M.apply(y.p, r.p); ++napp; }
  }
  copy_d2d(rt.p, r.p, n);
  res_norm = S.nrm2(r.p);
  if (left) { norm_b = res_norm; tol = norm_b * reltol; }
  double next_rho = 0.0;
  bool have_rho = false;
  const int nb = S.nb;
  double sums[2];
  for (it = 0.5f; it < itmax; it += 0.5f) {
    rho_old = rho;
    rho = have_rho ? next_rho : S.dot(r.p, rt.p);           // bicgstab.tpp:248
    if (it < 1) copy_d2d(p.p, r.p, n);
    else {
      beta = (std::fabs(res_norm) < 1e-40) ? 0.0 : (rho / rho_old) * (alpha / omega);
      hipLaunchKernelGGL(k_bicg_p, dim3(S.g256()), dim3(256), 0, stream(), n, beta, -beta * omega,
                         (const double*)r.p, (const double*)v.p, p.p);
      DECO_CHECK_LAUNCH();
      count_bytes(32.0 * n);
    }
    if (!left) { M.apply(p.p, y.p); A.apply(y.p, v.p); }
    else { A.apply(p.p, y.p); M.apply(y.p, v.p); }
    ++napp; ++nspmv;
    alpha = (std::fabs(res_norm) < 1e-40) ? 0.0 : rho / S.dot(v.p, rt.p);
    if (!left)
      hipLaunchKernelGGL((k_bicg_update<true, false>), dim3(nb), dim3(kRedBlock), 0, stream(), n, alpha,
                         (const double*)v.p, r.p, alpha, (const double*)y.p, x, (const double*)nullptr, S.rs.part.p);
    else
      hipLaunchKernelGGL((k_bicg_update<false, false>), dim3(nb), dim3(kRedBlock), 0, stream(), n, alpha,
                         (const double*)v.p, r.p, 0.0, (const double*)nullptr, (double*)nullptr,
                         (const double*)nullptr, S.rs.part.p);
    DECO_CHECK_LAUNCH();
    count_bytes((left ? 16.0 : 48.0) * n + 8.0 * nb);
    S.host_sums(1, sums);
    res_norm = std::sqrt(sums[0]);
    if (res_norm <= tol) {
      if (left) S.axpby(alpha, y.p, 1.0, x);                  // lbicgstab.tpp:193 (y = A p)
      flag = 0;
      break;
    }
    it += 0.5f;
    if (!left) { M.apply(r.p, y.p); A.apply(y.p, t.p); }
    else { A.apply(r.p, y.p); M.apply(y.p, t.p); }
    ++napp; ++nspmv;
    if (std::fabs(res_norm) < 1e-40) omega = 0.0;
    else {
      S.dot2(t.p, r.p, t.p, t.p, sums);
      omega = sums[0] / sums[1];
    }
    if (!left)
      hipLaunchKernelGGL((k_bicg_update<true, true>), dim3(nb), dim3(kRedBlock), 0, stream(), n, omega,
                         (const double*)t.p, r.p, omega, (const double*)y.p, x, (const double*)rt.p, S.rs.part.p);
    else
      hipLaunchKernelGGL((k_bicg_update<true, true>), dim3(nb), dim3(kRedBlock), 0, stream(), n, omega,
                         (const double*)t.p, r.p, alpha * omega, (const double*)y.p, x, (const double*)rt.p,
                         S.rs.part.p);
    DECO_CHECK_LAUNCH();
    count_bytes(56.0 * n + 16.0 * nb);
    S.host_sums(2, sums);
    res_norm = std::sqrt(sums[0]);
    next_rho = sums[1];
    have_rho = true;
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
static void pcg(const deco_sol_params& sp, const AOp& A, const MOp& M, const double* b, double* x,
                deco_sol_info& info) {
  const int n = A.transpose ? A.A->m : A.A->n;
  Solver S(n);
  WsRef r = ws(0, n), p = ws(2, n), z = ws(1, n), g = ws(7, n);
  const double norm_b = S.nrm2(b), tol = norm_b * sp.reltol;
  int it = 0, napp = 0, nspmv = 0;
  if (!sp.use_x0) { zero_async(x, sizeof(double) * n); copy_d2d(r.p, b, n); }
  else { A.residual(x, b, r.p); ++nspmv; }
  M.apply(r.p, p.p); ++napp;
  double rho1 = S.dot(r.p, p.p);
  double res_norm = S.nrm2(r.p);
  while (res_norm > tol && it < sp.itmax) {
    A.apply(p.p, z.p); ++nspmv;
    const double alpha = rho1 / S.dot(z.p, p.p);
    S.axpby(alpha, p.p, 1.0, x);
    S.axpby(-alpha, z.p, 1.0, r.p);
    M.apply(r.p, g.p); ++napp;
    const double rho2 = S.dot(r.p, g.p);
    const double beta = rho2 / rho1;
    rho1 = rho2;
    S.axpby(1.0, g.p, beta, p.p);
    res_norm = S.nrm2(r.p);
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

static void release_workspace() {
  for (int k = 0; k < kWsSlots; ++k) g_ws[k].release();
}

static void run_solver(const deco_sol_params& sp, const AOp& A, const MOp& M, const double* b, double* x,
                       deco_sol_info& info) {
  Event ev0, ev1;
  hipEvent_t e0 = ev0.e, e1 = ev1.e;
  DECO_HIP(hipEventRecord(e0, stream()));
  switch (sp.type) {
    case DECO_SOL_RGMRES:
    case DECO_SOL_LGMRES: {
      const bool left = sp.type == DECO_SOL_LGMRES;
      if (sp.ortho == DECO_ORTHO_MGS || sp.restart > kDcgs2MaxRestart) gmres(left, sp, A, M, b, x, info);
      else if (sp.ortho == DECO_ORTHO_DCGS2) gmres_dcgs2(left, sp, A, M, b, x, info);
      else fail(DECO_ERR_ARG, "unknown Gram-Schmidt variant %d", sp.ortho);
      break;
    }
    case DECO_SOL_BICGSTAB: bicgstab(false, sp, A, M, b, x, info); break;
    case DECO_SOL_LBICGSTAB: bicgstab(true, sp, A, M, b, x, info); break;
    case DECO_SOL_PCG: pcg(sp, A, M, b, x, info); break;
    default: fail(DECO_ERR_ARG, "Invalid solver type: %d", sp.type);
  }
  DECO_HIP(hipEventRecord(e1, stream()));
  DECO_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  DECO_HIP(hipEventElapsedTime(&ms, e0, e1));
  info.time_ms = ms;
  if (sp.verbosity > 0) {
    printf("\n--------------------------------\n");
    printf("Profile of %s:\n", sp.type == DECO_SOL_RGMRES ? "rGMRES" : sp.type == DECO_SOL_LGMRES ? "lGMRES"
                                : sp.type == DECO_SOL_PCG ? "PCG" : "BICGstab");
    printf("iteration count =  %d\n", info.iter);
    printf("residual = %.2e\n", info.res_rel);
  }
}

// ------------------------------------------------------------------ Jacobi -------------
// per-cell LU with partial pivoting (dgetf2) and inverse by unit solves (dgetrf + dgetri,
// Jacobi.tpp:64-65); the block is read as in Jacobi.tpp:311-323
constexpr int kMaxBs = 8;
__global__ void k_jacobi_setup(int sn, int bs, const int* __restrict__ r, const int* __restrict__ c,
                               const double* __restrict__ v, double* __restrict__ J) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sn) return;
  double B[kMaxBs * kMaxBs];  // column-major
  int piv[kMaxBs];
  for (int q = 0; q < bs * bs; ++q) B[q] = 0.0;
  for (int p = 0; p < bs; ++p) {
    const int i = s * bs + p;
    const int ia = r[i], ib = r[i + 1];
    int mn = ia, mx = ib, mid = (ia + ib) >> 1;  // core::binarysearch
    while (mx - mn > 1) { if (c[mid] > i) mx = mid; else mn = mid; mid = (mn + mx) >> 1; }
    for (int k = 0; k < bs; ++k) {
      const int jj = mid + k - p;
      if (jj < ib && jj >= ia && c[jj] / bs == s) B[p + (c[jj] % bs) * bs] = v[jj];
    }
  }
  for (int j = 0; j < bs; ++j) {
    int pv = j;
    double mxv = fabs(B[j + j * bs]);
    for (int i = j + 1; i < bs; ++i) { double t = fabs(B[i + j * bs]); if (t > mxv) { mxv = t; pv = i; } }
    piv[j] = pv;
    if (B[pv + j * bs] != 0.0) {
      if (pv != j) for (int q = 0; q < bs; ++q) { double t = B[j + q * bs]; B[j + q * bs] = B[pv + q * bs]; B[pv + q * bs] = t; }
      const double pivv = B[j + j * bs];
      if (fabs(pivv) >= 2.2250738585072014e-308) { const double rp = 1.0 / pivv; for (int i = j + 1; i < bs; ++i) B[i + j * bs] *= rp; }
      else for (int i = j + 1; i < bs; ++i) B[i + j * bs] /= pivv;
    }
    for (int q = j + 1; q < bs; ++q) {
      const double t = B[j + q * bs];
      if (t != 0.0) for (int i = j + 1; i < bs; ++i) B[i + q * bs] -= B[i + j * bs] * t;
    }
  }
  for (int k = 0; k < bs; ++k) {
    double col[kMaxBs];
    for (int q = 0; q < bs; ++q) col[q] = (q == k) ? 1.0 : 0.0;
    for (int j = 0; j < bs; ++j) if (piv[j] != j) { double t = col[j]; col[j] = col[piv[j]]; col[piv[j]] = t; }
    for (int j = 0; j < bs; ++j) { const double t = col[j]; if (t != 0.0) for (int i = j + 1; i < bs; ++i) col[i] -= t * B[i + j * bs]; }
    for (int j = bs - 1; j >= 0; --j) {
      if (col[j] != 0.0) { col[j] /= B[j + j * bs]; const double t = col[j]; for (int i = 0; i < j; ++i) col[i] -= t * B[i + j * bs]; }
    }
    for (int p = 0; p < bs; ++p) J[(size_t)s * bs * bs + p * bs + k] = col[p];
  }
}

__global__ void k_jacobi_point_setup(int n, const int* __restrict__ r, const int* __restrict__ c,
                                     const double* __restrict__ v, double* __restrict__ J) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int mn = r[i], mx = r[i + 1], mid = (mn + mx) >> 1;
  while (mx - mn > 1) { if (c[mid] > i) mx = mid; else mn = mid; mid = (mn + mx) >> 1; }
  J[i] = 1 / v[mid];
}

// b = J x with J in T; vectors in IO (IO = T, or FP64 vectors around FP32 inverses)
template <class T, class IO, bool TR>
__global__ void k_jacobi_apply(int n, int bs, const T* __restrict__ J, const IO* __restrict__ x, IO* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (bs == 1) { b[i] = (IO)(J[i] * (T)x[i]); return; }
  const int s = i / bs, p = i - s * bs;
  const T* Ap = J + (size_t)s * bs * bs;
  T sum = (TR ? Ap[p] : Ap[p * bs]) * (T)x[s * bs];
  for (int k = 1; k < bs; ++k) sum += (TR ? Ap[k * bs + p] : Ap[p * bs + k]) * (T)x[s * bs + k];
  b[i] = (IO)sum;
}

Jacobi::Jacobi(const deco_csr* A, int bs_) {
  kind = JACOBI;
  n = A->n;
  bs = bs_;
  if (bs < 1 || bs > kMaxBs || n % bs != 0) fail(DECO_ERR_ARG, "Jacobi block size %d invalid for n=%d", bs, n);
  if (bs == 1) {
    J.alloc(n);
    hipLaunchKernelGGL(k_jacobi_point_setup, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, A->r.p, A->c.p,
                       A->v.p, J.p);
  } else {
    const int sn = n / bs;
    J.alloc((size_t)sn * bs * bs);
    hipLaunchKernelGGL(k_jacobi_setup, dim3(grid_for(sn, 128)), dim3(128), 0, stream(), sn, bs, A->r.p, A->c.p,
                       A->v.p, J.p);
  }
  DECO_CHECK_LAUNCH();
  Jf.alloc(J.n);
  convert(J.p, Jf.p, J.n);
}
void Jacobi::apply(const double* r, double* x) {
  if (precision == DECO_PREC_FP32)
    hipLaunchKernelGGL((k_jacobi_apply<float, double, false>), dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, bs,
                       (const float*)Jf.p, r, x);
  else
    hipLaunchKernelGGL((k_jacobi_apply<double, double, false>), dim3(grid_for(n, 256)), dim3(256), 0, stream(), n,
                       bs, (const double*)J.p, r, x);
  DECO_CHECK_LAUNCH();
  count_bytes((double)n * (bs * (precision == DECO_PREC_FP32 ? 4 : 8) + 16));
}
void Jacobi::apply_t(const double* r, double* x) {
  if (precision == DECO_PREC_FP32)
    hipLaunchKernelGGL((k_jacobi_apply<float, double, true>), dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, bs,
                       (const float*)Jf.p, r, x);
  else
    hipLaunchKernelGGL((k_jacobi_apply<double, double, true>), dim3(grid_for(n, 256)), dim3(256), 0, stream(), n,
                       bs, (const double*)J.p, r, x);
  DECO_CHECK_LAUNCH();
  count_bytes((double)n * (bs * (precision == DECO_PREC_FP32 ? 4 : 8) + 16));
}
void Jacobi::apply32(const float* r, float* x) {
  hipLaunchKernelGGL((k_jacobi_apply<float, float, false>), dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, bs,
                     (const float*)Jf.p, r, x);
  DECO_CHECK_LAUNCH();
  count_bytes((double)n * (bs * 4 + 8));
}
void Jacobi::apply32_t(const float* r, float* x) {
  hipLaunchKernelGGL((k_jacobi_apply<float, float, true>), dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, bs,
                     (const float*)Jf.p, r, x);
  DECO_CHECK_LAUNCH();
  count_bytes((double)n * (bs * 4 + 8));
}

// ------------------------------------------------------------------ adjoint helpers ----
__global__ void k_grad_A(int n, const int* __restrict__ r, const int* __restrict__ c, const double* __restrict__ lam,
                         const double* __restrict__ x, double* __restrict__ dA) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double li = -lam[i];
  for (int j = r[i]; j < r[i + 1]; ++j) dA[j] = li * x[c[j]];
}

}  // namespace deco

using namespace deco;

extern "C" {

int deco_identity_create(int n, deco_prec** out) {
  return guard([&] {
    if (n < 0 || !out) fail(DECO_ERR_ARG, "bad argument");
    *out = new Identity(n);
  });
}

int deco_jacobi_create(const deco_cpr_params* p, const deco_csr* A, deco_prec** out) {
  return guard([&] {
    if (!p || !A || !out) fail(DECO_ERR_ARG, "null argument");
    *out = new Jacobi(A, p->glo_block_size);
    sync();
  });
}

int deco_prec_destroy(deco_prec* M) {
  return guard([&] { delete M; });
}

int deco_prec_apply(deco_prec* M, const double* r, double* x) {
  return guard([&] {
    if (!M) fail(DECO_ERR_ARG, "null preconditioner");
    need(r, M->n, "r");
    need(x, M->n, "x");
    M->apply(r, x);
    sync();
  });
}

int deco_prec_apply_transpose(deco_prec* M, const double* r, double* x) {
  return guard([&] {
    if (!M) fail(DECO_ERR_ARG, "null preconditioner");
    need(r, M->n, "r");
    need(x, M->n, "x");
    M->apply_t(r, x);
    sync();
  });
}

int deco_prec_size(const deco_prec* M, int* n) {
  return guard([&] {
    if (!M || !n) fail(DECO_ERR_ARG, "null argument");
    *n = M->n;
  });
}

int deco_prec_set_precision(deco_prec* M, int precision) {
  return guard([&] {
    if (!M) fail(DECO_ERR_ARG, "null preconditioner");
    if (precision != DECO_PREC_FP64 && precision != DECO_PREC_FP32) fail(DECO_ERR_ARG, "bad precision %d", precision);
    M->set_precision(precision);
    sync();
  });
}

int deco_prec_set_layout(deco_prec* M, int layout) {
  return guard([&] {
    if (!M) fail(DECO_ERR_ARG, "null preconditioner");
    if (layout != DECO_LAYOUT_BLOCK && layout != DECO_LAYOUT_CSR) fail(DECO_ERR_ARG, "bad layout %d", layout);
    M->set_layout(layout);
    sync();
  });
}

int deco_time_outer_products(int on) {
  return guard([&] {
    OuterTiming& t = outer_timing();
    t.on = on != 0;
    t.used = 0;
  });
}

int deco_outer_product_times(double* total_ms, long long* count) {
  return guard([&] {
    OuterTiming& t = outer_timing();
    double tot = 0.0;
    if (t.used >= 2) {
      DECO_HIP(hipEventSynchronize(t.ev[t.used - 1]));
      for (size_t k = 0; k + 1 < t.used; k += 2) {
        float ms = 0.f;
        DECO_HIP(hipEventElapsedTime(&ms, t.ev[k], t.ev[k + 1]));
        tot += ms;
      }
    }
    if (total_ms) *total_ms = tot;
    if (count) *count = (long long)(t.used / 2);
  });
}

int deco_workspace_release(void) {
  return guard([&] {
    sync();
    release_workspace();
  });
}

int deco_solve(const deco_sol_params* sp, const deco_csr* A, deco_prec* M, const double* b, double* x,
               deco_sol_info* info) {
  return guard([&] {
    if (!sp || !A || !b || !x) fail(DECO_ERR_ARG, "null argument");
    if (A->n != A->m) fail(DECO_ERR_ARG, "matrix is not square");
    if (M && M->n != A->n) fail(DECO_ERR_ARG, "preconditioner size %d != %d", M->n, A->n);
    deco_sol_info inf{};
    run_solver(*sp, AOp{A, false}, MOp{M, false, A->n}, b, x, inf);
    if (info) *info = inf;
  });
}

int deco_solve_transpose(const deco_sol_params* sp, deco_csr* A, deco_prec* M, const double* g, double* y,
                         deco_sol_info* info) {
  return guard([&] {
    if (!sp || !A || !g || !y) fail(DECO_ERR_ARG, "null argument");
    if (A->n != A->m) fail(DECO_ERR_ARG, "matrix is not square");
    if (!A->At) A->At = csr_transpose(A);
    deco_sol_info inf{};
    run_solver(*sp, AOp{A, true}, MOp{M, true, A->n}, g, y, inf);
    if (info) *info = inf;
  });
}

int deco_grad_A_pattern(const deco_csr* A, const double* lam, const double* x, double* dA) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    need(lam, A->n, "lambda");
    need(x, A->m, "x");
    need(dA, A->nnz, "dA");
    if (A->n == 0) return;
    hipLaunchKernelGGL(k_grad_A, dim3(grid_for(A->n, 256)), dim3(256), 0, stream(), A->n, A->r.p, A->c.p, lam, x, dA);
    DECO_CHECK_LAUNCH();
    count_bytes((double)A->nnz * 12 + 4.0 * (A->n + 1) + 16.0 * A->n);
    sync();
  });
}

}  // extern "C"
