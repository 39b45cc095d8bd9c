// BLAS-1 entry points (alg::dot core/alg.hpp:278-307, alg::l2norm :334-352, alg::axpby :185-190).
#include "blas.hpp"
#include <cmath>
#include <mutex>

namespace deco {

__global__ void k_finish(const double* __restrict__ part, int count, double* __restrict__ out, int take_sqrt) {
  double s = reduce_partials(part, count);
  if (threadIdx.x == 0) out[0] = take_sqrt ? sqrt(s) : s;
}

__global__ void k_axpby(int n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = b * y[i] + a * x[i];
}

void dot_partials(int n, const double* x, const double* y, double* part) {
  hipLaunchKernelGGL(k_dot_part<double>, dim3(red_grid(n)), dim3(kRedBlock), 0, stream(), n, x, y, part);
  DECO_CHECK_LAUNCH();
}

void finish(const double* part, int count, double* dst, bool take_sqrt) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, stream(), part, count, dst, take_sqrt ? 1 : 0);
  DECO_CHECK_LAUNCH();
}

double dot_host(int n, const double* x, const double* y, RedScratch& rs) {
  dot_partials(n, x, y, rs.slot(0));
  finish(rs.slot(0), red_grid(n), rs.scal.p, false);
  double h = 0.0;
  copy_d2h(&h, rs.scal.p, 1);
  return h;
}

double nrm2_host(int n, const double* x, RedScratch& rs) {
  dot_partials(n, x, x, rs.slot(0));
  finish(rs.slot(0), red_grid(n), rs.scal.p, true);
  double h = 0.0;
  copy_d2h(&h, rs.scal.p, 1);
  return h;
}

// ------------------------------------------------------------------ allocation-free dot --
// deco_dot_noalloc keeps the block partials of its reduction and the reduced value in this
// file-scope device array instead of library buffers: red_grid() caps any length at kRedMaxGrid
// partials, so kNoAllocDoubles (1056 doubles, 8448 bytes) holds every call. Each GPU the library
// runs on holds its own copy until the process ends. In host code the name `scratch` is not a
// device address: each device's address comes from hipGetSymbolAddress (scratch_dev) and only
// that address goes to kernels. Nothing in it needs clearing: every partial a call reads is
// written by the same call. Calls that overlap on one GPU share the array and race.
constexpr int kNoAllocDoubles = kRedMaxGrid + 32;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kNoAllocDevices = 64;
static double* g_scratch_dev[kNoAllocDevices] = {};
static std::mutex g_scratch_mu;

// the library device's address of `scratch`, looked up once per device with that device current
static double* scratch_dev() {
  const int dev = ctx().device;
  if (dev < 0 || dev >= kNoAllocDevices)
    fail(DECO_ERR_DEVICE, "device %d is beyond the %d the allocation-free dot tracks", dev, kNoAllocDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    int cur = -1;
    DECO_HIP(hipGetDevice(&cur));
    if (cur != dev) DECO_HIP(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      fail(DECO_ERR_DEVICE, "hipGetSymbolAddress of the dot scratch failed: %s", hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  return g_scratch_dev[dev];
}

}  // namespace deco

using namespace deco;

extern "C" {

int deco_dot(int n, const double* x, const double* y, double* result) {
  return guard([&] {
    if (n < 0 || !result) fail(DECO_ERR_ARG, "bad argument");
    need(x, n, "x");
    need(y, n, "y");
    RedScratch rs;
    *result = dot_host(n, x, y, rs);
  });
}

int deco_dot_noalloc(int n, const double* x, const double* y, double* result) {
  return guard([&] {
    if (n < 0 || !result) fail(DECO_ERR_ARG, "bad argument");
    need(x, n, "x");
    need(y, n, "y");
    if (!ctx().stream) fail(DECO_ERR_STATE, "deco_init has not run (the allocation-free dot creates nothing)");
    const int nb = red_grid(n);
    if (nb > kNoAllocDoubles - 1)
      fail(DECO_ERR_ARG, "n = %d needs %d partials, more than the %d the scratch holds", n, nb, kNoAllocDoubles - 1);
    double* part = scratch_dev();
    double* value = part + (kNoAllocDoubles - 1);
    hipLaunchKernelGGL(k_dot_part<double>, dim3(nb), dim3(kRedBlock), 0, stream(), n, x, y, part);
    DECO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, stream(), (const double*)part, nb, value, 0);
    DECO_CHECK_LAUNCH();
    // the init-time pinned host buffer's last slot receives the value (no staging allocation)
    double* host = ctx().pinned + 4095;
    DECO_HIP(hipMemcpyAsync(host, value, sizeof(double), hipMemcpyDeviceToHost, stream()));
    DECO_HIP(hipStreamSynchronize(stream()));
    *result = *host;
  });
}

int deco_nrm2(int n, const double* x, double* result) {
  return guard([&] {
    if (n < 0 || !result) fail(DECO_ERR_ARG, "bad argument");
    need(x, n, "x");
    RedScratch rs;
    *result = nrm2_host(n, x, rs);
  });
}

int deco_axpby(int n, double a, const double* x, double b, double* y) {
  return guard([&] {
    if (n < 0) fail(DECO_ERR_ARG, "bad argument");
    need(x, n, "x");
    need(y, n, "y");
    if (n == 0) return;
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, a, x, b, y);
    DECO_CHECK_LAUNCH();
    sync();
  });
}

}  // extern "C"
