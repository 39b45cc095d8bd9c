// BLAS-1 entry points (alg::dot core/alg.hpp:278-307, alg::l2norm :334-352, alg::axpby :185-190).
#include "blas.hpp"
#include <cmath>

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
