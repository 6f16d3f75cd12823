// Streaming stage updates on flat FP64 arrays: the vField BLAS-1 operations
// the reference's integrators use (src/vField.hpp:1002-1031), HBM-bound.
#include "dgx_kernels.hpp"

namespace dgx {

// z = a*x + y
__global__ void __launch_bounds__(256) k_axpy(double a, const double *__restrict__ x,
                                              const double *__restrict__ y,
                                              double *__restrict__ z, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    z[i] = fma(a, x[i], y[i]);
}

// z = a*x + b*y, as fma(a, x, b*y) (the batched serving kernels' fused update rounds alike)
__global__ void __launch_bounds__(256) k_axpby(double a, const double *__restrict__ x, double b,
                                               const double *__restrict__ y,
                                               double *__restrict__ z, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    z[i] = fma(a, x[i], b * y[i]);
}

__global__ void __launch_bounds__(256) k_scale(double a, double *__restrict__ x, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = a * x[i];
}

static unsigned grid_for(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  if (g == 0) g = 1;
  return (unsigned)g;
}

hipError_t blas_axpy(double a, const double *x, const double *y, double *z, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, s, a, x, y, z, n);
  return hipGetLastError();
}
hipError_t blas_axpby(double a, const double *x, double b, const double *y, double *z, size_t n,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, s, a, x, b, y, z, n);
  return hipGetLastError();
}
hipError_t blas_scale(double a, double *x, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, s, a, x, n);
  return hipGetLastError();
}

}  // namespace dgx
