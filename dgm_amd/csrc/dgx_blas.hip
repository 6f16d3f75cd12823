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

// Channel::holdBulkMassFlow (src/Channel.hpp:117-163): subtract from the
// density, x-momentum and energy residuals their domain means.  On affine
// elements with the normalized Legendre basis a constant c is the mode-0
// coefficient c / phi_0^d, and the element integral of a field is
// J_e u_0 / phi_0^d, so the mean is removed by subtracting the J-weighted mean
// of the mode-0 coefficients from every element's mode 0.  One workgroup; the
// reduction order is fixed, so the result does not depend on the launch.
constexpr int kBulkThreads = 1024;
__global__ void __launch_bounds__(kBulkThreads) k_bulk_hold(double *__restrict__ R, const double *__restrict__ J,
                                                            int ne, int nm, int nf) {
  __shared__ double red[4][kBulkThreads];
  const int tid = threadIdx.x;
  const int fld[3] = {0, 1, nf - 1};
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int e = tid; e < ne; e += kBulkThreads) {
    const double j = J[e];
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = fma(j, R[((size_t)fld[k] * ne + e) * nm], s[k]);
    s[3] += j;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) red[k][tid] = s[k];
  __syncthreads();
  for (int w = kBulkThreads / 2; w > 0; w >>= 1) {
    if (tid < w)
#pragma unroll
      for (int k = 0; k < 4; ++k) red[k][tid] += red[k][tid + w];
    __syncthreads();
  }
  const double inv = 1.0 / red[3][0];
  const double m[3] = {red[0][0] * inv, red[1][0] * inv, red[2][0] * inv};
  for (int e = tid; e < ne; e += kBulkThreads)
#pragma unroll
    for (int k = 0; k < 3; ++k) R[((size_t)fld[k] * ne + e) * nm] -= m[k];
}

hipError_t bulk_hold(double *R, const double *J, int ne, int nm, int nf, hipStream_t s) {
  hipLaunchKernelGGL(k_bulk_hold, dim3(1), dim3(kBulkThreads), 0, s, R, J, ne, nm, nf);
  return hipGetLastError();
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
