// Launch interface between the host engine (dgx_api.cpp) and the templated
// kernels (dgx_kernels2d.hip, dgx_kernels3d.hip, dgx_stage3d.hip, dgx_mixed3d.hip).
#pragma once
#include <string>

#include "dgx_device.hpp"

namespace dgx {

enum Op {
  OP_TRACES = 0,    // modal U -> face traces T
  OP_RHS = 1,       // U, T [, V, TV] -> R = -dU/dt
  OP_GRAD = 2,      // U, T -> lifted gradients G          (Navier-Stokes)
  OP_VISC = 3,      // U, G -> viscous volume V, traces TV  (Navier-Stokes)
  OP_FORWARD = 4,   // nodal X -> modal Y
  OP_BACKWARD = 5   // modal X -> nodal Y
};

struct LaunchArgs {
  int op = 0, p = 0, ne = 0, nfield = 0;
  bool ns = false;
  const double *U = nullptr;
  double *T = nullptr, *R = nullptr, *G = nullptr, *V = nullptr, *TV = nullptr;
  const double *X = nullptr;
  double *Y = nullptr;
  int *flag = nullptr;
  Tab tab;
  Geo geo;
  Phys phys;
  hipStream_t stream = nullptr;
};

// Fill each translation unit's constant-memory basis tables (kTabTotal doubles).
hipError_t upload_tables_2d(const double *h);
hipError_t upload_tables_3d(const double *h);
hipError_t upload_tables_s3d(const double *h);
hipError_t upload_tables_m3d(const double *h);

hipError_t launch_2d(const LaunchArgs &a);
hipError_t launch_3d(const LaunchArgs &a);
int lds_bytes2(int p, int which);
int lds_bytes3(int p, int which);

// Fused hex stage kernels (dgx_stage3d.hip), p <= 9.
struct StageLaunch {
  int op = 1;                    // 0: BR1 viscous terms (k_vterm3), 1: stage (k_stage3)
  int p = 0, ne = 0;
  bool ns = false, update = false;
  const double *T = nullptr, *V = nullptr, *TV = nullptr;
  double *U = nullptr, *Ut = nullptr, *R = nullptr, *Tn = nullptr, *Vw = nullptr, *TVw = nullptr;
  const float *Vf = nullptr, *TVf = nullptr;   // mixed precision: FP32 viscous terms
  float *Vfw = nullptr, *TVfw = nullptr;
  double ca = 0.0, cb = 0.0, dt = 0.0;
  int *flag = nullptr;
  Tab tab;
  Geo geo;
  Phys phys;
  hipStream_t stream = nullptr;
};
bool stage3_supported(int p);
hipError_t launch_stage3(const StageLaunch &s);

// Mixed-precision fused hex stage kernels (dgx_mixed3d.hip), p <= 8.
bool mixed3_supported(int p);
hipError_t launch_mixed3(const StageLaunch &s);

// Error message of the calling thread's last failure (dgx_api.cpp).
int set_error(int code, const std::string &msg);

// Flat-array stage updates (dgx_blas.hip)
hipError_t blas_axpy(double a, const double *x, const double *y, double *z, size_t n, hipStream_t s);
hipError_t blas_axpby(double a, const double *x, double b, const double *y, double *z, size_t n,
                      hipStream_t s);
hipError_t blas_scale(double a, double *x, size_t n, hipStream_t s);

}  // namespace dgx
