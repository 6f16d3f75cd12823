// CFL time step of the 2D Euler, Navier-Stokes and shallow-water domains
// (Euler::set_time_step, src/Euler.cpp:804-877; Shallow::set_time_step,
// src/Shallow.cpp:573-637) with the stability metrics of tensor-product quads
// (Element::compute_metrics_2d_tensor_prod, src/Element.cpp:485-555).
//
// At every GLL node of every element, from the nodal state,
//   idt = (|u dxr + v dyr| + |u dxs + v dys| + c acoustic) ivol,
// with dxr = x_r dr_i, dyr = y_r dr_i, dxs = x_s ds_j, dys = y_s ds_j (dr_i,
// ds_j the GLL node spacings: one-sided at the ends, centred inside),
// ivol = 1 / (dr_i ds_j J), acoustic = sqrt((dxr + dyr)^2 + (dxs + dys)^2),
// and c the sound speed sqrt(gamma (gamma-1) (E - 1/2 (rho u u + rho v v)) / rho)
// (Euler) or sqrt(phi) (shallow water).  The CFL number of dt is
// dt * max idt over the domain.
//
// k_cfl2 evaluates the nodal state from the modes in LDS (one wavefront per
// element) and writes each element's largest idt to a scratch array of ne
// values; k_max1 reduces that array in one workgroup, in a fixed order.
#include <cmath>
#include <mutex>
#include <string>

#include "dgx_domain.hpp"
#include "dgx_kernels.hpp"

namespace dgx {

__device__ constexpr double kNaN = __builtin_nan("");

template <int P>
__global__ void __launch_bounds__(64)
k_cfl2(CflArgs a) {
  constexpr int L = P + 1, Q = P + 2, L2 = L * L, Q2 = Q * Q, NT = 64;
  constexpr int NF = 4;   // at most 4 fields are read (shallow water: 3)
  __shared__ double A[NF * L2], X[NF * Q * L], O[NF * Q2], red[NT];
  const int e = a.elems ? a.elems[blockIdx.x] : blockIdx.x, tid = threadIdx.x, ne = a.ne;
  const int nf = a.shallow ? 3 : 4, NM = a.elems ? a.NM : L2;
  for (int i = tid; i < NF * L2; i += NT) {
    const int v = i / L2;
    A[i] = v < nf ? a.U[((size_t)v * ne + e) * NM + (i - v * L2)] : 0.0;
  }
  __syncthreads();
  // modal -> GLL nodes, all fields at once: [NF][L][L] -> [NF][Q][L] -> [NF][Q][Q]
  contract<NF, L, L, 1, Q, NT>(CT<P>::BT(), A, X, tid);
  __syncthreads();
  contract<NF, Q, L, 2, Q, NT>(CT<P>::BT(), X, O, tid);
  __syncthreads();
  const double g = a.gamma;
  double m = 0.0;
  bool bad = false;   // a NaN idt (non-realizable state) makes the element's value NaN
  for (int n = tid; n < Q2; n += NT) {
    const int i = n / Q, j = n - i * Q;
    const int S = a.curved ? Q2 : 1, o = a.curved ? n : 0;
    const double *cg = a.cgeo + (size_t)e * 5 * S + o;
    const double xr = cg[0], xs = cg[S], yr = cg[2 * S], ys = cg[3 * S], J = cg[4 * S];
    const double dr = a.dz[i], ds = a.dz[j];
    const double dxr = xr * dr, dxs = xs * ds, dyr = yr * dr, dys = ys * ds;
    const double ivol = 1.0 / (dr * ds * J);
    const double acoustic = sqrt((dxr + dyr) * (dxr + dyr) + (dxs + dys) * (dxs + dys));
    const double r = O[n], ru = O[Q2 + n], rv = O[2 * Q2 + n];
    const double u = ru / r, v = rv / r;
    double c;
    if (a.shallow) {
      c = sqrt(r);
    } else {
      c = u * ru;
      c = fma(v, rv, c);
      c = -0.5 * c + O[3 * Q2 + n];
      c = (g * (g - 1.0)) * c;
      c = sqrt(c / r);
    }
    double idt = fabs(fma(v, dyr, u * dxr));
    idt += fabs(fma(v, dys, u * dxs));
    idt = fma(c, acoustic, idt);
    idt = ivol * idt;
    bad |= !(idt == idt);
    m = fmax(m, idt);
  }
  red[tid] = bad ? kNaN : m;
  __syncthreads();
  if (tid == 0) {
    double b = 0.0;
    bool nan = false;
    for (int k = 0; k < NT; ++k) { nan |= !(red[k] == red[k]); b = fmax(b, red[k]); }
    a.part[e] = nan ? kNaN : b;
  }
}

// out[0] = max of part[0 .. n-1] (NaN if any is NaN); one workgroup.
constexpr int kMaxThreads = 1024;
__global__ void __launch_bounds__(kMaxThreads) k_max1(const double *__restrict__ part, int n, double *out) {
  __shared__ double red[kMaxThreads];
  __shared__ int bad[kMaxThreads];
  const int tid = threadIdx.x;
  double m = 0.0;
  int nan = 0;
  for (int i = tid; i < n; i += kMaxThreads) {
    const double x = part[i];
    nan |= !(x == x);
    m = fmax(m, x);
  }
  red[tid] = m;
  bad[tid] = nan;
  __syncthreads();
  for (int w = kMaxThreads / 2; w > 0; w >>= 1) {
    if (tid < w) {
      red[tid] = fmax(red[tid], red[tid + w]);
      bad[tid] |= bad[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) out[0] = bad[0] ? kNaN : red[0];
}

template <int P>
static hipError_t launch_cfl2_p(const CflArgs &a) {
  hipLaunchKernelGGL(k_cfl2<P>, dim3(a.elems ? a.nel : a.ne), dim3(64), 0, a.stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || a.elems) return e;
  hipLaunchKernelGGL(k_max1, dim3(1), dim3(kMaxThreads), 0, a.stream, a.part, a.ne, a.out);
  return hipGetLastError();
}

hipError_t launch_cfl_max1(const double *part, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_max1, dim3(1), dim3(kMaxThreads), 0, s, part, n, out);
  return hipGetLastError();
}

hipError_t launch_cfl2(const CflArgs &a) {
  if (a.ne <= 0 || !a.part || !a.out || !a.U || !a.cgeo) return hipErrorInvalidValue;
  if (a.elems && (a.nel < 1 || a.NM < (a.p + 1) * (a.p + 1))) return hipErrorInvalidValue;
  switch (a.p) {
#define DGX_CASE(PP) case PP: return launch_cfl2_p<PP>(a);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5)
    DGX_CASE(6) DGX_CASE(7) DGX_CASE(8) DGX_CASE(9) DGX_CASE(10)
#undef DGX_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t upload_tables_cfl(const double *h) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), h, sizeof(double) * kTabTotal);
}

// ---------------------------------------------------------------------------
// Allocation-free CFL evaluation (dgx_set_time_step_noalloc).  The kernels'
// scratch (ne per-element maxima and the domain maximum) is this file-scope
// device array instead of a buffer the library allocates; every GPU the
// library runs on holds its own copy until the process ends.  Its device
// address is looked up once per device, with that device current, and kept.
// Calls on one GPU share the array: calls that overlap on one GPU race.  The
// kernels write every value they later read, so nothing needs clearing.
constexpr size_t kCflScratch = size_t(1) << 19;   // doubles (4 MiB): ne + 1 <= 524288
static __device__ __align__(256) double scratch[kCflScratch];

namespace {
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_ptr(int dev, double **out) {
  if (dev < 0 || dev >= kMaxDevices) return set_error(DGX_ERR_ARGUMENT, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(DGX_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double *>(p);
  }
  *out = g_scratch[dev];
  return DGX_OK;
}
}  // namespace

}  // namespace dgx

extern "C" int dgx_set_time_step_noalloc(dgx_handle h, double dt, double CFLmax, double dtmax, double *cfl,
                                         double *dt_new) {
  using dgx::set_error;
  if (!h || !cfl || !dt_new) return set_error(DGX_ERR_ARGUMENT, "null argument");
  if (!(dt > 0) || !std::isfinite(dt) || !(CFLmax >= 0) || !(dtmax >= 0))
    return set_error(DGX_ERR_ARGUMENT, "dt must be positive and CFLmax, dtmax >= 0");
  if (h->nsd == 3) {   // Euler3d::set_time_step keeps dt and returns it
    *cfl = dt;
    *dt_new = dt;
    return DGX_OK;
  }
  if ((size_t)h->ne + 1 > dgx::kCflScratch)
    return set_error(DGX_ERR_ARGUMENT, "dgx_set_time_step_noalloc: " + std::to_string(h->ne) +
                                           " elements exceed its scratch of " +
                                           std::to_string(dgx::kCflScratch - 1) + "; use dgx_set_time_step");
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return set_error(DGX_ERR_DEVICE, "hipGetDevice failed");
  if (cur != h->device && hipSetDevice(h->device) != hipSuccess)
    return set_error(DGX_ERR_DEVICE, "hipSetDevice failed");
  double *ptr = nullptr;
  double m = 0.0;
  int rc = dgx::scratch_ptr(h->device, &ptr);
  if (!rc) rc = dgx::cfl_max(h, ptr, &m);
  if (cur != h->device) (void)hipSetDevice(cur);
  if (rc) return rc;
  return dgx::cfl_finish(dt, CFLmax, dtmax, m, cfl, dt_new);
}
