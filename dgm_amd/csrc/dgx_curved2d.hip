// Curved quads (Euler 2D and shallow water, p <= 10): the residual with
// metrics at the nodes and the projection through the element mass matrix,
// the quad counterpart of dgx_curved3d.hip.
//
//   volume term   Euler::convection / Shallow::convection with Quad::gradient
//                 on curved geometry: dF_x/dx = sum_d (dxi_d/dx)(node) D_d F_x
//                 at every node (src/Euler.cpp:296-390, src/Shallow.cpp:276-336,
//                 src/Quad.cpp:1309-1363, 2309-2330)
//   edges         the Riemann flux with the master side's normal at each Gauss
//                 point (the reference evaluates a pair once, on its master
//                 side), lifted at the edge nodes:
//                 G2GL((Fhat - F-.n) J_side/J_vol) / w_end (src/Quad.cpp:1367-1452, 1971-2038)
//   project       b = B^T (wJ f) and M x = b with M = B^T diag(wJ) B
//                 (src/Quad.cpp:263-300), by conjugate gradients inside the
//                 workgroup, all fields at once.
#include "dgx_kernels.hpp"

namespace dgx {

template <int P, int NF> struct C2 {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1;
  static constexpr int L2 = L * L, Q2 = Q * Q;
  static constexpr int NT = 128, NW = NT / 64;
  static constexpr int NPT = (Q2 + NT - 1) / NT;            // nodes per thread
  static constexpr int NML = (NF * L2 + NT - 1) / NT;       // modal entries per thread
  // LDS (doubles): Un, Fw [NF][Q2]; scratch S1, S2 [SZ]; gF [NF][4][QF]; Fc [2][Q2]; reductions [NW][8]
  static constexpr int SZ = (NF * Q2 > 4 * NF * Q) ? NF * Q2 : 4 * NF * Q;
  static constexpr int lds = (2 * NF * Q2 + 2 * SZ + 4 * NF * QF + 2 * Q2 + 8 * NW) * 8;
};

// per-field sums over the workgroup
template <int NF, int NW>
__device__ __forceinline__ void field_dots2(const double (&part)[NF], double *red, double (&out)[NF], int tid) {
  double v[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double t = part[f];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    v[f] = t;
  }
  const int w = tid >> 6;
  if ((tid & 63) == 0) {
#pragma unroll
    for (int f = 0; f < NF; ++f) red[w * 8 + f] = v[f];
  }
  __syncthreads();
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) t += red[k * 8 + f];
    out[f] = t;
  }
  __syncthreads();
}

// PROJ: project the nodal fields a.U [NF][ne][Q^2] instead (the forward
// transform of a curved element: x = M^-1 B^T (wJ u)).
template <int P, bool SHALLOW, bool PROJ>
__global__ void __launch_bounds__(128)
k_rhs2c(CurvedArgs a) {
  constexpr int NF = SHALLOW ? 3 : 4;
  using S = C2<P, NF>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, Q2 = S::Q2;
  constexpr int NT = S::NT, NPT = S::NPT, NML = S::NML, NW = S::NW;
  extern __shared__ double sm[];
  double *Un = sm, *Fw = Un + NF * Q2, *S1 = Fw + NF * Q2, *S2 = S1 + S::SZ, *gF = S2 + S::SZ;
  double *Fc = gF + 4 * NF * QF, *red = Fc + 2 * Q2;
  static_assert(NF * L * Q <= S::SZ && NF * L2 <= S::SZ, "scratch too small");
  static_assert(S::lds <= 64 * 1024, "LDS");
  const int e = blockIdx.x, tid = threadIdx.x, ne = a.ne;
  const Geo &geo = a.geo;
  const double *ge = a.gm + (size_t)e * 5 * Q2;   // wJ dxi_d/dx_x (rows r, s), wJ
  const double gm1 = a.phys.gamma - 1.0;

  if (PROJ) {
    for (int i = tid; i < NF * Q2; i += NT) {
      const int v = i / Q2, n = i - v * Q2;
      Fw[i] = ge[4 * Q2 + n] * a.U[((size_t)v * ne + e) * Q2 + n];
    }
    __syncthreads();
  } else {
  // modal state -> nodal
  for (int i = tid; i < NF * L2; i += NT) {
    const int v = i / L2;
    S1[i] = a.U[((size_t)v * ne + e) * L2 + (i - v * L2)];
  }
  __syncthreads();
  contract<NF, L, L, 1, Q, NT>(CT<P>::BT(), S1, S2, tid);   // [NF][Q][L]
  __syncthreads();
  contract<NF, Q, L, 2, Q, NT>(CT<P>::BT(), S2, Un, tid);   // [NF][Q][Q]

  // edge corrections g = (J_side/J_vol)(Fhat - F(U-).n) at the Gauss points
  bool bad = false;
  for (int i = tid; i < 4 * QF; i += NT) {
    const int s = i / QF, q = i - s * QF;
    const int es = e * 4 + s;
    const double *fg = a.fgeo + ((size_t)es * QF + q) * 4;
    double nx = fg[0], ny = fg[1];
    const double jr = fg[3];
    const int nb = geo.nbr[es];
    double UL[4], UR[4], FL[4], FR[4], Fh[4], lL, lR;
    const double *tl = a.T + ((size_t)e * 4 + s) * NF * QF;
#pragma unroll
    for (int v = 0; v < NF; ++v) UL[v] = tl[v * QF + q];
    if (nb >= 0) {
      const int qn = geo.perm[geo.permid[es] * QF + q];
      if (!geo.master[es]) {   // the pair's flux uses the master side's normal (both sides' normals
                               // come from their own interpolated metrics and differ slightly)
        const double *fm = a.fgeo + (((size_t)nb * 4 + geo.nside[es]) * QF + qn) * 4;
        nx = -fm[0]; ny = -fm[1];
      }
      const double *tr = a.T + ((size_t)nb * 4 + geo.nside[es]) * NF * QF;
#pragma unroll
      for (int v = 0; v < NF; ++v) UR[v] = tr[v * QF + qn];
    } else if (SHALLOW) {
      ghost_shallow2(-1 - nb, UL, nx, ny, a.phys.bcU, UR);
    } else {   // no Euler boundaries on quads (rejected at creation)
#pragma unroll
      for (int v = 0; v < 4; ++v) UR[v] = UL[v];
    }
    if (SHALLOW) {
      bad |= shallow_normal_flux2(UL, nx, ny, FL, lL);
      bad |= shallow_normal_flux2(UR, nx, ny, FR, lR);
      const double lmax = fmax(lL, lR);
#pragma unroll
      for (int v = 0; v < 3; ++v) Fh[v] = 0.5 * (FL[v] + FR[v] + lmax * (UL[v] - UR[v]));
    } else {
      bool bl, br;
      euler_normal_flux2(UL, nx, ny, gm1, FL, lL, bl);
      euler_normal_flux2(UR, nx, ny, gm1, FR, lR, br);
      bad |= bl | br;
      if (a.phys.flux == 1) {
        roe_flux2(UL, UR, nx, ny, a.phys.gamma, FL, Fh);
      } else {
        const double lmax = fmax(lL, lR);
#pragma unroll
        for (int v = 0; v < 4; ++v) Fh[v] = 0.5 * (FL[v] + FR[v] + lmax * (UL[v] - UR[v]));
      }
    }
#pragma unroll
    for (int v = 0; v < NF; ++v) gF[(v * 4 + s) * QF + q] = jr * (Fh[v] - FL[v]);
  }
  if (bad) atomicOr(a.flag, 1);

  // volume: wJ * sum_x sum_d (dxi_d/dx_x) D_d F_x at every node, field by field
  for (int v = 0; v < NF; ++v) {
    __syncthreads();
    for (int n = tid; n < Q2; n += NT) {
      double fx, fy;
      if (SHALLOW) {
        const double phi = Un[n], mx = Un[Q2 + n], my = Un[2 * Q2 + n];
        const double u = mx / phi, vv = my / phi, p = 0.5 * (phi * phi);
        if (v == 0) { fx = mx; fy = my; }
        else if (v == 1) { fx = mx * u + p; fy = mx * vv; }
        else { fx = my * u; fy = my * vv + p; }
      } else {
        const double r = Un[n], ru = Un[Q2 + n], rv = Un[2 * Q2 + n], rE = Un[3 * Q2 + n];
        const double ir = 1.0 / r, u = ru * ir, vv = rv * ir;
        const double p = gm1 * (rE - 0.5 * (ru * u + rv * vv));
        if (v == 0) { fx = ru; fy = rv; }
        else if (v == 1) { fx = ru * u + p; fy = ru * vv; }
        else if (v == 2) { fx = rv * u; fy = rv * vv + p; }
        else { const double h = rE + p; fx = u * h; fy = vv * h; }
      }
      Fc[n] = fx; Fc[Q2 + n] = fy;
    }
    __syncthreads();
    const double *Dn = CT<P>::DN();
#pragma unroll
    for (int rr = 0; rr < NPT; ++rr) {
      const int n = tid + rr * NT;
      if (n < Q2) {
        const int i = n / Q, j = n - i * Q;
        double acc = 0.0;
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          const double *F = Fc + x * Q2;
          double dr = 0.0, ds = 0.0;
#pragma unroll
          for (int m = 0; m < Q; ++m) {
            dr = fma(Dn[i * Q + m], F[m * Q + j], dr);
            ds = fma(Dn[j * Q + m], F[i * Q + m], ds);
          }
          acc = fma(ge[x * Q2 + n], dr, acc);
          acc = fma(ge[(2 + x) * Q2 + n], ds, acc);
        }
        Fw[v * Q2 + n] = acc;
      }
    }
  }
  __syncthreads();
  // edges: Gauss -> GLL edge nodes for NF x 4 blocks, then wJ / w_end at the edge nodes
  contract<4 * NF, QF, 1, 1, Q, NT>(CT<P>::GL(), gF, S1, tid);   // [NF 4][Q]
  __syncthreads();
  {
    const double wend = CT<P>::BW()[0] / CT<P>::BT()[0];   // GLL end weight: phi_0 w_0 / phi_0
#pragma unroll
    for (int rr = 0; rr < NPT; ++rr) {
      const int n = tid + rr * NT;
      if (n < Q2) {
        const int i = n / Q, j = n - i * Q;
        const double c = ge[4 * Q2 + n] / wend;
        // edge 0: j = 0 (along i), 1: i = Q-1 (along j), 2: j = Q-1 (along i), 3: i = 0 (along j)
        int es[2], t[2], cnt = 0;
        if (j == 0) { es[cnt] = 0; t[cnt++] = i; }
        if (i == Q - 1) { es[cnt] = 1; t[cnt++] = j; }
        if (j == Q - 1) { es[cnt] = 2; t[cnt++] = i; }
        if (i == 0) { es[cnt] = 3; t[cnt++] = j; }
        for (int f = 0; f < cnt; ++f)
#pragma unroll
          for (int v = 0; v < NF; ++v)
            Fw[v * Q2 + n] = fma(c, S1[(v * 4 + es[f]) * Q + t[f]], Fw[v * Q2 + n]);
      }
    }
  }
  __syncthreads();
  }   // !PROJ
  // b = B^T (wJ f): into Un (compact [NF][L2])
  contract<NF, Q, Q, 1, L, NT>(CT<P>::BL(), Fw, S1, tid);   // [NF][L][Q]
  __syncthreads();
  contract<NF, L, Q, 2, L, NT>(CT<P>::BL(), S1, Un, tid);   // [NF][L][L]
  __syncthreads();

  // conjugate gradients on M x = b, all fields at once; x0 = b / mean(J)
  double *Pv = Fw, *AP = Un;
  double x[NML], res[NML], bvec[NML];
  double part[NF], bb[NF], rs[NF], pap[NF];
  const double jinv = 1.0 / a.jav[e];
#pragma unroll
  for (int f = 0; f < NF; ++f) part[f] = 0.0;
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int g = tid + k * NT;
    x[k] = 0.0; res[k] = 0.0; bvec[k] = 0.0;
    if (g < NF * L2) {
      bvec[k] = Un[g];
      x[k] = bvec[k] * jinv;
      part[g / L2] = fma(bvec[k], bvec[k], part[g / L2]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int g = tid + k * NT;
    if (g < NF * L2) Pv[g] = x[k];
  }
  field_dots2<NF, NW>(part, red, bb, tid);
  auto mass = [&]() {   // AP = B^T diag(wJ) B Pv
    contract<NF, L, L, 1, Q, NT>(CT<P>::BT(), Pv, S1, tid);   // [NF][Q][L]
    __syncthreads();
    contract<NF, Q, L, 2, Q, NT>(CT<P>::BT(), S1, S2, tid);   // [NF][Q][Q]
    __syncthreads();
    for (int i = tid; i < NF * Q2; i += NT) S2[i] *= ge[4 * Q2 + (i % Q2)];
    __syncthreads();
    contract<NF, Q, Q, 1, L, NT>(CT<P>::BL(), S2, S1, tid);   // [NF][L][Q]
    __syncthreads();
    contract<NF, L, Q, 2, L, NT>(CT<P>::BL(), S1, AP, tid);   // [NF][L][L]
    __syncthreads();
  };
  mass();
#pragma unroll
  for (int f = 0; f < NF; ++f) part[f] = 0.0;
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int g = tid + k * NT;
    if (g < NF * L2) {
      res[k] = bvec[k] - AP[g];
      part[g / L2] = fma(res[k], res[k], part[g / L2]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int g = tid + k * NT;
    if (g < NF * L2) Pv[g] = res[k];
  }
  field_dots2<NF, NW>(part, red, rs, tid);
  constexpr double tol2 = 1e-30;   // (1e-15 relative)^2
  int it = 0;
  for (; it < 60; ++it) {
    bool conv = true;
#pragma unroll
    for (int f = 0; f < NF; ++f) conv = conv && (rs[f] <= tol2 * bb[f]);
    if (conv) break;
    mass();
#pragma unroll
    for (int f = 0; f < NF; ++f) part[f] = 0.0;
#pragma unroll
    for (int k = 0; k < NML; ++k) {
      const int g = tid + k * NT;
      if (g < NF * L2) part[g / L2] = fma(Pv[g], AP[g], part[g / L2]);
    }
    field_dots2<NF, NW>(part, red, pap, tid);
    double alpha[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) alpha[f] = (rs[f] > 0.0 && pap[f] > 0.0) ? rs[f] / pap[f] : 0.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) part[f] = 0.0;
#pragma unroll
    for (int k = 0; k < NML; ++k) {
      const int g = tid + k * NT;
      if (g < NF * L2) {
        const int f = g / L2;
        x[k] = fma(alpha[f], Pv[g], x[k]);
        res[k] = fma(-alpha[f], AP[g], res[k]);
        part[f] = fma(res[k], res[k], part[f]);
      }
    }
    double rn[NF];
    field_dots2<NF, NW>(part, red, rn, tid);
#pragma unroll
    for (int k = 0; k < NML; ++k) {
      const int g = tid + k * NT;
      if (g < NF * L2) {
        const int f = g / L2;
        const double beta = rs[f] > 0.0 ? rn[f] / rs[f] : 0.0;
        Pv[g] = fma(beta, Pv[g], res[k]);
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) rs[f] = rn[f];
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int g = tid + k * NT;
    if (g < NF * L2) {
      const int v = g / L2;
      a.R[((size_t)v * ne + e) * L2 + (g - v * L2)] = x[k];
    }
  }
  if (tid == 0 && a.iters) atomicMax(a.iters, it);
}

template <int P, bool SHALLOW>
static hipError_t launch_c2s(const CurvedArgs &a) {
  constexpr int sh = C2<P, SHALLOW ? 3 : 4>::lds;
  if (a.project) {
    (void)hipFuncSetAttribute((const void *)k_rhs2c<P, SHALLOW, true>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
    hipLaunchKernelGGL((k_rhs2c<P, SHALLOW, true>), dim3(a.ne), dim3(128), sh, a.stream, a);
  } else {
    (void)hipFuncSetAttribute((const void *)k_rhs2c<P, SHALLOW, false>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
    hipLaunchKernelGGL((k_rhs2c<P, SHALLOW, false>), dim3(a.ne), dim3(128), sh, a.stream, a);
  }
  return hipGetLastError();
}

template <int P>
static hipError_t launch_c2(const CurvedArgs &a) {
  return a.shallow ? launch_c2s<P, true>(a) : launch_c2s<P, false>(a);
}

hipError_t launch_curved2(const CurvedArgs &a) {
  switch (a.p) {
#define DGX_CASE(PP) case PP: return launch_c2<PP>(a);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5)
    DGX_CASE(6) DGX_CASE(7) DGX_CASE(8) DGX_CASE(9) DGX_CASE(10)
#undef DGX_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t upload_tables_c2d(const double *h) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), h, sizeof(double) * kTabTotal);
}

}  // namespace dgx
