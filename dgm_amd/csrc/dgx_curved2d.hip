// Curved quads (Euler 2D, Navier-Stokes 2D and shallow water, p <= 10): the
// residual with metrics at the nodes and the projection through the element
// mass matrix, the quad counterpart of dgx_curved3d.hip.
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
//   viscous       BR1 (Navier_Stokes::diffusion, src/Navier_Stokes.cpp:113-258):
//                 k_grad2c lifts and projects the nodal gradients of (u, v, e)
//                 with each side's own normal, k_visc2c forms the stresses and
//                 heat flux at the nodes, their weighted divergence and each
//                 side's normal flux, and the residual adds -div F_v and the
//                 lifted 1/2 (F_v.n own + F_v.n neighbour) before projecting.
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

// x = M^-1 B^T Fw for NF fields, Fw [NF][Q2] the weighted nodal values (wJ f),
// M = B^T diag(wJ) B, by conjugate gradients on all fields at once from
// x0 = b / mean(J); x goes to out[v * ostride + k].  Scratch: S1, S2 [NF Q2],
// AP [NF L2]; Fw is overwritten.
template <int P, int NF, int NT>
__device__ void curved_project2(double *Fw, double *S1, double *S2, double *AP, double *red, const double *ge,
                                double jav, double *out, size_t ostride, int *iters, int tid,
                                double *Bal = nullptr, int bal = 0, const double *cf = nullptr) {
  constexpr int L = P + 1, Q = P + 2, L2 = L * L, Q2 = Q * Q;
  constexpr int NW = NT / 64, NML = (NF * L2 + NT - 1) / NT;
  // b = B^T (wJ f): into AP (compact [NF][L2])
  contract<NF, Q, Q, 1, L, NT>(CT<P>::BL(), Fw, S1, tid);   // [NF][L][Q]
  __syncthreads();
  contract<NF, L, Q, 2, L, NT>(CT<P>::BL(), S1, AP, tid);   // [NF][L][L]
  __syncthreads();
  if (cf) {
    // M x = b with the element's Cholesky factor M = C C^T (rows of C packed),
    // forward then backward substitution, one thread per field, in the order
    // of the reference's factored solve (DGM_CHOLESKY_MASS_MATRIX,
    // src/Quad.cpp:263-300) and without contraction of its products
    if (tid < NF) {
#pragma clang fp contract(off)
      double *x = AP + tid * L2;
      for (int i = 0; i < L2; ++i) {
        const double *row = cf + (size_t)i * (i + 1) / 2;
        double t = x[i];
        for (int k = 0; k < i; ++k) t -= row[k] * x[k];
        x[i] = t / row[i];
      }
      for (int i = L2 - 1; i >= 0; --i) {
        double t = x[i];
        for (int k = i + 1; k < L2; ++k) t -= cf[(size_t)k * (k + 1) / 2 + i] * x[k];
        x[i] = t / cf[(size_t)i * (i + 1) / 2 + i];
      }
    }
    __syncthreads();
    for (int g = tid; g < NF * L2; g += NT) {
      const int v = g / L2;
      const size_t o = v * ostride + (g - v * L2);
      double r = AP[g];
      if (bal == 1) { Bal[o] = r; r = -1.0 * r + r; }   // the balance term (src/Shallow.cpp:251-266)
      else if (bal == 2) r = -1.0 * Bal[o] + r;
      out[o] = r;
    }
    return;
  }

  double *Pv = Fw;
  double x[NML], res[NML], bvec[NML];
  double part[NF], bb[NF], rs[NF], pap[NF];
  const double jinv = 1.0 / jav;
#pragma unroll
  for (int f = 0; f < NF; ++f) part[f] = 0.0;
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int g = tid + k * NT;
    x[k] = 0.0; res[k] = 0.0; bvec[k] = 0.0;
    if (g < NF * L2) {
      bvec[k] = AP[g];
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
      const size_t o = v * ostride + (g - v * L2);
      double r = x[k];
      if (bal == 1) { Bal[o] = r; r = -1.0 * r + r; }   // the balance term (src/Shallow.cpp:251-266)
      else if (bal == 2) r = -1.0 * Bal[o] + r;
      out[o] = r;
    }
  }
  if (tid == 0 && iters) atomicMax(iters, it);
}

// edge corrections gF [NF][4][QF] (Gauss points) lifted to the edge nodes and
// added to Fw [NF][Q2] with weight wJ / w_end: G2GL(g) / w_end at the edge
// nodes (src/Quad.cpp:1971-2038), before the wJ of the inner product
template <int P, int NF, int NT>
__device__ __forceinline__ void lift_edges2(const double *gF, double *S1, double *Fw, const double *ge, int tid) {
  constexpr int Q = P + 2, Q2 = Q * Q, QF = P + 1, NPT = (Q2 + NT - 1) / NT;
  contract<4 * NF, QF, 1, 1, Q, NT>(CT<P>::GL(), gF, S1, tid);   // [NF 4][Q]
  __syncthreads();
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
  __syncthreads();
}

// modal fields M [NF][L2] (LDS) -> nodal Out [NF][Q2], through S [NF][Q][L]
template <int P, int NF, int NT>
__device__ __forceinline__ void nodal2(const double *M, double *S, double *Out, int tid) {
  constexpr int L = P + 1, Q = P + 2;
  contract<NF, L, L, 1, Q, NT>(CT<P>::BT(), M, S, tid);   // [NF][Q][L]
  __syncthreads();
  contract<NF, Q, L, 2, Q, NT>(CT<P>::BT(), S, Out, tid);   // [NF][Q][Q]
  __syncthreads();
}

// weighted divergence wJ (d/dx Fx + d/dy Fy) at every node of an element:
// sum_x sum_d (wJ dxi_d/dx_x) D_d F_x (src/Quad.cpp:2309-2330 with the curved metrics)
template <int P, int NT>
__device__ __forceinline__ void divergence2c(const double *Fc, const double *ge, double *out, int tid) {
  constexpr int Q = P + 2, Q2 = Q * Q, NPT = (Q2 + NT - 1) / NT;
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
      out[n] = acc;
    }
  }
}

// PROJ: project the nodal fields a.U [NF][ne][Q^2] instead (the forward
// transform of a curved element: x = M^-1 B^T (wJ u)).
// NS: add the viscous terms of V and TV (k_visc2c) to the Euler residual.
template <int P, bool SHALLOW, bool PROJ, bool NS>
__global__ void __launch_bounds__(128)
k_rhs2c(CurvedArgs a) {
  constexpr int NF = SHALLOW ? 3 : 4;
  using S = C2<P, NF>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, Q2 = S::Q2;
  constexpr int NT = S::NT;
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
    } else if (SHALLOW && -1 - nb == 2) {   // function boundary (FuncBC) at the Gauss point and time t
      const double *xy = a.fn.fxy + ((size_t)es * QF + q) * 2;
#pragma unroll
      for (int v = 0; v < 3; ++v) UR[v] = eval_program(a.fn.code, a.fn.bc[v], a.fn.bc[v + 1], a.fn.consts, xy[0], xy[1], a.t);
    } else if (SHALLOW) {
      ghost_shallow2(-1 - nb, UL, nx, ny, a.phys.bcU, UR);
    } else {   // Euler boundary (Navier-Stokes meshes have none: rejected at creation)
      ghost2(-1 - nb, UL, nx, ny, a.phys.bcU, UR);
    }
    double gv[4] = {0.0, 0.0, 0.0, 0.0};
    if (NS && nb >= 0) {   // viscous: own normal flux less the average 1/2 (own - neighbour's), negated
      const int qn = geo.perm[geo.permid[es] * QF + q];
      const double *to = a.TV + ((size_t)es * 3) * QF;
      const double *tn = a.TV + (((size_t)nb * 4 + geo.nside[es]) * 3) * QF;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double own = to[c * QF + q], avg = 0.5 * (own + -1.0 * tn[c * QF + qn]);
        gv[1 + c] = jr * (avg - own);
      }
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
    for (int v = 0; v < NF; ++v) gF[(v * 4 + s) * QF + q] = NS ? jr * (Fh[v] - FL[v]) - gv[v] : jr * (Fh[v] - FL[v]);
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
    divergence2c<P, NT>(Fc, ge, Fw + v * Q2, tid);
    if (SHALLOW && a.fn.has_src) {   // FuncSource::apply: F -= s at the nodes (src/Source.cpp:11-24)
      const double *xx = a.fn.xn + (size_t)e * Q2, *yy = a.fn.xn + ((size_t)ne + e) * Q2;
      for (int n = tid; n < Q2; n += NT)
        Fw[v * Q2 + n] -= ge[4 * Q2 + n] * eval_program(a.fn.code, a.fn.src[v], a.fn.src[v + 1], a.fn.consts,
                                                         xx[n], yy[n], a.t);
    }
    if (NS && v > 0) {
      const double *vn = a.V + ((size_t)e * 3 + (v - 1)) * Q2;
      for (int n = tid; n < Q2; n += NT) Fw[v * Q2 + n] -= vn[n];
    }
  }
  __syncthreads();
  lift_edges2<P, NF, NT>(gF, S1, Fw, ge, tid);
  }   // !PROJ
  curved_project2<P, NF, NT>(Fw, S1, S2, Un, red, ge, a.jav[e], a.R + (size_t)e * L2, (size_t)ne * L2, a.iters, tid,
                             (SHALLOW && !PROJ && a.bal) ? a.Bal + (size_t)e * L2 : nullptr,
                             (SHALLOW && !PROJ) ? a.bal : 0,
                             a.chol ? a.chol + (size_t)e * (L2 * (L2 + 1) / 2) : nullptr);
}

// Lifted, projected gradients of the primitives (u, v, e) of a curved quad
// (Euler::gradient_lift path of Navier_Stokes::diffusion, src/Euler.cpp:1406-1456):
// at the nodes g_x = dxi_d/dx_x D_d q, plus on each side G2GL(J_side/J_vol
// ({q} - q-) n_x) / w_end with the side's own normal, then projected through
// the curved mass matrix; G [ne][6][L2] in the order (u_x, u_y, v_x, v_y, e_x, e_y).
template <int P>
__global__ void __launch_bounds__(128)
k_grad2c(CurvedArgs a) {
  constexpr int L = P + 1, Q = P + 2, QF = P + 1, L2 = L * L, Q2 = Q * Q, NT = 128;
  constexpr int NPT = (Q2 + NT - 1) / NT;
  extern __shared__ double sm[];
  double *Un = sm, *Fw = Un + 4 * Q2, *S1 = Fw + 6 * Q2, *S2 = S1 + 6 * Q2, *AP = S2 + 6 * Q2;
  double *gF = AP + 6 * L2, *Pq = gF + 6 * 4 * QF, *red = Pq + Q2;
  const int e = blockIdx.x, tid = threadIdx.x, ne = a.ne;
  const Geo &geo = a.geo;
  const double *ge = a.gm + (size_t)e * 5 * Q2;
  for (int i = tid; i < 4 * L2; i += NT) {
    const int v = i / L2;
    S2[i] = a.U[((size_t)v * ne + e) * L2 + (i - v * L2)];
  }
  __syncthreads();
  nodal2<P, 4, NT>(S2, S1, Un, tid);
  // jumps of the primitives at the Gauss points: J_side/J_vol ({q} - q-) n
  for (int i = tid; i < 4 * QF; i += NT) {
    const int s = i / QF, q = i - s * QF;
    const int es = e * 4 + s;
    const double *fg = a.fgeo + ((size_t)es * QF + q) * 4;
    const double nx = fg[0], ny = fg[1], jr = fg[3];
    const int nb = geo.nbr[es];   // Navier-Stokes quads have no boundary sides (rejected at creation)
    const int qn = nb >= 0 ? geo.perm[geo.permid[es] * QF + q] : q;
    const double *tl = a.T + ((size_t)es) * 4 * QF;
    const double *tr = nb >= 0 ? a.T + ((size_t)nb * 4 + geo.nside[es]) * 4 * QF : tl;
    double pl[3], pr[3];
    {
      const double r = tl[q], u = tl[QF + q] / r, v = tl[2 * QF + q] / r;
      double en = tl[3 * QF + q] / r;
      double wk = u * u; wk = v * v + wk;
      pl[0] = u; pl[1] = v; pl[2] = -0.5 * wk + en;
    }
    {
      const double r = tr[qn], u = tr[QF + qn] / r, v = tr[2 * QF + qn] / r;
      double en = tr[3 * QF + qn] / r;
      double wk = u * u; wk = v * v + wk;
      pr[0] = u; pr[1] = v; pr[2] = -0.5 * wk + en;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double jmp = jr * (0.5 * (pl[c] + pr[c]) - pl[c]);
      gF[((2 * c) * 4 + s) * QF + q] = nx * jmp;
      gF[((2 * c + 1) * 4 + s) * QF + q] = ny * jmp;
    }
  }
  // nodal gradients, weighted: wJ dq/dx_x = sum_d (wJ dxi_d/dx_x) D_d q
  const double *Dn = CT<P>::DN();
  for (int c = 0; c < 3; ++c) {
    for (int n = tid; n < Q2; n += NT) {
      const double r = Un[n], u = Un[Q2 + n] / r, v = Un[2 * Q2 + n] / r;
      double val;
      if (c == 0) val = u;
      else if (c == 1) val = v;
      else {
        double en = Un[3 * Q2 + n] / r;
        double wk = u * u; wk = v * v + wk;
        val = -0.5 * wk + en;
      }
      Pq[n] = val;
    }
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < NPT; ++rr) {
      const int n = tid + rr * NT;
      if (n < Q2) {
        const int i = n / Q, j = n - i * Q;
        double dr = 0.0, ds = 0.0;
#pragma unroll
        for (int m = 0; m < Q; ++m) {
          dr = fma(Dn[i * Q + m], Pq[m * Q + j], dr);
          ds = fma(Dn[j * Q + m], Pq[i * Q + m], ds);
        }
        Fw[(2 * c) * Q2 + n] = fma(ge[2 * Q2 + n], ds, ge[n] * dr);
        Fw[(2 * c + 1) * Q2 + n] = fma(ge[3 * Q2 + n], ds, ge[Q2 + n] * dr);
      }
    }
    __syncthreads();
  }
  lift_edges2<P, 6, NT>(gF, S1, Fw, ge, tid);
  curved_project2<P, 6, NT>(Fw, S1, S2, AP, red, ge, a.jav[e], a.G + (size_t)e * 6 * L2, L2, a.iters, tid, nullptr, 0,
                            a.chol ? a.chol + (size_t)e * (L2 * (L2 + 1) / 2) : nullptr);
}

// Viscous terms of a curved quad from the state and the lifted gradients G:
// the stresses and heat flux at the nodes (src/Navier_Stokes.cpp:113-258),
// V [ne][3][Q2] = wJ div F_v for (rho u, rho v, rho E), and each side's normal
// flux F_v.n at its Gauss points with its own normal into TV [ne][4][3][QF]
// (interpolated first, then dotted, as the reference's side flux is formed).
template <int P>
__global__ void __launch_bounds__(128)
k_visc2c(CurvedArgs a) {
  constexpr int L = P + 1, Q = P + 2, QF = P + 1, L2 = L * L, Q2 = Q * Q, NT = 128;
  extern __shared__ double sm[];
  double *Un = sm, *Gn = Un + 4 * Q2, *Fv = Gn + 6 * Q2, *M = Fv + 6 * Q2, *S = M + 6 * L2;
  double *Lay = S + 6 * Q * L, *Wf = Lay + 24 * Q;
  const int e = blockIdx.x, tid = threadIdx.x, ne = a.ne;
  const double *ge = a.gm + (size_t)e * 5 * Q2;
  for (int i = tid; i < 4 * L2; i += NT) {
    const int v = i / L2;
    M[i] = a.U[((size_t)v * ne + e) * L2 + (i - v * L2)];
  }
  __syncthreads();
  nodal2<P, 4, NT>(M, S, Un, tid);
  for (int i = tid; i < 6 * L2; i += NT) M[i] = a.G[(size_t)e * 6 * L2 + i];
  __syncthreads();
  nodal2<P, 6, NT>(M, S, Gn, tid);
  const double mu = a.phys.vis, gp = a.phys.gamma / a.phys.Pr;
  for (int n = tid; n < Q2; n += NT) {
    const double r = Un[n], u = Un[Q2 + n] / r, v = Un[2 * Q2 + n] / r;
    const double ux = Gn[n], uy = Gn[Q2 + n], vx = Gn[2 * Q2 + n], vy = Gn[3 * Q2 + n];
    const double ex = Gn[4 * Q2 + n], ey = Gn[5 * Q2 + n];
    const double t12 = mu * (uy + vx);
    const double t11 = mu * ((2.0 / 3.0) * (2.0 * ux - vy));
    const double t22 = mu * ((2.0 / 3.0) * (2.0 * vy - ux));
    Fv[n] = t11; Fv[Q2 + n] = t12;                      // x momentum: (t11, t12)
    Fv[2 * Q2 + n] = t12; Fv[3 * Q2 + n] = t22;         // y momentum: (t12, t22)
    Fv[4 * Q2 + n] = gp * (mu * ex) + fma(t12, v, t11 * u);
    Fv[5 * Q2 + n] = gp * (mu * ey) + fma(t22, v, t12 * u);
  }
  __syncthreads();
  // side values at the edge nodes -> Gauss points, then n . F with the side's normal
  for (int i = tid; i < 4 * Q; i += NT) {
    const int s = i / Q, j = i - s * Q;
    int k;
    switch (s) {
      case 0: k = j * Q; break;
      case 1: k = (Q - 1) * Q + j; break;
      case 2: k = j * Q + Q - 1; break;
      default: k = j; break;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) Lay[(s * 6 + c) * Q + j] = Fv[c * Q2 + k];
  }
  __syncthreads();
  contract<1, 24, Q, 2, QF, NT>(CT<P>::Ig(), Lay, Wf, tid);
  __syncthreads();
  for (int i = tid; i < 4 * 3 * QF; i += NT) {
    const int s = i / (3 * QF), r = i - s * 3 * QF, c = r / QF, q = r - c * QF;
    const double *fg = a.fgeo + (((size_t)e * 4 + s) * QF + q) * 4;
    a.TV[((size_t)e * 4 + s) * 3 * QF + r] =
        fma(fg[1], Wf[(s * 6 + 2 * c + 1) * QF + q], fg[0] * Wf[(s * 6 + 2 * c) * QF + q]);
  }
  for (int c = 0; c < 3; ++c) divergence2c<P, NT>(Fv + 2 * c * Q2, ge, a.V + ((size_t)e * 3 + c) * Q2, tid);
}

template <int P> struct C2NS {   // LDS bytes of the viscous kernels
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1, L2 = L * L, Q2 = Q * Q;
  static constexpr int grad = (4 * Q2 + 18 * Q2 + 6 * L2 + 24 * QF + Q2 + 8 * 2) * 8;
  static constexpr int visc = (16 * Q2 + 6 * L2 + 6 * Q * L + 24 * Q + 24 * QF) * 8;
  static_assert(grad <= 64 * 1024 && visc <= 64 * 1024, "LDS");
};

template <int P, bool SHALLOW, bool PROJ, bool NS>
static void launch_c2k(const CurvedArgs &a) {
  constexpr int sh = C2<P, SHALLOW ? 3 : 4>::lds;
  (void)hipFuncSetAttribute((const void *)k_rhs2c<P, SHALLOW, PROJ, NS>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
  hipLaunchKernelGGL((k_rhs2c<P, SHALLOW, PROJ, NS>), dim3(a.ne), dim3(128), sh, a.stream, a);
}

template <int P>
static hipError_t launch_c2(const CurvedArgs &a) {
  if (a.project) {
    if (a.shallow) launch_c2k<P, true, true, false>(a);
    else launch_c2k<P, false, true, false>(a);
  } else if (a.shallow) {
    launch_c2k<P, true, false, false>(a);
  } else if (!a.ns) {
    launch_c2k<P, false, false, false>(a);
  } else if (a.op == 1) {
    constexpr int sh = C2NS<P>::grad;
    (void)hipFuncSetAttribute((const void *)k_grad2c<P>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
    hipLaunchKernelGGL(k_grad2c<P>, dim3(a.ne), dim3(128), sh, a.stream, a);
  } else if (a.op == 2) {
    constexpr int sh = C2NS<P>::visc;
    (void)hipFuncSetAttribute((const void *)k_visc2c<P>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
    hipLaunchKernelGGL(k_visc2c<P>, dim3(a.ne), dim3(128), sh, a.stream, a);
  } else {
    launch_c2k<P, false, false, true>(a);
  }
  return hipGetLastError();
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
