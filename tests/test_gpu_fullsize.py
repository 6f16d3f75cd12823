"""The Taylor-Green workload at its full size (32^3 hexes, p = 7, 83.9 M DOF)
through properties that do not depend on size.

The oracle finishes only small meshes in seconds, so the full-size runs are
checked against what the discretization guarantees on a periodic mesh:
* conservation: the domain integral of every field's residual vanishes, since
  each face flux enters the two elements it joins with opposite signs.  On
  this uniform mesh the integral is proportional to the sum over elements of
  the residual's constant mode.  The CPU oracle shows the property on small
  meshes to 1e-14 (tests/test_oracle_invariants.py covers it for the oracle);
* the same for the mixed path, to FP32 rounding;
* LSRK14 steps keep the total mass, momentum and energy (the sums of the
  state's constant modes) to rounding.
A dropped face term, a wrong sign on one side of a face or a one-sided
viscous flux breaks these at O(1).

Measured: FP64 residuals conserve to 5e-15 of the residual's L1 norm (bar
1e-12), the mixed one to 7e-11 (bar 1e-9: its FP32 volume contractions
cancel only to FP32 rounding), and 3 LSRK14 steps move the totals by 2e-18
of their scale (bar 1e-12).
"""
import numpy as np
import pytest

from dgm_amd import dgx, ic, mesh as meshgen

pytestmark = pytest.mark.gpu

N, P = 32, 7


def _tgv(precision, flux, viscous):
    m = meshgen.periodic_hexes(N, N, N, (0.0, 0.0, 0.0), (2 * np.pi,) * 3)
    gpu = dgx.GPUDomain(m, P, dgx.DGX_NAVIER_STOKES3D, gamma=1.4, vis=1.0 / 1600.0, Pr=0.71, precision=precision,
                        flux=flux, viscous=viscous)
    x = gpu.coords()
    gpu.upload(ic.taylor_green(x[0], x[1], x[2]))
    return gpu


def _conservation(R):
    """per field: |sum of the constant modes| over the L1 norm of the whole
    residual (all fields: the Taylor-Green start is divergence-free, so the
    density residual alone is itself at rounding level)"""
    return np.abs(R[:, :, 0].sum(axis=1)) / np.abs(R).sum()


@pytest.mark.parametrize("precision,flux,viscous,bar", [
    (dgx.DGX_FP64, dgx.DGX_FLUX_ROE, dgx.DGX_VISCOUS_BR2, 1e-12),
    (dgx.DGX_FP64, dgx.DGX_FLUX_LF, dgx.DGX_VISCOUS_BR1, 1e-12),
    (dgx.DGX_MIXED, dgx.DGX_FLUX_ROE, dgx.DGX_VISCOUS_BR2, 1e-9),
])
def test_full_size_residual_is_conservative(precision, flux, viscous, bar):
    print(f"[fullsize] TGV {N}^3 p={P} precision={precision} flux={flux} viscous={viscous}", flush=True)
    gpu = _tgv(precision, flux, viscous)
    R = gpu.rhs(dgx.DGX_MODAL)
    gpu.close()
    c = _conservation(R)
    print(f"[fullsize] |sum of constant modes| / |residual|_1 per field: {c}", flush=True)
    assert np.all(np.isfinite(R))
    assert np.all(c <= bar), c


def test_full_size_lsrk14_keeps_the_totals():
    print(f"[fullsize] TGV {N}^3 p={P} FP64 Roe + BR2, 3 LSRK14 steps", flush=True)
    gpu = _tgv(dgx.DGX_FP64, dgx.DGX_FLUX_ROE, dgx.DGX_VISCOUS_BR2)
    U0 = gpu.download(dgx.DGX_MODAL)
    # relative to all fields' constant modes together (the Taylor-Green start has w = 0 everywhere)
    t0, s0 = U0[:, :, 0].sum(axis=1), np.abs(U0[:, :, 0]).sum()
    del U0
    h = 2 * np.pi / N
    gpu.advance(dgx.DGX_LSRK14, 0.5 * h / ((P + 1) ** 2 * 11.0), 3)
    U = gpu.download(dgx.DGX_MODAL)
    gpu.close()
    drift = np.abs(U[:, :, 0].sum(axis=1) - t0) / s0
    print(f"[fullsize] drift of the totals over 3 steps, relative to the sum of |constant modes|: {drift}", flush=True)
    assert np.all(np.isfinite(U))
    assert np.all(drift <= 1e-12), drift
