"""The step, vField and optimisation seams through the C ABI on the GPU.

* dgx_stage: running the stages of a step one call at a time is bitwise the
  same as dgx_advance (LSRK14 fused and unfused, RK4, TVD-RK3), and the stage
  order is enforced.
* dgx_vfield_sizes: vField::size/qtot/nmodes/ndof/sdof/nsd of the state
  (src/vField.hpp:330-376) for quads and hexes, through the pydgm mirror.
* ReoOptProblem over the acoustic problem: computeState / computeAdjoint /
  getGradient agree with the direct entry points, and a few descent steps
  lower the objective.
"""
import numpy as np
import pytest

from dgm_amd import dgx, ic, mesh as meshgen
from dgm_amd import problem as dgm

pytestmark = pytest.mark.gpu


def _tgv_domain(p=3, eqn=9, **kw):
    m = meshgen.periodic_hexes(2, 2, 2, (0, 0, 0), (2 * np.pi,) * 3)
    d = dgx.GPUDomain(m, p, eqn, vis=0.01 if eqn == 9 else 0.0, **kw)
    x = d.coords()
    U = ic.taylor_green(x[0], x[1], x[2])
    U[1:4] *= 1.0 + 0.02 * np.sin(x[0] + 2 * x[1])
    d.upload(U)
    return d


@pytest.mark.parametrize("inttype,eqn,kw", [(14, 9, {}), (14, 9, {"bulk_forcing": True}), (4, 8, {}), (3, 9, {})])
def test_stages_equal_advance(inttype, eqn, kw):
    print(f"[stage] inttype={inttype} eqn={eqn} {kw}", flush=True)
    a, b = _tgv_domain(eqn=eqn, **kw), _tgv_domain(eqn=eqn, **kw)
    dt, ns = 1e-3, {3: 3, 4: 4, 14: 14}[inttype]
    a.advance(inttype, dt, 2)
    for _ in range(2):
        for s in range(ns):
            b.stage(inttype, s, dt)
    a.synchronize()
    b.synchronize()
    assert np.array_equal(a.download(dgx.DGX_MODAL), b.download(dgx.DGX_MODAL))
    assert a.time == b.time and a.stage_count == b.stage_count
    a.close()
    b.close()


def test_stage_order_is_enforced():
    d = _tgv_domain()
    with pytest.raises(dgx.DgxError):
        d.stage(14, 1, 1e-3)                 # a step starts at stage 0
    d.stage(14, 0, 1e-3)
    with pytest.raises(dgx.DgxError):
        d.advance(14, 1e-3, 1)               # the begun step is unfinished
    with pytest.raises(dgx.DgxError):
        d.stage(4, 1, 1e-3)                  # same integrator through the step
    for s in range(1, 14):
        d.stage(14, s, 1e-3)
    d.advance(14, 1e-3, 1)
    d.synchronize()
    d.close()


def test_vfield_sizes():
    for m, p, eqn, nf, nq, nm, ns, qs in (
            (meshgen.periodic_quads(3, 2), 4, 1, 4, 36, 25, 4, 5),
            (meshgen.periodic_hexes(2, 2, 3), 3, 8, 5, 125, 64, 6, 16),
            (meshgen.periodic_quads(2, 2), 2, 30, 3, 16, 9, 4, 3)):
        d = dgx.GPUDomain(m, p, eqn, balance=False)
        v = dgm.vField(d)
        ne = m.ne
        assert v.size() == nf and v.nsd() == m.nsd
        assert v.qtot() == nf * ne * nq and v.nmodes() == v.ndof() == nf * ne * nm
        assert v.sdof() == nf * ne * ns * qs
        assert v.numpy().shape == (nf, ne, nq)
        v.zero()
        assert not np.any(v.numpy(dgx.DGX_MODAL))
        d.close()


def test_opt_problem_seam():
    from dgm_amd.acoustic import AcousticProblem, ReoOptProblem
    m = meshgen.periodic_quads(8, 8, 0.0, 1.0, 0.0, 1.0)
    rng = np.random.default_rng(4)
    rec = rng.uniform(0.1, 0.9, (4, 2))
    prob = AcousticProblem(m, 3, (1.0, 0.2, 0.0, 1.0 / 1.4), rec, 8.0, 0.12, 2e-3, 120)
    true = rng.standard_normal(m.ne)
    data = prob.forward(true)
    opt = ReoOptProblem(prob, data)
    J0 = opt.computeState()
    opt.computeAdjoint()
    Jd, gd = prob.gradient(opt.getControl(), data)
    assert abs(J0 - Jd) <= 1e-13 * Jd and np.array_equal(opt.getGradient(), gd)
    J3 = opt.solve(iterations=3, step=1.0 / max(1e-300, float(np.abs(gd).max())))
    assert J3 < J0
    prob.close()
