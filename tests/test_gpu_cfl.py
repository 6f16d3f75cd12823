"""CFL time step on the GPU (dgx_set_time_step) against the oracle's
Euler::set_time_step / Shallow::set_time_step (src/Euler.cpp:804-877,
src/Shallow.cpp:573-637), whose closed forms tests/test_oracle_cfl.py pins.

Bar: relative difference <= 1e-13 (a maximum over nodes of the same formula;
only the order of a few roundings differs)."""
import numpy as np
import pytest

from dgm_amd import dgx, ic, mesh as meshgen
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def _compare(m, p, eqn, U_fn, xnodes=None, **kw):
    orc = Oracle(m, p, eqn)
    if xnodes is not None:
        orc.set_nodes(xnodes)
    x = orc.coords()
    orc.set_state(U_fn(x))
    orc.project_state()
    Up = orc.state()
    gpu = dgx.GPUDomain(m, p, eqn, xnodes=xnodes, balance=False)
    gpu.upload(Up)
    orc.set_state(gpu.download())          # both evaluate the same (projected) nodal state
    for cflmax, dtmax in ((0.0, 0.0), (0.2, 0.0), (0.2, 1e-4)):
        want = orc.set_time_step(1e-3, cflmax, dtmax)
        got = gpu.set_time_step(1e-3, cflmax, dtmax)
        print(f"[cfl] eqn={eqn} p={p} CFLmax={cflmax} dtmax={dtmax}: gpu {got} oracle {want}", flush=True)
        assert abs(got[0] - want[0]) <= 1e-13 * want[0] and abs(got[1] - want[1]) <= 1e-13 * want[1]
    gpu.close()


@pytest.mark.parametrize("p", [2, 4, 8, 10])
def test_cfl_euler_vortex(p):
    """The isentropic vortex on periodic [0,10]^2 (the 2D Euler workload's IC)."""
    m = meshgen.periodic_quads(8, 8, 0.0, 10.0, 0.0, 10.0)
    _compare(m, p, 1, lambda x: ic.isentropic_vortex(x[0], x[1], 4.2, 5.7))


def test_cfl_shallow_and_curved():
    m = meshgen.periodic_quads(5, 4, 0.0, 5.0, 0.0, 4.0)
    rng = np.random.default_rng(2)
    _compare(m, 5, 30, lambda x: np.stack([1.0 + 0.05 * np.sin(x[0]) * np.cos(x[1]),
                                           0.1 * np.cos(x[1]), -0.05 * np.sin(x[0])]))
    # curved quads: per-node metrics
    p = 4
    xn = meshgen.quad_nodes(m, p)
    xn[1] = xn[1] + 0.08 * np.sin(2 * np.pi * xn[0] / 5.0) * np.sin(np.pi * xn[1] / 4.0) ** 2
    _compare(m, p, 1, lambda x: ic.isentropic_vortex(x[0], x[1], 2.5, 2.0, strength=1.0), xnodes=xn)


def test_cfl_hexes_keep_dt():
    gpu = dgx.GPUDomain(meshgen.periodic_hexes(2, 2, 2), 3, 8)
    assert gpu.set_time_step(2e-3, 0.5) == (2e-3, 2e-3)
    gpu.close()


def test_problem_runs_at_a_cfl_number(tmp_path):
    """ReoProblem with -CFLmax: every step takes the step the CFL condition gives."""
    import io
    import os
    import shutil
    from dgm_amd import problem as dgm
    from fixture_runs import GOLDEN
    for ext in (".inp", ".msh", ".con", ".ic", ".bc"):
        shutil.copy(os.path.join(GOLDEN, "vquad" + ext), tmp_path / ("vquad" + ext))
    log = io.StringIO()
    prob = dgm.ReoProblem(dgm.ArgList(["-CFLmax", "0.2", "-Nt", "4", "-ntout", "1", str(tmp_path / "vquad")]),
                          out=log)
    try:
        assert prob.solve() == dgm.DGM_SUCCESS
    finally:
        prob.close()
    rows = [ln.split() for ln in log.getvalue().splitlines() if ln.strip()[:1].isdigit()]
    print(log.getvalue(), flush=True)
    assert len(rows) == 4
    for r in rows:
        assert abs(float(r[3]) - 0.2) < 1e-12
