"""End to end on the GPU through the pydgm-style driver (dgm_amd.problem).

``ReoProblem(args, table, comm).solve()`` runs the reference's own regression
run directories (runs/vquad, runs/vhex, runs/cvquad with its arc sides,
runs/shallow_mms with its function boundaries and source, runs/chan with its
isothermal walls, bulk forcing and restart initial condition, runs/vortex
with its variable order per element: mesh,
connectivity, IC, parameters)
on libdgx.so and writes ``root.rst``; that file must match the reference's
golden ``.ref`` under the reference's own diff rule and tolerances
(runs/vquad.tst:23-24, runs/vhex.tst:23-24; src/Domain.cpp:1000-1071).
"""
from __future__ import annotations

import io
import os
import shutil

import numpy as np
import pytest

from dgm_amd import formats
from dgm_amd import problem as dgm
from fixture_runs import GOLDEN
from oracle_lib import Oracle, VarOracle, dgm_diff

pytestmark = pytest.mark.gpu
REPS_M = np.sqrt(np.finfo(float).eps)


def _run(tmp_path, root, extra=()):
    for ext in (".inp", ".msh", ".con", ".ic", ".crv", ".bc", ".bct", ".src", ".chn", ".ini", ".ord"):
        if os.path.exists(os.path.join(GOLDEN, root + ext)):
            shutil.copy(os.path.join(GOLDEN, root + ext), tmp_path / (root + ext))
    base = str(tmp_path / root)
    log = io.StringIO()
    prob = dgm.ReoProblem(dgm.ArgList(list(extra) + [base]), dgm.Table(), dgm.SerialComm(), out=log)
    try:
        assert prob.solve() == dgm.DGM_SUCCESS
    finally:
        prob.close()
    print(log.getvalue(), flush=True)
    return base, prob


@pytest.mark.parametrize("root,aeps,reps", [("vquad", 5e-12, 5e-12), ("vhex", 5e-13, 5e-10),
                                            ("cvquad", 5e-13, 5e-13), ("shallow_mms", 1e-9, 1e-10),
                                            ("chan", 5e-10, 1e-8), ("vortex", 5e-12, 5e-12)])
def test_golden_run_directory(tmp_path, root, aeps, reps):
    print(f"[problem] {root}: full run through ReoProblem.solve()", flush=True)
    base, prob = _run(tmp_path, root)
    prm = prob.params
    ne = prob.mesh.ne
    out = formats.read_restart(base + ".rst", ne)
    ref = formats.read_restart(os.path.join(GOLDEN, root + ".ref"), ne)
    assert out.header["Step"] == int(prm["Nt"])
    if prob.porder is not None:                           # variable order: each element's own basis
        o = VarOracle(prob.mesh, prob.porder)
    else:
        o = Oracle(prob.mesh, int(prm["p"]), int(prm["eqntype"]), vis=float(prm.get("vis", 0.0)))
    if prob.xnodes is not None:                           # the norms integrate over the curved elements
        o.set_nodes(prob.xnodes)
    rn, dn = dgm_diff(o, ref.data, out.data)
    assert np.all(dn <= aeps), f"{root}: absolute diff {dn}"
    mask = rn > REPS_M
    rel = np.where(mask, dn / np.where(mask, rn, 1.0), 0.0)
    assert np.all(rel <= reps), f"{root}: relative diff {rel}"
    # intermediate restart files every Ntout steps
    for s in range(0, int(prm["Nt"]) + 1, int(prm["Ntout"])):
        assert os.path.exists(f"{base}.{s}.rst")


def test_command_line_overrides(tmp_path):
    base, prob = _run(tmp_path, "vquad", ["-Nt", "20", "-Ntout", "10"])
    assert prob.params["Nt"] == 20 and prob.step == 20
    assert os.path.exists(base + ".10.rst") and os.path.exists(base + ".20.rst")
    out = formats.read_restart(base + ".rst", prob.mesh.ne)
    assert out.header["Step"] == 20 and abs(out.header["Time"] - 20 * prob.params["dt"]) < 1e-15
