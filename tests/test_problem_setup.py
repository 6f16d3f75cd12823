"""Host side of the pydgm-style driver (dgm_amd.problem), no GPU: how a run
directory's boundary files, restart initial condition, variable-order file and
channel file are read before a domain is built."""
from __future__ import annotations

import os
import shutil

import numpy as np
import pytest

from dgm_amd import dgx, formats
from dgm_amd import problem as dgm

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _copy(tmp_path, root, exts):
    for ext in exts:
        src = os.path.join(GOLD, root + ext)
        if os.path.exists(src):
            shutil.copy(src, tmp_path / (root + ext))
    return str(tmp_path / root)


@pytest.mark.parametrize("root,eqn", [("vquad", dgx.DGX_NAVIER_STOKES), ("vhex", dgx.DGX_EULER3D),
                                      ("cvquad", dgx.DGX_NAVIER_STOKES)])
def test_empty_bc_file_leaves_the_connectivity(tmp_path, root, eqn):
    # the reference's vquad/vhex/cvquad.bc hold "0 nbc": no boundary sides
    base = _copy(tmp_path, root, (".inp", ".msh", ".con", ".bc", ".ic"))
    assert formats.read_bc(base + ".bc") == []
    prob = dgm.ReoProblem(dgm.ArgList([base]))
    m = formats.read_mesh(base)
    kw = {}
    out = prob._boundaries(m, eqn, kw)
    assert kw == {} and np.array_equal(out.con, m.con)


def test_chan_boundaries_restart_and_forcing(tmp_path):
    base = _copy(tmp_path, "chan", (".inp", ".msh", ".con", ".bc", ".bct", ".ic", ".chn"))
    prob = dgm.ReoProblem(dgm.ArgList([base]))
    m = formats.read_mesh(base)
    kw = {}
    out = prob._boundaries(m, dgx.DGX_NAVIER_STOKES3D, kw)
    assert kw["wall"] == (1.0, 5.513081886073858e+03)
    bnd = out.con[out.con[:, 2] == -1]
    assert len(bnd) == 128 and np.all(bnd[:, 3] == dgx.DGX_BC_ISOTHERMAL)
    # the walls sit at z = 0 and z = 2 (sides 0 and 5 of the hexes there)
    assert set(bnd[:, 1].tolist()) == {0, 5}
    icd = formats.read_ic(base + ".ic")
    assert icd.code.lower() == "restart" and icd.file == "chan.ini"
    # for Euler3d the same Isothermal_Wall code is a slip wall (Euler3d::make_BC)
    kw3 = {}
    out3 = prob._boundaries(m, dgx.DGX_EULER3D, kw3)
    assert np.all(out3.con[out3.con[:, 2] == -1][:, 3] == dgx.DGX_BC_WALL) and "wall" not in kw3


def test_bc_types_reader():
    t = formats.read_bc_types(os.path.join(GOLD, "chan.bct"))
    assert t == {"T": ("isothermal_wall", [1.0, 5.513081886073858e+03]), "A": ("isoflux_wall", [0.0])}


def test_unknown_boundary_tag_raises(tmp_path):
    base = _copy(tmp_path, "chan", (".inp", ".msh", ".con", ".bc", ".ic"))
    with open(base + ".bct", "w") as f:
        f.write("Isoflux_Wall T\n{\n 0.0\n}\n")
    prob = dgm.ReoProblem(dgm.ArgList([base]))
    with pytest.raises(NotImplementedError):
        prob._boundaries(formats.read_mesh(base), dgx.DGX_NAVIER_STOKES3D, {})


def test_variable_order_without_a_path_raises(tmp_path):
    # a root.ord is read; where the library has no variable-order path (here
    # Navier-Stokes with viscosity) the run stops instead of taking the uniform p
    base = _copy(tmp_path, "vquad", (".inp", ".msh", ".con", ".bc", ".ic"))
    with open(base + ".ord", "w") as f:
        f.write("**ORDER DATA**\n0 4 0\n")
    prob = dgm.ReoProblem(dgm.ArgList([base]))
    with pytest.raises(dgx.DgxError, match="porder"):
        prob.setup()
