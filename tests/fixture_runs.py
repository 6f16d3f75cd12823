"""Drive a DGM regression case (root.inp/.msh/.con/.ic) through the oracle."""
from __future__ import annotations

import os

import numpy as np

from dgm_amd import formats, ic, mesh as meshgen
from oracle_lib import Oracle, VarOracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def initial_state(icd, x, nsd, gamma=1.4):
    return ic.from_code(icd, x, gamma)


def load_case(root):
    base = os.path.join(GOLDEN, root)
    prm = formats.read_inp(base + ".inp")
    mesh = formats.read_mesh(base)
    icd = formats.read_ic(base + ".ic")
    return prm, mesh, icd


def case_setup(root, prm, mesh):
    """Boundary sides, function boundaries and sources, and curved geometry of
    a case: returns (mesh with boundary rows, bc function, source function,
    nodal coordinates or None).  Function boundaries are the ``root.bc`` tags
    that ``root.bct`` defines as functions (FuncBC); the source is the first
    function of ``root.src`` (FuncSource).  Quads with a ``root.crv`` side,
    and meshes holding any non-parallelogram quad, take the curved path."""
    import dataclasses
    base = os.path.join(GOLDEN, root)
    bc_fn = src_fn = None
    rows = formats.read_bc(base + ".bc") if os.path.exists(base + ".bc") else []
    if rows:                                      # "0 nbc" files hold no boundary sides
        funcs = formats.read_functions(base + ".bct") if os.path.exists(base + ".bct") else {}
        tags = sorted({t for t, _, _ in rows})
        unknown = [t for t in tags if t not in funcs]
        if unknown or len(tags) > 1:
            raise ValueError(f"{root}: boundary tags {tags} other than one function boundary")
        bc_fn = formats.field_function(funcs[tags[0]])
        mesh = dataclasses.replace(mesh, con=formats.boundary_con(mesh.con, rows, {tags[0]: 2}))
    if os.path.exists(base + ".src"):
        src_fn = formats.field_function(next(iter(formats.read_functions(base + ".src").values())))
    p = int(prm["p"])
    xn = None
    if os.path.exists(base + ".crv"):             # curved sides (Domain::setup reads root.crv)
        xn, _ = meshgen.curved_quad_nodes(mesh, p, formats.read_crv(base + ".crv"))
    elif mesh.nsd == 2 and not np.all(meshgen.affine_quads(mesh)):
        xn = meshgen.quad_nodes(mesh, p)
    return mesh, bc_fn, src_fn, xn


def run_oracle_case(root, nsteps=None):
    """Run the case as Domain::solve does (restart writes every Ntout steps
    project the state, src/Domain.cpp:898-947) and return (oracle, modal state).
    A case with a root.ord runs on the oracle's variable-order quads (Euler;
    Navier-Stokes only without viscosity, whose diffusion term is then zero)."""
    prm, mesh, icd = load_case(root)
    p = int(prm["p"])
    mesh, bc_fn, src_fn, xn = case_setup(root, prm, mesh)
    if os.path.exists(os.path.join(GOLDEN, root + ".ord")):
        assert int(prm["eqntype"]) in (1, 2) and float(prm.get("vis", 0.0)) == 0.0 and xn is None
        orders = formats.read_ord(os.path.join(GOLDEN, root + ".ord"), mesh.ne, p)
        o = VarOracle(mesh, orders)
    else:
        o = Oracle(mesh, p, int(prm["eqntype"]), vis=float(prm.get("vis", 0.0)))
    if bc_fn is not None or src_fn is not None:
        o.set_functions(bc=bc_fn, src=src_fn)
    if xn is not None:
        o.set_nodes(xn)
    x = o.coords()
    o.set_state(initial_state(icd, x, mesh.nsd))
    o.project_state()
    Nt = int(prm["Nt"]) if nsteps is None else nsteps
    Ntout, dt, it = int(prm["Ntout"]), float(prm["dt"]), int(prm["inttype"])
    done = 0
    while done < Nt:
        n = min(Ntout, Nt - done)
        o.advance(it, dt, n)
        done += n
        o.project_state()
    return o, o.modal()
