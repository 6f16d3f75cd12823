"""Drive a DGM regression case (root.inp/.msh/.con/.ic) through the oracle."""
from __future__ import annotations

import os

import numpy as np

from dgm_amd import formats, ic, mesh as meshgen
from oracle_lib import Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def initial_state(icd, x, nsd, gamma=1.4):
    return ic.from_code(icd, x, gamma)


def load_case(root):
    base = os.path.join(GOLDEN, root)
    prm = formats.read_inp(base + ".inp")
    mesh = formats.read_mesh(base)
    icd = formats.read_ic(base + ".ic")
    return prm, mesh, icd


def run_oracle_case(root, nsteps=None):
    """Run the case as Domain::solve does (restart writes every Ntout steps
    project the state, src/Domain.cpp:898-947) and return (oracle, modal state)."""
    prm, mesh, icd = load_case(root)
    p = int(prm["p"])
    o = Oracle(mesh, p, int(prm["eqntype"]), vis=float(prm.get("vis", 0.0)))
    crv = os.path.join(GOLDEN, root + ".crv")
    if os.path.exists(crv):                       # curved sides (Domain::setup reads root.crv)
        xn, _ = meshgen.curved_quad_nodes(mesh, p, formats.read_crv(crv))
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
