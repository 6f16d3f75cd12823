"""pydgm-compatible driver: ``ReoProblem(args, table, comm).solve()`` on the GPU.

Mirrors the Python seam of the reference (python/dgm.pyx:64-207,
python/reo.py): a script written against pydgm

    problem = dgm.ReoProblem(dgm.ArgList(sys.argv[1:]), dgm.Table(), dgm.SerialComm())
    problem.solve()

runs unchanged with ``import dgm_amd.problem as dgm``.  ``solve`` follows
Domain::solve (src/Domain.cpp:883-998): parameters from ``root.inp`` with the
reference defaults (src/Domain.cpp:236-277) and command-line overrides
(src/Domain.cpp:369-530), mesh ``root.msh``/``root.con``, initial condition
``root.ic``; restart files ``root.<step>.rst`` at step 0 and every ``Ntout``
steps and ``root.rst`` at the end (src/Domain.cpp:1226-1250).  The residual
and the Runge-Kutta steps run in libdgx.so; the host only moves the state at
output steps.

Supported equation types are those of the GPU path: 1 Euler, 2
Navier_Stokes, 8 Euler3d, 9 Navier_Stokes3d (Reo::Problem::setup,
src/ReoProblem.cpp:55-161); others raise NotImplementedError.
"""
from __future__ import annotations

import math
import os
import sys
import time
from typing import Dict, List, Optional

import numpy as np

from . import dgx, formats, ic
from . import mesh as meshgen

DGM_SUCCESS = 0

_DEFAULTS: Dict[str, object] = {
    "inttype": 3, "p": 5, "Nt": 100, "Ntout": 50, "ntout": 1, "dt": 0.01,
    "gamma": 1.4, "vis": 0.0, "Pr": 0.72, "flux": 0, "precision": 0,
}

# command-line options that take one value (src/Domain.cpp:375-470)
_OPTIONS = {"-p": int, "-dt": float, "-Nt": int, "-Ntout": int, "-ntout": int,
            "-inttype": int, "-vis": float}

_FIELDS = {2: ["r", "ru", "rv", "E"], 3: ["r", "ru", "rv", "rw", "E"]}


class SerialComm:
    """Single-rank communicator (python/dgm.pyx SerialComm)."""

    def MyPID(self) -> int:
        return 0

    def NumProc(self) -> int:
        return 1

    def Master(self) -> bool:
        return True

    def exit(self, code: int = DGM_SUCCESS):
        sys.exit(code)

    def error(self, msg: str, code: int = 1):
        raise RuntimeError(msg)


class Table(dict):
    """Parameter table (DGM::Table): values set here override ``root.inp``."""

    def size(self) -> int:
        return len(self)


class ArgList(list):
    """Argument list (DGM::ArgList): ``[options...] root``."""

    def size(self) -> int:
        return len(self)


class ReoProblem:
    """Forward problem of a DGM run directory, solved on one GPU."""

    def __init__(self, args, table: Optional[Table] = None, comm=None, out=None):
        self.args = ArgList(args)
        self.table = Table() if table is None else table
        self.comm = comm or SerialComm()
        self.out = out or sys.stdout
        if self.args.size() == 0 or str(self.args[-1]).startswith("-"):
            self.args.append("root")
        self.root = str(self.args[-1])
        self.params = self._parameters()
        self.domain: Optional[dgx.GPUDomain] = None
        self.step = 0
        self.t = 0.0

    def _parameters(self) -> Dict[str, object]:
        prm = dict(_DEFAULTS)
        if os.path.exists(self.root + ".inp"):
            prm.update(formats.read_inp(self.root + ".inp"))
        else:
            print(f"WARNING:  using defaults as {self.root}.inp file does not exist", file=self.out)
        a = list(self.args[:-1])
        i = 0
        while i < len(a):
            name = a[i]
            if name in _OPTIONS and i + 1 < len(a):
                prm[name[1:]] = _OPTIONS[name](a[i + 1])
                i += 2
            else:
                i += 1
        prm.update(self.table)
        return prm

    def _boundaries(self, mesh, kw):
        """root.bc sides whose tag root.bct defines as a function (FuncBC),
        and the first function of root.src as the source (FuncSource)."""
        import dataclasses
        if os.path.exists(self.root + ".bc"):
            rows = formats.read_bc(self.root + ".bc")
            funcs = formats.read_functions(self.root + ".bct") if os.path.exists(self.root + ".bct") else {}
            tags = sorted({t for t, _, _ in rows})
            if any(t not in funcs for t in tags) or len(tags) > 1:
                raise NotImplementedError(f"Reo::Problem: boundary tags {tags}: one function boundary "
                                          "defined in root.bct is read from run directories")
            kw["bc_function"] = funcs[tags[0]]
            mesh = dataclasses.replace(mesh, con=formats.boundary_con(mesh.con, rows,
                                                                      {tags[0]: dgx.DGX_BC_FUNCTION}))
        if os.path.exists(self.root + ".src"):
            kw["source"] = next(iter(formats.read_functions(self.root + ".src").values()))
        return mesh

    def setup(self) -> dgx.GPUDomain:
        prm = self.params
        eqn = int(prm.get("eqntype", 0))
        if eqn not in (dgx.DGX_EULER, dgx.DGX_NAVIER_STOKES, dgx.DGX_EULER3D,
                       dgx.DGX_NAVIER_STOKES3D, dgx.DGX_SHALLOW):
            raise NotImplementedError(f"Reo::Problem: equation type {eqn} has no GPU path")
        p = int(prm["p"])
        kw = {}
        mesh = self._boundaries(formats.read_mesh(self.root), kw)
        self.xnodes = None
        if os.path.exists(self.root + ".crv"):     # curved sides (Field::read_curve)
            if mesh.nsd != 2:
                raise NotImplementedError("Reo::Problem: root.crv is read for quads; pass hex nodes as xnodes")
            self.xnodes, _ = meshgen.curved_quad_nodes(mesh, p, formats.read_crv(self.root + ".crv"))
        elif mesh.nsd == 2 and not np.all(meshgen.affine_quads(mesh)):
            self.xnodes = meshgen.quad_nodes(mesh, p)   # what the library does by itself, kept for the caller
        self.mesh = mesh
        self.eqn = eqn
        vis = float(prm["vis"]) if eqn in (dgx.DGX_NAVIER_STOKES, dgx.DGX_NAVIER_STOKES3D) else 0.0
        dom = dgx.GPUDomain(mesh, p, eqn, gamma=float(prm["gamma"]), vis=vis,
                            Pr=float(prm["Pr"]), flux=int(prm["flux"]),
                            precision=int(prm["precision"]), xnodes=self.xnodes, **kw)
        dom.set_time(self.t)
        icd = formats.read_ic(self.root + ".ic")
        x = dom.coords()
        dom.upload(ic.from_code(icd, x, float(prm["gamma"])))
        self.domain = dom
        return dom

    def _header(self) -> Dict[str, object]:
        p = self.params
        return {"Name": os.path.basename(self.root), "Nsd": self.mesh.nsd,
                "Lmax": int(p["p"]) + 1, "Step": self.step, "Time": self.t,
                "Time step": float(p["dt"]), "Viscosity": float(p["vis"]),
                "Field names": ["phi", "uphi", "vphi"] if self.eqn == dgx.DGX_SHALLOW else _FIELDS[self.mesh.nsd]}

    def write(self, fname: str) -> None:
        """Restart file of the current state (modal, JSON header)."""
        formats.write_restart(fname, self.domain.download(dgx.DGX_MODAL), self._header())

    def solve(self) -> int:
        if self.domain is None:
            self.setup()
        prm, dom = self.params, self.domain
        Nt, Ntout, ntout = int(prm["Nt"]), int(prm["Ntout"]), max(1, int(prm["ntout"]))
        dt, inttype = float(prm["dt"]), int(prm["inttype"])
        if self.step == 0:
            self.write(f"{self.root}.{self.step}.rst")
        print(f"Advancing solution {Nt} time steps", file=self.out)
        print("    Step      t           dt         CFL         dCPU", file=self.out)
        chunk = ntout if Ntout == 0 else math.gcd(ntout, Ntout)
        wall = time.perf_counter()
        done = 0
        while done < Nt:
            n = min(chunk, Nt - done)
            t0 = time.perf_counter()
            dom.advance(inttype, dt, n)
            dom.synchronize()
            el = (time.perf_counter() - t0) / n
            done += n
            self.step += n
            self.t += n * dt
            if done % ntout == 0 or done == Nt:
                print(f"{self.step:8d}{self.t:12.4e}{dt:12.4e}{0.0:12.4e}{el:12.4e}", file=self.out)
            if Ntout == 0 or self.step % Ntout == 0:
                self.write(f"{self.root}.{self.step}.rst")
        print(f"Finished solve in {time.perf_counter() - wall:.4e} sec", file=self.out)
        self.write(self.root + ".rst")
        return DGM_SUCCESS

    def close(self) -> None:
        if self.domain is not None:
            self.domain.close()
            self.domain = None


def main(argv: List[str]) -> int:
    """``python -m dgm_amd.problem [options] root`` (python/reo.py)."""
    if len(argv) < 1:
        print("dgm: not enough arguments", file=sys.stderr)
        return 1
    prob = ReoProblem(ArgList(argv), Table(), SerialComm())
    try:
        return prob.solve()
    finally:
        prob.close()


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
