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
    "gamma": 1.4, "vis": 0.0, "Pr": 0.72, "flux": 0, "precision": 0, "CFLmax": 0.0, "dtmax": 0.0,
}

# command-line options that take one value (src/Domain.cpp:375-470)
_OPTIONS = {"-p": int, "-dt": float, "-Nt": int, "-Ntout": int, "-ntout": int,
            "-inttype": int, "-vis": float, "-CFLmax": float, "-dtmax": float}

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


class vField:
    """Sizes and contents of a domain's conserved state, as pydgm's vField
    binds DGM::vField (python/dgm.pyx:297-318; src/vField.hpp:330-376).  The
    data stay on the GPU; ``numpy`` copies them to the host."""

    def __init__(self, domain: "dgx.GPUDomain"):
        self.domain = domain
        self._s = domain.vfield_sizes()

    def size(self) -> int:
        return self._s["size"]

    def qtot(self) -> int:
        return self._s["qtot"]

    def nmodes(self) -> int:
        return self._s["nmodes"]

    def ndof(self) -> int:
        return self._s["nmodes"]        # modal bases: vField::ndof is nmodes

    def sdof(self) -> int:
        return self._s["sdof"]

    def nsd(self) -> int:
        return self._s["nsd"]

    def zero(self) -> None:
        self.domain.upload(np.zeros((self.domain.nf, self.domain.ne, self.domain.nm)), dgx.DGX_MODAL)

    def numpy(self, layout: int = dgx.DGX_NODAL) -> np.ndarray:
        return self.domain.download(layout)


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
        self.dt = None       # the step Domain::set_time_step last chose (CFLmax runs)

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

    def _bc_database(self, eqn):
        """Boundary types by tag, as each domain's setup_BCdb builds its BC
        database: the defaults, then root.bct (Euler::setup_BCdb
        src/Euler.cpp:1199-1215, Euler3d::setup_BCdb src/Euler3d.cpp:824-841,
        Navier_Stokes3d::setup_BCdb src/Navier_Stokes3d.cpp:398-416,
        Shallow::setup_BCdb src/Shallow.cpp:745-760).  Returns {tag: (type,
        parameters)}.  Every code containing "wall" is a slip wall for Euler,
        Euler3d and Shallow (Euler3d::make_BC, src/Euler3d.cpp:857-860); for
        Navier_Stokes3d an Isothermal_Wall is isothermal, as its own make_BC
        reads it (src/Navier_Stokes3d.cpp:438-441)."""
        ma03 = 20.34126984126985                  # Ma = 0.3 default state energy
        db: Dict[str, tuple] = {}
        if eqn == dgx.DGX_EULER:
            db = {"S": (dgx.DGX_BC_STATE, [1.0, 1.0, 0.0, ma03]), "W": (dgx.DGX_BC_WALL, [])}
        elif eqn == dgx.DGX_EULER3D:
            db = {"S": (dgx.DGX_BC_STATE, [1.0, 1.0, 0.0, 0.0, ma03]), "W": (dgx.DGX_BC_WALL, [])}
        elif eqn == dgx.DGX_NAVIER_STOKES3D:
            db = {"S": (dgx.DGX_BC_STATE, [1.0, 1.0, 0.0, 0.0, ma03]),
                  "W": (dgx.DGX_BC_ISOTHERMAL, [1.0, ma03])}
        elif eqn == dgx.DGX_SHALLOW:
            db = {"W": (dgx.DGX_BC_WALL, [])}
        bct = self.root + ".bct"
        if os.path.exists(bct):
            for tag, comps in formats.read_functions(bct).items():
                db[tag] = (dgx.DGX_BC_FUNCTION, comps)
            for tag, (code, vals) in formats.read_bc_types(bct).items():
                if code == "state":
                    db[tag] = (dgx.DGX_BC_STATE, vals)
                elif eqn == dgx.DGX_NAVIER_STOKES3D and code == "isothermal_wall":
                    db[tag] = (dgx.DGX_BC_ISOTHERMAL, vals[:2])
                elif "wall" in code and eqn != dgx.DGX_NAVIER_STOKES3D:
                    db[tag] = (dgx.DGX_BC_WALL, [])
                else:
                    db[tag] = (None, code)
        return db

    def _boundaries(self, mesh, eqn, kw):
        """Boundary sides of root.bc by the type their tag has in the domain's
        BC database (state, slip wall, isothermal wall, function), and the
        first function of root.src as the source (FuncSource).  A root.bc
        without rows (``0 nbc``) leaves every side to root.con."""
        import dataclasses
        if os.path.exists(self.root + ".bc"):
            rows = formats.read_bc(self.root + ".bc")
            if rows:
                db = self._bc_database(eqn)
                types = {}
                states, walls, fns = set(), set(), set()
                for t in sorted({t for t, _, _ in rows}):
                    if t not in db or db[t][0] is None:
                        what = db[t][1] if t in db else "not defined"
                        raise NotImplementedError(f"Reo::Problem: boundary tag {t} ({what}) has no GPU path")
                    kind, prm = db[t]
                    types[t] = kind
                    if kind == dgx.DGX_BC_STATE:
                        states.add(tuple(prm))
                    elif kind == dgx.DGX_BC_ISOTHERMAL:
                        walls.add(tuple(prm))
                    elif kind == dgx.DGX_BC_FUNCTION:
                        fns.add(t)
                if len(states) > 1 or len(walls) > 1 or len(fns) > 1:
                    raise NotImplementedError("Reo::Problem: one state, one isothermal wall and one function "
                                              "boundary type per domain")
                if states:
                    kw["bc_state"] = list(next(iter(states)))
                if walls:
                    kw["wall"] = next(iter(walls))
                if fns:
                    kw["bc_function"] = db[next(iter(fns))][1]
                mesh = dataclasses.replace(mesh, con=formats.boundary_con(mesh.con, rows, types))
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
        mesh = self._boundaries(formats.read_mesh(self.root), eqn, kw)
        self.porder = None
        if os.path.exists(self.root + ".ord"):
            # variable order per element (Field::Field, src/Field.cpp:2290-2320):
            # the domain's p is the largest; the library rejects what it has
            # no variable-order path for (curved, viscous, 3D, functions)
            self.porder = formats.read_ord(self.root + ".ord", mesh.ne, p)
            p = int(self.porder.max())
            kw["porder"] = self.porder
        if os.path.exists(self.root + ".chn"):   # Channel::exists (src/Euler3d.cpp:75-77)
            kw["bulk_forcing"] = True
        self.xnodes = None
        crv = formats.read_crv(self.root + ".crv") if os.path.exists(self.root + ".crv") else None
        if crv is not None and crv[1]:             # curved sides (Field::read_curve)
            if mesh.nsd != 2:
                raise NotImplementedError("Reo::Problem: root.crv is read for quads; pass hex nodes as xnodes")
            self.xnodes, _ = meshgen.curved_quad_nodes(mesh, p, crv)
        elif mesh.nsd == 2 and not np.all(meshgen.affine_quads(mesh)):
            self.xnodes = meshgen.quad_nodes(mesh, p)   # what the library does by itself, kept for the caller
        self.mesh = mesh
        self.eqn = eqn
        vis = float(prm["vis"]) if eqn in (dgx.DGX_NAVIER_STOKES, dgx.DGX_NAVIER_STOKES3D) else 0.0
        dom = dgx.GPUDomain(mesh, p, eqn, gamma=float(prm["gamma"]), vis=vis,
                            Pr=float(prm["Pr"]), flux=int(prm["flux"]),
                            precision=int(prm["precision"]), xnodes=self.xnodes, **kw)
        icd = formats.read_ic(self.root + ".ic")
        if icd.code.lower() == "restart":
            # Euler3d::set_ic "Restart" (src/Euler3d.cpp:957-965): the state, step
            # and time of a restart file (Domain::read_header, src/Domain.cpp:1652-1653)
            path = os.path.join(os.path.dirname(self.root) or ".", icd.file)
            rst = formats.read_restart(path, mesh.ne)
            if rst.data.shape[2] != dom.nm:
                raise ValueError(f"{path}: {rst.data.shape[2]} modes per element, the domain has {dom.nm}")
            self.step = int(rst.header.get("Step", 0))
            self.t = float(rst.header.get("Time", 0.0))
            dom.set_time(self.t)
            dom.upload(rst.data, dgx.DGX_MODAL)
        else:
            dom.set_time(self.t)
            x = dom.coords()
            dom.upload(ic.from_code(icd, x, float(prm["gamma"])))
        self.domain = dom
        return dom

    def getState(self) -> vField:
        """The conserved state as a vField (sizes; host copies on request)."""
        if self.domain is None:
            self.setup()
        return vField(self.domain)

    def _header(self) -> Dict[str, object]:
        p = self.params
        pmax = int(p["p"]) if self.porder is None else int(self.porder.max())
        return {"Name": os.path.basename(self.root), "Nsd": self.mesh.nsd,
                "Lmax": pmax + 1, "Step": self.step, "Time": self.t,
                "Time step": float(self.dt if self.dt is not None else p["dt"]), "Viscosity": float(p["vis"]),
                "Field names": ["phi", "uphi", "vphi"] if self.eqn == dgx.DGX_SHALLOW else _FIELDS[self.mesh.nsd]}

    def write(self, fname: str) -> None:
        """Restart file of the current state (modal, JSON header); with
        variable order each element keeps its own number of modes."""
        nmodes = None if self.porder is None else (self.porder.astype(np.int64) + 1) ** self.mesh.nsd
        formats.write_restart(fname, self.domain.download(dgx.DGX_MODAL), self._header(), nmodes=nmodes)

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
        # Domain::solve calls set_time_step before every step (src/Domain.cpp:931):
        # with CFLmax the step follows the CFL condition and the run goes step by
        # step; otherwise dt is fixed and the CFL number is evaluated only for
        # the steps whose line is printed
        CFLmax, dtmax = float(prm["CFLmax"]), float(prm["dtmax"])
        chunk = 1 if CFLmax != 0 else (ntout if Ntout == 0 else math.gcd(ntout, Ntout))
        wall = time.perf_counter()
        done = 0
        cfl = 0.0
        while done < Nt:
            n = min(chunk, Nt - done)
            printed = (done + n) % ntout == 0 or done + n == Nt
            t0 = time.perf_counter()
            if CFLmax != 0:
                cfl, dt = dom.set_time_step(dt, CFLmax, dtmax)
                self.dt = dt
                dom.advance(inttype, dt, n)
            elif printed:
                if n > 1:
                    dom.advance(inttype, dt, n - 1)
                cfl, _ = dom.set_time_step(dt)
                dom.advance(inttype, dt, 1)
            else:
                dom.advance(inttype, dt, n)
            dom.synchronize()
            el = (time.perf_counter() - t0) / n
            done += n
            self.step += n
            self.t = dom.time
            if done % ntout == 0 or done == Nt:
                print(f"{self.step:8d}{self.t:12.4e}{dt:12.4e}{cfl:12.4e}{el:12.4e}", file=self.out)
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
