"""Pin the CPU oracle against the reference's golden restart files.

Each case is run exactly as the reference's regression test does
(runs/vquad.tst, runs/vhex.tst, runs/vortex.tst, runs/cvquad.tst,
runs/shallow_mms/shallow_mms.tst) and compared with ``dgm_diff`` semantics
(src/Domain.cpp:1000-1071): every field must pass the absolute tolerance and,
where the reference norm exceeds repsM = sqrt(eps), the relative tolerance.
"""
import numpy as np
import pytest

from dgm_amd import formats
from fixture_runs import GOLDEN, run_oracle_case
from oracle_lib import dgm_diff

REPS_M = np.sqrt(np.finfo(float).eps)


def _check(root, aeps, reps):
    o, Uh = run_oracle_case(root)
    ref = formats.read_restart(f"{GOLDEN}/{root}.ref", o.ne)
    rn, dn = dgm_diff(o, ref.data, Uh)
    assert np.all(dn <= aeps), f"{root}: absolute diff {dn}"
    mask = rn > REPS_M
    rel = np.where(mask, dn / np.where(mask, rn, 1.0), 0.0)
    assert np.all(rel <= reps), f"{root}: relative diff {rel}"
    return dn, rel


def test_vquad_2d_ns_br1_tvd():
    # runs/vquad.tst:23-24 aeps = reps = 5e-12
    _check("vquad", 5e-12, 5e-12)


def test_vhex_3d_euler_lf_tvd():
    # runs/vhex.tst:23-24 aeps = 5e-13, reps = 5e-10
    _check("vhex", 5e-13, 5e-10)


def test_vortex_variable_order():
    """runs/vortex: 2D Navier-Stokes without viscosity (so the Euler residual)
    on 25 periodic quads, p = 5 with a band of 15 elements at p = 10
    (vortex.ord), RK4, 100 steps; sides between orders carry the larger
    order's Gauss points.  Tolerances of runs/vortex.tst: 5e-12 / 5e-12."""
    _check("vortex", 5e-12, 5e-12)


def test_cvquad_2d_ns_curved_arcs():
    # runs/cvquad.tst:22-23 aeps = reps = 5e-13; four quads with circular-arc
    # sides (runs/cvquad.crv), p = 10, so the curved metrics and the curved
    # mass-matrix projection are both on the path
    _check("cvquad", 5e-13, 5e-13)


def test_shallow_mms_function_boundaries_and_source():
    # runs/shallow_mms.tst:23-24 aeps = 1e-9, reps = 1e-10: shallow water on
    # 400 non-parallelogram quads (the curved path), p = 1, RK4, 3 steps, with
    # function boundaries (shallow_mms.bct) and a function source
    # (shallow_mms.src) evaluated at each stage's time, and the balance term as
    # the reference compiles it; measured 4e-13 absolute
    dn, rel = _check("shallow_mms", 1e-9, 1e-10)
    assert dn.max() < 1e-11


def test_function_file_parser():
    fns = formats.read_functions(f"{GOLDEN}/shallow_mms.bct")
    f = formats.field_function(fns["dansMMSFunction"])
    x, y, t = np.array([0.3, 1.2]), np.array([-0.5, 2.0]), 0.1
    a = np.sin(np.pi / 2 + t) * np.cos(x) * np.sin(y)
    want = np.stack([49.0 * a / 5.0 + 539.0 / 5.0, 49.0 * x * y * (a + 11.0) / (20.0 * np.pi ** 2),
                     -49.0 * x * y * (a + 11.0) / (20.0 * np.pi ** 2)])
    assert np.allclose(f(t, x, y), want, rtol=1e-15, atol=0)
    ev = lambda text: formats.field_function({"v": formats._ExprParser(formats._tokens(text)).expr()})(0.0, x, y)[0]
    assert np.all(ev("-2^2") == -4.0) and np.all(ev("2^3^2") == 512.0) and np.all(ev("2*3-4/2+1") == 5.0)
    assert np.allclose(ev("atan2(y, x) + sqrt(abs(-x))"), np.arctan2(y, x) + np.sqrt(x))
    with pytest.raises(ValueError):
        ev("__import__('os')")


def test_affine_detection():
    from dgm_amd import mesh as meshgen
    prm, m, _ = __import__("fixture_runs").load_case("shallow_mms")
    assert not np.any(meshgen.affine_quads(m))
    assert np.all(meshgen.affine_quads(meshgen.periodic_quads(3, 2, 0.0, 1.0, 0.0, 2.0)))


def test_crv_reader_and_arc_nodes():
    from dgm_amd import mesh as meshgen
    from fixture_runs import load_case
    types, sides = formats.read_crv(f"{GOLDEN}/cvquad.crv")
    assert types == {"arc": ("arc", [-4.0]), "arc2": ("arc", [4.0])}
    assert sides == [(0, 0, "arc"), (1, 0, "arc2"), (2, 2, "arc2"), (3, 2, "arc")]
    prm, m, _ = load_case("cvquad")
    p = int(prm["p"])
    q = p + 2
    xn, curved = meshgen.curved_quad_nodes(m, p, (types, sides))
    assert curved == [0, 1, 2, 3]
    # element 0, side 0 runs (0,0) -> (2,0) on a circle of radius 4; the
    # polynomial side of degree p = 10 follows it to the projection error
    e0 = xn[:, 0].reshape(2, q, q)
    sx, sy = e0[0, :, 0], e0[1, :, 0]
    dist = np.sqrt(16.0 - 1.0)
    centre = (1.0, -dist)                    # radius < 0 puts the centre outside: the side bows inward
    r = np.hypot(sx - centre[0], sy - centre[1])
    assert np.max(np.abs(r - 4.0)) < 1e-10
    sagitta = 4.0 - dist
    assert np.all(sy[1:-1] > 0.0) and 0.95 * sagitta < np.max(sy) <= sagitta
    # sides shared across the periodic seam coincide after the shift by 4
    e2 = xn[:, 2].reshape(2, q, q)
    assert np.allclose(e2[0, :, -1], sx, atol=1e-12) and np.allclose(e2[1, :, -1] - 4.0, sy, atol=1e-12)


def test_restart_reader_consumes_whole_files():
    for root, ne in (("vquad", 4), ("vhex", 8), ("nvortex", 25), ("vortex", 25), ("cvquad", 4), ("shallow_mms", 400)):
        r = formats.read_restart(f"{GOLDEN}/{root}.ref", ne)
        assert r.data.shape[1] == ne
        assert np.all(np.isfinite(r.data))


def _chan_oracle():
    """runs/chan as the reference runs it (runs/chan/chan.tst): 3D Navier-Stokes
    (BR1, LF), p = 3 on 512 hexes periodic in x and y, isothermal walls at
    z = 0 and 2 (``Isothermal_Wall T {Tw Eref}`` of chan.bct), the channel's bulk
    forcing (Channel::holdBulkMassFlow, src/Channel.hpp:117-163), restarted
    from chan.ini and stepped 100 TVD-RK3 steps with a restart write (a
    projection) every Ntout = 50 steps."""
    import dataclasses
    from oracle_lib import Oracle
    prm = formats.read_inp(f"{GOLDEN}/chan.inp")
    mesh = formats.read_mesh(f"{GOLDEN}/chan")
    rows = formats.read_bc(f"{GOLDEN}/chan.bc")
    assert {t for t, _, _ in rows} == {"T"}
    code, (tw, eref) = formats.read_bc_types(f"{GOLDEN}/chan.bct")["T"]
    assert code == "isothermal_wall"
    mesh = dataclasses.replace(mesh, con=formats.boundary_con(mesh.con, rows, {"T": 3}))
    p = int(prm["p"])
    o = Oracle(mesh, p, int(prm["eqntype"]), vis=float(prm["vis"]))
    o.set_isothermal(tw, eref)
    o.set_channel(True)
    ini = formats.read_restart(f"{GOLDEN}/chan.ini", mesh.ne)
    o.set_modal(ini.data)
    Nt, Ntout, dt = int(prm["Nt"]), int(prm["Ntout"]), float(prm["dt"])
    done = 0
    while done < Nt:
        n = min(Ntout, Nt - done)
        o.advance(int(prm["inttype"]), dt, n)
        done += n
        o.project_state()
    return o, o.modal()


def test_chan_3d_ns_isothermal_walls_bulk_forcing():
    # runs/chan/chan.tst: aeps = 5e-10, reps = 1e-8.  The inviscid wall flux
    # takes the slip ghost of Euler3d::WallBC (Domain::BC_db) and only the
    # viscous terms take the isothermal ghost (Navier_Stokes3d::BC_db); with
    # the isothermal ghost in both the difference is up to 7e-2
    o, Uh = _chan_oracle()
    ref = formats.read_restart(f"{GOLDEN}/chan.ref", o.ne)
    rn, dn = dgm_diff(o, ref.data, Uh)
    assert np.all(dn <= 5e-10), f"chan: absolute diff {dn}"
    mask = rn > REPS_M
    rel = np.where(mask, dn / np.where(mask, rn, 1.0), 0.0)
    assert np.all(rel <= 1e-8), f"chan: relative diff {rel}"
