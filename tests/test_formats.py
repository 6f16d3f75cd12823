"""DGM file formats and the periodic mesh generators (no GPU)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from dgm_amd import formats, mesh as meshgen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("root,nsd,p", [("vquad", 2, 10), ("vhex", 3, 10)])
def test_run_directory_reads(root, nsd, p):
    prm = formats.read_inp(os.path.join(GOLD, root + ".inp"))
    assert prm["p"] == p
    m = formats.read_mesh(os.path.join(GOLD, root))
    assert m.nsd == nsd
    nv = 4 if nsd == 2 else 8
    assert m.vert.shape == (m.ne, nsd, nv)
    ns = 4 if nsd == 2 else 6
    # every .con row names an existing element side, and links are symmetric
    links = {(int(r[0]), int(r[1])): (int(r[2]), int(r[3])) for r in m.con}
    for (e, s), (le, ls) in links.items():
        assert 0 <= e < m.ne and 0 <= s < ns and 0 <= le < m.ne and 0 <= ls < ns
        assert links.get((le, ls)) == (e, s)
    ref = formats.read_restart(os.path.join(GOLD, root + ".ref"), m.ne)
    assert ref.data.shape == (nsd + 2, m.ne, (p + 1) ** nsd)
    assert np.all(np.isfinite(ref.data))


def test_restart_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    data = rng.standard_normal((5, 7, 27))
    path = str(tmp_path / "x.rst")
    formats.write_restart(path, data, {"Nsd": 3, "Step": 12, "Time": 0.25,
                                       "Field names": ["r", "ru", "rv", "rw", "E"]})
    back = formats.read_restart(path, 7)
    assert np.array_equal(back.data, data)
    assert back.header["Step"] == 12 and back.header["Nel"] == 7
    assert list(back.nmodes) == [27] * 7
    with open(path, "ab") as f:                         # trailing bytes are an error
        f.write(b"\0")
    with pytest.raises(ValueError):
        formats.read_restart(path, 7)


def test_old_header_restart_reads():
    m = formats.read_mesh(os.path.join(GOLD, "vquad"))
    ref = formats.read_restart(os.path.join(GOLD, "vquad.ref"), m.ne)
    assert ref.header["Nel"] == m.ne and ref.header["Nsd"] == 2


@pytest.mark.parametrize("n", [(2, 3, 4), (3, 3, 3), (1, 2, 2)])
def test_periodic_hexes_connectivity(n):
    nx, ny, nz = n
    lo, hi = (0.0, -1.0, 2.0), (2.0, 1.0, 5.0)
    m = meshgen.periodic_hexes(nx, ny, nz, lo, hi)
    assert m.ne == nx * ny * nz and len(m.con) == 6 * m.ne
    links = {(int(r[0]), int(r[1])): (int(r[2]), int(r[3])) for r in m.con}
    assert len(links) == 6 * m.ne
    L = np.subtract(hi, lo)
    # hex side -> vertex ids (bottom 0-3, top 4-7)
    faces = [(0, 1, 2, 3), (0, 1, 5, 4), (1, 2, 6, 5), (3, 2, 6, 7), (0, 3, 7, 4), (4, 5, 6, 7)]
    for (e, s), (le, ls) in links.items():
        assert links[(le, ls)] == (e, s)
        c0 = m.vert[e][:, faces[s]].mean(axis=1)
        c1 = m.vert[le][:, faces[ls]].mean(axis=1)
        d = (c0 - c1) / L
        assert np.allclose(d - np.round(d), 0.0)       # same face modulo the period
    vol = sum(np.prod(m.vert[e].max(axis=1) - m.vert[e].min(axis=1)) for e in range(m.ne))
    assert np.isclose(vol, np.prod(L))


def test_periodic_quads_connectivity():
    m = meshgen.periodic_quads(3, 2, 0.0, 3.0, 0.0, 1.0)
    links = {(int(r[0]), int(r[1])): (int(r[2]), int(r[3])) for r in m.con}
    edges = [(0, 1), (1, 2), (2, 3), (3, 0)]
    for (e, s), (le, ls) in links.items():
        assert links[(le, ls)] == (e, s)
        c0 = m.vert[e][:, edges[s]].mean(axis=1)
        c1 = m.vert[le][:, edges[ls]].mean(axis=1)
        d = (c0 - c1) / np.array([3.0, 1.0])
        assert np.allclose(d - np.round(d), 0.0)


@pytest.mark.parametrize("p", [2, 5, 8])
def test_quad_nodes_match_oracle(p):
    from oracle_lib import Oracle
    m = meshgen.periodic_quads(3, 2, 0.0, 3.0, -1.0, 1.5)
    x = meshgen.quad_nodes(m, p)
    ref = Oracle(m, p, 1).coords()
    assert np.allclose(x, ref, rtol=0, atol=1e-13)


def test_variable_order_restart_round_trip(tmp_path):
    """A variable-order restart (per-element nModes, as runs/vortex.ref) reads
    into padded slots and writes back to the same bytes of data."""
    ref = formats.read_restart(f"{GOLDEN}/vortex.ref", 25)
    assert set(ref.nmodes.tolist()) == {36, 121} and ref.data.shape == (4, 25, 121)
    for e in range(25):
        assert np.all(ref.data[:, e, ref.nmodes[e]:] == 0.0)
    out = tmp_path / "v.rst"
    formats.write_restart(str(out), ref.data, {"Step": 100}, nmodes=ref.nmodes)
    back = formats.read_restart(str(out), 25)
    assert np.array_equal(back.nmodes, ref.nmodes) and np.array_equal(back.data, ref.data)


def test_order_file():
    orders = formats.read_ord(f"{GOLDEN}/vortex.ord", 25, 5)
    assert orders.tolist() == [5, 10, 10, 10, 5] * 5
