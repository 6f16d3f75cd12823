"""Periodic structured meshes in DGM's .msh/.con conventions.

Element numbering is lexicographic (x fastest).  Vertex order follows the
reference's quad (counter-clockwise from the lower-left corner) and hex
(bottom face 0-3, top face 4-7) conventions (src/BaseQuad.cpp:20-27,
src/Field.cpp:316-330); side numbering is that of Quad (0 bottom, 1 right,
2 top, 3 left) and Hex (0 back z-, 1 y-, 2 x+, 3 y+, 4 x-, 5 front z+).
Connectivity rows list every side of every element in element order, which
makes the lower-numbered element of a pair the master (src/Field.cpp:3478-3556),
as the reference's periodic generator does (src/Mesh.cpp:538-541, 879-884).
"""
from __future__ import annotations

import numpy as np

from .formats import Mesh


def periodic_quads(nx, ny, x0=0.0, x1=1.0, y0=0.0, y1=1.0) -> Mesh:
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    vert = np.zeros((nx * ny, 2, 4))
    con = []
    for j in range(ny):
        for i in range(nx):
            e = i + nx * j
            vert[e, 0] = [xs[i], xs[i + 1], xs[i + 1], xs[i]]
            vert[e, 1] = [ys[j], ys[j], ys[j + 1], ys[j + 1]]
            nb = [(i + nx * ((j - 1) % ny), 2), ((i + 1) % nx + nx * j, 3),
                  (i + nx * ((j + 1) % ny), 0), ((i - 1) % nx + nx * j, 1)]
            for s, (le, ls) in enumerate(nb):
                con.append((e, s, le, ls))
    return Mesh(nsd=2, vert=vert, con=np.asarray(con, dtype=np.int32))


def box_quads(nx, ny, x0=0.0, x1=1.0, y0=0.0, y1=1.0, bc=(None, None, None, None)) -> Mesh:
    """Quads on [x0, x1] x [y0, y1]; bc[k] is the boundary type of the bottom
    (k=0), right (1), top (2) and left (3) edge, or None for periodic in that
    direction (bottom/top and left/right must agree on None).  Boundary rows
    are (eid, sid, -1, type): 0 wall, 1 state, 2 function of (x, y, t)."""
    if (bc[0] is None) != (bc[2] is None) or (bc[1] is None) != (bc[3] is None):
        raise ValueError("opposite edges must both be periodic or both be boundaries")
    m = periodic_quads(nx, ny, x0, x1, y0, y1)
    con = m.con.copy()
    for r in range(len(con)):
        e, s = int(con[r, 0]), int(con[r, 1])
        i, j = e % nx, e // nx
        edge = {0: j == 0, 1: i == nx - 1, 2: j == ny - 1, 3: i == 0}[s]
        if edge and bc[s] is not None:
            con[r, 2], con[r, 3] = -1, bc[s]
    return Mesh(nsd=2, vert=m.vert, con=con)


def periodic_hexes(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)) -> Mesh:
    xs = np.linspace(lo[0], hi[0], nx + 1)
    ys = np.linspace(lo[1], hi[1], ny + 1)
    zs = np.linspace(lo[2], hi[2], nz + 1)
    ne = nx * ny * nz
    vert = np.zeros((ne, 3, 8))
    con = np.zeros((ne * 6, 4), dtype=np.int32)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ii, jj, kk = (a.transpose(2, 1, 0).ravel() for a in (ii, jj, kk))   # x fastest
    e = ii + nx * (jj + ny * kk)
    vx = np.stack([xs[ii], xs[ii + 1], xs[ii + 1], xs[ii]] * 2, axis=1)
    vy = np.stack([ys[jj], ys[jj], ys[jj + 1], ys[jj + 1]] * 2, axis=1)
    vz = np.stack([zs[kk]] * 4 + [zs[kk + 1]] * 4, axis=1)
    vert[e, 0], vert[e, 1], vert[e, 2] = vx, vy, vz

    def eid(i, j, k):
        return (i % nx) + nx * ((j % ny) + ny * (k % nz))

    nbrs = [(eid(ii, jj, kk - 1), 5), (eid(ii, jj - 1, kk), 3), (eid(ii + 1, jj, kk), 4),
            (eid(ii, jj + 1, kk), 1), (eid(ii - 1, jj, kk), 2), (eid(ii, jj, kk + 1), 0)]
    for s, (le, ls) in enumerate(nbrs):
        rows = e * 6 + s
        con[rows, 0] = e
        con[rows, 1] = s
        con[rows, 2] = le
        con[rows, 3] = ls
    return Mesh(nsd=3, vert=vert, con=con)


def channel_hexes(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), bc_bottom=0, bc_top=1) -> Mesh:
    """Hexes periodic in x and y with boundary sides on z = lo[2] (side 0 of
    the bottom layer) and z = hi[2] (side 5 of the top layer).  Boundary rows
    are (eid, sid, -1, bc type): 0 slip wall (Euler3d::WallBC,
    src/Euler3d.cpp:876-909), 1 state (StateBC, src/BC.cpp:17-28)."""
    m = periodic_hexes(nx, ny, nz, lo, hi)
    con = m.con.copy()
    for r in range(len(con)):
        e, s = int(con[r, 0]), int(con[r, 1])
        k = e // (nx * ny)
        if s == 0 and k == 0:
            con[r, 2], con[r, 3] = -1, bc_bottom
        elif s == 5 and k == nz - 1:
            con[r, 2], con[r, 3] = -1, bc_top
    return Mesh(nsd=3, vert=m.vert, con=con)


def gll_points(q):
    """The q Gauss-Lobatto-Legendre points on [-1, 1], ascending: +-1 and the
    roots of P'_{q-1} (src/Polylib.cpp zwgll)."""
    from numpy.polynomial import legendre
    inner = legendre.Legendre.basis(q - 1).deriv().roots() if q > 2 else np.array([])
    return np.concatenate([[-1.0], np.sort(inner.real), [1.0]])


def quad_nodes(mesh: Mesh, p: int) -> np.ndarray:
    """Nodal coordinates [2][elem][q*q] of affine quads at the q = p+2 GLL
    points, node i*q + j at (xi_i, eta_j), bilinear vertex map with vertices
    counter-clockwise from (-1, -1) (src/Quad.cpp:993-1013)."""
    z = gll_points(p + 2)
    xi, eta = np.meshgrid(z, z, indexing="ij")
    xi, eta = xi.ravel(), eta.ravel()
    N = np.stack([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)]) / 4
    return np.einsum("edv,vq->deq", mesh.vert, N)


def hex_nodes(mesh: Mesh, p: int) -> np.ndarray:
    """Nodal coordinates [3][elem][q^3] of hexes at the q = p+2 GLL points,
    node (i*q + j)*q + k at (xi_i, eta_j, zeta_k), trilinear vertex map (bottom
    face 0-3 counter-clockwise, top face 4-7; src/Hex.cpp:817-866)."""
    z = gll_points(p + 2)
    a, b, c = np.meshgrid(z, z, z, indexing="ij")
    a, b, c = a.ravel(), b.ravel(), c.ravel()
    sa = [-1, 1, 1, -1, -1, 1, 1, -1]
    sb = [-1, -1, 1, 1, -1, -1, 1, 1]
    sc = [-1, -1, -1, -1, 1, 1, 1, 1]
    N = np.stack([(1 + sa[n] * a) * (1 + sb[n] * b) * (1 + sc[n] * c) / 8 for n in range(8)])
    return np.einsum("edv,vq->deq", mesh.vert, N)


def mapped_nodes(mesh: Mesh, p: int, fn) -> np.ndarray:
    """Curved elements: the GLL nodes of the affine mesh moved by fn(x, y, z)
    -> (x', y', z') (hexes) or fn(x, y) -> (x', y') (quads), for the
    curved-element path (dgx_domain_desc.xnodes)."""
    x = hex_nodes(mesh, p) if mesh.nsd == 3 else quad_nodes(mesh, p)
    return np.stack(fn(*x))


def affine_quads(mesh: Mesh) -> np.ndarray:
    """Per quad, whether it is a parallelogram within the reference's
    tolerances (Quad::is_affine, src/Quad.cpp:1184-1205: each half edge vector
    equal to its opposite within REL = sqrt(eps) or ABS = eps^(2/3)).  The
    reference gives every other quad a straight-sided curve, so it takes the
    curved path with nodal metrics (Field::read_curve_pc, src/Field.cpp:1226-1240)."""
    eps = np.finfo(float).eps
    rel, tol = np.sqrt(eps), eps ** 0.6666666666666666667

    def close(u, v):
        d = np.abs(u - v)
        return (u == v) | (d <= tol) | (d <= rel * np.maximum(np.abs(u), np.abs(v)))

    v = np.asarray(mesh.vert, dtype=np.float64)           # [ne][2][4]
    out = np.ones(v.shape[0], dtype=bool)
    for c in range(2):
        a = v[:, c]
        out &= close((a[:, 1] - a[:, 0]) * 0.5, (a[:, 2] - a[:, 3]) * 0.5)
        out &= close((a[:, 3] - a[:, 0]) * 0.5, (a[:, 2] - a[:, 1]) * 0.5)
    return out


def _arc_points(p0, p1, radius, orientation, z):
    """Points of a circular arc through p0 and p1 at the local coordinates z,
    as Arc::generate (src/Curve.cpp:523-590) places them."""
    arclength = np.hypot(p1[0] - p0[0], p1[1] - p0[1])
    if abs(arclength / radius) > 2.0:
        raise ValueError("arc radius too small for its side")
    pm = ((p0[0] + p1[0]) * 0.5, (p0[1] + p1[1]) * 0.5)
    dist = np.sqrt(radius * radius - arclength * arclength * 0.25) * (-1.0 if radius > 0.0 else 1.0)
    pc = (pm[0] + dist * (p1[1] - p0[1]) / arclength, pm[1] + dist * (p0[0] - p1[0]) / arclength)
    adis, arad = abs(dist), abs(radius)
    pd = ((pm[0] - pc[0]) * (arad / adis), (pm[1] - pc[1]) * (arad / adis))
    theta = np.arctan(0.5 * arclength / adis) * (1.0 if radius > 0 else -1.0) * orientation
    beta = theta * z
    return (pc[0] + pd[0] * np.cos(beta) - pd[1] * np.sin(beta),
            pc[1] + pd[0] * np.sin(beta) + pd[1] * np.cos(beta))


def _edge_deviation(f, z, w, L):
    """Element::generate_curve for one coordinate of a curved side
    (Quad::forward_transform_side and Quad::get_1d_mass_matrix,
    src/Quad.cpp:1455-1515): the side values less their linear interpolant,
    projected with the GLL rule onto the L - 2 edge modes (1-z)(1+z)/4
    P_i^(1,1)(z); returned at the points z."""
    from numpy.polynomial import polynomial as npoly
    lin = f[0] * (1 - z) / 2 + f[-1] * (1 + z) / 2
    d = f - lin
    if L <= 2:
        return np.zeros_like(z)
    from scipy.special import eval_jacobi
    E = np.stack([(1 - z) * (1 + z) / 4 * eval_jacobi(i, 1.0, 1.0, z) for i in range(L - 2)])
    M = (E[:, 1:-1] * w[1:-1]) @ E[:, 1:-1].T
    rhs = E @ (w * d)
    return np.linalg.solve(M, rhs) @ E


def curved_quad_nodes(mesh: Mesh, p: int, crv):
    """GLL nodal coordinates [2][elem][q*q] of quads with curved sides from a
    ``root.crv`` description (formats.read_crv): Quad::get_curved_coordinates
    (src/Quad.cpp:1018-1130) evaluates the modified C0 expansion of the
    vertices and the curved side's edge modes at the nodes, so the element map
    is the bilinear vertex map plus the side's projected deviation from its
    chord, carried into the element by the linear vertex function across it.
    Returns (nodes, curved element ids)."""
    types, sides = crv
    q, L = p + 2, p + 1
    z = gll_points(q)
    from numpy.polynomial import legendre
    P = legendre.legval(z, np.eye(q)[q - 1])
    w = 2.0 / (q * (q - 1) * P * P)
    x = quad_nodes(mesh, p).copy()
    curve = {}
    for eid, sid, tag in sides:
        curve[eid] = (sid, tag)
    curved = []
    for eid, (sid, tag) in sorted(curve.items()):
        code, params = types[tag]
        curved.append(eid)
        if code not in ("arc", "circle"):
            continue                     # straight (non-affine) sides: the vertex map already
        v = mesh.vert[eid]
        p0 = (v[0][sid], v[1][sid])
        p1 = (v[0][(sid + 1) % 4], v[1][(sid + 1) % 4])
        orientation = 1 if sid in (0, 1) else -1
        ax, ay = _arc_points(p0, p1, params[0], orientation, z)
        dx, dy = _edge_deviation(ax, z, w, L), _edge_deviation(ay, z, w, L)
        a, b = np.meshgrid(z, z, indexing="ij")
        if sid == 0:
            fac, dev = (1 - b) / 2, (dx[:, None], dy[:, None])
        elif sid == 2:
            fac, dev = (1 + b) / 2, (dx[:, None], dy[:, None])
        elif sid == 1:
            fac, dev = (1 + a) / 2, (dx[None, :], dy[None, :])
        else:
            fac, dev = (1 - a) / 2, (dx[None, :], dy[None, :])
        x[0, eid] += (fac * dev[0]).ravel()
        x[1, eid] += (fac * dev[1]).ravel()
    return x, curved


def bump_map(lo, hi, amp):
    """The channel [lo, hi] with a smooth periodic bump of height `amp` on its
    bottom wall, decaying linearly to the flat top: z' = z + amp b(x, y) (hi_z - z) / (hi_z - lo_z),
    b = (1 - cos(2 pi x / Lx)) (1 - cos(2 pi y / Ly)) / 4."""
    Lx, Ly, H = hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]

    def fn(x, y, z):
        b = (1 - np.cos(2 * np.pi * (x - lo[0]) / Lx)) * (1 - np.cos(2 * np.pi * (y - lo[1]) / Ly)) / 4
        return x, y, z + amp * b * (hi[2] - z) / H
    return fn
