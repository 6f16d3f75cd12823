"""The oracle's variable-order quads (VarOracle), pinned beyond runs/vortex.ref
(test_oracle_golden.py) by invariants:
- with every element at one order it is the uniform-order oracle;
- a uniform state has a zero residual (free-stream preservation across sides
  whose two elements differ in order).
The side sizing itself (both elements of a link on the larger order's Gauss
points) is what runs/vortex.ref pins: sizing each side by its own element
instead misses the golden file by 6e-3.
"""
import numpy as np

from dgm_amd import mesh as meshgen
from oracle_lib import Oracle, VarOracle


def _state(x):
    rho = 1.0 + 0.1 * np.sin(x[0]) * np.cos(x[1])
    u = 0.4 + 0.1 * np.sin(x[0] + x[1])
    v = -0.2 * np.cos(x[0]) * np.sin(x[1])
    p = 1.0 / 1.4 + 0.05 * np.cos(2 * x[0]) * np.cos(x[1])
    return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])


def test_one_order_is_the_uniform_oracle():
    m = meshgen.periodic_quads(4, 3, 0.0, 2 * np.pi, 0.0, 2 * np.pi)
    var = VarOracle(m, np.full(m.ne, 5, dtype=np.int32))
    uni = Oracle(m, 5, 1)
    x = uni.coords()
    assert np.array_equal(var.coords(), x)
    U = _state(x)
    np.testing.assert_allclose(var.rhs(U), uni.rhs(U), rtol=0, atol=1e-13 * np.abs(uni.rhs(U)).max())


def test_free_stream_across_order_changes():
    m = meshgen.periodic_quads(4, 4, 0.0, 2.0, 0.0, 2.0)
    orders = np.array([1, 4, 7, 10] * 4, dtype=np.int32)
    o = VarOracle(m, orders)
    x = o.coords()
    U = np.stack([np.full_like(x[0], c) for c in (1.0, 0.5, -0.2, 2.0)])
    F = o.rhs(U)
    assert np.abs(F).max() < 1e-12, np.abs(F).max()
