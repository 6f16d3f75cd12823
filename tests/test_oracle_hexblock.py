"""The oracle's element-block transforms (GpGpu::hex_kernels inner_product and
backward_transform, src/GpGpu.cu:236-262 and 505-750; Hex::inner_product and
Hex::backward_transform, src/Hex.cpp:703-812), pinned by mathematics:

- brute-force sums over the full tensor product (numpy einsum) on small blocks
  with random, non-square bases and distinct qa, qb, qc, so that a transposed
  matrix or a swapped direction fails;
- the GLL mass identity: with orthonormal Legendre modes at q = p+2
  Gauss-Lobatto points and wJ = J w_a w_b w_c, inner_product(backward(uh)) =
  J uh exactly (the integrand has degree 2p <= 2q - 3), per element J too.
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

from oracle_lib import hexblock_backward_transform, hexblock_inner_product


def _gll(q):
    """Gauss-Lobatto-Legendre points and weights on [-1, 1] (q points)."""
    inner = npleg.Legendre.basis(q - 1).deriv().roots()
    z = np.concatenate([[-1.0], np.sort(inner.real), [1.0]])
    w = 2.0 / (q * (q - 1) * npleg.Legendre.basis(q - 1)(z) ** 2)
    return z, w


def _orthonormal(nm, z):
    return np.stack([np.sqrt((2 * n + 1) / 2.0) * npleg.Legendre.basis(n)(z) for n in range(nm)])


@pytest.mark.parametrize("nm,qa,qb,qc", [(2, 3, 4, 5), (3, 5, 4, 3), (4, 4, 4, 4)])
def test_inner_product_is_the_tensor_contraction(nm, qa, qb, qc):
    rng = np.random.default_rng(nm * 100 + qa)
    ne = 3
    Ba, Bb, Bc = rng.standard_normal((nm, qa)), rng.standard_normal((nm, qb)), rng.standard_normal((nm, qc))
    u = rng.standard_normal((ne, qa * qb * qc))
    for per_element in (False, True):
        wJ = rng.uniform(0.5, 2.0, (ne if per_element else 1, qa * qb * qc))
        got = hexblock_inner_product(u, Ba, Bb, Bc, wJ, per_element)
        f = (u * wJ).reshape(ne, qa, qb, qc)
        want = np.einsum("ia,jb,kc,eabc->eijk", Ba, Bb, Bc, f).reshape(ne, -1)
        np.testing.assert_allclose(got[:, :nm ** 3], want, rtol=1e-13, atol=1e-13)
        np.testing.assert_array_equal(got[:, nm ** 3:], u[:, nm ** 3:])   # the rest of the slot is kept


@pytest.mark.parametrize("nm,qa,qb,qc", [(2, 3, 4, 5), (3, 5, 4, 3), (5, 6, 6, 6)])
def test_backward_transform_is_the_tensor_expansion(nm, qa, qb, qc):
    rng = np.random.default_rng(nm * 7 + qc)
    ne = 4
    Ba, Bb, Bc = rng.standard_normal((nm, qa)), rng.standard_normal((nm, qb)), rng.standard_normal((nm, qc))
    u = rng.standard_normal((ne, qa * qb * qc))
    got = hexblock_backward_transform(u, Ba, Bb, Bc)
    uh = u[:, :nm ** 3].reshape(ne, nm, nm, nm)
    want = np.einsum("ia,jb,kc,eijk->eabc", Ba, Bb, Bc, uh).reshape(ne, -1)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("p", [1, 4, 7, 9])
def test_gll_mass_identity(p):
    nm, q = p + 1, p + 2
    z, w = _gll(q)
    B = _orthonormal(nm, z)
    rng = np.random.default_rng(p)
    ne = 5
    uh = rng.standard_normal((ne, nm ** 3))
    slots = np.zeros((ne, q ** 3))
    slots[:, :nm ** 3] = uh
    nodal = hexblock_backward_transform(slots, B, B, B)
    J = rng.uniform(0.5, 3.0, ne)
    w3 = np.einsum("a,b,c->abc", w, w, w).ravel()
    out = hexblock_inner_product(nodal, B, B, B, J[:, None] * w3[None, :], per_element=True)
    np.testing.assert_allclose(out[:, :nm ** 3], J[:, None] * uh, rtol=0, atol=5e-13 * np.abs(uh).max() * J.max())
    # one shared wJ (an affine block): J = 1
    out1 = hexblock_inner_product(nodal, B, B, B, w3, per_element=False)
    np.testing.assert_allclose(out1[:, :nm ** 3], uh, rtol=0, atol=5e-13 * np.abs(uh).max())
