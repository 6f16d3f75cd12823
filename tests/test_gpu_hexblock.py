"""The element-block seam on the GPU (dgx_hexblock_*, GpGpu::hex_kernels,
src/GpGpu.hpp:24-57) against the oracle's restatement of Hex::inner_product
and Hex::backward_transform (src/Hex.cpp:703-812).  Both sum in ascending
index order with separate multiplies and adds, so the bar is bitwise
equality, on host buffers (copied in and out, as the reference does) and on
device buffers, for a shared and a per-element wJ, cubic and non-cubic
point counts, and blocks of 1 element up to more elements than resident
workgroups."""
import numpy as np
import pytest

from dgm_amd.hexblock import DeviceBuffer, HexKernels
from oracle_lib import hexblock_backward_transform, hexblock_inner_product

pytestmark = pytest.mark.gpu


def _case(seed, ne, nm, qa, qb, qc, per_element):
    rng = np.random.default_rng(seed)
    B = [rng.standard_normal((nm, q)) for q in (qa, qb, qc)]
    wJ = rng.uniform(0.5, 2.0, (ne if per_element else 1, qa * qb * qc))
    u = rng.standard_normal((ne, qa * qb * qc))
    return B, wJ, u


@pytest.mark.parametrize("ne,nm,qa,qb,qc,per_element", [(1, 2, 3, 3, 3, False), (37, 3, 4, 5, 6, False),
                                                         (200, 8, 9, 9, 9, False), (64, 8, 9, 9, 9, True),
                                                         (5, 11, 12, 12, 12, True), (3, 16, 16, 16, 16, False),
                                                         (40000, 4, 5, 5, 5, False)])
def test_host_buffers_bitwise(ne, nm, qa, qb, qc, per_element):
    print(f"[hexblock] host ne={ne} nm={nm} q=({qa},{qb},{qc}) per_element={per_element}", flush=True)
    B, wJ, u = _case(ne + nm, ne, nm, qa, qb, qc, per_element)
    hk = HexKernels().initialize(ne, nm, qa, qb, qc)
    hk.Basis(*B)
    hk.wJ(wJ, per_element)
    got = u.copy()
    hk.inner_product(got)
    want = hexblock_inner_product(u, *B, wJ, per_element)
    assert np.array_equal(got, want)
    got_b = u.copy()
    hk.backward_transform(got_b)
    assert np.array_equal(got_b, hexblock_backward_transform(u, *B))
    hk.close()


def test_device_buffers_bitwise_and_round_trip():
    ne, nm, q = 4096, 8, 9
    print(f"[hexblock] device ne={ne} nm={nm} q={q}", flush=True)
    B, wJ, u = _case(3, ne, nm, q, q, q, False)
    hk = HexKernels().initialize(ne, nm, q, q, q)
    hk.Basis(*B)
    hk.wJ(wJ)
    d = DeviceBuffer(u.size)
    d.upload(u)
    hk.inner_product(d.address)
    hk.synchronize()
    want = hexblock_inner_product(u, *B, wJ)
    assert np.array_equal(d.download().reshape(u.shape), want)
    hk.backward_transform(d.address)
    hk.synchronize()
    assert np.array_equal(d.download().reshape(u.shape), hexblock_backward_transform(want, *B))
    d.free()
    hk.close()


def test_errors_before_setup():
    hk = HexKernels().initialize(4, 2, 3, 3, 3)
    u = np.zeros((4, 27))
    with pytest.raises(RuntimeError, match="set the basis"):
        hk.inner_product(u)
    hk.Basis(np.ones((2, 3)), np.ones((2, 3)), np.ones((2, 3)))
    with pytest.raises(RuntimeError, match="set wJ"):
        hk.inner_product(u)
    hk.backward_transform(u)   # needs no wJ
    with pytest.raises(ValueError):
        hk.inner_product(np.zeros((3, 27)))
    hk.close()
