"""The element-block seam: batched hex transforms over a block of elements
that share one basis, with the interface of the reference's
GpGpu::hex_kernels (src/GpGpu.hpp:24-57) as HexBlock drives it
(src/HexBlock.cpp:35-130): initialize, wJ, Basis, inner_product,
backward_transform.

Each element owns a slot of qa*qb*qc doubles ([a][b][c], c fastest); the
nmodes^3 modal coefficients ([i][j][k]) are the first values of the slot.
Both transforms work in place, on a float64 numpy array of ne slots (copied to
the GPU and back on every call, as the reference does) or on device memory
given by its address (no copies).  All arithmetic runs in libdgx.so
(dgx_hexblock_*); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import dgx


def _host(a, n):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float64 numpy array")
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class DeviceBuffer:
    """A float64 device buffer from the HIP runtime that libdgx.so itself links
    (hipMalloc / hipMemcpy / hipFree), for callers that keep element blocks
    resident on the GPU."""

    def __init__(self, n):
        dgx.load()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        rc = self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * self.n))
        if rc != 0:
            raise MemoryError(f"hipMalloc of {8 * self.n} bytes failed ({rc})")

    def _copy(self, dst, src, kind):
        rc = self.hip.hipMemcpy(dst, src, ctypes.c_size_t(8 * self.n), ctypes.c_int(kind))
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != self.n:
            raise ValueError(f"expected {self.n} values, got {a.size}")
        self._copy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), 1)   # hipMemcpyHostToDevice

    def download(self):
        out = np.empty(self.n)
        self._copy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, 2)   # hipMemcpyDeviceToHost
        return out

    @property
    def address(self):
        return int(self.ptr.value)

    def free(self):
        if self.ptr.value:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


class HexKernels:
    """GpGpu::hex_kernels on one GPU (libdgx's dgx_hexblock handle)."""

    def __init__(self):
        self.lib = dgx.load()
        self.h = None
        self.ne = self.nmodes = self.qa = self.qb = self.qc = 0

    def initialize(self, ne, nmodes, qa, qb, qc, stream=None):
        """hex_kernels::initialize(ne, nmodes, qa, qb, qc)."""
        self.close()
        h = ctypes.c_void_p()
        dgx._check(self.lib.dgx_hexblock_create(int(ne), int(nmodes), int(qa), int(qb), int(qc), stream,
                                                ctypes.byref(h)))
        self.h = h
        self.ne, self.nmodes, self.qa, self.qb, self.qc = int(ne), int(nmodes), int(qa), int(qb), int(qc)
        return self

    @property
    def qtot(self):
        return self.qa * self.qb * self.qc

    def wJ(self, wJ, per_element=False):
        """hex_kernels::wJ: weighted Jacobian [qa][qb][qc] for the block, or
        [ne][qa][qb][qc] with per_element."""
        wJ = np.ascontiguousarray(wJ, dtype=np.float64)
        n = self.qtot * (self.ne if per_element else 1)
        dgx._check(self.lib.dgx_hexblock_set_wJ(self.h, _host(wJ, n), int(bool(per_element))))

    def Basis(self, Ba, Bb, Bc):
        """hex_kernels::Basis: Ba [nmodes][qa], Bb [nmodes][qb], Bc [nmodes][qc]."""
        arrs = [np.ascontiguousarray(B, dtype=np.float64) for B in (Ba, Bb, Bc)]
        ptrs = [_host(B, self.nmodes * q) for B, q in zip(arrs, (self.qa, self.qb, self.qc))]
        dgx._check(self.lib.dgx_hexblock_set_basis(self.h, *ptrs))

    def _run(self, fn, u):
        if isinstance(u, int):   # device address of ne slots
            dgx._check(fn(self.h, ctypes.cast(u, ctypes.POINTER(ctypes.c_double)), 1))
            return u
        dgx._check(fn(self.h, _host(u, self.ne * self.qtot), 0))
        return u

    def inner_product(self, u):
        """hex_kernels::inner_product, in place: each slot's first nmodes^3
        values become the inner products against the basis.  `u` is a numpy
        array (synchronous) or a device address (asynchronous on the block's
        stream)."""
        return self._run(self.lib.dgx_hexblock_inner_product, u)

    def backward_transform(self, u):
        """hex_kernels::backward_transform, in place: modal (each slot's first
        nmodes^3 values) to the qa*qb*qc point values."""
        return self._run(self.lib.dgx_hexblock_backward_transform, u)

    def synchronize(self):
        dgx._check(self.lib.dgx_hexblock_synchronize(self.h))

    def close(self):
        if self.h is not None:
            self.lib.dgx_hexblock_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
