"""msx_symmetric_eigensystem_noalloc (the allocation-free entry point) against the entry point
it varies, msx_symmetric_eigensystem: two calls in a row on device buffers the test owns, the
second smaller than the first (the static scratch array keeps the first call's contents), each
matching the allocating entry point and the eigen-decomposition itself; sizes beyond the
maximum are rejected before any GPU work."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from multiscat_amd import _lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    _lib.lib()  # the library's HIP runtime first: the same one then serves these calls
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipDeviceSynchronize):
        f.restype = C.c_int
    return hip


class _DeviceArrays:
    """Device buffers owned by the test (the caller of an allocation-free entry point)."""

    def __init__(self, hip, nbytes):
        self.hip, self.ptrs = hip, []
        for nb in nbytes:
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), nb) == 0
            self.ptrs.append(p)

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def _graded_symmetric(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    A = 0.5 * (M + M.T) + np.diag(np.linspace(-3.0, 40.0, n))
    return np.ascontiguousarray(A)


def _noalloc(hip, n, A):
    lam, Q, info = np.zeros(n), np.zeros((n, n)), np.zeros(1, dtype=np.int32)
    bufs = _DeviceArrays(hip, [8 * n * n, 8 * n, 8 * n * n, 4])
    try:
        dA, dlam, dQ, dinfo = bufs.ptrs
        assert hip.hipMemcpy(dA, A.ctypes.data, 8 * n * n, H2D) == 0
        rc = _lib.lib().msx_symmetric_eigensystem_noalloc(n, dA, dlam, dQ, dinfo, None)
        assert rc == _lib.MSX_OK, _lib.last_error()
        assert hip.hipDeviceSynchronize() == 0
        for host, dev, nb in ((lam, dlam, 8 * n), (Q, dQ, 8 * n * n), (info, dinfo, 4)):
            assert hip.hipMemcpy(host.ctypes.data, dev, nb, D2H) == 0
        A_back = np.zeros_like(A)
        assert hip.hipMemcpy(A_back.ctypes.data, dA, 8 * n * n, D2H) == 0
        np.testing.assert_array_equal(A_back, A)  # the input is only read
    finally:
        bufs.close()
    return lam, Q, int(info[0])


def _allocating(n, A):
    lam, Q = np.zeros(n), np.zeros((n, n))
    _lib.check(_lib.lib().msx_symmetric_eigensystem(n, _lib.ptr(A), 0, _lib.ptr(lam),
                                                    _lib.ptr(Q)), "msx_symmetric_eigensystem")
    return lam, Q


def test_noalloc_eigensystem_matches_allocating_entry_point_twice():
    hip = _hip()
    for n, seed in ((300, 1), (120, 2)):  # the second call smaller than the first
        A = _graded_symmetric(n, seed)
        lam, Q, info = _noalloc(hip, n, A)
        assert info == 0
        lam_ref, Q_ref = _allocating(n, A)
        scale = np.abs(A).max()
        np.testing.assert_allclose(lam, lam_ref, rtol=0, atol=1e-13 * scale)
        np.testing.assert_allclose(Q, Q_ref, rtol=0, atol=1e-12)
        # and the decomposition itself, against numpy's eigenvalues
        assert np.abs(A @ Q - Q * lam[None, :]).max() <= 1e-12 * scale
        assert np.abs(Q.T @ Q - np.eye(n)).max() <= 1e-12
        np.testing.assert_allclose(lam, np.linalg.eigvalsh(A), rtol=0, atol=1e-12 * scale)


def test_noalloc_rejects_sizes_beyond_the_maximum():
    hip = _hip()
    bufs = _DeviceArrays(hip, [8, 8, 8, 4])
    try:
        p = bufs.ptrs
        for n in (1, 513):
            rc = _lib.lib().msx_symmetric_eigensystem_noalloc(n, p[0], p[1], p[2], p[3], None)
            assert rc == _lib.MSX_ERR_DIMS
            assert "512" in _lib.last_error()
    finally:
        bufs.close()
