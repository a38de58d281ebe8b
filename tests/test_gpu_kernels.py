"""Kernel-level GPU parity: the all-channel diagonal blocks D^-1 / D in both kernel forms
(matrix-core GEMMs and vector ALU) and both precisions against a dense FP64 solve of each
block, D_j = T + diag(Re V_00) + e_j - c_j e_N e_N^T (multiscat/multiscat/_scipy.py:249-354)."""

from __future__ import annotations

import numpy as np
import pytest

from multiscat_amd import _lib
from multiscat_amd import basis as B

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nz,nc", [(24, 16), (40, 51), (40, 7), (150, 40), (300, 33)])
@pytest.mark.parametrize("inverse", [1, 0])
def test_diagonal_blocks(nz, nc, inverse):
    rng = np.random.default_rng(nz * 100 + nc)
    z, _ = B.lobatto_rule(nz, -1.5, 8.0)
    T = np.ascontiguousarray(B.kinetic_matrix(nz, -1.5, 8.0))
    V00 = np.ascontiguousarray((-10 * np.exp(-z * z)).astype(np.complex128))
    e = np.ascontiguousarray(rng.uniform(-20, 50, nc))
    c = np.ascontiguousarray(rng.standard_normal(nc) + 1j * rng.standard_normal(nc))
    x = np.ascontiguousarray(rng.standard_normal((nc, nz)) + 1j * rng.standard_normal((nc, nz)))
    H0 = T + np.diag(V00.real)
    Ds = []
    ref = np.zeros_like(x)
    for j in range(nc):
        D = (H0 + e[j] * np.eye(nz)).astype(complex)
        D[-1, -1] -= c[j]
        Ds.append(D)
        ref[j] = np.linalg.solve(D, x[j]) if inverse else D @ x[j]
    scale = np.abs(ref).max()
    # Forward error of any FP64 solve (numpy's included) is bounded by ~cond(D) eps; the
    # random blocks here reach cond ~ 2e8 at nz = 300, so the solve is checked by its
    # normwise backward error, which is ~eps for a correct solver whatever the conditioning,
    # and its forward error against the dense solve is held to cond * eps.
    cond = max(np.linalg.cond(D) for D in Ds) if inverse else 1.0
    for prec, bwd_tol, fwd_tol in ((0, 1e-13, max(1e-10, 1e-15 * cond)), (1, 1e-5, 2e-5 * max(1.0, cond / 1e7))):
        for variant in (0, 1, 2):
            if prec == 1 and variant == 1 and not inverse:
                continue  # the forward FP32 blocks exist in the matrix-core form only
            if variant == 2 and (prec == 1 or nz > 160):
                continue  # 8 channels per wavefront: an FP64 shape of the autotuner, nz <= 160
            y = np.zeros_like(x)
            rc = _lib.lib().msx_debug_apply_diagonal(nz, _lib.ptr(V00), _lib.ptr(T), nc,
                                                     _lib.ptr(e), _lib.ptr(c), _lib.ptr(x),
                                                     _lib.ptr(y), inverse, prec, variant)
            _lib.check(rc, "msx_debug_apply_diagonal")
            err = np.abs(y - ref).max() / scale
            assert err < fwd_tol, (prec, variant, err, cond)
            if inverse:
                bwd = max(np.abs(D @ y[j] - x[j]).max()
                          / (np.abs(D).sum(axis=1).max() * np.abs(y[j]).max() + np.abs(x[j]).max())
                          for j, D in enumerate(Ds))
                assert bwd < bwd_tol, (prec, variant, bwd)


@pytest.mark.parametrize("nz,nc", [(24, 16), (40, 51), (96, 7), (150, 40), (160, 33), (170, 21),
                                   (200, 45), (224, 9)])
def test_split_fp16_diagonal_blocks(nz, nc):
    """D^-1 with split-FP16 operands (each FP32 operand a scaled pair of FP16 values, three
    FP16 products accumulated in FP32; variant 3): within the FP32 bars against the dense FP64
    solve, within a small factor of the FP32-operand kernel's own error, and per channel,
    so that channels 50 orders of magnitude apart keep their relative accuracy (the split
    scales every column by its own power of two)."""
    rng = np.random.default_rng(nz * 7 + nc)
    z, _ = B.lobatto_rule(nz, -1.5, 8.0)
    T = np.ascontiguousarray(B.kinetic_matrix(nz, -1.5, 8.0))
    V00 = np.ascontiguousarray((-10 * np.exp(-z * z)).astype(np.complex128))
    e = np.ascontiguousarray(rng.uniform(-20, 50, nc))
    c = np.ascontiguousarray(rng.standard_normal(nc) + 1j * rng.standard_normal(nc))
    mag = 10.0 ** rng.uniform(-25, 25, nc)
    x = (rng.standard_normal((nc, nz)) + 1j * rng.standard_normal((nc, nz))) * mag[:, None]
    x[nc // 2] = 0.0  # an all-zero channel stays zero
    x = np.ascontiguousarray(x)
    H0 = T + np.diag(V00.real)
    Ds, ref = [], np.zeros_like(x)
    for j in range(nc):
        D = (H0 + e[j] * np.eye(nz)).astype(complex)
        D[-1, -1] -= c[j]
        Ds.append(D)
        ref[j] = np.linalg.solve(D, x[j])
    cond = max(np.linalg.cond(D) for D in Ds)
    out = {}
    for variant in (0, 3, 4):  # FP32 operands; split FP16 with 16 and 8 channels per wavefront
        y = np.zeros_like(x)
        rc = _lib.lib().msx_debug_apply_diagonal(nz, _lib.ptr(V00), _lib.ptr(T), nc, _lib.ptr(e),
                                                 _lib.ptr(c), _lib.ptr(x), _lib.ptr(y), 1, 1,
                                                 variant)
        _lib.check(rc, "msx_debug_apply_diagonal")
        out[variant] = y
    assert not out[3][nc // 2].any() and not out[4][nc // 2].any()
    fwd = {}
    for variant, y in out.items():
        per = [np.abs(y[j] - ref[j]).max() / np.abs(ref[j]).max()
               for j in range(nc) if j != nc // 2]
        fwd[variant] = max(per)
        bwd = max(np.abs(D @ y[j] - x[j]).max()
                  / (np.abs(D).sum(axis=1).max() * np.abs(y[j]).max() + np.abs(x[j]).max())
                  for j, D in enumerate(Ds) if j != nc // 2)
        assert bwd < 1e-5, (variant, bwd)
        assert fwd[variant] < 2e-5 * max(1.0, cond / 1e7), (variant, fwd[variant], cond)
    assert fwd[3] < 4.0 * fwd[0] + 1e-6, fwd
    assert fwd[4] < 4.0 * fwd[0] + 1e-6, fwd


def test_split_fp16_rejects_shapes_outside_the_kernel():
    nz, nc = 240, 4
    T = np.ascontiguousarray(B.kinetic_matrix(nz, -1.5, 8.0))
    V00 = np.zeros(nz, np.complex128)
    e = np.ones(nc)
    c = np.ones(nc, np.complex128)
    x = np.ones((nc, nz), np.complex128)
    y = np.zeros_like(x)
    for prec, variant in ((1, 3), (0, 3), (1, 4)):
        rc = _lib.lib().msx_debug_apply_diagonal(nz, _lib.ptr(V00), _lib.ptr(T), nc, _lib.ptr(e),
                                                 _lib.ptr(c), _lib.ptr(x), _lib.ptr(y), 1, prec,
                                                 variant)
        assert rc != 0
