"""The GPU symmetric eigensolver (Householder + divide and conquer, and Householder + QL) on
matrices chosen to break a faulty implementation: clustered and repeated eigenvalues (every
deflation path of the merges), glued Wilkinson matrices, graded spectra, zero couplings, and
H0 of the benchmark tables. Checked against what the mathematics fixes: A Q = Q diag(lam),
Q^T Q = I, and the eigenvalues of numpy.linalg.eigvalsh (LAPACK)."""

from __future__ import annotations

import numpy as np
import pytest

from multiscat_amd import batch as BT
from multiscat_amd import workloads as W

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def _tridiag(d, e):
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


def _orthogonal(n, rng):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


def _matrices():
    rng = np.random.default_rng(7)
    out = {}
    # 173: the largest size whose packed triangle (with a copy of u per wavefront) fits LDS;
    # 174 .. 189 with one shared copy; from 190 the first n - 189 steps run in global memory
    for n in (2, 3, 5, 17, 64, 150, 173, 174, 189, 190, 200, 300):
        a = rng.standard_normal((n, n))
        out[f"random{n}"] = a + a.T
    m = 10
    wp = np.abs(np.arange(2 * m + 1) - m).astype(float)
    out["wilkinson21"] = _tridiag(wp, np.ones(2 * m))
    e = np.tile(np.append(np.ones(2 * m), 1e-8), 5)[:-1]
    out["glued_wilkinson105"] = _tridiag(np.tile(wp, 5), e)
    out["zero40"] = np.zeros((40, 40))
    e = rng.standard_normal(59)
    e[::4] = 0.0
    out["ones_diag_some_zero_e60"] = _tridiag(np.ones(60), e)
    q = _orthogonal(90, rng)
    out["triple_eigenvalues90"] = (q * np.repeat([-2.0, 0.5, 3.0], 30)) @ q.T
    out["graded80"] = _tridiag(10.0 ** np.linspace(-10, 10, 80), 10.0 ** np.linspace(-10, 10, 79))
    out["tiny_coupling100"] = _tridiag(np.arange(100.0), np.full(99, 1e-12))
    for nz in (100, 150, 200, 300):
        w = W.hex_batch(n_conditions=1, nz=nz, nk=3, cond_seed=2, channel_counts=[5])
        out[f"H0_nz{nz}"] = w.kinetic + np.diag(w.potential[0, 0].real)
    return out


MATS = _matrices()


def _check(A, lam, Q):
    n = A.shape[0]
    ref = np.linalg.eigvalsh(A)
    scale = max(np.abs(ref).max(), np.finfo(float).tiny)
    tol = 40 * EPS * np.sqrt(n)
    assert np.abs(lam - ref).max() <= tol * scale
    assert np.all(np.diff(lam) >= 0)
    assert np.abs(Q.T @ Q - np.eye(n)).max() <= tol
    assert np.abs(A @ Q - Q * lam).max() <= tol * scale


@pytest.mark.parametrize("name", sorted(MATS))
def test_divide_and_conquer(name):
    A = MATS[name]
    lam, Q = BT.symmetric_eigensystem(A, "divide_conquer")
    _check(A, lam, Q)


@pytest.mark.parametrize("name", ["random150", "glued_wilkinson105", "triple_eigenvalues90",
                                  "H0_nz150"])
def test_ql_agrees(name):
    """The QL solver (used above n = 1024) on the same matrices, and against divide and
    conquer (eigenvectors up to sign where the eigenvalues are simple)."""
    A = MATS[name]
    lam, Q = BT.symmetric_eigensystem(A, "ql")
    _check(A, lam, Q)
    lam2, Q2 = BT.symmetric_eigensystem(A, "divide_conquer")
    np.testing.assert_allclose(lam, lam2, rtol=0, atol=40 * EPS * np.sqrt(len(lam)) * np.abs(lam).max())


def test_divide_and_conquer_largest():
    """n = 1024, the largest size the divide-and-conquer path takes."""
    rng = np.random.default_rng(3)
    a = rng.standard_normal((1024, 1024))
    A = a + a.T
    lam, Q = BT.symmetric_eigensystem(A)
    _check(A, lam, Q)


def test_table_uses_divide_and_conquer():
    """The table's eigensystem comes from divide and conquer at nz <= 1024 (no QL sweeps) and
    matches LAPACK."""
    w = W.hex_batch(n_conditions=1, nz=150, nk=3, cond_seed=2, channel_counts=[5])
    tab = BT.Table(w.potential, w.kinetic)
    ph = tab.eig_phases()
    assert ph[5] == 0 and ph[6] == 0
    lam, Q = tab.eigensystem()
    _check(w.kinetic + np.diag(w.potential[0, 0].real), lam, Q)
