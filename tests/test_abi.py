"""The C-ABI library: it loads, exports every symbol include/multiscat_amd.h declares, and its
host-side entry points (the reference's scatsub.f90 helpers and the argument checks that run
before any GPU work) behave like the reference's boundary. No GPU is touched here."""

from __future__ import annotations

import re

import numpy as np
import pytest

from conftest import ROOT, golden
from multiscat_amd import _lib, fortran_compat as fc


def header_symbols():
    text = (ROOT / "include" / "multiscat_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msx_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.lib()
    declared = header_symbols()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_lib.EXPORTED) == declared


@pytest.fixture(scope="module")
def helpers():
    return golden("helpers.npz")


@pytest.mark.parametrize("i", range(4))
def test_get_lobatto_weights(helpers, i):
    zmin, zmax, n = helpers[f"lobatto{i}"]
    w, x = fc.get_lobatto_weights(zmin, zmax, int(n))
    np.testing.assert_allclose(x, helpers[f"lobatto{i}_x"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(w, helpers[f"lobatto{i}_w"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("i", range(3))
def test_get_parallel_kinetic_energy(helpers, i):
    zmin, zmax, n = helpers[f"kinetic{i}"]
    T = fc.get_parallel_kinetic_energy(zmin, zmax, int(n))
    ref = helpers[f"kinetic{i}_T"]
    np.testing.assert_allclose(T, ref, rtol=0, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("i", range(3))
def test_get_perpendicular_kinetic_difference_and_abc(helpers, i):
    ax, ay, bx, by, kx, ky, kz, nx, ny = helpers[f"lattice{i}"]
    e = fc.get_perpendicular_kinetic_difference(kx, ky, kz, int(nx), int(ny), ax, ay, bx, by)
    np.testing.assert_allclose(e, helpers[f"lattice{i}_e"], rtol=1e-12, atol=1e-12)
    a, b, c = fc.get_abc_arrays(-1.0, 7.0, int(nx), int(ny), e, 40)
    abc = helpers[f"lattice{i}_abc"]
    for got, ref in zip((a, b, c), abc):
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_error_codes_mirror_reference():
    """ierr = 1 for non-positive dimensions, raised as the reference's RuntimeError
    (lib/multiscat_fortran/src/multiscat_fortran/__init__.py:33-36)."""
    with pytest.raises(RuntimeError, match="get_parallel_kinetic_energy failed with error code 1"):
        fc.get_parallel_kinetic_energy(0.0, 1.0, 0)
    with pytest.raises(RuntimeError, match="get_lobatto_weights failed with error code 1"):
        fc.get_lobatto_weights(0.0, 1.0, 0)
    lib = _lib.lib()
    z = np.zeros(4, dtype=np.complex128)
    assert lib.msx_run_multiscat_fortran(1, 5, 0, 2, 2, _lib.ptr(z), _lib.ptr(z), _lib.ptr(z),
                                         _lib.ptr(z), _lib.ptr(z), _lib.ptr(z), _lib.ptr(z)) == 1
    with pytest.raises(RuntimeError, match="run_multiscat_fortran failed with error code 1"):
        fc.run_multiscat_fortran(1, 5, np.zeros((0, 2, 2)), np.zeros((0, 2)), [], [], [],
                                 np.zeros((2, 2)))


def test_default_options_are_the_reference_settings():
    o = _lib.default_options(7)
    assert o.method == _lib.METHOD_GAUSS_SEIDEL and o.precision == _lib.PRECISION_FP64
    assert o.maxiter == 2000 and o.window == 3 and o.neumann == 1
    assert abs(o.eps - 0.5e-7) < 1e-22  # multiscat.f90:35


def test_shape_checks_before_gpu():
    with pytest.raises(ValueError):
        fc.run_multiscat_fortran(1, 5, np.zeros((2, 2, 3)), np.zeros((2, 3)), np.zeros(4),
                                 np.zeros(4), np.zeros(4), np.zeros((3, 3)))
