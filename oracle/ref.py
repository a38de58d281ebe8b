"""ctypes driver for the reference Fortran built from source into ``oracle/_ref``.

Test infrastructure only (checker, never product): tests, the smoke check and
bench.py's cpu_baseline leg may load it. ``make -C oracle ref`` builds the
library from the reference's own five Fortran files; nothing of the reference
is copied into this repository.

The native ABI is the reference's explicit-shape Fortran one: every argument by
reference, 2-D/3-D arrays column-major, the 1-D channel waves in C-ravel order
(idx = ix*nky + iy) (lib/multiscat_fortran/fortran/multiscat_python_bindings.f90:1-61;
the f2py interface multiscat_python_bindings.pyf:3-97). This module mirrors
the Python face of that boundary (lib/multiscat_fortran/src/multiscat_fortran/
__init__.py:33-169): same function names and argument names, RuntimeError on a
non-zero ierr.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
REF_PATH = _HERE / "_ref" / "libmultiscat_ref.so"

_lib = None


def available() -> bool:
    return REF_PATH.exists()


def _load():
    global _lib
    if _lib is None:
        if not REF_PATH.exists():
            raise FileNotFoundError(f"reference build missing: {REF_PATH} (run make -C oracle ref)")
        _lib = C.CDLL(str(REF_PATH))
    return _lib


def _i(v: int):
    return C.byref(C.c_int(int(v)))


def _d(v: float):
    return C.byref(C.c_double(float(v)))


def _f(a, dtype):
    """Column-major copy of ``a`` (the Fortran side reads column-major)."""
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _raise(name: str, ierr: int) -> None:
    # lib/multiscat_fortran/src/multiscat_fortran/__init__.py:33-36
    if ierr != 0:
        raise RuntimeError(f"Fortran {name} failed with error code {ierr}")


def run_multiscat_fortran(
    gmres_preconditioner_flag: int,
    convergence_significant_figures: int,
    potential_values,
    perpendicular_kinetic_difference,
    wave_a,
    wave_b,
    wave_c,
    parallel_kinetic_energy,
):
    """Reference solve: V_G(z) table -> scattered state (nkx, nky, nz) complex128."""
    lib = _load()
    V = _f(potential_values, np.complex128)
    nkx, nky, nz = V.shape
    e = _f(perpendicular_kinetic_difference, np.float64)
    a = np.ascontiguousarray(wave_a, dtype=np.complex128).ravel()
    b = np.ascontiguousarray(wave_b, dtype=np.complex128).ravel()
    c = np.ascontiguousarray(wave_c, dtype=np.complex128).ravel()
    T = _f(parallel_kinetic_energy, np.float64)
    out = np.zeros((nkx, nky, nz), dtype=np.complex128, order="F")
    ierr = C.c_int(0)
    lib.run_multiscat_fortran_(
        _i(gmres_preconditioner_flag), _i(convergence_significant_figures),
        _i(nkx), _i(nky), _i(nz), _p(V), _p(e), _p(a), _p(b), _p(c), _p(T), _p(out),
        C.byref(ierr),
    )
    _raise("run_multiscat_fortran", ierr.value)
    return np.ascontiguousarray(out)


def get_perpendicular_kinetic_difference(incident_kx, incident_ky, incident_kz, nx, ny,
                                         unit_cell_ax, unit_cell_ay, unit_cell_bx, unit_cell_by):
    lib = _load()
    out = np.zeros((nx, ny), dtype=np.float64, order="F")
    ierr = C.c_int(0)
    lib.get_perpendicular_kinetic_difference_(
        _d(incident_kx), _d(incident_ky), _d(incident_kz), _i(nx), _i(ny),
        _d(unit_cell_ax), _d(unit_cell_ay), _d(unit_cell_bx), _d(unit_cell_by),
        _p(out), C.byref(ierr))
    _raise("get_perpendicular_kinetic_difference", ierr.value)
    return np.ascontiguousarray(out)


def get_parallel_kinetic_energy(zmin, zmax, nz):
    lib = _load()
    out = np.zeros((nz, nz), dtype=np.float64, order="F")
    ierr = C.c_int(0)
    lib.get_parallel_kinetic_energy_(_d(zmin), _d(zmax), _i(nz), _p(out), C.byref(ierr))
    _raise("get_parallel_kinetic_energy", ierr.value)
    return np.ascontiguousarray(out)


def get_lobatto_weights(zmin, zmax, node_count):
    lib = _load()
    w = np.zeros(node_count, dtype=np.float64)
    x = np.zeros(node_count, dtype=np.float64)
    ierr = C.c_int(0)
    lib.get_lobatto_weights_(_d(zmin), _d(zmax), _i(node_count), _p(w), _p(x), C.byref(ierr))
    _raise("get_lobatto_weights", ierr.value)
    return w, x


def get_abc_arrays(zmin, zmax, nx, ny, perpendicular_kinetic_difference, n_z_points):
    lib = _load()
    e = _f(perpendicular_kinetic_difference, np.float64)
    a = np.zeros(nx * ny, dtype=np.complex128)
    b = np.zeros(nx * ny, dtype=np.complex128)
    c = np.zeros(nx * ny, dtype=np.complex128)
    ierr = C.c_int(0)
    lib.get_abc_arrays_(_d(zmin), _d(zmax), _i(nx), _i(ny), _p(e), _i(n_z_points),
                        _p(a), _p(b), _p(c), C.byref(ierr))
    _raise("get_abc_arrays", ierr.value)
    return a, b, c


def debug_build_preconditioner_fortran(potential_values, perpendicular_kinetic_difference,
                                       parallel_kinetic_energy):
    lib = _load()
    V = _f(potential_values, np.complex128)
    nkx, nky, nz = V.shape
    e = _f(perpendicular_kinetic_difference, np.float64)
    T = _f(parallel_kinetic_energy, np.float64)
    lam = np.zeros(nz, dtype=np.float64)
    fac = np.zeros((nz, nkx * nky), dtype=np.float64, order="F")
    vec = np.zeros((nz, nz), dtype=np.float64, order="F")
    ierr = C.c_int(0)
    lib.debug_build_preconditioner_fortran_(_i(nkx), _i(nky), _i(nz), _p(V), _p(e), _p(T),
                                            _p(lam), _p(fac), _p(vec), C.byref(ierr))
    _raise("debug_build_preconditioner_fortran", ierr.value)
    return lam, np.ascontiguousarray(fac), np.ascontiguousarray(vec)


def debug_apply_upper_block_fortran(potential_values, state_in):
    lib = _load()
    V = _f(potential_values, np.complex128)
    nkx, nky, nz = V.shape
    x = _f(state_in, np.complex128)
    out = np.zeros_like(x, order="F")
    ierr = C.c_int(0)
    lib.debug_apply_upper_block_fortran_(_i(nkx), _i(nky), _i(nz), _p(V), _p(x), _p(out),
                                         C.byref(ierr))
    _raise("debug_apply_upper_block_fortran", ierr.value)
    return np.ascontiguousarray(out)


def debug_solve_lower_block_fortran(potential_values, wave_c, perpendicular_kinetic_difference,
                                    parallel_kinetic_energy, state_in):
    lib = _load()
    V = _f(potential_values, np.complex128)
    nkx, nky, nz = V.shape
    c = np.ascontiguousarray(wave_c, dtype=np.complex128).ravel()
    e = _f(perpendicular_kinetic_difference, np.float64)
    T = _f(parallel_kinetic_energy, np.float64)
    x = _f(state_in, np.complex128)
    out = np.zeros_like(x, order="F")
    ierr = C.c_int(0)
    lib.debug_solve_lower_block_fortran_(_i(nkx), _i(nky), _i(nz), _p(V), _p(c), _p(e), _p(T),
                                         _p(x), _p(out), C.byref(ierr))
    _raise("debug_solve_lower_block_fortran", ierr.value)
    return np.ascontiguousarray(out)
