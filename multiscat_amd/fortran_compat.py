"""Drop-in replacement of the reference's ``multiscat_fortran`` module.

Same eight functions, same argument names and meaning, same return shapes and the same
``RuntimeError`` on a non-zero error code as
lib/multiscat_fortran/src/multiscat_fortran/__init__.py:33-181 (interface:
lib/multiscat_fortran/fortran/multiscat_python_bindings.pyf:3-97). Inputs may be any
array-likes; they are converted to the layouts of the native boundary (column-major 2-D and
3-D arrays, C-ravel-ordered 1-D channel waves) exactly as f2py does. The solve and the three
debug operators run on the GPU through libmultiscat_amd.so.

Usage: replace ``from multiscat_fortran import run_multiscat_fortran`` with
``from multiscat_amd.fortran_compat import run_multiscat_fortran``.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

__all__ = [
    "run_multiscat_fortran",
    "get_perpendicular_kinetic_difference",
    "get_parallel_kinetic_energy",
    "get_lobatto_weights",
    "get_abc_arrays",
    "debug_build_preconditioner_fortran",
    "debug_apply_upper_block_fortran",
    "debug_solve_lower_block_fortran",
]


def _raise_error(function_name: str, ierr: int) -> None:
    # lib/multiscat_fortran/src/multiscat_fortran/__init__.py:33-36
    if ierr != 0:
        detail = _lib.last_error() if ierr > 2 else ""
        msg = f"Fortran {function_name} failed with error code {ierr}"
        raise RuntimeError(msg + (f" ({detail})" if detail else ""))


def _fortran(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _flat(a, dtype, n, name):
    out = np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel())
    if out.size != n:
        raise ValueError(f"{name} has {out.size} elements, expected {n}")
    return out


def run_multiscat_fortran(
    gmres_preconditioner_flag: int,
    convergence_significant_figures: int,
    potential_values,
    perpendicular_kinetic_difference,
    wave_a,
    wave_b,
    wave_c,
    parallel_kinetic_energy,
    *,
    return_info: bool = False,
):
    """Scattered state (nkx, nky, nz) complex128 of the close-coupled system.

    ``return_info=True`` additionally returns ``{"iterations", "residual"}`` (an extension).
    """
    V = _fortran(potential_values, np.complex128)
    if V.ndim != 3:
        raise ValueError("potential_values must be 3-D (nkx, nky, nz)")
    nkx, nky, nz = V.shape
    e = _fortran(perpendicular_kinetic_difference, np.float64)
    if e.shape != (nkx, nky):
        raise ValueError("perpendicular_kinetic_difference must be (nkx, nky)")
    n = nkx * nky
    a = _flat(wave_a, np.complex128, n, "wave_a")
    b = _flat(wave_b, np.complex128, n, "wave_b")
    c = _flat(wave_c, np.complex128, n, "wave_c")
    T = _fortran(parallel_kinetic_energy, np.float64)
    if T.shape != (nz, nz):
        raise ValueError("parallel_kinetic_energy must be (nz, nz)")
    out = np.zeros((nkx, nky, nz), dtype=np.complex128, order="F")
    iters = C.c_int(0)
    resid = C.c_double(0.0)
    ierr = _lib.lib().msx_run_multiscat_fortran_ex(
        int(gmres_preconditioner_flag), int(convergence_significant_figures), nkx, nky, nz,
        _lib.ptr(V), _lib.ptr(e), _lib.ptr(a), _lib.ptr(b), _lib.ptr(c), _lib.ptr(T),
        _lib.ptr(out), C.byref(iters), C.byref(resid))
    _raise_error("run_multiscat_fortran", int(ierr))
    if return_info:
        return out, {"iterations": iters.value, "residual": resid.value}
    return out


def get_perpendicular_kinetic_difference(incident_kx, incident_ky, incident_kz, nx, ny,
                                         unit_cell_ax, unit_cell_ay, unit_cell_bx,
                                         unit_cell_by):
    out = np.zeros((nx, ny), dtype=np.float64, order="F") if nx > 0 and ny > 0 else \
        np.zeros((max(nx, 0), max(ny, 0)), order="F")
    ierr = _lib.lib().msx_get_perpendicular_kinetic_difference(
        float(incident_kx), float(incident_ky), float(incident_kz), int(nx), int(ny),
        float(unit_cell_ax), float(unit_cell_ay), float(unit_cell_bx), float(unit_cell_by),
        _lib.ptr(out))
    _raise_error("get_perpendicular_kinetic_difference", int(ierr))
    return out


def get_parallel_kinetic_energy(zmin, zmax, nz):
    out = np.zeros((max(nz, 0), max(nz, 0)), dtype=np.float64, order="F")
    ierr = _lib.lib().msx_get_parallel_kinetic_energy(float(zmin), float(zmax), int(nz),
                                                      _lib.ptr(out))
    _raise_error("get_parallel_kinetic_energy", int(ierr))
    return out


def get_lobatto_weights(zmin, zmax, node_count):
    w = np.zeros(max(node_count, 0), dtype=np.float64)
    x = np.zeros(max(node_count, 0), dtype=np.float64)
    ierr = _lib.lib().msx_get_lobatto_weights(float(zmin), float(zmax), int(node_count),
                                              _lib.ptr(w), _lib.ptr(x))
    _raise_error("get_lobatto_weights", int(ierr))
    return w, x


def get_abc_arrays(zmin, zmax, nx, ny, perpendicular_kinetic_difference, n_z_points):
    n = max(nx, 0) * max(ny, 0)
    e = _fortran(perpendicular_kinetic_difference, np.float64)
    a = np.zeros(n, dtype=np.complex128)
    b = np.zeros(n, dtype=np.complex128)
    c = np.zeros(n, dtype=np.complex128)
    ierr = _lib.lib().msx_get_abc_arrays(float(zmin), float(zmax), int(nx), int(ny),
                                         _lib.ptr(e), int(n_z_points), _lib.ptr(a), _lib.ptr(b),
                                         _lib.ptr(c))
    _raise_error("get_abc_arrays", int(ierr))
    return a, b, c


def debug_build_preconditioner_fortran(potential_values, perpendicular_kinetic_difference,
                                       parallel_kinetic_energy):
    V = _fortran(potential_values, np.complex128)
    nkx, nky, nz = V.shape
    e = _fortran(perpendicular_kinetic_difference, np.float64)
    T = _fortran(parallel_kinetic_energy, np.float64)
    lam = np.zeros(nz, dtype=np.float64)
    fac = np.zeros((nz, nkx * nky), dtype=np.float64, order="F")
    vec = np.zeros((nz, nz), dtype=np.float64, order="F")
    ierr = _lib.lib().msx_debug_build_preconditioner(nkx, nky, nz, _lib.ptr(V), _lib.ptr(e),
                                                     _lib.ptr(T), _lib.ptr(lam), _lib.ptr(fac),
                                                     _lib.ptr(vec))
    _raise_error("debug_build_preconditioner_fortran", int(ierr))
    return lam, fac, vec


def debug_apply_upper_block_fortran(potential_values, state_in):
    V = _fortran(potential_values, np.complex128)
    nkx, nky, nz = V.shape
    x = _fortran(state_in, np.complex128)
    if x.shape != (nz, nkx * nky):
        _raise_error("debug_apply_upper_block_fortran", 1)
    out = np.zeros_like(x, order="F")
    ierr = _lib.lib().msx_debug_apply_upper_block(nkx, nky, nz, _lib.ptr(V), _lib.ptr(x),
                                                  _lib.ptr(out))
    _raise_error("debug_apply_upper_block_fortran", int(ierr))
    return out


def debug_solve_lower_block_fortran(potential_values, wave_c, perpendicular_kinetic_difference,
                                    parallel_kinetic_energy, state_in):
    V = _fortran(potential_values, np.complex128)
    nkx, nky, nz = V.shape
    c = _flat(wave_c, np.complex128, nkx * nky, "wave_c")
    e = _fortran(perpendicular_kinetic_difference, np.float64)
    T = _fortran(parallel_kinetic_energy, np.float64)
    x = _fortran(state_in, np.complex128)
    if x.shape != (nz, nkx * nky):
        _raise_error("debug_solve_lower_block_fortran", 4)
    out = np.zeros_like(x, order="F")
    ierr = _lib.lib().msx_debug_solve_lower_block(nkx, nky, nz, _lib.ptr(V), _lib.ptr(c),
                                                  _lib.ptr(e), _lib.ptr(T), _lib.ptr(x),
                                                  _lib.ptr(out))
    _raise_error("debug_solve_lower_block_fortran", int(ierr))
    return out
