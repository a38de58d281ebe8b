"""ctypes binding of libmultiscat_amd.so (declarations in include/multiscat_amd.h).

The library is built in-tree (``python -m multiscat_amd.build``). There is no fallback:
if the shared library is missing or cannot be loaded, every entry point raises.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmultiscat_amd.so"

MSX_OK = 0
MSX_ERR_DIMS = 1
MSX_ERR_NOT_CONVERGED = 2
MSX_ERR_INVALID = 3
MSX_ERR_NO_SPECULAR = 4
MSX_ERR_EIGEN = 5
MSX_ERR_DEVICE = 6
MSX_ERR_OUT_OF_MEMORY = 7

METHOD_GAUSS_SEIDEL = 0
METHOD_JACOBI = 1
PRECISION_FP64 = 0
PRECISION_MIXED = 1

# Every symbol include/multiscat_amd.h declares.
EXPORTED = (
    "msx_last_error", "msx_device_info", "msx_synchronize", "msx_run_multiscat_fortran",
    "msx_run_multiscat_fortran_ex", "msx_get_perpendicular_kinetic_difference",
    "msx_get_parallel_kinetic_energy", "msx_get_lobatto_weights", "msx_get_abc_arrays",
    "msx_debug_build_preconditioner", "msx_debug_apply_upper_block",
    "msx_debug_solve_lower_block", "msx_debug_apply_coupling", "msx_debug_apply_diagonal",
    "msx_default_options", "msx_table_create",
    "msx_table_prepare", "msx_table_set_potential", "msx_table_destroy", "msx_table_eigensystem", "msx_table_eig_phases",
    "msx_symmetric_eigensystem", "msx_symmetric_eigensystem_noalloc", "msx_batch_create",
    "msx_batch_destroy", "msx_batch_solve", "msx_batch_state", "msx_batch_set_state",
    "msx_batch_intensities",
    "msx_batch_stats", "msx_batch_apply_lower", "msx_batch_tuning", "msx_batch_device_state",
    "msx_batch_total_channels",
    "msx_batch_adjoint", "msx_profile_enable", "msx_profile_reset", "msx_profile_read",
    "msx_profile_names", "msx_profile_select",
)


class SolveOptions(C.Structure):
    """Mirror of msx_solve_options."""

    _fields_ = [
        ("method", C.c_int),
        ("precision", C.c_int),
        ("neumann", C.c_int),
        ("eps", C.c_double),
        ("maxiter", C.c_int),
        ("restart", C.c_int),
        ("window", C.c_int),
        ("inner_eps", C.c_double),
        ("max_refine", C.c_int),
        ("coupling", C.c_int),
        ("diag", C.c_int),
        ("warm_start", C.c_int),
        ("inner_orth", C.c_int),
        ("autotune", C.c_int),
        ("graphs", C.c_int),
        ("inner_products", C.c_int),
    ]


class MsxError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"multiscat_amd error {code}: {message}")
        self.code = code


_lock = threading.Lock()
_lib = None


def _declare(lib):
    vp, i32, f64, ll = C.c_void_p, C.c_int, C.c_double, C.c_longlong
    pp = C.POINTER(C.c_void_p)
    sig = {
        "msx_last_error": (C.c_char_p, []),
        "msx_device_info": (i32, [C.c_char_p, i32, C.POINTER(i32)]),
        "msx_synchronize": (i32, []),
        "msx_run_multiscat_fortran": (i32, [i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp]),
        "msx_run_multiscat_fortran_ex": (
            i32, [i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, C.POINTER(i32),
                  C.POINTER(f64)]),
        "msx_get_perpendicular_kinetic_difference": (
            i32, [f64, f64, f64, i32, i32, f64, f64, f64, f64, vp]),
        "msx_get_parallel_kinetic_energy": (i32, [f64, f64, i32, vp]),
        "msx_get_lobatto_weights": (i32, [f64, f64, i32, vp, vp]),
        "msx_get_abc_arrays": (i32, [f64, f64, i32, i32, vp, i32, vp, vp, vp]),
        "msx_debug_build_preconditioner": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp]),
        "msx_debug_apply_upper_block": (i32, [i32, i32, i32, vp, vp, vp]),
        "msx_debug_solve_lower_block": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp]),
        "msx_debug_apply_coupling": (i32, [i32, i32, i32, vp, i32, vp, vp, vp, vp, i32, i32, i32]),
        "msx_debug_apply_diagonal": (i32, [i32, vp, vp, i32, vp, vp, vp, vp, i32, i32, i32]),
        "msx_default_options": (None, [C.POINTER(SolveOptions), i32]),
        "msx_table_create": (i32, [i32, i32, i32, vp, vp, pp]),
        "msx_table_prepare": (i32, [vp]),
        "msx_table_set_potential": (i32, [vp, vp, i32]),
        "msx_table_destroy": (i32, [vp]),
        "msx_table_eigensystem": (i32, [vp, vp, vp]),
        "msx_table_eig_phases": (i32, [vp, vp]),
        "msx_symmetric_eigensystem": (i32, [i32, vp, i32, vp, vp]),
        "msx_symmetric_eigensystem_noalloc": (i32, [i32, vp, vp, vp, vp, vp]),
        "msx_batch_create": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, vp, pp]),
        "msx_batch_destroy": (i32, [vp]),
        "msx_batch_solve": (i32, [vp, C.POINTER(SolveOptions)]),
        "msx_batch_state": (i32, [vp, vp]),
        "msx_batch_set_state": (i32, [vp, vp]),
        "msx_batch_intensities": (i32, [vp, vp]),
        "msx_batch_stats": (i32, [vp, vp, vp, vp]),
        "msx_batch_apply_lower": (i32, [vp, i32, vp, vp]),
        "msx_batch_tuning": (i32, [vp, C.c_char_p, i32]),
        "msx_batch_device_state": (i32, [vp, pp]),
        "msx_batch_total_channels": (i32, [vp, C.POINTER(ll)]),
        "msx_batch_adjoint": (i32, [vp, C.POINTER(SolveOptions), vp, vp]),
        "msx_profile_enable": (i32, [i32]),
        "msx_profile_reset": (i32, []),
        "msx_profile_read": (i32, [C.c_char_p, C.POINTER(ll), C.POINTER(f64), C.POINTER(f64),
                                   C.POINTER(f64)]),
        "msx_profile_names": (i32, [C.c_char_p, i32]),
        "msx_profile_select": (i32, [C.c_char_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """The loaded library; raises if it is missing (no CPU fallback exists)."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -m multiscat_amd.build`")
            handle = C.CDLL(str(LIB_PATH))
            _declare(handle)
            _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().msx_last_error()
    return msg.decode() if msg else ""


def check(code: int, what: str) -> None:
    if code != MSX_OK:
        raise MsxError(code, f"{what}: {last_error()}")


def ptr(a) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def default_options(nsig: int) -> SolveOptions:
    o = SolveOptions()
    lib().msx_default_options(C.byref(o), int(nsig))
    return o


def profile_enable(on: bool = True) -> None:
    lib().msx_profile_enable(1 if on else 0)
    lib().msx_profile_reset()


def profile_select(names=None) -> None:
    """Instrument only the listed kernel names (None: every kernel)."""
    lib().msx_profile_select(",".join(names).encode() if names else None)


def profile_report() -> dict:
    """{kernel: {"count", "ms", "flops", "bytes"}} summed over the instrumented launches."""
    buf = C.create_string_buffer(4096)
    lib().msx_profile_names(buf, 4096)
    out = {}
    for name in filter(None, buf.value.decode().split(",")):
        n, ms, fl, by = C.c_longlong(0), C.c_double(0), C.c_double(0), C.c_double(0)
        lib().msx_profile_read(name.encode(), C.byref(n), C.byref(ms), C.byref(fl), C.byref(by))
        out[name] = {"count": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}
    return out


def device_info():
    buf = C.create_string_buffer(64)
    n = C.c_int(0)
    check(lib().msx_device_info(buf, 64, C.byref(n)), "msx_device_info")
    return buf.value.decode(), n.value
