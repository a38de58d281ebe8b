"""Batched GPU solves: a potential table in HBM and many incidence conditions sharing it.

This is the extension of the drop-in boundary (no counterpart in the reference): one call
solves every condition of a batch with the reference's Gauss-Seidel-split GMRES or with the
all-channel block-Jacobi (eigenbasis) preconditioner, in FP64 or mixed precision, and the
adjoint solve gives d(loss)/dV_G(z). Channel subsets follow the reference's SciPy
``n_channels`` semantics (multiscat/multiscat/_scipy.py:414-443).
"""

from __future__ import annotations

import ctypes as C
import sys
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class SolveOptions:
    """Python face of msx_solve_options. Defaults are the reference's (nsig 5)."""

    method: str = "gauss_seidel"  # or "jacobi"
    precision: str = "fp64"  # or "mixed"
    neumann: int = 1  # 0/False none, 1/True the reference's factor, p >= 2 degree p (Jacobi)
    eps: float = 0.5e-5
    maxiter: int = 2000
    restart: int = 0
    window: int = 3
    inner_eps: float = 0.0  # 0: sqrt(eps)/3 within [1e-5, 2.5e-4]
    max_refine: int = 6
    coupling: str = "auto"  # or "direct"
    diag: str = "auto"  # "auto" (fastest), "vector" or "mfma" (matrix cores, nz <= 320)
    warm_start: bool = False  # start from the batch's current state (Batch.set_state)
    inner_orth: int = 0  # FP32 inner Arnoldi: 0 one pass + selective second, 1 one pass, 2 CGS2
    autotune: bool = True  # time the main kernels' launch shapes at a batch shape's first solve
    graphs: bool = True  # replay each Arnoldi iteration from a captured HIP graph
    # mixed precision, operands of the inner operator's products on the matrix cores (D^-1
    # and the DFT coupling): "auto" (split FP16 pairs where the kernels cover the shape, else
    # FP32), "fp32", or "split_fp16" (an error where the D^-1 kernel does not cover the shape)
    inner_products: str = "auto"

    @classmethod
    def from_nsig(cls, nsig: int, **kw) -> "SolveOptions":
        return cls(eps=0.5 * 10.0 ** (-int(nsig)), **kw)

    def native(self) -> _lib.SolveOptions:
        o = _lib.SolveOptions()
        methods = {"gauss_seidel": _lib.METHOD_GAUSS_SEIDEL, "jacobi": _lib.METHOD_JACOBI}
        precisions = {"fp64": _lib.PRECISION_FP64, "mixed": _lib.PRECISION_MIXED}
        if self.method not in methods:
            raise ValueError(f"unknown method {self.method!r}")
        if self.precision not in precisions:
            raise ValueError(f"unknown precision {self.precision!r}")
        o.method = methods[self.method]
        o.precision = precisions[self.precision]
        o.neumann = int(self.neumann)
        o.eps = float(self.eps)
        o.maxiter = int(self.maxiter)
        o.restart = int(self.restart)
        o.window = int(self.window)
        o.inner_eps = float(self.inner_eps)
        o.max_refine = int(self.max_refine)
        if self.coupling not in ("auto", "direct"):
            raise ValueError(f"unknown coupling {self.coupling!r}")
        o.coupling = 1 if self.coupling == "direct" else 0
        if self.diag not in ("auto", "vector", "mfma"):
            raise ValueError(f"unknown diag {self.diag!r}")
        o.diag = {"auto": 0, "vector": 1, "mfma": 2}[self.diag]
        o.warm_start = 1 if self.warm_start else 0
        if self.inner_orth not in (0, 1, 2):
            raise ValueError(f"inner_orth must be 0, 1 or 2, not {self.inner_orth!r}")
        o.inner_orth = self.inner_orth
        o.autotune = 1 if self.autotune else 0
        o.graphs = 1 if self.graphs else 0
        products = {"auto": 0, "fp32": 1, "split_fp16": 2}
        if self.inner_products not in products:
            raise ValueError(f"unknown inner_products {self.inner_products!r}")
        o.inner_products = products[self.inner_products]
        return o


_EIG_METHODS = {"auto": 0, "ql": 1, "divide_conquer": 2}


def symmetric_eigensystem(A, method: str = "auto"):
    """(lam ascending, Q) with A = Q diag(lam) Q^T for a symmetric matrix, on the GPU in FP64:
    Householder tridiagonalisation, then divide and conquer ("divide_conquer", n <= 1024) or
    implicit QL ("ql"); "auto" picks divide and conquer where it applies."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError("A must be square")
    n = A.shape[0]
    lam = np.zeros(n)
    Q = np.zeros((n, n))
    _lib.check(_lib.lib().msx_symmetric_eigensystem(n, _lib.ptr(A), _EIG_METHODS[method],
                                                    _lib.ptr(lam), _lib.ptr(Q)),
               "msx_symmetric_eigensystem")
    return lam, Q


class Table:
    """A potential table V (nkx, nky, nz) and kinetic matrix T resident on the GPU."""

    def __init__(self, potential, kinetic):
        V = np.ascontiguousarray(potential, dtype=np.complex128)
        T = np.ascontiguousarray(kinetic, dtype=np.float64)
        if V.ndim != 3 or T.shape != (V.shape[2], V.shape[2]):
            raise ValueError("potential must be (nkx, nky, nz) and kinetic (nz, nz)")
        self.shape = V.shape
        self._h = C.c_void_p()
        _lib.check(_lib.lib().msx_table_create(V.shape[0], V.shape[1], V.shape[2], _lib.ptr(V),
                                               _lib.ptr(T), C.byref(self._h)),
                   "msx_table_create")

    def prepare(self) -> None:
        _lib.check(_lib.lib().msx_table_prepare(self._h), "msx_table_prepare")

    def set_potential(self, potential) -> None:
        """Replace V (same shape) and prepare the table again. `potential` is a NumPy array or
        a complex128 torch tensor; a contiguous tensor on the GPU is read in place (after its
        stream's pending work), anything else through host memory."""
        torch = sys.modules.get("torch")
        if torch is not None and isinstance(potential, torch.Tensor):
            if tuple(potential.shape) != tuple(self.shape) or potential.dtype != torch.complex128:
                raise ValueError(f"potential must be complex128 of shape {self.shape}")
            if potential.is_cuda and potential.is_contiguous():
                torch.cuda.current_stream(potential.device).synchronize()
                _lib.check(_lib.lib().msx_table_set_potential(
                    self._h, C.c_void_p(potential.data_ptr()), 1), "msx_table_set_potential")
                return
            potential = potential.detach().cpu().numpy()
        V = np.ascontiguousarray(potential, dtype=np.complex128)
        if V.shape != tuple(self.shape):
            raise ValueError(f"potential must have shape {self.shape}, not {V.shape}")
        _lib.check(_lib.lib().msx_table_set_potential(self._h, _lib.ptr(V), 0),
                   "msx_table_set_potential")

    def eigensystem(self):
        nz = self.shape[2]
        lam = np.zeros(nz)
        Q = np.zeros((nz, nz))
        _lib.check(_lib.lib().msx_table_eigensystem(self._h, _lib.ptr(lam), _lib.ptr(Q)),
                   "msx_table_eigensystem")
        return lam, Q

    def eig_phases(self) -> np.ndarray:
        """Last eigensolve: ms of its phases (tridiagonalisation, reflector accumulation,
        tridiagonal solve, back-transformation or reordering), then for the QL solver the ms
        of bulge chasing and the numbers of sweeps and rotations."""
        ms = np.zeros(7)
        _lib.check(_lib.lib().msx_table_eig_phases(self._h, _lib.ptr(ms)), "msx_table_eig_phases")
        return ms

    def close(self) -> None:
        if self._h:
            _lib.lib().msx_table_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Conditions on one table. ``conditions``: iterable of dicts or objects with
    ``channels`` ((chx, chy) int arrays, solve order), ``e`` (nkx, nky), ``wave_b``,
    ``wave_c`` (nkx*nky, C-ravel) and ``wave_a`` (its [0] is a_0)."""

    def __init__(self, table: Table, conditions):
        self.table = table
        nkx, nky, nz = table.shape
        chx, chy, e, b, c, a0, counts = [], [], [], [], [], [], []
        for cond in conditions:
            get = (lambda k: cond[k]) if isinstance(cond, dict) else (lambda k: getattr(cond, k))
            cx_, cy_ = (np.asarray(v, dtype=np.int32) for v in get("channels"))
            ev = np.asarray(get("e"), dtype=np.float64).reshape(nkx, nky)
            bv = np.asarray(get("wave_b"), dtype=np.complex128).reshape(nkx, nky)
            cv = np.asarray(get("wave_c"), dtype=np.complex128).reshape(nkx, nky)
            av = np.asarray(get("wave_a"), dtype=np.complex128).ravel()
            chx.append(cx_)
            chy.append(cy_)
            e.append(ev[cx_, cy_])
            b.append(bv[cx_, cy_])
            c.append(cv[cx_, cy_])
            a0.append(av[0])
            counts.append(cx_.size)
        self.counts = np.asarray(counts, dtype=np.int32)
        self.offsets = np.concatenate(([0], np.cumsum(self.counts)[:-1])).astype(np.int64)
        self.chx = np.ascontiguousarray(np.concatenate(chx), dtype=np.int32)
        self.chy = np.ascontiguousarray(np.concatenate(chy), dtype=np.int32)
        self._e = np.ascontiguousarray(np.concatenate(e))
        self._b = np.ascontiguousarray(np.concatenate(b))
        self._c = np.ascontiguousarray(np.concatenate(c))
        self._a0 = np.ascontiguousarray(np.asarray(a0, dtype=np.complex128))
        self.n = len(counts)
        self.nz = nz
        self._h = C.c_void_p()
        _lib.check(_lib.lib().msx_batch_create(
            table._h, self.n, _lib.ptr(self.counts), _lib.ptr(self.chx), _lib.ptr(self.chy),
            _lib.ptr(self._e), _lib.ptr(self._b), _lib.ptr(self._c), _lib.ptr(self._a0),
            C.byref(self._h)), "msx_batch_create")

    @property
    def total_channels(self) -> int:
        return int(self.counts.sum())

    def solve(self, options: SolveOptions | None = None, *, check: bool = True) -> int:
        o = (options or SolveOptions()).native()
        rc = _lib.lib().msx_batch_solve(self._h, C.byref(o))
        if rc not in (_lib.MSX_OK, _lib.MSX_ERR_NOT_CONVERGED) or (check and rc != _lib.MSX_OK):
            raise _lib.MsxError(rc, f"msx_batch_solve: {_lib.last_error()}")
        return rc

    def state(self):
        """psi per condition: list of (n_channels, nz) complex arrays."""
        flat = np.zeros((self.total_channels, self.nz), dtype=np.complex128)
        _lib.check(_lib.lib().msx_batch_state(self._h, _lib.ptr(flat)), "msx_batch_state")
        return [flat[o:o + n] for o, n in zip(self.offsets, self.counts)]

    def set_state(self, psi) -> None:
        """Initial guess for a solve with warm_start: a list of (n_channels, nz) arrays (one
        per condition, channel order as given) or one flat array of the whole batch."""
        if isinstance(psi, (list, tuple)):
            flat = np.concatenate([np.asarray(p, dtype=np.complex128).reshape(-1, self.nz)
                                   for p in psi])
        else:
            flat = np.asarray(psi, dtype=np.complex128).reshape(-1, self.nz)
        if flat.shape[0] != self.total_channels:
            raise ValueError("state must have one row of nz values per channel of the batch")
        flat = np.ascontiguousarray(flat)
        _lib.check(_lib.lib().msx_batch_set_state(self._h, _lib.ptr(flat)), "msx_batch_set_state")

    def apply_lower(self, psi, *, inverse: bool = False):
        """L psi (or L^-1 psi with inverse) per condition, L = D + V_{i<j} the lower block
        operator of the Gauss-Seidel split in each condition's channel order, on the GPU
        (msx_batch_apply_lower; the reference's `lower` / `inverse_lower` operators,
        multiscat/multiscat/_scipy.py:499-527). psi as Batch.set_state takes it; returns a
        list of (n_channels, nz) arrays."""
        if isinstance(psi, (list, tuple)):
            flat = np.concatenate([np.asarray(p, dtype=np.complex128).reshape(-1, self.nz)
                                   for p in psi])
        else:
            flat = np.asarray(psi, dtype=np.complex128).reshape(-1, self.nz)
        if flat.shape[0] != self.total_channels:
            raise ValueError("state must have one row of nz values per channel of the batch")
        flat = np.ascontiguousarray(flat)
        out = np.empty_like(flat)
        _lib.check(_lib.lib().msx_batch_apply_lower(self._h, 1 if inverse else 0,
                                                    _lib.ptr(flat), _lib.ptr(out)),
                   "msx_batch_apply_lower")
        return [out[o:o + n] for o, n in zip(self.offsets, self.counts)]

    def intensities(self):
        """P per condition: list of (n_channels,) arrays, channel order as given."""
        P = np.zeros(self.total_channels)
        _lib.check(_lib.lib().msx_batch_intensities(self._h, _lib.ptr(P)), "msx_batch_intensities")
        return [P[o:o + n] for o, n in zip(self.offsets, self.counts)]

    def intensities_flat(self):
        P = np.zeros(self.total_channels)
        _lib.check(_lib.lib().msx_batch_intensities(self._h, _lib.ptr(P)), "msx_batch_intensities")
        return P

    def stats(self):
        it = np.zeros(self.n, dtype=np.int32)
        info = np.zeros(self.n, dtype=np.int32)
        res = np.zeros(self.n)
        _lib.check(_lib.lib().msx_batch_stats(self._h, _lib.ptr(it), _lib.ptr(info),
                                              _lib.ptr(res)), "msx_batch_stats")
        return {"iterations": it, "info": info, "residual": res}

    def tuning(self) -> dict:
        """The autotuner's choices for this batch (empty before a tuned solve)."""
        import json

        buf = C.create_string_buffer(8192)
        _lib.check(_lib.lib().msx_batch_tuning(self._h, buf, 8192), "msx_batch_tuning")
        return json.loads(buf.value.decode() or "{}")

    def device_state_ptr(self) -> int:
        p = C.c_void_p()
        _lib.check(_lib.lib().msx_batch_device_state(self._h, C.byref(p)),
                   "msx_batch_device_state")
        return int(p.value or 0)

    def adjoint(self, dL_dP, options: SolveOptions | None = None):
        """dL/dV (nkx, nky, nz) complex (d/dRe V + i d/dIm V) for a loss of the intensities."""
        g = np.ascontiguousarray(np.asarray(dL_dP, dtype=np.float64).ravel())
        if g.size != self.total_channels:
            raise ValueError("dL_dP must have one entry per channel of the batch")
        grad = np.zeros(self.table.shape, dtype=np.complex128)
        o = (options or SolveOptions()).native()
        _lib.check(_lib.lib().msx_batch_adjoint(self._h, C.byref(o), _lib.ptr(g), _lib.ptr(grad)),
                   "msx_batch_adjoint")
        return grad

    def close(self) -> None:
        if self._h:
            _lib.lib().msx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
