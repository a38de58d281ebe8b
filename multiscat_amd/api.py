"""Python face of the solver in the shape of the reference's public API, without slate.

Mirrors multiscat/config.py:41-227 (`OptimizationConfig`, `UnitSystem`,
`ScatteringCondition`, `momentum_from_angles`) and multiscat/multiscat/_multiscat.py:40-246
(`get_scattering_matrix`, `get_scattering_matrix_von_neumann`, `get_scattering_state`,
`get_scattering_matrix_from_state`), with the same argument names and meanings. The
reference's potential is a slate operator; here it is a `ScatteringPotential`: the Fourier
components V_G(z_k) (G on the FFT-ordered nkx x nky grid of the unit cell's reciprocal
lattice, z_k the Gauss-Lobatto nodes of [z_min, z_max]) in the condition's energy unit,
together with the two lattice vectors and the z range. Results are numpy arrays instead of
slate Arrays: intensities (nkx, nky) complex128 (the reference's dtype,
_multiscat.py:171-175) and states (nkx, nky, nz) complex128.

Every solve runs on the GPU through libmultiscat_amd; nothing here computes a solution on
the host. Backends:
  "fortran"  the reference algorithm (block Gauss-Seidel GMRES, FP64) through the drop-in
             entry point, exactly what backend="fortran" runs in the reference;
  "gpu"      the batched FP64 solver with the same preconditioner (also takes n_channels);
  "mixed"    block-Jacobi (von Neumann order 1) GMRES in mixed precision, FP64 refinement.
"""

from __future__ import annotations

import hashlib
import math
from collections import OrderedDict
from dataclasses import dataclass, field
from typing import Literal, Sequence

import numpy as np

from multiscat_amd import basis as B

# The reference takes its constants from scipy.constants (multiscat/config.py:8-19).
HBAR_SI = B.HBAR_SI
ATOMIC_MASS_SI = B.ATOMIC_MASS_SI
ANGSTROM_SI = B.ANGSTROM_SI
ELECTRON_VOLT_SI = B.ELECTRON_VOLT_SI


@dataclass(frozen=True, kw_only=True)
class OptimizationConfig:
    """Options of the GMRES solve (multiscat/config.py:41-58)."""

    precision: float = 1e-5
    max_iterations: int = 1000
    use_neumann_preconditioner: bool = True
    n_channels: int | None = None

    def fortran_parameters(self) -> tuple[int, int]:
        """(gmres_preconditioner_flag, convergence_significant_figures) as the reference
        derives them (multiscat/multiscat/_fortran.py:92-102): nsig = int(log10(1/precision)),
        so the tolerance 0.5 * 10^-nsig is the precision rounded down to a power of ten."""
        if self.precision <= 0:
            raise ValueError("Optimization precision must be greater than zero")
        flag = 1 if self.use_neumann_preconditioner else 0
        return flag, int(np.log10(1 / self.precision))


@dataclass(frozen=True, kw_only=True)
class UnitSystem:
    """Values of hbar, the atomic mass unit and the angstrom in this system
    (multiscat/config.py:61-84)."""

    hbar: float = HBAR_SI
    atomic_mass: float = ATOMIC_MASS_SI
    angstrom: float = ANGSTROM_SI

    @classmethod
    def si(cls) -> "UnitSystem":
        return cls()

    @property
    def kinetic_energy_unit(self) -> float:
        return (self.hbar ** 2) / (2 * self.atomic_mass * self.angstrom ** 2)

    @property
    def electron_volt(self) -> float:
        return ELECTRON_VOLT_SI * (UnitSystem().kinetic_energy_unit / self.kinetic_energy_unit)


SI_UNITS = UnitSystem.si()


@dataclass(frozen=True)
class ScatteringPotential:
    """V_G(z_k) on the FFT-ordered G grid at the GLL nodes of [z_min, z_max].

    values[ix, iy, k] is the Fourier component of the potential for the reciprocal vector
    G = m_x g_1 + m_y g_2 (m = FFT mode of the storage index) at node k, in energy units;
    lattice = (a_1, a_2) are the in-plane unit-cell vectors and z_range = (z_min, z_max) the
    Lobatto domain, both in length units. V_{-G} = conj(V_G) for a real potential.
    """

    values: np.ndarray
    lattice: tuple[tuple[float, float], tuple[float, float]]
    z_range: tuple[float, float]

    def __post_init__(self):
        v = np.asarray(self.values)
        if v.ndim != 3:
            raise ValueError("potential values must be (nkx, nky, nz)")
        object.__setattr__(self, "values", np.ascontiguousarray(v, dtype=np.complex128))

    @property
    def shape(self) -> tuple[int, int, int]:
        return self.values.shape

    @property
    def nodes(self) -> np.ndarray:
        """The z_k of the values (GLL nodes of z_range)."""
        return B.lobatto_rule(self.shape[2], *self.z_range)[0]

    def scaled(self, energy: float, length: float) -> "ScatteringPotential":
        (ax, ay), (bx, by) = self.lattice
        z0, z1 = self.z_range
        return ScatteringPotential(self.values * energy,
                                   ((ax * length, ay * length), (bx * length, by * length)),
                                   (z0 * length, z1 * length))

    def __eq__(self, other):
        return (isinstance(other, ScatteringPotential) and self.lattice == other.lattice
                and self.z_range == other.z_range and np.array_equal(self.values, other.values))


def momentum_from_angles(theta: float, phi: float, energy: float, mass: float, *,
                         units: UnitSystem = SI_UNITS) -> tuple[float, float, float]:
    """Incident k from angles (radians) and energy (multiscat/config.py:213-227)."""
    abs_k = (2 * mass * energy) ** 0.5 / units.hbar
    return (abs_k * math.sin(theta) * math.cos(phi), abs_k * math.sin(theta) * math.sin(phi),
            abs_k * math.cos(theta))


@dataclass(frozen=True)
class ScatteringCondition:
    """A particle of `mass` incident with wavevector `incident_k` on `potential`
    (multiscat/config.py:90-210), all in `units`."""

    mass: float
    incident_k: tuple[float, float, float]
    potential: ScatteringPotential
    units: UnitSystem = field(default_factory=UnitSystem.si)
    # Initial guess (nkx, nky, nz) for the GMRES solve, the reference's
    # ScatteringCondition.initial_state (multiscat/config.py:115-121; zero when None), in the
    # form get_scattering_state returns (L psi): the "gpu" and "mixed" backends and
    # get_scattering_matrices start GMRES from x0 = L^-1 initial_state as
    # multiscat/multiscat/_gmres.py:106-108 does, so the state of a nearby condition (energy
    # continuation) can be passed on; get_scattering_matrix_von_neumann takes it as x0 itself
    # (multiscat/multiscat/_scipy_von_neumann.py:161-170).
    initial_state: np.ndarray | None = field(default=None, compare=False)

    @staticmethod
    def from_angles(*, mass: float, theta: float = 0, phi: float = 0, energy: float,
                    potential: ScatteringPotential,
                    units: UnitSystem = SI_UNITS) -> "ScatteringCondition":
        return ScatteringCondition(mass, momentum_from_angles(theta, phi, energy, mass,
                                                              units=units), potential, units)

    @property
    def incident_energy(self) -> float:
        kx, ky, kz = self.incident_k
        return (self.units.hbar ** 2 * (kx ** 2 + ky ** 2 + kz ** 2)) / (2 * self.mass)

    @property
    def theta(self) -> float:
        kx, ky, kz = self.incident_k
        return float(np.arccos(kz / np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)))

    @property
    def phi(self) -> float:
        kx, ky, _ = self.incident_k
        return float(np.arctan2(ky, kx))

    def with_units(self, units: UnitSystem) -> "ScatteringCondition":
        """The same physical condition expressed in `units` (multiscat/config.py:230-296):
        lengths scale by units.angstrom / self.units.angstrom, masses by the atomic-mass
        ratio, and hbar^2 / (m L^2) fixes the energy scale."""
        length = units.angstrom / self.units.angstrom
        mass = units.atomic_mass / self.units.atomic_mass
        hbar = units.hbar / self.units.hbar
        energy = hbar ** 2 / (mass * length ** 2)
        kx, ky, kz = self.incident_k
        return ScatteringCondition(self.mass * mass, (kx / length, ky / length, kz / length),
                                   self.potential.scaled(energy, length), units,
                                   self.initial_state)

    def natural(self) -> "ScatteringCondition":
        """hbar = 1, m = 1/2, lengths in angstrom: energies become k^2 in 1/A^2
        (multiscat/multiscat/_multiscat.py:102-128)."""
        out = self.with_units(UnitSystem(angstrom=1.0, hbar=1.0,
                                         atomic_mass=0.5 * self.units.atomic_mass / self.mass))
        assert np.isclose(out.mass, 0.5)
        return out


def solver_inputs(condition: ScatteringCondition) -> B.SolverInputs:
    """The boundary arrays of the drop-in entry point for `condition`, in natural units
    (multiscat/multiscat/_fortran.py:105-156)."""
    nat = condition.natural()
    p = nat.potential
    return B.solver_inputs(p.values, nat.incident_k, p.lattice[0], p.lattice[1], *p.z_range)


def _intensities_from_state(psi: np.ndarray, inputs: B.SolverInputs) -> np.ndarray:
    """S_G = 2i b_G psi_G(z_N), S_0 += a_0, P = |S|^2 (_multiscat.py:40-60); a few numbers,
    formed on the host from the GPU state exactly as the reference forms them."""
    nkx, nky, _ = psi.shape
    b = inputs.wave_b.reshape(nkx, nky)
    S = 2.0j * b * psi[:, :, -1]
    S[0, 0] += inputs.wave_a[0]
    return (np.abs(S) ** 2).astype(np.complex128)


# Prepared tables of recent (potential, kinetic) pairs, keyed by a digest of their bytes: the
# reference's API solves one condition per call, and a sweep of conditions over one potential
# then computes the table's eigensystem once instead of per call.
_TABLES: "OrderedDict[bytes, object]" = OrderedDict()
_TABLES_KEPT = 2


def _shared_table(V: np.ndarray, T: np.ndarray):
    from multiscat_amd import batch as BT

    V = np.ascontiguousarray(V, dtype=np.complex128)
    T = np.ascontiguousarray(T, dtype=np.float64)
    h = hashlib.blake2b(digest_size=20)
    h.update(np.asarray(V.shape + T.shape, dtype=np.int64).tobytes())
    h.update(memoryview(V).cast("B"))
    h.update(memoryview(T).cast("B"))
    key = h.digest()
    table = _TABLES.get(key)
    if table is not None:
        _TABLES.move_to_end(key)
        return table
    table = BT.Table(V, T)
    _TABLES[key] = table
    while len(_TABLES) > _TABLES_KEPT:
        _TABLES.popitem(last=False)[1].close()
    return table


class _OneCondition:
    """A one-condition batch on the GPU for `inputs` over the channel set `channels` (solve
    order), on the prepared table of its potential (shared between calls); the batch is
    closed on exit."""

    def __init__(self, inputs: B.SolverInputs, channels):
        from multiscat_amd import batch as BT

        chx, chy = channels
        if not np.any((chx == 0) & (chy == 0)):
            raise ValueError("n_channels too small: the specular channel is not included")
        self.channels = channels
        self.shape = inputs.potential_values.shape
        self.table = _shared_table(inputs.potential_values, inputs.parallel_kinetic_energy)
        self.batch = BT.Batch(self.table, [dict(
            e=inputs.perpendicular_kinetic_difference, wave_a=inputs.wave_a,
            wave_b=inputs.wave_b, wave_c=inputs.wave_c, channels=channels)])

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.batch.close()

    def restrict(self, full: np.ndarray) -> np.ndarray:
        return np.asarray(full, dtype=np.complex128).reshape(self.shape)[self.channels]

    def expand(self, rows: np.ndarray, fill=None) -> np.ndarray:
        out = (np.zeros(self.shape, dtype=np.complex128) if fill is None
               else np.array(fill, dtype=np.complex128).reshape(self.shape))
        out[self.channels] = rows
        return out


def _scipy_channels(inputs: B.SolverInputs, n_channels):
    """The channel set and order of the reference's SciPy path: argsort of e_G, cut at
    n_channels (multiscat/multiscat/_scipy.py:414-416)."""
    return B.lowest_channels(inputs.perpendicular_kinetic_difference, n_channels)


def _batch_solve(inputs: B.SolverInputs, config: OptimizationConfig, *, method: str,
                 precision: str, neumann: int, initial_state=None,
                 initial_is_lower_state: bool = True):
    """psi (nkx, nky, nz) from the batched GPU solver on the SciPy path's channel set.

    initial_state: a state in the reference's convention, L psi (what get_scattering_state
    returns), entering GMRES as x0 = L^-1 initial_state as multiscat/multiscat/_gmres.py:106-108
    does; with initial_is_lower_state=False it is x0 itself (the von Neumann path,
    multiscat/multiscat/_scipy_von_neumann.py:161-170)."""
    from multiscat_amd import batch as BT

    flag, nsig = config.fortran_parameters()
    eps = 0.5 * 10.0 ** (-nsig)
    with _OneCondition(inputs, _scipy_channels(inputs, config.n_channels)) as one:
        warm = initial_state is not None
        if warm:
            x0 = one.restrict(initial_state)
            if initial_is_lower_state:
                x0 = one.batch.apply_lower([x0], inverse=True)[0]
            one.batch.set_state([x0])
        one.batch.solve(BT.SolveOptions(method=method, precision=precision, neumann=int(neumann),
                                        eps=eps, maxiter=max(config.max_iterations, 1),
                                        warm_start=warm))
        return one.expand(one.batch.state()[0])


def _solve_state(condition: ScatteringCondition, config: OptimizationConfig,
                 backend: str) -> tuple[np.ndarray, B.SolverInputs]:
    inputs = solver_inputs(condition)
    if backend == "fortran":
        if config.n_channels is not None:
            # the reference raises here too (multiscat/multiscat/_fortran.py:145-147)
            raise NotImplementedError("Channel filtering is not yet implemented in fortran code")
        from multiscat_amd import fortran_compat as F

        flag, nsig = config.fortran_parameters()
        psi = F.run_multiscat_fortran(
            flag, nsig, potential_values=inputs.potential_values,
            perpendicular_kinetic_difference=inputs.perpendicular_kinetic_difference,
            wave_a=inputs.wave_a, wave_b=inputs.wave_b, wave_c=inputs.wave_c,
            parallel_kinetic_energy=inputs.parallel_kinetic_energy)
    elif backend == "gpu":
        psi = _batch_solve(inputs, config, method="gauss_seidel", precision="fp64",
                           neumann=config.use_neumann_preconditioner,
                           initial_state=condition.initial_state)
    elif backend == "mixed":
        psi = _batch_solve(inputs, config, method="jacobi", precision="mixed", neumann=1,
                           initial_state=condition.initial_state)
    else:
        raise ValueError(f"Unknown backend '{backend}'. Expected 'fortran', 'gpu' or 'mixed'.")
    return psi, inputs


def get_scattering_matrix(condition: ScatteringCondition, config: OptimizationConfig, *,
                          backend: Literal["fortran", "gpu", "mixed"] = "fortran") -> np.ndarray:
    """Intensities P_G (nkx, nky) complex128 (multiscat/multiscat/_multiscat.py:131-175)."""
    psi, inputs = _solve_state(condition, config, backend)
    return _intensities_from_state(psi, inputs)


def get_scattering_matrix_von_neumann(condition: ScatteringCondition,
                                      config: OptimizationConfig, *,
                                      order: int = 1) -> np.ndarray:
    """Intensities with the block-Jacobi preconditioner and its von Neumann series
    M = sum_{k <= order} (-D^-1 V)^k D^-1 (multiscat/multiscat/_multiscat.py:178-224,
    _scipy_von_neumann.py:40-58,131-184), FP64 on the GPU, for any order >= 0; a negative
    order raises ValueError as the reference does (_scipy_von_neumann.py:145-147). The
    condition's initial_state is GMRES's x0 as given (_scipy_von_neumann.py:161-170)."""
    order = int(order)
    if order < 0:
        raise ValueError(f"Invalid von Neumann order {order}. Expected a non-negative integer.")
    inputs = solver_inputs(condition)
    psi = _batch_solve(inputs, config, method="jacobi", precision="fp64", neumann=order,
                       initial_state=condition.initial_state, initial_is_lower_state=False)
    return _intensities_from_state(psi, inputs)


def get_scattering_state(condition: ScatteringCondition, config: OptimizationConfig, *,
                         backend: Literal["fortran", "gpu", "mixed"] = "gpu") -> np.ndarray:
    """The full scattering state (nkx, nky, nz), interior included, in the reference's
    convention: L psi, with L = D + V_{i<j} the lower block operator of the Gauss-Seidel split
    in the SciPy path's channel order (argsort of e_G, cut at config.n_channels), as
    multiscat/multiscat/_multiscat.py:227-246 -> _scipy.py:580-613 -> _gmres.py:127 return it.
    psi itself comes from `backend`; L is applied on the GPU. Channels outside the set are 0.
    get_scattering_matrix_from_state and ScatteringCondition.initial_state take this form."""
    psi, inputs = _solve_state(condition, config, backend)
    with _OneCondition(inputs, _scipy_channels(inputs, config.n_channels)) as one:
        return one.expand(one.batch.apply_lower([one.restrict(psi)])[0], fill=psi)


def get_scattering_matrix_from_state(state: np.ndarray, condition: ScatteringCondition, *,
                                     n_channels: int | None = None) -> np.ndarray:
    """Intensities from a state in the form get_scattering_state returns (L psi): psi =
    L^-1 state over the channel set of `n_channels`, then P (multiscat/multiscat/
    _multiscat.py:63-99); L^-1 runs on the GPU."""
    inputs = solver_inputs(condition)
    with _OneCondition(inputs, _scipy_channels(inputs, n_channels)) as one:
        rows = one.batch.apply_lower([one.restrict(state)], inverse=True)[0]
        psi = one.expand(rows, fill=state)
    return _intensities_from_state(psi, inputs)


def get_scattering_matrices(conditions: Sequence[ScatteringCondition],
                            config: OptimizationConfig, *,
                            precision: Literal["fp64", "mixed"] = "mixed",
                            method: Literal["jacobi", "gauss_seidel"] = "jacobi") -> list:
    """Extension: many conditions on one potential table in a single batched GPU solve.
    The conditions must share the potential (same values, lattice and z range after the
    conversion to natural units); each keeps its own channel set (config.n_channels)."""
    from multiscat_amd import batch as BT

    if not conditions:
        return []
    inputs = [solver_inputs(c) for c in conditions]
    V0, T0 = inputs[0].potential_values, inputs[0].parallel_kinetic_energy
    for inp in inputs[1:]:
        if inp.potential_values.shape != V0.shape or not np.allclose(inp.potential_values, V0,
                                                                     rtol=1e-12, atol=0):
            raise ValueError("all conditions of a batch must share one potential table")
    flag, nsig = config.fortran_parameters()
    conds = []
    for inp in inputs:
        e = inp.perpendicular_kinetic_difference
        chx, chy = B.lowest_channels(e, config.n_channels)
        if not np.any((chx == 0) & (chy == 0)):
            raise ValueError("n_channels too small: the specular channel is not included")
        conds.append(dict(e=e, wave_a=inp.wave_a, wave_b=inp.wave_b, wave_c=inp.wave_c,
                          channels=(chx, chy)))
    tab = BT.Table(V0, T0)
    bt = BT.Batch(tab, conds)
    try:
        warm = any(c.initial_state is not None for c in conditions)
        if warm:
            # states in the get_scattering_state form (L psi) enter as x0 = L^-1 state
            # (multiscat/multiscat/_gmres.py:106-108); L^-1 0 = 0 for conditions without one
            x0 = [np.zeros((len(d["channels"][0]), V0.shape[2]), complex)
                  if c.initial_state is None else
                  np.asarray(c.initial_state, dtype=np.complex128)[d["channels"][0],
                                                                   d["channels"][1]]
                  for c, d in zip(conditions, conds)]
            bt.set_state(bt.apply_lower(x0, inverse=True))
        bt.solve(BT.SolveOptions(method=method, precision=precision,
                                 neumann=config.use_neumann_preconditioner,
                                 eps=0.5 * 10.0 ** (-nsig), maxiter=max(config.max_iterations, 1),
                                 warm_start=warm))
        out = []
        for c, P in zip(conds, bt.intensities()):
            full = np.zeros(V0.shape[:2], dtype=np.complex128)
            full[c["channels"][0], c["channels"][1]] = P
            out.append(full)
        return out
    finally:
        bt.close()
        tab.close()


__all__ = [
    "OptimizationConfig", "UnitSystem", "SI_UNITS", "ScatteringPotential",
    "ScatteringCondition", "momentum_from_angles", "solver_inputs", "get_scattering_matrix",
    "get_scattering_matrix_von_neumann", "get_scattering_state",
    "get_scattering_matrix_from_state", "get_scattering_matrices",
]
