"""Host-side input builders for the close-coupling solve (numpy, FP64).

These produce the arrays the solver boundary takes -- the Lobatto (GLL) rule,
the DVR kinetic matrix T, the channel energies e_G, the boundary waves a/b/c
and the natural-unit potential table -- restated from closed forms:

* GLL rule: nodes are +-1 and the roots of P'_{N}(x), N = n-1, weights
  w_i = 2 / (N (N+1) P_N(x_i)^2), mapped to [z_min, z_max]
  (lib/multiscat_fortran/fortran/scatsub.f90:229-305; slate's
  LobattoSpacedLengthMetadata as used at multiscat/multiscat/_util.py:74).
* Kinetic matrix T_ij = sum_k w_k u_i'(z_k) u_j'(z_k) with u_i = l_i / sqrt(w_i),
  over all n nodes including both ends (multiscat/polynomial.py:54-128;
  multiscat/multiscat/_fortran.py:127 passes minus slate's operator).
  The spectral-element derivative matrix of Lagrange polynomials at GLL nodes
  is used: D_ki = P_N(x_k) / (P_N(x_i) (x_k - x_i)), D_00 = -N(N+1)/4,
  D_NN = N(N+1)/4.
* e_G = |K + G|^2 - |k|^2 over the FFT-ordered (nkx, nky) grid
  (scatsub.f90:54-86; _util.py:58-75).
* a = exp(-2i theta), b = sqrt(dk) exp(-i theta) / sqrt(w_N) for open channels,
  theta = dk (L - z_min) (_util.py:78-113); c = i sqrt(-e) / w_N
  (_util.py:135-147).
* Natural units hbar = 1, m = 1/2: energies in 1/A^2 via 2m/hbar^2
  (multiscat/multiscat/_multiscat.py:102-128; multiscat/config.py:61-84).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# CODATA values as used by scipy.constants (the reference's unit source,
# multiscat/config.py:8-19).
HBAR_SI = 1.054571817e-34
ATOMIC_MASS_SI = 1.66053906660e-27
ELECTRON_VOLT_SI = 1.602176634e-19
ANGSTROM_SI = 1e-10
ALPHA_PARTICLE_MASS_SI = 6.6446573357e-27


def mev_to_natural(mass_kg: float) -> float:
    """Factor 2 m / hbar^2 converting meV to 1/A^2 for a particle of ``mass_kg``."""
    return 2.0 * mass_kg * (1e-3 * ELECTRON_VOLT_SI) * ANGSTROM_SI**2 / HBAR_SI**2


def _legendre(N: int, x: np.ndarray):
    """P_N(x) and P_{N-1}(x) by Bonnet's recursion."""
    p_prev = np.ones_like(x)
    p = x.copy()
    if N == 0:
        return p_prev, np.zeros_like(x)
    for k in range(2, N + 1):
        p_prev, p = p, ((2 * k - 1) * x * p - (k - 1) * p_prev) / k
    return p, p_prev


def gll_reference_rule(n: int):
    """GLL nodes (ascending) and weights on [-1, 1], plus P_N at the nodes."""
    if n < 2:
        raise ValueError("a Lobatto rule needs at least 2 nodes")
    N = n - 1
    # Chebyshev-Gauss-Lobatto start, Newton on (1 - x^2) P_N'(x) = N (P_{N-1} - x P_N).
    x = -np.cos(np.pi * np.arange(n) / N)
    interior = x[1:-1].copy()
    for _ in range(100):
        p, pm = _legendre(N, interior)
        one_m = 1.0 - interior**2
        dp = N * (pm - interior * p) / one_m
        ddp = (2.0 * interior * dp - N * (N + 1) * p) / one_m
        step = dp / ddp
        interior -= step
        if np.max(np.abs(step), initial=0.0) < 1e-16:
            break
    # Enforce the exact mirror symmetry of the rule.
    interior = 0.5 * (interior - interior[::-1])
    x = np.concatenate(([-1.0], interior, [1.0]))
    pN, _ = _legendre(N, x)
    w = 2.0 / (N * (N + 1) * pN**2)
    return x, w, pN


def lobatto_rule(n: int, z_min: float, z_max: float):
    """GLL nodes and weights on [z_min, z_max] (ascending, both ends included)."""
    t, w, _ = gll_reference_rule(n)
    half = 0.5 * (z_max - z_min)
    return z_min + half * (t + 1.0), w * half


def derivative_matrix(n: int, z_min: float, z_max: float) -> np.ndarray:
    """D[k, i] = l_i'(z_k) for the Lagrange polynomials on the GLL nodes."""
    t, _, pN = gll_reference_rule(n)
    N = n - 1
    diff = t[:, None] - t[None, :]
    np.fill_diagonal(diff, 1.0)
    D = (pN[:, None] / pN[None, :]) / diff
    np.fill_diagonal(D, 0.0)
    D[0, 0] = -N * (N + 1) / 4.0
    D[N, N] = N * (N + 1) / 4.0
    return D * (2.0 / (z_max - z_min))


def kinetic_matrix(n: int, z_min: float, z_max: float) -> np.ndarray:
    """T_ij = sum_k w_k u_i'(z_k) u_j'(z_k), u_i = l_i / sqrt(w_i) (polynomial.py:104-128)."""
    _, w = lobatto_rule(n, z_min, z_max)
    D = derivative_matrix(n, z_min, z_max)
    U = D / np.sqrt(w)[None, :]  # U[k, i] = u_i'(z_k)
    T = U.T @ (w[:, None] * U)
    return 0.5 * (T + T.T)


def fft_modes(n: int) -> np.ndarray:
    """Storage index -> FFT mode (multiscat_gmres.f90:50-56)."""
    i = np.arange(n)
    return np.where(i > (n - 1) // 2, i - n, i)


def reciprocal_vectors(a1, a2):
    """2-D reciprocal basis (scatsub.f90:38-52)."""
    ax, ay = float(a1[0]), float(a1[1])
    bx, by = float(a2[0]), float(a2[1])
    area = abs(ax * by - ay * bx)
    if area <= 0:
        raise ValueError("unit cell area must be positive")
    f = 2.0 * math.pi / area
    return np.array([by * f, -bx * f]), np.array([-ay * f, ax * f])


def perpendicular_kinetic_difference(incident_k, a1, a2, nkx: int, nky: int) -> np.ndarray:
    """e[ix, iy] = |K + G|^2 - |k|^2 for FFT-ordered G (scatsub.f90:54-86)."""
    ga, gb = reciprocal_vectors(a1, a2)
    mx = fft_modes(nkx)[:, None]
    my = fft_modes(nky)[None, :]
    gx = ga[0] * mx + gb[0] * my
    gy = ga[1] * mx + gb[1] * my
    kx, ky, kz = (float(v) for v in incident_k)
    return (kx + gx) ** 2 + (ky + gy) ** 2 - (kx * kx + ky * ky + kz * kz)


def ab_waves(e: np.ndarray, z_min: float, z_max: float, n: int):
    """Incoming/outgoing amplitude factors a, b (multiscat/multiscat/_util.py:78-113)."""
    _, w = lobatto_rule(n, z_min, z_max)
    e = np.asarray(e, dtype=np.float64)
    dk = np.sqrt(np.abs(e))
    a = np.zeros(e.shape, dtype=np.complex128)
    b = np.zeros(e.shape, dtype=np.complex128)
    open_ = e < 0.0
    theta = dk[open_] * ((z_max - z_min) - z_min)
    a[open_] = np.exp(-2.0j * theta)
    b[open_] = np.sqrt(dk[open_]) * np.exp(-1.0j * theta)
    return a, b / math.sqrt(w[-1])


def outgoing_log_derivative_wave(e: np.ndarray, z_min: float, z_max: float, n: int):
    """c = i sqrt(-e) / w_N (multiscat/multiscat/_util.py:135-147)."""
    _, w = lobatto_rule(n, z_min, z_max)
    return 1j * np.emath.sqrt(-np.asarray(e, dtype=np.float64)) / w[-1]


def momentum_from_angles(theta: float, phi: float, energy_natural: float):
    """Incident k in 1/A from angles (radians) and E in 1/A^2 (multiscat/config.py:213-227)."""
    k = math.sqrt(energy_natural)
    return (k * math.sin(theta) * math.cos(phi), k * math.sin(theta) * math.sin(phi),
            k * math.cos(theta))


def lowest_channels(e: np.ndarray, n_channels: int | None):
    """The n lowest-e channels in argsort order, with numpy's default (not stable) sort exactly
    as multiscat/multiscat/_scipy.py:414-416 calls it: where e_G ties (normal incidence, mirror
    symmetry) and the cut falls inside the tie, the set kept is the one the reference keeps."""
    flat = np.asarray(e).ravel()
    order = np.argsort(flat)
    if n_channels is not None:
        order = order[:n_channels]
    ix, iy = np.divmod(order, e.shape[1])
    return ix.astype(np.int32), iy.astype(np.int32)


@dataclass(frozen=True)
class SolverInputs:
    """Everything the solver boundary takes, in natural units."""

    potential_values: np.ndarray  # (nkx, nky, nz) complex128
    perpendicular_kinetic_difference: np.ndarray  # (nkx, nky) float64
    wave_a: np.ndarray  # (nkx*nky,) complex128
    wave_b: np.ndarray
    wave_c: np.ndarray
    parallel_kinetic_energy: np.ndarray  # (nz, nz) float64


def solver_inputs(potential_natural, incident_k, a1, a2, z_min, z_max) -> SolverInputs:
    """Assemble the boundary inputs the way the reference caller does
    (multiscat/multiscat/_fortran.py:105-156)."""
    V = np.ascontiguousarray(potential_natural, dtype=np.complex128)
    nkx, nky, nz = V.shape
    e = perpendicular_kinetic_difference(incident_k, a1, a2, nkx, nky)
    T = kinetic_matrix(nz, z_min, z_max)
    a, b = ab_waves(e, z_min, z_max, nz)
    c = outgoing_log_derivative_wave(e.ravel(), z_min, z_max, nz)
    return SolverInputs(V, e, a.ravel(), b.ravel(), c.ravel(), T)
