"""Synthetic scattering workloads (potential tables, incidence conditions, channel sets).

All of them build inputs in the reference's conventions: a (nkx, nky, nz) table
of Fourier components V_G(z_k) on GLL nodes, FFT-ordered G, natural units
(energies times 2m/hbar^2), channels either the full grid in storage order
(the Fortran path) or the n lowest e_G in argsort order (the SciPy path's
``n_channels``, multiscat/multiscat/_scipy.py:414-416).

Named workloads:

* ``square_morse``  -- He (4 amu) on a square lattice a = 2.5 A, Morse
  D = 8 meV, alpha = 1.2 1/A, z0 = 0, corrugation 0.05 D exp(-2 alpha z)
  [cos(2 pi x/a) + cos(2 pi y/a)], 11 x 11 table, E = 20 meV normal
  incidence, channels with e_G < 4 k^2, 100 GLL nodes on [-1.5, 8] A.
* ``hex_table``     -- hexagonal a = 3.16 A, 31 x 31 (or 65 x 65) table: Morse
  V_00 (D = 8 meV, alpha = 1.2, z0 = 0) plus V_G = c_G exp(-2 alpha z)
  exp(-|G|^2 sigma^2), complex Gaussian c_G of r.m.s. 0.3 meV drawn from a
  seed on a half plane and V_-G = conj(V_G); z in [-1.5, 8] A.
* ``hex_single``    -- hex_table (sigma 0.4 A, 200 nodes), E = 60 meV at
  30 deg polar angle, the 700 lowest channels.
* ``hex_batch``     -- hex_table (sigma 0.4 A, 150 nodes) at 256 seeded
  incidence conditions E ~ U(10, 100) meV, polar ~ U(0, 60) deg,
  azimuth ~ U(0, 60) deg, each with its own channel count: the channels with
  |K+G|^2 <= 10 k^2 (the rule that gives hex_single's point, 60 meV at 30 deg,
  706 channels), clipped to [200, 1200]: 200..876, mean 618 at the defaults.
* ``hex_large``     -- 65 x 65 table, sigma 0.25 A, 300 nodes, 32 energies
  E ~ U(20, 150) meV at normal incidence, the 4000 lowest channels.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import basis as B

HE4_MASS_KG = 4.0 * B.ATOMIC_MASS_SI
Z_MIN, Z_MAX = -1.5, 8.0


@dataclass
class Condition:
    """One incidence condition on a shared potential table (natural units)."""

    energy_mev: float
    theta: float
    phi: float
    incident_k: tuple
    e: np.ndarray  # (nkx, nky)
    wave_a: np.ndarray  # (nkx*nky,)
    wave_b: np.ndarray
    wave_c: np.ndarray
    channels: tuple  # (chx, chy) int32, solve order


@dataclass
class Workload:
    name: str
    potential: np.ndarray  # (nkx, nky, nz) complex128, natural units
    kinetic: np.ndarray  # (nz, nz)
    a1: tuple
    a2: tuple
    z_min: float
    z_max: float
    conditions: list = field(default_factory=list)

    @property
    def shape(self):
        return self.potential.shape

    def channel_counts(self):
        return [c.channels[0].size for c in self.conditions]


def morse(z, depth, alpha, z0):
    return depth * (np.exp(-2.0 * alpha * (z - z0)) - 2.0 * np.exp(-alpha * (z - z0)))


def _condition(V, a1, a2, z_min, z_max, energy_mev, theta, phi, n_channels, factor):
    nkx, nky, nz = V.shape
    k = B.momentum_from_angles(theta, phi, energy_mev * factor)
    e = B.perpendicular_kinetic_difference(k, a1, a2, nkx, nky)
    a, b = B.ab_waves(e, z_min, z_max, nz)
    c = B.outgoing_log_derivative_wave(e.ravel(), z_min, z_max, nz)
    if n_channels is None:
        ix, iy = np.divmod(np.arange(nkx * nky), nky)
        channels = (ix.astype(np.int32), iy.astype(np.int32))
    else:
        # the lowest-n set must reach the specular channel (e_00 = -kz^2 is not always the
        # lowest e_G at oblique incidence)
        rank_spec = int(np.nonzero(np.argsort(e.ravel()) == 0)[0][0])
        channels = B.lowest_channels(e, max(int(n_channels), rank_spec + 1))
    return Condition(energy_mev, theta, phi, k, e, a.ravel(), b.ravel(), c.ravel(), channels)


def square_morse(nz: int = 100, full_grid: bool = False) -> Workload:
    factor = B.mev_to_natural(HE4_MASS_KG)
    a = 2.5
    a1, a2 = (a, 0.0), (0.0, a)
    z, _ = B.lobatto_rule(nz, Z_MIN, Z_MAX)
    nk = 11
    V = np.zeros((nk, nk, nz), dtype=np.complex128)
    D, alpha = 8.0, 1.2
    V[0, 0] = morse(z, D, alpha, 0.0)
    corr = 0.025 * D * np.exp(-2.0 * alpha * z)
    for (ix, iy) in ((1, 0), (nk - 1, 0), (0, 1), (0, nk - 1)):
        V[ix, iy] = corr
    V *= factor
    T = B.kinetic_matrix(nz, Z_MIN, Z_MAX)
    w = Workload("square_morse", V, T, a1, a2, Z_MIN, Z_MAX)
    k2 = 20.0 * factor
    e = B.perpendicular_kinetic_difference(B.momentum_from_angles(0, 0, k2), a1, a2, nk, nk)
    n_ch = None if full_grid else int(np.count_nonzero(e < 4.0 * k2))
    w.conditions.append(_condition(V, a1, a2, Z_MIN, Z_MAX, 20.0, 0.0, 0.0, n_ch, factor))
    return w


def hex_table(nk: int, nz: int, sigma: float, seed: int = 1234, rms_mev: float = 0.3,
              cg_noise: float = 0.0, noise_seed: int = 4321):
    """Corrugated hexagonal table in meV: (V (nk, nk, nz), a1, a2, z). cg_noise > 0 multiplies
    every c_G by (1 + cg_noise xi), xi standard normal from noise_seed (the perturbed potential
    of a fitting target); the c_G themselves are the same draws as without noise."""
    a = 3.16
    a1, a2 = (a, 0.0), (0.5 * a, 0.5 * math.sqrt(3.0) * a)
    z, _ = B.lobatto_rule(nz, Z_MIN, Z_MAX)
    ga, gb = B.reciprocal_vectors(a1, a2)
    modes = B.fft_modes(nk)
    rng = np.random.default_rng(seed)
    noise = np.random.default_rng(noise_seed)
    V = np.zeros((nk, nk, nz), dtype=np.complex128)
    D, alpha = 8.0, 1.2
    V[0, 0] = morse(z, D, alpha, 0.0)
    decay = np.exp(-2.0 * alpha * z)
    half = rms_mev / math.sqrt(2.0)
    for ix in range(nk):
        for iy in range(nk):
            mx, my = int(modes[ix]), int(modes[iy])
            if not (mx > 0 or (mx == 0 and my > 0)):
                continue
            g = mx * ga + my * gb
            cg = complex(rng.normal(0.0, half), rng.normal(0.0, half))
            if cg_noise:
                cg *= 1.0 + cg_noise * float(noise.standard_normal())
            val = cg * math.exp(-float(g @ g) * sigma * sigma) * decay
            V[ix, iy] = val
            V[(-ix) % nk, (-iy) % nk] = np.conj(val)
    return V, a1, a2, z


def hex_single(nz: int = 200, n_channels: int = 700, seed: int = 1234) -> Workload:
    factor = B.mev_to_natural(HE4_MASS_KG)
    Vm, a1, a2, _ = hex_table(31, nz, 0.4, seed)
    V = Vm * factor
    T = B.kinetic_matrix(nz, Z_MIN, Z_MAX)
    w = Workload("hex_single", V, T, a1, a2, Z_MIN, Z_MAX)
    w.conditions.append(_condition(V, a1, a2, Z_MIN, Z_MAX, 60.0, math.radians(30.0), 0.0,
                                   n_channels, factor))
    return w


BATCH_CUTOFF_RATIO = 10.0


def batch_channel_count(e: np.ndarray, k2: float, lo: int, hi: int,
                        ratio: float = BATCH_CUTOFF_RATIO) -> int:
    """Channels with |K+G|^2 <= ratio k^2 (e_G <= (ratio-1) k^2), clipped to [lo, hi]."""
    n = int(np.count_nonzero(e <= (ratio - 1.0) * k2))
    return int(min(max(n, lo), hi, e.size))


def hex_batch(n_conditions: int = 256, nz: int = 150, seed: int = 1234, cond_seed: int = 99,
              nk: int = 31, lo: int = 200, hi: int = 1200, channel_counts=None,
              cg_noise: float = 0.0) -> Workload:
    """``channel_counts`` (one per condition) overrides the |K+G|^2 <= 10 k^2 rule;
    ``cg_noise`` perturbs the table's c_G (hex_table)."""
    factor = B.mev_to_natural(HE4_MASS_KG)
    Vm, a1, a2, _ = hex_table(nk, nz, 0.4, seed, cg_noise=cg_noise)
    V = Vm * factor
    T = B.kinetic_matrix(nz, Z_MIN, Z_MAX)
    w = Workload("hex_batch", V, T, a1, a2, Z_MIN, Z_MAX)
    rng = np.random.default_rng(cond_seed)
    for i in range(n_conditions):
        E = float(rng.uniform(10.0, 100.0))
        th = math.radians(float(rng.uniform(0.0, 60.0)))
        ph = math.radians(float(rng.uniform(0.0, 60.0)))
        k = B.momentum_from_angles(th, ph, E * factor)
        e = B.perpendicular_kinetic_difference(k, a1, a2, nk, nk)
        if channel_counts is not None:
            n = int(channel_counts[i % len(channel_counts)])
        else:
            n = batch_channel_count(e, E * factor, lo, hi)
        w.conditions.append(_condition(V, a1, a2, Z_MIN, Z_MAX, E, th, ph, n, factor))
    return w


def hex_large(n_conditions: int = 32, nz: int = 300, n_channels: int = 4000, seed: int = 1234,
              cond_seed: int = 7) -> Workload:
    factor = B.mev_to_natural(HE4_MASS_KG)
    Vm, a1, a2, _ = hex_table(65, nz, 0.25, seed)
    V = Vm * factor
    T = B.kinetic_matrix(nz, Z_MIN, Z_MAX)
    w = Workload("hex_large", V, T, a1, a2, Z_MIN, Z_MAX)
    rng = np.random.default_rng(cond_seed)
    for _ in range(n_conditions):
        E = float(rng.uniform(20.0, 150.0))
        w.conditions.append(_condition(V, a1, a2, Z_MIN, Z_MAX, E, 0.0, 0.0, n_channels, factor))
    return w
