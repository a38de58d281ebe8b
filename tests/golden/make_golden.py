"""Regenerate the golden fixtures in tests/golden/ from the reference Fortran.

    make -C oracle ref && python tests/golden/make_golden.py

The reference's five Fortran files are compiled from /root/reference (oracle/Makefile)
and driven through oracle/ref.py with the reference's own ABI; this script stores inputs
and outputs only (small .npz files). The reference's own fixture data
(tests/data/expected_potential.npy and expected_intensities.txt of the reference
repository) are stored alongside as data: the potential compressed, the intensities as text.
"""

from __future__ import annotations

import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from multiscat_amd import basis as B  # noqa: E402
from multiscat_amd import workloads as W  # noqa: E402
from oracle import oracle, ref  # noqa: E402

REF_DATA = Path("/root/reference/tests/data")


def wolken():
    """Reference fixture: Wolken He-LiF table + expected intensities, and Fortran outputs."""
    V_mev = np.load(REF_DATA / "expected_potential.npy", allow_pickle=False).reshape(10, 10, 550)
    np.savez_compressed(HERE / "wolken_potential.npz", potential_mev=V_mev)
    shutil.copyfile(REF_DATA / "expected_intensities.txt", HERE / "wolken_expected_intensities.txt")
    factor = B.mev_to_natural(B.ALPHA_PARTICLE_MASS_SI)
    V = V_mev * factor
    k = B.momentum_from_angles(np.deg2rad(30.0), 0.0, 20.0 * factor)
    inp = B.solver_inputs(V, k, (2.84, 0.0), (0.0, 2.84), 0.0, 8.0)
    out = {}
    for flag, nsig in ((1, 5), (0, 5), (1, 10)):
        psi = ref.run_multiscat_fortran(flag, nsig, V, inp.perpendicular_kinetic_difference,
                                        inp.wave_a, inp.wave_b, inp.wave_c,
                                        inp.parallel_kinetic_energy)
        _, info = oracle.solve(flag, nsig, V, inp.perpendicular_kinetic_difference, inp.wave_a,
                               inp.wave_b, inp.wave_c, inp.parallel_kinetic_energy,
                               return_info=True)
        out[f"psi_last_f{flag}_n{nsig}"] = psi[:, :, -1]
        out[f"P_f{flag}_n{nsig}"] = oracle.intensities(psi, inp.wave_a, inp.wave_b)
        out[f"iters_f{flag}_n{nsig}"] = np.int32(info["iterations"])
        print("wolken", flag, nsig, info["iterations"], flush=True)
    np.savez_compressed(HERE / "wolken_fortran.npz", **out)


def square_small():
    """Square-lattice Morse problem on a 7 x 6 full grid (nkx != nky), Nz = 48."""
    nz = 48
    factor = B.mev_to_natural(W.HE4_MASS_KG)
    a1, a2 = (2.5, 0.0), (0.0, 2.5)
    z, _ = B.lobatto_rule(nz, W.Z_MIN, W.Z_MAX)
    V = np.zeros((7, 6, nz), dtype=np.complex128)
    V[0, 0] = W.morse(z, 8.0, 1.2, 0.0)
    corr = 0.025 * 8.0 * np.exp(-2.4 * z)
    for g in ((1, 0), (6, 0), (0, 1), (0, 5)):
        V[g] = corr
    V[2, 1] = 0.3 * (1 + 0.5j) * np.exp(-2.4 * z)  # a complex, non-symmetric component
    V[5, 5] = 0.3 * (1 - 0.5j) * np.exp(-2.4 * z)
    V *= factor
    k = B.momentum_from_angles(np.deg2rad(20.0), np.deg2rad(10.0), 20.0 * factor)
    inp = B.solver_inputs(V, k, a1, a2, W.Z_MIN, W.Z_MAX)
    out = dict(potential=V, e=inp.perpendicular_kinetic_difference, wave_a=inp.wave_a,
               wave_b=inp.wave_b, wave_c=inp.wave_c, kinetic=inp.parallel_kinetic_energy)
    for flag, nsig in ((1, 10), (0, 10), (1, 5), (0, 4)):
        psi = ref.run_multiscat_fortran(flag, nsig, V, inp.perpendicular_kinetic_difference,
                                        inp.wave_a, inp.wave_b, inp.wave_c,
                                        inp.parallel_kinetic_energy)
        _, info = oracle.solve(flag, nsig, V, inp.perpendicular_kinetic_difference, inp.wave_a,
                               inp.wave_b, inp.wave_c, inp.parallel_kinetic_energy,
                               return_info=True)
        out[f"psi_f{flag}_n{nsig}"] = psi
        out[f"iters_f{flag}_n{nsig}"] = np.int32(info["iterations"])
        print("square", flag, nsig, info["iterations"], flush=True)
    rng = np.random.default_rng(2024)
    nc = 7 * 6
    state = rng.standard_normal((nz, nc)) + 1j * rng.standard_normal((nz, nc))
    out["state_in"] = state
    out["upper_out"] = ref.debug_apply_upper_block_fortran(V, state)
    out["lower_out"] = ref.debug_solve_lower_block_fortran(
        V, inp.wave_c, inp.perpendicular_kinetic_difference, inp.parallel_kinetic_energy, state)
    lam, fac, vec = ref.debug_build_preconditioner_fortran(
        V, inp.perpendicular_kinetic_difference, inp.parallel_kinetic_energy)
    out["eigenvalues"] = lam
    out["factors"] = fac
    out["abs_eigenvectors"] = np.abs(vec)
    np.savez_compressed(HERE / "square_small.npz", **out)


def helpers():
    """Outputs of the reference's four input-builder helpers."""
    out = {}
    for i, (zmin, zmax, n) in enumerate(((-0.4, 1.7, 57), (0.0, 8.0, 550), (-1.5, 8.0, 101), (0.0, 2.0, 3))):
        w, x = ref.get_lobatto_weights(zmin, zmax, n)
        out[f"lobatto{i}"] = np.array([zmin, zmax, n], dtype=np.float64)
        out[f"lobatto{i}_w"] = w
        out[f"lobatto{i}_x"] = x
    for i, (zmin, zmax, n) in enumerate(((0.0, 8.0, 49), (-1.5, 8.0, 100), (0.0, 2.0, 2))):
        out[f"kinetic{i}"] = np.array([zmin, zmax, n], dtype=np.float64)
        out[f"kinetic{i}_T"] = ref.get_parallel_kinetic_energy(zmin, zmax, n)
    lattices = (
        ((2.84, 0.0), (0.0, 2.84), (1.2, 0.0, 5.9), 10, 10),
        ((3.16, 0.0), (1.58, 2.7367864), (2.1, -0.7, 9.3), 9, 8),
        ((2.84 * np.cos(0.35), 2.84 * np.sin(0.35)), (-2.84 * np.sin(0.35), 2.84 * np.cos(0.35)),
         (0.5, 0.3, 4.0), 6, 7),
    )
    for i, (a1, a2, k, nx, ny) in enumerate(lattices):
        e = ref.get_perpendicular_kinetic_difference(k[0], k[1], k[2], nx, ny, a1[0], a1[1], a2[0],
                                                     a2[1])
        out[f"lattice{i}"] = np.array([*a1, *a2, *k, nx, ny], dtype=np.float64)
        out[f"lattice{i}_e"] = e
        a, b, c = ref.get_abc_arrays(-1.0, 7.0, nx, ny, e, 40)
        out[f"lattice{i}_abc"] = np.stack([a, b, c])
    np.savez_compressed(HERE / "helpers.npz", **out)


if __name__ == "__main__":
    helpers()
    square_small()
    wolken()
    for p in sorted(HERE.glob("*.npz")):
        print(p.name, p.stat().st_size)
