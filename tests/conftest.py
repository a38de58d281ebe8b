"""Shared test set-up: the `gpu` marker, repository paths and golden-fixture loaders."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden(name: str):
    with np.load(GOLDEN / name, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def wolken_expected_intensities(shape=(10, 10)):
    """The reference's golden P table (tests/data/expected_intensities.txt of the reference),
    FFT modes mapped to storage indices as test_fixed_example.py:82-112 does."""
    nx, ny = shape
    out = np.zeros(shape)
    for line in (GOLDEN / "wolken_expected_intensities.txt").read_text().splitlines():
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        hx, ky, v = s.split()
        out[int(hx) % nx, int(ky) % ny] = float(v)
    return out


def wolken_inputs():
    """Natural-unit solver inputs of the reference's Wolken He-LiF test
    (tests/test_fixed_example.py:68-79,135-170 of the reference)."""
    from multiscat_amd import basis as B

    V_mev = golden("wolken_potential.npz")["potential_mev"]
    factor = B.mev_to_natural(B.ALPHA_PARTICLE_MASS_SI)
    k = B.momentum_from_angles(np.deg2rad(30.0), 0.0, 20.0 * factor)
    return B.solver_inputs(V_mev * factor, k, (2.84, 0.0), (0.0, 2.84), 0.0, 8.0)


@pytest.fixture(scope="session")
def square_small():
    return golden("square_small.npz")
