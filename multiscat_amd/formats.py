"""Potential pre-processing and the on-disk formats of the reference's examples and tests.

* interpolate_potential: a potential sampled on an even real-space grid (nx, ny, nz_even)
  over one unit cell and an even z grid becomes the solver's table V_G(z_k): the in-plane
  Fourier components (cropped or zero-padded to the requested G grid, FFT storage order)
  linearly interpolated onto the Gauss-Lobatto nodes of [z_min, z_max]
  (multiscat/interpolate.py:38-116, which uses numpy.interp along z).
* The potential table file (tests/data/expected_potential.npy of the reference): the
  (nkx, nky, nz) table of V_G(z_k) in meV, complex128, flattened in C order
  (tests/test_fixed_example.py:115-132, 599-607).
* The intensity text file (tests/data/expected_intensities.txt): a `# kx ky intensity`
  header and one line per channel with the two FFT modes and P
  (tests/test_fixed_example.py:86-112 reads it).

Host-side numpy: these run once per potential, outside the solve.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from multiscat_amd import basis as B


def _crop_modes(a: np.ndarray, nkx: int, nky: int) -> np.ndarray:
    """Keep the FFT modes of an (Nx, Ny, ...) spectrum that an (nkx, nky) grid holds (modes
    absent from the source are zero)."""
    Nx, Ny = a.shape[:2]
    out = np.zeros((nkx, nky) + a.shape[2:], dtype=np.complex128)
    src_x, src_y = set(B.fft_modes(Nx).tolist()), set(B.fft_modes(Ny).tolist())
    for i, mx in enumerate(B.fft_modes(nkx)):
        if mx not in src_x:
            continue
        for j, my in enumerate(B.fft_modes(nky)):
            if my in src_y:
                out[i, j] = a[mx % Nx, my % Ny]
    return out


def interpolate_potential(values_xyz, z_values, *, shape, z_range, weighted: bool = False):
    """V_G(z_k) table from real-space samples.

    values_xyz[ix, iy, iz] = V(ix/Nx a_1 + iy/Ny a_2, z_values[iz]) over one unit cell,
    z_values ascending and evenly spaced (any spacing works for the linear interpolation);
    shape = (nkx, nky, nz) of the result, z_range = (z_min, z_max) of its Lobatto nodes.
    V_G = (1/(Nx Ny)) sum_r V(r) exp(-i G.r), so V(r) = sum_G V_G exp(i G.r).

    The result is the solver's table: the values at the nodes. weighted=True multiplies
    node k by the Lobatto basis weight 1/sqrt(w_k), as multiscat/interpolate.py:84-92 does
    when it stores the interpolated values as the raw data of a Lobatto-basis operator; that
    is the reference's operator storage, which its own test reads back by multiplying with
    the same weights (tests/test_fixed_example.py:115-132), not the table the solve takes."""
    v = np.asarray(values_xyz)
    z_values = np.asarray(z_values, dtype=np.float64)
    if v.ndim != 3 or v.shape[2] != z_values.size:
        raise ValueError("values_xyz must be (nx, ny, len(z_values))")
    if np.any(np.diff(z_values) <= 0):
        raise ValueError("z_values must be strictly ascending")
    nkx, nky, nz = shape
    spec = np.fft.fft2(v, axes=(0, 1)) / (v.shape[0] * v.shape[1])
    spec = _crop_modes(spec, nkx, nky)
    nodes, _ = B.lobatto_rule(nz, *z_range)
    out = np.empty((nkx, nky, nz), dtype=np.complex128)
    for i in range(nkx):
        for j in range(nky):
            out[i, j] = (np.interp(nodes, z_values, spec[i, j].real)
                         + 1j * np.interp(nodes, z_values, spec[i, j].imag))
    if weighted:
        _, w = B.lobatto_rule(nz, *z_range)
        out *= (1.0 / np.sqrt(w))[None, None, :]
    return out


def read_potential_npy(path, shape) -> np.ndarray:
    """The table (nkx, nky, nz) of V_G(z_k) in meV from the flattened .npy file."""
    data = np.load(Path(path), allow_pickle=False)
    return np.asarray(data, dtype=np.complex128).reshape(shape)


def write_potential_npy(path, values_mev) -> None:
    np.save(Path(path), np.ascontiguousarray(values_mev, dtype=np.complex128).ravel(),
            allow_pickle=False)


def write_intensities_txt(path, P) -> None:
    """`# kx ky intensity` and one line per grid point, modes in storage order."""
    P = np.asarray(P).real
    nx, ny = P.shape
    lines = ["# kx ky intensity"]
    for i, mx in enumerate(B.fft_modes(nx)):
        for j, my in enumerate(B.fft_modes(ny)):
            lines.append(f"{int(mx):4d} {int(my):4d} {P[i, j]:15.8e}")
    Path(path).write_text("\n".join(lines) + "\n")


def read_intensities_txt(path, shape) -> np.ndarray:
    """(nx, ny) intensities from the text format; modes map to storage indices mod n."""
    nx, ny = shape
    out = np.zeros(shape)
    for line in Path(path).read_text().splitlines():
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        kx, ky, val = s.split()[:3]
        out[int(kx) % nx, int(ky) % ny] = float(val)
    return out
