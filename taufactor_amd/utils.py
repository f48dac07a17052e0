"""Synthetic benchmark structures and the transmission-line impedance
(taufactor/utils.py:75-147, 179-221).

The generators the reference's benchmark harness registers (fcc packing,
stacked blocks, 2-D and 3-D diagonal channels, zigzag channels), with the
reference's names, arguments and integer output, so that harness rows are
computed on the same voxels; and the impedance of a resistor/capacitor
ladder that ElectrodeSolver turns into its tortuosity.
"""
from __future__ import annotations

import numpy as np


def add_voxel_sphere(array, center_x, center_y, center_z, radius):
    """Set to 1 every voxel whose centre lies within `radius` of the centre
    (taufactor/utils.py:101-111; voxel centres at index + 0.5)."""
    nx, ny, nz = array.shape
    x, y, z = np.ogrid[:nx, :ny, :nz]
    d2 = (x - center_x + 0.5) ** 2 + (y - center_y + 0.5) ** 2 + (z - center_z + 0.5) ** 2
    array[d2 <= radius ** 2] = 1


def create_fcc_cube(pixels, overlap=0.0):
    """Voxelised fcc unit cell: spheres on the 6 face centres and 8 corners,
    radius sqrt(2)/4 * pixels / (1 - overlap/2) (taufactor/utils.py:114-147)."""
    cube = np.zeros((pixels, pixels, pixels), dtype=int)
    c = 0.5 * pixels
    r = 0.25 * np.sqrt(2) * pixels / (1 - 0.5 * overlap)
    for axis in range(3):
        for sign in (-1, 1):
            pos = [c, c, c]
            pos[axis] = c + sign * c
            add_voxel_sphere(cube, *pos, r)
    n = pixels
    for corner in [(0, 0, 0), (0, 0, n), (0, n, 0), (0, n, n), (n, 0, 0), (n, 0, n), (n, n, 0), (n, n, n)]:
        add_voxel_sphere(cube, *corner, r)
    return cube


def _feature_size(Nx, features):
    if Nx % (2 * features) != 0:
        raise ValueError(f"Nx must be a multiple of 2*features; got Nx={Nx} and features={features}")
    return Nx // (2 * features)


def create_stacked_blocks(Nx, features=1):
    """Staggered blocks: every second x-row of blocks shifted by half a block
    in y and z (taufactor/utils.py:179-193)."""
    size = _feature_size(Nx, features)
    x, y, z = np.ogrid[:Nx, :Nx, :Nx]
    shift = ((x // size) % 2) * (size // 2)
    return ((((y + shift) // size) + ((z - shift) // size)) % 2).astype(int)


def create_2d_diagonals(Nx, features=1):
    """Diagonal channels in (x, y), extruded along z (taufactor/utils.py:196-203)."""
    size = _feature_size(Nx, features)
    x, y, z = np.ogrid[:Nx, :Nx, :Nx]
    return ((((x + y) // size) % 2) + z - z).astype(int)


def create_2d_zigzag(Nx, features=1):
    """Diagonal channels mirrored about the middle x plane (taufactor/utils.py:206-211)."""
    p = create_2d_diagonals(Nx, features)
    half = Nx // 2
    p[half:] = p[:half][::-1]
    return p


def create_3d_diagonals(Nx, features=1):
    """Diagonal layers in (x, y, z) (taufactor/utils.py:214-221)."""
    size = _feature_size(Nx, features)
    x, y, z = np.ogrid[:Nx, :Nx, :Nx]
    return (((x + y + z) // size) % 2).astype(int)


def compute_impedance(R, C, freq):
    """Input impedance of a transmission line of resistors R and capacitors C
    (excitation 1 at the left end, closed right end), per frequency
    (taufactor/utils.py:75-84): Z <- R_i + 1/(j w C_i + 1/Z) from the right."""
    Z = np.full_like(freq, 1e50, dtype=complex)
    for R_i, C_i in zip(R[::-1], C[::-1]):
        Z = R_i + 1.0 / (1j * freq * C_i + 1.0 / Z)
    return Z


def compute_impedance_batched(R, C, freq):
    """compute_impedance for R, C of shape (bs, Nx) and freq (F,) or (bs, F);
    returns (bs, F) (taufactor/utils.py:87-98)."""
    if R.shape != C.shape:
        raise ValueError(f"R and C must have same shape; got {R.shape} vs {C.shape}")
    bs, Nx = R.shape
    if freq.ndim == 1:
        w = freq[None, :].repeat(bs, axis=0)
    elif freq.ndim == 2 and freq.shape[0] == bs:
        w = freq
    else:
        raise ValueError("freq must be (F,) or (bs, F)")
    Z = np.full_like(w, 1e50, dtype=complex)
    for k in range(Nx - 1, -1, -1):
        Z = R[:, k][:, None] + 1.0 / (1j * w * C[:, k][:, None] + 1.0 / Z)
    return Z
