"""Oracle: synthetic structures of the reference's benchmark tables.

TEST INFRASTRUCTURE (see oracle/restate.py for who may import it).

Restated from taufactor/utils.py so that the stored outputs of
docs/notebooks/12-solver-benchmark.ipynb can be reproduced:
  add_voxel_sphere / create_fcc_cube  taufactor/utils.py:101-147
  create_2d_diagonals / create_2d_zigzag  taufactor/utils.py:196-211
  the "fcc" registry entry            taufactor/benchmark.py:29
  multi_zigzag                        12-solver-benchmark.ipynb (cell 13)
"""
from __future__ import annotations

import numpy as np


def _add_sphere(cube: np.ndarray, cx: float, cy: float, cz: float, radius: float) -> None:
    nx, ny, nz = cube.shape
    x, y, z = np.ogrid[:nx, :ny, :nz]
    d2 = (x - cx + 0.5) ** 2 + (y - cy + 0.5) ** 2 + (z - cz + 0.5) ** 2
    cube[d2 <= radius ** 2] = 1


def fcc_cube(n: int, overlap: float = 0.0) -> np.ndarray:
    """Voxelised fcc unit cell of edge n: 6 face-centred half spheres and 8
    corner spheres, radius sqrt(2)/4 * n / (1 - overlap/2); 1 = sphere."""
    cube = np.zeros((n, n, n), dtype=int)
    c = 0.5 * n
    r = 0.25 * np.sqrt(2) * n / (1 - 0.5 * overlap)
    for axis in range(3):
        for sign in (-1, 1):
            pos = [c, c, c]
            pos[axis] = c + sign * c
            _add_sphere(cube, *pos, r)
    for corner in [(0, 0, 0), (0, 0, n), (0, n, 0), (0, n, n), (n, 0, 0), (n, 0, n), (n, n, 0), (n, n, n)]:
        _add_sphere(cube, *corner, r)
    return cube


def fcc_pores(n: int) -> np.ndarray:
    """The benchmark's "fcc" structure: pore space of an fcc packing with 5 %
    sphere overlap (1 = pore)."""
    return (fcc_cube(n, overlap=0.05) == 0).astype(int)


def diagonals_2d(n: int, features: int = 1) -> np.ndarray:
    size = n // (2 * features)
    x, y, z = np.ogrid[:n, :n, :n]
    return ((((x + y) // size) % 2) + z - z).astype(int)


def zigzag_2d(n: int, features: int = 1) -> np.ndarray:
    p = diagonals_2d(n, features)
    half = n // 2
    p[half:] = p[:half][::-1]
    return p


def multi_zigzag(n: int) -> np.ndarray:
    """zigzag channels whose first half (in x) carries label 2 instead of 1."""
    g = zigzag_2d(n, 1)
    h = n // 2
    g[:h][g[:h] == 1] = 2
    return g


def staircase(n: int) -> np.ndarray:
    """Slanted 2x2 strip of ones (tests/test_taufactor.py:58-66)."""
    img = np.zeros((n, n + 1, n + 1))
    for i in range(n):
        img[i, i:i + 2, i:i + 2] = 1
    return img


def diagonals_2d_features(n: int, features: int) -> np.ndarray:
    """create_2d_diagonals(N, features) (taufactor/utils.py:196-203)."""
    return diagonals_2d(n, features)


def graded_2d(nx: int = 250, cube_size: int = 25) -> dict:
    """The six 2-D graded validation structures of
    docs/notebooks/01-tortuosity.ipynb (cell 4): diagonals, diag-pillars,
    large-small, small-large, smooth-rough, wedge; each [nx, nx, 1]."""
    x, y, z = np.ogrid[:nx, :nx, :1]
    fields = {}
    fields["diagonals"] = ((x + y + z) // cube_size) % 2
    pillars = x - x + ((y // cube_size) + (z // cube_size)) % 2
    p = np.zeros((nx, nx, 1))
    p[:125, :] = pillars[:125, :]
    p[125:, :] = fields["diagonals"][:125, :]
    fields["diag-pillars"] = np.roll(p[::-1, :, :], shift=10, axis=1)
    p = np.zeros((nx, nx, 1))
    p[:125, 0:40] = 1
    p[125:, 15:25] = 1
    p[:, 50:100] = p[:, :50]
    p[:, 100:200] = p[:, :100]
    p[:, 200:250] = p[:, :50]
    p = np.roll(p, shift=5, axis=1)
    fields["large-small"] = p
    fields["small-large"] = p[::-1, :, :].copy()
    p = pillars.copy()
    p[0::2, :, :] = np.roll(p[0::2, :, :], shift=1, axis=1)
    p[:125, :] = pillars[:125, :]
    fields["smooth-rough"] = np.roll(p, shift=10, axis=1)
    p = np.zeros((nx, nx, 1))
    mask = ((x + 0.5) / nx + 4 * (y + 0.5) / nx < 2) & ((x + 0.5) / nx - 4 * (y + 0.5) / nx < 0)
    p[mask] = 1
    p[:, 125:] = p[:, :125][:, ::-1]
    p[-1, :] = 0
    fields["wedge"] = p
    return fields


GRADED_ORDER = ["diagonals", "diag-pillars", "large-small", "small-large", "smooth-rough", "wedge"]


def graded_batch() -> np.ndarray:
    """np.stack of the six structures in the notebook's order (cell 7)."""
    f = graded_2d()
    return np.stack([f[k] for k in GRADED_ORDER], axis=0)
