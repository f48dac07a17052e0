"""Seeded synthetic volumes of the benchmark configurations.

Each generator is deterministic (numpy default_rng + scipy.ndimage on the
CPU), so a volume can be identified by the SHA-256 of its bytes.

  plain_100      100^3 binary, voxels i.i.d. Bernoulli(0.5), seed 0
  smoothed_512   512^3 binary: standard-normal noise, Gaussian-smoothed
                 (sigma 4, periodic), conductive below the 0.35 quantile
  uneven_batch   256 binary cubes, edge U[32,160], porosity U[0.3,0.7],
                 sigma U[1,5], one generator for all
  three_phase_320  320^3 labels {0,1,2}: label 0 below the 0.4 quantile of
                 one smoothed field, the rest split at the median of a second
  sigmoid_k_128  64 volumes of 128^3 with k = sigmoid(4 z), z standardised
                 smoothed noise (sigma 3)
"""
from __future__ import annotations

import hashlib

import numpy as np
import scipy.ndimage


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def field_sha256(a: np.ndarray) -> str:
    """SHA-256 of a float32 field's bytes with -0 folded into +0 (x + 0 is +0
    for x = -0 and x itself otherwise), so that two fields hash alike exactly
    when np.array_equal holds (NaN-free)."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    return hashlib.sha256((f + np.float32(0)).tobytes()).hexdigest()


def _smoothed(rng, shape, sigma):
    z = rng.standard_normal(shape, dtype=np.float32)
    return scipy.ndimage.gaussian_filter(z, sigma, mode="wrap")


def plain_100(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.random((100, 100, 100)) < 0.5).astype(np.uint8)


def smoothed_binary(n: int, porosity: float = 0.35, sigma: float = 4.0, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    z = _smoothed(rng, (n, n, n), sigma)
    return (z < np.quantile(z, porosity)).astype(np.uint8)


def smoothed_512(seed: int = 0) -> np.ndarray:
    return smoothed_binary(512, 0.35, 4.0, seed)


def uneven_batch(count: int = 256, seed: int = 0, edge=(32, 160)):
    """List of (img uint8 [e,e,e], porosity, sigma)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        e = int(rng.integers(edge[0], edge[1] + 1))
        phi = float(rng.uniform(0.3, 0.7))
        sigma = float(rng.uniform(1, 5))
        z = _smoothed(rng, (e, e, e), sigma)
        out.append(((z < np.quantile(z, phi)).astype(np.uint8), phi, sigma))
    return out


def three_phase(n: int = 320, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    z1 = _smoothed(rng, (n, n, n), 3.0)
    z2 = _smoothed(rng, (n, n, n), 3.0)
    lab = np.zeros((n, n, n), dtype=np.uint8)
    rest = z1 >= np.quantile(z1, 0.4)
    med = np.median(z2[rest])
    lab[rest & (z2 < med)] = 1
    lab[rest & (z2 >= med)] = 2
    return lab


def sigmoid_k(count: int = 64, n: int = 128, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out = np.empty((count, n, n, n), dtype=np.float32)
    for i in range(count):
        z = _smoothed(rng, (n, n, n), 3.0)
        z = (z - z.mean()) / z.std()
        out[i] = 1.0 / (1.0 + np.exp(-4.0 * z))
    return out
