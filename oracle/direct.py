"""Oracle: float64 sparse direct solve of the discrete operator the SOR
iteration converges to.

TEST INFRASTRUCTURE (see oracle/restate.py for who may import it).

Unknowns are the concentrations of conductive voxels.  Face conductances are
g = 2ab/(a+b) of the two voxel diffusivities (taufactor/taufactor.py:577-596;
g = 1 between two conductive voxels of a binary image, taufactor.py:402-410).
A voxel on an x face also couples to the Dirichlet value bc (-0.5 at x=0,
+0.5 at x=Nx-1) with conductance 2*k (the reference's ghost holds 2*bc and
counts twice in the prefactor, taufactor.py:291, 404, 601-602).  y and z faces
are closed.  The fixed point satisfies, for every unknown v,

    sum_u g_uv (c_v - c_u) + sum_{x faces of v} 2 k_v (c_v - bc) = 0.

6-connected clusters that touch neither x face are singular (pure Neumann);
they carry no flux at the fixed point and are held at 0.  The result is
D_rel = mean over internal x faces of the layer-mean flux, times Nx (the
reference's D_rel, taufactor.py:296-303, at exact convergence).
"""
from __future__ import annotations

import numpy as np
import scipy.ndimage
import scipy.sparse
import scipy.sparse.linalg


def _hm(a, b):
    s = a + b
    out = np.zeros_like(s)
    np.divide(2.0 * a * b, s, out=out, where=s > 0)
    return out


def solve_fixed_point(k: np.ndarray, rtol: float = 1e-13, direct_max: int = 40000):
    """k: float64 [X,Y,Z] voxel diffusivity (0 = non-conductive).
    Returns (c [X,Y,Z] float64, flux_1d [X-1] float64, D_rel float)."""
    k = np.asarray(k, dtype=np.float64)
    X, Y, Z = k.shape
    active = k > 0
    lab, n = scipy.ndimage.label(active, structure=scipy.ndimage.generate_binary_structure(3, 1))
    touching = set(np.unique(lab[0])) | set(np.unique(lab[-1]))
    touching.discard(0)
    keep = np.isin(lab, list(touching)) if touching else np.zeros_like(active)
    idx = -np.ones(k.shape, dtype=np.int64)
    nun = int(keep.sum())
    idx[keep] = np.arange(nun)
    c = np.zeros(k.shape)
    if nun == 0:
        return c, np.zeros(max(X - 1, 0)), 0.0

    rows, cols, vals = [], [], []
    diag = np.zeros(nun)
    rhs = np.zeros(nun)
    for axis in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[axis] = slice(0, -1)
        sl_b[axis] = slice(1, None)
        ka, kb = k[tuple(sl_a)], k[tuple(sl_b)]
        ia, ib = idx[tuple(sl_a)], idx[tuple(sl_b)]
        g = _hm(ka, kb)
        m = (ia >= 0) & (ib >= 0) & (g > 0)
        ga, a_, b_ = g[m], ia[m], ib[m]
        rows += [a_, b_]
        cols += [b_, a_]
        vals += [-ga, -ga]
        np.add.at(diag, a_, ga)
        np.add.at(diag, b_, ga)
    for xface, bc in ((0, -0.5), (X - 1, 0.5)):
        i = idx[xface]
        m = i >= 0
        gb = 2.0 * k[xface][m]
        np.add.at(diag, i[m], gb)
        np.add.at(rhs, i[m], gb * bc)
    ar = np.arange(nun)
    A = scipy.sparse.csr_matrix(
        (np.concatenate(vals + [diag]), (np.concatenate(rows + [ar]), np.concatenate(cols + [ar]))),
        shape=(nun, nun))
    if nun <= direct_max:
        sol = scipy.sparse.linalg.spsolve(A.tocsc(), rhs)
    else:
        dinv = 1.0 / diag
        M = scipy.sparse.linalg.LinearOperator((nun, nun), matvec=lambda v: dinv * v)
        sol, info = scipy.sparse.linalg.cg(A, rhs, rtol=rtol, atol=0.0, maxiter=100000, M=M)
        if info != 0:
            raise RuntimeError(f"CG did not converge (info={info})")
    c[keep] = sol
    gx = _hm(k[:-1], k[1:])
    flux = gx * (c[1:] - c[:-1])
    flux_1d = flux.reshape(X - 1, -1).mean(axis=1) if X > 1 else np.zeros(0)
    D_rel = float(flux_1d.mean() * X) if X > 1 else 0.0
    return c, flux_1d, D_rel


def solver_D_rel(img: np.ndarray, **kw) -> float:
    """D_rel of a binary image (conductive = 1)."""
    return solve_fixed_point((np.asarray(img) == 1).astype(np.float64), **kw)[2]


def multiphase_D_rel(img: np.ndarray, diffusivities: dict, **kw) -> float:
    """D_rel of a labelled image with per-label diffusivities (float32-rounded,
    as the reference stores them)."""
    img = np.asarray(img)
    k = np.zeros(img.shape, dtype=np.float64)
    for lbl, d in diffusivities.items():
        k[img == lbl] = float(np.float32(d))
    return solve_fixed_point(k, **kw)[2]
