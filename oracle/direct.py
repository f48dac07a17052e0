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


def _periodic_labels(lab: np.ndarray, n: int) -> np.ndarray:
    """Merge the 6-connected labels that meet across the y and z wraps."""
    parent = np.arange(n + 1)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for a_face, b_face in ((lab[:, 0, :], lab[:, -1, :]), (lab[:, :, 0], lab[:, :, -1])):
        for a, b in zip(a_face.ravel(), b_face.ravel()):
            if a and b:
                ra, rb = find(a), find(b)
                if ra != rb:
                    parent[max(ra, rb)] = min(ra, rb)
    roots = np.array([find(i) for i in range(n + 1)])
    return roots[lab]


def solve_fixed_point(k: np.ndarray, rtol: float = 1e-13, direct_max: int = 40000,
                      face_scale=(1.0, 1.0, 1.0), periodic_yz: bool = False):
    """k: float64 [X,Y,Z] voxel diffusivity (0 = non-conductive).
    face_scale: weights of the x, y and z faces (AnisotropicSolver's 1, Ky,
    Kz; the x boundary conductance 2k is not scaled).
    periodic_yz: y and z wrap round (PeriodicSolver / PeriodicMultiPhaseSolver,
    taufactor/taufactor.py:481-505, 623-656): the faces between the last and
    the first layer are faces like any other.
    Returns (c [X,Y,Z] float64, flux_1d [X-1] float64, D_rel float)."""
    k = np.asarray(k, dtype=np.float64)
    X, Y, Z = k.shape
    active = k > 0
    lab, n = scipy.ndimage.label(active, structure=scipy.ndimage.generate_binary_structure(3, 1))
    if periodic_yz:
        lab = _periodic_labels(lab, n)
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
    pairs = []
    for axis in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[axis] = slice(0, -1)
        sl_b[axis] = slice(1, None)
        pairs.append((axis, tuple(sl_a), tuple(sl_b)))
        if periodic_yz and axis > 0 and k.shape[axis] > 1:
            sl_a[axis], sl_b[axis] = slice(-1, None), slice(0, 1)  # the wrap face
            pairs.append((axis, tuple(sl_a), tuple(sl_b)))
    for axis, sl_a, sl_b in pairs:
        ka, kb = k[sl_a], k[sl_b]
        ia, ib = idx[sl_a], idx[sl_b]
        g = _hm(ka, kb) * face_scale[axis]
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


def anisotropic_D_rel(img: np.ndarray, spacing, **kw) -> float:
    """D_rel of a binary image with AnisotropicSolver's face weights."""
    dx, dy, dz = spacing
    return solve_fixed_point((np.asarray(img) == 1).astype(np.float64),
                             face_scale=(1.0, (dx / dy) ** 2, (dx / dz) ** 2), **kw)[2]


def multiphase_D_rel(img: np.ndarray, diffusivities: dict, **kw) -> float:
    """D_rel of a labelled image with per-label diffusivities (float32-rounded,
    as the reference stores them)."""
    img = np.asarray(img)
    k = np.zeros(img.shape, dtype=np.float64)
    for lbl, d in diffusivities.items():
        k[img == lbl] = float(np.float32(d))
    return solve_fixed_point(k, **kw)[2]


def dissipation(k: np.ndarray, c: np.ndarray) -> float:
    """W = sum_faces g (dc)^2 + sum_x-faces 2 k (c - bc)^2 of a field c."""
    k = np.asarray(k, dtype=np.float64)
    W = 0.0
    for axis in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[axis] = slice(0, -1)
        sl_b[axis] = slice(1, None)
        g = _hm(k[tuple(sl_a)], k[tuple(sl_b)])
        W += float(np.sum(g * (c[tuple(sl_b)] - c[tuple(sl_a)]) ** 2))
    W += float(np.sum(2 * k[0] * (c[0] + 0.5) ** 2)) + float(np.sum(2 * k[-1] * (c[-1] - 0.5) ** 2))
    return W


def torch_D_rel(k):
    """D_rel of a small voxel diffusivity field as a differentiable torch
    float64 function of k (dense direct solve; all k > 0 assumed, so every
    voxel is an unknown).  Its autograd gradient is the independent
    reference of the adjoint kernel: it differentiates through the solve."""
    import torch

    k = k.to(torch.float64)
    X, Y, Z = k.shape
    n = X * Y * Z
    idx = torch.arange(n).reshape(X, Y, Z)
    A = torch.zeros((n, n), dtype=torch.float64)
    b = torch.zeros(n, dtype=torch.float64)
    for axis in range(3):
        ia = idx.narrow(axis, 0, k.shape[axis] - 1).reshape(-1)
        ib = idx.narrow(axis, 1, k.shape[axis] - 1).reshape(-1)
        ka = k.narrow(axis, 0, k.shape[axis] - 1).reshape(-1)
        kb = k.narrow(axis, 1, k.shape[axis] - 1).reshape(-1)
        g = 2 * ka * kb / (ka + kb)
        A = A.index_put((ia, ia), g, accumulate=True)
        A = A.index_put((ib, ib), g, accumulate=True)
        A = A.index_put((ia, ib), -g, accumulate=True)
        A = A.index_put((ib, ia), -g, accumulate=True)
    i0 = idx[0].reshape(-1)
    i1 = idx[-1].reshape(-1)
    g0 = 2 * k[0].reshape(-1)
    g1 = 2 * k[-1].reshape(-1)
    A = A.index_put((i0, i0), g0, accumulate=True)
    A = A.index_put((i1, i1), g1, accumulate=True)
    b = b.index_put((i0,), g0 * -0.5, accumulate=True)
    b = b.index_put((i1,), g1 * 0.5, accumulate=True)
    c = torch.linalg.solve(A, b)
    # current through the x = 0 Dirichlet faces
    Q = torch.sum(g0 * (c[i0] + 0.5))
    return Q * X / (Y * Z)


def energy_gradient(k: np.ndarray, c: np.ndarray) -> np.ndarray:
    """dD_rel/dk from the dissipation identity (the formula the HIP kernel
    evaluates), numpy float64, including the one-sided derivative at k = 0."""
    k = np.asarray(k, dtype=np.float64)
    X, Y, Z = k.shape
    kp = np.pad(k, 1)
    cp = np.pad(c.astype(np.float64), 1)
    grad = np.zeros_like(k)
    offs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                kv, cv = k[x, y, z], c[x, y, z]
                nb = []
                for dx, dy, dz in offs:
                    xx = x + dx
                    if xx < 0 or xx >= X:
                        continue
                    ku = kp[xx + 1, y + dy + 1, z + dz + 1]
                    if ku > 0:
                        nb.append((ku, cp[xx + 1, y + dy + 1, z + dz + 1]))
                bnd = ([-0.5] if x == 0 else []) + ([0.5] if x == X - 1 else [])
                if kv > 0:
                    g = sum((cu - cv) ** 2 * 2 * ku * ku / (kv + ku) ** 2 for ku, cu in nb)
                    g += sum(2 * (cv - bc) ** 2 for bc in bnd)
                else:
                    vals = [cu for _, cu in nb] + bnd
                    if not vals:
                        continue
                    cs = sum(vals) / len(vals)
                    g = sum(2 * (v - cs) ** 2 for v in vals)
                grad[x, y, z] = g
    return grad * X / (Y * Z)
