"""CPU: the AnisotropicSolver restatement (oracle/restate.py) and the prefactor
table the HIP path uses.

Pinned by: spacing (1,1,1) reproduces Solver (the notebook's cross-check,
docs/notebooks/12-solver-benchmark.ipynb, where both stop at the same
iteration counts); uniform blocks give tau = 1 for any spacing (no y/z flux);
the converged field matches the float64 direct solve with face weights
(1, Ky, Kz), which distinguishes Ky from Kz and the weights from their
inverses; the float32 prefactor table equals the oracle's factor voxel by
voxel."""
import numpy as np
import pytest

from oracle import direct
from oracle.restate import OracleAnisotropicSolver, OracleSolver


def _img(shape, seed, p=0.6):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < p).astype(np.uint8)


def test_unit_spacing_matches_solver():
    img = _img((20, 14, 12), 0)
    a = OracleAnisotropicSolver(img, (1.0, 1.0, 1.0)).solve()
    s = OracleSolver(img).solve()
    assert a.iter == s.iter
    np.testing.assert_allclose(a.D_eff, s.D_eff, rtol=1e-6)


@pytest.mark.parametrize("spacing", [(1.0, 2.0, 3.0), (1.0, 1.0, 1.5), (1, 1, 4)])
def test_uniform_block_tau_one(spacing):
    img = np.ones((12, 9, 7), dtype=np.uint8)
    a = OracleAnisotropicSolver(img, spacing).solve(conv_crit=1e-5, iter_limit=100000)
    assert abs(float(a.tau[0]) - 1.0) < 1e-4


def test_flux_mask_of_weights_above_one():
    """Solver.vertical_flux zeroes the faces of voxels whose factor exceeds 8
    (taufactor/taufactor.py:417-418), a test meant for the neighbour count.
    With Ky = 4 (dx = 2 dy) interior factors reach 2 + 8 + ... > 8, so even a
    uniform block reports tau far from 1; the restatement keeps this."""
    img = np.ones((12, 9, 7), dtype=np.uint8)
    a = OracleAnisotropicSolver(img, (2.0, 1.0, 1.5))
    assert int((a.factor > 8).sum()) > 0
    a.run(400)
    tau, _ = a.compute_metrics()
    assert float(tau[0]) > 2.0


@pytest.mark.parametrize("spacing", [(1.0, 1.5, 2.0), (1.0, 2.0, 1.2)])
def test_converges_to_weighted_direct_solution(spacing):
    img = _img((18, 12, 10), 3, p=0.65)
    a = OracleAnisotropicSolver(img, spacing).solve(conv_crit=1e-5, iter_limit=200000)
    d = direct.anisotropic_D_rel(img, spacing)
    assert abs(float(a.D_rel[0]) - d) <= 2e-4 * d
    # the weights matter: swapping y and z spacing changes the answer
    dswap = direct.anisotropic_D_rel(img, (spacing[0], spacing[2], spacing[1]))
    assert abs(dswap - d) > 1e-3 * d


def _codes(mask):
    """The set-up kernel's anisotropic code of every interior voxel:
    1 + 48*xm + 16*xp + 8*ym + 4*yp + 2*zm + zp (x ghost state 2), 0 if inactive."""
    X, Y, Z = mask.shape
    p = np.zeros((X + 2, Y + 2, Z + 2), dtype=np.int64)
    p[1:-1, 1:-1, 1:-1] = mask
    p[0, 1:-1, 1:-1] = 2
    p[-1, 1:-1, 1:-1] = 2
    c = (1 + 48 * p[:-2, 1:-1, 1:-1] + 16 * p[2:, 1:-1, 1:-1] + 8 * p[1:-1, :-2, 1:-1]
         + 4 * p[1:-1, 2:, 1:-1] + 2 * p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:])
    return np.where(mask == 1, c, 0)


@pytest.mark.parametrize("spacing", [(1.0, 1.3, 0.7), (0.3, 1.0, 1.0), (1.0, 1.0, 1.0)])
def test_prefactor_table_equals_oracle_factor(spacing):
    from taufactor_amd.solver import AnisotropicSolver

    img = _img((9, 8, 7), 5)
    o = OracleAnisotropicSolver(img, spacing)
    s = AnisotropicSolver.__new__(AnisotropicSolver)
    dx, dy, dz = spacing
    s.Ky, s.Kz = (dx / dy) ** 2, (dx / dz) ** 2
    table = s._nn_table()
    got = table[_codes(img)]
    want = o.factor.numpy()[0]
    assert np.array_equal(got, want)
    assert table[145] == np.float32(s.Ky) and table[146] == np.float32(s.Kz)
