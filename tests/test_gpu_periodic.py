"""GPU: PeriodicSolver / PeriodicMultiPhaseSolver against the oracle, the
reference's periodic tests (tests/test_taufactor.py:80-106, 226-247) and the
stored notebook table.

With Y and Z even the sweep reads across the boundary and fuses half-sweeps;
the whole field, ghosts included, then equals the oracle's (the ghosts hold
the copy made before the last iteration).  With Y or Z odd the solve
alternates a ghost wrap and one half-sweep, and the interiors are compared
(the ghosts of the final field are a copy taken at another point)."""
import numpy as np
import pytest

import taufactor_amd as tau
from oracle.restate import OraclePeriodicMultiPhaseSolver, OraclePeriodicSolver
from oracle.structures import diagonals_2d_features, staircase

pytestmark = pytest.mark.gpu


def _interior(f):
    return f[:, 1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("shape", [(9, 10, 12), (7, 9, 11), (12, 1, 5), (5, 2, 3)])
def test_periodic_solver_iterates(shape):
    rng = np.random.default_rng(sum(shape))
    img = (rng.random(shape) < 0.6).astype(np.uint8)
    s = tau.PeriodicSolver(img)
    s.solve(iter_limit=37, verbose=False)
    o = OraclePeriodicSolver(img).run(37)
    assert np.array_equal(_interior(s.field.cpu().numpy()), _interior(o.field.numpy()))


@pytest.mark.parametrize("shape", [(9, 10, 12), (7, 9, 11)])
def test_periodic_multiphase_iterates(shape):
    rng = np.random.default_rng(sum(shape) + 1)
    img = rng.integers(0, 3, size=shape)
    Ds = {0: 0.0, 1: 1.0, 2: 0.2}
    s = tau.PeriodicMultiPhaseSolver(img, dict(Ds))
    s.solve(iter_limit=37, verbose=False)
    o = OraclePeriodicMultiPhaseSolver(img, dict(Ds)).run(37)
    assert np.array_equal(_interior(s.field.cpu().numpy()), _interior(o.field.numpy()))


def test_periodic_reference_known_answers():
    N = 20
    img = np.ones((N, N, N))
    img[:, 0] = 0
    S = tau.PeriodicSolver(img)
    S.solve()
    assert np.around(S.tau, decimals=5) == 1.0
    img = np.zeros((N, N, N))
    img[:2] = 1
    S = tau.PeriodicSolver(img)
    S.solve(verbose="per_iter", iter_limit=1000)
    assert S.tau == np.inf
    img = np.zeros((N, N, N))
    img[:, 0:10, 0:10] = 1
    S = tau.PeriodicSolver(img)
    S.solve()
    assert np.around(S.tau, decimals=5) == 1
    s = tau.PeriodicMultiPhaseSolver(np.ones((N, N, N)), {1: 1.0})
    s.solve(iter_limit=1000)
    assert np.around(s.tau, 4) == 1.0
    img = staircase(20)
    s_bin = tau.PeriodicSolver(img)
    s_bin.solve(iter_limit=1000)
    s_mp = tau.PeriodicMultiPhaseSolver(img, {0: 0.0, 1: 1.0})
    s_mp.solve(iter_limit=1000)
    assert np.isclose(float(s_bin.tau[0]), float(s_mp.tau[0]), atol=1e-3)


@pytest.mark.parametrize("which", ["PeriodicSolver", "PeriodicMultiPhaseSolver"])
def test_periodic_notebook_table(goldens, which):
    row = next(r for r in goldens["periodic_diagonal2d64_conv1e-3"]["rows"] if r["solver"] == which)
    S = getattr(tau, which)(diagonals_2d_features(64, 2))
    S.solve(conv_crit=1e-3, verbose=False)
    assert S.iter == row["iterations"]
    assert abs(float(S.tau[0]) - row["tau"]) <= 1.5e-6 * row["tau"]


def _binary(shape, seed, p=0.6):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < p).astype(np.uint8)


@pytest.mark.parametrize("shape", [(9, 10, 12), (17, 2, 4), (6, 4, 130), (13, 30, 2), (5, 6, 8)])
@pytest.mark.parametrize("H", [1, 2, 3, 5, 6])
def test_periodic_fused_bit_identical(shape, H):
    """Y smaller than the halo (the halo wraps several times), Z = 2, and
    z covered by more than one tile; two solve calls, so the second starts on
    the other colour."""
    img = _binary(shape, sum(shape) * 7 + H)
    s = tau.PeriodicSolver(img)
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=5, verbose=False)
    s.solve(iter_limit=5 + 2 * H + 3, verbose=False)
    o = OraclePeriodicSolver(img).run(5 + 2 * H + 3)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


@pytest.mark.parametrize("H,Z", [(4, 176), (4, 240), (5, 168), (6, 200)])
@pytest.mark.parametrize("tile", [(16, 2, 4), (8, 2, 0), (12, 2, 6)])
def test_periodic_fused_z_tiles(H, Z, tile):
    """Wide tiles, the narrow last strip (Y = 34 gives an odd count of
    16-row y tiles) and wraps at both ends of z."""
    img = _binary((9, 34, Z), Z + H)
    s = tau.PeriodicSolver(img)
    s.half_sweeps_per_pass = H
    s.tile = tile
    s.solve(iter_limit=2 * H + 1, verbose=False)
    o = OraclePeriodicSolver(img).run(2 * H + 1)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


@pytest.mark.parametrize("H", [2, 4, 6])
def test_periodic_multiphase_fused_bit_identical(H):
    rng = np.random.default_rng(40 + H)
    img = rng.integers(0, 3, size=(9, 10, 12))
    Ds = {0: 0.0, 1: 1.0, 2: 0.2}
    s = tau.PeriodicMultiPhaseSolver(img, dict(Ds))
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=37, verbose=False)
    o = OraclePeriodicMultiPhaseSolver(img, dict(Ds)).run(37)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


def test_periodic_fused_full_solve_and_autotune():
    """A full solve through the autotuned fused sweep stops where the oracle
    does with the same tau."""
    img = _binary((40, 32, 36), 3)
    s = tau.PeriodicSolver(img)
    s.solve(verbose=False)
    o = OraclePeriodicSolver(img).solve()
    assert s.iter == o.iter
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())
    np.testing.assert_allclose(s.tau, o.tau, rtol=2e-7)

