"""GPU: PeriodicSolver / PeriodicMultiPhaseSolver against the oracle
(interiors value for value; the ghosts of the final field hold the copy made
before the last iteration in both, but are compared only through the
interior they feed), the reference's periodic tests
(tests/test_taufactor.py:80-106, 226-247) and the stored notebook table."""
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
