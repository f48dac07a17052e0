"""GPU: AnisotropicSolver through the HIP path against the oracle
(oracle/restate.py OracleAnisotropicSolver): iterates value for value, the
stopping iteration and tau of full solves, and the flux mask of weights
above one.  Spacing (1,1,1) against Solver on the notebook's fcc structure."""
import numpy as np
import pytest

import taufactor_amd as tau
from oracle.restate import OracleAnisotropicSolver
from oracle.structures import fcc_pores

pytestmark = pytest.mark.gpu

SPACINGS = [(1.0, 1.0, 1.0), (1.0, 1.5, 2.0), (2.0, 1.0, 1.5), (0.7, 1.0, 1.0), (1, 3, 1)]


def _img(shape, seed, p=0.6):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < p).astype(np.uint8)


@pytest.mark.parametrize("spacing", SPACINGS)
@pytest.mark.parametrize("H", [1, 2, 3, 4])
def test_iterates_bit_identical(spacing, H):
    img = _img((21, 17, 133), sum(int(10 * v) for v in spacing) + H)
    s = tau.AnisotropicSolver(img, spacing)
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=23, verbose=False)
    o = OracleAnisotropicSolver(img, spacing).run(23)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


@pytest.mark.parametrize("spacing", [(1.0, 1.3, 2.0), (2.0, 1.0, 1.5)])
def test_full_solve_matches_oracle(spacing):
    img = _img((30, 22, 26), 7, p=0.7)
    s = tau.AnisotropicSolver(img, spacing)
    s.solve(verbose=False, iter_limit=3000)
    o = OracleAnisotropicSolver(img, spacing).solve(iter_limit=3000)
    assert s.iter == o.iter
    np.testing.assert_allclose(s.tau, o.tau, rtol=2e-7)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


def test_unit_spacing_matches_solver_on_fcc():
    img = fcc_pores(48)
    a = tau.AnisotropicSolver(img, (1.0, 1.0, 1.0))
    a.solve(verbose=False, conv_crit=1e-3)
    s = tau.Solver(img)
    s.solve(verbose=False, conv_crit=1e-3)
    assert a.iter == s.iter
    np.testing.assert_allclose(a.tau, s.tau, rtol=1e-5)


def test_spacing_validation():
    img = np.ones((4, 4, 4), dtype=np.uint8)
    with pytest.raises(ValueError):
        tau.AnisotropicSolver(img, (1, 1))
    with pytest.raises(ValueError):
        tau.AnisotropicSolver(img, (1, "a", 1))
    with pytest.warns(UserWarning):
        tau.AnisotropicSolver(img, (1, 1, 11))
