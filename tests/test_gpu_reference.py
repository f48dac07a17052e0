"""GPU: the reference's own tests (tests/test_taufactor.py:12-224, Solver and
MultiPhaseSolver) and the stored notebook tables
(docs/notebooks/12-solver-benchmark.ipynb) through the HIP path.  Same
inputs and the same assertions as the reference; device='cuda'."""
import numpy as np
import pytest

import taufactor_amd as tau
from oracle.structures import fcc_pores, multi_zigzag, staircase

pytestmark = pytest.mark.gpu


def test_solver_on_uniform_block():
    N = 20
    img = np.ones((N, N, N))
    img[:, 0] = 0
    S = tau.Solver(img)
    S.solve()
    assert np.around(S.tau, decimals=5) == 1.0


def test_solver_on_uniform_rectangular_block_solver_dim():
    N = 20
    img = np.ones((N * 2, N, N))
    img[:, 0] = 0
    S = tau.Solver(img)
    S.solve()
    assert np.around(S.tau, decimals=5) == 1.0


def test_solver_on_uniform_rectangular_block_non_solver_dim():
    N = 20
    img = np.ones((N, N, N * 2))
    img[:, 0] = 0
    S = tau.Solver(img)
    S.solve()
    assert np.around(S.tau, decimals=5) == 1.0


def test_solver_non_percolating(capsys):
    N = 20
    img = np.zeros((N, N, N))
    img[:2] = 1
    S = tau.Solver(img)
    S.solve(verbose="per_iter", iter_limit=1000)
    assert S.tau == np.inf
    assert "has no percolating path!" in capsys.readouterr().out


def test_solver_on_strip_of_ones():
    N = 20
    img = np.zeros((N, N, N))
    img[:, 0:10, 0:10] = 1
    S = tau.Solver(img)
    S.solve()
    assert np.around(S.tau, decimals=5) == 1


def test_solver_on_slanted_strip_of_ones():
    S = tau.Solver(staircase(20))
    S.solve()
    assert np.around(S.tau, decimals=5) == 7.51667


def test_deadend():
    solid = np.zeros((10, 50, 50))
    solid[:8, 25, 25] = 1
    S = tau.Solver(solid)
    S.solve()
    assert np.around(S.D_eff, decimals=5) == 0
    assert S.tau == np.inf


def test_mphsolver_non_percolating():
    N = 20
    img = np.zeros((N, N, N))
    img[:2] = 1
    S = tau.MultiPhaseSolver(img)
    S.solve(iter_limit=1000)
    assert S.tau == np.inf


def test_mphsolver_on_ones_block():
    N = 20
    S = tau.MultiPhaseSolver(np.ones((N, N, N)))
    S.solve(iter_limit=1000)
    assert np.around(S.tau, 4) == 1.0


def test_mphsolver_on_halves():
    N = 20
    S = tau.MultiPhaseSolver(np.ones((N, N, N)), {1: 0.5})
    S.solve(iter_limit=1000)
    assert np.around(S.tau, 4) == 1.0


def test_mphsolver_on_strip_of_ones():
    N = 20
    img = np.zeros((N, N, N))
    img[:, 0:10, 0:10] = 1
    S = tau.MultiPhaseSolver(img)
    S.solve()
    assert np.around(S.tau, 4) == 1.0


def test_mphsolver_on_strip_of_ones_and_twos():
    N = 20
    img = np.zeros((N, N, N))
    img[:, 0:10, 0:10] = 1
    img[:, 0:10, 10:N] = 2
    S = tau.MultiPhaseSolver(img, {0: 0, 1: 1, 2: 0.5})
    S.solve()
    assert np.around(S.tau, 4) == 1


def test_mphsolver_on_strip_of_ones_and_twos_and_threes():
    N = 20
    img = np.ones((N, N, N))
    img[:, 0:10, 0:10] = 2
    img[:, 0:10, 10:N] = 3
    S = tau.MultiPhaseSolver(img, {0: 0, 1: 1, 2: 0.5, 3: 2})
    S.solve()
    assert np.around(S.tau, 4) == 1


def test_mphsolver_allows_zero_Ds_and_label_zero_conductive():
    N = 20
    img = np.zeros((N, N, N))
    img[:, :2] = 1
    s = tau.MultiPhaseSolver(img, {0: 1.0, 1: 0.0})
    s.solve(iter_limit=1000)
    assert np.around(s.tau, 4) == 1


def test_mphsolver_matches_solver_for_binary_case():
    img = staircase(20)
    s_bin = tau.Solver(img)
    s_bin.solve(iter_limit=1000)
    s_mp = tau.MultiPhaseSolver(img)
    s_mp.solve(iter_limit=1000)
    assert np.isclose(float(np.asarray(s_bin.tau)[0]), float(np.asarray(s_mp.tau)[0]), atol=1e-3)


def test_mphsolver_supports_batched_inputs():
    N = 16
    img_a = np.ones((N, N, N))
    img_b = np.ones((N, N, N))
    img_b[:, :, : N // 2] = 2
    Ds = {0: 0.0, 1: 1.0, 2: 0.5}
    s_batch = tau.MultiPhaseSolver(np.stack([img_a, img_b], axis=0), Ds)
    s_batch.solve(iter_limit=1000)
    s_a = tau.MultiPhaseSolver(img_a, Ds)
    s_a.solve(iter_limit=1000)
    s_b = tau.MultiPhaseSolver(img_b, Ds)
    s_b.solve(iter_limit=1000)
    assert np.allclose(np.asarray(s_batch.tau), np.array([s_a.tau[0], s_b.tau[0]]), atol=1e-3)


def test_solve_returns_tau_x_and_prints_contract(capsys):
    S = tau.Solver(staircase(12))
    out = S.solve()
    assert out.shape == (1, 11) and out.dtype == np.float32
    lines = capsys.readouterr().out.splitlines()
    assert lines[0].startswith("converged to: [")
    ram = next(line for line in lines if "GPU-RAM" in line)
    # the benchmark harness tokenises this line (taufactor/benchmark.py:170-179)
    parts = ram.replace("(", "").replace(")", "").replace(",", "").split()
    float(parts[2]), float(parts[6]), float(parts[8])


def test_not_converged_warning_and_resume(capsys):
    S = tau.Solver(staircase(20))
    S.solve(iter_limit=150)
    assert not S.converged and "Warning: not converged" in capsys.readouterr().out
    S.solve()
    assert S.converged and S.iter == 300


# ---- stored notebook outputs ----
@pytest.mark.parametrize("N", [32, 64, 100, 128, 200, 256, 300])
def test_fcc_notebook_table(goldens, N):
    row = next(r for r in goldens["fcc_solver_conv1e-3"]["rows"] if r["N"] == N)
    S = tau.Solver(fcc_pores(N))
    S.solve(conv_crit=1e-3, verbose=False)
    assert S.iter == row["iterations"]
    assert abs(float(S.tau[0]) - row["tau"]) <= 1.5e-6 * row["tau"]


@pytest.mark.parametrize("D", [0.1, 0.5, 1.0, 2.0, 10.0])
def test_multiphase_zigzag_notebook_table(goldens, D):
    row = next(r for r in goldens["multi_zigzag64_conv1e-3"]["rows"] if r["D"] == D)
    S = tau.MultiPhaseSolver(multi_zigzag(64), {0: 0, 1: 1.0, 2: D})
    S.solve(conv_crit=1e-3, verbose=False)
    assert S.iter == row["iterations"]
    assert abs(float(S.tau[0]) - row["tau"]) <= 1.5e-6 * row["tau"]
