"""CPU tests that pin the oracle (oracle/restate.py, oracle/direct.py) to the
reference's own known answers, its stored notebook outputs, an independent
float64 direct solve and closed forms."""
import numpy as np
import pytest

from oracle import direct
from oracle.restate import OracleMultiPhaseSolver, OracleSolver, spans_x
from oracle.structures import fcc_pores, multi_zigzag, staircase


# ---- stored outputs of the reference (tests/golden/reference_goldens.json) ----
@pytest.mark.parametrize("reduce", ["f32", "f64"])
def test_staircase_known_answer(goldens, reduce):
    s = OracleSolver(staircase(20), reduce=reduce).solve()
    g = goldens["staircase_tau"]
    assert np.around(s.tau, g["decimals"]) == g["value"]


@pytest.mark.parametrize("N", [32, 64, 100, 128])
@pytest.mark.parametrize("reduce", ["f32", "f64"])
def test_fcc_notebook_table(goldens, N, reduce):
    row = next(r for r in goldens["fcc_solver_conv1e-3"]["rows"] if r["N"] == N)
    s = OracleSolver(fcc_pores(N), reduce=reduce).solve(conv_crit=1e-3)
    assert s.iter == row["iterations"]
    assert abs(float(s.tau[0]) - row["tau"]) <= 1.5e-6 * row["tau"]


@pytest.mark.parametrize("D", [0.1, 0.5, 1.0, 2.0, 10.0])
def test_multiphase_zigzag_notebook_table(goldens, D):
    row = next(r for r in goldens["multi_zigzag64_conv1e-3"]["rows"] if r["D"] == D)
    s = OracleMultiPhaseSolver(multi_zigzag(64), {0: 0, 1: 1.0, 2: D}, reduce="f32").solve(conv_crit=1e-3)
    assert s.iter == row["iterations"]
    assert abs(float(s.tau[0]) - row["tau"]) <= 1.5e-6 * row["tau"]


# ---- the reference's unit tests (tests/test_taufactor.py), run on the oracle ----
def test_uniform_block():
    N = 20
    img = np.ones((N, N, N))
    img[:, 0] = 0
    assert np.around(OracleSolver(img).solve().tau, 5) == 1.0
    img = np.ones((N * 2, N, N))
    img[:, 0] = 0
    assert np.around(OracleSolver(img).solve().tau, 5) == 1.0
    img = np.ones((N, N, N * 2))
    img[:, 0] = 0
    assert np.around(OracleSolver(img).solve().tau, 5) == 1.0


def test_non_percolating():
    img = np.zeros((20, 20, 20))
    img[:2] = 1
    s = OracleSolver(img).solve(iter_limit=1000)
    assert s.tau == np.inf


def test_strip_of_ones():
    img = np.zeros((20, 20, 20))
    img[:, 0:10, 0:10] = 1
    assert np.around(OracleSolver(img).solve().tau, 5) == 1


def test_deadend():
    solid = np.zeros((10, 50, 50))
    solid[:8, 25, 25] = 1
    s = OracleSolver(solid).solve()
    assert np.around(s.D_eff, 5) == 0
    assert s.tau == np.inf


def test_multiphase_known_answers():
    N = 20
    img = np.zeros((N, N, N))
    img[:2] = 1
    assert OracleMultiPhaseSolver(img).solve(iter_limit=1000).tau == np.inf
    img = np.ones((N, N, N))
    assert np.around(OracleMultiPhaseSolver(img).solve(iter_limit=1000).tau, 4) == 1.0
    assert np.around(OracleMultiPhaseSolver(img, {1: 0.5}).solve(iter_limit=1000).tau, 4) == 1.0
    img = np.zeros((N, N, N))
    img[:, 0:10, 0:10] = 1
    img[:, 0:10, 10:N] = 2
    assert np.around(OracleMultiPhaseSolver(img, {0: 0, 1: 1, 2: 0.5}).solve().tau, 4) == 1
    img = np.ones((N, N, N))
    img[:, 0:10, 0:10] = 2
    img[:, 0:10, 10:N] = 3
    assert np.around(OracleMultiPhaseSolver(img, {0: 0, 1: 1, 2: 0.5, 3: 2}).solve().tau, 4) == 1
    img = np.zeros((N, N, N))
    img[:, :2] = 1
    assert np.around(OracleMultiPhaseSolver(img, {0: 1.0, 1: 0.0}).solve(iter_limit=1000).tau, 4) == 1


def test_multiphase_matches_solver_binary():
    img = staircase(20)
    a = OracleSolver(img).solve(iter_limit=1000)
    b = OracleMultiPhaseSolver(img).solve(iter_limit=1000)
    assert np.isclose(float(a.tau[0]), float(b.tau[0]), atol=1e-3)


def test_multiphase_batch_matches_single():
    N = 16
    a = np.ones((N, N, N))
    b = np.ones((N, N, N))
    b[:, :, : N // 2] = 2
    Ds = {0: 0.0, 1: 1.0, 2: 0.5}
    sb = OracleMultiPhaseSolver(np.stack([a, b]), Ds).solve(iter_limit=1000)
    sa = OracleMultiPhaseSolver(a, Ds).solve(iter_limit=1000)
    s2 = OracleMultiPhaseSolver(b, Ds).solve(iter_limit=1000)
    assert np.allclose(sb.tau, [sa.tau[0], s2.tau[0]], atol=1e-3)


# ---- float64 direct solve: closed forms, then oracle vs direct ----
def test_direct_uniform_is_one():
    assert abs(direct.solver_D_rel(np.ones((7, 5, 4))) - 1.0) < 1e-12


def test_direct_series_layers_closed_form():
    X = 6
    d = np.array([1.0, 0.25, 0.5, 2.0, 1.0, 0.125])
    img = np.repeat(np.arange(X)[:, None, None], 4, axis=1).repeat(3, axis=2)
    Ds = {i: float(d[i]) for i in range(X)}
    k = np.array([float(np.float32(v)) for v in d])
    hm = 2 * k[:-1] * k[1:] / (k[:-1] + k[1:])
    R = 1 / (2 * k[0]) + np.sum(1 / hm) + 1 / (2 * k[-1])
    assert abs(direct.multiphase_D_rel(img, Ds) - X / R) < 1e-12


def test_direct_parallel_strips_closed_form():
    img = np.zeros((8, 6, 6))
    img[:, :3] = 1
    img[:, 3:] = 2
    # two independent straight channels: D_rel = mean diffusivity
    d = float(np.float32(0.3))
    assert abs(direct.multiphase_D_rel(img, {1: 1.0, 2: 0.3}) - (0.5 + 0.5 * d)) < 1e-12


def test_direct_dead_end_and_isolated_cluster():
    img = np.zeros((10, 6, 6))
    img[:, 1, 1] = 1   # straight channel
    img[:8, 4, 4] = 1  # dead end from x=0
    img[3:6, 3, 1] = 1  # isolated piece
    assert abs(direct.solver_D_rel(img) - 1.0 / 36) < 1e-12


def test_oracle_converges_to_direct_solution():
    rng = np.random.default_rng(3)
    img = (rng.random((24, 16, 16)) < 0.6).astype(np.uint8)
    s = OracleSolver(img).solve(conv_crit=1e-5, iter_limit=20000)
    D = direct.solver_D_rel(img)
    assert s.converged
    assert abs(float(s.D_rel[0]) - D) <= 2e-4 * D


def test_oracle_multiphase_converges_to_direct_solution():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 3, size=(20, 14, 12))
    Ds = {0: 0.0, 1: 1.0, 2: 0.1}
    # float32 iterates plateau near rel. error 1.5e-5 on this volume
    s = OracleMultiPhaseSolver(img, dict(Ds)).solve(conv_crit=1e-4, iter_limit=40000)
    D = direct.multiphase_D_rel(img, Ds)
    assert s.converged
    assert abs(float(s.D_rel[0]) - D) <= 2e-4 * D


# ---- percolation test against brute force ----
def _spans_bruteforce(m):
    X, Y, Z = m.shape
    seen = np.zeros_like(m, dtype=bool)
    stack = [(0, y, z) for y in range(Y) for z in range(Z) if m[0, y, z]]
    for p in stack:
        seen[p] = True
    while stack:
        x, y, z = stack.pop()
        if x == X - 1:
            return True
        for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            q = (x + dx, y + dy, z + dz)
            if all(0 <= q[i] < m.shape[i] for i in range(3)) and m[q] and not seen[q]:
                seen[q] = True
                stack.append(q)
    return False


def test_spans_x_matches_bruteforce():
    rng = np.random.default_rng(7)
    for _ in range(60):
        shape = tuple(rng.integers(1, 7, size=3))
        m = rng.random(shape) < rng.uniform(0.2, 0.7)
        assert spans_x(m) == _spans_bruteforce(m)


# ---- periodic solvers (tests/test_taufactor.py:80-106, 226-247 and the notebook table) ----
def test_periodic_known_answers():
    from oracle.restate import OraclePeriodicMultiPhaseSolver, OraclePeriodicSolver

    N = 20
    img = np.ones((N, N, N))
    img[:, 0] = 0
    assert np.around(OraclePeriodicSolver(img).solve().tau, 5) == 1.0
    img = np.zeros((N, N, N))
    img[:2] = 1
    assert OraclePeriodicSolver(img).solve(iter_limit=1000).tau == np.inf
    img = np.zeros((N, N, N))
    img[:, 0:10, 0:10] = 1
    assert np.around(OraclePeriodicSolver(img).solve().tau, 5) == 1
    assert np.around(OraclePeriodicMultiPhaseSolver(np.ones((N, N, N)), {1: 1.0}).solve(iter_limit=1000).tau, 4) == 1.0
    st = staircase(20)
    a = OraclePeriodicSolver(st).solve(iter_limit=1000)
    b = OraclePeriodicMultiPhaseSolver(st, {0: 0.0, 1: 1.0}).solve(iter_limit=1000)
    assert np.isclose(float(a.tau[0]), float(b.tau[0]), atol=1e-3)


@pytest.mark.parametrize("which", ["PeriodicSolver", "PeriodicMultiPhaseSolver"])
def test_periodic_notebook_table(goldens, which):
    from oracle.restate import OraclePeriodicMultiPhaseSolver, OraclePeriodicSolver
    from oracle.structures import diagonals_2d_features

    row = next(r for r in goldens["periodic_diagonal2d64_conv1e-3"]["rows"] if r["solver"] == which)
    img = diagonals_2d_features(64, 2)
    cls = OraclePeriodicSolver if which == "PeriodicSolver" else OraclePeriodicMultiPhaseSolver
    s = cls(img, reduce="f32").solve(conv_crit=1e-3)
    assert s.iter == row["iterations"]
    assert abs(float(s.tau[0]) - row["tau"]) <= 1.5e-6 * row["tau"]


def _wrap_channel(X=12, Y=8, Z=6):
    """A conductive channel that crosses the y boundary halfway along x: it
    connects the two x faces only across the wrap."""
    img = np.zeros((X, Y, Z), dtype=np.uint8)
    img[:X // 2, 0, 2:4] = 1
    img[X // 2 - 1:, Y - 1, 2:4] = 1
    return img


def test_periodic_direct_solve_matches_periodic_sor():
    """The periodic float64 direct solve against the periodic SOR oracle run to
    a tight criterion, on a random two-phase volume and on a channel that
    percolates only across the wrap (where the closed-boundary solve gives 0)."""
    from oracle.restate import OraclePeriodicMultiPhaseSolver, OraclePeriodicSolver

    chan = _wrap_channel()
    assert abs(direct.solver_D_rel(chan)) < 1e-12
    d_per = direct.solve_fixed_point(chan.astype(np.float64), periodic_yz=True)[2]
    o = OraclePeriodicSolver(chan).solve(conv_crit=1e-5, iter_limit=20000)
    assert d_per > 0
    assert abs(float(o.D_rel[0]) - d_per) <= 1e-3 * d_per
    rng = np.random.default_rng(3)
    img = rng.integers(0, 3, size=(10, 8, 8))
    Ds = {0: 0.0, 1: 1.0, 2: 0.2}
    k = np.zeros(img.shape)
    for lbl, d in Ds.items():
        k[img == lbl] = float(np.float32(d))
    d_per = direct.solve_fixed_point(k, periodic_yz=True)[2]
    d_closed = direct.solve_fixed_point(k)[2]
    o = OraclePeriodicMultiPhaseSolver(img, dict(Ds)).solve(conv_crit=1e-5, iter_limit=20000)
    assert abs(float(o.D_rel[0]) - d_per) <= 1e-3 * d_per
    assert abs(d_per - d_closed) > 1e-2 * d_per  # the wrap faces matter here
