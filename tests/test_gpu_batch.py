"""GPU: the uneven batch gives every volume the result of its own solo solve.

Each volume of a BatchSolver must stop at the same iteration as
Solver(img).solve() with the same tau (bit for bit: both evaluate the
check in float64 with the same operation order) and hold the same field.
Graph-captured periods are compared with eager periods, replayed several
times with host synchronisation between replays."""
import numpy as np
import pytest
import torch

import taufactor_amd as tau

pytestmark = pytest.mark.gpu


def _volumes(seed=0, n=9):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        shape = tuple(int(v) for v in rng.integers(6, 40, size=3))
        p = rng.uniform(0.45, 0.8)
        out.append((rng.random(shape) < p).astype(np.uint8))
    # special cases: non-percolating, straight channel, 2-D
    blocked = np.ones((12, 9, 10), dtype=np.uint8)
    blocked[5] = 0
    out.append(blocked)
    chan = np.zeros((20, 7, 7), dtype=np.uint8)
    chan[:, 3, 3] = 1
    out.append(chan)
    out.append((rng.random((25, 31)) < 0.7).astype(np.uint8))
    return out


def _solo(img):
    s = tau.Solver(img)
    s.solve(verbose=False)
    return s


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("H", [2, 3, 4, 5, 6])
def test_batch_equals_solo_solves(use_graph, H):
    imgs = _volumes(seed=H)
    b = tau.BatchSolver(imgs)
    b.half_sweeps_per_pass = H
    b.solve(use_graph=use_graph)
    for v, img in enumerate(imgs):
        s = _solo(img)
        assert b.iters[v] == s.iter, (v, b.iters[v], s.iter)
        assert bool(b.converged[v]) == s.converged
        assert b.tau[v] == s.tau[0] or (np.isinf(b.tau[v]) and np.isinf(s.tau[0])), (v, b.tau[v], s.tau[0])
        assert np.array_equal(b.field(v).cpu().numpy(), s.field[0].cpu().numpy()), v


def test_batch_volumes_match_oracle():
    """Independent of the solo GPU path: two volumes of a batch against the oracle."""
    from oracle.restate import OracleSolver

    imgs = _volumes(seed=5, n=3)
    b = tau.BatchSolver(imgs)
    b.solve()
    for v in (0, 2):
        o = OracleSolver(imgs[v]).solve()
        assert b.iters[v] == o.iter
        assert np.array_equal(b.field(v).cpu().numpy(), o.field.numpy()[0])


def test_batch_iteration_limit_and_reset():
    imgs = _volumes(seed=11, n=4)
    b = tau.BatchSolver(imgs)
    b.solve(iter_limit=200)
    assert (b.iters <= 200).all()
    first = b.tau.copy()
    b.reset()
    b.solve(iter_limit=200)
    assert np.array_equal(first, b.tau) or np.allclose(first, b.tau, equal_nan=True)


@pytest.mark.parametrize("use_graph", [False, True])
def test_work_queue_equals_full_dispatch(use_graph):
    """Passes as a work queue over the live tiles (the default) and passes
    that dispatch every tile give the same iterations and fields; volumes
    stop at different checks, so later passes run on fewer live tiles."""
    vols = _volumes(seed=5, n=14)
    a = tau.BatchSolver(vols)
    b = tau.BatchSolver(vols)
    b.use_queue = False
    for s in (a, b):
        s.half_sweeps_per_pass = 3
        s.tile = (8, 2, 16)
        s.solve(use_graph=use_graph)
    assert len(set(a.iters.tolist())) > 2
    assert np.array_equal(a.iters, b.iters) and np.array_equal(a.tau, b.tau)
    for v in range(len(vols)):
        assert np.array_equal(a.field(v).cpu().numpy(), b.field(v).cpu().numpy()), v


def test_solve_twice_gives_the_same_result():
    """Later solves start from reset(), which copies the initial field saved
    at construction into both field buffers: every reset leaves the buffers
    the set-up left (codes and counts untouched by the solves), and every
    solve gives the first one's result."""
    vols = _volumes(seed=6, n=5)
    b = tau.BatchSolver(vols)
    n = b._n_padded
    initial = [t[:n].clone() for t in (b._fa, b._fb, b._codes)] + [b._count.clone(), b._sweep_codes.clone()]
    b.solve()
    first = (b.iters.copy(), b.tau.copy(), [b.field(v).cpu().numpy() for v in range(len(vols))])
    for _ in range(2):
        b.reset()
        torch.cuda.synchronize()
        after = [t[:n] for t in (b._fa, b._fb, b._codes)] + [b._count, b._sweep_codes]
        for x, y in zip(initial, after):
            assert torch.equal(x, y)
        b.solve()
        assert np.array_equal(first[0], b.iters) and np.array_equal(first[1], b.tau)
        for v in range(len(vols)):
            assert np.array_equal(first[2][v], b.field(v).cpu().numpy())


@pytest.mark.parametrize("H", [1, 4, 6])
def test_batch_whole_z_windows(H):
    """Volumes whose rows one z window covers (TZ < Z <= 128: one wide tile
    per row tile from z = 1) next to ones with z strips: the batch equals the
    oracle's iterates and the solo solves stop where the batch does."""
    from oracle.restate import OracleSolver

    rng = np.random.default_rng(40 + H)
    shapes = [(9, 14, 128), (7, 30, 125), (6, 11, 121), (8, 12, 60), (5, 9, 130)]
    imgs = [(rng.random(sh) < 0.6).astype(np.uint8) for sh in shapes]
    b = tau.BatchSolver(imgs)
    b.half_sweeps_per_pass = H
    b.solve(iter_limit=100)  # no volume can stop at its first check
    for v, img in enumerate(imgs):
        assert b.iters[v] == 100
        o = OracleSolver(img).run(100)
        assert np.array_equal(b.field(v).cpu().numpy(), o.field.numpy()[0]), (v, shapes[v])
    b = tau.BatchSolver(imgs)
    b.half_sweeps_per_pass = H
    b.solve()
    for v, img in enumerate(imgs):
        s = _solo(img)
        assert b.iters[v] == s.iter and b.tau[v] == s.tau[0], v
        assert np.array_equal(b.field(v).cpu().numpy(), s.field[0].cpu().numpy()), v
