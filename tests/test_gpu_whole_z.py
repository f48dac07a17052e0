"""GPU: whole-z windows of the sweep against the oracle.

With closed y/z boundaries and TZ < Z <= 128 (TZ: the owned z of a tile at
the launch's depth, 116..124) one wave window covers z = 1..Z and the DPP
shifts stand in for the z ghosts (value 0, zero class, zero face).  Iterates
equal the oracle's value for value for every kind that sweeps: Solver,
MultiPhase through class codes and through face codes, FieldSolver,
AnisotropicSolver, ElectrodeSolver; at every fused depth, on the compiled
workgroup shapes, odd and even Z, passes that start on either colour, a
batch, and a full solve.  The launch covers each row with one z tile; the
periodic solvers keep the z strips (their ghosts hold wrapped values)."""
import math

import numpy as np
import pytest

import taufactor_amd as tau
from oracle.restate import (OracleAnisotropicSolver, OracleElectrodeSolver, OracleFieldSolver,
                            OracleMultiPhaseSolver, OraclePeriodicSolver, OracleSolver)

pytestmark = pytest.mark.gpu


class ClassCodeSolver(tau.MultiPhaseSolver):
    face_codes = False


def _binary(shape, seed, p=0.6):
    return (np.random.default_rng(seed).random(shape) < p).astype(np.uint8)


def _labels(shape, labels, seed):
    return np.random.default_rng(seed).choice(labels, size=shape)


def _same(s, o):
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


def _z_tiles_per_row(s):
    """Tiles per (x tile, row tile, volume) of the solver's next launch."""
    H, (nw, rpw, tx), n = s.resolved_launch()
    ty = nw * rpw - 2 * H
    rows = math.ceil(s.Ny / ty) * math.ceil(s.Nx / tx) * s.batch_size
    return n / rows


@pytest.mark.parametrize("shape", [(7, 19, 128), (6, 9, 127), (5, 30, 125), (4, 11, 121), (3, 8, 126)])
@pytest.mark.parametrize("H", [1, 2, 3, 4, 5, 6])
def test_solver_iterates(shape, H):
    img = _binary(shape, seed=sum(shape) + H)
    s = tau.Solver(img)
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=2 * H + 11, verbose=False)
    o = OracleSolver(img).run(2 * H + 11)
    _same(s, o)


@pytest.mark.parametrize("tile", [(8, 2, 4), (4, 4, 6), (12, 2, 0), (16, 2, 2), (8, 4, 2)])
@pytest.mark.parametrize("H", [2, 3, 4])
def test_solver_tiles_and_colours(tile, H):
    if tile[1] == 4 and H > 4:
        pytest.skip("four-row waves are compiled for n_half <= 4")
    img = _binary((9, 40, 128), seed=3 + H)
    s = tau.Solver(img)
    s.half_sweeps_per_pass = H
    s.tile = tile
    assert _z_tiles_per_row(s) == 1
    s.solve(iter_limit=3, verbose=False)       # later passes start on colour 1
    s.solve(iter_limit=3 + 4 * H + 5, verbose=False)
    o = OracleSolver(img).run(3 + 4 * H + 5)
    _same(s, o)


def test_solver_batch_and_full_solve():
    img = np.stack([_binary((10, 14, 128), seed=31), _binary((10, 14, 128), seed=32, p=0.7)])
    s = tau.Solver(img)
    s.half_sweeps_per_pass = 4
    s.solve(iter_limit=21, verbose=False)
    _same(s, OracleSolver(img).run(21))
    one = _binary((12, 16, 127), seed=33, p=0.7)
    s = tau.Solver(one)
    s.solve(verbose=False, conv_crit=1e-3)
    o = OracleSolver(one).solve(conv_crit=1e-3)
    assert s.iter == o.iter
    _same(s, o)
    np.testing.assert_allclose(s.tau, o.tau, rtol=2e-7)


@pytest.mark.parametrize("cls", [tau.MultiPhaseSolver, ClassCodeSolver])
@pytest.mark.parametrize("shape", [(6, 13, 128), (5, 21, 127)])
@pytest.mark.parametrize("H", [1, 3, 4, 6])
def test_multiphase_iterates(cls, shape, H):
    Ds = {0: 0.0, 1: 1.0, 2: 0.1, 3: 2.5}
    img = _labels(shape, sorted(Ds), seed=sum(shape) + H)
    s = cls(img, dict(Ds))
    assert (s._faces is not None) == (cls is tau.MultiPhaseSolver)
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=2 * H + 9, verbose=False)
    _same(s, OracleMultiPhaseSolver(img, dict(Ds)).run(2 * H + 9))


@pytest.mark.parametrize("shape", [(6, 13, 128), (5, 17, 126), (4, 9, 123)])
@pytest.mark.parametrize("H", [1, 2, 4, 5])
def test_field_iterates(shape, H):
    rng = np.random.default_rng(sum(shape) + H)
    k = (rng.random(shape) * (rng.random(shape) < 0.8)).astype(np.float32)
    s = tau.FieldSolver(k)
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=2 * H + 9, verbose=False)
    _same(s, OracleFieldSolver(k).run(2 * H + 9))


@pytest.mark.parametrize("H", [2, 4])
def test_anisotropic_iterates(H):
    img = _binary((6, 13, 128), seed=40 + H)
    s = tau.AnisotropicSolver(img, (1.0, 1.5, 0.7))
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=23, verbose=False)
    _same(s, OracleAnisotropicSolver(img, (1.0, 1.5, 0.7)).run(23))


@pytest.mark.parametrize("H", [1, 4])
def test_electrode_iterates(H):
    img = np.random.default_rng(50 + H).choice([0, 1, 5], size=(6, 13, 127), p=[0.3, 0.55, 0.15])
    s = tau.ElectrodeSolver(img, spacing=0.5)
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=23, verbose=False)
    _same(s, OracleElectrodeSolver(img, spacing=0.5).run(23))


def test_periodic_keeps_z_strips():
    img = _binary((6, 10, 128), seed=60)
    s = tau.PeriodicSolver(img)
    s.half_sweeps_per_pass = 4
    s.tile = (8, 2, 0)
    assert _z_tiles_per_row(s) > 1
    s.solve(iter_limit=17, verbose=False)
    _same(s, OraclePeriodicSolver(img).run(17))


def test_field_128_cube_uses_one_window():
    k = np.ones((2, 1, 130, 128), dtype=np.float32)
    s = tau.FieldSolver(k)
    for H in (2, 4, 6):
        s.half_sweeps_per_pass = H
        s.tile = (12, 2, 0)
        assert _z_tiles_per_row(s) == 1
