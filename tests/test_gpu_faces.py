"""GPU: the MultiPhase sweep through per-voxel face codes
(TF_KIND_MULTIPHASE_FACES), which MultiPhaseSolver uses whenever its
diffusivities fit the face tables (at most four distinct values with 0).

* iterates equal the oracle's (OracleMultiPhaseSolver, the restated
  taufactor/taufactor.py:535-620) value for value at every fused depth and
  compiled workgroup shape, on ragged shapes, X = 1, periodic y/z, and a
  batch;
* they equal the class-code sweep's (face_codes = False) bit for bit;
* a field of tiny and subnormal values takes the IEEE division of the rare
  path and still equals the oracle;
* diffusivities that do not fit fall back to the class-code sweep, which
  still equals the oracle;
* a full solve stops at the oracle's iteration with its tau;
* more than 63 labels with more than four diffusivities sweep as the
  per-voxel field kind and still equal the oracle;
* more than 255 labels take one class per distinct diffusivity and still
  equal the oracle."""
import numpy as np
import pytest
import torch

import taufactor_amd as tau
from oracle.restate import OracleMultiPhaseSolver, OraclePeriodicMultiPhaseSolver

pytestmark = pytest.mark.gpu


class ClassCodeSolver(tau.MultiPhaseSolver):
    face_codes = False


def _labels(shape, labels, seed):
    return np.random.default_rng(seed).choice(labels, size=shape)


DS = [{0: 0.0, 1: 1.0, 2: 0.1}, {0: 0.0, 1: 1.0, 2: 0.1, 3: 2.5}, {3: 0.7, 4: 0.0, 9: 1.0 / 3.0},
      {1: 1e-3, 2: 1e3}]


@pytest.mark.parametrize("shape", [(5, 7, 9), (33, 17, 70), (21, 40, 130), (1, 6, 5), (2, 3, 1)])
@pytest.mark.parametrize("H", [1, 2, 3, 4, 5, 6])
def test_face_code_iterates_equal_the_oracle(shape, H):
    Ds = DS[(sum(shape) + H) % len(DS)]
    img = _labels(shape, sorted(Ds), seed=sum(shape) * 7 + H)
    s = tau.MultiPhaseSolver(img, dict(Ds))
    assert s._faces is not None
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=2 * H + 9, verbose=False)
    o = OracleMultiPhaseSolver(img, dict(Ds)).run(2 * H + 9)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


@pytest.mark.parametrize("tile", [(8, 2, 6), (12, 2, 0), (16, 2, 4), (4, 4, 8)])
@pytest.mark.parametrize("H", [2, 4])
def test_face_code_tiles_and_batch(tile, H):
    Ds = DS[1]
    img = np.stack([_labels((14, 37, 66), sorted(Ds), seed=40 + H), _labels((14, 37, 66), sorted(Ds), seed=41)])
    s = tau.MultiPhaseSolver(img, dict(Ds))
    s.half_sweeps_per_pass = H
    s.tile = tile
    s.solve(iter_limit=3 * H + 1, verbose=False)
    o = OracleMultiPhaseSolver(img, dict(Ds)).run(3 * H + 1)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


@pytest.mark.parametrize("H", [1, 3, 6])
def test_face_codes_equal_class_codes(H):
    Ds = DS[0]
    img = _labels((40, 50, 60), sorted(Ds), seed=H)
    a = tau.MultiPhaseSolver(img, dict(Ds))
    b = ClassCodeSolver(img, dict(Ds))
    assert a._faces is not None and b._faces is None
    for s in (a, b):
        s.half_sweeps_per_pass = H
        s.solve(iter_limit=60, verbose=False)
    assert torch.equal(a.field.view(torch.int32), b.field.view(torch.int32))


@pytest.mark.parametrize("shape", [(6, 8, 10), (9, 4, 14)])
def test_periodic_face_codes_equal_the_oracle(shape):
    Ds = DS[0]
    img = _labels(shape, sorted(Ds), seed=3)
    s = tau.PeriodicMultiPhaseSolver(img, dict(Ds))
    assert s._faces is not None
    s.solve(iter_limit=23, verbose=False)
    o = OraclePeriodicMultiPhaseSolver(img, dict(Ds)).run(23)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


def test_periodic_odd_sizes_keep_the_class_codes():
    Ds = DS[0]
    img = _labels((6, 7, 8), sorted(Ds), seed=4)
    s = tau.PeriodicMultiPhaseSolver(img, dict(Ds))
    assert s._faces is None
    s.solve(iter_limit=9, verbose=False)
    o = OraclePeriodicMultiPhaseSolver(img, dict(Ds)).run(9)
    # interiors: the unfused odd-size solve leaves other values in the ghosts
    # (the wrap made before its last iteration is not repeated)
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    assert np.array_equal(s.field.cpu().numpy()[inner], o.field.numpy()[inner])


@pytest.mark.parametrize("H", [1, 4])
@pytest.mark.parametrize("ds", [3, 0])
def test_face_codes_tiny_and_subnormal_sums(H, ds):
    """Neighbour sums below 2^-100 take the IEEE division (the rare path);
    with the prefactors of {0, 1, 0.1} many sums are subnormal, where a
    residual can round to zero and only the guard on s is exact."""
    Ds = DS[ds]
    img = _labels((12, 10, 14), sorted(Ds), seed=21)
    rng = np.random.default_rng(5)
    vals = np.array([0.0, 2.0 ** -110, -(2.0 ** -112), 3e-40, -1e-42, 7e-45, 2.0 ** -101,
                     -(2.0 ** -126), 0.5], dtype=np.float32)
    s = tau.MultiPhaseSolver(img, dict(Ds))
    o = OracleMultiPhaseSolver(img, dict(Ds))
    f = vals[rng.integers(0, len(vals), size=tuple(o.field.shape))].astype(np.float32)
    f[:, :, [0, -1], :] = 0
    f[:, :, :, [0, -1]] = 0
    s.field = torch.from_numpy(f)
    o.field = torch.from_numpy(f.copy())
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=2 * H + 1, verbose=False)
    o.run(2 * H + 1)
    got, want = s.field.cpu().numpy(), o.field.numpy()
    assert np.array_equal(got, want)
    assert np.count_nonzero((np.abs(want) < 2.0 ** -126) & (want != 0)) > 0


def test_unfit_diffusivities_fall_back():
    Ds = {0: 0.0, 1: 1.0, 2: 0.1, 3: 2.5, 4: 7.0}
    img = _labels((10, 12, 14), sorted(Ds), seed=8)
    s = tau.MultiPhaseSolver(img, dict(Ds))
    assert s._faces is None
    s.solve(iter_limit=15, verbose=False)
    o = OracleMultiPhaseSolver(img, dict(Ds)).run(15)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


def test_face_code_full_solve():
    Ds = DS[0]
    img = _labels((30, 28, 26), sorted(Ds), seed=11)
    s = tau.MultiPhaseSolver(img, dict(Ds))
    s.solve(verbose=False, conv_crit=1e-3)
    o = OracleMultiPhaseSolver(img, dict(Ds)).solve(conv_crit=1e-3)
    assert s.iter == o.iter
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())
    assert float(s.tau[0]) == float(np.float32(o.tau[0]))


def test_autotuned_face_code_sweep_at_tuning_size():
    """A volume of at least autotune.MIN_VOXELS voxels is autotuned before its
    first iteration: the tuner times the face-code sweep the solver runs (its
    kind, face codes and pair table), and the tuned launches still equal the
    oracle."""
    from taufactor_amd import autotune

    Ds = DS[0]
    shape = (220, 200, 192)
    assert np.prod(shape) >= autotune.MIN_VOXELS
    img = _labels(shape, sorted(Ds), seed=13)
    s = tau.MultiPhaseSolver(img, dict(Ds))
    assert s._faces is not None
    s.solve(iter_limit=7, verbose=False)
    assert autotune.last_timings, "the tuner ran"
    o = OracleMultiPhaseSolver(img, dict(Ds)).run(7)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


def test_many_labels_with_few_diffusivities():
    """More than 63 labels (the class-code sweep's limit) when the labels
    take at most four diffusivities: the face-code sweep serves them, the
    layer sums read the per-label class codes, and a full solve equals the
    oracle's; their per-phase gradient is refused (63-label LDS table)."""
    rng = np.random.default_rng(17)
    labels = np.arange(100)
    Ds = {int(l): [0.0, 1.0, 0.1][l % 3] for l in labels}
    img = rng.choice(labels, size=(18, 16, 20))
    s = tau.MultiPhaseSolver(img, dict(Ds))
    assert s._faces is not None and s._n_classes == 101
    s.solve(verbose=False, conv_crit=1e-3)
    o = OracleMultiPhaseSolver(img, dict(Ds)).solve(conv_crit=1e-3)
    assert s.iter == o.iter
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())
    assert float(s.tau[0]) == float(np.float32(o.tau[0]))
    with pytest.raises(ValueError):
        s.grad_D_eff()


@pytest.mark.parametrize("H", [1, 4, 6])
def test_many_labels_with_many_diffusivities(H):
    """More than 63 labels taking more than four diffusivities: the per-voxel
    field sweep over the face conductances of the label diffusivities serves
    them, and its iterates equal the MultiPhase oracle's; a full solve stops
    at the oracle's iteration with its tau.  With periodic boundaries (no
    field sweep there) they are refused."""
    rng = np.random.default_rng(19 + H)
    labels = np.arange(100)
    many = {int(l): float(np.float32(0.1 * (l % 7))) for l in labels}
    img = rng.choice(labels, size=(14, 18, 22))
    s = tau.MultiPhaseSolver(img, dict(many))
    assert s._faces is None and s._field_faces is not None and s._n_classes == 101
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=2 * H + 9, verbose=False)
    o = OracleMultiPhaseSolver(img, dict(many)).run(2 * H + 9)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())
    if H == 4:
        s = tau.MultiPhaseSolver(img, dict(many))
        s.solve(verbose=False, conv_crit=1e-3)
        o = OracleMultiPhaseSolver(img, dict(many)).solve(conv_crit=1e-3)
        assert s.iter == o.iter
        assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())
        assert float(s.tau[0]) == float(np.float32(o.tau[0]))
        with pytest.raises(ValueError):
            tau.PeriodicMultiPhaseSolver(rng.choice(labels, size=(6, 8, 10)), dict(many))


@pytest.mark.parametrize("n_values", [3, 70])
def test_more_labels_than_class_codes(n_values):
    """More than 255 labels (the one-byte class codes' limit) take one class
    per distinct diffusivity: with three diffusivities the face-code sweep
    serves them, with seventy the per-voxel field sweep.  A full solve stops
    at the oracle's iteration with its field and tau; the per-label gradient
    is refused."""
    rng = np.random.default_rng(23 + n_values)
    labels = np.arange(1000, 1300)
    vals = [0.0, 1.0, 0.1] if n_values == 3 else [float(np.float32(0.05 * (i + 1))) for i in range(70)]
    Ds = {int(lbl): vals[i % len(vals)] for i, lbl in enumerate(labels)}
    img = rng.choice(labels, size=(16, 14, 18))
    s = tau.MultiPhaseSolver(img, dict(Ds))
    assert s._grouped and s._n_classes == len(set(vals)) + 1
    assert (s._faces is not None) if n_values == 3 else (s._field_faces is not None)
    s.solve(verbose=False, conv_crit=1e-3)
    o = OracleMultiPhaseSolver(img, dict(Ds)).solve(conv_crit=1e-3)
    assert s.iter == o.iter
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())
    assert float(s.tau[0]) == float(np.float32(o.tau[0]))
    with pytest.raises(ValueError):
        s.grad_D_eff()
