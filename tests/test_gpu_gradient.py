"""GPU: per-voxel diffusivity solve and the implicit-adjoint gradient.

* FieldSolver iterates equal the oracle's (MultiPhaseSolver restated with a
  per-voxel diffusivity map) value for value;
* the HIP gradient kernel equals the closed form evaluated on the same field
  (oracle/direct.energy_gradient, float64) to rtol 1e-5;
* at tight convergence the gradient matches torch autograd through a dense
  float64 direct solve, and the per-phase gradient of MultiPhaseSolver
  matches central finite differences of the float64 direct solve, both to
  rtol 1e-3 (the float32 iterate limits the agreement);
* the torch.autograd.Function returns the kernel's gradient."""
import numpy as np
import pytest
import torch

import taufactor_amd as tau
from oracle import direct
from oracle.restate import OracleFieldSolver

pytestmark = pytest.mark.gpu


def _k(shape, seed, zero_frac=0.0):
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.2, 2.0, size=shape).astype(np.float32)
    if zero_frac:
        k[rng.random(shape) < zero_frac] = 0.0
    return k


@pytest.mark.parametrize("shape", [(9, 11, 13), (30, 17, 140), (1, 5, 4)])
@pytest.mark.parametrize("H", [1, 2, 4])
def test_field_iterates_bit_identical(shape, H):
    k = _k(shape, sum(shape) + H, zero_frac=0.3)
    s = tau.FieldSolver(k)
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=21, verbose=False)
    o = OracleFieldSolver(k).run(21)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


@pytest.mark.parametrize("tile", [(8, 2, 6), (12, 2, 0), (16, 2, 4)])
@pytest.mark.parametrize("H", [2, 3, 4])
def test_field_iterates_with_tiles_and_batch(tile, H):
    """Two volumes at once, every compiled field shape: the register-blocked
    sweep reads the precomputed face conductances (tf_field_faces)."""
    k = np.stack([_k((11, 30, 70), 40 + H, zero_frac=0.25), _k((11, 30, 70), 41 + H)])
    s = tau.FieldSolver(k)
    s.half_sweeps_per_pass = H
    s.tile = tile
    s.solve(iter_limit=2 * H + 3, verbose=False)
    o = OracleFieldSolver(k).run(2 * H + 3)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


def test_field_full_solve_matches_oracle():
    k = _k((28, 20, 24), 3, zero_frac=0.2)
    s = tau.FieldSolver(k)
    s.solve(verbose=False)
    o = OracleFieldSolver(k).solve()
    assert s.iter == o.iter
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())
    np.testing.assert_allclose(s.tau, o.tau, rtol=2e-7)


@pytest.mark.parametrize("zero_frac", [0.0, 0.25])
def test_gradient_kernel_equals_closed_form_on_same_field(zero_frac):
    k = _k((10, 9, 8), 4, zero_frac=zero_frac)
    s = tau.FieldSolver(k)
    s.solve(verbose=False, conv_crit=1e-3)
    g = s.grad_D_eff().cpu().numpy()[0]
    c = s.field.cpu().numpy()[0, 1:-1, 1:-1, 1:-1].astype(np.float64)
    ge = direct.energy_gradient(k.astype(np.float64), c)
    np.testing.assert_allclose(g, ge, rtol=1e-5, atol=1e-7 * np.abs(ge).max())


@pytest.mark.parametrize("n_labels", [3, 6, 11])
def test_phase_gradient_is_the_voxel_closed_form_summed_per_label(n_labels):
    """MultiPhaseSolver.grad_D_eff (tf_grad_classes: register sums per class
    for up to 4 and up to 8 classes, per-voxel wave sums beyond) equals the
    float64 closed form of every voxel on the same field, summed per label."""
    rng = np.random.default_rng(60 + n_labels)
    labels = list(range(n_labels))
    Ds = {lbl: (0.0 if lbl == 0 else float(np.float32(rng.uniform(0.1, 2.0)))) for lbl in labels}
    img = rng.choice(labels, size=(12, 10, 9))
    s = tau.MultiPhaseSolver(img, dict(Ds))
    s.solve(verbose=False, conv_crit=1e-3)
    g = s.grad_D_eff()
    c = s.field.cpu().numpy()[0, 1:-1, 1:-1, 1:-1].astype(np.float64)
    kvox = np.array([np.float32(Ds[lbl]) for lbl in labels], dtype=np.float64)[img]
    ge = direct.energy_gradient(kvox, c)
    for lbl in labels:
        want = ge[img == lbl].sum() * s.D_0
        np.testing.assert_allclose(g[lbl][0], want, rtol=1e-7, atol=1e-12 * np.abs(ge).sum())


def test_gradient_matches_autograd_direct_solve():
    k = _k((10, 8, 8), 5)
    s = tau.FieldSolver(k)
    s.solve(verbose=False, conv_crit=1e-4, iter_limit=100000)
    g = s.grad_D_eff().cpu().numpy()[0]
    kt = torch.tensor(k.astype(np.float64), requires_grad=True)
    (ga,) = torch.autograd.grad(direct.torch_D_rel(kt), kt)
    ga = ga.numpy()
    assert np.max(np.abs(g - ga)) <= 1e-3 * np.max(np.abs(ga))
    D = float(direct.torch_D_rel(torch.tensor(k.astype(np.float64))))
    assert abs(float(s.D_rel[0]) - D) <= 1e-3 * D


def test_multiphase_phase_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    img = rng.integers(0, 3, size=(16, 12, 12))
    Ds = {0: 0.0, 1: 1.0, 2: 0.1}
    s = tau.MultiPhaseSolver(img, dict(Ds))
    s.solve(verbose=False, conv_crit=1e-4, iter_limit=100000)
    g = s.grad_D_eff()
    k = np.zeros(img.shape)
    for lbl, d in Ds.items():
        k[img == lbl] = float(np.float32(d))
    for lbl in (1, 2):
        h = 1e-3 * Ds[lbl]
        kp, km = k.copy(), k.copy()
        kp[img == lbl] += h
        km[img == lbl] -= h
        fd = (direct.solve_fixed_point(kp)[2] - direct.solve_fixed_point(km)[2]) / (2 * h)
        assert abs(g[lbl][0] - fd) <= 1e-3 * abs(fd), (lbl, g[lbl][0], fd)


def test_periodic_multiphase_phase_gradient_matches_finite_differences():
    """PeriodicMultiPhaseSolver.grad_D_eff: the gradient kernel reads the
    faces across the y/z boundary through the ghosts (wrapped class codes, the
    wrapped field of the last iteration); central differences of the periodic
    float64 direct solve."""
    rng = np.random.default_rng(8)
    img = rng.integers(0, 3, size=(16, 12, 10))
    Ds = {0: 0.0, 1: 1.0, 2: 0.1}
    s = tau.PeriodicMultiPhaseSolver(img, dict(Ds))
    s.solve(verbose=False, conv_crit=1e-4, iter_limit=100000)
    g = s.grad_D_eff()
    k = np.zeros(img.shape)
    for lbl, d in Ds.items():
        k[img == lbl] = float(np.float32(d))
    for lbl in (1, 2):
        h = 1e-3 * Ds[lbl]
        kp, km = k.copy(), k.copy()
        kp[img == lbl] += h
        km[img == lbl] -= h
        fd = (direct.solve_fixed_point(kp, periodic_yz=True)[2] -
              direct.solve_fixed_point(km, periodic_yz=True)[2]) / (2 * h)
        assert abs(g[lbl][0] - fd) <= 1e-3 * abs(fd), (lbl, g[lbl][0], fd)


def test_autograd_function_returns_kernel_gradient():
    k = torch.tensor(_k((2, 12, 10, 9), 7), device="cuda", requires_grad=True)
    D = tau.effective_diffusivity(k, conv_crit=1e-3)
    (gk,) = torch.autograd.grad((D * torch.tensor([1.0, 2.0], dtype=D.dtype, device=D.device)).sum(), k)
    s = tau.FieldSolver(k.detach())
    s.solve(verbose=False, conv_crit=1e-3)
    g = s.grad_D_eff()
    assert torch.equal(gk[0], g[0]) and torch.allclose(gk[1], 2 * g[1])
    np.testing.assert_allclose(D.detach().cpu().numpy(), s.D_rel, rtol=0, atol=0)
