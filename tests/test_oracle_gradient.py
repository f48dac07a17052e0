"""CPU: pin the gradient oracles (oracle/direct.py) that the HIP adjoint
kernel is checked against.

* the dissipation identity D_rel = (X/(Y Z)) * W(c*) holds at the direct solution;
* torch autograd through the dense float64 direct solve agrees with central
  finite differences of the sparse direct solve;
* the closed-form gradient the kernel evaluates (energy_gradient) agrees with
  that autograd gradient, and at k = 0 with a one-sided finite difference."""
import numpy as np
import torch

from oracle import direct


def _k(shape, seed, zero_frac=0.0):
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.2, 2.0, size=shape)
    if zero_frac:
        k[rng.random(shape) < zero_frac] = 0.0
    return k


def test_dissipation_identity():
    k = _k((7, 5, 6), 0, zero_frac=0.2)
    c, flux_1d, D = direct.solve_fixed_point(k)
    X, Y, Z = k.shape
    assert abs(D - X / (Y * Z) * direct.dissipation(k, c)) < 1e-10 * D


def test_two_direct_solves_agree():
    k = _k((6, 4, 5), 1)
    assert abs(direct.solve_fixed_point(k)[2] - float(direct.torch_D_rel(torch.tensor(k)))) < 1e-12


def test_autograd_direct_solve_matches_finite_differences():
    k = _k((6, 4, 5), 2)
    kt = torch.tensor(k, requires_grad=True)
    (g,) = torch.autograd.grad(direct.torch_D_rel(kt), kt)
    rng = np.random.default_rng(3)
    for _ in range(8):
        i = tuple(int(rng.integers(0, s)) for s in k.shape)
        h = 1e-5 * k[i]
        kp, km = k.copy(), k.copy()
        kp[i] += h
        km[i] -= h
        fd = (direct.solve_fixed_point(kp)[2] - direct.solve_fixed_point(km)[2]) / (2 * h)
        assert abs(fd - float(g[i])) <= 1e-6 * abs(fd) + 1e-12


def test_energy_gradient_matches_autograd():
    k = _k((6, 5, 4), 4)
    kt = torch.tensor(k, requires_grad=True)
    (g,) = torch.autograd.grad(direct.torch_D_rel(kt), kt)
    c = direct.solve_fixed_point(k)[0]
    ge = direct.energy_gradient(k, c)
    np.testing.assert_allclose(ge, g.numpy(), rtol=1e-8, atol=1e-14)


def test_energy_gradient_one_sided_at_zero():
    k = _k((6, 5, 5), 5, zero_frac=0.25)
    c = direct.solve_fixed_point(k)[0]
    ge = direct.energy_gradient(k, c)
    zeros = np.argwhere(k == 0)
    D0 = direct.solve_fixed_point(k)[2]
    for i in map(tuple, zeros[:6]):
        h = 1e-7
        kp = k.copy()
        kp[i] = h
        fd = (direct.solve_fixed_point(kp)[2] - D0) / h
        assert abs(fd - ge[i]) <= 1e-4 * max(abs(fd), 1e-8), (i, fd, ge[i])


def test_per_phase_gradient_is_sum_of_voxel_gradient():
    rng = np.random.default_rng(6)
    img = rng.integers(0, 3, size=(7, 6, 5))
    Ds = {0: 0.0, 1: 1.0, 2: 0.3}
    k = np.zeros(img.shape)
    for l, d in Ds.items():
        k[img == l] = float(np.float32(d))
    c = direct.solve_fixed_point(k)[0]
    ge = direct.energy_gradient(k, c)
    for lbl in (1, 2):
        h = 1e-6
        kp, km = k.copy(), k.copy()  # float64 steps (no float32 rounding of D +- h)
        kp[img == lbl] += h
        km[img == lbl] -= h
        fd = (direct.solve_fixed_point(kp)[2] - direct.solve_fixed_point(km)[2]) / (2 * h)
        assert abs(fd - ge[img == lbl].sum()) <= 1e-6 * abs(fd)
