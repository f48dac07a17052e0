"""CPU: input validation of the drop-in API raises what the reference raises
(taufactor/taufactor.py:194-204, 387-397, 535-558), before any device work;
there is no CPU fallback (a CPU device raises)."""
import numpy as np
import pytest
import torch

import taufactor_amd as tau


def test_non_ndarray_raises_type_error():
    with pytest.raises(TypeError, match="NumPy array"):
        tau.Solver([[0, 1], [1, 0]])


def test_bad_ndim_raises_value_error():
    with pytest.raises(ValueError, match=r"expected \[B, X, Y, Z\]"):
        tau.Solver(np.ones((2, 2, 2, 2, 2)))


def test_non_binary_labels_raise():
    img = np.zeros((4, 4, 4))
    img[0] = 2
    with pytest.raises(ValueError, match="Input image must only contain 0s and 1s"):
        tau.Solver(img)


def test_multiphase_rejects_negative_conductivity():
    with pytest.raises(ValueError):
        tau.MultiPhaseSolver(np.zeros([6, 6, 6]), {0: 1.0, 1: -0.1})


def test_multiphase_rejects_non_dict_and_non_int_labels():
    with pytest.raises(TypeError):
        tau.MultiPhaseSolver(np.zeros([6, 6, 6]), [1.0])
    with pytest.raises(TypeError, match="Phase label must be integer"):
        tau.MultiPhaseSolver(np.zeros([6, 6, 6]), {"a": 1.0})
    with pytest.raises(ValueError, match="finite"):
        tau.MultiPhaseSolver(np.zeros([6, 6, 6]), {0: float("inf")})


def test_multiphase_warns_on_missing_labels_then_needs_a_gpu():
    img = np.zeros([10, 10, 10])
    img[:, :, :5] = 1
    img[:, :, 5:] = 2
    Ds = {1: 1.0}
    with pytest.warns(UserWarning, match="assuming these phases are isolating."):
        try:
            tau.MultiPhaseSolver(img, Ds, device="cuda")
        except RuntimeError:
            pass  # no GPU in this container
    assert Ds[2] == 0.0


def test_cpu_device_is_refused():
    with pytest.raises(RuntimeError, match="HIP GPU only"):
        tau.Solver(np.ones((4, 4, 4)), device="cpu")


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error")
def test_missing_gpu_is_an_error_not_a_fallback():
    with pytest.raises(RuntimeError, match="needs a HIP GPU"):
        tau.Solver(np.ones((4, 4, 4)))


def test_small_volume_default_launch():
    """Volumes the tuner skips get the measured small-volume configurations:
    5 half-sweeps with 8 x 2 workgroups below 1.5 M voxels, 4 with 8 x 2
    below 3 M, then the library's default; every shape they name is one the
    sweep compiles for every kind (8 x 2 at every depth) and leaves rows to
    own (8 x 2 rows > 2 H)."""
    from taufactor_amd import autotune

    assert autotune.small_volume_default(100 ** 3) == (5, (8, 2, 0))
    assert autotune.small_volume_default(128 ** 3) == (4, (8, 2, 0))
    assert autotune.small_volume_default(160 ** 3) == (None, None)
    assert 160 ** 3 < autotune.MIN_VOXELS
    for v in (1, 10 ** 6, 2 * 10 ** 6, 5 * 10 ** 6):
        H, tile = autotune.small_volume_default(v)
        if tile is not None:
            assert tile[:2] == (8, 2) and 8 * 2 > 2 * H
