"""GPU: tf_layer_sums_noalloc (partial sums in the library's static device
array, no allocation) against tf_layer_sums_ws (caller's workspace) through
the C-ABI: identical float64 sums, bit for bit, on two calls in a row, the
second on a smaller volume (the array keeps the first call's partials, which
the second must not see), for the Solver and the MultiPhase kinds; sizes
beyond the array are refused before anything is launched."""
import ctypes

import numpy as np
import pytest
import torch

import taufactor_amd as tau
from taufactor_amd import _lib

pytestmark = pytest.mark.gpu


def _sums(s, noalloc: bool):
    lib = _lib.load()
    B, X, Y, Z = s.batch_size, s.Nx, s.Ny, s.Nz
    flux = torch.full((B * X,), np.nan, dtype=torch.float64, device=s.device)
    conc = torch.full((B * X,), np.nan, dtype=torch.float64, device=s.device)
    hm = _lib.ptr(s._hm_lut) if s._hm_lut is not None else None
    common = (_lib.ptr(s.field), _lib.ptr(s._codes), hm, s._n_classes, s.kind, B, X, Y, Z,
              _lib.ptr(flux), _lib.ptr(conc))
    stream = _lib.stream_handle(s.device)
    if noalloc:
        st = lib.tf_layer_sums_noalloc(*common, stream)
        _lib.check(st, "tf_layer_sums_noalloc")
    else:
        nbytes = lib.tf_layer_sums_workspace_bytes(B, X, Y, Z)
        assert nbytes > 0
        ws = torch.empty(nbytes // 8, dtype=torch.float64, device=s.device)
        st = lib.tf_layer_sums_ws(*common, _lib.ptr(ws), nbytes, stream)
        _lib.check(st, "tf_layer_sums_ws")
    torch.cuda.synchronize(s.device)
    n_flux = B * (X - 1)
    return flux[:n_flux].cpu().numpy(), conc.cpu().numpy()


def _solver(shape, seed, multiphase=False):
    rng = np.random.default_rng(seed)
    if multiphase:
        img = rng.integers(0, 3, size=shape)
        s = tau.MultiPhaseSolver(img, {0: 0.0, 1: 1.0, 2: 0.25}, device="cuda:0")
    else:
        img = (rng.random(shape) < 0.6).astype(np.uint8)
        s = tau.Solver(img, device="cuda:0")
    s.solve(iter_limit=30, verbose=False)  # a field that is not the initial profile
    return s


@pytest.mark.parametrize("multiphase", [False, True])
def test_noalloc_equals_workspace_two_calls(multiphase):
    big = _solver((2, 150, 130, 120), 3, multiphase)
    for s in (big, _solver((1, 40, 36, 30), 4, multiphase)):
        f_ws, c_ws = _sums(s, noalloc=False)
        f_na, c_na = _sums(s, noalloc=True)
        print(f"shape {tuple(s.field.shape)} kind {s.kind}: {f_ws.size} flux sums")
        assert np.isfinite(c_na).all() and np.isfinite(f_na).all()
        assert np.array_equal(f_ws.view(np.int64), f_na.view(np.int64))
        assert np.array_equal(c_ws.view(np.int64), c_na.view(np.int64))
    # the large volume again, after the small one left its partials in the array
    f1, c1 = _sums(big, noalloc=True)
    f2, c2 = _sums(big, noalloc=False)
    assert np.array_equal(f1.view(np.int64), f2.view(np.int64))
    assert np.array_equal(c1.view(np.int64), c2.view(np.int64))


def test_noalloc_rejects_sizes_beyond_the_array():
    lib = _lib.load()
    # 64 volumes of 512 x 256 x 256: 64 MiB of partial sums, each volume
    # within the 2^31-voxel limit
    B, X, Y, Z = 64, 512, 256, 256
    need = lib.tf_layer_sums_workspace_bytes(B, X, Y, Z)
    assert need > _lib.NOALLOC_SCRATCH_BYTES, need
    dummy = torch.zeros(16, dtype=torch.float64, device="cuda:0")
    p = _lib.ptr(dummy)
    st = lib.tf_layer_sums_noalloc(p, p, None, 0, _lib.KIND_SOLVER, B, X, Y, Z, p, p,
                                   _lib.stream_handle(dummy.device))
    assert st == -1
    msg = lib.tf_last_error().decode()
    assert "static scratch array" in msg and str(need) in msg, msg
    # a size that fits is not refused by the size check
    assert lib.tf_layer_sums_workspace_bytes(1, 512, 512, 512) <= _lib.NOALLOC_SCRATCH_BYTES
