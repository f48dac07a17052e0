"""ctypes binding of libtauhip.so (the C-ABI declared in include/tauhip.h).

The library is loaded once, from the package directory.  There is no fallback:
if it is missing or stale, every solver call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtauhip.so")
ABI_VERSION = 12

KIND_SOLVER = 0
KIND_MULTIPHASE = 1
KIND_FIELD = 2
KIND_ANISOTROPIC = 3
ANISO_TABLE = 147
KIND_ELECTRODE = 4
KIND_MULTIPHASE_FACES = 5
MPF_MAX_CLASSES = 4
MPF_FACE_SLOTS = 16
MPF_MAX_FACTORS = 2047
SWEEP_PERIODIC_YZ = 0x100  # flag OR-ed into the kind of tf_sweep
ELECTRODE_TABLE = 64
SETUP_PERIODIC_YZ = 1
SETUP_ANISOTROPIC = 2
MAX_FUSED_HALF_SWEEPS = 6
# size of the static partial-sum array of tf_layer_sums_noalloc (TF_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 16 << 20
# readable bytes the sweep needs before and after each field buffer (TF_FIELD_SLACK_BYTES)
FIELD_SLACK_BYTES = 16

_c_int = ctypes.c_int
_c_float = ctypes.c_float
_vp = ctypes.c_void_p


class Tile(ctypes.Structure):
    _fields_ = [("waves", ctypes.c_int), ("rows_per_wave", ctypes.c_int), ("tx", ctypes.c_int)]


# name -> (restype, argtypes); the exported surface of include/tauhip.h
SIGNATURES = {
    "tf_abi_version": (_c_int, []),
    "tf_last_error": (ctypes.c_char_p, []),
    "tf_setup_solver": (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int,
                                 _vp]),
    "tf_setup_multiphase": (_c_int, [_vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp,
                                     _c_int, _c_int, _c_int, _c_int, _c_int, _vp]),
    "tf_setup_electrode": (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int,
                                    _c_int, _vp]),
    "tf_wrap_ghosts": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp]),
    "tf_wrap_ghosts_from": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp]),
    "tf_sweep_codes_bytes": (ctypes.c_longlong, [_c_int, _c_int, _c_int]),
    "tf_pack_sweep_codes": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp]),
    "tf_multiphase_face_codes": (_c_int, [_vp, _vp, _c_int, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp]),
    "tf_sweep": (_c_int, [_vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                          _c_float, _c_int, _c_int, ctypes.POINTER(Tile), _vp]),
    "tf_sweep_half_inplace": (_c_int, [_vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int,
                                       _c_int, _c_float, _c_int, _vp]),
    "tf_layer_sums": (_c_int, [_vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                               _vp, _vp, _vp]),
    "tf_layer_sums_workspace_bytes": (ctypes.c_longlong, [_c_int, _c_int, _c_int, _c_int]),
    "tf_layer_sums_ws": (_c_int, [_vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                  _vp, _vp, _vp, ctypes.c_longlong, _vp]),
    "tf_layer_sums_noalloc": (_c_int, [_vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                       _vp, _vp, _vp]),
    "tf_spans_x": (_c_int, [_vp, _c_int, _c_int, _c_int]),
    "tf_setup_field": (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp]),
    "tf_field_faces": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp]),
    "tf_grad_field": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp]),
    "tf_grad_classes": (_c_int, [_vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp]),
    "tf_sweep_launch_config": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                        ctypes.POINTER(Tile), ctypes.POINTER(Tile),
                                        ctypes.POINTER(_c_int)]),
    "tf_sweep_tile_count": (_c_int, [_c_int, _c_int, _c_int, _c_int, ctypes.POINTER(Tile),
                                     ctypes.POINTER(ctypes.c_int)]),
    "tf_batch_sweep": (_c_int, [_vp, _vp, _vp, _vp, _vp, _c_int, _vp, _c_int, _c_int, _c_int,
                                ctypes.POINTER(Tile), _vp]),
    "tf_batch_compact": (_c_int, [_vp, _c_int, _vp, _vp, _vp, _c_int, _vp]),
    "tf_batch_sweep_queue": (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_int, _c_int,
                                      _c_int, ctypes.POINTER(Tile), _vp]),
    "tf_batch_layer_sums": (_c_int, [_vp, _vp, _vp, _vp, _vp, _c_int, _vp, _c_int, _vp, _vp, _vp]),
    "tf_batch_check": (_c_int, [_vp, _c_int, _vp, _vp, _c_int, _c_int, _c_int, ctypes.c_double, _vp]),
    "tf_hbm_copy": (_c_int, [_vp, _vp, ctypes.c_longlong, _vp]),
}

# numpy mirrors of the batch structs (layout checked in tests/test_abi.py)
import numpy as _np  # noqa: E402

VOLUME_DTYPE = _np.dtype([("offset", _np.int64), ("code_offset", _np.int64), ("X", _np.int32), ("Y", _np.int32), ("Z", _np.int32),
                          ("omega", _np.float32), ("sum_offset", _np.int32), ("tile_first", _np.int32),
                          ("spans", _np.int32), ("reserved", _np.int32), ("D_mean", _np.float64)])
STATE_DTYPE = _np.dtype([("active", _np.int32), ("converged", _np.int32), ("iters", _np.int32),
                         ("cur", _np.int32), ("tau", _np.float64), ("old_tau", _np.float64),
                         ("D_rel", _np.float64), ("rel_err", _np.float64)])

_lock = threading.Lock()
_lib = None


class TauHipError(RuntimeError):
    """A non-zero status from libtauhip."""


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libtauhip.so and declare its signatures (cached)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise TauHipError(
                f"{path} is missing: build it with `python -m taufactor_amd.csrc.build` "
                "(or __graft_entry__.build()); taufactor_amd has no non-HIP code path")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        ver = lib.tf_abi_version()
        if ver != ABI_VERSION:
            raise TauHipError(f"{path} has ABI version {ver}, expected {ABI_VERSION}; rebuild it")
        _lib = lib
        return lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = load().tf_last_error()
        msg = msg.decode(errors="replace") if msg else ""
        raise TauHipError(f"{what} failed with status {status}: {msg}")


def ptr(t) -> int:
    """Device (or host) address of a contiguous torch tensor / numpy array."""
    if hasattr(t, "data_ptr"):
        if not t.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return t.data_ptr()
    return t.ctypes.data


def slack_buffer(shape, dtype, device):
    """A contiguous tensor of `shape` inside a zero-filled allocation with
    FIELD_SLACK_BYTES before and after it (what tf_sweep needs of field
    buffers and of the per-voxel diffusivity)."""
    import torch
    itemsize = torch.empty((), dtype=dtype).element_size()
    pad = FIELD_SLACK_BYTES // itemsize
    n = 1
    for d in shape:
        n *= int(d)
    flat = torch.zeros(n + 2 * pad, dtype=dtype, device=device)
    return flat[pad:pad + n].view(*shape)


def sweep_code_buffer(B, X, Y, Z, device):
    """Uninitialised uint8 buffer for B volumes of the sweep-code layout."""
    import torch
    n = load().tf_sweep_codes_bytes(X, Y, Z)
    if n < 0:
        check(int(n), "tf_sweep_codes_bytes")
    return torch.empty(B * int(n), dtype=torch.uint8, device=device)


def stream_handle(device) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream
