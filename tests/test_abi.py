"""CPU: libtauhip.so loads and exports every function include/tauhip.h
declares, with the ABI version the header states.  No compute calls."""
import ctypes
import os
import re

from taufactor_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "tauhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tf_\w+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert "tf_sweep" in names and "tf_layer_sums" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_loader_declares_every_header_symbol():
    assert sorted(_lib.SIGNATURES) == header_functions()


def test_abi_version_matches_header():
    text = open(os.path.join(ROOT, "include", "tauhip.h")).read()
    ver = int(re.search(r"#define TF_ABI_VERSION (\d+)", text).group(1))
    lib = _lib.load()
    assert lib.tf_abi_version() == ver == _lib.ABI_VERSION


def test_argument_errors_are_reported_without_a_gpu():
    lib = _lib.load()
    st = lib.tf_sweep(None, None, None, None, 0, 0, 1, 4, 4, 4, 1.5, 0, 2, None, None)
    assert st < 0
    assert b"null pointer" in lib.tf_last_error()
    st = lib.tf_sweep(1, 1, 1, None, 0, 0, 1, 4, 4, 4, 1.5, 0, 2, None, None)
    assert st < 0 and b"distinct" in lib.tf_last_error()
    st = lib.tf_sweep(1, 2, 3, None, 0, 0, 1, 4, 4, 4, 1.5, 0, 99, None, None)
    assert st < 0 and b"n_half" in lib.tf_last_error()


def test_spans_x_host_entry_point():
    import numpy as np

    lib = _lib.load()
    m = np.zeros((5, 3, 3), dtype=np.uint8)
    m[:, 1, 1] = 1
    assert lib.tf_spans_x(m.ctypes.data, 5, 3, 3) == 1
    m[2, 1, 1] = 0
    assert lib.tf_spans_x(m.ctypes.data, 5, 3, 3) == 0
    rng = np.random.default_rng(11)
    from oracle.restate import spans_x

    for _ in range(100):
        shape = tuple(int(v) for v in rng.integers(1, 8, size=3))
        mm = np.ascontiguousarray(rng.random(shape) < rng.uniform(0.2, 0.7), dtype=np.uint8)
        assert lib.tf_spans_x(mm.ctypes.data, *shape) == int(spans_x(mm.astype(bool)))
