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
    # the periodic sweep flag: even Y and Z, code kinds only, no other flag bits
    per = _lib.SWEEP_PERIODIC_YZ
    for kind, Y, Z in [(_lib.KIND_SOLVER, 7, 8), (_lib.KIND_MULTIPHASE, 8, 9),
                       (_lib.KIND_FIELD, 8, 8), (_lib.KIND_ANISOTROPIC, 8, 8)]:
        st = lib.tf_sweep(1, 2, 3, None, 0, kind | per, 1, 4, Y, Z, 1.5, 0, 2, None, None)
        assert st < 0 and b"TF_SWEEP_PERIODIC_YZ" in lib.tf_last_error()
    st = lib.tf_sweep(1, 2, 3, None, 0, 0x200, 1, 4, 4, 4, 1.5, 0, 2, None, None)
    assert st < 0 and b"flags" in lib.tf_last_error()
    st = lib.tf_wrap_ghosts_from(None, 1, 1, 4, 4, 4, None)
    assert st < 0 and b"null pointer" in lib.tf_last_error()
    # set-up grids: 256 lanes per block and a block per 256 positions of each
    # padded x plane; 20M planes of 3 x 3 exceed the 2^32 lanes of a launch
    st = lib.tf_setup_solver(1, 2, 3, 4, 5, 6, 1, 20_000_000, 1, 1, 0, None)
    assert st < 0 and b"too long in x" in lib.tf_last_error()
    st = lib.tf_field_faces(1, 2, 1, 20_000_000, 1, 1, None)
    assert st < 0 and b"too long in x" in lib.tf_last_error()


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


def test_batch_struct_mirrors_match_c_layout():
    """The numpy mirrors of tf_volume / tf_volume_state have the field
    offsets and sizes C gives the header's structs."""
    c_int, c_ll, c_f, c_d = ctypes.c_int, ctypes.c_longlong, ctypes.c_float, ctypes.c_double

    class Volume(ctypes.Structure):
        _fields_ = [("offset", c_ll), ("code_offset", c_ll), ("X", c_int), ("Y", c_int),
                    ("Z", c_int), ("omega", c_f), ("sum_offset", c_int), ("tile_first", c_int),
                    ("spans", c_int), ("reserved", c_int), ("D_mean", c_d)]

    class State(ctypes.Structure):
        _fields_ = [("active", c_int), ("converged", c_int), ("iters", c_int), ("cur", c_int),
                    ("tau", c_d), ("old_tau", c_d), ("D_rel", c_d), ("rel_err", c_d)]

    for cs, dt in ((Volume, _lib.VOLUME_DTYPE), (State, _lib.STATE_DTYPE)):
        assert ctypes.sizeof(cs) == dt.itemsize
        for name, _ in cs._fields_:
            assert getattr(cs, name).offset == dt.fields[name][1], name
    text = open(os.path.join(ROOT, "include", "tauhip.h")).read()
    body = re.search(r"typedef struct tf_volume \{(.*?)\} tf_volume;", text, re.S).group(1)
    names = re.findall(r"\b(\w+);", re.sub(r"/\*.*?\*/", "", body, flags=re.S))
    assert names == [n for n, _ in Volume._fields_ if n not in ("X", "Y")]  # "int X, Y, Z;"


def test_layer_sum_workspace_size():
    """8 doubles per (volume, layer, 2048-position chunk): the concentration
    and flux partials of the four waves of a chunk."""
    lib = _lib.load()
    for B, X, Y, Z in [(1, 1, 1, 1), (2, 5, 7, 9), (1, 512, 512, 512), (3, 4, 64, 32), (1, 3, 64, 33)]:
        chunks = -(-(Y * Z) // 2048)
        assert lib.tf_layer_sums_workspace_bytes(B, X, Y, Z) == 8 * B * X * chunks * 8
    assert lib.tf_layer_sums_workspace_bytes(0, 1, 1, 1) < 0


def test_sweep_code_layout_size():
    lib = _lib.load()
    for X, Y, Z in [(1, 1, 1), (5, 7, 9), (512, 512, 512), (3, 4, 11)]:
        row = ((Z + 5 + 3) // 4) * 4
        assert row % 4 == 0 and row >= Z + 5
        assert lib.tf_sweep_codes_bytes(X, Y, Z) == (X + 2) * (Y + 2) * row
    assert lib.tf_sweep_codes_bytes(0, 1, 1) < 0
