"""CPU: the tables of the MultiPhase face-code sweep (solver.face_code_tables,
TF_KIND_MULTIPHASE_FACES in include/tauhip.h).

* Every voxel's neighbourhood key of random labelled volumes (closed and
  periodic y/z, the first and last layers, X = 1) selects a prefactor pair
  whose f is the restated reference's factor, bit for bit
  (OracleMultiPhaseSolver.factor, taufactor/taufactor.py:595-604; its inf is
  pair 0), and whose face offsets select the oracle's face conductances
  D_x / D_y / D_z (:585-594).
* 1/f of every pair is the correctly rounded reciprocal (checked against the
  exact rational 1/f).
* The sweep's division sequence, reciprocal product plus one fma
  correction, equals IEEE division for every s in the domain at the
  exponents around the guard (tests/csrc/divcheck_pairs.c; a sample of the
  pairs here, TAUHIP_DIVCHECK_FULL=1 for every pair and every s).
* Diffusivity sets that do not fit (more than four distinct values with 0)
  give None, and the solver then keeps the class-code sweep."""
import os
import shutil
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from oracle.restate import OracleMultiPhaseSolver, OraclePeriodicMultiPhaseSolver
from taufactor_amd.solver import face_code_tables

HERE = os.path.dirname(os.path.abspath(__file__))


def _classes(img, Ds):
    labels = np.unique(img)
    d_cls = np.concatenate([[0.0], [Ds.get(int(l), 0.0) for l in labels]]).astype(np.float32)
    cls = (np.searchsorted(labels, img) + 1).astype(np.int64)
    return d_cls, cls


def _padded(cls, periodic):
    """Padded class codes as tf_setup_multiphase writes them."""
    p = np.pad(cls, 1)
    p[0], p[-1] = p[1], p[-2]
    if periodic:
        p[:, 0], p[:, -1] = p[:, -2], p[:, 1]
        p[:, :, 0], p[:, :, -1] = p[:, :, -2], p[:, :, 1]
    return p


@pytest.mark.parametrize("shape,periodic", [((9, 7, 8), False), ((6, 8, 10), True), ((1, 5, 6), False),
                                            ((2, 6, 4), True)])
@pytest.mark.parametrize("Ds", [{0: 0.0, 1: 1.0, 2: 0.1}, {1: 1.0, 2: 0.5, 3: 2.0}, {0: 0.0, 5: 0.3, 7: 0.3}])
def test_keys_select_the_oracle_factor(shape, periodic, Ds):
    rng = np.random.default_rng(sum(shape) + len(Ds))
    img = rng.choice(sorted(Ds), size=shape)
    d_cls, cls = _classes(img, Ds)
    fmap, keys, table = face_code_tables(d_cls)
    P = _padded(cls, periodic)
    fc = fmap[P].astype(np.int64)
    X, Y, Z = shape
    c = fc[1:-1, 1:-1, 1:-1]
    key = c
    for sl in ((slice(0, -2), slice(1, -1), slice(1, -1)), (slice(2, None), slice(1, -1), slice(1, -1)),
               (slice(1, -1), slice(0, -2), slice(1, -1)), (slice(1, -1), slice(2, None), slice(1, -1)),
               (slice(1, -1), slice(1, -1), slice(0, -2)), (slice(1, -1), slice(1, -1), slice(2, None))):
        key = key * 4 + fc[sl]
    x = np.arange(X)[:, None, None]
    key = key * 4 + (x == 0) + 2 * (x == X - 1)
    pair = keys[key].astype(np.int64) // 8
    f = table[16 + 2 * pair + 1]
    r = table[16 + 2 * pair]
    Oc = OraclePeriodicMultiPhaseSolver if periodic else OracleMultiPhaseSolver
    o = Oc(img, dict(Ds))
    ref = o.factor.numpy()[0]
    got = np.where(pair == 0, np.float32(np.inf), f)
    assert np.array_equal(got, ref)
    assert np.all((pair == 0) == np.isinf(ref))
    assert np.all(r[pair > 0] * f[pair > 0] > 0)
    # faces towards x+1, y+1, z+1 by their offsets in the 4 x 4 face table
    hm = table[:16]
    gx = hm[(fc[:-1, 1:-1, 1:-1] * 4 + fc[1:, 1:-1, 1:-1])]
    assert np.array_equal(gx, o.D_x.numpy()[0])
    if not periodic:
        gy = hm[(fc[1:-1, :-1, 1:-1] * 4 + fc[1:-1, 1:, 1:-1])]
        gz = hm[(fc[1:-1, 1:-1, :-1] * 4 + fc[1:-1, 1:-1, 1:])]
        assert np.array_equal(gy, o.D_y.numpy()[0])
        assert np.array_equal(gz, o.D_z.numpy()[0])


def _rn32(q: Fraction) -> np.float32:
    """q > 0 rounded to the nearest float32 (ties to even), exactly."""
    lo = np.float32(float(q))
    cands = [np.nextafter(lo, np.float32(0)), lo, np.nextafter(lo, np.float32(np.inf))]
    errs = [abs(Fraction(float(c)) - q) for c in cands]
    best = min(errs)
    ties = [c for c, e in zip(cands, errs) if e == best]
    return ties[0] if len(ties) == 1 else [c for c in ties if (c.view(np.uint32) & 1) == 0][0]


@pytest.mark.parametrize("Ds", [[0, 1, 0.1], [0, 1, 0.5, 2], [0, 1e-3, 1e3]])
def test_reciprocals_are_correctly_rounded(Ds):
    _, _, table = face_code_tables(np.array(Ds, np.float32))
    pairs = table[18:].reshape(-1, 2)
    assert len(pairs) > 10
    for r, f in pairs:
        assert r == _rn32(Fraction(1) / Fraction(float(f))), f


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_division_sequence_is_ieee_for_the_pairs(tmp_path):
    exe = tmp_path / "divcheck_pairs"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                    os.path.join(HERE, "csrc", "divcheck_pairs.c"), "-o", str(exe), "-lm"], check=True)
    full = os.environ.get("TAUHIP_DIVCHECK_FULL") == "1"
    lines = []
    for Ds in ([0, 0, 1, 0.1], [0, 1, 0.5, 2], [0, 1e-3, 1e3]):
        pairs = face_code_tables(np.array(Ds, np.float32))[2][18:].reshape(-1, 2)
        if not full:
            # the extremes and a seeded sample
            pick = np.unique(np.concatenate([[0, len(pairs) - 1],
                                             np.random.default_rng(len(pairs)).integers(0, len(pairs), 4)]))
            pairs = pairs[pick]
        lines += [f"{r.view(np.uint32):08x} {f.view(np.uint32):08x}" for r, f in pairs]
    res = subprocess.run([str(exe), "full" if full else "quick"], input="\n".join(lines), capture_output=True,
                         text=True, timeout=3600)
    assert res.returncode == 0, res.stdout[-2000:]
    assert f"pairs={len(lines)}" in res.stdout


def test_unfit_diffusivities_give_none():
    assert face_code_tables(np.array([0, 1, 2, 3, 4], np.float32)) is None
    assert face_code_tables(np.array([0, 1, 2, 3], np.float32)) is not None
    # a prefactor below 2^-20
    assert face_code_tables(np.array([0, 1e-8], np.float32)) is None


def test_sweep_codes_bytes_per_kind():
    """The codes size the host checks before every sweep launch (the
    autotuner's launches included): the layout each kind's kernel addresses."""
    from taufactor_amd import _lib
    from taufactor_amd.solver import sweep_codes_bytes

    lib = _lib.load()
    B, X, Y, Z = 2, 5, 6, 7
    padded = B * (X + 2) * (Y + 2) * (Z + 2)
    row = ((Z + 5 + 3) // 4) * 4
    assert sweep_codes_bytes(lib, _lib.KIND_MULTIPHASE, B, X, Y, Z) == B * (X + 2) * (Y + 2) * row
    assert sweep_codes_bytes(lib, _lib.KIND_SOLVER, B, X, Y, Z) == B * (X + 2) * (Y + 2) * row
    assert sweep_codes_bytes(lib, _lib.KIND_MULTIPHASE_FACES, B, X, Y, Z) == 4 * padded
    assert sweep_codes_bytes(lib, _lib.KIND_FIELD, B, X, Y, Z) == 12 * padded
    # the padded class codes are smaller than the class sweep's layout: a
    # launch handed them instead would read past their end
    assert padded < sweep_codes_bytes(lib, _lib.KIND_MULTIPHASE, B, X, Y, Z)
