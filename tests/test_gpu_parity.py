"""GPU parity: the HIP path (through libtauhip's C-ABI) against the CPU oracle.

Iterates are compared value for value (np.array_equal, which also treats
+0 and -0 as equal: the reference leaves inactive voxels at a signed zero
whose sign carries no information).  Per-layer float64 sums are compared to
the oracle's float64 sums of the same float32 values (only the summation
order differs: rtol 1e-12).  Full solves must stop at the same iteration
with the same tau and D_eff (both decide in float64; the public
arrays are float32, hence rtol 2e-7 and 1e-6)."""
import numpy as np
import pytest
import torch

import taufactor_amd as tau
from oracle.restate import OracleMultiPhaseSolver, OracleSolver

pytestmark = pytest.mark.gpu


def _binary(shape, p=0.55, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < p).astype(np.uint8)


def _labels(shape, n=3, seed=0):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, size=shape)


def _field(s):
    return s.field.cpu().numpy()


SHAPES = [(5, 7, 9), (33, 17, 70), (20, 21, 21), (2, 1, 3), (1, 6, 5)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("H", [1, 2, 3, 4, 5, 6])
def test_solver_iterates_bit_identical(shape, H):
    img = _binary(shape, seed=sum(shape) + H)
    k = 23
    s = tau.Solver(img)
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=k, verbose=False)
    o = OracleSolver(img).run(k)
    assert s.iter == k
    assert np.array_equal(_field(s), o.field.numpy())


@pytest.mark.parametrize("tile", [(8, 2, 1), (4, 4, 9), (12, 2, 5), (8, 4, 2), (16, 2, 3)])
@pytest.mark.parametrize("H", [2, 3, 4, 6])
def test_solver_iterates_with_small_tiles(tile, H):
    if tile[1] == 4 and H > 4:
        pytest.skip("four-row waves are compiled for n_half <= 4")
    img = _binary((19, 23, 300), seed=5)
    s = tau.Solver(img)
    s.half_sweeps_per_pass = H
    s.tile = tile
    s.solve(iter_limit=3, verbose=False)        # passes then start on colour 1
    s.solve(iter_limit=3 + 31, verbose=False)
    o = OracleSolver(img).run(34)
    assert np.array_equal(_field(s), o.field.numpy())


def test_solver_batch_iterates():
    img = np.stack([_binary((12, 10, 14), seed=i) for i in range(3)])
    s = tau.Solver(img)
    s.solve(iter_limit=41, verbose=False)
    o = OracleSolver(img).run(41)
    assert np.array_equal(_field(s), o.field.numpy())


def test_isolated_voxels_and_decay():
    img = np.zeros((6, 7, 8), dtype=np.uint8)
    img[2, 3, 3] = 1          # isolated conductive voxel: decays as (1-w)^k
    img[:, 1, 1] = 1          # straight channel
    img[0, 5, 5] = 1          # touches only the x=0 ghost
    s = tau.Solver(img)
    s.solve(iter_limit=57, verbose=False)
    o = OracleSolver(img).run(57)
    assert np.array_equal(_field(s), o.field.numpy())


@pytest.mark.parametrize("shape", [(5, 7, 9), (33, 17, 70), (20, 21, 21), (1, 6, 5)])
@pytest.mark.parametrize("H", [1, 2, 4, 6])
def test_multiphase_iterates_bit_identical(shape, H):
    img = _labels(shape, n=4, seed=sum(shape))
    Ds = {0: 0.0, 1: 1.0, 2: 0.1, 3: 2.5}
    s = tau.MultiPhaseSolver(img, dict(Ds))
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=27, verbose=False)
    o = OracleMultiPhaseSolver(img, dict(Ds)).run(27)
    assert np.array_equal(_field(s), o.field.numpy())


def test_multiphase_label_zero_conductive_and_odd_values():
    img = _labels((14, 15, 16), n=3, seed=9)
    Ds = {0: 0.7, 1: 0.0, 2: 1.0 / 3.0}
    s = tau.MultiPhaseSolver(img, dict(Ds))
    s.tile = (16, 2, 7)
    s.solve(iter_limit=33, verbose=False)
    o = OracleMultiPhaseSolver(img, dict(Ds)).run(33)
    assert np.array_equal(_field(s), o.field.numpy())


@pytest.mark.parametrize("shape", [(3, 4, 5), (2, 3, 12), (4, 2, 1)])
def test_sweep_code_layout(shape):
    """tf_pack_sweep_codes: byte z+3 of each row holds code z (times 8 for the
    Solver kind: the byte offset of its table entry), the rest 0."""
    img = _binary(shape, seed=4)
    s = tau.Solver(np.stack([img, 1 - img]))
    X, Y, Z = shape
    row = ((Z + 5 + 3) // 4) * 4
    got = s._sweep_codes.cpu().numpy().reshape(2, X + 2, Y + 2, row)
    want = np.zeros_like(got)
    want[..., 3:Z + 5] = s._codes.cpu().numpy() * 8
    assert np.array_equal(got, want)


def test_inplace_half_sweep_kernel_matches_oracle():
    from taufactor_amd import _lib

    img = _binary((9, 11, 13), seed=2)
    s = tau.Solver(img)
    lib = _lib.load()
    f = s.field
    for it in range(7):
        st = lib.tf_sweep_half_inplace(_lib.ptr(f), _lib.ptr(s._codes), None, 0, _lib.KIND_SOLVER,
                                       1, 9, 11, 13, s._w32, it & 1, _lib.stream_handle(s.device))
        _lib.check(st, "tf_sweep_half_inplace")
    o = OracleSolver(img).run(7)
    assert np.array_equal(f.cpu().numpy(), o.field.numpy())


@pytest.mark.parametrize("kind", ["solver", "multiphase"])
def test_layer_sums_match_float64_oracle(kind):
    shape = (17, 19, 23)
    if kind == "solver":
        img = _binary(shape, seed=3)
        s, o = tau.Solver(img), OracleSolver(img)
    else:
        img = _labels(shape, seed=3)
        Ds = {0: 0.0, 1: 1.0, 2: 0.25}
        s, o = tau.MultiPhaseSolver(img, dict(Ds)), OracleMultiPhaseSolver(img, dict(Ds))
    s.solve(iter_limit=60, verbose=False)
    o.run(60)
    flux, conc = s._layer_sums()
    vf = o.vertical_flux().to(torch.float64).sum((2, 3)).numpy()
    cs = o.field[:, 1:-1, 1:-1, 1:-1].to(torch.float64).sum((2, 3)).numpy()
    np.testing.assert_allclose(flux, vf, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(conc, cs, rtol=1e-12, atol=1e-12)
    f2, c2 = s._layer_sums()
    assert np.array_equal(flux, f2) and np.array_equal(conc, c2)  # deterministic


@pytest.mark.parametrize("seed", [0, 1])
def test_solver_full_solve_matches_oracle(seed):
    img = _binary((40, 32, 36), p=0.6, seed=seed)
    s = tau.Solver(img)
    s.solve(verbose=False)
    o = OracleSolver(img).solve()
    assert s.iter == o.iter and s.converged == o.converged
    assert np.array_equal(_field(s), o.field.numpy())
    np.testing.assert_allclose(s.tau, o.tau, rtol=2e-7)
    np.testing.assert_allclose(s.D_eff, o.D_eff, rtol=1e-6)


def test_multiphase_full_solve_matches_oracle():
    img = _labels((36, 30, 28), n=3, seed=4)
    Ds = {0: 0.0, 1: 1.0, 2: 0.1}
    s = tau.MultiPhaseSolver(img, dict(Ds))
    s.solve(verbose=False, conv_crit=1e-3)
    o = OracleMultiPhaseSolver(img, dict(Ds)).solve(conv_crit=1e-3)
    assert s.iter == o.iter
    assert np.array_equal(_field(s), o.field.numpy())
    np.testing.assert_allclose(s.tau, o.tau, rtol=2e-7)


def test_integration_stub_runs():
    """The ctypes binding shown to reference maintainers works as printed."""
    import os
    import re

    from taufactor_amd import _lib

    text = open(os.path.join(os.path.dirname(os.path.dirname(__file__)), "INTEGRATION.md")).read()
    code = re.search(r"```python\nimport ctypes, math, torch\n(.*?)```", text, re.S).group(0)
    code = code.strip("`").removeprefix("python\n").replace('"libtauhip.so"', repr(_lib.LIB_PATH))
    ns = {}
    exec(compile(code, "INTEGRATION.md", "exec"), ns)
    img = _binary((11, 9, 13), seed=8)
    w = 2 - np.pi / (1.5 * 11)
    st = ns["HipSolverState"](torch.from_numpy(img[None]).cuda(), w)
    st.iterate(0, 10)
    st.iterate(10, 7)
    o = OracleSolver(img).run(17)
    assert np.array_equal(st.f[st.cur].cpu().numpy(), o.field.numpy())
    flux, conc = st.layer_means()
    vf = o.vertical_flux().to(torch.float64).mean((2, 3)).numpy()
    np.testing.assert_allclose(flux, vf, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("H,Z", [(4, 120), (4, 176), (4, 177), (4, 240), (5, 168), (5, 169), (3, 7)])
@pytest.mark.parametrize("Y", [40, 60])
def test_solver_z_tiling_boundaries(H, Z, Y):
    """The register-blocked sweep covers z with wide tiles of TZ voxels and, when
    the rest fits, one narrow strip of at most TZN (TZ = 120, TZN = 56 at
    n_half = 4; 116 and 52 at 5): Z = TZ (wide only), TZ + TZN (narrow strip of
    TZN), TZ + TZN + 1 (two wide), 2 TZ, and small Z (narrow only).  With 16x2
    workgroups Y = 60 gives an odd count of y tiles, so the last narrow tile has
    an empty half."""
    img = _binary((9, Y, Z), seed=Z + Y + H)
    s = tau.Solver(img)
    s.half_sweeps_per_pass = H
    s.tile = (16, 2, 4)
    s.solve(iter_limit=2 * H + 1, verbose=False)
    o = OracleSolver(img).run(2 * H + 1)
    assert np.array_equal(_field(s), o.field.numpy())


@pytest.mark.parametrize("which", ["Solver", "PeriodicSolver", "PeriodicSolver_odd", "MultiPhaseSolver"])
def test_checks_with_sweep_launched_ahead(which):
    """At each convergence check the first sweeps of the next block are
    launched before the host's decision and dropped if the solve stops.
    Checks that do not stop (conv_crit 0), blocks cut by iter_limit, a solve
    resumed after a check, and a check that stops must all leave the
    oracle's field."""
    from oracle.restate import OraclePeriodicSolver

    # odd periodic sizes: the reference's iteration grows without bound
    # (same-colour neighbours across the boundary); at X = 6 the values stay
    # finite through 420 iterations, so they still compare exactly
    shape = (6, 9, 11) if which.endswith("_odd") else (20, 10, 12)
    if which == "MultiPhaseSolver":
        img = _labels(shape, n=3, seed=12)
        Ds = {0: 0.0, 1: 1.0, 2: 0.3}
        make = lambda: tau.MultiPhaseSolver(img, dict(Ds))  # noqa: E731
        oracle = lambda: OracleMultiPhaseSolver(img, dict(Ds))  # noqa: E731
    else:
        img = _binary(shape, p=0.7, seed=12)
        cls, ocls = ((tau.Solver, OracleSolver) if which == "Solver" else
                     (tau.PeriodicSolver, OraclePeriodicSolver))
        make, oracle = (lambda: cls(img)), (lambda: ocls(img))
    s = make()
    # small closed volumes keep a third buffer and run three sweeps ahead;
    # the periodic solvers run one
    assert (s._spare is None) == which.startswith("Periodic")
    s.half_sweeps_per_pass = 6
    s.solve(iter_limit=250, verbose=False, conv_crit=0.0)   # checks at 100, 200; stops at 250
    assert s.iter == 250 and not s.converged
    s.solve(iter_limit=420, verbose=False, conv_crit=0.0)   # resumes; checks at 300, 400
    # odd Y or Z (periodic): one ghost wrap per half-sweep, whose final
    # ghosts are a copy taken at another point; compare the interiors
    cut = (slice(None),) + (slice(1, -1),) * 3 if which.endswith("_odd") else (slice(None),)
    o = oracle().run(420)
    assert np.array_equal(_field(s)[cut], o.field.numpy()[cut])
    if which.endswith("_odd"):
        return
    # a solve that stops at a check returns the checked field
    s2 = make()
    s2.half_sweeps_per_pass = 5
    s2.solve(verbose=False)
    o2 = oracle().solve()
    assert s2.converged and s2.iter == o2.iter
    assert np.array_equal(_field(s2)[cut], o2.field.numpy()[cut])


def _chunked_sum(terms):
    """Layer sum in the library's fixed order: chunks of 2048 positions;
    thread t of 256 adds positions t + 256k in k order; each 64-lane wave
    sums in a shuffle-down tree; wave sums in wave order; chunk partials in
    chunk order from 0."""
    total = 0.0
    for c0 in range(0, len(terms), 2048):
        seg = np.zeros(2048)
        part = terms[c0:c0 + 2048]
        seg[:len(part)] = part
        acc = np.zeros(256)
        for k in range(8):
            acc = acc + seg[256 * k:256 * (k + 1)]
        w = acc.reshape(4, 64).copy()
        for off in (32, 16, 8, 4, 2, 1):
            w[:, :off] = w[:, :off] + w[:, off:2 * off]
        chunk = w[0, 0]
        for i in range(1, 4):
            chunk = chunk + w[i, 0]
        total = total + chunk
    return total


def _sums_both_ways(s):
    from taufactor_amd import _lib

    lib = _lib.load()
    B, X, Y, Z = s.batch_size, s.Nx, s.Ny, s.Nz
    hm = _lib.ptr(s._hm_lut) if s._hm_lut is not None else None
    out = []
    for ws in (False, True):
        flux = torch.full((B * X,), np.nan, dtype=torch.float64, device="cuda")
        conc = torch.full((B * X,), np.nan, dtype=torch.float64, device="cuda")
        args = (_lib.ptr(s.field), _lib.ptr(s._codes), hm, s._n_classes, s.kind, B, X, Y, Z,
                _lib.ptr(flux), _lib.ptr(conc))
        if ws:
            n = lib.tf_layer_sums_workspace_bytes(B, X, Y, Z)
            buf = torch.full((n // 8,), np.nan, dtype=torch.float64, device="cuda")
            st = lib.tf_layer_sums_ws(*args, _lib.ptr(buf), n, _lib.stream_handle(s.device))
        else:
            st = lib.tf_layer_sums(*args, _lib.stream_handle(s.device))
        _lib.check(st, "layer sums")
        w = X if s.kind == _lib.KIND_ELECTRODE else X - 1
        out.append((flux.cpu().numpy()[:B * w].reshape(B, w), conc.cpu().numpy().reshape(B, X)))
    return out


@pytest.mark.parametrize("shape", [(17, 19, 23), (5, 50, 90), (40, 64, 64), (70, 3, 700)])
def test_layer_sums_order_is_pinned(shape):
    """Per-layer and streaming sums are bit-identical to a numpy evaluation
    of the documented order (Solver kind, two volumes, several chunks and x
    slabs, the last chunk partial)."""
    img = np.stack([_binary(shape, seed=sum(shape)), _binary(shape, p=0.7, seed=1)])
    s = tau.Solver(img)
    s.solve(iter_limit=30, verbose=False)
    (f1, c1), (f2, c2) = _sums_both_ways(s)
    assert np.array_equal(f1, f2) and np.array_equal(c1, c2)
    f = s.field.cpu().numpy()
    k = s._codes.cpu().numpy()
    X = shape[0]
    for b in range(2):
        for x in range(X):
            c = f[b, x + 1, 1:-1, 1:-1].ravel()
            assert c1[b, x] == _chunked_sum(c.astype(np.float64))
            if x < X - 1:
                cn = f[b, x + 2, 1:-1, 1:-1].ravel()
                d = (cn - c).astype(np.float32)
                d[(k[b, x + 1, 1:-1, 1:-1].ravel() == 0) | (k[b, x + 2, 1:-1, 1:-1].ravel() == 0)] = 0
                assert f1[b, x] == _chunked_sum(d.astype(np.float64))


@pytest.mark.parametrize("which", ["multiphase", "electrode", "field", "anisotropic", "periodic"])
def test_streaming_layer_sums_match_per_layer_sums(which):
    shape = (37, 41, 57)
    if which == "multiphase":
        s = tau.MultiPhaseSolver(_labels(shape, n=4, seed=3), {0: 0.0, 1: 1.0, 2: 0.25, 3: 3.0})
    elif which == "electrode":
        rng = np.random.default_rng(5)
        s = tau.ElectrodeSolver(rng.choice([0, 1, 5], size=shape, p=[0.3, 0.55, 0.15]))
    elif which == "field":
        rng = np.random.default_rng(6)
        s = tau.FieldSolver((rng.random(shape) * (rng.random(shape) < 0.8)).astype(np.float32))
    elif which == "anisotropic":
        s = tau.AnisotropicSolver(_binary(shape, p=0.6, seed=7), spacing=(1.0, 1.5, 0.7))
    else:
        s = tau.PeriodicSolver(_binary((36, 40, 58), p=0.6, seed=8))
    s.solve(iter_limit=40, verbose=False)
    (f1, c1), (f2, c2) = _sums_both_ways(s)
    assert np.array_equal(f1, f2) and np.array_equal(c1, c2)
    assert np.isfinite(f1).all() and np.isfinite(c1).all()


@pytest.mark.parametrize("H", [1, 3, 6])
def test_tiny_and_subnormal_sums_take_the_ieee_division(H):
    """A field of tiny normal and subnormal values (and zeros, negatives and a
    few ordinary values): most neighbour sums fall below 2^-100, where the
    Solver update leaves the reciprocal product for IEEE division.  The
    iterates must still equal the oracle's, subnormals included."""
    img = _binary((12, 10, 14), p=0.8, seed=21)
    rng = np.random.default_rng(5)
    vals = np.array([0.0, 2.0 ** -110, -(2.0 ** -112), 3e-40, -1e-42, 7e-45, 2.0 ** -101,
                     -(2.0 ** -126), 0.5], dtype=np.float32)
    s = tau.Solver(img)
    o = OracleSolver(img)
    f = vals[rng.integers(0, len(vals), size=tuple(o.field.shape))].astype(np.float32)
    bad = f.copy()
    f[:, :, [0, -1], :] = 0  # closed y/z boundaries: zero ghosts, as the set-up writes
    f[:, :, :, [0, -1]] = 0
    with pytest.raises(ValueError):
        s.field = torch.from_numpy(bad)
    s.field = torch.from_numpy(f)
    o.field = torch.from_numpy(f.copy())
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=2 * H + 1, verbose=False)
    o.run(2 * H + 1)
    got, want = _field(s), o.field.numpy()
    assert np.array_equal(got, want)
    assert np.count_nonzero((np.abs(want) < 2.0 ** -126) & (want != 0)) > 0  # subnormals present


def test_hbm_copy_kernel():
    """tf_hbm_copy (the bandwidth calibration kernel of bench.py) copies every
    byte, also when the size is not a multiple of a workgroup's 16 KiB, writes
    nothing past the end, and refuses sizes that are not a multiple of 16."""
    from taufactor_amd import _lib

    lib = _lib.load()
    stream = _lib.stream_handle(torch.device("cuda", 0))
    for n in (4, 4096, 4100, 3 * 2 ** 20 + 12):
        src = torch.randint(-2 ** 31, 2 ** 31 - 1, (n,), dtype=torch.int32, device="cuda")
        dst = torch.zeros(n + 64, dtype=torch.int32, device="cuda")
        assert lib.tf_hbm_copy(_lib.ptr(src), _lib.ptr(dst), 4 * n, stream) == 0
        torch.cuda.synchronize()
        assert torch.equal(dst[:n], src) and not torch.any(dst[n:])
    assert lib.tf_hbm_copy(_lib.ptr(src), _lib.ptr(dst), 20, stream) != 0


@pytest.mark.parametrize("which", ["Solver", "Solver_large", "MultiPhaseSolver", "FieldSolver", "PeriodicSolver",
                                   "PeriodicSolver_odd", "ElectrodeSolver"])
def test_reset_returns_to_the_initial_field(which):
    """reset() copies the initial field the set-up wrote back into the field
    buffers (and into the third buffer of small closed volumes): the field
    equals a fresh solver's bit for bit, and a second solve repeats the
    first.  Solver_large has no third buffer."""
    shape = {"PeriodicSolver_odd": (8, 9, 11), "Solver_large": (40, 300, 720)}.get(which, (16, 14, 18))
    if which == "MultiPhaseSolver":
        img = _labels(shape, n=3, seed=31)
        make = lambda: tau.MultiPhaseSolver(img, {0: 0.0, 1: 1.0, 2: 0.2})  # noqa: E731
    elif which == "FieldSolver":
        k = np.random.default_rng(31).uniform(0.2, 2.0, size=shape).astype(np.float32)
        make = lambda: tau.FieldSolver(k)  # noqa: E731
    elif which == "ElectrodeSolver":
        img = np.random.default_rng(31).choice([0, 1, 5], size=shape, p=[0.3, 0.55, 0.15])
        make = lambda: tau.ElectrodeSolver(img)  # noqa: E731
    else:
        img = _binary(shape, p=0.7, seed=31)
        cls = tau.PeriodicSolver if which.startswith("Periodic") else tau.Solver
        make = lambda: cls(img)  # noqa: E731
    s = make()
    assert (s._spare is not None) == (which in ("Solver", "MultiPhaseSolver", "FieldSolver", "ElectrodeSolver"))
    fresh = s.field.cpu().numpy().copy()
    s.solve(iter_limit=300, verbose=False)
    first = (s.iter, s.field.cpu().numpy().copy())
    s.reset()
    assert s.iter == 0 and np.array_equal(s.field.cpu().numpy(), fresh)
    s.solve(iter_limit=300, verbose=False)
    assert s.iter == first[0] and np.array_equal(s.field.cpu().numpy(), first[1])
