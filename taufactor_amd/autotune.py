"""Launch-configuration autotuner for the fused sweep.

Every launch configuration (fused half-sweeps per HBM pass H, workgroup
shape, x-tile length) gives identical iterates, so the choice is purely a
speed question.  The tuner times a few candidates on the solver's own volume
and keeps the one with the lowest time per reference iteration:

* plain solvers (Solver, MultiPhaseSolver, FieldSolver): a few launches of
  each candidate on scratch copies of the current field, so the solver's
  state is untouched;
* BatchSolver: a few passes of each candidate over the whole packed batch,
  after which the batch is re-packed (a depth that does not divide 100 ends
  each period with one shorter pass, not timed here).

Results are cached per (kind, batch, shape) for the life of the process.
Volumes below MIN_VOXELS are not tuned: the defaults (H = 4, library tile)
serve them and tuning would cost more than it saves.
"""
from __future__ import annotations

import torch

from . import _lib

MIN_VOXELS = 1 << 23
# even, so that a candidate with an odd H is timed on as many launches of
# each starting colour
TIMED_LAUNCHES = 6

# (H values, workgroup shapes, x-tile lengths), in order of preference: the
# first candidate within TIE of the fastest wins, which keeps the choice
# stable against timing noise.  x tile 0 is the library's occupancy-balanced
# length for the shape and volume.
SOLVER_SPACE = ((5, 4, 6, 3, 2), ((16, 2), (12, 2), (8, 2)), (0, 128, 64))
FIELD_SPACE = ((4, 3, 2), ((8, 2), (12, 2), (16, 2)), (0, 128, 64))
BATCH_SPACE = ((5, 4, 6, 2), ((8, 2), (4, 4), (12, 2), (16, 2)), (64, 32))
TIE = 0.02

_cache: dict = {}
# (ms per half-sweep, H, tile) of every candidate of the latest tuning run
last_timings: list = []


def small_volume_default(voxels: int):
    """(H, tile) for a volume the tuner skips (fewer than MIN_VOXELS voxels):
    smaller workgroups put more tiles on the 256 CUs, where a few large
    tiles would leave most of them idle.  Measured on MI355X (whole solves,
    Solver / MultiPhase / per-voxel field): below 1.5 M voxels 5 half-sweeps
    with 8 x 2 workgroups (100^3: 2.63 -> 1.72 ms Solver, 3.36 -> 2.23
    MultiPhase, 1.42 -> 1.23 field); up to 3 M voxels 4 with 8 x 2 (128^3:
    2.61 -> 1.92, 3.46 -> 2.56, 1.48 -> 1.50); above that the library's
    default (4 half-sweeps, 16 x 2) was as fast as either.  (None, None)
    leaves the library's default."""
    if voxels < 3 * (1 << 19):
        return (5, (8, 2, 0))
    if voxels < 3 * (1 << 20):
        return (4, (8, 2, 0))
    return (None, None)


def _candidates(space, fixed_h, fixed_tile):
    hs, shapes, txs = space
    if fixed_h is not None:
        hs = (int(fixed_h),)
    tiles = [(w, r, tx) for (w, r) in shapes for tx in txs]
    if fixed_tile is not None:
        tiles = [tuple(fixed_tile)]
    return [(h, t) for h in hs for t in tiles]


def _time(launch, n):
    """Mean ms of n calls of launch() (after one untimed call), or None if a
    call reports an error (e.g. a shape that needs too much LDS)."""
    if launch() != 0:
        return None
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(n):
        if launch() != 0:
            return None
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / n


def tune_solver(s, fixed_h=None, fixed_tile=None):
    """Best (H, tile) for a plain solver `s` (SORSolver subclass)."""
    B, X, Y, Z = s.batch_size, s.Nx, s.Ny, s.Nz
    # the kind, codes and table the solver's own launches use (MultiPhase
    # volumes may sweep through face codes instead of class codes)
    kind, codes, hm, n_table = s._sweep_pointers()
    key = ("plain", kind, n_table, B, X, Y, Z, fixed_h, None if fixed_tile is None else tuple(fixed_tile))
    if key in _cache:
        return _cache[key]
    lib = s._lib
    space = FIELD_SPACE if s.kind == _lib.KIND_FIELD else SOLVER_SPACE
    shape = tuple(s.field.shape)
    src = _lib.slack_buffer(shape, torch.float32, s.device)
    dst = _lib.slack_buffer(shape, torch.float32, s.device)
    src.copy_(s.field)
    dst.copy_(s.field)
    stream = s._stream()
    timings = []
    for H, tile in _candidates(space, fixed_h, fixed_tile):
        t = _lib.Tile(*tile)
        bufs = [src, dst]
        state = {"i": 0}

        def launch():
            # launch i starts at half-sweep i*H, so with H odd the starting
            # colour alternates, as in a solve (the two colours are separate
            # code paths of the kernel and need not run equally fast)
            i = state["i"]
            state["i"] += 1
            return lib.tf_sweep(_lib.ptr(bufs[i & 1]), _lib.ptr(bufs[1 - (i & 1)]), codes, hm,
                                n_table, kind, B, X, Y, Z, s._w32, (i * H) & 1, H, t, stream)

        ms = _time(launch, TIMED_LAUNCHES)
        if ms is not None:
            timings.append((ms / H, H, tile))
    del src, dst
    last_timings[:] = timings
    result = _pick(timings)
    _cache[key] = result
    return result


def _pick(timings):
    """First candidate (in preference order) within TIE of the fastest."""
    if not timings:
        return (None, None)
    fastest = min(t for t, _, _ in timings)
    for t, H, tile in timings:
        if t <= fastest * (1 + TIE):
            return (H, tile)


def tune_batch(b, fixed_h=None, fixed_tile=None):
    """Best (H, tile) for a BatchSolver `b`; re-packs the batch afterwards."""
    key = ("batch", tuple(b.shapes), fixed_h, None if fixed_tile is None else tuple(fixed_tile))
    if key in _cache:
        return _cache[key]
    timings = []
    try:
        for H, tile in _candidates(BATCH_SPACE, fixed_h, fixed_tile):
            b._make_plan(H, tile)

            def launch():
                # the full pass over every tile (a queue pass would find its
                # counter spent after the first of the timed launches)
                return b._pass_status(0, 0, H)

            ms = _time(launch, TIMED_LAUNCHES)
            if ms is not None:
                timings.append((ms / H, H, tile))
    finally:
        b._plan = None
        b.reset()
    last_timings[:] = timings
    result = _pick(timings)
    _cache[key] = result
    return result
