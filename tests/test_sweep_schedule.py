"""CPU: the register-blocked sweep's schedule, restated in numpy and run
against the oracle.

The restatement follows rsweep_tile (taufactor_amd/csrc/tauhip_sweep.h) step
by step: tiles of NW waves x RPW consecutive rows (TY = NW*RPW - 2H owned) by a
128-voxel z window (TZ owned) or, for the last strip of z when it fits, a
narrow tile whose two half-waves hold 64-voxel windows of two y tiles, or, for
TZ < Z <= 128 with closed boundaries, one whole-z window from z = 1; each
lane keeps its z-pair of every row in a ring of NR planes; half-sweep h of step
s updates plane s-h; z neighbours across lanes come from the adjacent lane
(the DPP wave shifts; 0 past the ends), the y neighbours of a wave's first and
last row from the edge values the neighbouring waves published at step s-1,
double-buffered by step parity; rows and z are updated unconditionally, planes
only inside each half-sweep's range.  With wrap (periodic y/z) every row
and z position loads the interior voxel it repeats.  A wrong ring slot, parity, publication
step or narrow-tile row offset changes the stored field, which must equal H
reference iterations of the oracle (Solver kind, IEEE division in place of the
exact reciprocal sequence, which gives the same quotient).
"""
import numpy as np
import pytest

from oracle.restate import OracleSolver

L = 64


def _wrap_at(p, n):
    return (p - 1) % n + 1


def _tile_sweep(f0, codes, H, NW, RPW, TX, colour0, w, bug=None, wrap=False):
    X, Y, Z = (n - 2 for n in f0.shape)
    sX, sY = (Y + 2) * (Z + 2), Z + 2
    flat = np.concatenate([np.zeros(4, np.float32), f0.ravel(), np.zeros(4, np.float32)])
    cflat = codes.ravel()
    dst = f0.copy()
    EY = NW * RPW
    TY = EY - 2 * H
    ML = H + (H & 1)
    TZ = ((128 - ML - H) // 2) * 2
    TZN = ((64 - ML - H) // 2) * 2
    NR = 8 if H == 5 else ((H + 5) // 2) * 2
    D = NR - H - 2
    nTy, nTx = -(-Y // TY), -(-X // TX)
    wide, rest = Z // TZ, Z % TZ
    wz = not wrap and TZ < Z <= 128
    tiles = [(1 + tz * TZ, ty, False) for tz in range(wide + (rest > TZN)) for ty in range(nTy)]
    if 0 < rest <= TZN:
        tiles += [(1 + wide * TZ, ty, True) for ty in range(-(-nTy // 2))]
    if wz:
        tiles = [(1, ty, False) for ty in range(nTy)]
    d_of = np.array([np.inf] + list(range(1, 9)), np.float32)
    lanes = np.arange(L)
    for tx in range(nTx):
        for z0, ty, nar in tiles:
            half = (lanes >> 5) if nar else np.zeros(L, int)
            y0 = 1 + ((2 * ty + half) if nar else ty) * TY
            x0 = 1 + tx * TX
            z1, x1 = min(z0 + (TZN if nar else TZ), Z + 1), min(x0 + TX, X + 1)
            ew = z0 - ML
            if wz:  # the window starts at z = 1; the DPP zero fill stands in for the ghosts
                z1, ew = Z + 1, 1
            zl = ew + 2 * (lanes & 31 if nar else lanes)
            lx1 = min(x1 + H, X + 2)
            F = np.zeros((NR, NW, RPW, L, 2), np.float32)
            K = np.zeros((NR, NW, RPW, L, 2), np.int64)
            E = np.zeros((H, 2, NW + 2, 2, L), np.float32)

            def rows(wv, j):
                return y0 - H + wv * RPW + j

            def load(p, sl):
                for wv in range(NW):
                    for j in range(RPW):
                        y = rows(wv, j)
                        for e in range(2):
                            z = zl + e
                            ok = (0 <= p < lx1) & (y >= 0) & (y <= Y + 1)
                            if wrap:  # periodic y/z: the interior voxel repeated
                                y, z, ok = _wrap_at(rows(wv, j), Y), _wrap_at(zl + e, Z), 0 <= p < lx1
                            idx = 4 + p * sX + y * sY + z
                            inb = ok & (idx >= 0) & (idx < len(flat))
                            F[sl, wv, j, :, e] = np.where(inb, flat[np.clip(idx, 0, len(flat) - 1)], 0)
                            cok = ok & (z >= 0) & (z <= Z + 1)
                            K[sl, wv, j, :, e] = np.where(cok, cflat[np.clip(p * sX + y * sY + z, 0, len(cflat) - 1)], 0)

            plo = [max(x0 - (H - 1 - h), 1) for h in range(H)]
            pn = [min(x1 + (H - 1 - h), X + 1) - plo[h] for h in range(H)]
            S0, S1 = x0 - (H - 1), x1 - 1 + H - 1
            A = (colour0 & 1) ^ (0 if bug == "colour" else 1)

            def act(u, j):
                return (A ^ (u + j)) & 1

            def par(u):
                return (H + u) & 1

            def publish(h, u, sl):
                for wv in range(NW):
                    E[h, par(u), wv + 1, 0] = F[sl, wv, 0, :, act(u, 0)]
                    E[h, par(u), wv + 1, 1] = F[sl, wv, RPW - 1, :, act(u, RPW - 1)]

            for i in range(D + 3):
                load(S0 - 1 + i, i)
            publish(0, NR - 1, 1)
            s0 = S0
            while s0 <= S1:
                for u in range(NR):
                    s = s0 + u
                    for h in range(H):
                        p = s - h
                        sl = (u - h + 1) % NR
                        if 0 <= p - plo[h] < pn[h]:
                            Eb = E[h, par(u if bug == "edge_parity" else u + 1)]
                            for wv in range(NW):
                                for j in range(RPW):
                                    a = act(u, j)
                                    c = F[sl, wv, j, :, a]
                                    xm = F[(sl - (2 if bug == "x_slot" else 1)) % NR, wv, j, :, a]
                                    xp = F[(sl + 1) % NR, wv, j, :, a]
                                    ym = F[sl, wv, j - 1, :, a] if j > 0 else Eb[wv, 1]
                                    yp = F[sl, wv, j + 1, :, a] if j < RPW - 1 else Eb[wv + 2, 0]
                                    if a == 0:
                                        zp = F[sl, wv, j, :, 1]
                                        zm = np.concatenate([[0], F[sl, wv, j, :-1, 1]]).astype(np.float32)
                                    else:
                                        zm = F[sl, wv, j, :, 0]
                                        zp = np.concatenate([F[sl, wv, j, 1:, 0], [0]]).astype(np.float32)
                                    code = K[sl, wv, j, :, a]
                                    with np.errstate(all="ignore"):
                                        t = xp + xm
                                        t = t + yp
                                        t = t + ym
                                        t = t + zp
                                        t = t + zm
                                        q = np.where(code == 0, np.float32(0), t / d_of[code]).astype(np.float32)
                                        q = q - c
                                        q = (np.float32(w) * q).astype(np.float32)
                                    F[sl, wv, j, :, a] = (c + q).astype(np.float32)
                        if h + 1 < H:
                            publish(h + 1, u, (u - h + 1) % NR)
                    publish(0, u, (u + 2) % NR)
                    q, sl = s - (H - 1), (u - H + 2) % NR
                    if x0 <= q < x1:
                        for wv in range(NW):
                            for j in range(RPW):
                                y = rows(wv, j)
                                own = (y >= y0) & (y < np.minimum(y0 + TY, Y + 1))
                                for e in range(2):
                                    # a lane writes its whole pair when it owns
                                    # the first voxel (with Z odd the second
                                    # can be the z = Z + 1 ghost)
                                    z = zl + e
                                    m = (zl >= z0) & (zl < z1) & own
                                    if np.any(m):
                                        dst[q, np.broadcast_to(y, m.shape)[m], z[m]] = F[sl, wv, j, m, e]
                    load(s + 2 + D, (u - H + 1) % NR)
                s0 += NR
    return dst


def _binary(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < 0.55).astype(np.uint8)


@pytest.mark.parametrize("shape,H,NW,RPW,TX", [
    ((5, 7, 9), 1, 8, 2, 6),      # all narrow (Z < TZ)
    ((6, 9, 11), 3, 8, 2, 4),     # several x tiles, first stream start below plane 1
    ((4, 20, 140), 4, 4, 4, 2),   # one wide tile + narrow strip, four-row waves
    ((4, 30, 250), 6, 12, 2, 4),  # two wide tiles + narrow strip, odd count of narrow y tiles
    ((3, 11, 128), 4, 8, 2, 2),   # whole-z window (TZ = 120 < Z = 128)
    ((3, 9, 125), 1, 4, 4, 3),    # whole-z window, odd Z
    ((2, 13, 121), 6, 8, 2, 2),   # whole-z window (TZ = 116)
    ((2, 7, 120), 4, 8, 2, 2),    # Z = TZ: one wide tile, not whole-z
])
def test_schedule_equals_reference_iterations(shape, H, NW, RPW, TX):
    img = _binary(shape, sum(shape) + H)
    o = OracleSolver(img)
    f0 = o.field.numpy()[0].copy()
    nn = o.factor.numpy()[0]
    codes = np.zeros(f0.shape, np.int64)
    codes[1:-1, 1:-1, 1:-1] = np.where(np.isinf(nn), 0, nn).astype(np.int64)
    got = _tile_sweep(f0, codes, H, NW, RPW, TX, 0, np.float32(o.omega))
    o.run(H)
    assert np.array_equal(got, o.field.numpy()[0])


@pytest.mark.parametrize("bug", ["colour", "edge_parity", "x_slot"])
def test_planted_schedule_mistakes_are_caught(bug):
    """The same comparison fails for a wrong launch-colour constant, edge
    values read from the buffer of the current step instead of the previous
    one, and the x-1 neighbour taken from the wrong ring slot."""
    img = _binary((6, 9, 11), 7)
    o = OracleSolver(img)
    f0 = o.field.numpy()[0].copy()
    nn = o.factor.numpy()[0]
    codes = np.zeros(f0.shape, np.int64)
    codes[1:-1, 1:-1, 1:-1] = np.where(np.isinf(nn), 0, nn).astype(np.int64)
    got = _tile_sweep(f0, codes, 3, 8, 2, 4, 0, np.float32(o.omega), bug=bug)
    o.run(3)
    assert not np.array_equal(got, o.field.numpy()[0])


def _periodic_case(shape, seed):
    from oracle.restate import OraclePeriodicSolver

    o = OraclePeriodicSolver(_binary(shape, seed))
    f0 = o.field.numpy()[0].copy()
    nn = o.factor.numpy()[0]
    codes = np.zeros(f0.shape, np.int64)
    codes[1:-1, 1:-1, 1:-1] = np.where(np.isinf(nn), 0, nn).astype(np.int64)
    return o, f0, codes


@pytest.mark.parametrize("shape,H,NW,RPW,TX", [
    ((5, 6, 8), 3, 8, 2, 4),       # narrow, y halo wraps
    ((4, 2, 4), 5, 8, 2, 2),       # Y and Z below the halo: several wraps
    ((4, 20, 140), 4, 4, 4, 2),    # wide + narrow strip
])
def test_periodic_wrap_equals_reference_iterations(shape, H, NW, RPW, TX):
    """Periodic y/z with Y and Z even: H fused half-sweeps equal H periodic
    reference iterations (ghost wrap before each) on the interior."""
    o, f0, codes = _periodic_case(shape, sum(shape) + H)
    got = _tile_sweep(f0, codes, H, NW, RPW, TX, 0, np.float32(o.omega), wrap=True)
    o.run(H)
    inner = (slice(1, -1),) * 3
    assert np.array_equal(got[inner], o.field.numpy()[0][inner])


def test_periodic_wrap_needs_even_sizes():
    """With Z odd the red/black colouring does not repeat across the
    boundary: the same fused schedule then departs from the reference
    (which is why the library refuses the periodic flag for odd Y or Z)."""
    o, f0, codes = _periodic_case((5, 6, 9), 3)
    got = _tile_sweep(f0, codes, 3, 8, 2, 4, 0, np.float32(o.omega), wrap=True)
    o.run(3)
    inner = (slice(1, -1),) * 3
    assert not np.array_equal(got[inner], o.field.numpy()[0][inner])
