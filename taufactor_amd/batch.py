"""Uneven batch of two-phase volumes with per-volume convergence.

The reference batches only volumes of one shape (``[B, X, Y, Z]``) and keeps
iterating until every element has converged (taufactor/taufactor.py:30-31,
143).  ``BatchSolver`` takes volumes of any shapes and gives each one exactly
the result of its own ``Solver(img).solve(...)``: its own default omega
(2 - pi/(1.5 Nx)), iteration colours, checks every 100 iterations, ``old_tau``
and stopping iteration.  All volumes advance together on the device; a
volume that stops drops out of the following passes.  The convergence
decision runs on the device, so a period of 100 iterations (the sweep passes,
the layer sums and the check) is captured once in a HIP graph and replayed
without host round trips.
"""
from __future__ import annotations

import ctypes
import math
from timeit import default_timer as timer

import numpy as np
import torch

from . import _lib, autotune
from .solver import CHECK_EVERY, Solver


def _as_3d(img: np.ndarray) -> np.ndarray:
    if not isinstance(img, np.ndarray):
        raise TypeError("Error: input image must be a NumPy array!")
    if img.ndim == 2:
        img = img[..., None]
    if img.ndim != 3:
        raise ValueError("each batch volume must be [X, Y, Z] (or [X, Y])")
    return img


class BatchSolver:
    """Solve many binary volumes of different shapes on one GPU.

    Args:
        imgs: sequence of binary numpy arrays ([X,Y,Z] or [X,Y]), labels {0, 1}.
        omega: over-relaxation factor for every volume; default per volume
            2 - pi/(1.5*Nx) as the reference's Solver.
        D_0: reference diffusivity (D_eff = D_0 * D_rel).
        device: a HIP device.

    After ``solve()``: ``tau``, ``D_eff`` (float32 [V]), ``iters`` (int [V]),
    ``converged`` (bool [V]), ``walltime``; ``field(v)`` is volume v's padded
    field on the device.
    """

    def __init__(self, imgs, omega=None, D_0=1, device="cuda"):
        imgs = [_as_3d(i) for i in imgs]
        if not imgs:
            raise ValueError("empty batch")
        for img in imgs:
            Solver._check_binary_labels(None, img)
        self.device = Solver._init_device(device)
        self._lib = _lib.load()
        self.D_0 = D_0
        # launch configuration: None lets the autotuner (large batches) or the
        # defaults (H = 4, 16-row tiles) decide; H = 1..6 (a period of 100
        # iterations ends with a shorter pass when H does not divide 100)
        self.half_sweeps_per_pass = None
        self.tile = None
        self.autotune = True
        # passes as a work queue over the live tiles (tf_batch_sweep_queue);
        # False dispatches every tile of every volume and lets the tiles of
        # stopped volumes return at once (tf_batch_sweep)
        self.use_queue = True
        self.shapes = [tuple(int(d) for d in i.shape) for i in imgs]
        V = self.n_volumes = len(imgs)
        vols = np.zeros(V, dtype=_lib.VOLUME_DTYPE)
        off = 0
        coff = 0
        moff = 0
        soff = 0
        mask_offsets = []
        for v, (X, Y, Z) in enumerate(self.shapes):
            if (X + 2) * (Y + 2) * (Z + 2) * 4 >= 2 ** 31:
                raise ValueError(f"volume {v} of shape {(X, Y, Z)} is too large for a batch (2 GiB per volume)")
            w = 2 - math.pi / (1.5 * X) if omega is None else omega
            vols[v]["offset"] = off
            vols[v]["code_offset"] = coff
            vols[v]["X"], vols[v]["Y"], vols[v]["Z"] = X, Y, Z
            vols[v]["omega"] = np.float32(w)
            vols[v]["sum_offset"] = soff
            mask_offsets.append(moff)
            off += (X + 2) * (Y + 2) * (Z + 2)
            coff += int(self._lib.tf_sweep_codes_bytes(X, Y, Z))
            moff += X * Y * Z
            soff += X
        self._vols_host = vols
        self._n_padded, self._n_layers = off, soff
        with torch.cuda.device(self.device):
            masks = np.concatenate([np.ascontiguousarray(i != 0, dtype=np.uint8).ravel() for i in imgs])
            self._mask_dev = torch.from_numpy(masks).to(self.device)
            xinit = torch.cat([torch.linspace(-0.5 + 1 / (2 * X), 0.5 - 1 / (2 * X), X, dtype=torch.float32)
                               for (X, _, _) in self.shapes])
            self._xinit_dev = xinit.to(self.device)
            self._fa = _lib.slack_buffer((off,), torch.float32, self.device)
            self._fb = _lib.slack_buffer((off,), torch.float32, self.device)
            self._codes = torch.empty(off, dtype=torch.uint8, device=self.device)
            self._sweep_codes = torch.empty(coff, dtype=torch.uint8, device=self.device)
            self._count = torch.empty(soff, dtype=torch.int32, device=self.device)
            self._flux = torch.zeros(soff, dtype=torch.float64, device=self.device)
            self._conc = torch.zeros(soff, dtype=torch.float64, device=self.device)
            self._mask_offsets = mask_offsets
            self._setup_all()
            counts = self._count.cpu().numpy()
            # a solve changes only the two field buffers, which the set-up
            # fills alike: reset() copies this initial field back into both
            # (a copy at HBM rate instead of three set-up launches per volume).
            # Whole 16-byte quads: the copy may take up to three floats of the
            # zero slack after the field, which this copy holds as zeros too
            n4 = (off + 3) // 4 * 4
            self._reset_bytes = 4 * n4
            self._field_init = torch.empty(n4, dtype=torch.float32, device=self.device)
            _lib.check(self._lib.tf_hbm_copy(_lib.ptr(self._fa), _lib.ptr(self._field_init), self._reset_bytes,
                                             self._stream()), "tf_hbm_copy")
        self.vol_x = []
        for v, (X, Y, Z) in enumerate(self.shapes):
            s0 = int(vols[v]["sum_offset"])
            vx = counts[s0:s0 + X].astype(np.float32) / np.float32(Y * Z)
            self.vol_x.append(vx)
            vols[v]["D_mean"] = float(np.mean(vx))
            m = np.ascontiguousarray(imgs[v] != 0, dtype=np.uint8)
            r = self._lib.tf_spans_x(m.ctypes.data, X, Y, Z)
            if r < 0:
                _lib.check(r, "tf_spans_x")
            vols[v]["spans"] = r
        self.D_mean = np.array([vols[v]["D_mean"] for v in range(V)], dtype=np.float32)
        self._plan = None
        self.tau = self.D_eff = self.iters = self.converged = None

    def _stream(self):
        return _lib.stream_handle(self.device)

    def _setup_all(self):
        for v, (X, Y, Z) in enumerate(self.shapes):
            o = int(self._vols_host[v]["offset"])
            s0 = int(self._vols_host[v]["sum_offset"])
            st = self._lib.tf_setup_solver(
                _lib.ptr(self._mask_dev) + self._mask_offsets[v], _lib.ptr(self._xinit_dev) + 4 * s0,
                _lib.ptr(self._codes) + o, _lib.ptr(self._fa) + 4 * o, _lib.ptr(self._fb) + 4 * o,
                _lib.ptr(self._count) + 4 * s0, 1, X, Y, Z, 0, self._stream())
            _lib.check(st, "tf_setup_solver")
            st = self._lib.tf_pack_sweep_codes(
                _lib.ptr(self._codes) + o, _lib.ptr(self._sweep_codes) + int(self._vols_host[v]["code_offset"]),
                _lib.KIND_SOLVER, 1, X, Y, Z, self._stream())
            _lib.check(st, "tf_pack_sweep_codes")

    def reset(self):
        """Back to the initial fields (volumes resident on the device)."""
        with torch.cuda.device(self.device):
            for dst in (self._fa, self._fb):
                _lib.check(self._lib.tf_hbm_copy(_lib.ptr(self._field_init), _lib.ptr(dst), self._reset_bytes,
                                                 self._stream()), "tf_hbm_copy")
        self._plan = None

    @staticmethod
    def _schedule(H):
        """Pass depths of one 100-iteration period: passes of H half-sweeps,
        then one shorter pass for the rest (H = 6: sixteen of 6, one of 4)."""
        return [H] * (CHECK_EVERY // H) + ([CHECK_EVERY % H] if CHECK_EVERY % H else [])

    def _make_plan(self, H, tile):
        """Tile tables of every pass depth of the period (the tiles of a
        volume depend on the depth), the per-layer volume table and fresh
        volume states."""
        tile_c = None if tile is None else _lib.Tile(*tile)
        dev = self.device
        depths = {}
        for h in sorted(set(self._schedule(H))):
            vols = self._vols_host.copy()
            tile_vol = []
            for v, (X, Y, Z) in enumerate(self.shapes):
                n = ctypes.c_int(0)
                st = self._lib.tf_sweep_tile_count(X, Y, Z, h, tile_c, ctypes.byref(n))
                _lib.check(st, "tf_sweep_tile_count")
                vols[v]["tile_first"] = len(tile_vol)
                tile_vol.extend([v] * n.value)
            depths[h] = dict(vols=torch.from_numpy(vols.view(np.uint8).copy()).to(dev),
                             tile_vol=torch.from_numpy(np.array(tile_vol, dtype=np.int32)).to(dev),
                             n_tiles=len(tile_vol),
                             # live tile list and [live count, one counter per pass]
                             live=torch.zeros(len(tile_vol), dtype=torch.int32, device=dev),
                             queue=torch.zeros(1 + len(self._schedule(H)), dtype=torch.int32, device=dev))
        # the checks read only the volume geometry, the same in every table
        self._vols_dev = depths[H]["vols"]
        layer_vol = np.repeat(np.arange(self.n_volumes, dtype=np.int32),
                              [X for (X, _, _) in self.shapes])
        self._layer_vol_dev = torch.from_numpy(layer_vol).to(dev)
        state = np.zeros(self.n_volumes, dtype=_lib.STATE_DTYPE)
        state["active"] = 1
        self._state_dev = torch.from_numpy(state.view(np.uint8).copy()).to(dev)
        self._plan = dict(H=H, tile=tile_c, depths=depths, schedule=self._schedule(H))
        self._compact()

    def _compact(self):
        """Live tile lists of every pass depth from the current volume states,
        and the pass counters cleared (after each check)."""
        n_counters = len(self._plan["schedule"])
        for d in self._plan["depths"].values():
            st = self._lib.tf_batch_compact(_lib.ptr(d["tile_vol"]), d["n_tiles"], _lib.ptr(self._state_dev),
                                            _lib.ptr(d["live"]), _lib.ptr(d["queue"]), n_counters, self._stream())
            _lib.check(st, "tf_batch_compact")

    def _pass(self, src_b, colour, h, index=0):
        _lib.check(self._pass_status(src_b, colour, h, index, self.use_queue), "tf_batch_sweep")

    def _pass_status(self, src_b, colour, h, index=0, queue=False):
        """One pass of depth h from buffer src_b; with queue, the work-queue form
        on the live tiles, counting with the counter of pass `index`."""
        p = self._plan
        d = p["depths"][h]
        if queue:
            return self._lib.tf_batch_sweep_queue(
                _lib.ptr(self._fa), _lib.ptr(self._fb), _lib.ptr(self._sweep_codes), _lib.ptr(d["vols"]),
                _lib.ptr(d["tile_vol"]), _lib.ptr(d["live"]), _lib.ptr(d["queue"]), index, d["n_tiles"],
                _lib.ptr(self._state_dev), src_b, colour, h, p["tile"], self._stream())
        return self._lib.tf_batch_sweep(_lib.ptr(self._fa), _lib.ptr(self._fb), _lib.ptr(self._sweep_codes),
                                      _lib.ptr(d["vols"]), _lib.ptr(d["tile_vol"]), d["n_tiles"],
                                      _lib.ptr(self._state_dev), src_b, colour, h, p["tile"], self._stream())

    def _check(self, src_b, iter_limit, conv_crit):
        st = self._lib.tf_batch_layer_sums(_lib.ptr(self._fa), _lib.ptr(self._fb), _lib.ptr(self._codes),
                                           _lib.ptr(self._vols_dev), _lib.ptr(self._layer_vol_dev),
                                           self._n_layers, _lib.ptr(self._state_dev), src_b,
                                           _lib.ptr(self._flux), _lib.ptr(self._conc), self._stream())
        _lib.check(st, "tf_batch_layer_sums")
        st = self._lib.tf_batch_check(_lib.ptr(self._vols_dev), self.n_volumes, _lib.ptr(self._flux),
                                      _lib.ptr(self._state_dev), src_b, CHECK_EVERY, iter_limit,
                                      conv_crit, self._stream())
        _lib.check(st, "tf_batch_check")

    def _period(self, src_b, iter_limit, conv_crit):
        """100 iterations (passes + check) starting from buffer src_b; returns the
        buffer holding the result."""
        done = 0
        for i, h in enumerate(self._plan["schedule"]):
            # a period starts on colour 0 (100 is even); a pass starts on the
            # colour of the half-sweeps before it
            self._pass(src_b, done & 1, h, i)
            src_b ^= 1
            done += h
        self._check(src_b, iter_limit, conv_crit)
        if self.use_queue:
            self._compact()
        return src_b

    def _launch_config(self):
        """(H, tile): explicit settings win, then the autotuner for batches of
        at least autotune.MIN_VOXELS voxels, then H = 4 and the default tile."""
        H, tile = self.half_sweeps_per_pass, self.tile
        if H is not None and tile is not None:
            return int(H), tuple(tile)
        aH, atile = None, None
        if self.autotune and sum(X * Y * Z for X, Y, Z in self.shapes) >= autotune.MIN_VOXELS:
            aH, atile = autotune.tune_batch(self, H, tile)
        H = int(H) if H is not None else (aH or 4)
        tile = tuple(tile) if tile is not None else atile
        return H, tile

    def _state(self):
        return self._state_dev.cpu().numpy().view(_lib.STATE_DTYPE)

    def solve(self, iter_limit=10000, verbose=False, conv_crit=1e-2, use_graph=True):
        """Iterate every volume until its own convergence (or iter_limit, a
        multiple of 100).  Returns tau (float32 [V]).  Every call solves from
        the initial fields: a second call resets them first (volumes that
        stopped keep their last iterate in either buffer, so continuing from
        the device state is not an option)."""
        if self._plan is not None:
            self.reset()
        H, tile = self._launch_config()
        if not 1 <= H <= _lib.MAX_FUSED_HALF_SWEEPS:
            raise ValueError(f"half_sweeps_per_pass must be 1..{_lib.MAX_FUSED_HALF_SWEEPS}")
        if iter_limit % CHECK_EVERY:
            raise ValueError("BatchSolver needs iter_limit to be a multiple of 100")
        with torch.no_grad(), torch.cuda.device(self.device):
            start = timer()
            self._make_plan(H, tile)
            n_periods = iter_limit // CHECK_EVERY
            src_b = 0
            done = 0
            graph = None
            # passes per period: the buffer parity repeats after an even count
            n_pass = len(self._plan["schedule"])
            per_graph = 1 if n_pass % 2 == 0 else 2
            while done < n_periods:
                if graph is None and use_graph and done >= 1 and n_periods - done >= per_graph:
                    graph = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(graph):
                        b = src_b
                        for _ in range(per_graph):
                            b = self._period(b, iter_limit, conv_crit)
                if graph is not None and n_periods - done >= per_graph:
                    graph.replay()
                    done += per_graph
                    src_b ^= (per_graph * n_pass) & 1
                else:
                    src_b = self._period(src_b, iter_limit, conv_crit)
                    done += 1
                if not self._state()["active"].any():
                    break
            torch.cuda.synchronize(self.device)
            self.walltime = timer() - start
        S = self._state()
        self._final_state = S
        tau = S["tau"].copy()
        self.tau = tau.astype(np.float32)
        self.D_rel = S["D_rel"].copy()
        self.D_eff = (self.D_0 * S["D_rel"]).astype(np.float32)
        self.iters = S["iters"].astype(np.int64)
        self.converged = S["converged"].astype(bool)
        if verbose:
            print(f"batch of {self.n_volumes} volumes: {int(self.converged.sum())} converged, "
                  f"iterations {int(self.iters.min())}..{int(self.iters.max())} in "
                  f"{np.around(self.walltime, 4)} s")
        return self.tau

    def field(self, v: int) -> torch.Tensor:
        """Padded final field of volume v, [X+2, Y+2, Z+2] on the device."""
        X, Y, Z = self.shapes[v]
        o = int(self._vols_host[v]["offset"])
        buf = self._fb if self._final_state["cur"][v] else self._fa
        return buf[o:o + (X + 2) * (Y + 2) * (Z + 2)].view(X + 2, Y + 2, Z + 2)
