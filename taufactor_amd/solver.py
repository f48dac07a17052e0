"""Host side of the MI355X steady-state diffusion solve.

The classes mirror the reference's solver API (taufactor/taufactor.py):
constructor signatures, attributes, ``solve()`` arguments, return values,
printed messages and input errors are those of ``SORSolver`` (:15-269),
``ThroughTransportSolver`` (:272-352), ``Solver`` (:355-419) and
``MultiPhaseSolver`` (:508-620).  The iteration itself runs in libtauhip
(hand-written gfx950 kernels, include/tauhip.h); Python only validates
input, launches, and evaluates the per-layer sums every 100 iterations.

Differences from the reference, all deliberate:
  * there is no CPU code path: ``device`` must name a HIP GPU (``'cuda'`` on
    ROCm); a CPU device, or no GPU, raises ``RuntimeError`` instead of
    falling back to CPU with a warning (taufactor/taufactor.py:207-216);
  * the convergence metrics are evaluated from float64 per-layer sums
    (the reference averages in float32, whose flux-uniformity test cancels
    catastrophically); public result arrays are float32 as in the reference;
  * a NaN mean flux is still treated as converged (taufactor/taufactor.py:329)
    but raises a ``RuntimeWarning`` instead of passing silently;
  * ``walltime`` includes the completion of the last kernel.
"""
from __future__ import annotations

import ctypes

import math
import warnings
from abc import ABC, abstractmethod
from timeit import default_timer as timer

import numpy as np
import torch

from . import _lib, autotune

__all__ = ["SORSolver", "ThroughTransportSolver", "Solver", "PeriodicSolver", "AnisotropicSolver",
           "MultiPhaseSolver", "PeriodicMultiPhaseSolver", "FieldSolver"]

# half-sweeps fused per HBM pass unless a solver is told otherwise
DEFAULT_HALF_SWEEPS_PER_PASS = 4
# reference check cadence (taufactor/taufactor.py:184) and tau tolerance (:148)
CHECK_EVERY = 100
TAU_TOL = 2e-3


class SORSolver(ABC):
    """Template of the red-black SOR solve (taufactor/taufactor.py:15-269).

    Args:
        img: labelled input image, [X,Y], [X,Y,Z] or [B,X,Y,Z] numpy array.
        omega: over-relaxation factor; default 2 - pi/(1.5*Nx).
        precision: must be torch.float32 (the kernels compute in float32).
        device: a HIP device ('cuda', 'cuda:N').
    """

    kind: int = _lib.KIND_SOLVER
    _setup_flags: int = 0
    _sweep_flags: int = 0  # OR-ed into the kind of tf_sweep (SWEEP_PERIODIC_YZ)
    _zero_yz_ghosts: bool = True  # closed y/z boundaries: the ghosts stay 0

    def __init__(self, img: np.ndarray, omega: float | None = None, precision=None, device="cuda"):
        self.cpu_img = self._expand_to_4d(img)
        self.batch_size, self.Nx, self.Ny, self.Nz = self.cpu_img.shape
        self.device = self._init_device(device)
        self.precision = precision or torch.float
        if self.precision != torch.float32:
            raise ValueError("taufactor_amd computes in float32 only (precision=torch.float)")

        if omega is None:
            omega = 2 - torch.pi / (1.5 * self.Nx)
        self.omega = omega
        # the reference rounds omega to float32 when building its chequerboard
        # (taufactor/taufactor.py:224-225)
        self._w32 = float(np.float32(omega))

        self._lib = _lib.load()
        # launch configuration: None lets the autotuner (large volumes) or the
        # defaults decide; any choice gives identical iterates
        self.half_sweeps_per_pass = None
        self.tile = None  # (waves, rows_per_wave, tx) or None
        self.autotune = True
        self._auto_cfg = None

        self._fields = None  # two padded float32 buffers (ping-pong)
        self._cur = 0
        self._spare = None  # third buffer: more sweeps ahead of a check
        self._checked = None  # (buffers, index) of the checked field
        self._launch_events = None  # list -> (start, end, n_launches) per _advance
        self._codes = None
        self._sweep_codes = None
        self._hm_lut = None
        self._n_classes = 0
        self._spans_cache = {}
        with torch.cuda.device(self.device):
            self._setup_device()
            self._run_setup()
            # a solve changes only the field buffers, which the set-up fills
            # alike: reset() copies this initial field back into them (at the
            # copy rate) instead of running the set-up again
            shape = tuple(self._fields[0].shape)
            self._field_init = _lib.slack_buffer(shape, torch.float32, self.device)
            self._copy_field(self._fields[0], self._field_init)
            if self._zero_yz_ghosts and self._fields[0].numel() <= self.SPARE_MAX_VOXELS:
                self._spare = _lib.slack_buffer(shape, torch.float32, self.device)
                self._copy_field(self._field_init, self._spare)
            # per-layer float64 sums, flux (B x (X-1), or B x X for the
            # electrode kind) then concentration (B x X), in one buffer so
            # that a check copies them to the host in one transfer
            B, X = self.batch_size, self.Nx
            self._sums = torch.zeros(2 * B * X, dtype=torch.float64, device=self.device)
            self._sums_host = torch.empty(2 * B * X, dtype=torch.float64, pin_memory=True)
            # per-chunk partial sums of tf_layer_sums_ws
            ws = self._lib.tf_layer_sums_workspace_bytes(B, X, self.Ny, self.Nz)
            if ws < 0:
                _lib.check(int(ws), "tf_layer_sums_workspace_bytes")
            self._sums_ws = torch.empty(max(ws // 8, 1), dtype=torch.float64, device=self.device)
        self._sums_ready = None  # event: sums enqueued, not yet read
        self._ahead = 0  # half-sweeps already launched for the next block

        self.converged = False
        self.old_tau = 0
        self.iter = 0
        self.tau = None
        self.tau_x = None
        self.D_eff = None

    # ---------------- hooks the subclasses implement ----------------
    @abstractmethod
    def _setup_device(self) -> None:
        """Upload the input, allocate codes and field buffers, set self.vol_x."""

    @abstractmethod
    def _run_setup(self) -> None:
        """Write codes and both field buffers from the device-resident input."""

    def _copy_field(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        """dst := src for two field buffers (tf_hbm_copy in whole 16-byte
        quads: up to three floats of the zero slack after the field go
        along, zeros on both sides)."""
        n = (src.numel() + 3) // 4 * 16
        _lib.check(self._lib.tf_hbm_copy(_lib.ptr(src), _lib.ptr(dst), n, self._stream()), "tf_hbm_copy")

    def reset(self) -> None:
        """Return to the initial field and iteration 0 without touching the
        host image: the initial field the set-up wrote is copied back (the
        codes and tables the set-up built do not change during a solve)."""
        with torch.cuda.device(self.device):
            for f in self._fields + ([self._spare] if self._spare is not None else []):
                self._copy_field(self._field_init, f)
        self._cur = 0
        self._ahead = 0
        self._checked = None
        self._sums_ready = None
        self.converged = False
        self.old_tau = 0
        self.iter = 0
        self.tau = None
        self.tau_x = None
        self.D_eff = None

    @abstractmethod
    def compute_metrics(self):
        """Defines tau and relative error"""

    def return_mask(self, img):
        raise NotImplementedError

    def apply_boundary_conditions(self):
        """Dirichlet in x and no-flux in y and z: held by the ghost layers."""

    # ---------------- device state ----------------
    @property
    def field(self) -> torch.Tensor:
        """Current padded field [B, Nx+2, Ny+2, Nz+2] (float32, on device)."""
        return self._fields[self._cur]

    @field.setter
    def field(self, value: torch.Tensor) -> None:
        """Set both field buffers.  With closed y/z boundaries the y/z ghost
        layers must hold 0, as the set-up writes them (the reference pads with
        zeros, taufactor/taufactor.py:227-238): the fused sweep updates halo
        rows and z voxels unconditionally and relies on a zero ghost staying
        zero."""
        value = torch.as_tensor(value, dtype=torch.float32, device=self.device)
        if self._zero_yz_ghosts and value.dim() == 4 and (
                torch.any(value[:, :, [0, -1], :] != 0) or torch.any(value[:, :, :, [0, -1]] != 0)):
            raise ValueError("the y/z ghost layers of the field must hold 0 (closed boundaries)")
        for f in self._fields + ([self._spare] if self._spare is not None else []):
            f.copy_(value)

    @property
    def cb(self):
        """[omega*cb, omega*(1-cb)] float32 chequerboards (taufactor/taufactor.py:218-225)."""
        x = torch.arange(self.Nx, device=self.device).view(-1, 1, 1)
        y = torch.arange(self.Ny, device=self.device).view(1, -1, 1)
        z = torch.arange(self.Nz, device=self.device).view(1, 1, -1)
        even = ((x + y + z) % 2 == 0).to(torch.float64)
        return [(self.omega * even).to(torch.float32), (self.omega * (1 - even)).to(torch.float32)]

    def _stream(self) -> int:
        return _lib.stream_handle(self.device)

    def _xinit(self) -> torch.Tensor:
        # linear initial profile (ThroughTransportSolver.init_field,
        # taufactor/taufactor.py:284-286), computed by torch.linspace in float32
        sh = 1 / (2 * self.Nx)
        vec = torch.linspace(self.top_bc + sh, self.bot_bc - sh, self.Nx, dtype=torch.float32)
        return vec.to(self.device)

    def _pack_sweep_codes(self) -> None:
        """Repack the codes into the aligned layout the sweep reads."""
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        st = self._lib.tf_pack_sweep_codes(_lib.ptr(self._codes), _lib.ptr(self._sweep_codes),
                                           self.kind, B, X, Y, Z, self._stream())
        _lib.check(st, "tf_pack_sweep_codes")

    _tunable = True

    def _launch_config(self):
        """(H, tile) of the next launches: explicit settings win; otherwise the
        autotuner's choice for volumes of at least autotune.MIN_VOXELS voxels
        (made once, before the first iteration), for smaller ones with
        neither set autotune.small_volume_default, else H = 4 and the
        library's default tile."""
        H, tile = self.half_sweeps_per_pass, self.tile
        if H is not None and tile is not None:
            return int(H), tuple(tile)
        voxels = self.batch_size * self.Nx * self.Ny * self.Nz
        big = voxels >= autotune.MIN_VOXELS
        if H is None and tile is None and self.autotune and self._tunable and not big:
            sH, stile = autotune.small_volume_default(voxels)
            return (sH or DEFAULT_HALF_SWEEPS_PER_PASS), stile
        if self._auto_cfg is None:
            if self.autotune and self._tunable and big and self.iter == 0:
                self._auto_cfg = autotune.tune_solver(self, H, tile)
            else:
                self._auto_cfg = (None, None)
        aH, atile = self._auto_cfg
        H = int(H) if H is not None else (aH or DEFAULT_HALF_SWEEPS_PER_PASS)
        tile = tuple(tile) if tile is not None else atile
        return H, tile

    def _sweep_target(self):
        """(kind, codes, table, table length) the sweep reads: this kind's
        aligned sweep codes (or its codes) and its table."""
        codes = self._sweep_codes if self._sweep_codes is not None else self._codes
        return self.kind, codes, self._hm_lut, self._n_classes

    def _sweep_pointers(self):
        """(kind with the sweep flags, codes address, table address, table
        length) for tf_sweep, after checking that the codes buffer holds the
        bytes the kind's layout addresses (the kernel reads it by that layout,
        whatever the buffer)."""
        kind, codes, table, n_table = self._sweep_target()
        need = sweep_codes_bytes(self._lib, kind, self.batch_size, self.Nx, self.Ny, self.Nz)
        have = codes.numel() * codes.element_size()
        if have < need:
            raise RuntimeError(f"sweep codes of kind {kind} need {need} bytes, the buffer holds {have}")
        if n_table and (table is None or table.numel() < n_table):
            raise RuntimeError(f"the sweep table of kind {kind} needs {n_table} entries")
        return kind | self._sweep_flags, _lib.ptr(codes), _lib.ptr(table) if table is not None else None, n_table

    def resolved_launch(self):
        """(H, (waves, rows_per_wave, tx), n_tiles) of the next sweep launches
        with every default resolved by the library (tf_sweep_launch_config)."""
        H, tile = self._launch_config()
        H = max(1, min(H, _lib.MAX_FUSED_HALF_SWEEPS))
        out, n = _lib.Tile(), ctypes.c_int(0)
        req = _lib.Tile(*tile) if tile is not None else None
        kind, _, _, n_table = self._sweep_pointers()
        st = self._lib.tf_sweep_launch_config(kind, n_table, self.batch_size, self.Nx,
                                              self.Ny, self.Nz, H, req, ctypes.byref(out),
                                              ctypes.byref(n))
        _lib.check(st, "tf_sweep_launch_config")
        return H, (out.waves, out.rows_per_wave, out.tx), n.value

    def _sweep_args(self):
        """(per, tile, codes, hm) of the next sweep launches."""
        per, tile = self._launch_config()
        per = max(1, min(per, _lib.MAX_FUSED_HALF_SWEEPS))
        if tile is not None:
            tile = _lib.Tile(*tile)
        # the sweep reads the aligned sweep codes (the face conductances for
        # the field kind, the face codes for MultiPhase volumes that fit them)
        _, codes, hm, _ = self._sweep_pointers()
        return per, tile, codes, hm

    def _launch(self, h: int, it: int, tile, codes, hm) -> None:
        """One tf_sweep launch of h half-sweeps from iteration it (its colour)."""
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        src = self._fields[self._cur]
        dst = self._fields[1 - self._cur]
        kind, _, _, n_table = self._sweep_pointers()
        st = self._lib.tf_sweep(_lib.ptr(src), _lib.ptr(dst), codes, hm, n_table,
                                kind, B, X, Y, Z, self._w32, it & 1, h,
                                tile, self._stream())
        _lib.check(st, "tf_sweep")
        self._cur = 1 - self._cur

    def _advance(self, n_half: int) -> None:
        """Run n_half reference iterations (half-sweeps) on the device; the
        first self._ahead of them were launched ahead of the last check."""
        done, self._ahead = self._ahead, 0
        it = self.iter + done
        n_half -= done
        per, tile, codes, hm = self._sweep_args()
        # timing (bench.py): HIP events on the launch stream around the
        # launches that fuse the full `per` half-sweeps (a shorter last
        # launch is another kernel and is left out)
        ev = self._launch_events
        full = n_half // per
        if ev is not None and full:
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
        launches = 0
        while n_half > 0:
            h = min(per, n_half)
            self._launch(h, it, tile, codes, hm)
            it += h
            n_half -= h
            launches += 1
            if ev is not None and launches == full:
                t1.record()
                ev.append((t0, t1, full))

    def _ahead_room(self, n_next: int) -> int:
        """Half-sweeps of the next block that may run before its check
        decision: its first launch (the periodic solvers override)."""
        return n_next

    # volumes of at most this many padded voxels keep a third field buffer,
    # so that up to MAX_AHEAD sweeps run while the host decides on a check
    # (its decision outlasts one short sweep); larger volumes run one ahead
    SPARE_MAX_VOXELS = 1 << 23
    MAX_AHEAD = 3

    def _launch_ahead(self, n_next: int) -> None:
        """Launch the first sweeps of the next block of n_next iterations
        while the host decides on the check just enqueued.  The first reads
        the checked field and writes the other buffer; with a third buffer the
        second writes that one, which takes the checked field's place in the
        pair, so no sweep writes the checked field and, if the solve stops
        there, _drop_ahead returns to it unchanged."""
        room = self._ahead_room(n_next)
        if room <= 0:
            return
        per, tile, codes, hm = self._sweep_args()
        launches = 1
        if self._spare is not None:
            launches = self.MAX_AHEAD
            self._checked = (list(self._fields), self._cur)
        done = 0
        for i in range(launches):
            h = min(per, room - done)
            if h <= 0:
                break
            if i == 1:
                k = 1 - self._cur  # the checked field, after the first launch
                self._fields[k], self._spare = self._spare, self._fields[k]
            self._launch(h, self.iter + done, tile, codes, hm)
            done += h
        self._ahead = done

    def _drop_ahead(self) -> None:
        if self._ahead:
            if self._checked is not None:
                fields, cur = self._checked
                pool = self._fields + [self._spare]
                self._spare = next(t for t in pool if all(t is not f for f in fields))
                self._fields, self._cur = list(fields), cur
            else:
                self._cur = 1 - self._cur
            self._ahead = 0
        self._checked = None

    def _enqueue_layer_sums(self) -> None:
        """Layer sums of the current field and their copy to pinned host
        memory, enqueued on the launch stream; _layer_sums waits for them."""
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        flux = self._sums[:B * X]
        conc = self._sums[B * X:]
        st = self._lib.tf_layer_sums_ws(
            _lib.ptr(self.field), _lib.ptr(self._codes),
            _lib.ptr(self._hm_lut) if self._hm_lut is not None else None,
            self._n_classes, self.kind, B, X, Y, Z,
            _lib.ptr(flux), _lib.ptr(conc), _lib.ptr(self._sums_ws), self._sums_ws.numel() * 8,
            self._stream())
        _lib.check(st, "tf_layer_sums_ws")
        self._sums_host.copy_(self._sums, non_blocking=True)
        self._sums_ready = torch.cuda.Event()
        self._sums_ready.record()

    def _layer_sums(self):
        """Per-layer float64 flux and concentration sums on the host (those
        enqueued by _enqueue_layer_sums, else of the current field)."""
        if self._sums_ready is None:
            self._enqueue_layer_sums()
        self._sums_ready.synchronize()
        self._sums_ready = None
        B, X = self.batch_size, self.Nx
        h = self._sums_host.numpy()
        # X - 1 internal faces; the electrode fluxes also have the left face
        width = X if self.kind == _lib.KIND_ELECTRODE else X - 1
        flux = h[:B * width].reshape(B, width).copy()
        conc = h[B * X:].reshape(B, X).copy()
        return flux, conc

    # ---------------- convergence ----------------
    def check_convergence(self, verbose, conv_crit, plot_interval):
        """taufactor/taufactor.py:109-153; decisions are taken in float64."""
        tau, relative_error = self.compute_metrics()

        if verbose == "per_iter":
            i = np.argmax(relative_error)
            print(f"Iter: {self.iter}, conv error: {abs(relative_error[i]):.3E}, "
                  f"tau: {tau[i]:.5f} (batch element {i})")
        if verbose == "debug":
            self.tau_t.append(tau.astype(np.float32))

        if not np.all(relative_error < conv_crit):
            self.old_tau = tau
            self.tau = tau.astype(np.float32)
            return False

        tau_error = np.max(np.abs(tau - self.old_tau))
        if not tau_error < TAU_TOL:
            self.old_tau = tau
            self.tau = tau.astype(np.float32)
            return False

        tau = tau.copy()
        tau[tau == 0] = np.inf
        self.tau = tau.astype(np.float32)
        return True

    # ---------------- main loop -------------------
    def solve(self, iter_limit=10000, verbose=True, conv_crit=1e-2, plot_interval=10):
        """
        Solve steady-state with SOR solver (taufactor/taufactor.py:156-191)

        :param iter_limit: max iterations before aborting
        :param verbose: Set to 'True', 'per_iter' or 'debug' for more feedback
        :param conv_crit: convergence criteria, minimum percent difference between
        max and min flux through a given layer
        :return: tau_x, or tau if no convergence check ran
        """
        if verbose:
            torch.cuda.reset_peak_memory_stats(device=self.device)
        if verbose == "debug":
            self.tau_t = []

        with torch.no_grad(), torch.cuda.device(self.device):
            start = timer()
            # apply_boundary_conditions before every iteration (reference
            # loop) is part of _advance: the ghosts of the open solvers are
            # constant, the periodic solvers wrap (or read across) y and z
            while not self.converged and self.iter < iter_limit:
                next_check = (self.iter // CHECK_EVERY + 1) * CHECK_EVERY
                n = min(next_check, iter_limit) - self.iter
                self._advance(n)
                self.iter += n
                if self.iter % CHECK_EVERY == 0:
                    # the device runs the first sweep of the next block while
                    # the host takes the decision; dropped if the solve stops
                    self._enqueue_layer_sums()
                    self._launch_ahead(min(self.iter + CHECK_EVERY, iter_limit) - self.iter)
                    try:
                        self.converged = self.check_convergence(verbose, conv_crit, plot_interval)
                    except BaseException:
                        self._drop_ahead()
                        self._sums_ready = None
                        raise
                    if self.converged:
                        self._drop_ahead()
            torch.cuda.synchronize(self.device)
            self.walltime = timer() - start
            self._end_simulation(self.iter, verbose)
            if self.tau_x is None:
                return self.tau
            return self.tau_x

    # ---------------- helpers ----------------
    @staticmethod
    def _expand_to_4d(img: np.ndarray) -> np.ndarray:
        if not isinstance(img, np.ndarray):
            raise TypeError("Error: input image must be a NumPy array!")
        if img.ndim == 2:
            img = img[..., None]
        if img.ndim == 3:
            img = img[None, ...]
        if img.ndim != 4:
            raise ValueError("expected [B, X, Y, Z]")
        return img

    @staticmethod
    def _init_device(device) -> torch.device:
        dev = torch.device(device)
        if dev.type != "cuda":
            raise RuntimeError(
                f"taufactor_amd runs on a HIP GPU only (device='cuda' on ROCm); got device='{device}'")
        if not torch.cuda.is_available():
            raise RuntimeError("taufactor_amd needs a HIP GPU, but torch.cuda.is_available() is False")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        return dev

    def _end_simulation(self, iterations: int, verbose: bool):
        if self.converged:
            msg = "converged to"
        else:
            print("Warning: not converged")
            msg = "unconverged value of tau"

        if verbose:
            per_iter = self.walltime / iterations if iterations else float("nan")
            print(f"{msg}: {self.tau} after: {iterations} iterations in: "
                  f"{np.around(self.walltime, 4)} s "
                  f"({np.around(per_iter, 4)} s/iter)")
            print(f"GPU-RAM currently {torch.cuda.memory_allocated(device=self.device) / 1e6:.2f} MB "
                  f"(max allocated {torch.cuda.max_memory_allocated(device=self.device) / 1e6:.2f} MB; "
                  f"{torch.cuda.max_memory_reserved(device=self.device) / 1e6:.2f} MB reserved)")


class ThroughTransportSolver(SORSolver):
    """Dirichlet x boundaries at (-0.5, +0.5), zero flux on y and z
    (taufactor/taufactor.py:272-352)."""

    def __init__(self, img, omega=None, precision=None, device="cuda"):
        self.top_bc, self.bot_bc = (-0.5, 0.5)
        super().__init__(img, omega, precision, device)

    def _spans(self, b: int) -> bool:
        """Does a 6-connected conductive cluster of batch element b span x?
        (extract_through_feature(...)[1] != 0, taufactor/taufactor.py:320-322)"""
        if b not in self._spans_cache:
            mask = np.ascontiguousarray(np.isin(self.cpu_img[b], self.conductive_labels), dtype=np.uint8)
            X, Y, Z = mask.shape
            r = self._lib.tf_spans_x(mask.ctypes.data, X, Y, Z)
            if r < 0:
                _lib.check(r, "tf_spans_x")
            self._spans_cache[b] = bool(r)
        return self._spans_cache[b]

    def compute_metrics(self):
        """taufactor/taufactor.py:293-331, from float64 per-layer sums."""
        flux_sum, conc_sum = self._layer_sums()
        n2 = self.Ny * self.Nz
        flux_1d = flux_sum / n2
        fl_max = np.max(flux_1d, axis=1)
        fl_min = np.min(flux_1d, axis=1)
        # sequential sum, the order the device-side batch check uses
        mean_fl = np.cumsum(flux_1d, axis=1)[:, -1] / flux_1d.shape[1]
        relative_error = np.divide(fl_max - fl_min, fl_max,
                                   out=np.full_like(fl_max, np.nan), where=fl_max != 0)

        D_rel = mean_fl * self.Nx / abs(self.top_bc - self.bot_bc)
        D_mean = np.asarray(self.D_mean, dtype=np.float64)
        tau = np.divide(D_mean, D_rel, out=np.full_like(D_rel, np.nan), where=D_rel != 0)

        vol_x = self.vol_x.astype(np.float64)
        c_x = conc_sum / n2
        c_x = np.divide(c_x, vol_x, out=np.zeros_like(vol_x), where=vol_x != 0)
        flux_from_c = c_x[:, 1:] - c_x[:, :-1]
        flux_from_c[vol_x[:, 1:] == 0] = 0
        flux_from_c[vol_x[:, :-1] == 0] = 0
        eps = 0.5 * (vol_x[:, :-1] + vol_x[:, 1:])
        tau_x = np.divide(eps * flux_from_c, flux_1d,
                          out=np.full_like(flux_from_c, np.nan), where=flux_1d != 0)

        for b in range(self.batch_size):
            if (fl_min[b] == 0) or (fl_max[b] == 0) or (mean_fl[b] == 0):
                if not self._spans(b):
                    print(f"Warning: batch element {b} has no percolating path!")
                    relative_error[b] = 0
                    D_rel[b] = 0
                    tau[b] = 0
                    tau_x[b, :] = 0
        nan_mean = np.isnan(mean_fl)
        if np.any(nan_mean):
            warnings.warn(f"NaN mean flux in batch element(s) {np.nonzero(nan_mean)[0].tolist()}; "
                          "treated as converged as in the reference", RuntimeWarning)
        relative_error[nan_mean] = 0

        self.flux_1d = flux_1d.astype(np.float32)
        self.c_x = c_x.astype(np.float32)
        self.tau_x = tau_x.astype(np.float32)
        self.D_rel = D_rel
        self.D_eff = (self.D_0 * D_rel).astype(np.float32)
        return tau, relative_error


def sweep_codes_bytes(lib, kind: int, B: int, X: int, Y: int, Z: int) -> int:
    """Bytes of codes tf_sweep reads for B volumes of the given kind: the
    sweep-code layout of the one-byte kinds, three float32 face sections of
    the padded layout for the per-voxel field, one uint32 per padded voxel
    for the MultiPhase face codes."""
    padded = B * (X + 2) * (Y + 2) * (Z + 2)
    if kind == _lib.KIND_FIELD:
        return 3 * padded * 4
    if kind == _lib.KIND_MULTIPHASE_FACES:
        return padded * 4
    n = lib.tf_sweep_codes_bytes(X, Y, Z)
    if n < 0:
        _lib.check(int(n), "tf_sweep_codes_bytes")
    return B * int(n)


def _reciprocal_rn(f: np.ndarray) -> np.ndarray:
    """1/f rounded correctly to float32 (f: positive normal float32).  The
    float64 quotient rounded to float32 can be a double rounding away from
    it, so the neighbours of that value are compared by their exact residual
    1 - f r (f r has at most 48 significant bits: exact in float64, and so is
    the difference)."""
    f64 = f.astype(np.float64)
    best = (1.0 / f64).astype(np.float32)
    err = np.abs(1.0 - f64 * best.astype(np.float64))
    for cand in (np.nextafter(best, np.float32(0)), np.nextafter(best, np.float32(np.inf))):
        e = np.abs(1.0 - f64 * cand.astype(np.float64))
        better = e < err
        best = np.where(better, cand, best)
        err = np.where(better, e, err)
    return best.astype(np.float32)


def face_code_tables(d_cls: np.ndarray):
    """Tables of the face-code sweep (TF_KIND_MULTIPHASE_FACES, include/tauhip.h)
    for the class diffusivities d_cls (float32, class 0 = the zero class), or
    None when they do not fit it: more than MPF_MAX_CLASSES distinct
    diffusivities (0 included), more than MPF_MAX_FACTORS distinct
    prefactors, or a prefactor outside [2^-20, 2^20].

    Returns (face class of each class, uint8; key table, uint16 [4^7 * 4]:
    byte offset of each neighbourhood's prefactor pair; float32 table: the
    4 x 4 face conductances hm(D_a, D_b), then the pairs (1/f, f), pair 0 =
    (0, 0) for a zero prefactor).

    A voxel's prefactor is the reference's float32 sum of its six face
    conductances, ((((x- + x+) + y-) + y+) + z-) + z+, with the x- face added
    again in the first layer and the x+ face in the last, infinite where
    zero (MultiPhaseSolver.init_conductive_neighbours,
    taufactor/taufactor.py:595-604).  It depends only on the face classes of
    the voxel and its six neighbours and on those two layer flags, so every
    such neighbourhood is evaluated here once, with the same float32
    operations.  The harmonic means follow MultiPhaseSolver._harmonic_mean
    (:577-583) in float32, as the class-code table does."""
    M = _lib.MPF_MAX_CLASSES
    d = np.asarray(d_cls, dtype=np.float32)
    dvals = np.unique(np.concatenate([np.zeros(1, np.float32), d]))
    n = len(dvals)
    if n > M:
        return None
    fmap = np.searchsorted(dvals, d).astype(np.uint8)
    dv = np.zeros(M, np.float32)
    dv[:n] = dvals
    a, b = dv[:, None], dv[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        hm = np.where(a + b > 0, (2 * a * b) / (a + b), np.float32(0)).astype(np.float32)
    key = np.arange(M ** 7)
    digit = [(key // M ** (6 - i)) % M for i in range(7)]  # c, xm, xp, ym, yp, zm, zp
    g = [hm[digit[0], digit[i]] for i in range(1, 7)]
    f = g[0] + g[1]
    for k in range(2, 6):
        f = f + g[k]
    F = np.stack([f, f + g[0], f + g[1], (f + g[0]) + g[1]], axis=1).reshape(-1)
    valid = np.repeat(np.all(np.stack(digit) < n, axis=0), 4)
    vals = np.unique(F[valid & (F != 0)])
    if len(vals) > _lib.MPF_MAX_FACTORS:
        return None
    if vals.size and (vals.min() < 2.0 ** -20 or vals.max() > 2.0 ** 20):
        return None
    idx = np.searchsorted(vals, F) + 1
    idx[(F == 0) | ~valid] = 0
    keys = (8 * idx).astype(np.uint16)
    pairs = np.stack([_reciprocal_rn(vals), vals], axis=1).reshape(-1) if vals.size else np.zeros(0, np.float32)
    table = np.concatenate([hm.reshape(-1), np.zeros(2, np.float32), pairs]).astype(np.float32)
    return fmap, keys, table


def _label_table(img: np.ndarray):
    """(labels, counts[B, L]) of the label values of a [B,X,Y,Z] image."""
    flat = img.reshape(img.shape[0], -1)
    if img.dtype.kind in "iub" or (img.dtype.kind == "f" and img.size and
                                   np.array_equal(img, np.floor(img))):
        lo = img.min() if img.size else 0
        hi = img.max() if img.size else 0
        if lo >= 0 and hi < 65536:
            idx = flat.astype(np.int64, copy=False)
            counts = np.stack([np.bincount(r, minlength=int(hi) + 1) for r in idx])
            present = np.nonzero(counts.sum(0))[0]
            return present.astype(img.dtype), counts[:, present]
    labels, inv = np.unique(flat, return_inverse=True)
    inv = inv.reshape(flat.shape)
    counts = np.stack([np.bincount(r, minlength=len(labels)) for r in inv])
    return labels, counts


class Solver(ThroughTransportSolver):
    """Two-phase (binary) through-transport solver (taufactor/taufactor.py:355-419).

    Args:
        img (numpy.ndarray): Binary image with labels in ``{0, 1}``.
        omega (float, optional): over-relaxation factor.
        D_0 (float, optional): Reference (mean) diffusivity. Defaults to ``1``.
        device (str | torch.device, optional): HIP device. Defaults to ``'cuda'``.

    Raises:
        ValueError: If labels are not strictly in ``{0, 1}``.
    """

    kind = _lib.KIND_SOLVER

    def __init__(self, img, omega=None, D_0=1, device="cuda"):
        self._check_binary_labels(img)
        self.conductive_labels = [1]
        super().__init__(img, omega=omega, device=device)
        self.D_0 = D_0
        self.D_mean = np.mean(self.vol_x, axis=1)

    def _check_binary_labels(self, img):
        arr = np.asarray(img)
        if arr.size and bool(np.all((arr == 0) | (arr == 1))):
            return
        if len(np.unique(img)) > 2 or \
           np.unique(img).max() not in [0, 1] or \
           np.unique(img).min() not in [0, 1]:
            raise ValueError(
                "Input image must only contain 0s and 1s. "
                "Your image must be segmented to use this tool. "
                "If your image has been segmented, ensure your labels are "
                "0 for non-conductive and 1 for conductive phase. "
                f"Your image has the following labels: {np.unique(img)}. "
                "If you have more than one conductive phase, use the multi-phase solver.")

    def _setup_device(self):
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        self._input_dev = torch.from_numpy(
            np.ascontiguousarray(self.cpu_img != 0, dtype=np.uint8)).to(self.device)
        shape = (B, X + 2, Y + 2, Z + 2)
        self._codes = torch.empty(shape, dtype=torch.uint8, device=self.device)
        self._sweep_codes = _lib.sweep_code_buffer(B, X, Y, Z, self.device)
        self._fields = [_lib.slack_buffer(shape, torch.float32, self.device) for _ in range(2)]
        self._count = torch.empty((B, X), dtype=torch.int32, device=self.device)
        self._xinit_dev = self._xinit()

    def _run_setup(self):
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        st = self._lib.tf_setup_solver(_lib.ptr(self._input_dev), _lib.ptr(self._xinit_dev),
                                       _lib.ptr(self._codes), _lib.ptr(self._fields[0]),
                                       _lib.ptr(self._fields[1]), _lib.ptr(self._count),
                                       B, X, Y, Z, self._setup_flags, self._stream())
        _lib.check(st, "tf_setup_solver")
        self._pack_sweep_codes()
        self.vol_x = self._count.cpu().numpy().astype(np.float32) / np.float32(Y * Z)

    @property
    def factor(self) -> torch.Tensor:
        """Neighbour count, inf where the update is void (taufactor/taufactor.py:402-410)."""
        c = self._codes[:, 1:-1, 1:-1, 1:-1].to(torch.float32)
        return torch.where(c == 0, torch.full_like(c, math.inf), c)

    def vertical_flux(self) -> torch.Tensor:
        """taufactor/taufactor.py:412-419"""
        f = self.field
        vert_flux = f[:, 2:-1, 1:-1, 1:-1] - f[:, 1:-2, 1:-1, 1:-1]
        fac = self.factor
        vert_flux[fac[:, 0:-1] > 8] = 0
        vert_flux[fac[:, 1:] > 8] = 0
        return vert_flux


class AnisotropicSolver(Solver):
    """Two-phase solver with voxel-spacing weights (taufactor/taufactor.py:422-478).

    y neighbours are weighted by Ky = (dx/dy)^2 and z neighbours by
    Kz = (dx/dz)^2, in the prefactor and in the neighbour sum.

    Args:
        img (numpy.ndarray): Binary image with labels in ``{0, 1}``.
        spacing (tuple[float, float, float]): voxel spacing ``(dx, dy, dz)``.
        omega, D_0, device: as Solver.

    Raises:
        ValueError: If ``spacing`` is not three numbers.
        UserWarning: If the spacings differ by more than a factor 10.
    """

    kind = _lib.KIND_ANISOTROPIC
    _setup_flags = _lib.SETUP_ANISOTROPIC
    def __init__(self, img, spacing, omega=None, D_0=1, device="cuda"):
        if not isinstance(spacing, (list, tuple)) or len(spacing) != 3:
            raise ValueError("spacing must be a list or tuple with three elements (dx, dy, dz)")
        if not all(isinstance(x, (int, float)) for x in spacing):
            raise ValueError("All elements in spacing must be integers or floats")
        if np.max(spacing) / np.min(spacing) > 10:
            warnings.warn("This computation is very questionable for largely different spacings e.g. dz >> dx.")
        dx, dy, dz = spacing
        self.Ky = (dx / dy) ** 2
        self.Kz = (dx / dz) ** 2
        super().__init__(img, omega=omega, D_0=D_0, device=device)

    def _nn_table(self) -> np.ndarray:
        """Prefactor of every neighbour state, summed in float32 in the
        reference's order: x-, x+, Ky*y-, Ky*y+, Kz*z-, Kz*z+ (the rolls of
        AnisotropicSolver.init_conductive_neighbours, taufactor/taufactor.py:459-471),
        inf for zero; then Ky and Kz as float32."""
        one, ky, kz = np.float32(1.0), np.float32(self.Ky), np.float32(self.Kz)
        table = np.full(_lib.ANISO_TABLE, np.inf, dtype=np.float32)
        for code in range(1, 145):
            i = code - 1
            terms = (np.float32(i // 48) * one, np.float32((i // 16) % 3) * one,
                     np.float32((i >> 3) & 1) * ky, np.float32((i >> 2) & 1) * ky,
                     np.float32((i >> 1) & 1) * kz, np.float32(i & 1) * kz)
            nn = np.float32(0.0)
            for t in terms:
                nn = np.float32(nn + t)
            table[code] = nn if nn != 0 else np.inf
        table[145], table[146] = ky, kz
        return table

    def _setup_device(self):
        super()._setup_device()
        self._hm_lut = torch.from_numpy(self._nn_table()).to(self.device)
        self._n_classes = _lib.ANISO_TABLE

    @property
    def factor(self) -> torch.Tensor:
        """Weighted neighbour sum, inf where the update is void (taufactor/taufactor.py:459-471)."""
        c = self._codes[:, 1:-1, 1:-1, 1:-1].long()
        return self._hm_lut[:145][c]


class MultiPhaseSolver(ThroughTransportSolver):
    """Multi-phase SOR solver with per-phase diffusivity and harmonic-mean
    face conductances (taufactor/taufactor.py:508-620).

    Args:
        img (numpy.ndarray): Labeled image.
        diffusivities (dict[int, float], optional): label -> diffusivity.
            Labels not provided are assumed isolating. Defaults to ``{0: 0, 1: 1}``.
        D_scaling (float, optional): reference diffusivity ``D_0``.
        omega (float, optional): over-relaxation factor.
        device (str | torch.device, optional): HIP device. Defaults to ``'cuda'``.

    Raises:
        TypeError: non-dict diffusivities or non-integer labels.
        ValueError: If any diffusivity is negative or non-finite.
    """

    kind = _lib.KIND_MULTIPHASE
    # classes (labels + the zero ghost class) of the class-code sweep and of
    # grad_D_eff (LDS tables); the face-code sweep, the per-voxel field sweep
    # (closed boundaries) and the layer sums take up to MAX_LABELS classes:
    # one per label, or one per distinct diffusivity beyond MAX_LABELS labels
    MAX_CLASSES = 64
    MAX_LABELS = 255
    # sweep through per-voxel face codes when the diffusivities allow it
    # (face_code_tables); False keeps the class-code sweep
    face_codes = True

    def __init__(self, img, diffusivities=None, D_scaling=1, omega=None, device="cuda"):
        if diffusivities is None:
            diffusivities = {0: 0, 1: 1}
        if not isinstance(diffusivities, dict):
            raise TypeError("diffusivities must be a dictionary mapping phase labels to diffusivities")
        for phase, D_p in diffusivities.items():
            if not isinstance(phase, (int, np.integer)):
                raise TypeError(f"Phase label must be integer, got {type(phase).__name__}")
            D_p = float(D_p)
            if (not np.isfinite(D_p)) or (D_p < 0):
                raise ValueError(f"Diffusivity for label {phase} must be finite and >= 0, got {D_p}")
        self.Ds = diffusivities

        missing_labels = sorted(int(lbl) for lbl in np.unique(img) if lbl not in self.Ds)
        if missing_labels:
            warnings.warn(
                "No diffusivity provided for phase label(s) "
                f"{missing_labels}; assuming these phases are isolating.",
                UserWarning,
            )
            for lbl in missing_labels:
                self.Ds[lbl] = 0.0
        self.conductive_labels = [lbl for lbl, D_p in self.Ds.items() if D_p > 0]

        super().__init__(img, omega=omega, device=device)
        self.VF = {int(p): self._label_counts[:, i] / float(self.Nx * self.Ny * self.Nz)
                   for i, p in enumerate(self._labels)}
        self.D_0 = D_scaling
        self.D_mean = np.sum([self.VF[z] * self.Ds.get(z, 0.0) for z in self.VF], axis=0)

    def _setup_device(self):
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        labels, counts = _label_table(self.cpu_img)
        self._labels, self._label_counts = labels, counts
        # diffusivity per label as the reference's float32 diff_map holds it
        d_label = np.array([np.float32(self.Ds.get(lbl.item(), 0.0)) for lbl in labels], dtype=np.float32)
        # classes: 0 = the zero-diffusivity ghost class, then one per label
        # present in the image (labels stay separate so that gradients can be
        # reported per label); with more labels than the one-byte class codes
        # hold, one class per distinct diffusivity: the sweep, the layer sums
        # and tau see only the diffusivities (diff_map,
        # taufactor/taufactor.py:586-589), the per-label gradient is refused
        self._grouped = len(labels) > self.MAX_LABELS
        if self._grouped:
            d_values, cls_idx = np.unique(d_label, return_inverse=True)
            if len(d_values) > self.MAX_LABELS:
                raise ValueError(f"more than {self.MAX_LABELS} distinct labels are supported when their "
                                 f"diffusivities take at most {self.MAX_LABELS} values")
            d_cls = np.concatenate([np.zeros(1, dtype=np.float32), d_values]).astype(np.float32)
            cls_of_label = (cls_idx.ravel() + 1).astype(np.uint8)
        else:
            d_cls = np.concatenate([np.zeros(1, dtype=np.float32), d_label]).astype(np.float32)
            cls_of_label = np.arange(1, len(labels) + 1, dtype=np.uint8)
        # face conductance table: hm(a,b) = 2*a*b/(a+b), 0 where a+b == 0
        # (MultiPhaseSolver._harmonic_mean, taufactor/taufactor.py:577-583)
        a = d_cls[:, None]
        b = d_cls[None, :]
        denom = a + b
        with np.errstate(divide="ignore", invalid="ignore"):
            hm = np.where(denom > 0, (2 * a * b) / denom, np.float32(0)).astype(np.float32)
        self._d_cls = d_cls
        self._d_cls_dev = torch.from_numpy(d_cls).to(self.device)
        self._n_classes = len(d_cls)
        cls_img = self._class_image(labels, cls_of_label)
        self._input_dev = torch.from_numpy(cls_img).to(self.device)
        self._hm_lut = torch.from_numpy(np.ascontiguousarray(hm)).to(self.device)
        self._cls_cond = torch.from_numpy((d_cls > 0).astype(np.uint8)).to(self.device)
        shape = (B, X + 2, Y + 2, Z + 2)
        self._codes = torch.empty(shape, dtype=torch.uint8, device=self.device)
        # the face codes serve the fused sweeps (a periodic volume with odd Y
        # or Z alternates ghost wraps and single half-sweeps on the class codes)
        fused = getattr(self, "_fused_periodic", True)
        faces = face_code_tables(d_cls) if self.face_codes and fused else None
        self._faces = None
        self._field_faces = None
        if faces is None and len(d_cls) > self.MAX_CLASSES:
            if not self._zero_yz_ghosts:
                raise ValueError(f"at most {self.MAX_CLASSES - 1} distinct labels are supported with periodic "
                                 f"boundaries unless the diffusivities take at most {_lib.MPF_MAX_CLASSES} "
                                 "values (0 included)")
            # more classes than the class-code table holds: the per-voxel
            # field sweep over the face conductances of the padded
            # diffusivity map d_cls[class] (the same harmonic means, x ghosts
            # copying the first and last layer, and the same update)
            self._field_faces = {
                "kpad": _lib.slack_buffer(shape, torch.float32, self.device),
                "faces": _lib.slack_buffer((3,) + shape, torch.float32, self.device),
            }
        if self._field_faces is not None:
            self._sweep_codes = None
        elif faces is not None:
            cmap, keys, table = faces
            self._faces = {
                "cmap": torch.from_numpy(cmap).to(self.device),
                "keys": torch.from_numpy(keys.view(np.int16)).to(self.device),
                "table": torch.from_numpy(table).to(self.device),
                # uint32 codes in the field layout (int32 storage, same bits)
                "codes": _lib.slack_buffer(shape, torch.int32, self.device),
            }
            self._sweep_codes = None
        else:
            self._sweep_codes = _lib.sweep_code_buffer(B, X, Y, Z, self.device)
        self._fields = [_lib.slack_buffer(shape, torch.float32, self.device) for _ in range(2)]
        self._count = torch.empty((B, X), dtype=torch.int32, device=self.device)
        self._xinit_dev = self._xinit()

    def _run_setup(self):
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        st = self._lib.tf_setup_multiphase(_lib.ptr(self._input_dev), _lib.ptr(self._cls_cond),
                                           self._n_classes, _lib.ptr(self._xinit_dev),
                                           _lib.ptr(self._codes), _lib.ptr(self._fields[0]),
                                           _lib.ptr(self._fields[1]), _lib.ptr(self._count),
                                           B, X, Y, Z, self._setup_flags, self._stream())
        _lib.check(st, "tf_setup_multiphase")
        if self._faces is not None:
            f = self._faces
            st = self._lib.tf_multiphase_face_codes(_lib.ptr(self._codes), _lib.ptr(f["cmap"]), self._n_classes,
                                                    _lib.ptr(f["keys"]), _lib.ptr(f["codes"]), B, X, Y, Z,
                                                    self._stream())
            _lib.check(st, "tf_multiphase_face_codes")
        elif self._field_faces is not None:
            ff = self._field_faces
            d = self._d_cls_dev.index_select(0, self._codes.view(-1).to(torch.int32))
            ff["kpad"].copy_(d.view(ff["kpad"].shape))
            del d
            st = self._lib.tf_field_faces(_lib.ptr(ff["kpad"]), _lib.ptr(ff["faces"]), B, X, Y, Z,
                                          self._stream())
            _lib.check(st, "tf_field_faces")
        else:
            self._pack_sweep_codes()
        self.vol_x = self._count.cpu().numpy().astype(np.float32) / np.float32(Y * Z)

    def _sweep_target(self):
        if self._field_faces is not None:
            return _lib.KIND_FIELD, self._field_faces["faces"], None, 0
        if self._faces is None:
            return super()._sweep_target()
        t = self._faces["table"]
        return _lib.KIND_MULTIPHASE_FACES, self._faces["codes"], t, t.numel()

    def _class_image(self, labels, cls_of_label) -> np.ndarray:
        img = self.cpu_img
        if labels.size and labels.dtype.kind in "iubf" and labels.min() >= 0 and labels.max() < 65536 \
                and np.array_equal(labels, np.floor(labels)):
            lut = np.zeros(int(labels.max()) + 1, dtype=np.uint8)
            lut[labels.astype(np.int64)] = cls_of_label
            return np.ascontiguousarray(lut[img.astype(np.int64, copy=False)])
        idx = np.searchsorted(labels, img)
        return np.ascontiguousarray(cls_of_label[idx])

    def return_mask(self, img):
        if len(self.conductive_labels) == 0:
            return torch.zeros_like(img)
        conductive = torch.tensor(self.conductive_labels, dtype=self.precision, device=self.device)
        return torch.isin(img, conductive).to(self.precision)

    def grad_D_eff(self) -> dict:
        """dD_eff/dD_label for every label of the image, float64 [B] each, from
        the current (converged) field.  No reference counterpart (the
        reference solves under torch.no_grad).  The derivative of the discrete
        fixed point: see include/tauhip.h, tf_grad_classes."""
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        if self._n_classes > self.MAX_CLASSES or self._grouped:
            raise ValueError(f"grad_D_eff supports at most {self.MAX_CLASSES - 1} distinct labels")
        out = torch.empty((B, X, self._n_classes), dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            st = self._lib.tf_grad_classes(_lib.ptr(self.field), _lib.ptr(self._codes),
                                           _lib.ptr(self._d_cls_dev), self._n_classes, B, X, Y, Z,
                                           _lib.ptr(out), self._stream())
            _lib.check(st, "tf_grad_classes")
            per_cls = out.cpu().numpy().sum(axis=1) * self.D_0
        return {int(lbl): per_cls[:, i + 1] for i, lbl in enumerate(self._labels)}

    # ---- reference attributes, derived on demand from the class codes ----
    def _face_maps(self):
        c = self._codes.long()
        hm = self._hm_lut
        n = self._n_classes
        mid = c[:, :, 1:-1, 1:-1]
        D_x = hm.view(-1)[mid[:, :-1] * n + mid[:, 1:]]
        inner = c[:, 1:-1]
        D_y = hm.view(-1)[inner[:, :, :-1, 1:-1] * n + inner[:, :, 1:, 1:-1]]
        D_z = hm.view(-1)[inner[:, :, 1:-1, :-1] * n + inner[:, :, 1:-1, 1:]]
        return D_x, D_y, D_z

    @property
    def D_x(self):
        return self._face_maps()[0]

    @property
    def D_y(self):
        return self._face_maps()[1]

    @property
    def D_z(self):
        return self._face_maps()[2]

    @property
    def factor(self) -> torch.Tensor:
        """Sum of face conductances, Dirichlet faces doubled, inf where 0
        (taufactor/taufactor.py:598-604)."""
        D_x, D_y, D_z = self._face_maps()
        factor = D_x[:, :-1] + D_x[:, 1:] + D_y[:, :, :-1] + D_y[:, :, 1:] + \
            D_z[:, :, :, :-1] + D_z[:, :, :, 1:]
        factor[:, 0] += D_x[:, 0]
        factor[:, -1] += D_x[:, -1]
        factor[factor == 0] = torch.inf
        return factor

    def vertical_flux(self):
        """taufactor/taufactor.py:615-620"""
        f = self.field
        return self.D_x[:, 1:-1] * (f[:, 2:-1, 1:-1, 1:-1] - f[:, 1:-2, 1:-1, 1:-1])


class FieldSolver(ThroughTransportSolver):
    """Through-transport solve with one diffusivity per voxel.

    No reference counterpart as a class: it is the reference's
    ``MultiPhaseSolver`` (taufactor/taufactor.py:508-620) with the label ->
    diffusivity map replaced by a diffusivity per voxel, i.e. the same
    harmonic-mean face conductances, Dirichlet doubling, check and metrics;
    ``D_mean`` is the mean voxel diffusivity.  ``grad_D_eff()`` returns
    dD_eff/dk for every voxel (implicit adjoint, one pass).

    Args:
        k: voxel diffusivities >= 0 ([X,Y], [X,Y,Z] or [B,X,Y,Z]); a numpy
           array or a torch tensor (a float32 tensor on the solver's device is
           used in place).
        D_scaling: reference diffusivity D_0 (D_eff = D_0 * D_rel).
    """

    kind = _lib.KIND_FIELD

    def __init__(self, k, D_scaling=1, omega=None, device="cuda"):
        self._k_tensor = k if isinstance(k, torch.Tensor) else None
        k_np = k.detach().cpu().numpy() if isinstance(k, torch.Tensor) else k
        if not isinstance(k_np, np.ndarray):
            raise TypeError("Error: input image must be a NumPy array!")
        if k_np.size and (not np.all(np.isfinite(k_np)) or k_np.min() < 0):
            raise ValueError("voxel diffusivities must be finite and >= 0")
        self.conductive_labels = None
        super().__init__(k_np, omega=omega, device=device)
        self.D_0 = D_scaling
        self.D_mean = self.cpu_img.reshape(self.batch_size, -1).astype(np.float32).astype(np.float64).mean(axis=1)

    def _spans(self, b):
        if b not in self._spans_cache:
            mask = np.ascontiguousarray(self.cpu_img[b] > 0, dtype=np.uint8)
            X, Y, Z = mask.shape
            r = self._lib.tf_spans_x(mask.ctypes.data, X, Y, Z)
            if r < 0:
                _lib.check(r, "tf_spans_x")
            self._spans_cache[b] = bool(r)
        return self._spans_cache[b]

    def _setup_device(self):
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        t = self._k_tensor
        if t is not None and t.device == self.device and t.dtype == torch.float32:
            self._input_dev = t.detach().reshape(B, X, Y, Z).contiguous()
        else:
            self._input_dev = torch.from_numpy(
                np.ascontiguousarray(self.cpu_img, dtype=np.float32)).to(self.device)
        shape = (B, X + 2, Y + 2, Z + 2)
        self._codes = _lib.slack_buffer(shape, torch.float32, self.device)
        # the sweep reads precomputed face conductances (gx, gy, gz sections)
        self._sweep_codes = _lib.slack_buffer((3,) + shape, torch.float32, self.device)
        self._fields = [_lib.slack_buffer(shape, torch.float32, self.device) for _ in range(2)]
        self._count = torch.empty((B, X), dtype=torch.int32, device=self.device)
        self._xinit_dev = self._xinit()

    def _run_setup(self):
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        st = self._lib.tf_setup_field(_lib.ptr(self._input_dev), _lib.ptr(self._xinit_dev),
                                      _lib.ptr(self._codes), _lib.ptr(self._fields[0]),
                                      _lib.ptr(self._fields[1]), _lib.ptr(self._count),
                                      B, X, Y, Z, self._stream())
        _lib.check(st, "tf_setup_field")
        st = self._lib.tf_field_faces(_lib.ptr(self._codes), _lib.ptr(self._sweep_codes), B, X, Y, Z,
                                      self._stream())
        _lib.check(st, "tf_field_faces")
        self.vol_x = self._count.cpu().numpy().astype(np.float32) / np.float32(Y * Z)

    def grad_D_eff(self) -> torch.Tensor:
        """dD_eff/dk for every voxel, float32 [B, X, Y, Z] on the device, from
        the current (converged) field."""
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        g = torch.empty((B, X, Y, Z), dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            st = self._lib.tf_grad_field(_lib.ptr(self.field), _lib.ptr(self._codes), B, X, Y, Z,
                                         _lib.ptr(g), self._stream())
            _lib.check(st, "tf_grad_field")
        if self.D_0 != 1:
            g.mul_(self.D_0)
        return g

    def vertical_flux(self):
        k = self._codes
        kx = k[:, 1:-2, 1:-1, 1:-1]
        kn = k[:, 2:-1, 1:-1, 1:-1]
        den = kx + kn
        hm = torch.where(den > 0, (2 * kx * kn) / den, torch.zeros_like(den))
        f = self.field
        return hm * (f[:, 2:-1, 1:-1, 1:-1] - f[:, 1:-2, 1:-1, 1:-1])


class _PeriodicYZ:
    """Periodic y/z boundaries (apply_boundary_conditions before every
    iteration, taufactor/taufactor.py:501-505, 652-656).

    With Y and Z even the sweep reads every neighbour across the boundary
    from the opposite interior layer (TF_SWEEP_PERIODIC_YZ), which is what
    the refreshed ghost holds, so half-sweeps fuse as in the open solvers.
    The last half-sweep before each convergence check runs on its own and
    its source's interior is then copied into the ghosts of the result, so
    the ghosts too end as the reference's: the copy made before the last
    iteration.  With Y or Z odd the colouring does not repeat across the
    boundary and the solve alternates one ghost wrap and one half-sweep."""

    _setup_flags = _lib.SETUP_PERIODIC_YZ
    _zero_yz_ghosts = False  # never read: wrapped before use, or read across

    @property
    def _fused_periodic(self) -> bool:
        return self.Ny % 2 == 0 and self.Nz % 2 == 0

    @property
    def _tunable(self) -> bool:
        return self._fused_periodic

    @property
    def _sweep_flags(self) -> int:
        return _lib.SWEEP_PERIODIC_YZ if self._fused_periodic else 0

    def apply_boundary_conditions(self):
        st = self._lib.tf_wrap_ghosts(_lib.ptr(self.field), self.batch_size, self.Nx, self.Ny, self.Nz,
                                      self._stream())
        _lib.check(st, "tf_wrap_ghosts")

    def _ahead_room(self, n_next: int) -> int:
        # the fused block's last half-sweep runs on its own; unfused, every
        # half-sweep follows a ghost wrap of the field it reads
        return n_next - 1 if self._fused_periodic else 0

    def _advance(self, n_half: int) -> None:
        if self._fused_periodic:
            it = self.iter
            try:
                if n_half > 1:
                    SORSolver._advance(self, n_half - 1)
                    self.iter = it + n_half - 1  # colour of the last iteration
                src = self.field
                SORSolver._advance(self, 1)
            finally:
                self.iter = it
            st = self._lib.tf_wrap_ghosts_from(_lib.ptr(src), _lib.ptr(self.field), self.batch_size,
                                               self.Nx, self.Ny, self.Nz, self._stream())
            _lib.check(st, "tf_wrap_ghosts_from")
            return
        ev = self._launch_events
        if ev is not None:
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
        saved, self._launch_events = self._launch_events, None
        per = self.half_sweeps_per_pass
        self.half_sweeps_per_pass = 1
        try:
            it = self.iter
            for i in range(n_half):
                self.apply_boundary_conditions()
                self.iter = it + i  # colour of this iteration
                SORSolver._advance(self, 1)
            self.iter = it
        finally:
            self.half_sweeps_per_pass = per
            self._launch_events = saved
        if ev is not None:
            t1.record()
            ev.append((t0, t1, n_half))


class PeriodicSolver(_PeriodicYZ, Solver):
    """Two-phase solver with periodic y/z boundaries (taufactor/taufactor.py:481-505)."""


class PeriodicMultiPhaseSolver(_PeriodicYZ, MultiPhaseSolver):
    """Multi-phase solver with periodic y/z boundaries (taufactor/taufactor.py:623-656)."""
