"""Transport-reaction tortuosity of electrodes (taufactor/electrode.py:13-157).

``ElectrodeSolver`` solves the electrolyte potential in the conductive phase
with a reaction sink at every face shared with the reactive phase: the SOR
factor is cond_nn + k0 * reac_nn (taufactor/taufactor.py:47-56), the left x
face is held at 1 (ghost value 2, weight 2) and the right face is closed.  The
tortuosity comes from the impedance of a transmission line built from the
per-layer fluxes (taufactor/electrode.py:72-104).  The iteration runs in the
same gfx950 sweep kernel as Solver (kind TF_KIND_ELECTRODE: a per-volume
prefactor table indexed by the neighbour code); the per-layer sums behind
the metrics are float64 device sums.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .solver import SORSolver, _PeriodicYZ
from .utils import compute_impedance_batched

__all__ = ["ElectrodeSolver", "PeriodicElectrodeSolver"]


class ElectrodeSolver(SORSolver):
    """Electrode tortuosity solver (taufactor/electrode.py:13-129).

    Args:
        img (numpy.ndarray): labelled image [X,Y,Z] or [B,X,Y,Z].
        conductive_label: label of the transporting phase (default 1).
        reactive_label: label whose faces with the conductive phase react (default 0).
        omega (float, optional): over-relaxation factor.
        spacing (float, optional): voxel size dx (default 1).
        device: HIP device (default 'cuda').

    After ``solve()``: ``tau`` (per batch element), ``tau_x``, ``k_x``, ``c_x``,
    ``Z_sim``, ``Z_ideal``, ``a_x``, ``k_0``, ``vol_x``.
    """

    kind = _lib.KIND_ELECTRODE

    def __init__(self, img, conductive_label=1, reactive_label=0, omega=None, spacing=None,
                 device="cuda"):
        self.left_bc = 1.0
        self.electrode_bc = 0.0
        self.cond_label = conductive_label
        self.reac_label = reactive_label
        self.dx = spacing or 1
        super().__init__(img, omega=omega, device=device)
        self.c_x = 0

    # ---------------- set-up ----------------
    def _xinit(self) -> torch.Tensor:
        """electrode_bc + (left_bc - electrode_bc) cosh(1 - x/Nx)/cosh(1) at voxel
        centres, float64 rounded to float32 (taufactor/electrode.py:39-42)."""
        x = np.arange(self.Nx) + 0.5
        c = self.electrode_bc + (self.left_bc - self.electrode_bc) * np.cosh(1 - x / self.Nx) / np.cosh(1)
        return torch.tensor(c, dtype=torch.float32).to(self.device)

    def _setup_device(self):
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        img = self.cpu_img
        cls = np.where(img == self.cond_label, 1, np.where(img == self.reac_label, 2, 0)).astype(np.uint8)
        self._input_dev = torch.from_numpy(np.ascontiguousarray(cls)).to(self.device)
        shape = (B, X + 2, Y + 2, Z + 2)
        self._codes = torch.empty(shape, dtype=torch.uint8, device=self.device)
        self._sweep_codes = _lib.sweep_code_buffer(B, X, Y, Z, self.device)
        self._fields = [_lib.slack_buffer(shape, torch.float32, self.device) for _ in range(2)]
        self._count = torch.empty((B, X), dtype=torch.int32, device=self.device)
        self._reac_count = torch.empty((B, X), dtype=torch.int32, device=self.device)
        self._xinit_dev = self._xinit()
        self._n_classes = _lib.ELECTRODE_TABLE

    def _run_setup(self):
        B, X, Y, Z = self.batch_size, self.Nx, self.Ny, self.Nz
        st = self._lib.tf_setup_electrode(_lib.ptr(self._input_dev), _lib.ptr(self._xinit_dev),
                                          _lib.ptr(self._codes), _lib.ptr(self._fields[0]),
                                          _lib.ptr(self._fields[1]), _lib.ptr(self._count),
                                          _lib.ptr(self._reac_count), B, X, Y, Z, self._setup_flags,
                                          self._stream())
        _lib.check(st, "tf_setup_electrode")
        self._pack_sweep_codes()
        count = torch.from_numpy(self._count.cpu().numpy().astype(np.float32))
        reac = torch.from_numpy(self._reac_count.cpu().numpy().astype(np.float32))
        # the reference's float32 torch set-up arithmetic (taufactor/taufactor.py:36, 47-52)
        vol_x = count / (Y * Z)
        a_x = reac / (Y * Z * self.dx)
        k_0 = torch.mean(vol_x, 1) / torch.mean(a_x * self.dx, 1) / X ** 2
        self.vol_x = vol_x.numpy()
        self.a_x = a_x.numpy()
        self.k_0 = k_0.numpy()
        table = np.full((B, _lib.ELECTRODE_TABLE), np.inf, dtype=np.float32)
        for code in range(1, 57):
            cn, rn = np.float32((code - 1) // 7), np.float32((code - 1) % 7)
            f = cn + rn * self.k_0       # float32: product, then sum
            table[:, code] = np.where(f == 0, np.float32(np.inf), f)
        self._hm_lut = torch.from_numpy(table).to(self.device)

    def return_mask(self, img):
        return (img == self.cond_label).astype(img.dtype)

    @property
    def factor(self) -> torch.Tensor:
        """cond_nn + k0*reac_nn, inf where inert (taufactor/taufactor.py:45-54)."""
        c = self._codes[:, 1:-1, 1:-1, 1:-1].long()
        return torch.gather(self._hm_lut, 1, c.reshape(self.batch_size, -1)).reshape(c.shape)

    # ---------------- metrics ----------------
    def compute_metrics(self):
        """taufactor/electrode.py:72-104, from float64 per-layer sums."""
        flux_sum, conc_sum = self._layer_sums()
        n2 = self.Ny * self.Nz
        vol_x = self.vol_x.astype(np.float64)
        c_x = conc_sum / n2
        c_x = np.divide(c_x, vol_x, out=np.zeros_like(vol_x), where=vol_x != 0)
        relative_error = np.max(np.abs(c_x - self.c_x), axis=1)
        self.c_x = c_x

        fluxes = flux_sum / n2
        fluxes_1d = np.concatenate((2 * (self.left_bc - c_x[:, :1]), (-c_x[:, 1:] + c_x[:, :-1])), axis=1)
        fluxes_1d[:, 1:][vol_x[:, 1:] == 0] = 0
        fluxes_1d[:, 1:][vol_x[:, :-1] == 0] = 0
        eps = np.concatenate((vol_x[:, :1], 0.5 * (vol_x[:, :-1] + vol_x[:, 1:])), axis=1)
        self.tau_x = np.divide(eps * fluxes_1d, fluxes, out=np.full_like(fluxes_1d, np.nan),
                               where=fluxes != 0)
        fluxes[:, :-1] -= fluxes[:, 1:].copy()
        self.k_x = np.divide(fluxes, c_x - self.electrode_bc, out=np.zeros_like(c_x),
                             where=(c_x - self.electrode_bc) != 0) / self.k_0[:, None]
        a_dx = self.a_x.astype(np.float64) * self.dx
        freq = np.mean(eps, axis=1, keepdims=True) / np.mean(a_dx, axis=1, keepdims=True) / self.Nx ** 2 * 2 ** -3
        R = self.tau_x / eps
        R[eps == 0] = 1e30
        R[np.isnan(self.tau_x)] = 1e30
        self.Z_sim = compute_impedance_batched(R, a_dx, freq)
        R_ideal = 1 / np.mean(eps, axis=1)[:, None].repeat(self.Nx, axis=1)
        C_ideal = np.mean(a_dx, axis=1)[:, None].repeat(self.Nx, axis=1)
        self.Z_ideal = compute_impedance_batched(R_ideal, C_ideal, freq)
        tau = self.Z_sim[:, 0].real / self.Z_ideal[:, 0].real
        return tau, relative_error

    def check_convergence(self, verbose, conv_crit, plot_interval):
        """SORSolver.check_convergence (taufactor/taufactor.py:109-153); the
        electrode tau is kept in float64 as the reference's numpy result."""
        tau, relative_error = self.compute_metrics()
        if verbose == "per_iter":
            i = np.argmax(relative_error)
            print(f"Iter: {self.iter}, conv error: {abs(relative_error[i]):.3E}, "
                  f"tau: {tau[i]:.5f} (batch element {i})")
        if verbose == "debug":
            self.tau_t.append(tau)
        self.tau = tau
        if not np.all(relative_error < conv_crit):
            self.old_tau = tau
            return False
        if not np.max(np.abs(tau - self.old_tau)) < 2e-3:
            self.old_tau = tau
            return False
        self.tau = tau.copy()
        self.tau[self.tau == 0] = np.inf
        return True


class PeriodicElectrodeSolver(_PeriodicYZ, ElectrodeSolver):
    """ElectrodeSolver with periodic y/z boundaries (taufactor/electrode.py:131-157)."""
