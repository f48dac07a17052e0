"""Oracle: CPU restatement of the reference's Solver / MultiPhaseSolver solve.

TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module; the HIP path's results
are compared against it, and it is never part of that path.

It restates, op for op in float32 torch on the CPU, what the reference does:
  set-up      taufactor/taufactor.py:24-67 (SORSolver.__init__), 218-253
              (_init_chequerboard, _pad, _crop, _sum_by_rolling), 282-291
              (init_field), 380-410 (Solver), 422-478 (AnisotropicSolver),
              535-604 (MultiPhaseSolver); taufactor/electrode.py:13-157
              (ElectrodeSolver, PeriodicElectrodeSolver)
  iteration   taufactor/taufactor.py:95-103, 174-182, 606-613
  check       taufactor/taufactor.py:109-153 (check_convergence),
              293-331 (compute_metrics), 412-419 / 615-620 (vertical_flux)
  percolation taufactor/metrics/connectivity.py:83-115, 138-213
Because torch's elementwise float32 operations are IEEE-rounded one by one,
the field after k iterations here is the reference's field after k
iterations on the CPU.

The per-layer means of the check are computed either in float32 with
torch.mean, exactly as the reference does (reduce="f32"), or as float64 sums
of the same float32 face values (reduce="f64"), the form the HIP path uses.

Pinned by (tests/test_oracle_*.py): the reference's own known answers
(tests/test_taufactor.py:12-247), the stored outputs of
docs/notebooks/12-solver-benchmark.ipynb (tau and iteration counts for fcc
N=32..128 and the multiphase zigzag), an FP64 sparse direct solve of the
same discrete operator (oracle/direct.py) and closed forms.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.ndimage
import torch

TOP_BC, BOT_BC = -0.5, 0.5


def expand_to_4d(img: np.ndarray) -> np.ndarray:
    """taufactor/taufactor.py:194-204"""
    if img.ndim == 2:
        img = img[..., None]
    if img.ndim == 3:
        img = img[None, ...]
    assert img.ndim == 4
    return img


def default_omega(nx: int) -> float:
    """taufactor/taufactor.py:36-37"""
    return 2 - math.pi / (1.5 * nx)


def pad(t: torch.Tensor, vals) -> torch.Tensor:
    """Pad the three spatial dims by one voxel and set the ghost values
    (x faces, then y faces, then z faces; taufactor/taufactor.py:227-238)."""
    vals = list(vals) + [0] * (6 - len(vals))
    out = torch.nn.functional.pad(t, [1, 1, 1, 1, 1, 1], "constant")
    out[:, 0], out[:, -1] = vals[0], vals[1]
    out[:, :, 0], out[:, :, -1] = vals[2], vals[3]
    out[:, :, :, 0], out[:, :, :, -1] = vals[4], vals[5]
    return out


def chequerboards(nx, ny, nz, omega):
    """[omega*cb, omega*(1-cb)] in float32, cb = 1 where (i+j+k) is even
    (taufactor/taufactor.py:218-225)."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cb = ((i + j + k) % 2 == 0).astype(np.float64)
    return [torch.tensor(omega * cb, dtype=torch.float32), torch.tensor(omega * (1 - cb), dtype=torch.float32)]


def init_field(mask: torch.Tensor) -> torch.Tensor:
    """mask * linear profile, padded with 2*bc on the x faces
    (taufactor/taufactor.py:282-291)."""
    B, X, Y, Z = mask.shape
    sh = 1 / (2 * X)
    vec = torch.linspace(TOP_BC + sh, BOT_BC - sh, X, dtype=torch.float32)
    vec = vec.view(1, X, 1, 1).repeat(B, 1, Y, Z)
    return pad(mask * vec, [2 * TOP_BC, 2 * BOT_BC])


def spans_x(mask3d: np.ndarray) -> bool:
    """Is any 6-connected cluster present on both x faces?
    (extract_through_feature(...)[1] != 0; taufactor/metrics/connectivity.py:138-213)"""
    if not mask3d.any():
        return False
    lab, _ = scipy.ndimage.label(mask3d, structure=scipy.ndimage.generate_binary_structure(3, 1))
    first = set(np.unique(lab[0])) - {0}
    last = set(np.unique(lab[-1])) - {0}
    return len(first & last) > 0


class OracleSolve:
    """Common loop; subclasses provide factor / weighted sum / flux."""

    def __init__(self, img, omega=None, reduce="f64"):
        self.cpu_img = expand_to_4d(np.asarray(img))
        self.B, self.Nx, self.Ny, self.Nz = self.cpu_img.shape
        self.omega = default_omega(self.Nx) if omega is None else omega
        self.reduce = reduce
        torch_img = torch.tensor(self.cpu_img, dtype=torch.float32)
        mask = self.mask(torch_img)
        self.vol_x = torch.mean(mask, (2, 3)).numpy()
        self.field = init_field(mask)
        self.factor = self.init_factor(torch_img, mask)
        self.cb = chequerboards(self.Nx, self.Ny, self.Nz, self.omega)
        self.iter = 0
        self.old_tau = 0
        self.converged = False
        self.tau = None
        self.tau_x = None
        self.D_eff = None
        self.n_checks = 0

    # one reference iteration (taufactor/taufactor.py:175-182)
    def step(self):
        inc = self.weighted_sum()
        inc /= self.factor
        inc -= self.field[:, 1:-1, 1:-1, 1:-1]
        inc *= self.cb[self.iter % 2]
        self.field[:, 1:-1, 1:-1, 1:-1] += inc
        self.iter += 1

    def run(self, n):
        with torch.no_grad():
            for _ in range(n):
                self.step()
        return self

    def _layer_mean(self, t: torch.Tensor) -> np.ndarray:
        if self.reduce == "f32":
            return torch.mean(t, (2, 3)).numpy()
        return (t.to(torch.float64).sum((2, 3)) / (t.shape[2] * t.shape[3])).numpy()

    def compute_metrics(self):
        """taufactor/taufactor.py:293-331"""
        vf = self.vertical_flux()
        flux_1d = self._layer_mean(vf)
        fl_max = np.max(flux_1d, axis=1)
        fl_min = np.min(flux_1d, axis=1)
        mean_fl = np.mean(flux_1d, axis=1)
        rel = np.divide(fl_max - fl_min, fl_max, out=np.full_like(fl_max, np.nan), where=fl_max != 0)
        D_rel = mean_fl * self.Nx / abs(TOP_BC - BOT_BC)
        tau = np.divide(self.D_mean, D_rel, out=np.full_like(D_rel, np.nan), where=D_rel != 0)
        c_x = self._layer_mean(self.field[:, 1:-1, 1:-1, 1:-1])
        vol_x = self.vol_x.astype(c_x.dtype)
        c_x = np.divide(c_x, vol_x, out=np.zeros_like(vol_x), where=vol_x != 0)
        self.c_x = c_x
        dfc = c_x[:, 1:] - c_x[:, :-1]
        dfc[vol_x[:, 1:] == 0] = 0
        dfc[vol_x[:, :-1] == 0] = 0
        eps = 0.5 * (vol_x[:, :-1] + vol_x[:, 1:])
        self.tau_x = np.divide(eps * dfc, flux_1d, out=np.full_like(dfc, np.nan), where=flux_1d != 0)
        for b in range(self.B):
            if fl_min[b] == 0 or fl_max[b] == 0 or mean_fl[b] == 0:
                if not _spans_any(self.cpu_img[b], self.conductive_labels):
                    rel[b] = 0
                    D_rel[b] = 0
                    tau[b] = 0
                    self.tau_x[b, :] = 0
        rel[np.isnan(mean_fl)] = 0
        self.flux_1d = flux_1d
        self.D_rel = D_rel
        self.D_eff = self.D_0 * D_rel
        return tau, rel

    def check_convergence(self, conv_crit):
        """taufactor/taufactor.py:109-153"""
        self.n_checks += 1
        self.tau, rel = self.compute_metrics()
        self.rel_err = rel
        if not np.all(rel < conv_crit):
            self.old_tau = self.tau
            return False
        if not np.max(np.abs(self.tau - self.old_tau)) < 2e-3:
            self.old_tau = self.tau
            return False
        self.tau[self.tau == 0] = np.inf
        return True

    def solve(self, iter_limit=10000, conv_crit=1e-2):
        """taufactor/taufactor.py:156-191 without the printing"""
        with torch.no_grad():
            while not self.converged and self.iter < iter_limit:
                self.step()
                if self.iter % 100 == 0:
                    self.converged = self.check_convergence(conv_crit)
        return self


class OracleSolver(OracleSolve):
    """taufactor/taufactor.py:380-419"""

    conductive_labels = [1]

    def __init__(self, img, omega=None, D_0=1, reduce="f64"):
        super().__init__(img, omega, reduce)
        self.D_0 = D_0
        self.D_mean = np.mean(self.vol_x, axis=1)

    def mask(self, torch_img):
        return torch_img

    def init_factor(self, torch_img, mask):
        """count of conductive neighbours, x ghosts counting 2 (taufactor/taufactor.py:402-410)"""
        p = pad(mask, [2, 2])
        nn = (p[:, 2:, 1:-1, 1:-1] + p[:, :-2, 1:-1, 1:-1] + p[:, 1:-1, 2:, 1:-1] +
              p[:, 1:-1, :-2, 1:-1] + p[:, 1:-1, 1:-1, 2:] + p[:, 1:-1, 1:-1, :-2])
        nn[mask == 0] = torch.inf
        nn[nn == 0] = torch.inf
        return nn

    def weighted_sum(self):
        f = self.field
        return (f[:, 2:, 1:-1, 1:-1] + f[:, :-2, 1:-1, 1:-1] + f[:, 1:-1, 2:, 1:-1] +
                f[:, 1:-1, :-2, 1:-1] + f[:, 1:-1, 1:-1, 2:] + f[:, 1:-1, 1:-1, :-2])

    def vertical_flux(self):
        f = self.field
        vf = f[:, 2:-1, 1:-1, 1:-1] - f[:, 1:-2, 1:-1, 1:-1]
        vf[self.factor[:, 0:-1] > 8] = 0
        vf[self.factor[:, 1:] > 8] = 0
        return vf


class OracleAnisotropicSolver(OracleSolver):
    """taufactor/taufactor.py:422-478: y and z neighbours weighted by
    Ky = (dx/dy)^2 and Kz = (dx/dz)^2 (Python floats, as the reference holds
    them; torch rounds them to float32 in each product)."""

    def __init__(self, img, spacing, omega=None, D_0=1, reduce="f64"):
        dx, dy, dz = spacing
        self.Ky = (dx / dy) ** 2
        self.Kz = (dx / dz) ** 2
        super().__init__(img, omega, D_0, reduce)

    def init_factor(self, torch_img, mask):
        """AnisotropicSolver.init_conductive_neighbours (taufactor/taufactor.py:459-471):
        rolls by +1 then -1 along x, y, z of the padded mask, each times its
        weight, accumulated in that order; the rolled value at an interior
        voxel is its neighbour at -1 / +1."""
        p = pad(mask, [2, 2])
        inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
        nn = torch.zeros_like(mask)
        weight = [1.0, self.Ky, self.Kz]
        for dim in range(1, 4):
            for dr in (1, -1):
                sl = list(inner)
                sl[dim] = slice(0, -2) if dr == 1 else slice(2, None)
                nn += p[tuple(sl)] * weight[dim - 1]
        nn[mask == 0] = torch.inf
        nn[nn == 0] = torch.inf
        return nn

    def weighted_sum(self):
        """AnisotropicSolver.sum_weighted_neighbours (taufactor/taufactor.py:473-478)"""
        f = self.field
        return f[:, 2:, 1:-1, 1:-1] + f[:, :-2, 1:-1, 1:-1] + \
            self.Ky * (f[:, 1:-1, 2:, 1:-1] + f[:, 1:-1, :-2, 1:-1]) + \
            self.Kz * (f[:, 1:-1, 1:-1, 2:] + f[:, 1:-1, 1:-1, :-2])


def harmonic_mean(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """2ab/(a+b), 0 where a+b == 0 (taufactor/taufactor.py:577-583)"""
    denom = a + b
    hm = torch.zeros_like(denom)
    valid = denom > 0
    hm[valid] = 2 * a[valid] * b[valid] / denom[valid]
    return hm


class OracleMultiPhaseSolver(OracleSolve):
    """taufactor/taufactor.py:535-620"""

    def __init__(self, img, diffusivities=None, D_scaling=1, omega=None, reduce="f64"):
        Ds = dict({0: 0, 1: 1} if diffusivities is None else diffusivities)
        for lbl in np.unique(img):
            if lbl not in Ds:
                Ds[int(lbl)] = 0.0
        self.Ds = Ds
        self.conductive_labels = [lbl for lbl, d in Ds.items() if d > 0]
        super().__init__(img, omega, reduce)
        self.VF = {int(p): np.mean(self.cpu_img == p, axis=(1, 2, 3)) for p in np.unique(self.cpu_img)}
        self.D_0 = D_scaling
        self.D_mean = np.sum([self.VF[z] * self.Ds.get(z, 0.0) for z in self.VF], axis=0)

    def mask(self, torch_img):
        if not self.conductive_labels:
            return torch.zeros_like(torch_img)
        return torch.isin(torch_img, torch.tensor(self.conductive_labels, dtype=torch.float32)).to(torch.float32)

    def init_factor(self, torch_img, mask):
        """face conductances and their sum (taufactor/taufactor.py:585-604)"""
        dmap = torch.zeros_like(torch_img)
        for phase, d in self.Ds.items():
            dmap[torch_img == phase] = d
        dmap = pad(dmap, [0, 0])
        dmap[:, 0] = dmap[:, 1]
        dmap[:, -1] = dmap[:, -2]
        self.D_x = harmonic_mean(dmap[:, :-1, 1:-1, 1:-1], dmap[:, 1:, 1:-1, 1:-1])
        self.D_y = harmonic_mean(dmap[:, 1:-1, :-1, 1:-1], dmap[:, 1:-1, 1:, 1:-1])
        self.D_z = harmonic_mean(dmap[:, 1:-1, 1:-1, :-1], dmap[:, 1:-1, 1:-1, 1:])
        factor = (self.D_x[:, :-1] + self.D_x[:, 1:] + self.D_y[:, :, :-1] + self.D_y[:, :, 1:] +
                  self.D_z[:, :, :, :-1] + self.D_z[:, :, :, 1:])
        factor[:, 0] += self.D_x[:, 0]
        factor[:, -1] += self.D_x[:, -1]
        factor[factor == 0] = torch.inf
        return factor

    def weighted_sum(self):
        f = self.field
        return (f[:, 2:, 1:-1, 1:-1] * self.D_x[:, 1:] + f[:, :-2, 1:-1, 1:-1] * self.D_x[:, :-1] +
                f[:, 1:-1, 2:, 1:-1] * self.D_y[:, :, 1:] + f[:, 1:-1, :-2, 1:-1] * self.D_y[:, :, :-1] +
                f[:, 1:-1, 1:-1, 2:] * self.D_z[:, :, :, 1:] + f[:, 1:-1, 1:-1, :-2] * self.D_z[:, :, :, :-1])

    def vertical_flux(self):
        f = self.field
        return self.D_x[:, 1:-1] * (f[:, 2:-1, 1:-1, 1:-1] - f[:, 1:-2, 1:-1, 1:-1])


class OracleFieldSolver(OracleMultiPhaseSolver):
    """MultiPhaseSolver restated with the diffusivity map given per voxel
    (taufactor/taufactor.py:585-620 with diff_map = k): the oracle of
    taufactor_amd.FieldSolver.  D_mean is the mean voxel diffusivity."""

    def __init__(self, k, omega=None, D_scaling=1, reduce="f64"):
        k = expand_to_4d(np.asarray(k, dtype=np.float32))
        self._k = k
        self.Ds = {}
        OracleSolve.__init__(self, k, omega, reduce)
        self.D_0 = D_scaling
        self.D_mean = k.reshape(k.shape[0], -1).astype(np.float64).mean(axis=1)

    def mask(self, torch_img):
        return (torch_img > 0).to(torch.float32)

    def init_factor(self, torch_img, mask):
        self.conductive_labels = None
        dmap = pad(torch_img.clone(), [0, 0])
        dmap[:, 0] = dmap[:, 1]
        dmap[:, -1] = dmap[:, -2]
        self.D_x = harmonic_mean(dmap[:, :-1, 1:-1, 1:-1], dmap[:, 1:, 1:-1, 1:-1])
        self.D_y = harmonic_mean(dmap[:, 1:-1, :-1, 1:-1], dmap[:, 1:-1, 1:, 1:-1])
        self.D_z = harmonic_mean(dmap[:, 1:-1, 1:-1, :-1], dmap[:, 1:-1, 1:-1, 1:])
        factor = (self.D_x[:, :-1] + self.D_x[:, 1:] + self.D_y[:, :, :-1] + self.D_y[:, :, 1:] +
                  self.D_z[:, :, :, :-1] + self.D_z[:, :, :, 1:])
        factor[:, 0] += self.D_x[:, 0]
        factor[:, -1] += self.D_x[:, -1]
        factor[factor == 0] = torch.inf
        return factor

    def compute_metrics(self):
        # the spanning test of a per-voxel map uses k > 0
        self.conductive_labels = None
        return super().compute_metrics()


def _spans_any(img3d, labels):
    return spans_x(img3d > 0) if labels is None else spans_x(np.isin(img3d, labels))


def _wrap_yz(field: torch.Tensor) -> None:
    """y/z ghosts := opposite interior layer (taufactor/taufactor.py:501-505)."""
    field[:, :, 0, :] = field[:, :, -2, :]
    field[:, :, -1, :] = field[:, :, 1, :]
    field[:, :, :, 0] = field[:, :, :, -2]
    field[:, :, :, -1] = field[:, :, :, 1]


class OraclePeriodicSolver(OracleSolver):
    """taufactor/taufactor.py:481-505: neighbour counts wrap round in y and z,
    y/z ghosts refreshed before every iteration."""

    def init_factor(self, torch_img, mask):
        p = pad(mask, [2, 2])[:, :, 1:-1, 1:-1]  # x ghosts only
        nn = torch.zeros_like(p)
        for dim in range(1, 4):
            for dr in (1, -1):
                nn += torch.roll(p, dr, dim)
        nn = nn[:, 1:-1]
        nn[mask == 0] = torch.inf
        nn[nn == 0] = torch.inf
        return nn

    def step(self):
        _wrap_yz(self.field)
        super().step()


class OraclePeriodicMultiPhaseSolver(OracleMultiPhaseSolver):
    """taufactor/taufactor.py:623-656"""

    def init_factor(self, torch_img, mask):
        dmap = torch.zeros_like(torch_img)
        for phase, d in self.Ds.items():
            dmap[torch_img == phase] = d
        dmap = pad(dmap, [0, 0])
        dmap[:, 0] = dmap[:, 1]
        dmap[:, -1] = dmap[:, -2]
        _wrap_yz(dmap)
        self.D_x = harmonic_mean(dmap[:, :-1, 1:-1, 1:-1], dmap[:, 1:, 1:-1, 1:-1])
        self.D_y = harmonic_mean(dmap[:, 1:-1, :-1, 1:-1], dmap[:, 1:-1, 1:, 1:-1])
        self.D_z = harmonic_mean(dmap[:, 1:-1, 1:-1, :-1], dmap[:, 1:-1, 1:-1, 1:])
        factor = (self.D_x[:, :-1] + self.D_x[:, 1:] + self.D_y[:, :, :-1] + self.D_y[:, :, 1:] +
                  self.D_z[:, :, :, :-1] + self.D_z[:, :, :, 1:])
        factor[:, 0] += self.D_x[:, 0]
        factor[:, -1] += self.D_x[:, -1]
        factor[factor == 0] = torch.inf
        return factor

    def step(self):
        _wrap_yz(self.field)
        super().step()


class OracleElectrodeSolver(OracleSolve):
    """taufactor/electrode.py:13-129 (ElectrodeSolver) with the reaction
    prefactor of SORSolver.__init__ (taufactor/taufactor.py:24-67): factor =
    cond_nn + k0 * reac_nn, left x ghost 2 with weight 2, right x face closed,
    tau from the transmission-line impedance of the layer fluxes."""

    periodic = False

    def __init__(self, img, conductive_label=1, reactive_label=0, omega=None, spacing=None, reduce="f64"):
        self.left_bc, self.electrode_bc = 1.0, 0.0
        self.cond_label, self.reac_label = conductive_label, reactive_label
        self.dx = spacing or 1
        self.cpu_img = expand_to_4d(np.asarray(img))
        self.B, self.Nx, self.Ny, self.Nz = self.cpu_img.shape
        self.omega = default_omega(self.Nx) if omega is None else omega
        self.reduce = reduce
        torch_img = torch.tensor(self.cpu_img, dtype=torch.float32)
        mask = torch.zeros_like(torch_img)
        mask[torch_img == self.cond_label] = 1
        vol_x = torch.mean(mask, (2, 3))
        self.field = self._init_field(mask)
        self.factor = self._cond_nn(mask)
        reac_nn = self._reac_nn(torch_img)
        a_x = torch.sum(reac_nn, (2, 3)) / (self.Ny * self.Nz * self.dx)
        k_0 = torch.mean(vol_x, 1) / torch.mean(a_x * self.dx, 1) / self.Nx ** 2
        reac_nn = reac_nn * k_0[:, None, None, None]
        self.factor += reac_nn
        self.factor[self.factor == 0] = torch.inf
        self.a_x, self.k_0 = a_x.numpy(), k_0.numpy()
        self.vol_x = vol_x.numpy()
        self.cb = chequerboards(self.Nx, self.Ny, self.Nz, self.omega)
        self.iter = 0
        self.old_tau = 0
        self.converged = False
        self.tau = None
        self.tau_x = None
        self.c_x = 0
        self.n_checks = 0

    def _init_field(self, mask):
        """taufactor/electrode.py:39-46"""
        x = np.arange(self.Nx) + 0.5
        c = self.electrode_bc + (self.left_bc - self.electrode_bc) * np.cosh(1 - x / self.Nx) / np.cosh(1)
        vec = torch.tensor(c, dtype=torch.float32)[None, :, None, None].repeat(self.B, 1, self.Ny, self.Nz)
        return pad(mask * vec, [self.left_bc * 2, 0])

    @staticmethod
    def _roll_sum(t):
        s = torch.zeros_like(t)
        for dim in range(1, 4):
            for dr in (1, -1):
                s += torch.roll(t, dr, dim)
        return s

    def _cond_nn(self, mask):
        """taufactor/electrode.py:48-54 (periodic: :133-142)"""
        if self.periodic:
            nn = self._roll_sum(pad(mask, [2, 0])[:, :, 1:-1, 1:-1])[:, 1:-1]
        else:
            nn = self._roll_sum(pad(mask, [2, 0]))[:, 1:-1, 1:-1, 1:-1]
        nn[mask == 0] = torch.inf
        return nn

    def _reac_nn(self, img):
        """taufactor/electrode.py:56-62 (periodic: :144-151)"""
        img2 = torch.zeros_like(img)
        img2[img == self.reac_label] = 1
        if self.periodic:
            nn = self._roll_sum(pad(img2, [0, 0])[:, :, 1:-1, 1:-1])[:, 1:-1]
        else:
            nn = self._roll_sum(pad(img2, [0, 0]))[:, 1:-1, 1:-1, 1:-1]
        nn[img != self.cond_label] = 0.0
        return nn

    def weighted_sum(self):
        """SORSolver.sum_weighted_neighbours (taufactor/taufactor.py:95-103)"""
        f = self.field
        return (f[:, 2:, 1:-1, 1:-1] + f[:, :-2, 1:-1, 1:-1] + f[:, 1:-1, 2:, 1:-1] +
                f[:, 1:-1, :-2, 1:-1] + f[:, 1:-1, 1:-1, 2:] + f[:, 1:-1, 1:-1, :-2])

    def step(self):
        if self.periodic:
            _wrap_yz(self.field)
        super().step()

    def compute_metrics(self):
        """taufactor/electrode.py:72-104"""
        f = self.field
        inner = f[:, 1:-1, 1:-1, 1:-1]
        vol_x = self.vol_x if self.reduce == "f32" else self.vol_x.astype(np.float64)
        c_x = self._layer_mean(inner)
        c_x = np.divide(c_x, vol_x, out=np.zeros_like(vol_x), where=vol_x != 0)
        relative_error = np.max(np.abs(c_x - self.c_x), axis=1)
        self.c_x = c_x
        fl = -inner + f[:, :-2, 1:-1, 1:-1]
        fl[:, 0, :, :] = (self.left_bc - f[:, 1, 1:-1, 1:-1]) / 0.5
        fl[inner == 0] = 0
        fl[f[:, :-2, 1:-1, 1:-1] == 0] = 0
        fluxes = self._layer_mean(fl)
        fluxes_1d = np.concatenate((2 * (self.left_bc - c_x[:, :1]), (-c_x[:, 1:] + c_x[:, :-1])), axis=1)
        fluxes_1d[:, 1:][vol_x[:, 1:] == 0] = 0
        fluxes_1d[:, 1:][vol_x[:, :-1] == 0] = 0
        eps = np.concatenate((vol_x[:, :1], 0.5 * (vol_x[:, :-1] + vol_x[:, 1:])), axis=1)
        self.tau_x = np.divide(eps * fluxes_1d, fluxes, out=np.full_like(fluxes_1d, np.nan),
                               where=fluxes != 0)
        fluxes[:, :-1] -= fluxes[:, 1:]
        self.k_x = np.divide(fluxes, c_x - self.electrode_bc, out=np.zeros_like(c_x),
                             where=(c_x - self.electrode_bc) != 0) / self.k_0[:, None]
        a_dx = self.a_x * self.dx if self.reduce == "f32" else self.a_x.astype(np.float64) * self.dx
        freq = np.mean(eps, axis=1, keepdims=True) / np.mean(a_dx, axis=1, keepdims=True) / self.Nx ** 2 * 2 ** -3
        R = self.tau_x / eps
        R[eps == 0] = 1e30
        R[np.isnan(self.tau_x)] = 1e30
        Z_sim = _impedance_batched(R, a_dx, freq)
        R_ideal = 1 / np.mean(eps, axis=1)[:, None].repeat(self.Nx, axis=1)
        C_ideal = np.mean(a_dx, axis=1)[:, None].repeat(self.Nx, axis=1)
        Z_ideal = _impedance_batched(R_ideal, C_ideal, freq)
        return Z_sim[:, 0].real / Z_ideal[:, 0].real, relative_error


class OraclePeriodicElectrodeSolver(OracleElectrodeSolver):
    """taufactor/electrode.py:131-157"""

    periodic = True


def _impedance_batched(R, C, w):
    """compute_impedance_batched (taufactor/utils.py:87-98) for w of shape (bs, F)."""
    Z = np.full_like(w, 1e50, dtype=complex)
    for k in range(R.shape[1] - 1, -1, -1):
        Z = R[:, k][:, None] + 1.0 / (1j * w * C[:, k][:, None] + 1.0 / Z)
    return Z
