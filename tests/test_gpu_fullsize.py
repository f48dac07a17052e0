"""GPU: the benchmark configurations at their full sizes against the oracle's
goldens (tests/golden/fullsize_<config>.json, written by
tests/golden/make_fullsize_goldens.py from the op-for-op restatement of the
reference loop, taufactor/taufactor.py:156-191, 109-153, 293-331).

* C2, the 512^3 Solver volume: the HIP solve stops at the oracle's iteration
  with the oracle's tau at every check, tau and D_eff equal in float32 and
  the final padded field equal bit for bit (SHA-256 of its bytes);
* C4, the 320^3 three-label MultiPhaseSolver volume: the same;
* C5, 64 volumes of 128^3 with per-voxel k, solved as one batch: the same
  per volume;
* C3, the 256 uneven volumes: one BatchSolver gives every volume its solo
  result (iterations, tau, D_eff, field bits) and that result is the
  oracle's; one more batch with 32-plane x tiles (several x tiles per volume)
  on the largest volumes;
* C5 gradient: on the first 24^3 corner of every one of the 64 volumes the
  per-voxel dD_eff/dk at tight convergence is within 1e-3 (of the largest)
  of central differences of the float64 direct solve at 32 voxels per
  volume; at the criteria 1e-3 (effective_diffusivity's default) and 1e-2
  (the solvers') the error is measured and bounded."""
import json
import os

import numpy as np
import pytest
import torch

import taufactor_amd as tau
from taufactor_amd import workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _gold(name):
    with open(os.path.join(GOLDEN, f"fullsize_{name}.json")) as f:
        return json.load(f)


def _field_hash(t: torch.Tensor) -> str:
    return workloads.field_sha256(t.cpu().numpy())


def _check_single(s, g):
    assert s.iter == g["iterations"] and bool(s.converged) == g["converged"]
    assert [float(t[0]) for t in s.tau_t] == g["tau_at_checks"]
    assert float(s.tau[0]) == g["tau"] and float(s.D_eff[0]) == g["D_eff"]
    rel = abs(float(s.D_rel[0]) - g["D_rel_oracle"]) / g["D_rel_oracle"]
    print(f"iterations {s.iter}, tau {float(s.tau[0]):.7f}, D_rel rel. difference to the oracle {rel:.1e}")
    assert rel < 1e-12
    assert _field_hash(s.field) == g["field_sha256"]


def test_c2_solver_512_full_solve():
    g = _gold("C2")
    img = workloads.smoothed_512(0)
    assert workloads.sha256(img) == g["sha256"]
    s = tau.Solver(img)
    s.solve(verbose="debug")
    _check_single(s, g)


def test_c4_multiphase_320_full_forward():
    g = _gold("C4")
    img = workloads.three_phase(320, 0)
    assert workloads.sha256(img) == g["sha256"]
    s = tau.MultiPhaseSolver(img, {int(k): v for k, v in g["diffusivities"].items()})
    s.solve(verbose="debug")
    _check_single(s, g)


def test_c5_field_64x128_full_forward():
    g = _gold("C5")
    k = workloads.sigmoid_k(64, 128, 0)
    assert workloads.sha256(k) == g["sha256"]
    s = tau.FieldSolver(torch.from_numpy(k).cuda())
    s.solve(verbose=False)
    assert s.iter == g["iterations"] and bool(s.converged) == g["converged"]
    assert [float(t) for t in s.tau] == g["tau"]
    assert [float(d) for d in s.D_eff] == g["D_eff"]
    field = s.field.cpu().numpy()
    assert [workloads.field_sha256(field[b]) for b in range(field.shape[0])] == g["field_sha256"]


@pytest.fixture(scope="module")
def batch256():
    return [v for v, _, _ in workloads.uneven_batch(256, seed=0)]


def test_c3_batch_256_equals_solo_and_oracle(batch256):
    g = _gold("C3")["per_volume"]
    vols = batch256
    assert [workloads.sha256(v) for v in vols] == [e["sha256"] for e in g]
    b = tau.BatchSolver(vols)
    b.solve()
    iters = [e["iterations"] for e in g]
    assert b.iters.tolist() == iters
    assert b.tau.tolist() == [e["tau"] for e in g]
    assert b.D_eff.tolist() == [e["D_eff"] for e in g]
    for v in range(len(vols)):
        assert _field_hash(b.field(v)) == g[v]["field_sha256"], v
    rel_max = 0.0
    for v, img in enumerate(vols):
        s = tau.Solver(img)
        s.solve(verbose=False)
        assert s.iter == b.iters[v], v
        assert float(s.tau[0]) == float(b.tau[v]) and float(s.D_eff[0]) == float(b.D_eff[v]), v
        assert torch.equal(s.field[0], b.field(v)), v
        rel_max = max(rel_max, abs(float(s.D_eff[0]) - float(b.D_eff[v])) / float(s.D_eff[0]))
    print(f"256 volumes, iterations {min(iters)}..{max(iters)}, largest relative D_eff difference {rel_max:.1e}")
    assert rel_max <= 1e-3


def test_c3_batch_with_32_plane_x_tiles(batch256):
    """The largest volumes (edge >= 128) with an explicit 32-plane x tile:
    several x tiles per volume (the multi-tile decode of the batch sweep)."""
    g = _gold("C3")["per_volume"]
    idx = [i for i, v in enumerate(batch256) if v.shape[0] >= 128][:12]
    assert len(idx) >= 4
    b = tau.BatchSolver([batch256[i] for i in idx])
    b.half_sweeps_per_pass = 4
    b.tile = (8, 2, 32)
    b.solve()
    for n, i in enumerate(idx):
        assert b.iters[n] == g[i]["iterations"] and float(b.tau[n]) == g[i]["tau"], i
        assert _field_hash(b.field(n)) == g[i]["field_sha256"], i


def test_c5_gradient_every_volume_against_finite_differences():
    g = _gold("C5fd")
    n = g["crop"][0]
    k = workloads.sigmoid_k(64, 128, 0)[:, :n, :n, :n].copy()
    entries = g["per_volume"]
    assert [workloads.sha256(np.ascontiguousarray(k[e["volume"]])) for e in entries] == [e["sha256"] for e in entries]
    errs = {}
    for crit in (1e-5, 1e-3, 1e-2):
        s = tau.FieldSolver(torch.from_numpy(k).cuda())
        s.solve(verbose=False, conv_crit=crit, iter_limit=200000)
        grad = s.grad_D_eff().cpu().numpy()
        worst = 0.0
        for e in entries:
            gv = grad[e["volume"]]
            got = np.array([gv[tuple(v)] for v in e["voxels"]], dtype=np.float64)
            fd = np.array(e["grad_fd"])
            worst = max(worst, float(np.max(np.abs(got - fd)) / np.max(np.abs(fd))))
            if crit == 1e-5:
                assert abs(float(s.D_rel[e["volume"]]) - e["D_rel_direct"]) <= 1e-4 * e["D_rel_direct"]
        errs[crit] = worst
        print(f"conv_crit {crit:g}: {s.iter} iterations, largest gradient error / largest |FD| over "
              f"64 volumes x 32 voxels: {worst:.2e}")
    assert errs[1e-5] <= 1e-3
    # looser criteria stop the forward early: the gradient of the
    # unconverged field is measured, with loose bounds (the default of
    # effective_diffusivity is 1e-3, the solvers' 1e-2)
    assert errs[1e-3] <= 2e-2
    assert errs[1e-2] <= 5e-2
