"""Write the oracle's goldens for the benchmark configurations at their full
sizes, one JSON file per section (tests/golden/fullsize_<section>.json).

Run from the repository root (each section is independent; the CPU times are
those of an 8-core host):

  python tests/golden/make_fullsize_goldens.py C2     # ~15 min
  python tests/golden/make_fullsize_goldens.py C4     # ~5 min
  python tests/golden/make_fullsize_goldens.py C5     # ~15 min
  python tests/golden/make_fullsize_goldens.py C3     # ~20 min
  python tests/golden/make_fullsize_goldens.py C5fd   # ~70 CPU-min, --workers N

  C2    smoothed_512 (512^3, seed 0), Solver at the reference's default
        criterion: stopping iteration, tau and D_eff (float32), the float64
        D_rel, tau at every check and the SHA-256 of the final padded field
  C4    three_phase(320), MultiPhaseSolver {0:0, 1:1, 2:0.1}: the same
  C5    sigmoid_k(64, 128), per-voxel k, all 64 volumes as one batch (the
        reference's batch semantics: every volume stops at the common
        iteration): iterations, per-volume tau / D_eff, field hash
  C3    uneven_batch(256): each of the 256 volumes solved alone by the oracle:
        iterations, tau, D_eff, D_rel and field hash per volume
  C5fd  the first 24^3 corner of every one of the 64 C5 volumes: float64
        direct-solve D_rel and central differences (h = 1e-3 k) of it at 32
        seeded voxels per volume

The field hash is workloads.field_sha256 (float32 bytes, -0 folded into +0).
The oracle is the op-for-op restatement of the reference loop
(oracle/restate.py), the direct solve is oracle/direct.py: both are test
infrastructure.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import direct  # noqa: E402
from oracle.restate import OracleFieldSolver, OracleMultiPhaseSolver, OracleSolver  # noqa: E402
from taufactor_amd import workloads  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def _write(section: str, data: dict) -> None:
    path = os.path.join(GOLDEN, f"fullsize_{section}.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", path, flush=True)


def _solve_logged(o, conv_crit=1e-2, iter_limit=10000, label=""):
    """The oracle's solve loop with tau recorded at every check."""
    import torch

    taus = []
    t0 = time.time()
    with torch.no_grad():
        while not o.converged and o.iter < iter_limit:
            o.step()
            if o.iter % 100 == 0:
                o.converged = o.check_convergence(conv_crit)
                taus.append([float(np.float32(t)) for t in o.tau])
                if label:
                    print(f"{label} iter {o.iter} tau {taus[-1][:4]} ({time.time() - t0:.0f} s)", flush=True)
    return taus, time.time() - t0


def _single(o, taus, seconds, extra):
    return {
        **extra,
        "iterations": int(o.iter), "converged": bool(o.converged),
        "tau": float(np.float32(o.tau[0])), "D_eff": float(np.float32(o.D_eff[0])),
        "D_rel_oracle": float(o.D_rel[0]), "tau_at_checks": [t[0] for t in taus],
        "field_sha256": workloads.field_sha256(o.field.numpy()), "seconds": round(seconds, 1),
    }


def c2():
    img = workloads.smoothed_512(0)
    o = OracleSolver(img)
    taus, dt = _solve_logged(o, label="C2")
    _write("C2", _single(o, taus, dt, {"volume": "smoothed_512(0)", "sha256": workloads.sha256(img),
                                      "shape": list(img.shape)}))


def c4():
    img = workloads.three_phase(320, 0)
    Ds = {0: 0.0, 1: 1.0, 2: 0.1}
    o = OracleMultiPhaseSolver(img, dict(Ds))
    taus, dt = _solve_logged(o, label="C4")
    _write("C4", _single(o, taus, dt, {"volume": "three_phase(320, 0)", "sha256": workloads.sha256(img),
                                      "shape": list(img.shape),
                                      "diffusivities": {str(k): v for k, v in Ds.items()}}))


def c5():
    k = workloads.sigmoid_k(64, 128, 0)
    o = OracleFieldSolver(k)
    taus, dt = _solve_logged(o, label="C5")
    field = o.field.numpy()
    _write("C5", {
        "volume": "sigmoid_k(64, 128, 0)", "sha256": workloads.sha256(k), "shape": list(k.shape),
        "iterations": int(o.iter), "converged": bool(o.converged),
        "tau": [float(np.float32(t)) for t in o.tau], "D_eff": [float(np.float32(d)) for d in o.D_eff],
        "D_rel_oracle": [float(d) for d in o.D_rel],
        "field_sha256": [workloads.field_sha256(field[b]) for b in range(field.shape[0])],
        "seconds": round(dt, 1),
    })


def c3():
    vols = workloads.uneven_batch(256, seed=0)
    out = []
    t0 = time.time()
    for i, (img, phi, sigma) in enumerate(vols):
        o = OracleSolver(img)
        taus, dt = _solve_logged(o)
        out.append({
            "edge": int(img.shape[0]), "sha256": workloads.sha256(img), "iterations": int(o.iter),
            "converged": bool(o.converged), "tau": float(np.float32(o.tau[0])),
            "D_eff": float(np.float32(o.D_eff[0])), "D_rel_oracle": float(o.D_rel[0]),
            "field_sha256": workloads.field_sha256(o.field.numpy()),
        })
        print(f"C3 volume {i} edge {img.shape[0]}: {o.iter} iterations tau {out[-1]['tau']:.6f} "
              f"({dt:.1f} s, total {time.time() - t0:.0f} s)", flush=True)
    _write("C3", {"volumes": "uneven_batch(256, seed=0)", "per_volume": out, "seconds": round(time.time() - t0, 1)})


def _fd_volume(args):
    i, k = args
    k64 = k.astype(np.float64)
    rng = np.random.default_rng(1000 + i)
    vox = [tuple(int(c) for c in rng.integers(0, k.shape[0], size=3)) for _ in range(32)]
    fd = []
    for v in vox:
        h = 1e-3 * k64[v]
        kp, km = k64.copy(), k64.copy()
        kp[v] += h
        km[v] -= h
        fd.append((direct.solve_fixed_point(kp)[2] - direct.solve_fixed_point(km)[2]) / (2 * h))
    return {"volume": i, "sha256": workloads.sha256(k), "D_rel_direct": direct.solve_fixed_point(k64)[2],
            "voxels": [list(v) for v in vox], "grad_fd": fd}


def c5fd(workers: int):
    n = 24
    k = workloads.sigmoid_k(64, 128, 0)
    crops = [(i, np.ascontiguousarray(k[i, :n, :n, :n])) for i in range(k.shape[0])]
    del k
    t0 = time.time()
    out = []
    with multiprocessing.Pool(workers) as pool:
        for r in pool.imap(_fd_volume, crops):
            out.append(r)
            print(f"C5fd volume {r['volume']} D_rel {r['D_rel_direct']:.6f} ({time.time() - t0:.0f} s)", flush=True)
    _write("C5fd", {"crop": [n, n, n], "step": "h = 1e-3 k, central", "per_volume": out,
                    "seconds": round(time.time() - t0, 1)})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("section", choices=["C2", "C3", "C4", "C5", "C5fd"])
    ap.add_argument("--workers", type=int, default=3)
    ap.add_argument("--threads", type=int, default=0, help="torch CPU threads (0: torch's default)")
    a = ap.parse_args()
    if a.threads:
        import torch

        torch.set_num_threads(a.threads)
    if a.section == "C5fd":
        c5fd(a.workers)
    else:
        {"C2": c2, "C3": c3, "C4": c4, "C5": c5}[a.section]()


if __name__ == "__main__":
    main()
