"""Benchmark of the MI355X steady-state diffusion solve.

Workload (one GPU): the 512^3 plain-Solver volume (seeded standard-normal
noise, Gaussian-smoothed with sigma 4, conductive below the 0.35 quantile),
solved to the reference's default convergence criterion (conv_crit 1e-2,
checks every 100 iterations).  One step = one complete solve from the
initial field with the volume already resident in HBM (Solver.reset()
re-packs it on the device).

value       GVoxel-updates/s = voxels x (iterations/2) x steps / time
            (one voxel-update = one voxel through both colours)
roofline    the sweep kernel (k_sweep): algorithmic HBM bytes per launch =
            9 B x voxels (4 B read + 4 B write of the float32 field, 1 B code,
            for however many half-sweeps the launch fuses), divided by the
            kernel's mean duration measured with HIP events on its stream
            during the timed steps; peak 8 TB/s
cpu_baseline  the CPU oracle (an op-for-op torch restatement of the
            reference loop) on this host, on a bounded sample: a fixed number
            of iterations of the same 512^3 volume, with one convergence check

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "GVoxel-updates/s + % HBM roofline; solves/s batched vs plain @ D_eff rel err≤1e-3"
HBM_PEAK_GBS = 8000.0
BYTES_PER_VOXEL_PASS = 9


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(img, n_iter: int):
    """Time the oracle on the host: n_iter reference iterations + one check."""
    import torch

    from oracle.restate import OracleSolver

    t0 = time.perf_counter()
    o = OracleSolver(img)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    with torch.no_grad():
        o.run(n_iter)
        o.check_convergence(1e-2)
    dt = time.perf_counter() - t0
    nvox = img.size
    return {
        "value": round(nvox * (n_iter / 2) / dt / 1e9, 4),
        "unit": "GVoxel-updates/s",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": (f"{n_iter} iterations + 1 convergence check of the same {img.shape[0]}^3 volume "
                   f"on the host ({dt:.1f} s; set-up {t_setup:.1f} s not timed)"),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--half-sweeps", type=int, default=None)
    ap.add_argument("--tile", type=str, default=None, help="ty,tz,tx")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this solve has no multi-GPU path in this build "
              "(see DESIGN.md for how it would shard); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import taufactor_amd as tau
    from taufactor_amd import workloads

    t0 = time.perf_counter()
    img = workloads.smoothed_binary(args.size, 0.35, 4.0, 0)
    log(f"volume {img.shape} porosity {img.mean():.4f} sha256 {workloads.sha256(img)[:16]} "
        f"({time.perf_counter() - t0:.1f} s)")

    dev = torch.device("cuda", 0)
    s = tau.Solver(img, device=dev)
    if args.half_sweeps:
        s.half_sweeps_per_pass = args.half_sweeps
    if args.tile:
        s.tile = tuple(int(v) for v in args.tile.split(","))
    H = s.half_sweeps_per_pass
    if 100 % H:
        log(f"note: {H} half-sweeps per pass does not divide 100; launches will be mixed")

    for w in range(args.warmup):
        s.reset()
        s.solve(verbose=False)
        log(f"warmup {w}: {s.iter} iterations, tau {float(s.tau[0]):.6f}")

    s._launch_events = []
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    iters = []
    for k in range(args.steps):
        s.reset()
        s.solve(verbose=False)
        iters.append(s.iter)
    torch.cuda.synchronize(dev)
    elapsed = time.perf_counter() - t0
    events = s._launch_events
    s._launch_events = None
    sweep_ms = sum(a.elapsed_time(b) for a, b, _ in events)
    launches = sum(n for _, _, n in events)
    nvox = img.size
    per_launch_s = sweep_ms / 1e3 / launches
    achieved = BYTES_PER_VOXEL_PASS * nvox / per_launch_s / 1e9
    gvus = nvox * sum(iters) / 2 / elapsed / 1e9
    log(f"{args.steps} solves, iterations {iters}, {elapsed:.3f} s; sweep kernel mean "
        f"{per_launch_s * 1e3:.3f} ms over {launches} launches -> {achieved:.0f} GB/s")

    out = {
        "metric": METRIC,
        "value": round(gvus, 3),
        "unit": "GVoxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded smoothed-noise volume, generated on the host)",
        "config": {
            "workload": f"Solver {args.size}^3 binary volume, smoothed noise sigma 4, porosity 0.35, "
                        "seed 0, converged at conv_crit 1e-2",
            "shape": list(img.shape),
            "iterations": iters[-1],
            "tau": float(s.tau[0]),
            "D_eff": float(s.D_eff[0]),
            "half_sweeps_per_pass": H,
            "tile": list(s.tile) if s.tile else "default",
            "volume_sha256": workloads.sha256(img),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_sweep",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_launch_algorithmic": BYTES_PER_VOXEL_PASS * nvox,
            "mean_launch_ms": round(per_launch_s * 1e3, 4),
        },
        "cpu_baseline": None,
    }
    traffic_file = os.path.join(ROOT, "profiles", f"k_sweep_solver_{args.size}_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            t = json.load(f)
        if t.get("half_sweeps_per_pass") == H:
            out["roofline"]["traffic"] = t["hbm_bytes_per_launch"]
    if not args.no_cpu_baseline:
        log(f"cpu baseline: {args.cpu_iters} oracle iterations at {args.size}^3 ...")
        out["cpu_baseline"] = cpu_baseline(img, args.cpu_iters)
        log(f"cpu baseline {out['cpu_baseline']['value']} GVoxel-updates/s")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
