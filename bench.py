"""Benchmark of the MI355X steady-state diffusion solve.

Workload (one GPU): the 512^3 plain-Solver volume (seeded standard-normal
noise, Gaussian-smoothed with sigma 4, conductive below the 0.35 quantile),
solved to the reference's default convergence criterion (conv_crit 1e-2,
checks every 100 iterations).  One step = one complete solve from the
initial field with the volume already resident in HBM (Solver.reset()
re-packs it on the device).

value       GVoxel-updates/s = voxels x (iterations/2) x steps / time
            (one voxel-update = one voxel through both colours)
roofline    the sweep kernel (k_sweep): algorithmic bytes per launch =
            9 B per voxel-update (4 B read + 4 B write of the float32 field,
            1 B code: one unblocked pass per voxel-update) x the voxel-updates
            one launch performs (voxels x H/2 for H fused half-sweeps),
            divided by the kernel's mean duration measured with HIP events on
            its stream during the timed steps, over the launches that fuse
            all H half-sweeps (with 100 % H != 0 each check interval ends
            with a shorter launch, another kernel, left out); peak 8 TB/s.  Temporal
            blocking lets this exceed the bytes actually moved: `traffic`
            holds the measured HBM bytes per launch (rocprofv3 FETCH_SIZE +
            WRITE_SIZE, calibrated), when a summary for this H and tile exists
cpu_baseline  the CPU oracle (an op-for-op torch restatement of the
            reference loop) on this host, on a bounded sample: a fixed number
            of iterations of the same 512^3 volume, with one convergence check

Other workloads (not part of the default line):
  --config batch256       256 binary cubes of edge U[32,160] (smoothed noise,
                          porosity U[0.3,0.7], sigma U[1,5]): one BatchSolver
                          vs 256 sequential Solver solves; solves/s and the
                          largest relative D_eff difference
  --config multiphase320  320^3 three-label volume {0:0, 1:1, 2:0.1}: forward
                          solve and per-phase dD_eff/dD (one gradient pass)
  --config field128x64    64 volumes of 128^3 with k = sigmoid(4 z): batched
                          forward solve and per-voxel dD_eff/dk
  --config periodic512    PeriodicSolver (periodic y/z, the reference harness's
                          default solver) on the same 512^3 volume, solved to
                          conv_crit 1e-2: whole-solve GVoxel-updates/s
  --config hbm_copy       device-to-device copy of one 512^3 float32 field
                          (537 MB read + 537 MB written): the achievable-
                          bandwidth yardstick, and the kernel whose
                          FETCH_SIZE / WRITE_SIZE counters calibrate the
                          sweep kernel's measured HBM traffic

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                       [--config solver512|batch256|multiphase320|field128x64|periodic512]
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
    ap.add_argument("--tile", type=str, default=None, help="waves,rows_per_wave,tx")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--config", default="solver512",
                    choices=["solver512", "batch256", "multiphase320", "field128x64", "periodic512",
                             "hbm_copy"])
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the solve has no multi-GPU path (it would "
              "shard as x slabs with H-plane ghost slabs exchanged per pass); run with --gpus 1",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import taufactor_amd as tau
    from taufactor_amd import autotune, workloads

    if args.config == "batch256":
        return bench_batch(args)
    if args.config == "multiphase320":
        return bench_multiphase(args)
    if args.config == "field128x64":
        return bench_field(args)
    if args.config == "periodic512":
        return bench_periodic(args)
    if args.config == "hbm_copy":
        return bench_copy(args)

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
    t0 = time.perf_counter()
    H, tile = s._launch_config()  # runs the autotuner unless both are given
    _, rtile, n_tiles = s.resolved_launch()
    for ms, h, t in autotune.last_timings:
        log(f"  autotune candidate H={h} tile {t}: {ms:.4f} ms per half-sweep")
    log(f"launch configuration: {H} half-sweeps per launch, tile {tile or 'library default'} "
        f"-> {rtile[0]}x{rtile[1]} waves x rows, x tile {rtile[2]}, {n_tiles} tiles "
        f"({'autotuned' if args.half_sweeps is None or args.tile is None else 'given'}, "
        f"{time.perf_counter() - t0:.2f} s)")
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
    # the timed launches all fuse H half-sweeps (the solver leaves a shorter
    # last launch of a check interval out of its events): H/2 voxel-updates
    # per voxel per launch
    updates_per_launch = nvox * H / 2
    achieved = BYTES_PER_VOXEL_PASS * updates_per_launch / per_launch_s / 1e9
    gvus = nvox * sum(iters) / 2 / elapsed / 1e9
    log(f"{args.steps} solves, iterations {iters}, {elapsed:.3f} s; sweep kernel mean "
        f"{per_launch_s * 1e3:.3f} ms over {launches} launches of k_sweep<{H}, ...> -> "
        f"{achieved:.0f} GB/s")

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
            "tile": list(rtile),
            "tiles_per_launch": n_tiles,
            "launch_config": "autotuned" if args.half_sweeps is None or args.tile is None else "given",
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
            "bytes_per_launch_algorithmic": int(BYTES_PER_VOXEL_PASS * updates_per_launch),
            "mean_launch_ms": round(per_launch_s * 1e3, 4),
        },
        "cpu_baseline": None,
    }
    traffic_file = os.path.join(ROOT, "profiles", f"k_sweep_solver_{args.size}_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            entries = json.load(f)
        for t in entries if isinstance(entries, list) else [entries]:
            if t.get("half_sweeps_per_pass") == H and t.get("tile") == list(rtile):
                out["roofline"]["traffic"] = t["hbm_bytes_per_launch"]
                out["roofline"]["traffic_source"] = (
                    "rocprofv3 FETCH_SIZE x %.1f + WRITE_SIZE x %.1f per launch of %s, "
                    "profiles/k_sweep_solver_%d_traffic.md"
                    % (t["fetch_factor"], t["write_factor"], t["kernel"], args.size))
    if not args.no_cpu_baseline:
        log(f"cpu baseline: {args.cpu_iters} oracle iterations at {args.size}^3 ...")
        out["cpu_baseline"] = cpu_baseline(img, args.cpu_iters)
        log(f"cpu baseline {out['cpu_baseline']['value']} GVoxel-updates/s")
    print(json.dumps(out), flush=True)
    return 0


def _log_launch(s) -> dict:
    """Log the autotuner's candidates and the resolved launch of solver s."""
    from taufactor_amd import autotune

    for ms, H, tile in autotune.last_timings:
        log(f"  autotune candidate H={H} tile {tile}: {ms:.4f} ms per half-sweep")
    H, tile, n_tiles = s.resolved_launch()
    log(f"launch configuration: {H} half-sweeps per launch, {tile[0]}x{tile[1]} waves x rows, "
        f"x tile {tile[2]}, {n_tiles} tiles")
    return {"half_sweeps_per_pass": H, "tile": list(tile), "tiles_per_launch": n_tiles}


def _line(metric, value, unit, steps, warmup, ms, config, higher=True, dtype="f32", **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded, generated on the host)",
           "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


def bench_copy(args) -> int:
    """Copy of a --size^3 float32 field on the device, timed with HIP events."""
    import torch

    dev = torch.device("cuda", 0)
    n = args.size ** 3
    src = torch.rand(n, device=dev)
    dst = torch.empty_like(src)
    for _ in range(max(args.warmup, 1)):
        dst.copy_(src)
    reps = 20 * args.steps
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize(dev)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize(dev)
    ms = a.elapsed_time(b) / reps
    assert torch.equal(dst, src)
    gbs = 2 * 4 * n / (ms / 1e3) / 1e9
    log(f"copy {4 * n / 1e6:.0f} MB: {ms:.4f} ms -> {gbs:.0f} GB/s (read + write)")
    _line("device copy GB/s (read + write)", round(gbs, 1), "GB/s", reps, args.warmup, round(ms, 5),
          {"workload": f"torch copy_ of {n} float32 ({4 * n} B read, {4 * n} B written)"},
          bytes_read=4 * n, bytes_written=4 * n)
    return 0


def bench_batch(args) -> int:
    """256 uneven volumes: one BatchSolver against 256 sequential Solver solves."""
    import numpy as np
    import torch

    import taufactor_amd as tau
    from taufactor_amd import autotune, workloads

    t0 = time.perf_counter()
    vols = [v for v, _, _ in workloads.uneven_batch(256, seed=0)]
    nvox = sum(v.size for v in vols)
    log(f"{len(vols)} volumes, {nvox / 1e6:.1f} M voxels ({time.perf_counter() - t0:.1f} s)")
    dev = torch.device("cuda", 0)
    b = tau.BatchSolver(vols, device=dev)
    b.solve()  # warm-up (graph capture, autotuned launches)
    torch.cuda.synchronize(dev)
    for ms, H, tile in autotune.last_timings:
        log(f"  autotune candidate H={H} tile {tile}: {ms:.4f} ms per half-sweep")
    t = b._plan["tile"]
    log(f"batch launch configuration: H={b._plan['H']}, {len(b._plan['schedule'])} passes per "
        f"100 iterations, tile {(t.waves, t.rows_per_wave, t.tx) if t is not None else 'default'}")
    t0 = time.perf_counter()
    b.reset()
    b.solve()
    torch.cuda.synchronize(dev)
    t_batch = time.perf_counter() - t0
    log(f"batch: {t_batch:.3f} s, iterations {int(b.iters.min())}..{int(b.iters.max())}")
    # sequential plain solves, volumes already resident on the device
    solvers = [tau.Solver(v, device=dev) for v in vols]
    for s in solvers[:4]:
        s.solve(verbose=False)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    d_seq, it_seq = [], []
    for s in solvers:
        s.reset()
        s.solve(verbose=False)
        d_seq.append(float(s.D_eff[0]))
        it_seq.append(s.iter)
    torch.cuda.synchronize(dev)
    t_seq = time.perf_counter() - t0
    d_seq = np.array(d_seq)
    rel = np.abs(b.D_eff.astype(np.float64) - d_seq) / np.maximum(np.abs(d_seq), 1e-30)
    same_iters = bool(np.array_equal(np.array(it_seq), b.iters))
    log(f"sequential: {t_seq:.3f} s; batch speed-up {t_seq / t_batch:.2f}x; max rel D_eff diff {rel.max():.2e}")
    vox_updates = sum(v.size * it / 2 for v, it in zip(vols, b.iters))
    _line("solves/s batched vs plain @ D_eff rel err<=1e-3", round(len(vols) / t_batch, 3), "solves/s",
          1, 1, round(t_batch * 1e3, 3),
          {"workload": "256 binary cubes, edge U[32,160], smoothed noise sigma U[1,5], porosity U[0.3,0.7], seed 0",
           "voxels": nvox},
          batch={"solves_per_s": round(len(vols) / t_batch, 3), "time_s": round(t_batch, 4),
                 "GVoxel_updates_per_s": round(vox_updates / t_batch / 1e9, 3)},
          sequential={"solves_per_s": round(len(vols) / t_seq, 3), "time_s": round(t_seq, 4)},
          speedup=round(t_seq / t_batch, 3), max_rel_D_eff_diff=float(rel.max()),
          same_stopping_iterations=same_iters)
    return 0


def bench_multiphase(args) -> int:
    """320^3 three-label volume: forward solve and per-phase gradient."""
    import torch

    import taufactor_amd as tau
    from taufactor_amd import workloads

    t0 = time.perf_counter()
    img = workloads.three_phase(320, seed=0)
    log(f"volume {img.shape} fractions {[(img == l).mean().round(3) for l in range(3)]} "
        f"({time.perf_counter() - t0:.1f} s)")
    dev = torch.device("cuda", 0)
    Ds = {0: 0.0, 1: 1.0, 2: 0.1}
    s = tau.MultiPhaseSolver(img, dict(Ds), device=dev)
    s.solve(verbose=False)
    launch = _log_launch(s)
    s._launch_events = []
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    s.reset()
    s.solve(verbose=False)
    torch.cuda.synchronize(dev)
    t_fwd = time.perf_counter() - t0
    ev, s._launch_events = s._launch_events, None
    t0 = time.perf_counter()
    g = s.grad_D_eff()
    torch.cuda.synchronize(dev)
    t_bwd = time.perf_counter() - t0
    nvox = img.size
    sweep_ms = sum(a.elapsed_time(b) for a, b, _ in ev) / sum(n for _, _, n in ev)
    gvus = nvox * s.iter / 2 / t_fwd / 1e9
    log(f"forward {t_fwd:.3f} s ({s.iter} iterations, D_eff {float(s.D_eff[0]):.6f}); "
        f"backward {t_bwd * 1e3:.2f} ms; dD_eff/dD {g[1][0]:.5f}, {g[2][0]:.5f}")
    _line("GVoxel-updates/s (MultiPhaseSolver forward) + backward/forward time", round(gvus, 3),
          "GVoxel-updates/s", 1, 1, round(t_fwd * 1e3, 3),
          {"workload": "MultiPhaseSolver 320^3, labels 0/1/2 at 0.4/0.3/0.3 from smoothed noise sigma 3, "
                       "D {0:0, 1:1.0, 2:0.1}, conv_crit 1e-2", "iterations": s.iter, **launch},
          D_eff=float(s.D_eff[0]), tau=float(s.tau[0]),
          grad_D_eff={"1": float(g[1][0]), "2": float(g[2][0])},
          backward_ms=round(t_bwd * 1e3, 3), backward_over_forward=round(t_bwd / t_fwd, 5),
          sweep_kernel_mean_ms=round(sweep_ms, 4),
          sweep_kernel_GBps_alg=round(9 * nvox / (sweep_ms / 1e3) / 1e9, 1))
    return 0


def bench_periodic(args) -> int:
    """PeriodicSolver on the 512^3 volume of the default line: whole solve."""
    import torch

    import taufactor_amd as tau
    from taufactor_amd import workloads

    img = workloads.smoothed_binary(512, 0.35, 4.0, 0)
    dev = torch.device("cuda", 0)
    s = tau.PeriodicSolver(img, device=dev)
    s.solve(verbose=False)  # autotune + warm-up
    H, tile, n_tiles = s.resolved_launch()
    times = []
    for _ in range(max(1, args.steps)):
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        s.reset()
        s.solve(verbose=False)
        torch.cuda.synchronize(dev)
        times.append(time.perf_counter() - t0)
    t = sum(times) / len(times)
    gvus = img.size * s.iter / 2 / t / 1e9
    log(f"{len(times)} solves, {s.iter} iterations, tau {float(s.tau[0]):.6f}, {t * 1e3:.1f} ms each")
    _line("GVoxel-updates/s (PeriodicSolver whole solve)", round(gvus, 3), "GVoxel-updates/s",
          len(times), 1, round(t * 1e3, 3),
          {"workload": "PeriodicSolver 512^3 binary volume, smoothed noise sigma 4, porosity 0.35, "
                       "seed 0, converged at conv_crit 1e-2", "iterations": s.iter,
           "half_sweeps_per_pass": H, "tile": list(tile), "tiles_per_launch": n_tiles},
          D_eff=float(s.D_eff[0]), tau=float(s.tau[0]))
    return 0


def bench_field(args) -> int:
    """64 volumes of 128^3 with per-voxel k: batched forward and per-voxel gradient."""
    import torch

    import taufactor_amd as tau
    from taufactor_amd import workloads

    t0 = time.perf_counter()
    k = workloads.sigmoid_k(64, 128, seed=0)
    log(f"k field {k.shape} ({time.perf_counter() - t0:.1f} s)")
    dev = torch.device("cuda", 0)
    kt = torch.from_numpy(k).to(dev)
    s = tau.FieldSolver(kt, device=dev)
    s.solve(verbose=False)
    launch = _log_launch(s)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    s.reset()
    s.solve(verbose=False)
    torch.cuda.synchronize(dev)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = s.grad_D_eff()
    torch.cuda.synchronize(dev)
    t_bwd = time.perf_counter() - t0
    nvox = k.size
    gvus = nvox * s.iter / 2 / t_fwd / 1e9
    log(f"forward {t_fwd:.3f} s ({s.iter} iterations); backward {t_bwd * 1e3:.2f} ms")
    _line("GVoxel-updates/s (per-voxel k forward, 64 volumes) + backward/forward time", round(gvus, 3),
          "GVoxel-updates/s", 1, 1, round(t_fwd * 1e3, 3),
          {"workload": "64 volumes of 128^3, k = sigmoid(4 z), z standardised smoothed noise sigma 3, "
                       "conv_crit 1e-2 (all volumes stop together, reference batch semantics)",
           "iterations": s.iter, **launch},
          D_eff_mean=float(s.D_eff.mean()), backward_ms=round(t_bwd * 1e3, 3),
          backward_over_forward=round(t_bwd / t_fwd, 5),
          grad_abs_mean=float(g.abs().mean()))
    return 0


if __name__ == "__main__":
    sys.exit(main())
