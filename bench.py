"""Benchmark of the MI355X steady-state diffusion solve.

Default run (one GPU, one JSON line on stdout):

  headline  the 512^3 plain-Solver volume (seeded standard-normal noise,
            Gaussian-smoothed with sigma 4, conductive below the 0.35
            quantile) solved to the reference's default criterion (conv_crit
            1e-2, checks every 100 iterations).  One step = one complete
            solve from the initial field with the volume already resident in
            HBM (Solver.reset() copies the initial field back on the device).
            value = GVoxel-updates/s = voxels x (iterations/2) x steps / time
            (one voxel-update = one voxel through both colours).
  configs   the other workloads of the benchmark table, each measured the
            same way and reported in a sub-object:
            C3  256 binary cubes of edge U[32,160]: one BatchSolver against
                256 sequential Solver solves (solves/s, speed-up, largest
                relative D_eff difference, same stopping iterations)
            C4  MultiPhaseSolver 320^3 {0:0, 1:1, 2:0.1}: forward at conv
                1e-2, and the per-phase gradient after a conv 1e-3 forward
                (1e-4 is below what the float32 field resolves at this size)
            C5  64 volumes of 128^3, per-voxel k = sigmoid(4 z): batched
                forward at conv 1e-2, and the per-voxel gradient after a
                conv 1e-4 forward
            C1  100^3 Bernoulli(0.5) volume, whole solve (cache-resident:
                reported as throughput only)

roofline (every sweep kernel, one definition): algorithmic bytes per launch =
  B_u x voxels x H/2 (B_u = 9 B per voxel-update for the one-byte-code kinds:
  4 B field read + 4 B field write + 1 B code; 12 B for per-voxel k; H fused
  half-sweeps = H/2 voxel-updates per voxel), divided by the kernel's mean
  launch duration, measured with HIP events on the launch stream over the
  timed solves (full-depth launches only), against 8 TB/s.  This is an
  effective bandwidth: temporal blocking lets it exceed the bytes actually
  moved.  `traffic` is the measured HBM bytes per launch (rocprofv3
  FETCH_SIZE + WRITE_SIZE, calibrated) from the committed summary whose
  kernel-source hash equals this build's; `dram_frac` = traffic / launch time
  / 8 TB/s is the fraction of HBM peak actually moved.

cpu_baseline: the CPU oracle (an op-for-op torch restatement of the
  reference loop) on this host: 20 iterations + one check of the 512^3 volume
  (value), plus legs for C1 (full solve), C3 (a subset of the volumes solved,
  extrapolated by voxel-iterations), C4 and C5 (a few iterations,
  extrapolated to the GPU's iteration count).

Single workloads (one line each, for profiling):
  --config solver512 | batch256 | multiphase320 | field128x64 | periodic512 | hbm_copy

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--config NAME]
"""
from __future__ import annotations

import argparse
import contextlib
import hashlib
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "GVoxel-updates/s + % HBM roofline; solves/s batched vs plain @ D_eff rel err≤1e-3"
HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {"code": 9, "field": 12}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sweep_traffic.json")
# the compilation unit of each workload's sweep kernel (its traffic entry is
# keyed to the hash of that unit, the shared sweep header and the flags)
SWEEP_UNITS = {"C2": "sweep_solver.hip", "C4": "sweep_mpfaces.hip", "C5": "sweep_field.hip"}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sweep_source_sha256(unit: str = "sweep_solver.hip") -> str:
    """Hash of what a sweep kernel is built from: the shared sweep header,
    the kind's compilation unit and the compiler flags.  Keys the committed
    rocprof traffic summary to a build."""
    from taufactor_amd.csrc import build

    h = hashlib.sha256()
    for name in ("tauhip_sweep.h", unit):
        with open(os.path.join(build.HERE, name), "rb") as f:
            h.update(f.read())
    h.update(" ".join(build.HIPCC_FLAGS).encode())
    return h.hexdigest()


def roofline(kind: str, nvox: int, H: int, mean_launch_s: float, kernel: str) -> dict:
    alg = BYTES_PER_UPDATE[kind] * nvox * H / 2
    achieved = alg / mean_launch_s / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "frac_kind": "effective",
            "bytes_per_update": BYTES_PER_UPDATE[kind], "bytes_per_launch_algorithmic": int(alg),
            "mean_launch_ms": round(mean_launch_s * 1e3, 4), "traffic": None}


def attach_traffic(rl: dict, workload: str, H: int, tile, mean_launch_s: float) -> None:
    """Measured HBM bytes per launch from the committed summary of this
    kernel build, when there is one for this workload, depth and tile."""
    if not os.path.exists(TRAFFIC_FILE):
        rl["traffic_note"] = "no traffic summary committed"
        return
    with open(TRAFFIC_FILE) as f:
        entries = json.load(f)
    src = sweep_source_sha256(SWEEP_UNITS[workload])
    for t in entries:
        if (t.get("workload") == workload and t.get("half_sweeps_per_pass") == H and t.get("tile") == list(tile)
                and t.get("source_sha256") == src):
            b = t["hbm_bytes_per_launch"]
            rl["traffic"] = b
            rl["dram_GBps"] = round(b / mean_launch_s / 1e9, 1)
            rl["dram_frac"] = round(b / mean_launch_s / 1e9 / HBM_PEAK_GBS, 4)
            rl["traffic_source"] = (f"rocprofv3 FETCH_SIZE x {t['fetch_factor']} + WRITE_SIZE x "
                                    f"{t['write_factor']} per launch of {t['kernel']}, profiles/{t['profile']}, "
                                    "same kernel-source hash")
            return
    rl["traffic_note"] = (f"no traffic summary for this kernel build (source sha256 {src[:16]}), H={H}, "
                          f"tile {list(tile)}")


def kernel_mean_s(events) -> float:
    ms = sum(a.elapsed_time(b) for a, b, _ in events)
    return ms / 1e3 / sum(n for _, _, n in events)


def host_info() -> dict:
    import torch

    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "affinity_cores": len(os.sched_getaffinity(0)),
            "torch_threads": torch.get_num_threads()}


# ---------------------------------------------------------------------------
# GPU workloads
# ---------------------------------------------------------------------------
def run_solver512(args, dev):
    """C2: whole solves of the 512^3 volume; returns (line fields, img, H)."""
    import torch

    import taufactor_amd as tau
    from taufactor_amd import autotune, workloads

    t0 = time.perf_counter()
    img = workloads.smoothed_binary(args.size, 0.35, 4.0, 0)
    log(f"volume {img.shape} porosity {img.mean():.4f} sha256 {workloads.sha256(img)[:16]} "
        f"({time.perf_counter() - t0:.1f} s)")
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
    log(f"launch configuration: {H} half-sweeps per launch, {rtile[0]}x{rtile[1]} waves x rows, "
        f"x tile {rtile[2]}, {n_tiles} tiles ({time.perf_counter() - t0:.2f} s)")
    for w in range(args.warmup):
        s.reset()
        s.solve(verbose=False)
        log(f"warmup {w}: {s.iter} iterations, tau {float(s.tau[0]):.6f}")
    s._launch_events = []
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        s.reset()
        s.solve(verbose=False)
        iters.append(s.iter)
    torch.cuda.synchronize(dev)
    elapsed = time.perf_counter() - t0
    mean_s = kernel_mean_s(s._launch_events)
    s._launch_events = None
    nvox = img.size
    gvus = nvox * sum(iters) / 2 / elapsed / 1e9
    rl = roofline("code", nvox, H, mean_s, f"k_sweep<{H}, Solver, {rtile[0]}, {rtile[1]}>")
    attach_traffic(rl, "C2", H, rtile, mean_s)
    log(f"{args.steps} solves, iterations {iters}, {elapsed:.3f} s -> {gvus:.1f} GVoxel-updates/s; "
        f"sweep kernel {mean_s * 1e3:.4f} ms -> effective frac {rl['frac']}, dram frac {rl.get('dram_frac')}")
    out = {
        "value": round(gvus, 3),
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "config": {
            "workload": f"Solver {args.size}^3 binary volume, smoothed noise sigma 4, porosity 0.35, "
                        "seed 0, converged at conv_crit 1e-2",
            "shape": list(img.shape), "iterations": iters[-1], "tau": float(s.tau[0]),
            "D_eff": float(s.D_eff[0]), "half_sweeps_per_pass": H, "tile": list(rtile),
            "tiles_per_launch": n_tiles,
            "launch_config": "autotuned" if args.half_sweeps is None or args.tile is None else "given",
            "volume_sha256": workloads.sha256(img),
        },
        "roofline": rl,
    }
    del s
    torch.cuda.empty_cache()
    return out, img


def run_batch256(dev) -> dict:
    """C3: one BatchSolver against 256 sequential Solver solves."""
    import numpy as np
    import torch

    import taufactor_amd as tau
    from taufactor_amd import workloads

    t0 = time.perf_counter()
    vols = [v for v, _, _ in workloads.uneven_batch(256, seed=0)]
    nvox = sum(v.size for v in vols)
    log(f"C3: {len(vols)} volumes, {nvox / 1e6:.1f} M voxels ({time.perf_counter() - t0:.1f} s)")
    b = tau.BatchSolver(vols, device=dev)
    b.solve()  # warm-up: autotune, graph capture
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    b.reset()
    b.solve()
    torch.cuda.synchronize(dev)
    t_batch = time.perf_counter() - t0
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
    same = bool(np.array_equal(np.array(it_seq), b.iters))
    updates = float(sum(v.size * it / 2 for v, it in zip(vols, b.iters)))
    plan = b._plan
    t = plan["tile"]
    log(f"C3: batch {t_batch:.3f} s, sequential {t_seq:.3f} s -> {t_seq / t_batch:.2f}x; "
        f"max rel D_eff diff {rel.max():.2e}, same iterations {same}")
    out = {
        "workload": "256 binary cubes, edge U[32,160], smoothed noise sigma U[1,5], porosity U[0.3,0.7], "
                    "seed 0, conv_crit 1e-2, each volume stopping at its own check",
        "voxels": nvox, "batched_solves_per_s": round(len(vols) / t_batch, 3),
        "sequential_solves_per_s": round(len(vols) / t_seq, 3), "batched_s": round(t_batch, 4),
        "sequential_s": round(t_seq, 4), "speedup": round(t_seq / t_batch, 3),
        "max_rel_D_eff_diff": float(rel.max()), "same_stopping_iterations": same,
        "iterations_min_max": [int(b.iters.min()), int(b.iters.max())],
        "batched_GVoxel_updates_per_s": round(updates / t_batch / 1e9, 3),
        "half_sweeps_per_pass": plan["H"],
        "tile": None if t is None else [t.waves, t.rows_per_wave, t.tx],
        "_iters": b.iters.tolist(),
    }
    del b, solvers
    torch.cuda.empty_cache()
    return out


def _with_launch(s, launch):
    """s with the launch configuration (H, tile) set, if one is given."""
    if launch:
        s.half_sweeps_per_pass, s.tile = launch
    return s


def _grad_timing(make, forward_crit: float, grad):
    """Forward at forward_crit, then one gradient; (solver, forward s, backward s)."""
    import torch

    s = make()
    dev = s.device
    s.solve(verbose=False, conv_crit=forward_crit, iter_limit=100000)  # warm-up and autotune
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    s.reset()
    s.solve(verbose=False, conv_crit=forward_crit, iter_limit=100000)
    torch.cuda.synchronize(dev)
    t_fwd = time.perf_counter() - t0
    grad(s)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    g = grad(s)
    torch.cuda.synchronize(dev)
    t_bwd = time.perf_counter() - t0
    return s, g, t_fwd, t_bwd


def run_multiphase320(dev, launch=None) -> dict:
    """C4: MultiPhaseSolver 320^3 forward; per-phase gradient.  launch: (H,
    tile) of the forward instead of the autotuner's choice."""
    import torch

    import taufactor_amd as tau
    from taufactor_amd import workloads

    t0 = time.perf_counter()
    img = workloads.three_phase(320, seed=0)
    log(f"C4: volume {img.shape} ({time.perf_counter() - t0:.1f} s)")
    Ds = {0: 0.0, 1: 1.0, 2: 0.1}
    s = _with_launch(tau.MultiPhaseSolver(img, dict(Ds), device=dev), launch)
    s.solve(verbose=False)
    H, tile, n_tiles = s.resolved_launch()
    s._launch_events = []
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    s.reset()
    s.solve(verbose=False)
    torch.cuda.synchronize(dev)
    t_fwd = time.perf_counter() - t0
    mean_s = kernel_mean_s(s._launch_events)
    s._launch_events = None
    nvox = img.size
    gvus = nvox * s.iter / 2 / t_fwd / 1e9
    fwd = {"iterations": s.iter, "tau": float(s.tau[0]), "D_eff": float(s.D_eff[0]), "ms": round(t_fwd * 1e3, 3),
           "GVoxel_updates_per_s": round(gvus, 3)}
    faces = s._faces is not None
    rl = roofline("code", nvox, H, mean_s,
                  f"k_sweep<{H}, {'MultiPhaseFaces' if faces else 'MultiPhase'}, {tile[0]}, {tile[1]}>")
    if faces:
        attach_traffic(rl, "C4", H, tile, mean_s)
    del s
    s, g, t_f, t_b = _grad_timing(lambda: _with_launch(tau.MultiPhaseSolver(img, dict(Ds), device=dev), launch),
                                  1e-3, lambda x: x.grad_D_eff())
    log(f"C4: forward {t_fwd * 1e3:.1f} ms ({fwd['iterations']} iterations, {gvus:.1f} GVU/s, kernel "
        f"{mean_s * 1e3:.4f} ms, frac {rl['frac']}); conv 1e-3 forward {t_f * 1e3:.1f} ms, "
        f"gradient {t_b * 1e3:.2f} ms")
    out = {
        "workload": "MultiPhaseSolver 320^3, labels 0/1/2 at 0.4/0.3/0.3 from smoothed noise sigma 3, "
                    "D {0:0, 1:1.0, 2:0.1}",
        "forward_conv_1e-2": fwd, "roofline": rl,
        "half_sweeps_per_pass": H, "tile": list(tile), "tiles_per_launch": n_tiles,
        "gradient": {"forward_conv_crit": 1e-3, "forward_iterations": s.iter, "forward_converged": bool(s.converged),
                     "forward_ms": round(t_f * 1e3, 3),
                     "backward_ms": round(t_b * 1e3, 3), "backward_over_forward": round(t_b / t_f, 5),
                     "dD_eff_dD": {"1": float(g[1][0]), "2": float(g[2][0])}},
    }
    del s
    torch.cuda.empty_cache()
    return out


def run_field128x64(dev, launch=None) -> dict:
    """C5: 64 volumes of 128^3 with per-voxel k, batched forward; per-voxel gradient."""
    import torch

    import taufactor_amd as tau
    from taufactor_amd import workloads

    t0 = time.perf_counter()
    k = workloads.sigmoid_k(64, 128, seed=0)
    log(f"C5: k {k.shape} ({time.perf_counter() - t0:.1f} s)")
    kt = torch.from_numpy(k).to(dev)
    s = _with_launch(tau.FieldSolver(kt, device=dev), launch)
    s.solve(verbose=False)
    H, tile, n_tiles = s.resolved_launch()
    s._launch_events = []
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    s.reset()
    s.solve(verbose=False)
    torch.cuda.synchronize(dev)
    t_fwd = time.perf_counter() - t0
    mean_s = kernel_mean_s(s._launch_events)
    s._launch_events = None
    nvox = k.size
    gvus = nvox * s.iter / 2 / t_fwd / 1e9
    fwd = {"iterations": s.iter, "D_eff_mean": float(s.D_eff.mean()), "ms": round(t_fwd * 1e3, 3),
           "GVoxel_updates_per_s": round(gvus, 3)}
    rl = roofline("field", nvox, H, mean_s, f"k_sweep<{H}, Field, {tile[0]}, {tile[1]}>")
    attach_traffic(rl, "C5", H, tile, mean_s)
    del s
    s, g, t_f, t_b = _grad_timing(lambda: _with_launch(tau.FieldSolver(kt, device=dev), launch), 1e-4,
                                  lambda x: x.grad_D_eff())
    log(f"C5: forward {t_fwd * 1e3:.1f} ms ({fwd['iterations']} iterations, {gvus:.1f} GVU/s, kernel "
        f"{mean_s * 1e3:.4f} ms, frac {rl['frac']}); conv 1e-4 forward {t_f * 1e3:.1f} ms, "
        f"gradient {t_b * 1e3:.2f} ms")
    out = {
        "workload": "64 volumes of 128^3, k = sigmoid(4 z), z standardised smoothed noise sigma 3 "
                    "(all volumes stop together, the reference's batch semantics)",
        "forward_conv_1e-2": fwd, "roofline": rl,
        "half_sweeps_per_pass": H, "tile": list(tile), "tiles_per_launch": n_tiles,
        "gradient": {"forward_conv_crit": 1e-4, "forward_iterations": s.iter, "forward_converged": bool(s.converged),
                     "forward_ms": round(t_f * 1e3, 3),
                     "backward_ms": round(t_b * 1e3, 3), "backward_over_forward": round(t_b / t_f, 5),
                     "grad_abs_mean": float(g.abs().mean())},
    }
    del s, kt, g
    torch.cuda.empty_cache()
    return out


def run_plain100(dev) -> dict:
    """C1: the 100^3 Bernoulli volume, whole solve."""
    import torch

    import taufactor_amd as tau
    from taufactor_amd import workloads

    img = workloads.plain_100(0)
    s = tau.Solver(img, device=dev)
    s.solve(verbose=False)
    torch.cuda.synchronize(dev)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        s.reset()
        s.solve(verbose=False)
        torch.cuda.synchronize(dev)
        times.append(time.perf_counter() - t0)
    t = min(times)
    return {"workload": "Solver 100^3, voxels Bernoulli(0.5) from seed 0, conv_crit 1e-2",
            "iterations": s.iter, "tau": float(s.tau[0]), "D_eff": float(s.D_eff[0]), "ms": round(t * 1e3, 3),
            "GVoxel_updates_per_s": round(img.size * s.iter / 2 / t / 1e9, 3)}


# ---------------------------------------------------------------------------
# CPU oracle legs
# ---------------------------------------------------------------------------
def _timed_oracle(o, n_iter: int) -> float:
    import torch

    t0 = time.perf_counter()
    with torch.no_grad():
        o.run(n_iter)
        o.check_convergence(1e-2)
    return time.perf_counter() - t0


def cpu_baseline(img, n_iter: int, configs: dict) -> dict:
    """The oracle on this host.  value: n_iter iterations + one check of the
    512^3 volume; legs for C1 (full solve), C3, C4 and C5."""
    from oracle.restate import OracleFieldSolver, OracleMultiPhaseSolver, OracleSolver
    from taufactor_amd import workloads

    t0 = time.perf_counter()
    o = OracleSolver(img)
    t_setup = time.perf_counter() - t0
    dt = _timed_oracle(o, n_iter)
    del o
    out = {
        "value": round(img.size * (n_iter / 2) / dt / 1e9, 4), "unit": "GVoxel-updates/s",
        "kind": "port",
        "sample": (f"{n_iter} iterations + 1 convergence check of the same {img.shape[0]}^3 volume "
                   f"on the host ({dt:.1f} s; set-up {t_setup:.1f} s not timed)"),
        **host_info(),
    }
    out["cores"] = out["torch_threads"]
    legs = {}
    # C1: a full solve
    img1 = workloads.plain_100(0)
    t0 = time.perf_counter()
    o = OracleSolver(img1).solve()
    t = time.perf_counter() - t0
    legs["C1"] = {"sample": "full solve of the 100^3 volume", "iterations": int(o.iter), "tau": float(o.tau[0]),
                  "s": round(t, 3), "GVoxel_updates_per_s": round(img1.size * o.iter / 2 / t / 1e9, 4)}
    # C3: volumes solved in full in batch order until 15 s are spent,
    # extrapolated to all 256 by voxel-updates
    c3 = configs.get("C3")
    if c3:
        vols = [v for v, _, _ in workloads.uneven_batch(256, seed=0)]
        n_sub = 0
        t0 = time.perf_counter()
        sub_updates = 0.0
        while n_sub < len(vols) and (n_sub == 0 or time.perf_counter() - t0 < 15.0):
            o = OracleSolver(vols[n_sub]).solve()
            sub_updates += vols[n_sub].size * o.iter / 2
            n_sub += 1
        t = time.perf_counter() - t0
        all_updates = float(sum(v.size * it / 2 for v, it in zip(vols, c3["_iters"])))
        t_all = t * all_updates / sub_updates
        legs["C3"] = {"sample": f"the first {n_sub} of the 256 volumes solved in full ({t:.1f} s), "
                                "extrapolated to all 256 by voxel-updates (GPU iteration counts)",
                      "solves_per_s_extrapolated": round(256 / t_all, 4), "s_extrapolated": round(t_all, 1)}
    c4 = configs.get("C4")
    if c4:
        lab = workloads.three_phase(320, seed=0)
        o = OracleMultiPhaseSolver(lab, {0: 0.0, 1: 1.0, 2: 0.1})
        n = 10
        t = _timed_oracle(o, n)
        it = c4["forward_conv_1e-2"]["iterations"]
        legs["C4"] = {"sample": f"{n} iterations + 1 check of the 320^3 forward, extrapolated to the GPU's "
                                f"{it} iterations", "s_per_iteration": round(t / n, 4),
                      "forward_s_extrapolated": round(t / n * it, 1),
                      "GVoxel_updates_per_s": round(lab.size * n / 2 / t / 1e9, 4)}
        del o, lab
    c5 = configs.get("C5")
    if c5:
        k = workloads.sigmoid_k(64, 128, seed=0)
        o = OracleFieldSolver(k)
        n = 4
        t = _timed_oracle(o, n)
        it = c5["forward_conv_1e-2"]["iterations"]
        legs["C5"] = {"sample": f"{n} iterations + 1 check of the 64 x 128^3 forward, extrapolated to the "
                                f"GPU's {it} iterations", "s_per_iteration": round(t / n, 4),
                      "forward_s_extrapolated": round(t / n * it, 1),
                      "GVoxel_updates_per_s": round(k.size * n / 2 / t / 1e9, 4)}
        del o, k
    out["legs"] = legs
    return out


def run_default(args, dev) -> dict:
    """The default line: the 512^3 headline, the other configurations and the CPU legs."""
    t_start = time.perf_counter()
    head, img = run_solver512(args, dev)
    configs = {}
    if args.config is None and not args.headline_only:
        configs["C1"] = run_plain100(dev)
        configs["C3"] = run_batch256(dev)
        configs["C4"] = run_multiphase320(dev)
        configs["C5"] = run_field128x64(dev)
    out = {
        "metric": METRIC, "value": head["value"], "unit": "GVoxel-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded smoothed-noise volumes, generated on the host)",
        "config": head["config"], "roofline": head["roofline"], "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log(f"cpu baseline: {args.cpu_iters} oracle iterations at {args.size}^3 and the config legs ...")
        out["cpu_baseline"] = cpu_baseline(img, args.cpu_iters, configs)
        log(f"cpu baseline {out['cpu_baseline']['value']} GVoxel-updates/s")
    if configs:
        configs["C3"].pop("_iters", None)
        out["configs"] = configs
    log(f"total {time.perf_counter() - t_start:.1f} s")
    return out


# ---------------------------------------------------------------------------
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
    ap.add_argument("--headline-only", action="store_true", help="the 512^3 line without the other configs")
    ap.add_argument("--config", default=None,
                    choices=["solver512", "batch256", "multiphase320", "field128x64", "periodic512", "hbm_copy"])
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the solve has no multi-GPU path (it would "
              "shard as x slabs with H-plane ghost slabs exchanged per pass); run with --gpus 1",
              file=sys.stderr)
        return 2

    import torch

    dev = torch.device("cuda", 0)
    single = {"batch256": run_batch256, "multiphase320": run_multiphase320, "field128x64": run_field128x64}
    if args.config in single:
        kw = {}
        if args.config != "batch256" and args.half_sweeps and args.tile:
            # a fixed launch configuration (profiling runs: the autotuner's
            # timings under counter collection need not pick the default's)
            kw["launch"] = (args.half_sweeps, tuple(int(v) for v in args.tile.split(",")))
        with contextlib.redirect_stdout(sys.stderr):
            r = single[args.config](dev, **kw)
        r.pop("_iters", None)
        print(json.dumps({"config": args.config, **r}), flush=True)
        return 0
    if args.config == "periodic512":
        return bench_periodic(args, dev)
    if args.config == "hbm_copy":
        return bench_copy(args, dev)

    # the solvers' own messages go to stderr: stdout carries the one JSON line
    with contextlib.redirect_stdout(sys.stderr):
        out = run_default(args, dev)
    print(json.dumps(out), flush=True)
    return 0


def bench_copy(args, dev) -> int:
    """Copy of a --size^3 float32 field on the device, timed with HIP events
    on the launch stream: the library's streaming copy kernel (tf_hbm_copy,
    the device's copy rate and the traffic-counter calibration) and torch's
    copy (the runtime's copy kernel)."""
    import torch

    from taufactor_amd import _lib

    lib = _lib.load()
    n = args.size ** 3
    src = torch.rand(n, device=dev)
    dst = torch.empty_like(src)
    stream = _lib.stream_handle(dev)
    reps = 20 * args.steps

    def timed(copy):
        for _ in range(max(args.warmup, 1)):
            copy()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize(dev)
        a.record()
        for _ in range(reps):
            copy()
        b.record()
        torch.cuda.synchronize(dev)
        return a.elapsed_time(b) / reps

    def kernel_copy():
        _lib.check(lib.tf_hbm_copy(_lib.ptr(src), _lib.ptr(dst), 4 * n, stream), "tf_hbm_copy")

    out = {"config": "hbm_copy", "bytes_read": 4 * n, "bytes_written": 4 * n}
    for name, copy in (("tf_hbm_copy", kernel_copy), ("torch_copy", lambda: dst.copy_(src))):
        dst.zero_()
        ms = timed(copy)
        assert torch.equal(dst, src)
        gbs = 2 * 4 * n / (ms / 1e3) / 1e9
        log(f"{name}: copy {4 * n / 1e6:.0f} MB: {ms:.4f} ms -> {gbs:.0f} GB/s (read + write)")
        out[name] = {"ms": round(ms, 5), "GBps": round(gbs, 1), "frac_of_8TBps": round(gbs / HBM_PEAK_GBS, 4)}
    print(json.dumps(out), flush=True)
    return 0


def bench_periodic(args, dev) -> int:
    """PeriodicSolver on the 512^3 volume of the default line: whole solve."""
    import torch

    import taufactor_amd as tau
    from taufactor_amd import workloads

    img = workloads.smoothed_binary(512, 0.35, 4.0, 0)
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
    print(json.dumps({"config": "periodic512", "GVoxel_updates_per_s": round(gvus, 3), "ms": round(t * 1e3, 3),
                      "iterations": s.iter, "tau": float(s.tau[0]), "half_sweeps_per_pass": H,
                      "tile": list(tile), "tiles_per_launch": n_tiles}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
