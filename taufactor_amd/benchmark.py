"""Convergence-study harness with the reference's interface and output
(taufactor/benchmark.py:1-244).

``run_benchmark_study`` solves synthetic structures of growing size with a
solver from this package and returns (and appends to a text file) one row
per case: size, structure, solver, device, criterion, total and solve wall
times, iterations, tau and the GPU memory figures the solver prints.  The
row format, the file header and the parsing of the solver's
``GPU-RAM currently ...`` line are the reference's, so tables produced here
line up with the reference's stored ones (docs/notebooks/12-solver-benchmark.ipynb).
"""
from __future__ import annotations

import gc
import inspect
import io
import itertools
import os
import time
from collections.abc import Callable
from contextlib import redirect_stdout

import torch

from . import solver as _solver
from .utils import create_2d_diagonals, create_2d_zigzag, create_3d_diagonals, create_fcc_cube, \
    create_stacked_blocks

DEFAULT_OUTFILE = "taufactor_benchmark_results.txt"


def _registry():
    import taufactor_amd as tau

    return {name: cls for name, cls in vars(tau).items()
            if name.endswith("Solver") and isinstance(cls, type) and not inspect.isabstract(cls)
            and issubclass(cls, _solver.SORSolver)}


# Solvers of the package namespace (taufactor/benchmark.py:22-26); BatchSolver
# takes a list of volumes and is not a harness solver.
SOLVER_REGISTRY = _registry()

STRUCTURE_REGISTRY = {
    "fcc": lambda N, features=None: (create_fcc_cube(N, overlap=0.05) == 0).astype(int),
    "blocks": create_stacked_blocks,
    "diagonal2d": create_2d_diagonals,
    "zigzag": create_2d_zigzag,
    "diagonal3d": create_3d_diagonals,
}


def resolve_solver(solver: str | type | None) -> type:
    """Solver class from a class or a registry name; None is PeriodicSolver
    (taufactor/benchmark.py:37-50)."""
    if solver is None:
        return _solver.PeriodicSolver
    if isinstance(solver, str):
        if solver not in SOLVER_REGISTRY:
            available = ", ".join(sorted(SOLVER_REGISTRY))
            raise ValueError(f"Unknown solver '{solver}'. Available solvers: {available}")
        return SOLVER_REGISTRY[solver]
    if isinstance(solver, type):
        return solver
    raise TypeError("solver must be None, a solver class, or a solver name string")


def _call_structure_hook(structure_fn: Callable, N: int, features: int | None):
    """Call a custom structure hook with N or Nx, with or without features
    (taufactor/benchmark.py:52-73)."""
    attempts = [
        lambda: structure_fn(N=N, features=features),
        lambda: structure_fn(Nx=N, features=features),
        lambda: structure_fn(N, features=features),
        lambda: structure_fn(N),
    ]
    last = None
    for attempt in attempts:
        try:
            return attempt()
        except TypeError as exc:
            last = exc
    raise TypeError("Unable to call custom structure hook. Expected a callable that accepts "
                    "N or Nx (optionally features).") from last


def resolve_structure(structure: str | Callable, N: int, features: int | None = None) -> tuple:
    """(volume, name) from a registry name or a callable (taufactor/benchmark.py:76-96)."""
    if isinstance(structure, str):
        if structure not in STRUCTURE_REGISTRY:
            available = ", ".join(sorted(STRUCTURE_REGISTRY))
            raise ValueError(f"Unknown structure '{structure}'. Supported: {available}")
        return STRUCTURE_REGISTRY[structure](N, features=features), structure
    if callable(structure):
        cube = _call_structure_hook(structure, N=N, features=features)
        return cube, getattr(structure, "__name__", structure.__class__.__name__)
    raise TypeError("structure must be a predefined structure name or a callable hook")


def write_header_if_missing(outfile: str = DEFAULT_OUTFILE) -> None:
    """Create the results file with the reference's header (taufactor/benchmark.py:99-108)."""
    if not os.path.exists(outfile):
        with open(outfile, "w", encoding="utf-8") as f:
            f.write(f"{'N':>4} {'struct':>10} {'solver':>16} {'dev':>4} {'conv':>6} "
                    f"{'Ttime(s)':>9} {'Wtime(s)':>9} {'iters':>6} {'tau':>8} "
                    f"{'VRAM(cur)':>10} {'VRAM(max)':>10} {'VRAM(res)':>10}\n")
            f.write("=" * 120 + "\n")


def format_row(row: dict) -> str:
    """One results line (taufactor/benchmark.py:111-120)."""
    return (f"{row['N']:4d} {row['structure'][:10]:>10} {row['solver'][:16]:>16} {row['device'][:4]:>4} "
            f"{row['conv_crit']:.4f} {row['total_time']:9.3f} {row['solve_time']:9.3f} "
            f"{row['iterations']:6d} {row['taufactor']:8.3f} {row['torch_cur']:10.2f} "
            f"{row['torch_max']:10.2f} {row['torch_res']:10.2f}\n")


def append_row_to_file(row: dict, outfile: str = DEFAULT_OUTFILE) -> None:
    with open(outfile, "a", encoding="utf-8") as f:
        f.write(format_row(row))


def parse_gpu_ram(lines) -> tuple[float, float, float]:
    """(current, max allocated, reserved) MB from the solver's GPU-RAM line,
    tokenised as the reference does (taufactor/benchmark.py:170-179)."""
    line = next((ln for ln in lines if "GPU-RAM" in ln), "")
    if not line:
        return 0.0, 0.0, 0.0
    parts = line.replace("(", "").replace(")", "").replace(",", "").split()
    return float(parts[2]), float(parts[6]), float(parts[8])


def run_benchmark_case(N: int, device: str, conv_crit: float, structure: str | Callable = "fcc",
                       features: int | None = None, iter_limit: int = 10000,
                       solver: str | type | None = None, solver_kwargs: dict | None = None,
                       solve_kwargs: dict | None = None) -> dict:
    """Solve one structure and return its row (taufactor/benchmark.py:123-193)."""
    cube, structure_name = resolve_structure(structure, N=N, features=features)
    solver_cls = resolve_solver(solver)
    solver_kwargs = dict(solver_kwargs or {})
    solve_kwargs = dict(solve_kwargs or {})
    if device == "cuda":
        torch.cuda.empty_cache()
    gc.collect()
    if device == "cuda":
        torch.cuda.synchronize()
    start = time.perf_counter()
    buf = io.StringIO()
    with redirect_stdout(buf):
        s = solver_cls(cube, device=device, **solver_kwargs)
        s.solve(iter_limit=iter_limit, conv_crit=conv_crit, **solve_kwargs)
        if device == "cuda":
            torch.cuda.synchronize()
        end = time.perf_counter()
    cur, mx, res = parse_gpu_ram(buf.getvalue().splitlines())
    return {
        "N": N, "structure": structure_name, "solver": solver_cls.__name__, "device": device,
        "conv_crit": conv_crit, "total_time": end - start, "solve_time": float(s.walltime),
        "iterations": int(s.iter), "taufactor": float(s.tau[0]),
        "torch_cur": cur, "torch_max": mx, "torch_res": res,
    }


def run_benchmark_study(Ns=(100, 128, 200, 256, 300, 384, 400), devices=("cuda",),
                        conv_crit_values=(1e-3,), structure: str | Callable = "fcc", features: int = 1,
                        outfile: str = DEFAULT_OUTFILE, write_file: bool = True, iter_limit: int = 10000,
                        solver: str | type | None = None, solver_kwargs: dict | None = None,
                        solve_kwargs: dict | None = None) -> list[dict]:
    """Rows for every (N, device, conv_crit) (taufactor/benchmark.py:196-244)."""
    rows = []
    if write_file:
        write_header_if_missing(outfile=outfile)
    for N, device, conv_crit in itertools.product(Ns, devices, conv_crit_values):
        if device == "cuda" and not torch.cuda.is_available():
            print(f"Skipping N={N} on CUDA (not available)")
            continue
        row = run_benchmark_case(N=N, device=device, conv_crit=conv_crit, structure=structure,
                                 features=features, iter_limit=iter_limit, solver=solver,
                                 solver_kwargs=solver_kwargs, solve_kwargs=solve_kwargs)
        rows.append(row)
        if write_file:
            append_row_to_file(row, outfile=outfile)
    return rows
