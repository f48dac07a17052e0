"""Convergence studies: solve synthetic structures of growing size with one of
this package's solvers and tabulate iterations, tau, wall times and the GPU
memory the solver reports.

What matches the reference (taufactor/benchmark.py), because tables made
here must line up with its stored ones (docs/notebooks/12-solver-benchmark.ipynb)
and scripts written against it must keep working:
  * the public names: DEFAULT_OUTFILE, SOLVER_REGISTRY, STRUCTURE_REGISTRY
    and its keys, resolve_solver, resolve_structure, write_header_if_missing,
    append_row_to_file, run_benchmark_case, run_benchmark_study, with their
    arguments, defaults, row keys and error types;
  * the text table: header and row widths (:99-120);
  * the memory figures, read from the solver's printed
    "GPU-RAM currently X MB (max allocated Y MB; Z MB reserved)" line
    (:170-179);
  * how a custom structure hook is called: by N or by Nx, with `features` when
    it takes it (:52-73).

How it is built here: a study is a list of `Case`s (one per size, device and
criterion), each run by `Case.run()` into a `Row` that renders its own table
line; a hook's call is worked out from its signature instead of by trial
calls, so a TypeError raised inside a hook reaches the caller unchanged.
"""
from __future__ import annotations

import gc
import inspect
import io
import re
import time
from collections.abc import Callable, Iterable
from contextlib import redirect_stdout
from dataclasses import asdict, dataclass, field

import torch

from . import solver as _solver
from .utils import create_2d_diagonals, create_2d_zigzag, create_3d_diagonals, create_fcc_cube, \
    create_stacked_blocks

DEFAULT_OUTFILE = "taufactor_benchmark_results.txt"

# table layout: (header, width) per column, then the rule under the header
_COLUMNS = (("N", 4), ("struct", 10), ("solver", 16), ("dev", 4), ("conv", 6), ("Ttime(s)", 9),
            ("Wtime(s)", 9), ("iters", 6), ("tau", 8), ("VRAM(cur)", 10), ("VRAM(max)", 10),
            ("VRAM(res)", 10))
_RULE = "=" * 120

_RAM_LINE = re.compile(r"GPU-RAM currently\s+(\S+)\s+MB\s+\(max allocated\s+(\S+)\s+MB;\s+(\S+)\s+MB reserved\)")


def _solver_classes() -> dict:
    """Every concrete solver class exported by the package (the names a study
    may ask for); BatchSolver takes a list of volumes and is not one."""
    import taufactor_amd

    out = {}
    for name, obj in vars(taufactor_amd).items():
        if (name.endswith("Solver") and inspect.isclass(obj) and issubclass(obj, _solver.SORSolver)
                and not inspect.isabstract(obj)):
            out[name] = obj
    return out


SOLVER_REGISTRY = _solver_classes()


def _fcc_pores(N, features=None):
    """Pore space of an fcc packing of spheres overlapping by 5 %."""
    return (create_fcc_cube(N, overlap=0.05) == 0).astype(int)


STRUCTURE_REGISTRY = {
    "fcc": _fcc_pores,
    "blocks": create_stacked_blocks,
    "diagonal2d": create_2d_diagonals,
    "zigzag": create_2d_zigzag,
    "diagonal3d": create_3d_diagonals,
}


def resolve_solver(solver: str | type | None) -> type:
    """A solver class from a class, a registry name, or None (PeriodicSolver,
    the reference's default)."""
    if solver is None:
        return _solver.PeriodicSolver
    if isinstance(solver, type):
        return solver
    if not isinstance(solver, str):
        raise TypeError("solver must be None, a solver class, or a solver name string")
    try:
        return SOLVER_REGISTRY[solver]
    except KeyError:
        raise ValueError(f"Unknown solver '{solver}'. Available solvers: "
                         f"{', '.join(sorted(SOLVER_REGISTRY))}") from None


def _hook_arguments(fn: Callable, N: int, features) -> tuple[tuple, dict] | None:
    """The (args, kwargs) to call a structure hook with: the size as keyword N,
    else as keyword Nx, else positionally; `features` as a keyword when the hook
    takes it (by name or through **kwargs).  None when the signature cannot be
    read or does not fit."""
    try:
        sig = inspect.signature(fn)
    except (TypeError, ValueError):
        return None
    params = sig.parameters
    keyword_ok = (inspect.Parameter.POSITIONAL_OR_KEYWORD, inspect.Parameter.KEYWORD_ONLY)
    takes_kwargs = any(p.kind is inspect.Parameter.VAR_KEYWORD for p in params.values())
    extra = {"features": features} if ("features" in params and params["features"].kind in keyword_ok) \
        or takes_kwargs else {}
    for name in ("N", "Nx"):
        if name in params and params[name].kind in keyword_ok:
            candidate = ((), {name: N, **extra})
            break
    else:
        candidate = ((N,), extra)
    try:
        sig.bind(*candidate[0], **candidate[1])
    except TypeError:
        return None
    return candidate


def resolve_structure(structure: str | Callable, N: int, features: int | None = None) -> tuple:
    """(volume, name) from a registry name or a callable hook taking N or Nx
    (and optionally features)."""
    if isinstance(structure, str):
        make = STRUCTURE_REGISTRY.get(structure)
        if make is None:
            raise ValueError(f"Unknown structure '{structure}'. Supported: "
                             f"{', '.join(sorted(STRUCTURE_REGISTRY))}")
        return make(N, features=features), structure
    if not callable(structure):
        raise TypeError("structure must be a predefined structure name or a callable hook")
    call = _hook_arguments(structure, N, features)
    if call is None:
        # no readable signature (a builtin or an extension callable): the size alone
        try:
            inspect.signature(structure)
        except (TypeError, ValueError):
            call = ((N,), {})
    if call is None:
        raise TypeError("Unable to call custom structure hook. Expected a callable that accepts "
                        "N or Nx (optionally features).")
    name = getattr(structure, "__name__", type(structure).__name__)
    return structure(*call[0], **call[1]), name


def _read_gpu_ram(report: str) -> tuple[float, float, float]:
    """(current, max allocated, reserved) MB from the solver's printed report;
    zeros when it printed no GPU-RAM line."""
    for line in report.splitlines():
        m = _RAM_LINE.search(line)
        if m:
            return tuple(float(v) for v in m.groups())
    return 0.0, 0.0, 0.0


@dataclass
class Row:
    """One solved case (the keys of the reference's row dictionaries)."""

    N: int
    structure: str
    solver: str
    device: str
    conv_crit: float
    total_time: float
    solve_time: float
    iterations: int
    taufactor: float
    torch_cur: float
    torch_max: float
    torch_res: float

    def line(self) -> str:
        text = (self.structure[:10], self.solver[:16], self.device[:4])
        return (f"{self.N:4d} {text[0]:>10} {text[1]:>16} {text[2]:>4} {self.conv_crit:.4f} "
                f"{self.total_time:9.3f} {self.solve_time:9.3f} {self.iterations:6d} {self.taufactor:8.3f} "
                f"{self.torch_cur:10.2f} {self.torch_max:10.2f} {self.torch_res:10.2f}\n")


@dataclass
class Case:
    """One solve of a study: a structure at size N, a solver, a device and a
    convergence criterion."""

    N: int
    device: str
    conv_crit: float
    structure: str | Callable = "fcc"
    features: int | None = None
    iter_limit: int = 10000
    solver: str | type | None = None
    solver_kwargs: dict = field(default_factory=dict)
    solve_kwargs: dict = field(default_factory=dict)

    def run(self) -> Row:
        volume, struct_name = resolve_structure(self.structure, N=self.N, features=self.features)
        cls = resolve_solver(self.solver)
        on_gpu = self.device == "cuda"
        if on_gpu:
            torch.cuda.empty_cache()
        gc.collect()
        if on_gpu:
            torch.cuda.synchronize()
        report = io.StringIO()
        t0 = time.perf_counter()
        with redirect_stdout(report):
            s = cls(volume, device=self.device, **self.solver_kwargs)
            s.solve(iter_limit=self.iter_limit, conv_crit=self.conv_crit, **self.solve_kwargs)
            if on_gpu:
                torch.cuda.synchronize()
            t1 = time.perf_counter()
        cur, peak, reserved = _read_gpu_ram(report.getvalue())
        return Row(N=self.N, structure=struct_name, solver=cls.__name__, device=self.device,
                   conv_crit=self.conv_crit, total_time=t1 - t0, solve_time=float(s.walltime),
                   iterations=int(s.iter), taufactor=float(s.tau[0]), torch_cur=cur, torch_max=peak,
                   torch_res=reserved)


def write_header_if_missing(outfile: str = DEFAULT_OUTFILE) -> None:
    """Start the results file with the table header unless it exists."""
    try:
        with open(outfile, "x", encoding="utf-8") as f:
            f.write(" ".join(f"{title:>{width}}" for title, width in _COLUMNS) + "\n")
            f.write(_RULE + "\n")
    except FileExistsError:
        pass


def append_row_to_file(row: dict, outfile: str = DEFAULT_OUTFILE) -> None:
    """Append one result (a row dictionary) to the results file."""
    with open(outfile, "a", encoding="utf-8") as f:
        f.write(Row(**row).line())


def run_benchmark_case(N: int, device: str, conv_crit: float, structure: str | Callable = "fcc",
                       features: int | None = None, iter_limit: int = 10000,
                       solver: str | type | None = None, solver_kwargs: dict | None = None,
                       solve_kwargs: dict | None = None) -> dict:
    """Solve one case and return its row dictionary."""
    case = Case(N, device, conv_crit, structure, features, iter_limit, solver,
                dict(solver_kwargs or {}), dict(solve_kwargs or {}))
    return asdict(case.run())


def _cases(Ns: Iterable[int], devices: Iterable[str], conv_crit_values: Iterable[float], **common):
    """Every (N, device, criterion) combination, sizes outermost."""
    devices, conv_crit_values = list(devices), list(conv_crit_values)
    for N in Ns:
        for device in devices:
            for conv in conv_crit_values:
                yield Case(N, device, conv, **common)


def run_benchmark_study(Ns=(100, 128, 200, 256, 300, 384, 400), devices=("cuda",),
                        conv_crit_values=(1e-3,), structure: str | Callable = "fcc", features: int = 1,
                        outfile: str = DEFAULT_OUTFILE, write_file: bool = True, iter_limit: int = 10000,
                        solver: str | type | None = None, solver_kwargs: dict | None = None,
                        solve_kwargs: dict | None = None) -> list[dict]:
    """Solve every size x device x criterion combination and return the rows
    (also appended to `outfile` as they finish when write_file is set).
    Cases on "cuda" are skipped with a message when no GPU is visible."""
    if write_file:
        write_header_if_missing(outfile=outfile)
    rows = []
    for case in _cases(Ns, devices, conv_crit_values, structure=structure, features=features,
                       iter_limit=iter_limit, solver=solver, solver_kwargs=dict(solver_kwargs or {}),
                       solve_kwargs=dict(solve_kwargs or {})):
        if case.device == "cuda" and not torch.cuda.is_available():
            print(f"Skipping N={case.N} on CUDA (not available)")
            continue
        row = asdict(case.run())
        rows.append(row)
        if write_file:
            append_row_to_file(row, outfile=outfile)
    return rows
