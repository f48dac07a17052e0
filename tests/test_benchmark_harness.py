"""CPU: the benchmark harness (taufactor_amd.benchmark) keeps the reference's
interface (taufactor/benchmark.py) and output format, and its structures are
the reference's (compared with the oracle's restatement, which the notebook
tables pin through tau)."""
import numpy as np
import pytest

import taufactor_amd as tau
from oracle import structures
from taufactor_amd import benchmark as bm
from taufactor_amd.utils import create_stacked_blocks


def test_resolve_structure_predefined_name():
    cube, name = bm.resolve_structure("blocks", N=12, features=1)
    assert name == "blocks"
    assert cube.shape == (12, 12, 12)


def test_resolve_structure_custom_hook():
    def my_structure(Nx, features=None):
        arr = np.zeros((Nx, Nx, Nx), dtype=int)
        arr[:, :, : Nx // 2] = 1
        return arr

    cube, name = bm.resolve_structure(my_structure, N=10, features=3)
    assert name == "my_structure"
    assert cube.shape == (10, 10, 10) and cube.dtype == int


def test_resolve_structure_rejects_invalid_input():
    with pytest.raises(TypeError):
        bm.resolve_structure(123, N=10, features=1)
    with pytest.raises(ValueError):
        bm.resolve_structure("nope", N=10)


def test_structures_equal_the_restated_reference_generators():
    for n in (16, 32, 48):
        assert np.array_equal(bm.STRUCTURE_REGISTRY["fcc"](n), structures.fcc_pores(n))
        assert np.array_equal(bm.STRUCTURE_REGISTRY["diagonal2d"](n, features=2),
                              structures.diagonals_2d_features(n, 2))
        assert np.array_equal(bm.STRUCTURE_REGISTRY["zigzag"](n), structures.zigzag_2d(n, 1))
    with pytest.raises(ValueError):
        create_stacked_blocks(10, features=3)
    b = create_stacked_blocks(8, features=1)
    assert b.shape == (8, 8, 8) and set(np.unique(b)) <= {0, 1}


def test_solver_registry_and_resolution():
    names = set(bm.SOLVER_REGISTRY)
    assert {"Solver", "PeriodicSolver", "AnisotropicSolver", "MultiPhaseSolver",
            "PeriodicMultiPhaseSolver", "FieldSolver"} <= names
    assert "BatchSolver" not in names and "SORSolver" not in names
    assert bm.resolve_solver(None) is tau.PeriodicSolver
    assert bm.resolve_solver("Solver") is tau.Solver
    with pytest.raises(ValueError):
        bm.resolve_solver("NoSuchSolver")
    with pytest.raises(TypeError):
        bm.resolve_solver(3)


def test_output_format_and_ram_parsing(tmp_path):
    out = tmp_path / "rows.txt"
    bm.write_header_if_missing(str(out))
    row = {"N": 64, "structure": "fcc", "solver": "Solver", "device": "cuda", "conv_crit": 1e-3,
           "total_time": 1.5, "solve_time": 1.25, "iterations": 500, "taufactor": 2.348118,
           "torch_cur": 10.0, "torch_max": 20.5, "torch_res": 30.25}
    bm.append_row_to_file(row, str(out))
    lines = out.read_text().splitlines()
    assert lines[0].split() == ["N", "struct", "solver", "dev", "conv", "Ttime(s)", "Wtime(s)", "iters",
                                "tau", "VRAM(cur)", "VRAM(max)", "VRAM(res)"]
    assert lines[1] == "=" * 120
    assert lines[2] == ("  64        fcc           Solver cuda 0.0010     1.500     1.250    500"
                        "    2.348      10.00      20.50      30.25")
    line = "GPU-RAM currently 339.74 MB (max allocated 540.80 MB; 547.36 MB reserved)"
    assert bm._read_gpu_ram("x\n" + line + "\n") == (339.74, 540.80, 547.36)
    assert bm._read_gpu_ram("no ram line") == (0.0, 0.0, 0.0)
    # the reference's tokenisation of the same line (taufactor/benchmark.py:170-179)
    parts = line.replace("(", "").replace(")", "").replace(",", "").split()
    assert (float(parts[2]), float(parts[6]), float(parts[8])) == bm._read_gpu_ram(line)
    # the header is written once
    bm.write_header_if_missing(str(out))
    assert out.read_text().splitlines()[0] == lines[0] and len(out.read_text().splitlines()) == 3


def test_structure_hook_calling_conventions():
    """A hook is called with N or Nx as a keyword when it names one, else
    positionally, and with features when it takes them (the reference's
    taufactor/benchmark.py:52-73 behaviour)."""
    seen = []

    def by_n(N, features=None):
        seen.append(("N", N, features))
        return np.zeros((N, N, N))

    def by_nx(Nx, features):
        seen.append(("Nx", Nx, features))
        return np.zeros((Nx, 2, 2))

    def positional(size):
        seen.append(("pos", size, None))
        return np.zeros((size, 1, 1))

    def with_kwargs(n, **kw):
        seen.append(("kw", n, kw.get("features")))
        return np.zeros((n, 3, 3))

    def n_only(N):
        seen.append(("N-only", N, None))
        return np.zeros((N, 1, 1))

    class Maker:
        def __call__(self, Nx, features=2):
            seen.append(("obj", Nx, features))
            return np.zeros((Nx, 1, 1))

    for fn in (by_n, by_nx, positional, with_kwargs, n_only, Maker()):
        cube, name = bm.resolve_structure(fn, N=7, features=3)
        assert cube.shape[0] == 7
    assert seen == [("N", 7, 3), ("Nx", 7, 3), ("pos", 7, None), ("kw", 7, 3), ("N-only", 7, None),
                    ("obj", 7, 3)]
    assert bm.resolve_structure(Maker(), N=4)[1] == "Maker"

    def two_required(a, b):
        return np.zeros((a, b, 1))

    with pytest.raises(TypeError, match="Unable to call custom structure hook"):
        bm.resolve_structure(two_required, N=5)

    def raises_inside(N):
        raise TypeError("from inside the hook")

    with pytest.raises(TypeError, match="from inside the hook"):
        bm.resolve_structure(raises_inside, N=5)


def test_case_rows_and_study_order(monkeypatch):
    """run_benchmark_study runs sizes outermost, then devices, then criteria,
    and returns row dictionaries with the reference's keys."""
    calls = []

    def fake_run(self):
        calls.append((self.N, self.device, self.conv_crit))
        return bm.Row(self.N, "fcc", "Solver", self.device, self.conv_crit, 1.0, 0.5, 100, 2.0,
                      1.0, 2.0, 3.0)

    monkeypatch.setattr(bm.Case, "run", fake_run)
    rows = bm.run_benchmark_study(Ns=[8, 16], devices=("cpu", "cpux"), conv_crit_values=(1e-2, 1e-3),
                                  write_file=False)
    assert calls == [(8, "cpu", 1e-2), (8, "cpu", 1e-3), (8, "cpux", 1e-2), (8, "cpux", 1e-3),
                     (16, "cpu", 1e-2), (16, "cpu", 1e-3), (16, "cpux", 1e-2), (16, "cpux", 1e-3)]
    assert set(rows[0]) == {"N", "structure", "solver", "device", "conv_crit", "total_time", "solve_time",
                            "iterations", "taufactor", "torch_cur", "torch_max", "torch_res"}
