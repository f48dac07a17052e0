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
    assert bm.parse_gpu_ram(["x", line]) == (339.74, 540.80, 547.36)
    assert bm.parse_gpu_ram(["no ram line"]) == (0.0, 0.0, 0.0)
