"""GPU: the benchmark harness reproduces the reference's stored table, and the
reference's electrode.tiff (read with taufactor_amd.io) solves to the values
the reference's overview notebook prints (docs/notebooks/00-overview.ipynb,
cells 2, 4 and 8: 700 iterations and tau 2.0776296 at the default criterion,
1300 iterations and tau 2.0771587 at 1e-3, tau 2.0766 along y)."""
import json
import os

import numpy as np
import pytest

import taufactor_amd as tau
from taufactor_amd import benchmark as bm
from taufactor_amd import io as tio
from taufactor_amd.utils import create_stacked_blocks

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_harness_reproduces_fcc_table(goldens, tmp_path):
    rows = bm.run_benchmark_study(Ns=[32, 64, 100], solver="Solver", structure="fcc",
                                  conv_crit_values=[1e-3], outfile=str(tmp_path / "r.txt"))
    want = {r["N"]: r for r in goldens["fcc_solver_conv1e-3"]["rows"]}
    for r in rows:
        assert r["iterations"] == want[r["N"]]["iterations"]
        assert abs(r["taufactor"] - want[r["N"]]["tau"]) < 1.5e-6
        assert r["torch_max"] > 0
    assert len((tmp_path / "r.txt").read_text().splitlines()) == 2 + 3


def test_harness_blocks_string_and_hook_match():
    def structure_fn(N):
        return create_stacked_blocks(N)

    r1 = bm.run_benchmark_study(Ns=[16], solver="Solver", structure="blocks", write_file=False)
    r2 = bm.run_benchmark_study(Ns=[16], solver="Solver", structure=structure_fn, write_file=False)
    assert np.isclose(r1[0]["taufactor"], r2[0]["taufactor"])


@pytest.fixture(scope="module")
def electrode():
    return tio.imread(os.path.join(HERE, "golden", "electrode.tiff"))


def test_electrode_default_criterion(electrode):
    s = tau.Solver(electrode == 0)
    s.solve(verbose=False)
    assert s.iter == 700
    assert abs(float(s.tau[0]) - 2.0776296) <= 2e-7 * 2.0776296
    assert f"{s.D_eff[0]:.4f}" == "0.2166"


def test_electrode_tight_criterion_and_y_direction(electrode):
    s = tau.Solver(electrode == 0)
    s.solve(verbose=False, conv_crit=1e-3)
    assert s.iter == 1300
    assert abs(float(s.tau[0]) - 2.0771587) <= 2e-7 * 2.0771587
    t = tau.Solver(np.transpose(electrode, (1, 2, 0)) == 0)
    t.solve(verbose=False)
    assert f"{t.tau[0]:.4f}" == "2.0766"


def test_tiff_goldens_file_is_consistent():
    g = json.load(open(os.path.join(HERE, "golden", "tiff_goldens.json")))
    assert g["electrode.tiff"]["labels"]["0"] == 7548361
