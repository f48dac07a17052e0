"""GPU: ElectrodeSolver / PeriodicElectrodeSolver through the HIP path against
the oracle (iterates value for value, stopping iteration and tau) and
against the reference's stored outputs: the periodic graded batch of
docs/notebooks/01-tortuosity.ipynb (1900 iterations, six tau values, cell 9)
and electrode.tiff with pore/NMC labels (3800 iterations, tau 2.07291984,
cell 26)."""
import os

import numpy as np
import pytest

import taufactor_amd as tau
from oracle.restate import OracleElectrodeSolver, OraclePeriodicElectrodeSolver
from oracle.structures import graded_batch
from taufactor_amd import io as tio

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _labels(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.choice([0, 1, 5], size=shape, p=[0.3, 0.55, 0.15])


@pytest.mark.parametrize("shape", [(13, 9, 17), (40, 21, 130), (7, 5, 1)])
@pytest.mark.parametrize("H", [1, 2, 3, 4])
def test_iterates_bit_identical(shape, H):
    img = _labels(shape, sum(shape) + H)
    s = tau.ElectrodeSolver(img, spacing=0.5)
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=23, verbose=False)
    o = OracleElectrodeSolver(img, spacing=0.5).run(23)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


def test_periodic_iterates_and_labels():
    # interiors value for value: the y/z ghosts of the final field are wrap
    # copies taken at different points of the last iteration by the two
    # implementations and feed nothing afterwards
    img = _labels((11, 8, 9), 4)
    s = tau.PeriodicElectrodeSolver(img, conductive_label=5, reactive_label=1)
    s.solve(iter_limit=31, verbose=False)
    o = OraclePeriodicElectrodeSolver(img, conductive_label=5, reactive_label=1).run(31)
    assert np.array_equal(s.field.cpu().numpy()[:, :, 1:-1, 1:-1], o.field.numpy()[:, :, 1:-1, 1:-1])


@pytest.mark.parametrize("H", [1, 3, 6])
def test_periodic_fused_iterates_bit_identical(H):
    # Y and Z even: the sweep reads across the boundary and fuses; the whole
    # field, ghosts included, equals the oracle's
    img = _labels((11, 8, 10), 5 + H)
    s = tau.PeriodicElectrodeSolver(img, conductive_label=5, reactive_label=1)
    s.half_sweeps_per_pass = H
    s.solve(iter_limit=31, verbose=False)
    o = OraclePeriodicElectrodeSolver(img, conductive_label=5, reactive_label=1).run(31)
    assert np.array_equal(s.field.cpu().numpy(), o.field.numpy())


def test_full_solve_matches_oracle():
    img = _labels((24, 16, 18), 9)
    s = tau.ElectrodeSolver(img)
    s.solve(verbose=False, conv_crit=1e-3)
    o = OracleElectrodeSolver(img).solve(conv_crit=1e-3)
    assert s.iter == o.iter
    np.testing.assert_allclose(s.tau, o.tau, rtol=1e-6)


def test_graded_batch_notebook_values(goldens):
    g = goldens["periodic_electrode_graded_conv1e-3"]
    s = tau.PeriodicElectrodeSolver(graded_batch(), spacing=100e-6 / 250)
    s.solve(iter_limit=50000, conv_crit=1e-3, verbose=False)
    assert s.iter == g["iterations"]
    np.testing.assert_allclose(s.tau, g["tau"], rtol=1e-6)


def test_electrode_tiff_notebook_value(goldens):
    g = goldens["electrode_tiff_pore_nmc_conv1e-4"]
    img = tio.imread(os.path.join(HERE, "golden", "electrode.tiff"))
    s = tau.ElectrodeSolver(img, spacing=100e-6 / img.shape[0], conductive_label=0, reactive_label=85)
    s.solve(iter_limit=50000, conv_crit=1e-4, verbose=False)
    assert s.iter == g["iterations"]
    np.testing.assert_allclose(s.tau, [g["tau"]], rtol=1e-6)
