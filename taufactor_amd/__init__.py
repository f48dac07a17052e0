"""taufactor_amd — MI355X-native steady-state diffusion solve with the
TauFactor solver API (``Solver``, ``MultiPhaseSolver``).

The iteration runs in hand-written gfx950 HIP kernels (libtauhip.so, C-ABI in
include/tauhip.h) loaded through ctypes; PyTorch-ROCm holds the device
buffers.  There is no CPU code path.
"""
from .batch import BatchSolver
from .electrode import ElectrodeSolver, PeriodicElectrodeSolver
from .autograd import effective_diffusivity
from .solver import (AnisotropicSolver, FieldSolver, MultiPhaseSolver, PeriodicMultiPhaseSolver,
                     PeriodicSolver, SORSolver, Solver, ThroughTransportSolver)

__all__ = ["Solver", "MultiPhaseSolver", "PeriodicSolver", "PeriodicMultiPhaseSolver", "AnisotropicSolver",
           "ElectrodeSolver", "PeriodicElectrodeSolver", "BatchSolver", "FieldSolver", "effective_diffusivity", "SORSolver", "ThroughTransportSolver"]
__version__ = "0.1.0"
