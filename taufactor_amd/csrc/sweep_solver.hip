// Sweep kernels of the Solver kind (register-blocked): every compiled (n_half, shape) of k_sweep.
#include "tauhip_sweep.h"

int sweep_dispatch_solver(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st) {
    return dispatch_sweep<TF_KIND_SOLVER>(H, a, B, nw, rpw, st);
}
