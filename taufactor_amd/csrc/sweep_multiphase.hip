// Sweep kernels of the MultiPhase kind (LDS ring): every compiled (n_half, shape) of k_sweep.
#include "tauhip_sweep.h"

int sweep_dispatch_multiphase(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st) {
    return dispatch_sweep<TF_KIND_MULTIPHASE>(H, a, B, nw, rpw, st);
}
