// Sweep kernels of the MultiPhase kind read through per-voxel face codes:
// every compiled (n_half, shape) of k_sweep.
#include "tauhip_sweep.h"

int sweep_dispatch_mpfaces(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st) {
    return dispatch_sweep<TF_KIND_MULTIPHASE_FACES>(H, a, B, nw, rpw, st);
}
