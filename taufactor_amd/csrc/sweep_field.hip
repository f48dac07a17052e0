// Sweep kernels of the per-voxel diffusivity kind (LDS ring): every compiled (n_half, shape) of k_sweep.
#include "tauhip_sweep.h"

int sweep_dispatch_field(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st) {
    return dispatch_sweep<TF_KIND_FIELD>(H, a, B, nw, rpw, st);
}
