// Sweep kernels of the anisotropic kind (register-blocked): every compiled (n_half, shape) of k_sweep.
#include "tauhip_sweep.h"

int sweep_dispatch_anisotropic(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st) {
    return dispatch_sweep<TF_KIND_ANISOTROPIC>(H, a, B, nw, rpw, st);
}
