// Sweep kernels of the uneven batch (register-blocked, Solver kind).
#include "tauhip_sweep.h"

int sweep_dispatch_batch(int H, BatchArgs& a, int nw, int rpw, hipStream_t st) {
    switch (H) {
        case 1: return dispatch_batch_shape<1>(a, nw, rpw, st);
        case 2: return dispatch_batch_shape<2>(a, nw, rpw, st);
        case 3: return dispatch_batch_shape<3>(a, nw, rpw, st);
        case 4: return dispatch_batch_shape<4>(a, nw, rpw, st);
        case 5: return dispatch_batch_shape<5>(a, nw, rpw, st);
        case 6: return dispatch_batch_shape<6>(a, nw, rpw, st);
        default: return fail(-1, "n_half=%d out of range 1..%d", H, TF_MAX_FUSED_HALF_SWEEPS);
    }
}
