// tauhip.hip — gfx950 kernels and C-ABI for the steady-state diffusion solve
// behind taufactor_amd.Solver / taufactor_amd.MultiPhaseSolver.
//
// Built with -ffp-contract=off: every float32 operation of the update is
// rounded on its own, in the order the reference's eager torch ops use
// (taufactor/taufactor.py:97-102, 176-181, 606-613), so iterates match the
// reference value for value.  Fused multiply-adds appear only where written
// explicitly (the exact-division sequence in div_by_count).
//
// Layout and conventions are documented in include/tauhip.h.  The sweep
// kernels live in tauhip_sweep.h and are instantiated in sweep_<kind>.hip.

#include "tauhip_sweep.h"

#include <mutex>

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    return fail((int)e, "%s: %s", what, hipGetErrorString(e));
}

extern "C" int tf_abi_version(void) { return TF_ABI_VERSION; }
extern "C" const char* tf_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------
// set-up kernels
// ---------------------------------------------------------------------------
// One thread per padded voxel.  Writes codes and both field buffers.
// mode 0: mask is 0/1 conductive flag, code = neighbour count (anisotropic:
//         the six neighbour states)
// mode 1: mask is a class index, code = class (x ghosts copy the adjacent
//         voxel's class, y/z ghosts class 0 or, periodic, the opposite
//         layer's), conductive = cls_cond[class]
// mode 2: mask is 0 / 1 conductive / 2 reactive, code = 1 + 7 x conductive
//         neighbours + reactive neighbours for conductive voxels
// Grid (set_up_grid): x = (padded x plane, block of the plane), y = volume
// b0 + blockIdx.y, so that the plane and the volume are uniform per block and
// only the row is divided out per thread.
template <int MODE>
__global__ void k_setup(const uint8_t* __restrict__ mask, const uint8_t* __restrict__ cls_cond,
                        const float* __restrict__ xinit, uint8_t* __restrict__ codes,
                        float* __restrict__ fa, float* __restrict__ fb, Vol v, int b0, int nbx,
                        int periodic, int aniso) {
    const int xp = (int)(blockIdx.x / (unsigned)nbx);
    const int r2 = (int)(blockIdx.x - (unsigned)xp * (unsigned)nbx) * (int)blockDim.x + (int)threadIdx.x;
    if (r2 >= v.sX) return;
    const int b = b0 + (int)blockIdx.y;
    const int r = xp * v.sX + r2;
    const int yp = r2 / v.sY;
    const int zp = r2 - yp * v.sY;
    const uint8_t* M = mask + (size_t)b * v.X * v.Y * v.Z;
    auto m_at = [&](int x, int y, int z) -> int {  // interior 0-based
        return (int)M[((size_t)x * v.Y + y) * v.Z + z];
    };
    bool yz_in = (yp >= 1 && yp <= v.Y && zp >= 1 && zp <= v.Z);
    bool in = yz_in && xp >= 1 && xp <= v.X;
    float f;
    uint8_t code = 0;
    if constexpr (MODE == 1) {
        // class codes: one unconditional load of the class the padded voxel
        // takes: its own; the adjacent layer's for an x ghost (diff_map[:,0] =
        // diff_map[:,1], taufactor/taufactor.py:591-592); with periodic y/z the
        // opposite interior layer's for a y/z ghost, after the x copy
        // (PeriodicMultiPhaseSolver, taufactor/taufactor.py:633-638); closed
        // y/z ghosts are class 0 (their clamped load goes unused)
        const int xs = xp == 0 ? 0 : (xp == v.X + 1 ? v.X - 1 : xp - 1);
        int ys, zs;
        if (periodic) {
            ys = (yp == 0 ? v.Y : (yp == v.Y + 1 ? 1 : yp)) - 1;
            zs = (zp == 0 ? v.Z : (zp == v.Z + 1 ? 1 : zp)) - 1;
        } else {
            ys = min(max(yp - 1, 0), v.Y - 1);
            zs = min(max(zp - 1, 0), v.Z - 1);
        }
        const int m = m_at(xs, ys, zs);
        const int cond = cls_cond[m];
        const float xi = xinit[xs];
        code = (periodic || yz_in) ? (uint8_t)m : (uint8_t)0;
        // mask * linspace in float32 (taufactor/taufactor.py:291); ghosts: x
        // planes hold 2*bc on their y/z interior, everything else 0
        // (SORSolver._pad, taufactor/taufactor.py:234-237)
        f = in ? (cond ? 1.0f : 0.0f) * xi
               : (yz_in && xp == 0) ? -1.0f : (yz_in && xp == v.X + 1) ? 1.0f : 0.0f;
    } else if (in) {
        const int x = xp - 1, y = yp - 1, z = zp - 1;
        const uint8_t* P = M + ((size_t)x * v.Y + y) * v.Z + z;
        const int m = P[0];
        const int cond = MODE == 0 ? (m != 0) : (m == 1);
        // mask * linspace in float32 (taufactor/taufactor.py:291)
        f = (cond ? 1.0f : 0.0f) * xinit[x];
        // the six neighbours, loaded unconditionally so that no load sits
        // under a divergent branch: y/z offsets wrap round when periodic;
        // a ghost neighbour's offset is 0 (the voxel itself) and its value
        // is replaced below
        const long long YZ = (long long)v.Y * v.Z;
        const bool xq_in = x + 1 < v.X, xm_in = x >= 1;
        const bool yq_in = y + 1 < v.Y, ym_in = y >= 1, zq_in = z + 1 < v.Z, zm_in = z >= 1;
        const long long oxq = xq_in ? YZ : 0, oxm = xm_in ? -YZ : 0;
        int oyq, oym, ozq, ozm;
        if (periodic) {
            oyq = (yq_in ? 1 : 1 - v.Y) * v.Z;
            oym = (ym_in ? -1 : v.Y - 1) * v.Z;
            ozq = zq_in ? 1 : 1 - v.Z;
            ozm = zm_in ? -1 : v.Z - 1;
        } else {
            oyq = yq_in ? v.Z : 0;
            oym = ym_in ? -v.Z : 0;
            ozq = zq_in ? 1 : 0;
            ozm = zm_in ? -1 : 0;
        }
        const int mxq = P[oxq], mxm = P[oxm], myq = P[oyq], mym = P[oym], mzq = P[ozq], mzm = P[ozm];
        // y/z neighbours inside the volume (all of them when periodic)
        const bool yq_ok = periodic || yq_in, ym_ok = periodic || ym_in;
        const bool zq_ok = periodic || zq_in, zm_ok = periodic || zm_in;
        if (MODE == 2) {
            if (cond) {
                // conductive neighbours with x ghosts 2 (x = 0) and 0 (x = X+1),
                // reactive neighbours with every ghost 0; y/z wrap round when
                // periodic (taufactor/electrode.py:47-62, 133-151)
                const int cn = (xq_in ? (mxq == 1) : 0) + (xm_in ? (mxm == 1) : 2) + (yq_ok && myq == 1) +
                               (ym_ok && mym == 1) + (zq_ok && mzq == 1) + (zm_ok && mzm == 1);
                const int rn = (xq_in && mxq == 2) + (xm_in && mxm == 2) + (yq_ok && myq == 2) +
                               (ym_ok && mym == 2) + (zq_ok && mzq == 2) + (zm_ok && mzm == 2);
                code = (uint8_t)(1 + 7 * cn + rn);
            }
        } else if (aniso) {
            if (cond) {
                // neighbour states x-,x+ (0/1, 2 for the x ghost), y-,y+,z-,z+
                // (0/1, ghosts 0): the terms AnisotropicSolver sums, in its
                // order (taufactor/taufactor.py:459-471)
                const int xm = xm_in ? (mxm != 0) : 2;
                const int xq = xq_in ? (mxq != 0) : 2;
                const int ym = ym_in && mym != 0, yq = yq_in && myq != 0;
                const int zm = zm_in && mzm != 0, zq = zq_in && mzq != 0;
                code = (uint8_t)(1 + 48 * xm + 16 * xq + 8 * ym + 4 * yq + 2 * zm + zq);
            }
        } else if (cond) {
            // padded mask: x ghosts 2, y/z ghosts 0 (taufactor/taufactor.py:404-410),
            // or y/z wrapped round (PeriodicSolver, taufactor/taufactor.py:493-499)
            const int nn = (xq_in ? (mxq != 0) : 2) + (xm_in ? (mxm != 0) : 2) + (yq_ok && myq != 0) +
                           (ym_ok && mym != 0) + (zq_ok && mzq != 0) + (zm_ok && mzm != 0);
            code = (uint8_t)nn;  // 0 stays 0 = inf
        }
    } else {
        // ghost: x planes hold 2*bc on their y/z interior, everything else 0
        // (SORSolver._pad, taufactor/taufactor.py:234-237)
        f = 0.0f;
        if (MODE == 2) {
            // ElectrodeSolver pads with [2*left_bc, 0] (taufactor/electrode.py:45)
            if (yz_in && xp == 0) f = 2.0f;
        } else {
            if (yz_in && xp == 0) f = -1.0f;
            if (yz_in && xp == v.X + 1) f = 1.0f;
        }
    }
    size_t o = (size_t)b * v.sB + r;
    codes[o] = code;
    fa[o] = f;
    fb[o] = f;
}

constexpr int kSetupThreads = 256;
constexpr int kCountThreads = 1024;
constexpr int kMaxGridY = 65535;

// Launch shape of k_setup: nbx blocks per padded x plane; the volumes go in
// chunks of at most kMaxGridY along the grid's y.
static int set_up_grid(const Vol& v, int* nbx, unsigned* gx) {
    *nbx = (v.sX + kSetupThreads - 1) / kSetupThreads;
    const long long blocks = (long long)*nbx * (v.X + 2);
    if (blocks * kSetupThreads >= (1LL << 32)) return fail(-1, "volume too long in x for the set-up grid");
    *gx = (unsigned)blocks;
    return 0;
}

// Sum over the workgroup (kCountThreads lanes); the total is valid in
// thread 0.  Integer counts: the order does not change the result.
__device__ inline uint32_t workgroup_sum(uint32_t a) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += (uint32_t)__shfl_xor((int)a, o);
    __shared__ uint32_t part[kCountThreads / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = a;
    __syncthreads();
    uint32_t s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kCountThreads / 64; ++i) s += part[i];
    return s;
}

// conductive flag of one mask byte, and the count over the four bytes of a
// word: MODE 0 counts nonzero bytes (bit 7 of ((b & 0x7f) + 0x7f) | b is set
// exactly when b != 0, and no byte carries into the next), MODE 2 bytes equal
// to 1 (the nonzero test of b ^ 1), MODE 1 looks each class up
template <int MODE>
__device__ inline uint32_t conductive_byte(uint32_t m, const uint8_t* __restrict__ cls_cond) {
    return MODE == 0 ? (m != 0) : MODE == 2 ? (m == 1) : cls_cond[m];
}
template <int MODE>
__device__ inline uint32_t conductive_word(uint32_t w, const uint8_t* __restrict__ cls_cond) {
    if (MODE == 1)
        return (uint32_t)cls_cond[w & 255u] + cls_cond[(w >> 8) & 255u] + cls_cond[(w >> 16) & 255u] +
               cls_cond[w >> 24];
    if (MODE == 2) w ^= 0x01010101u;
    const uint32_t nz = __popc((((w & 0x7f7f7f7fu) + 0x7f7f7f7fu) | w) & 0x80808080u);
    return MODE == 2 ? 4u - nz : nz;
}

// conductive voxels per x layer; one workgroup per (x, b0 + blockIdx.y): the
// layer's bytes up to a 16-byte boundary, then 16-byte loads, then the rest
template <int MODE>
__global__ void __launch_bounds__(kCountThreads)
k_layer_count(const uint8_t* __restrict__ mask, const uint8_t* __restrict__ cls_cond,
              uint32_t* __restrict__ cnt, Vol v, int b0) {
    const int x = blockIdx.x, b = b0 + (int)blockIdx.y;
    const long long n = (long long)v.Y * v.Z;
    const uint8_t* M = mask + ((size_t)b * v.X + x) * (size_t)n;
    const long long head = min(n, (long long)((16 - ((uintptr_t)M & 15)) & 15));
    const long long nv = (n - head) / 16;
    const long long tail = head + nv * 16;
    uint32_t acc = 0;
    if (threadIdx.x < head) acc += conductive_byte<MODE>(M[threadIdx.x], cls_cond);
    const uint4* W = reinterpret_cast<const uint4*>(M + head);
    for (long long i = threadIdx.x; i < nv; i += kCountThreads) {
        const uint4 w = W[i];
        acc += conductive_word<MODE>(w.x, cls_cond) + conductive_word<MODE>(w.y, cls_cond) +
               conductive_word<MODE>(w.z, cls_cond) + conductive_word<MODE>(w.w, cls_cond);
    }
    if (tail + threadIdx.x < n) acc += conductive_byte<MODE>(M[tail + threadIdx.x], cls_cond);
    const uint32_t total = workgroup_sum(acc);
    if (threadIdx.x == 0) cnt[(size_t)b * v.X + x] = total;
}

template <int MODE>
static int setup_common(const uint8_t* mask, const uint8_t* cls_cond, const float* xinit,
                        uint8_t* codes, float* fa, float* fb, uint32_t* cnt,
                        int B, int X, int Y, int Z, int flags, void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (flags & ~(TF_SETUP_PERIODIC_YZ | TF_SETUP_ANISOTROPIC))
        return fail(-1, "unknown set-up flags 0x%x", flags);
    if ((flags & TF_SETUP_ANISOTROPIC) && (MODE != 0 || (flags & TF_SETUP_PERIODIC_YZ)))
        return fail(-1, "TF_SETUP_ANISOTROPIC applies to tf_setup_solver without periodic boundaries");
    if (!mask || !xinit || !codes || !fa || !fb || !cnt || (MODE == 1 && !cls_cond))
        return fail(-1, "null pointer argument");
    if (fa == fb) return fail(-1, "field_a and field_b must be distinct buffers");
    Vol v = make_vol(X, Y, Z);
    hipStream_t st = (hipStream_t)stream;
    int nbx;
    unsigned gx;
    if (int rc = set_up_grid(v, &nbx, &gx)) return rc;
    if ((long long)X * kCountThreads >= (1LL << 32)) return fail(-1, "volume too long in x for the layer count");
    for (int b0 = 0; b0 < B; b0 += kMaxGridY) {
        const unsigned nb = (unsigned)min(B - b0, kMaxGridY);
        hipLaunchKernelGGL(k_setup<MODE>, dim3(gx, nb), dim3(kSetupThreads), 0, st, mask, cls_cond, xinit,
                           codes, fa, fb, v, b0, nbx, (flags & TF_SETUP_PERIODIC_YZ) ? 1 : 0,
                           (flags & TF_SETUP_ANISOTROPIC) ? 1 : 0);
        TF_LAUNCH_CHECK("k_setup");
        hipLaunchKernelGGL(k_layer_count<MODE>, dim3(X, nb), dim3(kCountThreads), 0, st, mask, cls_cond, cnt,
                           v, b0);
        TF_LAUNCH_CHECK("k_layer_count");
    }
    return 0;
}

// Sum over each layer of the reactive-neighbour part of the electrode codes;
// one block per (x, b)
__global__ void k_reactive_count(const uint8_t* __restrict__ codes, uint32_t* __restrict__ cnt, Vol v) {
    const int x = blockIdx.x, b = blockIdx.y;
    const uint8_t* C = codes + (size_t)b * v.sB + (size_t)(x + 1) * v.sX;
    uint32_t acc = 0;
    for (int i = threadIdx.x; i < v.Y * v.Z; i += blockDim.x) {
        const int y = i / v.Z, z = i - y * v.Z;
        const int c = C[(y + 1) * v.sY + z + 1];
        acc += c ? (uint32_t)((c - 1) % 7) : 0u;
    }
    __shared__ uint32_t red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[(size_t)b * v.X + x] = red[0];
}

extern "C" int tf_setup_electrode(const uint8_t* cls, const float* xinit, uint8_t* codes,
                                  float* field_a, float* field_b, uint32_t* layer_count,
                                  uint32_t* reactive_count, int B, int X, int Y, int Z, int flags,
                                  void* stream) {
    if (flags & ~TF_SETUP_PERIODIC_YZ) return fail(-1, "unknown set-up flags 0x%x", flags);
    if (!reactive_count) return fail(-1, "null pointer argument");
    if (int rc = setup_common<2>(cls, nullptr, xinit, codes, field_a, field_b, layer_count, B, X, Y, Z,
                                 flags, stream))
        return rc;
    hipLaunchKernelGGL(k_reactive_count, dim3(X, B), dim3(256), 0, (hipStream_t)stream, codes,
                       reactive_count, make_vol(X, Y, Z));
    TF_LAUNCH_CHECK("k_reactive_count");
    return 0;
}

extern "C" int tf_setup_solver(const uint8_t* mask, const float* xinit, uint8_t* codes,
                               float* field_a, float* field_b, uint32_t* layer_count,
                               int B, int X, int Y, int Z, int flags, void* stream) {
    return setup_common<0>(mask, nullptr, xinit, codes, field_a, field_b, layer_count,
                           B, X, Y, Z, flags, stream);
}

// y/z ghosts of every x plane := the opposite interior layer (ghost corners
// take both wraps), as SORSolver.apply_boundary_conditions of the periodic
// solvers does before each iteration (taufactor/taufactor.py:501-505, 652-656)
__global__ void k_wrap_ghosts(const float* S, float* D, Vol v, int B) {
    // one thread per ghost position of one (b, x) plane: the border of the
    // (Y+2) x (Z+2) plane, 2(Y+2) + 2Z positions
    const int per_plane = 2 * (v.Y + 2) + 2 * v.Z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per_plane) return;
    const int bx = blockIdx.y;
    int yp, zp;
    if (i < 2 * (v.Y + 2)) {
        yp = i >> 1;
        zp = (i & 1) ? v.Z + 1 : 0;
    } else {
        const int j = i - 2 * (v.Y + 2);
        zp = 1 + (j >> 1);
        yp = (j & 1) ? v.Y + 1 : 0;
    }
    const int ys = yp == 0 ? v.Y : (yp == v.Y + 1 ? 1 : yp);
    const int zs = zp == 0 ? v.Z : (zp == v.Z + 1 ? 1 : zp);
    const size_t plane = (size_t)bx * v.sX;
    D[plane + yp * v.sY + zp] = S[plane + ys * v.sY + zs];
}

// ghosts of dst from the interior of src (src == dst: in place; the ghosts
// are neither read nor written by the copies of other ghosts)
extern "C" int tf_wrap_ghosts_from(const float* src, float* dst, int B, int X, int Y, int Z,
                                   void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!src || !dst) return fail(-1, "null pointer argument");
    if ((long long)B * (X + 2) > 65535) return fail(-1, "grid too large for ghost wrap");
    Vol v = make_vol(X, Y, Z);
    const int per_plane = 2 * (Y + 2) + 2 * Z;
    hipLaunchKernelGGL(k_wrap_ghosts, dim3((per_plane + 255) / 256, B * (X + 2)), dim3(256), 0,
                       (hipStream_t)stream, src, dst, v, B);
    TF_LAUNCH_CHECK("k_wrap_ghosts");
    return 0;
}

extern "C" int tf_wrap_ghosts(float* field, int B, int X, int Y, int Z, void* stream) {
    return tf_wrap_ghosts_from(field, field, B, X, Y, Z, stream);
}

extern "C" int tf_setup_multiphase(const uint8_t* cls, const uint8_t* cls_conductive,
                                   int n_classes, const float* xinit, uint8_t* codes,
                                   float* field_a, float* field_b, uint32_t* layer_count,
                                   int B, int X, int Y, int Z, int flags, void* stream) {
    if (n_classes < 1 || n_classes > 256) return fail(-1, "n_classes=%d out of range", n_classes);
    return setup_common<1>(cls, cls_conductive, xinit, codes, field_a, field_b, layer_count,
                           B, X, Y, Z, flags, stream);
}

// one thread per 4-byte word of the sweep-code layout (rows are whole words)
__global__ void k_pack_sweep_codes(const uint8_t* __restrict__ codes, uint32_t* __restrict__ out,
                                   Vol v, int row_c, long long n_words, int shift) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n_words) return;
    const long long byte0 = gid * 4;
    const long long row = byte0 / row_c;  // row index over B*(X+2)*(Y+2)
    const int k0 = (int)(byte0 - row * row_c);
    const uint8_t* src = codes + row * (long long)v.sY;
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int z = k0 + i - 3;
        const uint32_t c = (z >= 0 && z <= v.Z + 1) ? src[z] : 0u;
        w |= (c << shift) << (8 * i);
    }
    out[gid] = w;
}

extern "C" long long tf_sweep_codes_bytes(int X, int Y, int Z) {
    if (X < 1 || Y < 1 || Z < 1) return fail(-1, "invalid shape %dx%dx%d", X, Y, Z);
    return sweep_code_volume(X, Y, Z);
}

extern "C" int tf_pack_sweep_codes(const uint8_t* codes, uint8_t* sweep_codes, int kind, int B, int X,
                                   int Y, int Z, void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (kind < TF_KIND_SOLVER || kind > TF_KIND_ELECTRODE) return fail(-1, "unknown kind %d", kind);
    if (!codes || !sweep_codes) return fail(-1, "null pointer argument");
    if (reinterpret_cast<uintptr_t>(sweep_codes) & 3) return fail(-1, "sweep_codes must be 4-byte aligned");
    const long long n_words = (long long)B * sweep_code_volume(X, Y, Z) / 4;
    const long long blocks = (n_words + 255) / 256;
    if (blocks >= (1LL << 31)) return fail(-1, "volume too large");
    hipLaunchKernelGGL(k_pack_sweep_codes, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       codes, reinterpret_cast<uint32_t*>(sweep_codes), make_vol(X, Y, Z),
                       sweep_code_row(Z), n_words, kind == TF_KIND_SOLVER ? 3 : 0);
    TF_LAUNCH_CHECK("k_pack_sweep_codes");
    return 0;
}

// ---------------------------------------------------------------------------
// unblocked in-place half-sweep: one thread per interior voxel
// ---------------------------------------------------------------------------
template <int KIND>
__global__ void k_half_inplace(float* __restrict__ F, const uint8_t* __restrict__ C,
                               const float* __restrict__ hm, int ncls, Vol v, float w, int colour) {
    int z = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    int bx = blockIdx.z;
    int b = bx / v.X, x = bx - b * v.X;
    if (z >= v.Z) return;
    if (((x + y + z) & 1) != colour) return;
    float* V = F + (size_t)b * v.sB;
    const uint8_t* Cv = C + (size_t)b * v.sB;
    int o = (x + 1) * v.sX + (y + 1) * v.sY + (z + 1);
    Nbr n;
    n.c = V[o];
    n.xp = V[o + v.sX];
    n.xm = V[o - v.sX];
    n.yp = V[o + v.sY];
    n.ym = V[o - v.sY];
    n.zp = V[o + 1];
    n.zm = V[o - 1];
    float out;
    if (KIND == TF_KIND_SOLVER) {
        out = update_solver(n, Cv[o], w);
    } else {
        const float* row = hm + Cv[o] * ncls;
        out = update_multiphase(n, row[Cv[o + v.sX]], row[Cv[o - v.sX]], row[Cv[o + v.sY]],
                                row[Cv[o - v.sY]], row[Cv[o + 1]], row[Cv[o - 1]], x == 0,
                                x == v.X - 1, w);
    }
    V[o] = out;
}

extern "C" int tf_sweep_half_inplace(float* field, const uint8_t* codes, const float* hm_lut,
                                     int n_classes, int kind, int B, int X, int Y, int Z,
                                     float omega, int colour, void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!field || !codes) return fail(-1, "null pointer argument");
    if (kind == TF_KIND_MULTIPHASE && (!hm_lut || n_classes < 1 || n_classes > 256))
        return fail(-1, "multiphase sweep needs hm_lut and 1..256 classes");
    if (colour != 0 && colour != 1) return fail(-1, "colour must be 0 or 1");
    if ((long long)B * X > 65535 || Y > 65535) return fail(-1, "grid too large for unblocked sweep");
    Vol v = make_vol(X, Y, Z);
    dim3 grid((Z + 255) / 256, Y, B * X);
    hipStream_t st = (hipStream_t)stream;
    if (kind == TF_KIND_SOLVER)
        hipLaunchKernelGGL(k_half_inplace<TF_KIND_SOLVER>, grid, dim3(256), 0, st, field, codes,
                           hm_lut, n_classes, v, omega, colour);
    else if (kind == TF_KIND_MULTIPHASE)
        hipLaunchKernelGGL(k_half_inplace<TF_KIND_MULTIPHASE>, grid, dim3(256), 0, st, field,
                           codes, hm_lut, n_classes, v, omega, colour);
    else
        return fail(-1, "unknown kind %d", kind);
    TF_LAUNCH_CHECK("k_half_inplace");
    return 0;
}


// Default workgroup shape.  32-row tiles (16 waves x 2) halve the y-halo
// overhead of 16-row tiles where their LDS ring fits (one-byte codes,
// n_half <= 4) and the volume is tall enough; the x-tile length is left to
// the launch (balanced_tx).  Any shape gives identical results.
static void default_tile(int H, int kind, int Y, tf_tile* t) {
    const bool tall = (own_code_kind(kind) || (kind != TF_KIND_FIELD && H <= 4)) && Y >= 64;
    t->waves = tall ? 16 : 8;
    t->rows_per_wave = 2;
    t->tx = 0;  // chosen at launch from the resident workgroups (balanced_tx)
}

// The kind argument of tf_sweep / tf_sweep_launch_config: the kind in the
// low byte, TF_SWEEP_PERIODIC_YZ above it.
static int split_sweep_kind(int& kind, int& periodic, int Y, int Z) {
    if (kind & ~(0xff | TF_SWEEP_PERIODIC_YZ)) return fail(-1, "unknown sweep flags in kind 0x%x", kind);
    periodic = (kind & TF_SWEEP_PERIODIC_YZ) ? 1 : 0;
    kind &= 0xff;
    if (periodic) {
        if (kind != TF_KIND_SOLVER && kind != TF_KIND_MULTIPHASE && kind != TF_KIND_ELECTRODE &&
            kind != TF_KIND_MULTIPHASE_FACES)
            return fail(-1, "TF_SWEEP_PERIODIC_YZ applies to the solver, multiphase and electrode kinds");
        if ((Y | Z) & 1)
            return fail(-1, "TF_SWEEP_PERIODIC_YZ needs even Y and Z (got %d, %d)", Y, Z);
    }
    return 0;
}

// Checks shared by tf_sweep and tf_sweep_launch_config; fills the shape.
static int sweep_prepare(const float* hm_lut, int n_classes, int kind, int B, int X, int Y, int Z,
                         int n_half, const tf_tile* tile, tf_tile* t) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (kind != TF_KIND_SOLVER && kind != TF_KIND_MULTIPHASE && kind != TF_KIND_FIELD &&
        kind != TF_KIND_ANISOTROPIC && kind != TF_KIND_ELECTRODE && kind != TF_KIND_MULTIPHASE_FACES)
        return fail(-1, "unknown kind %d", kind);
    if (kind == TF_KIND_MULTIPHASE_FACES &&
        (!hm_lut || n_classes < TF_MPF_FACE_SLOTS + 2 || n_classes > TF_MPF_FACE_SLOTS + 2 * TF_MPF_MAX_FACTORS + 2))
        return fail(-1, "face-code sweep needs a table of %d face slots and 1..%d prefactor pairs (got %d floats)",
                    TF_MPF_FACE_SLOTS, TF_MPF_MAX_FACTORS + 1, n_classes);
    if (kind == TF_KIND_ELECTRODE && (!hm_lut || n_classes != TF_ELECTRODE_TABLE))
        return fail(-1, "electrode sweep needs the %d-entry prefactor table per volume", TF_ELECTRODE_TABLE);
    if (kind == TF_KIND_MULTIPHASE && (!hm_lut || n_classes < 1 || n_classes > 64))
        return fail(-1, "multiphase sweep needs hm_lut and 1..64 classes (got %d)", n_classes);
    if (kind == TF_KIND_ANISOTROPIC && (!hm_lut || n_classes != TF_ANISO_TABLE))
        return fail(-1, "anisotropic sweep needs the %d-entry prefactor table", TF_ANISO_TABLE);
    if (n_half < 1 || n_half > TF_MAX_FUSED_HALF_SWEEPS)
        return fail(-1, "n_half=%d out of range 1..%d", n_half, TF_MAX_FUSED_HALF_SWEEPS);
    default_tile(n_half, kind, Y, t);
    if (tile) {
        if (tile->waves > 0) t->waves = tile->waves;
        if (tile->rows_per_wave > 0) t->rows_per_wave = tile->rows_per_wave;
        if (tile->tx > 0) t->tx = tile->tx;
    }
    if (t->tx < 0 || t->tx > 65536) return fail(-1, "invalid tile tx=%d", t->tx);
    if (sweep_code_volume(X, Y, Z) >= (1LL << 31))
        return fail(-1, "volume too large: %lld bytes of sweep codes", sweep_code_volume(X, Y, Z));
    // a tile's loads address its planes with 32-bit offsets
    if ((2LL * n_half + 2) * (Y + 2) * (Z + 2) * 4 >= (1LL << 31) - 64)
        return fail(-1, "a %dx%d plane is too large for the sweep", Y, Z);
    return 0;
}

static int sweep_dispatch(int kind, int n_half, SweepArgs& a, int B, const tf_tile& t,
                          hipStream_t st) {
    if (kind == TF_KIND_SOLVER)
        return sweep_dispatch_solver(n_half, a, B, t.waves, t.rows_per_wave, st);
    if (kind == TF_KIND_MULTIPHASE)
        return sweep_dispatch_multiphase(n_half, a, B, t.waves, t.rows_per_wave, st);
    if (kind == TF_KIND_MULTIPHASE_FACES)
        return sweep_dispatch_mpfaces(n_half, a, B, t.waves, t.rows_per_wave, st);
    if (kind == TF_KIND_ANISOTROPIC)
        return sweep_dispatch_anisotropic(n_half, a, B, t.waves, t.rows_per_wave, st);
    if (kind == TF_KIND_ELECTRODE)
        return sweep_dispatch_electrode(n_half, a, B, t.waves, t.rows_per_wave, st);
    return sweep_dispatch_field(n_half, a, B, t.waves, t.rows_per_wave, st);
}

extern "C" int tf_sweep(const float* src, float* dst, const void* codes, const float* hm_lut,
                        int n_classes, int kind, int B, int X, int Y, int Z, float omega,
                        int first_colour, int n_half, const tf_tile* tile, void* stream) {
    tf_tile t;
    int periodic = 0;
    if (int rc = split_sweep_kind(kind, periodic, Y, Z)) return rc;
    if (int rc = sweep_prepare(hm_lut, n_classes, kind, B, X, Y, Z, n_half, tile, &t)) return rc;
    if (!src || !dst || !codes) return fail(-1, "null pointer argument");
    if (src == dst) return fail(-1, "tf_sweep needs distinct src and dst buffers");
    if (first_colour != 0 && first_colour != 1) return fail(-1, "first_colour must be 0 or 1");
    SweepArgs a = {};
    a.src = src; a.dst = dst; a.codes = codes; a.hm = hm_lut; a.ncls = n_classes;
    a.v = make_vol(X, Y, Z);
    a.w = omega;
    a.colour0 = first_colour;
    a.TX = t.tx;
    a.fsec = (long long)B * a.v.sB;
    a.periodic = periodic;
    return sweep_dispatch(kind, n_half, a, B, t, (hipStream_t)stream);
}

extern "C" int tf_sweep_launch_config(int kind, int n_classes, int B, int X, int Y, int Z,
                                      int n_half, const tf_tile* tile, tf_tile* out,
                                      int* n_tiles) {
    if (!out) return fail(-1, "null pointer argument");
    // the tables are not read here: any non-null pointer passes the checks
    static const float dummy = 0.0f;
    tf_tile t;
    int periodic = 0;
    if (int rc = split_sweep_kind(kind, periodic, Y, Z)) return rc;
    if (int rc = sweep_prepare(&dummy, n_classes, kind, B, X, Y, Z, n_half, tile, &t)) return rc;
    SweepArgs a = {};
    a.v = make_vol(X, Y, Z);
    a.ncls = n_classes;
    a.TX = t.tx;
    a.periodic = periodic;
    a.resolve_only = 1;
    if (int rc = sweep_dispatch(kind, n_half, a, B, t, nullptr)) return rc;
    out->waves = t.waves;
    out->rows_per_wave = t.rows_per_wave;
    out->tx = a.TX;
    if (n_tiles) *n_tiles = a.total;
    return 0;
}

// ---------------------------------------------------------------------------
// MultiPhase face codes
// ---------------------------------------------------------------------------
// One thread per padded voxel of the class codes C (x ghosts copy the
// adjacent class, y/z ghosts class 0 or, periodic, the opposite layer's),
// mapped to face classes by fmap: byte offsets 4 (4 a + b) of the faces
// towards x+1, y+1, z+1 in the 4 x 4 face table (0 where that neighbour is
// outside the padded volume), and for interior voxels the prefactor's pair
// offset from the neighbourhood key; ghosts get pair 0.
__global__ void k_face_codes(const uint8_t* __restrict__ C, const uint8_t* __restrict__ fmap,
                             const uint16_t* __restrict__ keys, uint32_t* __restrict__ out, Vol v,
                             int b0, int nbx) {
    // grid as k_setup; the seven class loads (and then the seven face-class
    // lookups) are unconditional, from offsets that fall back to the voxel
    // itself where a neighbour lies outside the padded volume or is not needed
    const int xp = (int)(blockIdx.x / (unsigned)nbx);
    const int r2 = (int)(blockIdx.x - (unsigned)xp * (unsigned)nbx) * (int)blockDim.x + (int)threadIdx.x;
    if (r2 >= v.sX) return;
    const int b = b0 + (int)blockIdx.y;
    const int r = xp * v.sX + r2;
    const int yp = r2 / v.sY, zp = r2 - yp * v.sY;
    const uint8_t* P = C + (size_t)b * v.sB + r;
    const bool inner = xp >= 1 && xp <= v.X && yp >= 1 && yp <= v.Y && zp >= 1 && zp <= v.Z;
    const int oxq = xp <= v.X ? v.sX : 0, oyq = yp <= v.Y ? v.sY : 0, ozq = zp <= v.Z ? 1 : 0;
    const int oxm = inner ? -v.sX : 0, oym = inner ? -v.sY : 0, ozm = inner ? -1 : 0;
    const uint32_t c = fmap[P[0]];
    const uint32_t cxq = fmap[P[oxq]], cyq = fmap[P[oyq]], czq = fmap[P[ozq]];
    const uint32_t cxm = fmap[P[oxm]], cym = fmap[P[oym]], czm = fmap[P[ozm]];
    const uint32_t fx = xp <= v.X ? 4u * (4u * c + cxq) : 0u;
    const uint32_t fy = yp <= v.Y ? 4u * (4u * c + cyq) : 0u;
    const uint32_t fz = zp <= v.Z ? 4u * (4u * c + czq) : 0u;
    uint32_t key = c;
    key = key * 4 + cxm;
    key = key * 4 + cxq;
    key = key * 4 + cym;
    key = key * 4 + cyq;
    key = key * 4 + czm;
    key = key * 4 + czq;
    key = key * 4 + (xp == 1 ? 1u : 0u) + (xp == v.X ? 2u : 0u);
    const uint32_t pk = keys[inner ? key : 0u];
    const uint32_t pair = inner ? pk : 0u;
    out[(size_t)b * v.sB + r] = fx | (fy << 6) | (fz << 12) | (pair << 18);
}

extern "C" int tf_multiphase_face_codes(const uint8_t* codes, const uint8_t* face_class, int n_classes,
                                        const uint16_t* key_table, uint32_t* face_codes, int B, int X,
                                        int Y, int Z, void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!codes || !face_class || !key_table || !face_codes) return fail(-1, "null pointer argument");
    if (n_classes < 1 || n_classes > 256) return fail(-1, "1..256 classes expected (got %d)", n_classes);
    Vol v = make_vol(X, Y, Z);
    int nbx;
    unsigned gx;
    if (int rc = set_up_grid(v, &nbx, &gx)) return rc;
    for (int b0 = 0; b0 < B; b0 += kMaxGridY) {
        hipLaunchKernelGGL(k_face_codes, dim3(gx, (unsigned)min(B - b0, kMaxGridY)), dim3(kSetupThreads), 0,
                           (hipStream_t)stream, codes, face_class, key_table, face_codes, v, b0, nbx);
        TF_LAUNCH_CHECK("k_face_codes");
    }
    return 0;
}

// ---------------------------------------------------------------------------
// per-layer float64 sums for the convergence check
// ---------------------------------------------------------------------------
// Summation order (fixed, so repeated calls and every kernel below give the
// same bits): the Y*Z positions of a layer, flattened as y*Z + z, are cut into
// chunks of kSumChunk; thread t of 256 adds the terms of positions
// chunk*kSumChunk + t + 256k, k = 0..kSumPerThread-1, in that order; each wave
// sums its lanes in a fixed shuffle tree and the four wave sums are added in
// wave order (the chunk partial); the layer sum is 0 + the chunk partials in
// chunk order.  tf_layer_sums and tf_batch_layer_sums run one block per layer
// over its chunks; tf_layer_sums_ws streams x-slabs of one chunk per block,
// reading each plane once, and adds the partials in a second kernel.
constexpr int kSumThreads = 256;
constexpr int kSumPerThread = 8;
constexpr int kSumChunk = kSumThreads * kSumPerThread;

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// flux term of a face between padded voxels with values c (x side) and cn
// (x+1 side), codes k0 and k1: vertical_flux of each solver kind
// (taufactor/taufactor.py:412-419, 615-620; anisotropic: faces of voxels
// whose factor exceeds 8 are masked, :417-418)
template <int KIND>
__device__ __forceinline__ float face_flux(float c, float cn, code_t<KIND> k0, code_t<KIND> k1,
                                           const float* __restrict__ hm, int ncls) {
    float d = cn - c;
    if constexpr (KIND == TF_KIND_SOLVER) {
        if (k0 == 0 || k1 == 0) d = 0.0f;
    } else if constexpr (KIND == TF_KIND_ANISOTROPIC) {
        if (hm[k0] > 8.0f || hm[k1] > 8.0f) d = 0.0f;
    } else if constexpr (KIND == TF_KIND_MULTIPHASE) {
        d = hm[k0 * ncls + k1] * d;
    } else {
        d = hmean(k0, k1) * d;
    }
    return d;
}

// electrode flux of layer x: -c[x] + c[x-1], (1 - c)/0.5 at x = 0, zero where
// c or c[x-1] is 0 (ElectrodeSolver.compute_metrics, taufactor/electrode.py:80-83)
__device__ __forceinline__ float electrode_flux(float c, float cp, int x) {
    float d = x == 0 ? (1.0f - c) / 0.5f : -c + cp;
    if (c == 0.0f || cp == 0.0f) d = 0.0f;
    return d;
}

// The two chunk partials of a block (accC, accF per thread) into thread 0,
// in the fixed order.  `slot` alternates between consecutive calls, so one
// barrier per call suffices.
__device__ __forceinline__ void chunk_reduce(double accC, double accF, int slot, double& sc,
                                             double& sf) {
    __shared__ double red[2][2][kSumThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    accC = wave_sum(accC);
    accF = wave_sum(accF);
    if (lane == 0) {
        red[slot][0][wave] = accC;
        red[slot][1][wave] = accF;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double c = red[slot][0][0], f = red[slot][1][0];
        for (int i = 1; i < kSumThreads / 64; ++i) {
            c += red[slot][0][i];
            f += red[slot][1][i];
        }
        sc = c;
        sf = f;
    }
}

// Sums of layer x of one padded volume V, one block looping over the chunks;
// returns them in thread 0.
template <int KIND>
__device__ __forceinline__ void layer_sum_block(const float* __restrict__ V,
                                                const code_t<KIND>* __restrict__ Cv,
                                                const float* __restrict__ hm, int ncls,
                                                const Vol v, int x, double& sum_conc,
                                                double& sum_flux) {
    const bool has_face = KIND == TF_KIND_ELECTRODE || x < v.X - 1;
    const int n = v.Y * v.Z;
    const int plane = (x + 1) * v.sX;
    double totC = 0.0, totF = 0.0;
    for (int ch = 0, slot = 0; ch * kSumChunk < n; ++ch, slot ^= 1) {
        double accC = 0.0, accF = 0.0;
#pragma unroll
        for (int k = 0; k < kSumPerThread; ++k) {
            const int p = ch * kSumChunk + (int)threadIdx.x + kSumThreads * k;
            if (p >= n) break;
            const int y = p / v.Z, z = p - y * v.Z;
            const int o = plane + (y + 1) * v.sY + z + 1;
            const float c = V[o];
            accC += (double)c;
            if constexpr (KIND == TF_KIND_ELECTRODE) {
                accF += (double)electrode_flux(c, V[o - v.sX], x);
            } else if (has_face) {
                accF += (double)face_flux<KIND>(c, V[o + v.sX], Cv[o], Cv[o + v.sX], hm, ncls);
            }
        }
        double sc = 0.0, sf = 0.0;
        chunk_reduce(accC, accF, slot, sc, sf);
        if (threadIdx.x == 0) {
            totC += sc;
            totF += sf;
        }
    }
    sum_conc = totC;
    sum_flux = totF;
}

template <int KIND>
__global__ __launch_bounds__(kSumThreads) void k_layer_sums(const float* __restrict__ F,
                                                            const code_t<KIND>* __restrict__ C,
                                                            const float* __restrict__ hm, int ncls,
                                                            Vol v, double* __restrict__ flux,
                                                            double* __restrict__ conc) {
    const int x = blockIdx.x, b = blockIdx.y;
    double sc = 0.0, sf = 0.0;
    layer_sum_block<KIND>(F + (size_t)b * v.sB, C + (size_t)b * v.sB, hm, ncls, v, x, sc, sf);
    if (threadIdx.x == 0) {
        conc[(size_t)b * v.X + x] = sc;
        if (KIND == TF_KIND_ELECTRODE)
            flux[(size_t)b * v.X + x] = sf;
        else if (x < v.X - 1)
            flux[(size_t)b * (v.X - 1) + x] = sf;
    }
}

// Streaming form: block (chunk, x-slab, b) walks the layers [x0, x0 + XS) of
// one chunk along x, keeping the previous plane's values and codes in
// registers, so each plane is read once (plus one plane per slab); the loads
// of the next layer's plane are issued before the current layer's sums.  Each
// wave writes its tree sum of every layer, part[sel][b*X + x][chunk][wave]
// (sel 0 concentration, 1 flux); the waves never wait for each other, and
// k_layer_sums_finish adds the wave sums and chunk partials in the order of
// chunk_reduce and layer_sum_block.
template <int KIND>
__global__ __launch_bounds__(kSumThreads) void k_layer_sums_stream(
    const float* __restrict__ F, const code_t<KIND>* __restrict__ C, const float* __restrict__ hm,
    int ncls, Vol v, int XS, double* __restrict__ part) {
    const int ch = blockIdx.x, nch = gridDim.x, b = blockIdx.z;
    const int x0 = blockIdx.y * XS, x1 = min(x0 + XS, v.X);
    if (x0 >= x1) return;
    const float* V = F + (size_t)b * v.sB;
    const code_t<KIND>* Cv = C + (size_t)b * v.sB;
    const int n = v.Y * v.Z;
    int off[kSumPerThread];
#pragma unroll
    for (int k = 0; k < kSumPerThread; ++k) {
        const int p = ch * kSumChunk + (int)threadIdx.x + kSumThreads * k;
        const int y = p / v.Z, z = p - y * v.Z;
        off[k] = p < n ? (y + 1) * v.sY + z + 1 : -1;
    }
    // electrode: layer x pairs padded planes x+1 and x; the others x+1, x+2
    // (for x = X-1 that is the ghost plane X+1, read and not used)
    constexpr bool kE = KIND == TF_KIND_ELECTRODE;
    float cq[kSumPerThread], cn[kSumPerThread];
    code_t<KIND> kq[kSumPerThread], kn[kSumPerThread];
    auto load = [&](int plane, float* c, code_t<KIND>* kc) {
#pragma unroll
        for (int k = 0; k < kSumPerThread; ++k) {
            const int o = plane * v.sX + max(off[k], 0);
            c[k] = V[o];
            kc[k] = kE ? code_t<KIND>(0) : Cv[o];
        }
    };
    load(kE ? x0 : x0 + 1, cq, kq);
    load(x0 + (kE ? 1 : 2), cn, kn);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int W = kSumThreads / 64;
    const size_t BX = (size_t)gridDim.z * v.X;
    double* const pC = part + ((size_t)b * v.X * nch + ch) * W + wave;
    double* const pF = pC + BX * nch * W;
    for (int x = x0; x < x1; ++x) {
        float cn2[kSumPerThread];
        code_t<KIND> kn2[kSumPerThread];
        if (x + 1 < x1) load(x + 1 + (kE ? 1 : 2), cn2, kn2);
        const bool face = kE || x < v.X - 1;
        double accC = 0.0, accF = 0.0;
#pragma unroll
        for (int k = 0; k < kSumPerThread; ++k) {
            if (off[k] < 0) break;
            if constexpr (kE) {
                accC += (double)cn[k];
                accF += (double)electrode_flux(cn[k], cq[k], x);
            } else {
                accC += (double)cq[k];
                if (face) accF += (double)face_flux<KIND>(cq[k], cn[k], kq[k], kn[k], hm, ncls);
            }
        }
        accC = wave_sum(accC);
        accF = wave_sum(accF);
        if (lane == 0) {
            pC[(size_t)x * nch * W] = accC;
            pF[(size_t)x * nch * W] = accF;
        }
#pragma unroll
        for (int k = 0; k < kSumPerThread; ++k) {
            cq[k] = cn[k];
            kq[k] = kn[k];
            cn[k] = cn2[k];
            kn[k] = kn2[k];
        }
    }
}

// layer sums = 0 + the chunk partials in chunk order, each the wave sums in
// wave order: one 64-thread block per layer stages the partials in LDS and
// its first lane adds them
template <int KIND>
__global__ __launch_bounds__(64) void k_layer_sums_finish(const double* __restrict__ part, int B,
                                                          int X, int nch, double* __restrict__ flux,
                                                          double* __restrict__ conc) {
    constexpr int W = kSumThreads / 64, kStage = 1024;
    __shared__ double st[2][kStage];
    const int i = blockIdx.x;
    const size_t BX = (size_t)B * X;
    const int n = nch * W;
    const double* pc = part + (size_t)i * n;
    const double* pf = part + (BX + i) * (size_t)n;
    double sc = 0.0, sf = 0.0;
    for (int base = 0; base < n; base += kStage) {
        const int m = min(kStage, n - base);
        for (int j = threadIdx.x; j < m; j += 64) {
            st[0][j] = pc[base + j];
            st[1][j] = pf[base + j];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int j = 0; j < m; j += W) {
                double cc = st[0][j], cf = st[1][j];
                for (int w = 1; w < W; ++w) {
                    cc += st[0][j + w];
                    cf += st[1][j + w];
                }
                sc += cc;
                sf += cf;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int b = i / X, x = i - b * X;
        conc[i] = sc;
        if (KIND == TF_KIND_ELECTRODE)
            flux[i] = sf;
        else if (x < X - 1)
            flux[(size_t)b * (X - 1) + x] = sf;
    }
}

static int layer_sums_checks(const float* field, const void* codes, const float* hm_lut,
                             int n_classes, int kind, int B, int X, int Y, int Z,
                             const double* flux_sum, const double* conc_sum) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!field || !codes || !conc_sum || (X > 1 && !flux_sum)) return fail(-1, "null pointer argument");
    if (kind == TF_KIND_MULTIPHASE && (!hm_lut || n_classes < 1 || n_classes > 256))
        return fail(-1, "multiphase sums need hm_lut and 1..256 classes");
    if (kind == TF_KIND_ANISOTROPIC && (!hm_lut || n_classes != TF_ANISO_TABLE))
        return fail(-1, "anisotropic sums need the %d-entry prefactor table", TF_ANISO_TABLE);
    if (kind != TF_KIND_SOLVER && kind != TF_KIND_MULTIPHASE && kind != TF_KIND_FIELD &&
        kind != TF_KIND_ELECTRODE && kind != TF_KIND_ANISOTROPIC)
        return fail(-1, "unknown kind %d", kind);
    if (X > 65535 || B > 65535) return fail(-1, "grid too large for layer sums");
    return 0;
}

// kernel launch for every kind: L(kernel<KIND>, codes cast to its type)
template <class L>
static void by_kind(int kind, const void* codes, L&& launch) {
    switch (kind) {
    case TF_KIND_SOLVER: launch(std::integral_constant<int, TF_KIND_SOLVER>{}, static_cast<const uint8_t*>(codes)); break;
    case TF_KIND_MULTIPHASE: launch(std::integral_constant<int, TF_KIND_MULTIPHASE>{}, static_cast<const uint8_t*>(codes)); break;
    case TF_KIND_FIELD: launch(std::integral_constant<int, TF_KIND_FIELD>{}, static_cast<const float*>(codes)); break;
    case TF_KIND_ELECTRODE: launch(std::integral_constant<int, TF_KIND_ELECTRODE>{}, static_cast<const uint8_t*>(codes)); break;
    default: launch(std::integral_constant<int, TF_KIND_ANISOTROPIC>{}, static_cast<const uint8_t*>(codes)); break;
    }
}

extern "C" int tf_layer_sums(const float* field, const void* codes, const float* hm_lut,
                             int n_classes, int kind, int B, int X, int Y, int Z, double* flux_sum,
                             double* conc_sum, void* stream) {
    if (int rc = layer_sums_checks(field, codes, hm_lut, n_classes, kind, B, X, Y, Z, flux_sum, conc_sum))
        return rc;
    Vol v = make_vol(X, Y, Z);
    hipStream_t st = (hipStream_t)stream;
    by_kind(kind, codes, [&](auto kc, auto cp) {
        constexpr int K = decltype(kc)::value;
        hipLaunchKernelGGL(k_layer_sums<K>, dim3(X, B), dim3(kSumThreads), 0, st, field, cp, hm_lut,
                           n_classes, v, flux_sum, conc_sum);
    });
    TF_LAUNCH_CHECK("k_layer_sums");
    return 0;
}

static long long layer_sum_chunks(int Y, int Z) {
    return ((long long)Y * Z + kSumChunk - 1) / kSumChunk;
}

extern "C" long long tf_layer_sums_workspace_bytes(int B, int X, int Y, int Z) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    return 2LL * B * X * layer_sum_chunks(Y, Z) * (kSumThreads / 64) * (long long)sizeof(double);
}

// The two kernels of the streaming layer sums, partials in `part` (every
// entry the finish kernel reads is written by the stream kernel first, so
// `part` needs no clearing).
static int layer_sums_stream_launch(const float* field, const void* codes, const float* hm_lut,
                                    int n_classes, int kind, int B, int X, int Y, int Z,
                                    double* flux_sum, double* conc_sum, double* part,
                                    void* stream) {
    const long long nch = layer_sum_chunks(Y, Z);
    if (nch > 0x7fffffff) return fail(-1, "too many layer-sum chunks");
    // x slabs: long enough that the extra plane per slab is a few percent,
    // short enough for a few waves of blocks per CU
    int xs = 32;
    while (xs > 4 && nch * ((X + xs - 1) / xs) * B < 4096) xs >>= 1;
    Vol v = make_vol(X, Y, Z);
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((unsigned)nch, (X + xs - 1) / xs, B);
    const int fin_blocks = B * X;
    by_kind(kind, codes, [&](auto kc, auto cp) {
        constexpr int K = decltype(kc)::value;
        hipLaunchKernelGGL(k_layer_sums_stream<K>, grid, dim3(kSumThreads), 0, st, field, cp, hm_lut,
                           n_classes, v, xs, part);
        hipLaunchKernelGGL(k_layer_sums_finish<K>, dim3(fin_blocks), dim3(64), 0, st, part, B, X,
                           (int)nch, flux_sum, conc_sum);
    });
    TF_LAUNCH_CHECK("k_layer_sums_stream");
    return 0;
}

extern "C" int tf_layer_sums_ws(const float* field, const void* codes, const float* hm_lut,
                                int n_classes, int kind, int B, int X, int Y, int Z,
                                double* flux_sum, double* conc_sum, void* workspace,
                                long long workspace_bytes, void* stream) {
    if (int rc = layer_sums_checks(field, codes, hm_lut, n_classes, kind, B, X, Y, Z, flux_sum, conc_sum))
        return rc;
    const long long need = tf_layer_sums_workspace_bytes(B, X, Y, Z);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "layer-sum workspace of %lld bytes needed (got %lld)", need, workspace_bytes);
    return layer_sums_stream_launch(field, codes, hm_lut, n_classes, kind, B, X, Y, Z, flux_sum,
                                    conc_sum, static_cast<double*>(workspace), stream);
}

// ---------------------------------------------------------------------------
// layer sums without allocation: the partial sums live in a static device
// array of TF_NOALLOC_SCRATCH_BYTES, one copy per device, held until the
// process ends.  Its host-side name is not a device address: the address is
// fetched once per device with hipGetSymbolAddress (that device current) and
// only that address is handed to kernels.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[TF_NOALLOC_SCRATCH_BYTES / sizeof(double)];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices];

// device address of `scratch` on the current device (looked up on first use)
static int scratch_address(hipStream_t st, double** out) {
    int dev = 0;
    if (int rc = hip_status(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (dev < 0 || dev >= kMaxDevices) return fail(-1, "device ordinal %d out of range", dev);
    if (st) {
        hipDevice_t sdev = 0;
        if (int rc = hip_status(hipStreamGetDevice(st, &sdev), "hipStreamGetDevice")) return rc;
        if ((int)sdev != dev)
            return fail(-1, "stream belongs to device %d but device %d is current", (int)sdev, dev);
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (int rc = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                                "hipGetSymbolAddress(scratch)"))
            return rc;
        if (!p) return fail(-1, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" int tf_layer_sums_noalloc(const float* field, const void* codes, const float* hm_lut,
                                     int n_classes, int kind, int B, int X, int Y, int Z,
                                     double* flux_sum, double* conc_sum, void* stream) {
    if (int rc = layer_sums_checks(field, codes, hm_lut, n_classes, kind, B, X, Y, Z, flux_sum, conc_sum))
        return rc;
    const long long need = tf_layer_sums_workspace_bytes(B, X, Y, Z);
    if (need > TF_NOALLOC_SCRATCH_BYTES)
        return fail(-1, "tf_layer_sums_noalloc: these sizes need %lld bytes of partial sums, more than "
                        "the %lld-byte static scratch array (use tf_layer_sums_ws)",
                    need, (long long)TF_NOALLOC_SCRATCH_BYTES);
    double* part = nullptr;
    if (int rc = scratch_address((hipStream_t)stream, &part)) return rc;
    return layer_sums_stream_launch(field, codes, hm_lut, n_classes, kind, B, X, Y, Z, flux_sum,
                                    conc_sum, part, stream);
}

// ---------------------------------------------------------------------------
// per-voxel diffusivity field: set-up
// ---------------------------------------------------------------------------
// kpad[b][x][y][z] = k of the voxel, x ghosts copy the adjacent voxel's k,
// y/z ghosts 0 (the reference's diff_map padding, taufactor/taufactor.py:590-592);
// both field buffers get the initial field, conductive = k > 0.
// Same grid as k_setup.
__global__ void k_setup_field(const float* __restrict__ k, const float* __restrict__ xinit,
                              float* __restrict__ kpad, float* __restrict__ fa,
                              float* __restrict__ fb, Vol v, int b0, int nbx) {
    const int xp = (int)(blockIdx.x / (unsigned)nbx);
    const int r2 = (int)(blockIdx.x - (unsigned)xp * (unsigned)nbx) * (int)blockDim.x + (int)threadIdx.x;
    if (r2 >= v.sX) return;
    const int b = b0 + (int)blockIdx.y;
    const int r = xp * v.sX + r2;
    const int yp = r2 / v.sY;
    const int zp = r2 - yp * v.sY;
    const float* K = k + (size_t)b * v.X * v.Y * v.Z;
    const bool yz_in = (yp >= 1 && yp <= v.Y && zp >= 1 && zp <= v.Z);
    const bool x_in = xp >= 1 && xp <= v.X;
    // one unconditional load: the voxel, the adjacent layer's for an x ghost,
    // a clamped one (unused) for a y/z ghost
    const int xs = xp == 0 ? 0 : (x_in ? xp - 1 : v.X - 1);
    const int ys = min(max(yp - 1, 0), v.Y - 1), zs = min(max(zp - 1, 0), v.Z - 1);
    const float kl = K[((size_t)xs * v.Y + ys) * v.Z + zs];
    const float xi = xinit[xs];
    const float kk = yz_in ? kl : 0.0f;
    const float f = !yz_in ? 0.0f : x_in ? (kl > 0.0f ? 1.0f : 0.0f) * xi : (xp == 0 ? -1.0f : 1.0f);
    const size_t o = (size_t)b * v.sB + r;
    kpad[o] = kk;
    fa[o] = f;
    fb[o] = f;
}

// voxels with k > 0 per x layer, as k_layer_count (4-byte elements: up to
// three before a 16-byte boundary)
__global__ void __launch_bounds__(kCountThreads)
k_layer_count_field(const float* __restrict__ k, uint32_t* __restrict__ cnt, Vol v, int b0) {
    const int x = blockIdx.x, b = b0 + (int)blockIdx.y;
    const long long n = (long long)v.Y * v.Z;
    const float* K = k + ((size_t)b * v.X + x) * (size_t)n;
    const long long head = min(n, (long long)(((16 - ((uintptr_t)K & 15)) & 15) / 4));
    const long long nv = (n - head) / 4;
    const long long tail = head + nv * 4;
    uint32_t acc = 0;
    if (threadIdx.x < head) acc += K[threadIdx.x] > 0.0f;
    const float4* W = reinterpret_cast<const float4*>(K + head);
    for (long long i = threadIdx.x; i < nv; i += kCountThreads) {
        const float4 w = W[i];
        acc += (uint32_t)(w.x > 0.0f) + (w.y > 0.0f) + (w.z > 0.0f) + (w.w > 0.0f);
    }
    if (tail + threadIdx.x < n) acc += K[tail + threadIdx.x] > 0.0f;
    const uint32_t total = workgroup_sum(acc);
    if (threadIdx.x == 0) cnt[(size_t)b * v.X + x] = total;
}

extern "C" int tf_setup_field(const float* k, const float* xinit, float* kpad, float* field_a,
                              float* field_b, uint32_t* layer_count, int B, int X, int Y, int Z,
                              void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!k || !xinit || !kpad || !field_a || !field_b || !layer_count)
        return fail(-1, "null pointer argument");
    if (field_a == field_b) return fail(-1, "field_a and field_b must be distinct buffers");
    Vol v = make_vol(X, Y, Z);
    hipStream_t st = (hipStream_t)stream;
    int nbx;
    unsigned gx;
    if (int rc = set_up_grid(v, &nbx, &gx)) return rc;
    if ((long long)X * kCountThreads >= (1LL << 32)) return fail(-1, "volume too long in x for the layer count");
    for (int b0 = 0; b0 < B; b0 += kMaxGridY) {
        const unsigned nb = (unsigned)min(B - b0, kMaxGridY);
        hipLaunchKernelGGL(k_setup_field, dim3(gx, nb), dim3(kSetupThreads), 0, st, k, xinit, kpad, field_a,
                           field_b, v, b0, nbx);
        TF_LAUNCH_CHECK("k_setup_field");
        hipLaunchKernelGGL(k_layer_count_field, dim3(X, nb), dim3(kCountThreads), 0, st, k, layer_count, v, b0);
        TF_LAUNCH_CHECK("k_layer_count_field");
    }
    return 0;
}

// Face conductances of the per-voxel field for the sweep: for every padded
// voxel o, gx[o] = hm(k[o], k[o + x]), gy[o] = hm(k[o], k[o + y]),
// gz[o] = hm(k[o], k[o + z]) (0 where the neighbour is outside the padded
// volume).  hm(a, b) = ((2a) b)/(a + b) rounds the same either way round
// (doubling is exact, the product is rounded once), so the face of voxels u
// and v is what the update of either computes (MultiPhaseSolver's harmonic
// mean, taufactor/taufactor.py:577-583).
__global__ void k_field_faces(const float* __restrict__ kpad, float* __restrict__ faces, Vol v, int b0,
                              int nbx, long long fsec) {
    // grid as k_setup; the neighbour loads are unconditional (offset 0, the
    // voxel itself, where the neighbour lies outside the padded volume)
    const int xp = (int)(blockIdx.x / (unsigned)nbx);
    const int r2 = (int)(blockIdx.x - (unsigned)xp * (unsigned)nbx) * (int)blockDim.x + (int)threadIdx.x;
    if (r2 >= v.sX) return;
    const int yp = r2 / v.sY, zp = r2 - yp * v.sY;
    const size_t o = (size_t)(b0 + (int)blockIdx.y) * v.sB + (size_t)xp * v.sX + r2;
    const bool xq = xp <= v.X, yq = yp <= v.Y, zq = zp <= v.Z;
    const float* K = kpad + o;
    const float kc = K[0], kx = K[xq ? v.sX : 0], ky = K[yq ? v.sY : 0], kz = K[zq ? 1 : 0];
    faces[o] = xq ? hmean(kc, kx) : 0.0f;
    faces[fsec + o] = yq ? hmean(kc, ky) : 0.0f;
    faces[2 * fsec + o] = zq ? hmean(kc, kz) : 0.0f;
}

extern "C" int tf_field_faces(const float* kpad, float* faces, int B, int X, int Y, int Z,
                              void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!kpad || !faces) return fail(-1, "null pointer argument");
    const Vol v = make_vol(X, Y, Z);
    const long long total = (long long)B * v.sB;
    int nbx;
    unsigned gx;
    if (int rc = set_up_grid(v, &nbx, &gx)) return rc;
    for (int b0 = 0; b0 < B; b0 += kMaxGridY) {
        hipLaunchKernelGGL(k_field_faces, dim3(gx, (unsigned)min(B - b0, kMaxGridY)), dim3(kSetupThreads), 0,
                           (hipStream_t)stream, kpad, faces, v, b0, nbx, total);
        TF_LAUNCH_CHECK("k_field_faces");
    }
    return 0;
}

// ---------------------------------------------------------------------------
// implicit-adjoint gradient of D_rel with respect to voxel diffusivities
// ---------------------------------------------------------------------------
// At the fixed point c of the discrete operator (faces g = hm(k_a, k_b),
// Dirichlet half-faces 2*k_v to bc = -+0.5), D_rel = (X/(Y Z)) * W with
// W = sum_faces g (dc)^2 + sum_x-faces 2 k_v (c_v - bc)^2 the dissipation at
// unit potential difference; c minimises W, so dW/dk_v is the explicit
// derivative (no adjoint solve needed):
//   k_v > 0: sum over internal faces (c_u - c_v)^2 * 2 k_u^2/(k_v+k_u)^2
//            + sum over its x faces 2 (c_v - bc)^2
//   k_v = 0: one-sided derivative; v joins at c* = the mean of its
//            conductive neighbours and Dirichlet faces (weights all 2):
//            sum 2 (c_u - c*)^2 + sum 2 (c* - bc)^2
// Evaluated in float64 from the float32 field.
// Every value is loaded first (the six neighbours of an interior voxel lie in
// the padded volume; an x ghost's values go unused), so that no load waits
// behind a branch on an earlier one; the terms and their order are those of
// the formula above.
template <typename KFUN>
__device__ __forceinline__ double voxel_grad(const float* __restrict__ V, int o, const Vol v, int x,
                                             KFUN kat) {
    const int offs[6] = {v.sX, -v.sX, v.sY, -v.sY, 1, -1};
    const double kv = kat(o);
    const double cv = V[o];
    double ku[6], cu[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        ku[i] = kat(o + offs[i]);
        cu[i] = V[o + offs[i]];
    }
    // the x neighbours of the end layers are the Dirichlet faces, added below
    bool face[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) face[i] = !((i == 0 && x == v.X - 1) || (i == 1 && x == 0)) && ku[i] > 0.0;
    double g = 0.0;
    if (kv > 0.0) {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            if (!face[i]) continue;
            const double d = cu[i] - cv;
            const double s = kv + ku[i];
            g += d * d * (2.0 * ku[i] * ku[i]) / (s * s);
        }
        if (x == 0) g += 2.0 * (cv + 0.5) * (cv + 0.5);
        if (x == v.X - 1) g += 2.0 * (cv - 0.5) * (cv - 0.5);
    } else {
        double sum = 0.0;
        int n = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            if (!face[i]) continue;
            sum += cu[i];
            ++n;
        }
        if (x == 0) { sum += -0.5; ++n; }
        if (x == v.X - 1) { sum += 0.5; ++n; }
        if (n == 0) return 0.0;
        const double cs = sum / n;
#pragma unroll
        for (int i = 0; i < 6; ++i)
            if (face[i]) g += 2.0 * (cu[i] - cs) * (cu[i] - cs);
        if (x == 0) g += 2.0 * (cs + 0.5) * (cs + 0.5);
        if (x == v.X - 1) g += 2.0 * (cs - 0.5) * (cs - 0.5);
    }
    return g;
}

// one thread per voxel; a block is 4 rows of 64 z (full waves wherever Z is
// a multiple of 64)
__global__ void k_grad_field(const float* __restrict__ F, const float* __restrict__ K, Vol v, int B,
                             double scale, float* __restrict__ grad) {
    const int z = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int bx = blockIdx.z;
    if (y >= v.Y) return;
    const int b = bx / v.X, x = bx - b * v.X;
    if (z >= v.Z) return;
    const float* V = F + (size_t)b * v.sB;
    const float* Kb = K + (size_t)b * v.sB;
    const int o = (x + 1) * v.sX + (y + 1) * v.sY + (z + 1);
    const double g = voxel_grad(V, o, v, x, [&](int i) { return (double)Kb[i]; });
    grad[(((size_t)b * v.X + x) * v.Y + y) * v.Z + z] = (float)(g * scale);
}

extern "C" int tf_grad_field(const float* field, const float* kpad, int B, int X, int Y, int Z,
                             float* grad, void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!field || !kpad || !grad) return fail(-1, "null pointer argument");
    if ((long long)B * X > 65535 || (Y + 3) / 4 > 65535) return fail(-1, "grid too large for gradient");
    Vol v = make_vol(X, Y, Z);
    const double scale = (double)X / ((double)Y * (double)Z);
    hipLaunchKernelGGL(k_grad_field, dim3((Z + 63) / 64, (Y + 3) / 4, B * X), dim3(256), 0,
                       (hipStream_t)stream, field, kpad, v, B, scale, grad);
    TF_LAUNCH_CHECK("k_grad_field");
    return 0;
}

// per-class sums of the voxel gradient for one x layer; one block per (x, b)
constexpr int kMaxGradClasses = 64;
__global__ __launch_bounds__(256) void k_grad_classes(const float* __restrict__ F,
                                                      const uint8_t* __restrict__ C,
                                                      const float* __restrict__ dcls, int ncls,
                                                      Vol v, double scale,
                                                      double* __restrict__ out) {
    const int x = blockIdx.x, b = blockIdx.y;
    const float* V = F + (size_t)b * v.sB;
    const uint8_t* Cv = C + (size_t)b * v.sB;
    __shared__ double part[256 / 64][kMaxGradClasses];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < (256 / 64) * kMaxGradClasses; i += 256) (&part[0][0])[i] = 0.0;
    __syncthreads();
    auto kat = [&](int i) { return (double)dcls[Cv[i]]; };
    for (int y = 0; y < v.Y; ++y) {
        for (int z0 = 0; z0 < v.Z; z0 += 256) {
            const int z = z0 + threadIdx.x;
            double g = 0.0;
            int cl = 0;
            if (z < v.Z) {
                const int o = (x + 1) * v.sX + (y + 1) * v.sY + (z + 1);
                cl = Cv[o];
                g = voxel_grad(V, o, v, x, kat);
            }
            // classes are few: reduce each class present in the wave in turn
            for (int c = 0; c < ncls; ++c) {
                double gc = cl == c ? g : 0.0;
                gc = wave_sum(gc);
                if (lane == 0) part[wave][c] += gc;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < ncls) {
        double s2 = 0.0;
        for (int w = 0; w < 256 / 64; ++w) s2 += part[w][threadIdx.x];
        out[((size_t)b * v.X + x) * ncls + threadIdx.x] = s2 * scale;
    }
}

// The same sums for up to NC classes with 16 waves per layer: wave w takes
// the rows y = w, w + 16, ..., its lanes z = lane, lane + 64, ...; each lane
// adds its voxels' gradients into one register per class, and the per-class
// totals are reduced once at the end (wave shuffle tree, then the 16 waves in
// order), so the result does not depend on scheduling.
template <int NC>
__global__ __launch_bounds__(1024) void k_grad_classes_reg(const float* __restrict__ F,
                                                           const uint8_t* __restrict__ C,
                                                           const float* __restrict__ dcls, int ncls,
                                                           Vol v, double scale,
                                                           double* __restrict__ out) {
    constexpr int kWaves = 16;
    const int x = blockIdx.x, b = blockIdx.y;
    const float* V = F + (size_t)b * v.sB;
    const uint8_t* Cv = C + (size_t)b * v.sB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    auto kat = [&](int i) { return (double)dcls[Cv[i]]; };
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
    for (int y = wave; y < v.Y; y += kWaves) {
        for (int z = lane; z < v.Z; z += 64) {
            const int o = (x + 1) * v.sX + (y + 1) * v.sY + (z + 1);
            const int cl = Cv[o];
            const double g = voxel_grad(V, o, v, x, kat);
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[c] += cl == c ? g : 0.0;
        }
    }
    __shared__ double part[kWaves][NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const double s = wave_sum(acc[c]);
        if (lane == 0) part[wave][c] = s;
    }
    __syncthreads();
    if ((int)threadIdx.x < ncls) {
        double s2 = 0.0;
        for (int w = 0; w < kWaves; ++w) s2 += part[w][threadIdx.x];
        out[((size_t)b * v.X + x) * ncls + threadIdx.x] = s2 * scale;
    }
}

extern "C" int tf_grad_classes(const float* field, const uint8_t* codes, const float* d_cls,
                               int n_classes, int B, int X, int Y, int Z, double* grad_layers,
                               void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!field || !codes || !d_cls || !grad_layers) return fail(-1, "null pointer argument");
    if (n_classes < 1 || n_classes > kMaxGradClasses) return fail(-1, "n_classes=%d out of range", n_classes);
    if (X > 65535 || B > 65535) return fail(-1, "grid too large for gradient");
    Vol v = make_vol(X, Y, Z);
    const double scale = (double)X / ((double)Y * (double)Z);
    const hipStream_t st = (hipStream_t)stream;
    const dim3 grid(X, B);
    if (n_classes <= 4) {
        hipLaunchKernelGGL(k_grad_classes_reg<4>, grid, dim3(1024), 0, st, field, codes, d_cls, n_classes, v,
                           scale, grad_layers);
    } else if (n_classes <= 8) {
        hipLaunchKernelGGL(k_grad_classes_reg<8>, grid, dim3(1024), 0, st, field, codes, d_cls, n_classes, v,
                           scale, grad_layers);
    } else {
        hipLaunchKernelGGL(k_grad_classes, grid, dim3(256), 0, st, field, codes, d_cls, n_classes, v, scale,
                           grad_layers);
    }
    TF_LAUNCH_CHECK("k_grad_classes");
    return 0;
}

static int resolve_tile(int n_half, const tf_tile* tile, tf_tile* t) {
    // batches: volumes of different sizes share one shape; 16-row tiles
    t->waves = 8;
    t->rows_per_wave = 2;
    t->tx = 64;
    if (tile) {
        if (tile->waves > 0) t->waves = tile->waves;
        if (tile->rows_per_wave > 0) t->rows_per_wave = tile->rows_per_wave;
        if (tile->tx > 0) t->tx = tile->tx;
    }
    if (t->tx < 1 || t->tx > 65536) return fail(-1, "invalid tile tx=%d", t->tx);
    t->tx = (t->tx + 1) & ~1;  // the register-blocked kernel needs an even x tile
    if (t->waves * t->rows_per_wave <= 2 * n_half)
        return fail(-1, "workgroup shape %dx%d has no interior rows for n_half=%d", t->waves,
                    t->rows_per_wave, n_half);
    return 0;
}

extern "C" int tf_sweep_tile_count(int X, int Y, int Z, int n_half, const tf_tile* tile, int* count) {
    if (!count) return fail(-1, "null pointer argument");
    if (X < 1 || Y < 1 || Z < 1) return fail(-1, "invalid shape %dx%dx%d", X, Y, Z);
    if (n_half < 1 || n_half > TF_MAX_FUSED_HALF_SWEEPS) return fail(-1, "n_half=%d out of range", n_half);
    tf_tile t;
    if (int rc = resolve_tile(n_half, tile, &t)) return rc;
    // the tiles k_sweep_batch decodes: wide tiles, then a narrow strip (one
    // wide tile per row tile for whole-z windows)
    if ((long long)(t.tx + 2 * n_half) * (Y + 2) * (Z + 2) * 4 >= (1LL << 31) - 64)
        return fail(-1, "x tile %d of %dx%d planes exceeds the sweep's 2 GiB tile span", t.tx, Y, Z);
    const long long ty = (long long)t.waves * t.rows_per_wave - 2 * n_half;
    const ZTiling zt = whole_z_window(Z, reg_tile_tz(n_half), false)
                           ? ZTiling{1, false}
                           : z_tiling(Z, reg_tile_tz(n_half), reg_tile_tzn(n_half));
    const long long nty = (Y + ty - 1) / ty;
    long long n = ((long long)zt.wide * nty + (zt.narrow ? (nty + 1) / 2 : 0)) * ((X + t.tx - 1) / t.tx);
    if (n >= (1LL << 28)) return fail(-1, "too many tiles");
    *count = (int)n;
    return 0;
}

extern "C" int tf_batch_sweep(float* field_a, float* field_b, const uint8_t* codes,
                              const tf_volume* vols, const int* tile_vol, int n_tiles,
                              const tf_volume_state* state, int src_buffer, int first_colour,
                              int n_half, const tf_tile* tile, void* stream) {
    if (!field_a || !field_b || !codes || !vols || !tile_vol || !state)
        return fail(-1, "null pointer argument");
    if (field_a == field_b) return fail(-1, "field_a and field_b must be distinct buffers");
    if (n_tiles < 1 || n_tiles >= (1 << 28)) return fail(-1, "n_tiles=%d out of range", n_tiles);
    if (first_colour != 0 && first_colour != 1) return fail(-1, "first_colour must be 0 or 1");
    if (src_buffer != 0 && src_buffer != 1) return fail(-1, "src_buffer must be 0 or 1");
    if (n_half < 1 || n_half > TF_MAX_FUSED_HALF_SWEEPS)
        return fail(-1, "n_half=%d out of range 1..%d", n_half, TF_MAX_FUSED_HALF_SWEEPS);
    tf_tile t;
    if (int rc = resolve_tile(n_half, tile, &t)) return rc;
    BatchArgs a = {};
    a.fa = field_a; a.fb = field_b; a.codes = codes; a.vols = vols; a.tile_vol = tile_vol;
    a.state = state; a.n_tiles = n_tiles; a.src_b = src_buffer; a.colour0 = first_colour;
    a.TX = t.tx;
    hipStream_t st = (hipStream_t)stream;
    return sweep_dispatch_batch(n_half, a, t.waves, t.rows_per_wave, st);
}

// Live tile list of one pass depth: the tiles of active volumes in tile
// order (one block; chunks of 1024 tiles, a prefix sum of the active flags
// per chunk), their count in queue[0] and the pass counters queue[1..]
// cleared.
constexpr int kCompactThreads = 1024;
__global__ __launch_bounds__(kCompactThreads) void k_batch_compact(
    const int* __restrict__ tile_vol, int n_tiles, const tf_volume_state* __restrict__ state,
    int* __restrict__ live, int* __restrict__ queue, int n_counters) {
    __shared__ int s_scan[kCompactThreads];
    __shared__ int s_base;
    const int tid = threadIdx.x;
    if (tid == 0) s_base = 0;
    __syncthreads();
    for (int c0 = 0; c0 < n_tiles; c0 += kCompactThreads) {
        const int t = c0 + tid;
        const int f = (t < n_tiles && state[tile_vol[t]].active) ? 1 : 0;
        s_scan[tid] = f;
        __syncthreads();
        for (int off = 1; off < kCompactThreads; off <<= 1) {  // inclusive scan
            const int v = tid >= off ? s_scan[tid - off] : 0;
            __syncthreads();
            s_scan[tid] += v;
            __syncthreads();
        }
        if (f) live[s_base + s_scan[tid] - 1] = t;
        __syncthreads();
        if (tid == kCompactThreads - 1) s_base += s_scan[tid];
        __syncthreads();
    }
    if (tid == 0) queue[0] = s_base;
    for (int i = tid; i < n_counters; i += kCompactThreads) queue[1 + i] = 0;
}

extern "C" int tf_batch_compact(const int* tile_vol, int n_tiles, const tf_volume_state* state,
                                int* live_tiles, int* queue, int n_counters, void* stream) {
    if (!tile_vol || !state || !live_tiles || !queue) return fail(-1, "null pointer argument");
    if (n_tiles < 1 || n_tiles >= (1 << 28)) return fail(-1, "n_tiles=%d out of range", n_tiles);
    if (n_counters < 0 || n_counters > (1 << 20)) return fail(-1, "n_counters=%d out of range", n_counters);
    hipLaunchKernelGGL(k_batch_compact, dim3(1), dim3(kCompactThreads), 0, (hipStream_t)stream, tile_vol,
                       n_tiles, state, live_tiles, queue, n_counters);
    TF_LAUNCH_CHECK("k_batch_compact");
    return 0;
}

extern "C" int tf_batch_sweep_queue(float* field_a, float* field_b, const uint8_t* codes,
                                    const tf_volume* vols, const int* tile_vol, const int* live_tiles,
                                    int* queue, int pass, int n_tiles, const tf_volume_state* state,
                                    int src_buffer, int first_colour, int n_half, const tf_tile* tile,
                                    void* stream) {
    if (!field_a || !field_b || !codes || !vols || !tile_vol || !state || !live_tiles || !queue)
        return fail(-1, "null pointer argument");
    if (field_a == field_b) return fail(-1, "field_a and field_b must be distinct buffers");
    if (n_tiles < 1 || n_tiles >= (1 << 28)) return fail(-1, "n_tiles=%d out of range", n_tiles);
    if (pass < 0 || pass >= (1 << 20)) return fail(-1, "pass=%d out of range", pass);
    if (first_colour != 0 && first_colour != 1) return fail(-1, "first_colour must be 0 or 1");
    if (src_buffer != 0 && src_buffer != 1) return fail(-1, "src_buffer must be 0 or 1");
    if (n_half < 1 || n_half > TF_MAX_FUSED_HALF_SWEEPS)
        return fail(-1, "n_half=%d out of range 1..%d", n_half, TF_MAX_FUSED_HALF_SWEEPS);
    tf_tile t;
    if (int rc = resolve_tile(n_half, tile, &t)) return rc;
    BatchArgs a = {};
    a.fa = field_a; a.fb = field_b; a.codes = codes; a.vols = vols; a.tile_vol = tile_vol;
    a.state = state; a.n_tiles = n_tiles; a.src_b = src_buffer; a.colour0 = first_colour;
    a.TX = t.tx;
    a.live = live_tiles;
    a.queue = queue;
    a.pass = pass;
    return sweep_dispatch_batch(n_half, a, t.waves, t.rows_per_wave, (hipStream_t)stream);
}

__global__ __launch_bounds__(kSumThreads) void k_batch_layer_sums(
    const float* __restrict__ fa, const float* __restrict__ fb, const uint8_t* __restrict__ codes,
    const tf_volume* __restrict__ vols, const int* __restrict__ layer_vol,
    const tf_volume_state* __restrict__ state, int src_b, double* __restrict__ flux,
    double* __restrict__ conc) {
    const int l = blockIdx.x;
    const int vi = layer_vol[l];
    if (!state[vi].active) return;
    const tf_volume vol = vols[vi];
    const Vol v = make_vol(vol.X, vol.Y, vol.Z);
    const int x = l - vol.sum_offset;
    const float* F = (src_b ? fb : fa) + vol.offset;
    double sc = 0.0, sf = 0.0;
    layer_sum_block<TF_KIND_SOLVER>(F, codes + vol.offset, nullptr, 0, v, x, sc, sf);
    if (threadIdx.x == 0) {
        conc[l] = sc;
        flux[l] = x < v.X - 1 ? sf : 0.0;
    }
}

extern "C" int tf_batch_layer_sums(const float* field_a, const float* field_b, const uint8_t* codes,
                                   const tf_volume* vols, const int* layer_vol, int n_layers,
                                   const tf_volume_state* state, int src_buffer, double* flux_sum,
                                   double* conc_sum, void* stream) {
    if (!field_a || !field_b || !codes || !vols || !layer_vol || !state || !flux_sum || !conc_sum)
        return fail(-1, "null pointer argument");
    if (n_layers < 1 || n_layers > (1 << 30)) return fail(-1, "n_layers=%d out of range", n_layers);
    if (src_buffer != 0 && src_buffer != 1) return fail(-1, "src_buffer must be 0 or 1");
    hipLaunchKernelGGL(k_batch_layer_sums, dim3(n_layers), dim3(kSumThreads), 0, (hipStream_t)stream,
                       field_a, field_b, codes, vols, layer_vol, state, src_buffer, flux_sum, conc_sum);
    TF_LAUNCH_CHECK("k_batch_layer_sums");
    return 0;
}

// Convergence decision of one volume, in float64, with the arithmetic of
// ThroughTransportSolver.compute_metrics / SORSolver.check_convergence as
// taufactor_amd evaluates them on the host (taufactor/taufactor.py:109-153,
// 293-331): layer-mean fluxes, max/min/sequential mean, relative error,
// D_rel, tau, the non-percolation rule, the NaN rule, the 2e-3 tau rule.
__global__ void k_batch_check(const tf_volume* __restrict__ vols, int nvol,
                              const double* __restrict__ flux, tf_volume_state* __restrict__ st,
                              int src_b, int iters_per_check, int iter_limit, double conv_crit) {
    const int vi = blockIdx.x * blockDim.x + threadIdx.x;
    if (vi >= nvol) return;
    tf_volume_state S = st[vi];
    if (!S.active) return;
    const tf_volume vol = vols[vi];
    S.iters += iters_per_check;
    const double n2 = (double)vol.Y * (double)vol.Z;
    const double* F = flux + vol.sum_offset;
    const int nf = vol.X - 1;
    double fmax = -__builtin_inf(), fmin = __builtin_inf(), acc = 0.0;
    for (int x = 0; x < nf; ++x) {
        const double f = F[x] / n2;
        fmax = f > fmax ? f : fmax;
        fmin = f < fmin ? f : fmin;
        acc += f;
    }
    const double mean = acc / (double)nf;
    double rel = fmax != 0.0 ? (fmax - fmin) / fmax : __builtin_nan("");
    double D_rel = mean * (double)vol.X / 1.0;
    double tau = D_rel != 0.0 ? vol.D_mean / D_rel : __builtin_nan("");
    if ((fmin == 0.0 || fmax == 0.0 || mean == 0.0) && !vol.spans) {
        rel = 0.0;
        D_rel = 0.0;
        tau = 0.0;
    }
    if (mean != mean) rel = 0.0;
    S.D_rel = D_rel;
    S.rel_err = rel;
    bool conv = false;
    if (!(rel < conv_crit)) {
        S.old_tau = tau;
    } else if (!(__builtin_fabs(tau - S.old_tau) < 2e-3)) {
        S.old_tau = tau;
    } else {
        conv = true;
        if (tau == 0.0) tau = __builtin_inf();
    }
    S.tau = tau;
    if (conv || S.iters >= iter_limit) {
        S.active = 0;
        S.converged = conv ? 1 : 0;
        S.cur = src_b;
    }
    st[vi] = S;
}

extern "C" int tf_batch_check(const tf_volume* vols, int nvol, const double* flux_sum,
                              tf_volume_state* state, int src_buffer, int iters_per_check,
                              int iter_limit, double conv_crit, void* stream) {
    if (!vols || !flux_sum || !state) return fail(-1, "null pointer argument");
    if (nvol < 1) return fail(-1, "nvol=%d out of range", nvol);
    if (src_buffer != 0 && src_buffer != 1) return fail(-1, "src_buffer must be 0 or 1");
    hipLaunchKernelGGL(k_batch_check, dim3((nvol + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       vols, nvol, flux_sum, state, src_buffer, iters_per_check, iter_limit, conv_crit);
    TF_LAUNCH_CHECK("k_batch_check");
    return 0;
}

// ---------------------------------------------------------------------------
// host: does a 6-connected conductive cluster span x?
// ---------------------------------------------------------------------------
// Breadth-first search from every conductive voxel of the x=0 face; the
// volume spans iff the search reaches the x=X-1 face.  Equivalent to asking
// whether some 6-connected component label occurs on both x faces.
extern "C" int tf_spans_x(const uint8_t* mask, int X, int Y, int Z) {
    if (!mask) return fail(-1, "null mask");
    if (X < 1 || Y < 1 || Z < 1) return fail(-1, "invalid shape %dx%dx%d", X, Y, Z);
    const long long n = (long long)X * Y * Z;
    if (n >= (1LL << 31)) return fail(-1, "volume too large for spans test");
    const int sy = Z, sx = Y * Z;
    std::vector<uint8_t> seen((size_t)n, 0);
    std::vector<int> queue;
    queue.reserve(1024);
    for (int i = 0; i < sx; ++i)
        if (mask[i]) {
            seen[i] = 1;
            queue.push_back(i);
        }
    size_t head = 0;
    while (head < queue.size()) {
        const int o = queue[head++];
        const int x = o / sx, r = o - x * sx, y = r / sy, z = r - y * sy;
        if (x == X - 1) return 1;
        int nb[6];
        int k = 0;
        if (x + 1 < X) nb[k++] = o + sx;
        if (x >= 1) nb[k++] = o - sx;
        if (y + 1 < Y) nb[k++] = o + sy;
        if (y >= 1) nb[k++] = o - sy;
        if (z + 1 < Z) nb[k++] = o + 1;
        if (z >= 1) nb[k++] = o - 1;
        for (int j = 0; j < k; ++j) {
            const int q = nb[j];
            if (mask[q] && !seen[q]) {
                seen[q] = 1;
                queue.push_back(q);
            }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// streaming copy (HBM calibration)
// ---------------------------------------------------------------------------
// Each lane moves four 16-byte vectors, 256 lanes apart, so a workgroup
// covers 16 KiB with every load issued before the first store; one pass,
// no grid-stride loop (the grid is n / 1024 workgroups).  Nontemporal loads
// and stores: the data passes through once (5.97 TB/s for 2 GiB on MI355X,
// 5.67 with the default cache policy).
constexpr int kCopyThreads = 256;
constexpr int kCopyPerLane = 4;
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kCopyThreads) void k_hbm_copy(const u32x4_t* __restrict__ s,
                                                           u32x4_t* __restrict__ d, long long n16) {
    const long long base = (long long)blockIdx.x * (kCopyThreads * kCopyPerLane) + threadIdx.x;
    u32x4_t v[kCopyPerLane];
#pragma unroll
    for (int k = 0; k < kCopyPerLane; ++k) {
        const long long i = base + (long long)k * kCopyThreads;
        if (i < n16) v[k] = __builtin_nontemporal_load(s + i);
    }
#pragma unroll
    for (int k = 0; k < kCopyPerLane; ++k) {
        const long long i = base + (long long)k * kCopyThreads;
        if (i < n16) __builtin_nontemporal_store(v[k], d + i);
    }
}

extern "C" int tf_hbm_copy(const void* src, void* dst, long long bytes, void* stream) {
    if (!src || !dst) return fail(-1, "null pointer argument");
    if (bytes < 0 || (bytes & 15)) return fail(-1, "tf_hbm_copy needs a multiple of 16 bytes (got %lld)", bytes);
    if (((uintptr_t)src | (uintptr_t)dst) & 15) return fail(-1, "tf_hbm_copy needs 16-byte aligned pointers");
    if (bytes == 0) return 0;
    const long long n16 = bytes / 16;
    const long long per = (long long)kCopyThreads * kCopyPerLane;
    const long long grid = (n16 + per - 1) / per;
    if (grid > 0x7fffffffLL) return fail(-1, "tf_hbm_copy: %lld bytes is too large", bytes);
    hipLaunchKernelGGL(k_hbm_copy, dim3((unsigned)grid), dim3(kCopyThreads), 0, (hipStream_t)stream,
                       static_cast<const u32x4_t*>(src), static_cast<u32x4_t*>(dst), n16);
    TF_LAUNCH_CHECK("k_hbm_copy");
    return 0;
}
