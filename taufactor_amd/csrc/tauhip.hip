// tauhip.hip — gfx950 kernels and C-ABI for the steady-state diffusion solve
// behind taufactor_amd.Solver / taufactor_amd.MultiPhaseSolver.
//
// Built with -ffp-contract=off: every float32 operation of the update is
// rounded on its own, in the order the reference's eager torch ops use
// (taufactor/taufactor.py:97-102, 176-181, 606-613), so iterates match the
// reference value for value.  Fused multiply-adds appear only where written
// explicitly (the exact-division sequence in div_by_count).
//
// Layout and conventions are documented in include/tauhip.h.

#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "tauhip.h"

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

static int hip_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    return fail((int)e, "%s: %s", what, hipGetErrorString(e));
}

#define TF_LAUNCH_CHECK(what)                                   \
    do {                                                        \
        hipError_t e_ = hipGetLastError();                      \
        if (e_ != hipSuccess) return hip_status(e_, what);      \
    } while (0)

extern "C" int tf_abi_version(void) { return TF_ABI_VERSION; }
extern "C" const char* tf_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------
// geometry of one padded volume
// ---------------------------------------------------------------------------
struct Vol {
    int X, Y, Z;     // interior extents
    int sY, sX;      // padded strides: sY = Z+2, sX = (Y+2)(Z+2)
    long long sB;    // padded volume size (X+2)(Y+2)(Z+2)
};

__host__ __device__ static inline Vol make_vol(int X, int Y, int Z) {
    Vol v;
    v.X = X; v.Y = Y; v.Z = Z;
    v.sY = Z + 2;
    v.sX = (Y + 2) * (Z + 2);
    v.sB = (long long)(X + 2) * v.sX;
    return v;
}

// sweep-code layout: rows of sweep_code_row(Z) bytes (a multiple of 4),
// byte z+3 of a row = code of padded voxel z (z in [0, Z+1]), other bytes 0
__host__ __device__ inline int sweep_code_row(int Z) { return ((Z + 5 + 3) / 4) * 4; }
__host__ __device__ inline long long sweep_code_volume(int X, int Y, int Z) {
    return (long long)(X + 2) * (Y + 2) * sweep_code_row(Z);
}

static int check_dims(int B, int X, int Y, int Z) {
    if (B < 1 || X < 1 || Y < 1 || Z < 1)
        return fail(-1, "invalid volume shape B=%d X=%d Y=%d Z=%d", B, X, Y, Z);
    // offsets inside one padded volume are 32-bit
    long long n = (long long)(X + 2) * (Y + 2) * (Z + 2);
    if (n >= (1LL << 31))
        return fail(-1, "padded volume of %lld voxels exceeds 2^31", n);
    return 0;
}

// ---------------------------------------------------------------------------
// raw buffer access: 32-bit offsets from a uniform base; an offset at or past
// the record count reads 0 and drops a write
// ---------------------------------------------------------------------------
typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr int kBufferFlags = 0x00020000;  // dword 3 of a raw buffer descriptor on gfx9

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kBufferFlags);
}
// byte_off: per-lane offset (range-checked against the record count);
// soff: wave-uniform offset added to the address
__device__ __forceinline__ float ld_f32(rsrc_t r, int byte_off, int soff = 0) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, soff, 0));
}
__device__ __forceinline__ uint32_t ld_u8(rsrc_t r, int byte_off, int soff = 0) {
    return __builtin_amdgcn_raw_buffer_load_b8(r, byte_off, soff, 0);
}
__device__ __forceinline__ void st_f32(float val, rsrc_t r, int byte_off, int soff = 0) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, val), r, byte_off, soff, 0);
}

// ---------------------------------------------------------------------------
// exact float32 division by a conductive-neighbour count
// ---------------------------------------------------------------------------
// q = s / d for d in {inf, 1..8}, correctly rounded (the value IEEE division
// gives).  One reciprocal product plus one residual correction with the
// correctly rounded reciprocal r = RN(1/d); numerators below 2^-100 (where
// the correction can be off for d = 6, and where the sign of zero matters)
// take the IEEE division.  Exhaustively checked against s/d over all finite
// float32 s for each d (tests/test_division.py samples that check).
__device__ __forceinline__ float div_by_count(float s, float r, float d, int code) {
    float q = s * r;
    if (code != 0) {
        float e = __builtin_fmaf(-q, d, s);
        q = __builtin_fmaf(e, r, q);
    }
    if (__builtin_expect(__builtin_fabsf(s) < 0x1p-100f, 0)) q = s / d;
    return q;
}

// table of (1/d, d) per Solver code; code 0 means factor = inf
__constant__ float c_rcp[9] = {0.0f, 1.0f, 0.5f, 1.0f / 3.0f, 0.25f,
                               0.2f, 1.0f / 6.0f, 1.0f / 7.0f, 0.125f};

// ---------------------------------------------------------------------------
// set-up kernels
// ---------------------------------------------------------------------------
// One thread per padded voxel.  Writes codes and both field buffers.
// mode 0: mask is 0/1 conductive flag, code = neighbour count
// mode 1: mask is a class index, code = class (x ghosts copy the adjacent
//         voxel's class, y/z ghosts class 0), conductive = cls_cond[class]
template <int MODE>
__global__ void k_setup(const uint8_t* __restrict__ mask, const uint8_t* __restrict__ cls_cond,
                        const float* __restrict__ xinit, uint8_t* __restrict__ codes,
                        float* __restrict__ fa, float* __restrict__ fb, Vol v, int B,
                        int periodic, int aniso) {
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)B * v.sB;
    if (gid >= total) return;
    int b = (int)(gid / v.sB);
    int r = (int)(gid - (long long)b * v.sB);
    int xp = r / v.sX;
    int r2 = r - xp * v.sX;
    int yp = r2 / v.sY;
    int zp = r2 - yp * v.sY;
    const uint8_t* M = mask + (size_t)b * v.X * v.Y * v.Z;
    auto m_at = [&](int x, int y, int z) -> int {  // interior 0-based
        return (int)M[((size_t)x * v.Y + y) * v.Z + z];
    };
    bool yz_in = (yp >= 1 && yp <= v.Y && zp >= 1 && zp <= v.Z);
    bool in = yz_in && xp >= 1 && xp <= v.X;
    float f;
    uint8_t code = 0;
    if (in) {
        int x = xp - 1, y = yp - 1, z = zp - 1;
        int m = m_at(x, y, z);
        int cond = MODE == 0 ? (m != 0) : MODE == 2 ? (m == 1) : (int)cls_cond[m];
        // mask * linspace in float32 (taufactor/taufactor.py:291)
        f = (cond ? 1.0f : 0.0f) * xinit[x];
        if (MODE == 2) {
            if (cond) {
                // conductive neighbours with x ghosts 2 (x = 0) and 0 (x = X+1),
                // reactive neighbours with every ghost 0; y/z wrap round when
                // periodic (taufactor/electrode.py:47-62, 133-151)
                auto at = [&](int xx, int yy, int zz) -> int {
                    if (periodic) {
                        yy = (yy + v.Y) % v.Y;
                        zz = (zz + v.Z) % v.Z;
                    } else if (yy < 0 || yy >= v.Y || zz < 0 || zz >= v.Z) {
                        return 0;
                    }
                    return m_at(xx, yy, zz);
                };
                int cn = 0, rn = 0;
                const int nb[6][3] = {{x + 1, y, z}, {x - 1, y, z}, {x, y + 1, z},
                                      {x, y - 1, z}, {x, y, z + 1}, {x, y, z - 1}};
                for (int k = 0; k < 6; ++k) {
                    const int xx = nb[k][0];
                    if (xx < 0) { cn += 2; continue; }
                    if (xx >= v.X) continue;
                    const int mm = at(xx, nb[k][1], nb[k][2]);
                    cn += mm == 1;
                    rn += mm == 2;
                }
                code = (uint8_t)(1 + 7 * cn + rn);
            }
        } else if (MODE == 0 && aniso) {
            if (cond) {
                // neighbour states x-,x+ (0/1, 2 for the x ghost), y-,y+,z-,z+
                // (0/1, ghosts 0): the terms AnisotropicSolver sums, in its
                // order (taufactor/taufactor.py:459-471)
                const int xm = (x >= 1) ? (m_at(x - 1, y, z) != 0) : 2;
                const int xq = (x + 1 < v.X) ? (m_at(x + 1, y, z) != 0) : 2;
                const int ym = (y >= 1) ? (m_at(x, y - 1, z) != 0) : 0;
                const int yq = (y + 1 < v.Y) ? (m_at(x, y + 1, z) != 0) : 0;
                const int zm = (z >= 1) ? (m_at(x, y, z - 1) != 0) : 0;
                const int zq = (z + 1 < v.Z) ? (m_at(x, y, z + 1) != 0) : 0;
                code = (uint8_t)(1 + 48 * xm + 16 * xq + 8 * ym + 4 * yq + 2 * zm + zq);
            }
        } else if (MODE == 0) {
            if (cond) {
                // padded mask: x ghosts 2, y/z ghosts 0 (taufactor/taufactor.py:404-410),
                // or y/z wrapped round (PeriodicSolver, taufactor/taufactor.py:493-499)
                int nn = 0;
                nn += (x + 1 < v.X) ? (m_at(x + 1, y, z) != 0) : 2;
                nn += (x >= 1) ? (m_at(x - 1, y, z) != 0) : 2;
                if (periodic) {
                    nn += m_at(x, (y + 1) % v.Y, z) != 0;
                    nn += m_at(x, (y + v.Y - 1) % v.Y, z) != 0;
                    nn += m_at(x, y, (z + 1) % v.Z) != 0;
                    nn += m_at(x, y, (z + v.Z - 1) % v.Z) != 0;
                } else {
                    nn += (y + 1 < v.Y) ? (m_at(x, y + 1, z) != 0) : 0;
                    nn += (y >= 1) ? (m_at(x, y - 1, z) != 0) : 0;
                    nn += (z + 1 < v.Z) ? (m_at(x, y, z + 1) != 0) : 0;
                    nn += (z >= 1) ? (m_at(x, y, z - 1) != 0) : 0;
                }
                code = (uint8_t)nn;  // 0 stays 0 = inf
            }
        } else {
            code = (uint8_t)m;
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
        if (MODE == 1 && yz_in && (xp == 0 || xp == v.X + 1)) {
            // diff_map[:,0] = diff_map[:,1] (taufactor/taufactor.py:591-592)
            int x = xp == 0 ? 0 : v.X - 1;
            code = (uint8_t)m_at(x, yp - 1, zp - 1);
        }
        if (MODE == 1 && periodic) {
            // y/z ghosts copy the opposite interior layer, after the x copy
            // (PeriodicMultiPhaseSolver, taufactor/taufactor.py:633-638)
            const int y = (yp == 0 ? v.Y : (yp == v.Y + 1 ? 1 : yp)) - 1;
            const int z = (zp == 0 ? v.Z : (zp == v.Z + 1 ? 1 : zp)) - 1;
            const int x = xp == 0 ? 0 : (xp == v.X + 1 ? v.X - 1 : xp - 1);
            code = (uint8_t)m_at(x, y, z);
        }
    }
    size_t o = (size_t)b * v.sB + r;
    codes[o] = code;
    fa[o] = f;
    fb[o] = f;
}

// conductive voxels per x layer; one block per (x, b)
template <int MODE>
__global__ void k_layer_count(const uint8_t* __restrict__ mask, const uint8_t* __restrict__ cls_cond,
                              uint32_t* __restrict__ cnt, Vol v) {
    int x = blockIdx.x, b = blockIdx.y;
    const uint8_t* M = mask + ((size_t)b * v.X + x) * v.Y * v.Z;
    int n = v.Y * v.Z;
    uint32_t acc = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int m = M[i];
        acc += MODE == 0 ? (m != 0) : MODE == 2 ? (m == 1) : cls_cond[m];
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
    long long total = (long long)B * v.sB;
    int threads = 256;
    long long blocks = (total + threads - 1) / threads;
    hipLaunchKernelGGL(k_setup<MODE>, dim3((unsigned)blocks), dim3(threads), 0, st,
                       mask, cls_cond, xinit, codes, fa, fb, v, B,
                       (flags & TF_SETUP_PERIODIC_YZ) ? 1 : 0, (flags & TF_SETUP_ANISOTROPIC) ? 1 : 0);
    TF_LAUNCH_CHECK("k_setup");
    hipLaunchKernelGGL(k_layer_count<MODE>, dim3(X, B), dim3(256), 0, st, mask, cls_cond, cnt, v);
    TF_LAUNCH_CHECK("k_layer_count");
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
__global__ void k_wrap_ghosts(float* __restrict__ F, Vol v, int B) {
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
    float* P = F + (size_t)bx * v.sX;
    P[yp * v.sY + zp] = P[ys * v.sY + zs];
}

extern "C" int tf_wrap_ghosts(float* field, int B, int X, int Y, int Z, void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!field) return fail(-1, "null pointer argument");
    if ((long long)B * (X + 2) > 65535) return fail(-1, "grid too large for ghost wrap");
    Vol v = make_vol(X, Y, Z);
    const int per_plane = 2 * (Y + 2) + 2 * Z;
    hipLaunchKernelGGL(k_wrap_ghosts, dim3((per_plane + 255) / 256, B * (X + 2)), dim3(256), 0,
                       (hipStream_t)stream, field, v, B);
    TF_LAUNCH_CHECK("k_wrap_ghosts");
    return 0;
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
                                   Vol v, int row_c, long long n_words) {
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
        w |= c << (8 * i);
    }
    out[gid] = w;
}

extern "C" long long tf_sweep_codes_bytes(int X, int Y, int Z) {
    if (X < 1 || Y < 1 || Z < 1) return fail(-1, "invalid shape %dx%dx%d", X, Y, Z);
    return sweep_code_volume(X, Y, Z);
}

extern "C" int tf_pack_sweep_codes(const uint8_t* codes, uint8_t* sweep_codes, int B, int X, int Y,
                                   int Z, void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!codes || !sweep_codes) return fail(-1, "null pointer argument");
    if (reinterpret_cast<uintptr_t>(sweep_codes) & 3) return fail(-1, "sweep_codes must be 4-byte aligned");
    const long long n_words = (long long)B * sweep_code_volume(X, Y, Z) / 4;
    const long long blocks = (n_words + 255) / 256;
    if (blocks >= (1LL << 31)) return fail(-1, "volume too large");
    hipLaunchKernelGGL(k_pack_sweep_codes, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       codes, reinterpret_cast<uint32_t*>(sweep_codes), make_vol(X, Y, Z),
                       sweep_code_row(Z), n_words);
    TF_LAUNCH_CHECK("k_pack_sweep_codes");
    return 0;
}

// ---------------------------------------------------------------------------
// the update of one voxel (shared by both sweep kernels)
// ---------------------------------------------------------------------------
// Solver (taufactor/taufactor.py:97-102, 177-181):
//   s = ((((c[x+1]+c[x-1])+c[y+1])+c[y-1])+c[z+1])+c[z-1]
//   s = s / nn ; s = s - c ; s = s * w ; c = c + s
// MultiPhase (taufactor/taufactor.py:598-613):
//   s = c[x+1]*gx+ + c[x-1]*gx- + c[y+1]*gy+ + c[y-1]*gy- + c[z+1]*gz+ + c[z-1]*gz-
//   f = gx- + gx+ + gy- + gy+ + gz- + gz+ ; f += gx- at x=0 ; f += gx+ at x=X-1 ;
//   f = inf if f == 0 ;  s = s / f ; s = s - c ; s = s * w ; c = c + s
struct Nbr {
    float c, xp, xm, yp, ym, zp, zm;
};

__device__ __forceinline__ float update_solver(const Nbr& n, int code, float w) {
    float s = n.xp + n.xm;
    s = s + n.yp;
    s = s + n.ym;
    s = s + n.zp;
    s = s + n.zm;
    float r = c_rcp[code];
    float d = code ? (float)code : __builtin_inff();
    s = div_by_count(s, r, d, code);
    s = s - n.c;
    s = s * w;
    return n.c + s;
}

__device__ __forceinline__ float update_multiphase(const Nbr& n, float gxp, float gxm, float gyp,
                                                   float gym, float gzp, float gzm, bool xfirst,
                                                   bool xlast, float w) {
    float s = n.xp * gxp + n.xm * gxm;
    s = s + n.yp * gyp;
    s = s + n.ym * gym;
    s = s + n.zp * gzp;
    s = s + n.zm * gzm;
    float f = gxm + gxp;
    f = f + gym;
    f = f + gyp;
    f = f + gzm;
    f = f + gzp;
    if (xfirst) f = f + gxm;
    if (xlast) f = f + gxp;
    if (f == 0.0f) f = __builtin_inff();
    s = s / f;
    s = s - n.c;
    s = s * w;
    return n.c + s;
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

// ---------------------------------------------------------------------------
// temporally blocked sweep: H half-sweeps per HBM pass
// ---------------------------------------------------------------------------
// A workgroup owns an interior tile [y0,y1) x [z0,z1) x [x0,x1) (padded
// coordinates) and streams it plane by plane along x through a ring of H+3
// planes in LDS, each plane extended by a halo in y and z.  Half-sweep h
// (h = 0..H-1, colour first_colour^(h&1)) runs on plane s-h at step s, on
// the tile grown by H-1-h voxels, so that after the last half-sweep the
// owned tile holds exactly the values of H successive reference iterations.
// Red-black colouring makes each half-sweep a Jacobi step within one colour,
// so reading src and writing dst (no in-place hazard between workgroups
// that recompute each other's halos) gives the in-place result.
//
// Compute mapping: one wavefront per tile row (y), one lane per pair of
// voxels in z, so an LDS row is a window of 128 voxels of z starting at ew
// and the colour pattern of a row is uniform across the wave.  An LDS row
// stores the first voxel of each pair (64), then the second (64): every LDS
// access of a half-sweep is lane-contiguous (no bank conflicts).  Each wave
// owns RPW rows (interleaved, NW even so that all rows of a wave share their
// colour parity), NW waves make EY = NW*RPW rows (tile interior TY = EY-2H).
// The window starts ML >= H voxels before the tile, with ML and the tile
// width TZ multiples of 4 so that z0 and ew are 1 mod 4.
//
// Memory mapping: a plane of the tile is read with one 16-byte load per lane
// covering two rows (lanes 0-31: row 2m, lanes 32-63: row 2m+1, four z
// each), codes with one 4-byte load per lane from the aligned sweep-code
// layout, and written back with one 8-byte store per lane and row (one pair
// each).  The x loop is unrolled by the ring length NP, and plane q lives in
// LDS slot and register slot (q - S0 + 1) mod NP: every slot is a
// compile-time constant, so each LDS access is one instruction with an
// immediate offset from one of a few per-lane base addresses.  Planes are
// loaded NP-1 steps ahead of their use.
//
// Workgroups are numbered so that consecutive tiles land on the same XCD
// (hardware dispatch sends workgroup i to XCD i mod 8): neighbouring tiles
// then share their halo lines in that XCD's L2.
struct SweepArgs {
    const float* src;
    float* dst;
    const void* codes;
    const float* hm;
    int ncls;
    Vol v;
    float w;
    int colour0;
    int TX;
    int nTz, nTy, nTx, total;
    // register-blocked kinds: tiles [0, nWide) are wide, the rest narrow
    // (the last strip of z, from z0N, nTyN tiles of two y tiles each)
    int nWide, nTyN, z0N;
    int resolve_only;  // set TX and the tile counts, launch nothing
};

// compile-time loop: f(std::integral_constant<int, I>{}) for I in [B, E)
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

template <int I>
using ic = std::integral_constant<int, I>;

typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

constexpr int kRowPairs = 64;           // lanes per row
constexpr int kRowLen = 2 * kRowPairs;  // voxels per LDS row
constexpr int kLdsPad = kRowLen * 4;    // bytes before the plane ring: keeps the row-below
                                        // bias of the base addresses non-negative
constexpr int kSlackBytes = TF_FIELD_SLACK_BYTES;

template <int H, int NW, int RPW>
struct SweepShape {
    static constexpr int EY = NW * RPW;
    static constexpr int TY = EY - 2 * H;
    static constexpr int ML = 4 * ((H + 3) / 4);                // z0 - ew
    static constexpr int TZ = ((kRowLen - ML - H) / 4) * 4;     // owned voxels of z
    static constexpr int NP = H + 3;         // planes in the LDS ring and in the register ring
    static constexpr int PS = EY * kRowLen;  // floats (and codes) per plane
    static constexpr int RPL = RPW / 2;      // row pairs loaded per wave and plane
};

// per-voxel auxiliary data: a one-byte code (Solver, MultiPhase) or the
// voxel diffusivity (per-voxel field)
template <int KIND>
using code_t = std::conditional_t<KIND == TF_KIND_FIELD, float, uint8_t>;

// 2ab/(a+b), 0 when a+b == 0, with the reference's float32 operation order
// ((2*a)*b)/(a+b) (MultiPhaseSolver._harmonic_mean, taufactor/taufactor.py:577-583)
__device__ __forceinline__ float hmean(float a, float b) {
    const float den = a + b;
    return den > 0.0f ? ((2.0f * a) * b) / den : 0.0f;
}

// (1/d, d) per Solver code for the sweep.  Code 0 (factor inf) uses (0, 0):
// the sequence of div_by_count then yields a zero whose sign can differ from
// s/inf only when s is -0, and c + w*(q - c) is the same float for either
// zero q whatever c is (c = +-0 gives +0, c != 0 gives c - w*c).
__constant__ float c_rd[18] = {0.0f, 0.0f, 1.0f, 1.0f, 0.5f, 2.0f, 1.0f / 3.0f, 3.0f, 0.25f, 4.0f,
                               0.2f, 5.0f, 1.0f / 6.0f, 6.0f, 1.0f / 7.0f, 7.0f, 0.125f, 8.0f};

template <int H, int KIND, int NW, int RPW>
__device__ __forceinline__ void sweep_tile(const float* __restrict__ S, float* __restrict__ D,
                                           const code_t<KIND>* __restrict__ Cg,
                                           const float* __restrict__ hm, int ncls, const Vol v,
                                           float w, int colour0, int TX, int tz, int ty, int tx) {
    using Sh = SweepShape<H, NW, RPW>;
    using CT = code_t<KIND>;
    constexpr int TY = Sh::TY, TZ = Sh::TZ, ML = Sh::ML, NP = Sh::NP, PS = Sh::PS, RPL = Sh::RPL;
    static_assert(TY >= 1, "tile has no interior rows");
    static_assert(NW % 2 == 0, "the rows of a wave must share their colour parity");
    static_assert(RPW % 2 == 0, "rows are loaded in pairs");
    extern __shared__ __align__(16) unsigned char smem[];
    float* sF = reinterpret_cast<float*>(smem + kLdsPad);
    CT* sC = reinterpret_cast<CT*>(sF + NP * PS);
    float* sT = reinterpret_cast<float*>(sC + NP * PS);

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int y0 = 1 + ty * TY, z0 = 1 + tz * TZ, x0 = 1 + tx * TX;
    const int y1 = min(y0 + TY, v.Y + 1), z1 = min(z0 + TZ, v.Z + 1), x1 = min(x0 + TX, v.X + 1);
    const int ey0 = y0 - H;
    const int ew = z0 - ML;         // z of window element 0; ew = 1 mod 4
    const int zl = ew + 2 * lane;   // z of the first voxel of this lane's pair
    const int lx1 = min(x1 + H, v.X + 2);

    if (KIND == TF_KIND_SOLVER) {
        if (threadIdx.x < 18) sT[threadIdx.x] = c_rd[threadIdx.x];
    } else if (KIND == TF_KIND_MULTIPHASE) {
        for (int i = threadIdx.x; i < ncls * ncls; i += NW * 64) sT[i] = hm[i];
    } else if (KIND == TF_KIND_ANISOTROPIC) {
        if (threadIdx.x < TF_ANISO_TABLE) sT[threadIdx.x] = hm[threadIdx.x];
    } else if (KIND == TF_KIND_ELECTRODE) {
        if (threadIdx.x < TF_ELECTRODE_TABLE) sT[threadIdx.x] = hm[threadIdx.x];
    }
    // spacing weights of TF_KIND_ANISOTROPIC (wave-uniform)
    const float ky = KIND == TF_KIND_ANISOTROPIC ? hm[TF_ANISO_TABLE - 2] : 0.0f;
    const float kz = KIND == TF_KIND_ANISOTROPIC ? hm[TF_ANISO_TABLE - 1] : 0.0f;

    // ---- global memory ----
    // Loads: one descriptor per array for the whole tile.  The field (and
    // the per-voxel diffusivity) descriptor starts kSlackBytes before the
    // volume and spans it plus the slack on both sides, so the window's
    // first quad (z = ew >= -3) is never at a negative offset; the lane
    // offset holds row and z, the wave-uniform offset the plane.  Rows
    // outside [0, Y+1] and planes past the tile's last plane use a lane
    // offset past any record count: loads read 0, no memory traffic.  The
    // range check is per dword, so quads running past the volume read 0
    // there.
    // Stores: one descriptor per row and plane, based at z = 0 of that row
    // with record count 4*z1 (0 outside the owned rows and planes), so a
    // pair straddling z1 = Z+1 stores only its first voxel; pairs before z0
    // use the out-of-range lane offset.
    // No branches around memory instructions, so the waitcnt pass keeps NP-1
    // planes in flight.
    constexpr int kOff = 0x7fffff00;  // lane offset past any record count
    const int plane_f = v.sX * 4;
    const int vol_f = (int)(v.sB * 4);
    const int half = lane >> 5, qd = lane & 31;
    const rsrc_t rF = make_rsrc(reinterpret_cast<const char*>(S) - kSlackBytes,
                                (uint32_t)vol_f + 2u * kSlackBytes);
    int vf[RPL], vc[RPL];
    rsrc_t rC;
    int plane_c;
    if constexpr (KIND == TF_KIND_FIELD) {
        plane_c = plane_f;
        rC = make_rsrc(reinterpret_cast<const char*>(Cg) - kSlackBytes,
                       (uint32_t)vol_f + 2u * kSlackBytes);
    } else {
        const int row_c = sweep_code_row(v.Z);
        plane_c = (v.Y + 2) * row_c;
        rC = make_rsrc(Cg, (uint32_t)((v.X + 2) * plane_c));
    }
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
        const int y = ey0 + 2 * (wave + i * NW) + half;
        const bool rok = y >= 0 && y <= v.Y + 1;
        vf[i] = rok ? kSlackBytes + (y * v.sY + ew + 4 * qd) * 4 : kOff;
        if constexpr (KIND == TF_KIND_FIELD)
            vc[i] = vf[i];
        else
            vc[i] = rok ? y * sweep_code_row(v.Z) + (ew + 3) + 4 * qd : kOff;
    }

    u32x4_t pf[NP][RPL];
    std::conditional_t<KIND == TF_KIND_FIELD, u32x4_t, uint32_t> pc[NP][RPL];
    auto issue_load = [&](int p, auto reg_c) {
        constexpr int rg = decltype(reg_c)::value;
        const bool pok = p < lx1;
        const int sf = pok ? p * plane_f : 0, sc = pok ? p * plane_c : 0;
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            pf[rg][i] = __builtin_amdgcn_raw_buffer_load_b128(rF, pok ? vf[i] : kOff, sf, 0);
            if constexpr (KIND == TF_KIND_FIELD)
                pc[rg][i] = __builtin_amdgcn_raw_buffer_load_b128(rC, pok ? vc[i] : kOff, sc, 0);
            else
                pc[rg][i] = __builtin_amdgcn_raw_buffer_load_b32(rC, pok ? vc[i] : kOff, sc, 0);
        }
    };
    // this lane's first quad slot (pairs 2qd, 2qd+1) in row 2*wave+half of slot 0
    float* const Fl = sF + (2 * wave + half) * kRowLen + 2 * qd;
    CT* const Cl = sC + (2 * wave + half) * kRowLen + 2 * qd;
    auto commit_load = [&](auto slot_c, auto reg_c) {
        constexpr int sl = decltype(slot_c)::value, rg = decltype(reg_c)::value;
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const int o = sl * PS + 2 * i * NW * kRowLen;
            const u32x4_t q = pf[rg][i];
            *reinterpret_cast<u32x2_t*>(Fl + o) = u32x2_t{q.x, q.z};               // first voxels
            *reinterpret_cast<u32x2_t*>(Fl + o + kRowPairs) = u32x2_t{q.y, q.w};   // second voxels
            if constexpr (KIND == TF_KIND_FIELD) {
                const u32x4_t k = pc[rg][i];
                *reinterpret_cast<u32x2_t*>(Cl + o) = u32x2_t{k.x, k.z};
                *reinterpret_cast<u32x2_t*>(Cl + o + kRowPairs) = u32x2_t{k.y, k.w};
            } else {
                const uint32_t c = pc[rg][i];  // codes of z = 4qd+ew .. +3, byte 0 first
                *reinterpret_cast<uint16_t*>(Cl + o) =
                    (uint16_t)((c & 0xffu) | ((c >> 8) & 0xff00u));
                *reinterpret_cast<uint16_t*>(Cl + o + kRowPairs) =
                    (uint16_t)(((c >> 8) & 0xffu) | ((c >> 16) & 0xff00u));
            }
        }
    };
    // stores: rows of the compute mapping, one pair per lane
    const float* Fw = sF + wave * kRowLen + lane;
    const int zs = zl >= z0 ? zl * 4 : kOff;
    bool own_row[RPW];
    int ys[RPW];
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
        const int y = ey0 + wave + j * NW;
        own_row[j] = y >= y0 && y < y1;
        ys[j] = own_row[j] ? y : 0;
    }
    auto store_plane = [&](int q, auto slot_c) {
        constexpr int sl = decltype(slot_c)::value;
        const bool qok = q >= x0 && q < x1;
        const int qs = qok ? q : 0;
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
            const int o = sl * PS + j * NW * kRowLen;
            const rsrc_t r = make_rsrc(D + qs * v.sX + ys[j] * v.sY,
                                       (qok && own_row[j]) ? (uint32_t)(4 * z1) : 0u);
            const u32x2_t val{__builtin_bit_cast(uint32_t, Fw[o]),
                              __builtin_bit_cast(uint32_t, Fw[o + kRowPairs])};
            __builtin_amdgcn_raw_buffer_store_b64(val, r, zs, 0, 0);
        }
    };

    // ---- per-tile bounds of each half-sweep ----
    // planes [plo[h], plo[h] + pn[h]) and, as bit h*RPW+j of rowbits, the
    // active rows.  In z every lane updates its active element: outside the
    // half-sweep's region the result is never read by a later half-sweep of
    // this pass nor stored, and on the z ghosts (value 0, code 0 / class 0 /
    // k 0) the update returns the ghost's value.
    int plo[H], pn[H];
    int rowbits = 0;
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int grow = H - 1 - h;
        plo[h] = max(x0 - grow, 1);
        pn[h] = min(x1 + grow, v.X + 1) - plo[h];
        const int ry0 = max(y0 - grow, 1), ry1 = min(y1 + grow, v.Y + 1);
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
            const int y = ey0 + wave + j * NW;
            rowbits |= (y >= ry0 && y < ry1) ? 1 << (h * RPW + j) : 0;
        }
    }

    // The colour of (p, y, z) is (p+y+z+1)&1 and half-sweep h of step s
    // updates colour colour0^(h&1) on plane s-h, so the active element of a
    // lane's pair, act = (s + y + ew + 1 + colour0) & 1, is the same for all
    // half-sweeps of a step and all rows of a wave; it alternates with s.
    const int S0 = max(x0 - (H - 1), 1);
    const int S1 = (x1 - 1) + (H - 1);
    int A = (S0 + ey0 + wave + ew + 1 + colour0) & 1;  // act of the even steps of an iteration

    // base addresses of the own element and its z neighbours for either
    // parity of the step, biased one row down (the row below is at offset 0)
    const float* Fo[2];
    const float* Fp[2];
    const float* Fm[2];
    const CT* Co[2];
    const CT* Cp[2];
    const CT* Cm[2];
    auto set_bases = [&](int a0) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int a = a0 ^ q;
            const int r0 = wave * kRowLen - kRowLen + lane;
            const int io = r0 + (a ? kRowPairs : 0);
            const int ip = r0 + (a ? 1 : kRowPairs);      // z+1: next pair's first / own second
            const int im = r0 + (a ? 0 : kRowPairs - 1);  // z-1: own first / previous pair's second
            Fo[q] = sF + io; Fp[q] = sF + ip; Fm[q] = sF + im;
            Co[q] = sC + io; Cp[q] = sC + ip; Cm[q] = sC + im;
        }
    };

    // ---- one half-sweep of one plane ----
    auto stage = [&](auto h_c, auto q_c, auto sp_c, int p) {
        constexpr int h = decltype(h_c)::value, q = decltype(q_c)::value;
        constexpr int sp = decltype(sp_c)::value;
        constexpr int spm = (sp + NP - 1) % NP, spp = (sp + 1) % NP;
        if ((unsigned)(p - plo[h]) >= (unsigned)pn[h]) return;
        float* const Fq = const_cast<float*>(Fo[q]);
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
            if (!((rowbits >> (h * RPW + j)) & 1)) continue;
            const int o = sp * PS + j * NW * kRowLen + kRowLen;
            const int om = spm * PS + j * NW * kRowLen + kRowLen;
            const int op = spp * PS + j * NW * kRowLen + kRowLen;
            Nbr n;
            n.c = Fq[o];
            n.xm = Fq[om];
            n.xp = Fq[op];
            n.ym = Fq[o - kRowLen];
            n.yp = Fq[o + kRowLen];
            n.zp = Fp[q][o];
            n.zm = Fm[q][o];
            float out;
            if constexpr (KIND == TF_KIND_SOLVER) {
                const int code = Co[q][o];
                const float2 rdv = reinterpret_cast<const float2*>(sT)[code];
                float s = n.xp + n.xm;
                s = s + n.yp;
                s = s + n.ym;
                s = s + n.zp;
                s = s + n.zm;
                float qt = s * rdv.x;
                const float e = __builtin_fmaf(-qt, rdv.y, s);
                qt = __builtin_fmaf(e, rdv.x, qt);
                // below 2^-100 the residual correction can be off (d = 6):
                // IEEE division there (s = +-0 needs none: any zero quotient
                // gives the same update)
                if (__builtin_expect(s != 0.0f && __builtin_fabsf(s) < 0x1p-100f, 0))
                    qt = s / (code ? (float)code : __builtin_inff());
                s = qt - n.c;
                s = s * w;
                out = n.c + s;
            } else if constexpr (KIND == TF_KIND_ELECTRODE) {
                // SORSolver.sum_weighted_neighbours / factor with the reaction
                // term in the factor (taufactor/taufactor.py:47-56, 95-103, 175-181)
                float s = n.xp + n.xm;
                s = s + n.yp;
                s = s + n.ym;
                s = s + n.zp;
                s = s + n.zm;
                s = s / sT[Co[q][o]];
                s = s - n.c;
                s = s * w;
                out = n.c + s;
            } else if constexpr (KIND == TF_KIND_ANISOTROPIC) {
                // ((x+ + x-) + Ky*(y+ + y-)) + Kz*(z+ + z-), / nn
                // (AnisotropicSolver.sum_weighted_neighbours, taufactor/taufactor.py:473-478)
                float s = n.xp + n.xm;
                float t = n.yp + n.ym;
                t = ky * t;
                s = s + t;
                t = n.zp + n.zm;
                t = kz * t;
                s = s + t;
                s = s / sT[Co[q][o]];
                s = s - n.c;
                s = s * w;
                out = n.c + s;
            } else if constexpr (KIND == TF_KIND_MULTIPHASE) {
                const float* row = sT + (int)Co[q][o] * ncls;
                out = update_multiphase(n, row[Co[q][op]], row[Co[q][om]], row[Co[q][o + kRowLen]],
                                        row[Co[q][o - kRowLen]], row[Cp[q][o]], row[Cm[q][o]],
                                        p == 1, p == v.X, w);
            } else {
                const float kc = Co[q][o];
                out = update_multiphase(n, hmean(kc, Co[q][op]), hmean(kc, Co[q][om]),
                                        hmean(kc, Co[q][o + kRowLen]), hmean(kc, Co[q][o - kRowLen]),
                                        hmean(kc, Cp[q][o]), hmean(kc, Cm[q][o]), p == 1, p == v.X,
                                        w);
            }
            Fq[o] = out;
        }
    };

    // planes S0-1 .. S0+1 go straight to LDS (slots 0..2); S0+2 .. S0+NP are put in flight
    issue_load(S0 - 1, ic<0>{});
    commit_load(ic<0>{}, ic<0>{});
    issue_load(S0, ic<1>{});
    commit_load(ic<1>{}, ic<1>{});
    issue_load(S0 + 1, ic<2>{});
    commit_load(ic<2>{}, ic<2>{});
    static_for<2, NP + 1>([&](auto d_c) {
        constexpr int d = decltype(d_c)::value;
        issue_load(S0 + d, ic<(d + 1) % NP>{});
    });
    __syncthreads();

    // step s = s0 + u: put plane s+1+NP in flight, run half-sweep h on plane
    // s-h, write plane s-H+1 out, move plane s+2 from registers to LDS.
    // Steps past S1 (to complete the last iteration) do no work but keep the
    // memory instructions unconditional.
    for (int s0 = S0; s0 <= S1; s0 += NP) {
        set_bases(A);
        static_for<0, NP>([&](auto u_c) {
            constexpr int u = decltype(u_c)::value;
            const int s = s0 + u;
            issue_load(s + 1 + NP, ic<(u + 2) % NP>{});
            static_for<0, H>([&](auto h_c) {
                constexpr int h = decltype(h_c)::value;
                stage(h_c, ic<(u & 1)>{}, ic<((u + 1 - h) % NP + NP) % NP>{}, s - h);
                if constexpr (h + 1 < H) __syncthreads();
            });
            store_plane(s - (H - 1), ic<((u + 2 - H) % NP + NP) % NP>{});
            commit_load(ic<(u + 3) % NP>{}, ic<(u + 3) % NP>{});
            __syncthreads();
        });
        A ^= NP & 1;
    }
}

// ---------------------------------------------------------------------------
// register-blocked sweep for the kinds whose update reads only the voxel's
// own code (Solver, anisotropic, electrode)
// ---------------------------------------------------------------------------
// Same compute mapping as sweep_tile (a lane holds the z-pair zl, zl+1 of
// each of its rows; the red/black pattern is wave-uniform), but a wave owns
// RPW consecutive rows and keeps every plane of the ring in registers:
// x neighbours and the y neighbours inside the wave are registers, the z
// neighbour across lanes comes by a DPP wave shift, and only the first and
// last row of each wave go through LDS.  Half-sweep h of step s reads the
// edge rows that the neighbouring waves published at step s-1 (the values
// of plane s-h after half-sweep h-1, or as loaded for h = 0), double
// buffered by step parity, so one barrier per step suffices.
__host__ __device__ constexpr bool own_code_kind(int kind) {
    return kind == TF_KIND_SOLVER || kind == TF_KIND_ANISOTROPIC || kind == TF_KIND_ELECTRODE;
}

template <int H, int NW, int RPW>
struct RegShape {
    static constexpr int EY = NW * RPW;
    static constexpr int TY = EY - 2 * H;
    static constexpr int ML = H + (H & 1);                    // z0 - ew: even, so ew is odd
    static constexpr int TZ = ((kRowLen - ML - H) / 2) * 2;   // owned voxels of z (even)
    static constexpr int TZN = ((kRowPairs - ML - H) / 2) * 2;  // the same for a narrow tile
    // planes in registers (even): H + 2 in use and D in flight; D = 1 at
    // H = 5 keeps a 16-wave workgroup within 128 VGPRs per lane
    static constexpr int NR = H == 5 ? 8 : ((H + 5) / 2) * 2;
    static constexpr int D = NR - H - 2;
    static constexpr int TBL = 640;                           // LDS bytes for the table
    static constexpr int EB = (NW + 2) * 2 * kRowPairs;       // floats per edge buffer
    static constexpr size_t lds_bytes() { return TBL + (size_t)H * 2 * EB * sizeof(float); }
};

// (the component goes through a scalar first: hipcc 7.2 compiles
// __builtin_bit_cast of an ext_vector component as a cast of component 0)
__device__ __forceinline__ float lo_f(u32x2_t v) {
    const uint32_t t = v.x;
    return __builtin_bit_cast(float, t);
}
__device__ __forceinline__ float hi_f(u32x2_t v) {
    const uint32_t t = v.y;
    return __builtin_bit_cast(float, t);
}
// lane l receives lane l-1's value (lane 0: 0) / lane l+1's value (lane 63: 0)
__device__ __forceinline__ float from_prev_lane(uint32_t v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_next_lane(uint32_t v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, false));
}

// Tile geometry requirements (checked by the launch): TX and TY even, so that
// with the unclipped stream start S0 = x0 - (H-1) the active element of a
// lane depends only on the launch's colour and the step and row parities.
//
// A narrow tile (nar) covers the last strip of z when it is at most TZN wide:
// each half-wave holds a 64-voxel window of z, half 0 for the rows of y tile
// 2*ty and half 1 for those of y tile 2*ty + 1 (TY rows further, an even
// count, so the active element stays wave-uniform).  The DPP shifts mix the
// halves only at lanes 31 and 32, which are halo lanes of both windows, and
// the edge rows go through LDS per lane as before; only the addresses and
// the stores (one per half) differ.
template <int H, int KIND, int NW, int RPW>
__device__ __forceinline__ void rsweep_tile(const float* __restrict__ S, float* __restrict__ D,
                                            const uint8_t* __restrict__ Cg,
                                            const float* __restrict__ hm, const Vol v, float w,
                                            int colour0, int TX, int z0, int ty, int tx, bool nar) {
    using Sh = RegShape<H, NW, RPW>;
    constexpr int TY = Sh::TY, ML = Sh::ML, NR = Sh::NR, DP = Sh::D, EB = Sh::EB;
    static_assert(TY >= 2 && TY % 2 == 0, "tile rows must be even and positive");
    static_assert(RPW % 2 == 0, "rows per wave must be even");
    extern __shared__ __align__(16) unsigned char smem[];
    float* const sT = reinterpret_cast<float*>(smem);
    float* const sE = reinterpret_cast<float*>(smem + Sh::TBL);

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // y tile of each half (for a wide tile half 1 owns nothing)
    const int y0h[2] = {1 + (nar ? 2 * ty : ty) * TY, 1 + (nar ? 2 * ty + 1 : ty) * TY};
    const int x0 = 1 + tx * TX;
    const int x1 = min(x0 + TX, v.X + 1);
    const int z1 = min(z0 + (nar ? Sh::TZN : Sh::TZ), v.Z + 1);
    const int half = nar ? lane >> 5 : 0;                 // this lane's half
    const int y0 = half ? y0h[1] : y0h[0];                  // this lane's y tile
    const int ey0 = y0 - H;
    const int ew = z0 - ML;                                 // odd
    const int zl = ew + 2 * (nar ? (lane & 31) : lane);     // z of the lane's first voxel
    const int lx1 = min(x1 + H, v.X + 2);

    if (KIND == TF_KIND_SOLVER) {
        if (threadIdx.x < 18) sT[threadIdx.x] = c_rd[threadIdx.x];
    } else if (KIND == TF_KIND_ANISOTROPIC) {
        if (threadIdx.x < TF_ANISO_TABLE) sT[threadIdx.x] = hm[threadIdx.x];
    } else {
        if (threadIdx.x < TF_ELECTRODE_TABLE) sT[threadIdx.x] = hm[threadIdx.x];
    }
    const float ky = KIND == TF_KIND_ANISOTROPIC ? hm[TF_ANISO_TABLE - 2] : 0.0f;
    const float kz = KIND == TF_KIND_ANISOTROPIC ? hm[TF_ANISO_TABLE - 1] : 0.0f;

    // ---- global memory (see sweep_tile for the descriptor conventions) ----
    // loads: one 8-byte field load (the lane's pair) and one 2-byte code
    // load per row and plane, straight into the register ring
    constexpr int kOff = 0x7fffff00;
    const int plane_f = v.sX * 4;
    const rsrc_t rF = make_rsrc(reinterpret_cast<const char*>(S) - kSlackBytes,
                                (uint32_t)(v.sB * 4) + 2u * kSlackBytes);
    const int row_c = sweep_code_row(v.Z);
    const int plane_c = (v.Y + 2) * row_c;
    const rsrc_t rC = make_rsrc(Cg, (uint32_t)((v.X + 2) * plane_c));
    int vf[RPW], vc[RPW];
    bool own_row[2][RPW];   // wave-uniform, per half
    int ys[2][RPW];
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
        const int y = ey0 + wave * RPW + j;
        const bool rok = y >= 0 && y <= v.Y + 1;
        vf[j] = rok ? kSlackBytes + (y * v.sY + zl) * 4 : kOff;
        vc[j] = rok ? y * row_c + zl + 3 : kOff;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int yk = y0h[k] - H + wave * RPW + j;
            own_row[k][j] = (k == 0 || nar) && yk >= y0h[k] && yk < min(y0h[k] + TY, v.Y + 1);
            ys[k][j] = own_row[k][j] ? yk : 0;
        }
    }
    u32x2_t F[NR][RPW];
    uint32_t K[NR][RPW];
    // slots the prologue does not load hold planes before the stream start
    // in the first steps: defined values, read by no active half-sweep
#pragma unroll
    for (int i = 0; i < NR; ++i)
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
            F[i][j] = u32x2_t{0u, 0u};
            K[i][j] = 0u;
        }
    auto issue_load = [&](int p, auto slot_c) {
        constexpr int sl = decltype(slot_c)::value;
        const bool pok = p >= 0 && p < lx1;
        const int sf = pok ? p * plane_f : 0, sc = pok ? p * plane_c : 0;
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
            F[sl][j] = __builtin_amdgcn_raw_buffer_load_b64(rF, pok ? vf[j] : kOff, sf, 0);
            K[sl][j] = __builtin_amdgcn_raw_buffer_load_b16(rC, pok ? vc[j] : kOff, sc, 0);
        }
    };
    // store k writes the lanes of half k: one descriptor per row, plane and
    // half, based at z = 0 of that row with record count 4*z1
    const int zs0 = (zl >= z0 && half == 0) ? zl * 4 : kOff;
    const int zs1 = (zl >= z0 && half == 1) ? zl * 4 : kOff;
    auto store_plane = [&](int q, auto slot_c) {
        constexpr int sl = decltype(slot_c)::value;
        const bool qok = q >= x0 && q < x1;
        const int qs = qok ? q : 0;
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const rsrc_t r = make_rsrc(D + qs * v.sX + ys[k][j] * v.sY,
                                           (qok && own_row[k][j]) ? (uint32_t)(4 * z1) : 0u);
                __builtin_amdgcn_raw_buffer_store_b64(F[sl][j], r, k ? zs1 : zs0, 0, 0);
            }
        }
    };

    // ---- per-tile bounds of each half-sweep (planes only: rows and z are
    // updated unconditionally; outside a half-sweep's region the result is
    // never read by a later half-sweep of this pass nor stored, and ghosts
    // (value 0 with code 0, whose table entry gives a zero quotient) keep
    // their value) ----
    int plo[H], pn[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int grow = H - 1 - h;
        plo[h] = max(x0 - grow, 1);
        pn[h] = min(x1 + grow, v.X + 1) - plo[h];
    }
    // unclipped stream start: with TX and TY even, S0 + ey0 is odd, so the
    // active element of row j at step u of an unrolled block,
    // (s + y + colour0) & 1, is A ^ ((u + j) & 1) with A = colour0 ^ 1 for
    // every tile of the launch; the parity of S0 is that of H
    const int S0 = x0 - (H - 1);
    const int S1 = (x1 - 1) + (H - 1);
    const int A = (colour0 & 1) ^ 1;

    // edge buffers: [h][parity][NW + 2][top, bottom][64 lanes]; wave w
    // publishes in slot w + 1, so the missing neighbours of the first and
    // last wave read a slot nobody writes (their rows are halo rows)
    float* const eb = sE + (wave + 1) * 2 * kRowPairs + lane;
    // parity of step s is that of S0 + u, S0 having the parity of H
    auto ebuf = [&](int h, int par) { return eb + (h * 2 + par) * EB; };

    auto body = [&](auto a_c) {
        constexpr int AA = decltype(a_c)::value;
        // active element (0: z = zl, 1: z = zl + 1) of row j at step u
        auto act = [](int u, int j) { return (AA ^ (u + j)) & 1; };
        auto par = [](int u) { return (H + u) & 1; };

        auto publish = [&](int h, int u, auto slot_c) {
            constexpr int sl = decltype(slot_c)::value;
            float* e = ebuf(h, par(u));
            const int a0 = act(u, 0), a1 = act(u, RPW - 1);
            e[0] = a0 ? hi_f(F[sl][0]) : lo_f(F[sl][0]);
            e[kRowPairs] = a1 ? hi_f(F[sl][RPW - 1]) : lo_f(F[sl][RPW - 1]);
        };

        auto stage = [&](auto h_c, auto u_c, int p) {
            constexpr int h = decltype(h_c)::value, u = decltype(u_c)::value;
            constexpr int sl = ((u - h + 1) % NR + NR) % NR;
            constexpr int slm = (sl + NR - 1) % NR, slp = (sl + 1) % NR;
            if ((unsigned)(p - plo[h]) >= (unsigned)pn[h]) return;
            const float* e = ebuf(h, par(u + 1));
            const float ym0 = e[-kRowPairs];        // bottom row of the wave above
            const float ypL = e[2 * kRowPairs];     // top row of the wave below
#pragma unroll
            for (int j = 0; j < RPW; ++j) {
                const int a = act(u, j);
                const u32x2_t m = F[sl][j];
                const float c = a ? hi_f(m) : lo_f(m);
                const float xm = a ? hi_f(F[slm][j]) : lo_f(F[slm][j]);
                const float xp = a ? hi_f(F[slp][j]) : lo_f(F[slp][j]);
                const float ym = j > 0 ? (a ? hi_f(F[sl][j - 1]) : lo_f(F[sl][j - 1])) : ym0;
                const float yp = j < RPW - 1 ? (a ? hi_f(F[sl][j + 1]) : lo_f(F[sl][j + 1])) : ypL;
                const float zp = a ? from_next_lane(m.x) : hi_f(m);
                const float zm = a ? lo_f(m) : from_prev_lane(m.y);
                const uint32_t code = a ? (K[sl][j] >> 8) : (K[sl][j] & 0xffu);
                float out;
                if constexpr (KIND == TF_KIND_SOLVER) {
                    const float2 rdv = reinterpret_cast<const float2*>(sT)[code];
                    float s = xp + xm;
                    s = s + yp;
                    s = s + ym;
                    s = s + zp;
                    s = s + zm;
                    float qt = s * rdv.x;
                    const float e2 = __builtin_fmaf(-qt, rdv.y, s);
                    qt = __builtin_fmaf(e2, rdv.x, qt);
                    if (__builtin_expect(s != 0.0f && __builtin_fabsf(s) < 0x1p-100f, 0))
                        qt = s / (code ? (float)code : __builtin_inff());
                    s = qt - c;
                    s = s * w;
                    out = c + s;
                } else if constexpr (KIND == TF_KIND_ELECTRODE) {
                    float s = xp + xm;
                    s = s + yp;
                    s = s + ym;
                    s = s + zp;
                    s = s + zm;
                    s = s / sT[code];
                    s = s - c;
                    s = s * w;
                    out = c + s;
                } else {
                    float s = xp + xm;
                    float t = yp + ym;
                    t = ky * t;
                    s = s + t;
                    t = zp + zm;
                    t = kz * t;
                    s = s + t;
                    s = s / sT[code];
                    s = s - c;
                    s = s * w;
                    out = c + s;
                }
                if (a)
                    F[sl][j].y = __builtin_bit_cast(uint32_t, out);
                else
                    F[sl][j].x = __builtin_bit_cast(uint32_t, out);
            }
        };

        // planes S0-1 .. S0+1+DP in flight; plane S0's edges published as
        // "step S0-1" for the first half-sweep
        static_for<0, DP + 3>([&](auto i_c) {
            constexpr int i = decltype(i_c)::value;
            issue_load(S0 - 1 + i, ic<i>{});
        });
        publish(0, -1 + NR, ic<1>{});  // u = -1 has the parities of u = NR - 1
        __syncthreads();

        for (int s0 = S0; s0 <= S1; s0 += NR) {
            static_for<0, NR>([&](auto u_c) {
                constexpr int u = decltype(u_c)::value;
                const int s = s0 + u;
                static_for<0, H>([&](auto h_c) {
                    constexpr int h = decltype(h_c)::value;
                    stage(h_c, u_c, s - h);
                    if constexpr (h + 1 < H) publish(h + 1, u, ic<((u - h + 1) % NR + NR) % NR>{});
                });
                publish(0, u, ic<(u + 2) % NR>{});
                store_plane(s - (H - 1), ic<((u - H + 2) % NR + NR) % NR>{});
                issue_load(s + 2 + DP, ic<((u - H + 1) % NR + NR) % NR>{});
                __syncthreads();
            });
        }
    };
    if (A)
        body(ic<1>{});
    else
        body(ic<0>{});
}

template <int H, int KIND, int NW, int RPW>
__global__ __launch_bounds__(NW * 64) void k_sweep(SweepArgs a) {
    // XCD-aware tile numbering (grid is a multiple of 8)
    const int per = gridDim.x >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= a.total) return;
    int t = logical;
    const bool nar = logical >= a.nWide;
    int tz = 0, ty, tx, b;
    if (!nar) {
        tz = t % a.nTz; t /= a.nTz;
        ty = t % a.nTy; t /= a.nTy;
    } else {
        t -= a.nWide;
        ty = t % a.nTyN; t /= a.nTyN;
    }
    tx = t % a.nTx;
    b = t / a.nTx;
    const size_t off = (size_t)b * a.v.sB;
    const size_t coff = KIND == TF_KIND_FIELD ? off
                                              : (size_t)b * sweep_code_volume(a.v.X, a.v.Y, a.v.Z);
    const float* hm = KIND == TF_KIND_ELECTRODE ? a.hm + (size_t)b * TF_ELECTRODE_TABLE : a.hm;
    if constexpr (own_code_kind(KIND))
        rsweep_tile<H, KIND, NW, RPW>(a.src + off, a.dst + off,
                                      static_cast<const uint8_t*>(a.codes) + coff, hm, a.v, a.w,
                                      a.colour0, a.TX,
                                      nar ? a.z0N : 1 + tz * RegShape<H, NW, RPW>::TZ, ty, tx,
                                      nar);
    else
        sweep_tile<H, KIND, NW, RPW>(a.src + off, a.dst + off,
                                     static_cast<const code_t<KIND>*>(a.codes) + coff, hm, a.ncls,
                                     a.v, a.w, a.colour0, a.TX, tz, ty, tx);
}

// Tile geometry of the kernel k_sweep<H, KIND, NW, RPW> runs: owned rows
// and z voxels per tile, steps per unrolled block, LDS bytes.
template <int H, int KIND, int NW, int RPW>
struct SweepGeom {
    static constexpr bool reg = own_code_kind(KIND);
    static constexpr int TY = reg ? RegShape<H, NW, RPW>::TY : SweepShape<H, NW, RPW>::TY;
    static constexpr int TZ = reg ? RegShape<H, NW, RPW>::TZ : SweepShape<H, NW, RPW>::TZ;
    static constexpr int NP = reg ? RegShape<H, NW, RPW>::NR : SweepShape<H, NW, RPW>::NP;
};

template <int H, int NW, int RPW>
static size_t sweep_lds_bytes(int kind, int ncls) {
    if (own_code_kind(kind)) return RegShape<H, NW, RPW>::lds_bytes();
    using Sh = SweepShape<H, NW, RPW>;
    const size_t code_bytes = kind == TF_KIND_FIELD ? sizeof(float) : 1;
    size_t bytes = kLdsPad + (size_t)Sh::NP * Sh::PS * (sizeof(float) + code_bytes);
    bytes += kind == TF_KIND_MULTIPHASE    ? (size_t)ncls * ncls * sizeof(float)
             : kind == TF_KIND_ANISOTROPIC ? TF_ANISO_TABLE * sizeof(float)
             : kind == TF_KIND_ELECTRODE   ? TF_ELECTRODE_TABLE * sizeof(float)
                                           : 18 * sizeof(float);
    return bytes;
}

// Workgroups of one sweep kernel resident on the current device at once
// (CUs x workgroups per CU); 0 if the runtime cannot say.  `cache` holds the
// answer + 1 per device (0: not asked yet).
constexpr int kMaxDevices = 64;
static int resident_workgroups(const void* fn, int threads, size_t lds, int* cache) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return 0;
    }
    if (cache[dev] > 0) return cache[dev] - 1;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds) != hipSuccess) {
        (void)hipGetLastError();
        cus = per_cu = 0;
    }
    cache[dev] = std::max(0, cus) * std::max(0, per_cu) + 1;
    return cache[dev] - 1;
}

// x-tile length that minimises the modelled duration of one launch: the
// tiles are spread evenly over the 8 XCDs (k_sweep's numbering) and run in
// rounds of the resident workgroups; a tile costs its streamed planes,
// TX + 2(H-1) rounded up to the unrolled step count NP, plus a fixed start.
// A grid with fewer tiles than resident slots is one round, so for small
// volumes the shortest tile that still fits one round wins.
static int balanced_tx(int X, long long yz_tiles, int H, int NP, int slots, bool even) {
    auto fit = [even](int tx) { return even ? (tx + 1) & ~1 : tx; };
    if (slots <= 0) return fit(std::min(X, 64));
    const long long per_xcd = std::max(1, slots / 8);
    long long best_cost = -1;
    int best = fit(X);
    for (int n = 1; n <= X; ++n) {
        const int tx = fit((X + n - 1) / n);
        const int nx = (X + tx - 1) / tx;
        if (nx != n) continue;  // same tiling as a smaller n
        const long long tiles = yz_tiles * nx;
        const long long rounds = ((tiles + 7) / 8 + per_xcd - 1) / per_xcd;
        const long long steps = (tx + 2 * (H - 1) + NP - 1) / NP * NP + 4;
        const long long cost = rounds * steps;
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = tx;
        }
        if (tx <= 8) break;
    }
    return best;
}

template <int H, int KIND, int NW, int RPW>
static int launch_sweep(SweepArgs& a, int B, hipStream_t st) {
    using Sh = SweepGeom<H, KIND, NW, RPW>;
    if constexpr (Sh::TY < (Sh::reg ? 2 : 1)) {
        return fail(-1, "workgroup shape %dx%d has no interior rows for n_half=%d", NW, RPW, H);
    } else {
        const size_t lds = sweep_lds_bytes<H, NW, RPW>(KIND, a.ncls);
        if (lds > 160 * 1024)
            return fail(-1, "sweep shape %dx%d with n_half=%d needs %zu B of LDS", NW, RPW, H, lds);
        static bool attr_set = false;
        if (!attr_set && lds > 64 * 1024) {
            // opt in to more than 64 KiB of dynamic LDS; if the runtime
            // refuses, the launch below reports the error
            if (hipFuncSetAttribute((const void*)k_sweep<H, KIND, NW, RPW>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
                (void)hipGetLastError();
            attr_set = true;
        }
        a.nTy = (a.v.Y + Sh::TY - 1) / Sh::TY;
        a.nTz = (a.v.Z + Sh::TZ - 1) / Sh::TZ;
        a.nTyN = 0;
        if constexpr (Sh::reg) {
            // wide tiles while a full one fits, then a narrow strip if the
            // rest fits one (RegShape::TZN), else one more wide tile
            const int tzn = RegShape<H, NW, RPW>::TZN;
            const int wide = a.v.Z / Sh::TZ, rest = a.v.Z - wide * Sh::TZ;
            a.nTz = wide + (rest > tzn ? 1 : 0);
            if (rest > 0 && rest <= tzn) a.nTyN = (a.nTy + 1) / 2;
            a.z0N = 1 + wide * Sh::TZ;
        }
        const long long yz = (long long)a.nTz * a.nTy + a.nTyN;
        if (a.TX <= 0) {
            static int slots_cache[kMaxDevices] = {};
            const int slots = resident_workgroups((const void*)k_sweep<H, KIND, NW, RPW>, NW * 64,
                                                  lds, slots_cache);
            a.TX = balanced_tx(a.v.X, yz * B, H, Sh::NP, slots, Sh::reg);
        }
        // the register-blocked kernel needs an even x tile (rsweep_tile)
        if (Sh::reg) a.TX = (a.TX + 1) & ~1;
        a.nTx = (a.v.X + a.TX - 1) / a.TX;
        const long long wide_tiles = (long long)a.nTz * a.nTy * a.nTx * B;
        const long long total = yz * a.nTx * B;
        if (total * 8 >= (1LL << 31)) return fail(-1, "too many tiles (%lld)", total);
        a.nWide = (int)wide_tiles;
        a.total = (int)total;
        if (a.resolve_only) return 0;
        const int grid = (int)(((total + 7) / 8) * 8);
        hipLaunchKernelGGL((k_sweep<H, KIND, NW, RPW>), dim3(grid), dim3(NW * 64), lds, st, a);
        TF_LAUNCH_CHECK("k_sweep");
        return 0;
    }
}

// workgroup shapes (waves x rows per wave) that are compiled
// compiled workgroup shapes: all five for TF_KIND_SOLVER, 8x2 / 4x4 / 16x2
// for the other kinds
template <int H, int KIND>
static int dispatch_shape(SweepArgs& a, int B, int nw, int rpw, hipStream_t st) {
    if (nw == 8 && rpw == 2) return launch_sweep<H, KIND, 8, 2>(a, B, st);
    if (nw == 16 && rpw == 2) return launch_sweep<H, KIND, 16, 2>(a, B, st);
    if constexpr (H <= 4) {
        if (nw == 4 && rpw == 4) return launch_sweep<H, KIND, 4, 4>(a, B, st);
    }
    if constexpr (KIND == TF_KIND_SOLVER) {
        if (nw == 12 && rpw == 2) return launch_sweep<H, KIND, 12, 2>(a, B, st);
        if constexpr (H <= 4) {
            if (nw == 8 && rpw == 4) return launch_sweep<H, KIND, 8, 4>(a, B, st);
        }
    }
    return fail(-1, "workgroup shape %d waves x %d rows is not compiled for kind %d and n_half %d",
                nw, rpw, KIND, H);
}

template <int KIND>
static int dispatch_sweep(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st) {
    switch (H) {
        case 1: return dispatch_shape<1, KIND>(a, B, nw, rpw, st);
        case 2: return dispatch_shape<2, KIND>(a, B, nw, rpw, st);
        case 3: return dispatch_shape<3, KIND>(a, B, nw, rpw, st);
        case 4: return dispatch_shape<4, KIND>(a, B, nw, rpw, st);
        case 5: return dispatch_shape<5, KIND>(a, B, nw, rpw, st);
        case 6: return dispatch_shape<6, KIND>(a, B, nw, rpw, st);
        default: return fail(-1, "n_half=%d out of range 1..%d", H, TF_MAX_FUSED_HALF_SWEEPS);
    }
}

// owned voxels of z per tile of the register-blocked kernel (RegShape::TZ)
static int reg_tile_tz(int H) { return ((kRowLen - (H + (H & 1)) - H) / 2) * 2; }

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

// Checks shared by tf_sweep and tf_sweep_launch_config; fills the shape.
static int sweep_prepare(const float* hm_lut, int n_classes, int kind, int B, int X, int Y, int Z,
                         int n_half, const tf_tile* tile, tf_tile* t) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (kind != TF_KIND_SOLVER && kind != TF_KIND_MULTIPHASE && kind != TF_KIND_FIELD &&
        kind != TF_KIND_ANISOTROPIC && kind != TF_KIND_ELECTRODE)
        return fail(-1, "unknown kind %d", kind);
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
    if ((long long)(X + 2) * (Y + 2) * (Z + 2) * 4 >= (1LL << 31) - 64 ||
        sweep_code_volume(X, Y, Z) >= (1LL << 31))
        return fail(-1, "padded volume of %lld bytes: the sweep addresses at most 2 GiB per volume",
                    (long long)(X + 2) * (Y + 2) * (Z + 2) * 4);
    return 0;
}

static int sweep_dispatch(int kind, int n_half, SweepArgs& a, int B, const tf_tile& t,
                          hipStream_t st) {
    if (kind == TF_KIND_SOLVER)
        return dispatch_sweep<TF_KIND_SOLVER>(n_half, a, B, t.waves, t.rows_per_wave, st);
    if (kind == TF_KIND_MULTIPHASE)
        return dispatch_sweep<TF_KIND_MULTIPHASE>(n_half, a, B, t.waves, t.rows_per_wave, st);
    if (kind == TF_KIND_ANISOTROPIC)
        return dispatch_sweep<TF_KIND_ANISOTROPIC>(n_half, a, B, t.waves, t.rows_per_wave, st);
    if (kind == TF_KIND_ELECTRODE)
        return dispatch_sweep<TF_KIND_ELECTRODE>(n_half, a, B, t.waves, t.rows_per_wave, st);
    return dispatch_sweep<TF_KIND_FIELD>(n_half, a, B, t.waves, t.rows_per_wave, st);
}

extern "C" int tf_sweep(const float* src, float* dst, const void* codes, const float* hm_lut,
                        int n_classes, int kind, int B, int X, int Y, int Z, float omega,
                        int first_colour, int n_half, const tf_tile* tile, void* stream) {
    tf_tile t;
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
    return sweep_dispatch(kind, n_half, a, B, t, (hipStream_t)stream);
}

extern "C" int tf_sweep_launch_config(int kind, int n_classes, int B, int X, int Y, int Z,
                                      int n_half, const tf_tile* tile, tf_tile* out,
                                      int* n_tiles) {
    if (!out) return fail(-1, "null pointer argument");
    // the tables are not read here: any non-null pointer passes the checks
    static const float dummy = 0.0f;
    tf_tile t;
    if (int rc = sweep_prepare(&dummy, n_classes, kind, B, X, Y, Z, n_half, tile, &t)) return rc;
    SweepArgs a = {};
    a.v = make_vol(X, Y, Z);
    a.ncls = n_classes;
    a.TX = t.tx;
    a.resolve_only = 1;
    if (int rc = sweep_dispatch(kind, n_half, a, B, t, nullptr)) return rc;
    out->waves = t.waves;
    out->rows_per_wave = t.rows_per_wave;
    out->tx = a.TX;
    if (n_tiles) *n_tiles = a.total;
    return 0;
}

// ---------------------------------------------------------------------------
// per-layer float64 sums for the convergence check
// ---------------------------------------------------------------------------
// One block per (x, b).  Each thread accumulates a fixed strided subset of
// the layer in float64; the block then reduces in a fixed tree, so the result
// does not depend on scheduling.
constexpr int kSumThreads = 1024;

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sums of layer x of one padded volume V; returns them in thread 0.
template <int KIND>
__device__ __forceinline__ void layer_sum_block(const float* __restrict__ V,
                                                const code_t<KIND>* __restrict__ Cv,
                                                const float* __restrict__ hm, int ncls,
                                                const Vol v, int x, double& sum_conc,
                                                double& sum_flux) {
    const bool has_face = KIND == TF_KIND_ELECTRODE || x < v.X - 1;
    double accC = 0.0, accF = 0.0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int y = wave; y < v.Y; y += kSumThreads / 64) {
        const int row = (x + 1) * v.sX + (y + 1) * v.sY + 1;
        for (int z = lane; z < v.Z; z += 64) {
            const int o = row + z;
            const float c = V[o];
            accC += (double)c;
            if constexpr (KIND == TF_KIND_ELECTRODE) {
                // fluxes = -c[x] + c[x-1]; at x = 0 (left_bc - c)/0.5; zero
                // where c or c[x-1] is 0 (taufactor/electrode.py:80-83)
                const float cp = V[o - v.sX];
                float d = x == 0 ? (1.0f - c) / 0.5f : -c + cp;
                if (c == 0.0f || cp == 0.0f) d = 0.0f;
                accF += (double)d;
            } else if (has_face) {
                const float cn = V[o + v.sX];
                float d = cn - c;
                if constexpr (KIND == TF_KIND_SOLVER) {
                    if (Cv[o] == 0 || Cv[o + v.sX] == 0) d = 0.0f;
                } else if constexpr (KIND == TF_KIND_ANISOTROPIC) {
                    // Solver.vertical_flux masks faces of voxels whose factor
                    // exceeds 8 (taufactor/taufactor.py:417-418)
                    if (hm[Cv[o]] > 8.0f || hm[Cv[o + v.sX]] > 8.0f) d = 0.0f;
                } else if constexpr (KIND == TF_KIND_MULTIPHASE) {
                    d = hm[Cv[o] * ncls + Cv[o + v.sX]] * d;
                } else {
                    d = hmean(Cv[o], Cv[o + v.sX]) * d;
                }
                accF += (double)d;
            }
        }
    }
    __shared__ double redC[kSumThreads / 64], redF[kSumThreads / 64];
    accC = wave_sum(accC);
    accF = wave_sum(accF);
    if (lane == 0) {
        redC[wave] = accC;
        redF[wave] = accF;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sc = 0.0, sf = 0.0;
        for (int i = 0; i < kSumThreads / 64; ++i) {
            sc += redC[i];
            sf += redF[i];
        }
        sum_conc = sc;
        sum_flux = sf;
    }
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

extern "C" int tf_layer_sums(const float* field, const void* codes, const float* hm_lut,
                             int n_classes, int kind, int B, int X, int Y, int Z, double* flux_sum,
                             double* conc_sum, void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!field || !codes || !conc_sum || (X > 1 && !flux_sum)) return fail(-1, "null pointer argument");
    if (kind == TF_KIND_MULTIPHASE && (!hm_lut || n_classes < 1 || n_classes > 256))
        return fail(-1, "multiphase sums need hm_lut and 1..256 classes");
    if (X > 65535 || B > 65535) return fail(-1, "grid too large for layer sums");
    Vol v = make_vol(X, Y, Z);
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(X, B);
    if (kind == TF_KIND_SOLVER)
        hipLaunchKernelGGL(k_layer_sums<TF_KIND_SOLVER>, grid, dim3(kSumThreads), 0, st, field,
                           static_cast<const uint8_t*>(codes), hm_lut, n_classes, v, flux_sum, conc_sum);
    else if (kind == TF_KIND_MULTIPHASE)
        hipLaunchKernelGGL(k_layer_sums<TF_KIND_MULTIPHASE>, grid, dim3(kSumThreads), 0, st, field,
                           static_cast<const uint8_t*>(codes), hm_lut, n_classes, v, flux_sum, conc_sum);
    else if (kind == TF_KIND_FIELD)
        hipLaunchKernelGGL(k_layer_sums<TF_KIND_FIELD>, grid, dim3(kSumThreads), 0, st, field,
                           static_cast<const float*>(codes), hm_lut, n_classes, v, flux_sum, conc_sum);
    else if (kind == TF_KIND_ELECTRODE)
        hipLaunchKernelGGL(k_layer_sums<TF_KIND_ELECTRODE>, grid, dim3(kSumThreads), 0, st, field,
                           static_cast<const uint8_t*>(codes), hm_lut, n_classes, v, flux_sum, conc_sum);
    else if (kind == TF_KIND_ANISOTROPIC && hm_lut && n_classes == TF_ANISO_TABLE)
        hipLaunchKernelGGL(k_layer_sums<TF_KIND_ANISOTROPIC>, grid, dim3(kSumThreads), 0, st, field,
                           static_cast<const uint8_t*>(codes), hm_lut, n_classes, v, flux_sum, conc_sum);
    else
        return fail(-1, "unknown kind %d", kind);
    TF_LAUNCH_CHECK("k_layer_sums");
    return 0;
}

// ---------------------------------------------------------------------------
// per-voxel diffusivity field: set-up
// ---------------------------------------------------------------------------
// kpad[b][x][y][z] = k of the voxel, x ghosts copy the adjacent voxel's k,
// y/z ghosts 0 (the reference's diff_map padding, taufactor/taufactor.py:590-592);
// both field buffers get the initial field, conductive = k > 0.
__global__ void k_setup_field(const float* __restrict__ k, const float* __restrict__ xinit,
                              float* __restrict__ kpad, float* __restrict__ fa,
                              float* __restrict__ fb, Vol v, int B) {
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)B * v.sB;
    if (gid >= total) return;
    int b = (int)(gid / v.sB);
    int r = (int)(gid - (long long)b * v.sB);
    int xp = r / v.sX;
    int r2 = r - xp * v.sX;
    int yp = r2 / v.sY;
    int zp = r2 - yp * v.sY;
    const float* K = k + (size_t)b * v.X * v.Y * v.Z;
    const bool yz_in = (yp >= 1 && yp <= v.Y && zp >= 1 && zp <= v.Z);
    float f = 0.0f, kk = 0.0f;
    if (yz_in && xp >= 1 && xp <= v.X) {
        kk = K[((size_t)(xp - 1) * v.Y + (yp - 1)) * v.Z + (zp - 1)];
        f = (kk > 0.0f ? 1.0f : 0.0f) * xinit[xp - 1];
    } else if (yz_in) {
        const int x = xp == 0 ? 0 : v.X - 1;
        kk = K[((size_t)x * v.Y + (yp - 1)) * v.Z + (zp - 1)];
        f = xp == 0 ? -1.0f : 1.0f;
    }
    const size_t o = (size_t)b * v.sB + r;
    kpad[o] = kk;
    fa[o] = f;
    fb[o] = f;
}

__global__ void k_layer_count_field(const float* __restrict__ k, uint32_t* __restrict__ cnt, Vol v) {
    int x = blockIdx.x, b = blockIdx.y;
    const float* K = k + ((size_t)b * v.X + x) * v.Y * v.Z;
    int n = v.Y * v.Z;
    uint32_t acc = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += K[i] > 0.0f;
    __shared__ uint32_t red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[(size_t)b * v.X + x] = red[0];
}

extern "C" int tf_setup_field(const float* k, const float* xinit, float* kpad, float* field_a,
                              float* field_b, uint32_t* layer_count, int B, int X, int Y, int Z,
                              void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!k || !xinit || !kpad || !field_a || !field_b || !layer_count)
        return fail(-1, "null pointer argument");
    if (field_a == field_b) return fail(-1, "field_a and field_b must be distinct buffers");
    if (X > 65535 || B > 65535) return fail(-1, "grid too large");
    Vol v = make_vol(X, Y, Z);
    hipStream_t st = (hipStream_t)stream;
    long long total = (long long)B * v.sB;
    hipLaunchKernelGGL(k_setup_field, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, k, xinit,
                       kpad, field_a, field_b, v, B);
    TF_LAUNCH_CHECK("k_setup_field");
    hipLaunchKernelGGL(k_layer_count_field, dim3(X, B), dim3(256), 0, st, k, layer_count, v);
    TF_LAUNCH_CHECK("k_layer_count_field");
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
template <typename KFUN>
__device__ __forceinline__ double voxel_grad(const float* __restrict__ V, int o, const Vol v, int x,
                                             KFUN kat) {
    const int offs[6] = {v.sX, -v.sX, v.sY, -v.sY, 1, -1};
    const double kv = kat(o);
    const double cv = V[o];
    double g = 0.0;
    if (kv > 0.0) {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            if ((i == 0 && x == v.X - 1) || (i == 1 && x == 0)) continue;  // Dirichlet faces below
            const double ku = kat(o + offs[i]);
            if (ku <= 0.0) continue;
            const double d = (double)V[o + offs[i]] - cv;
            const double s = kv + ku;
            g += d * d * (2.0 * ku * ku) / (s * s);
        }
        if (x == 0) g += 2.0 * (cv + 0.5) * (cv + 0.5);
        if (x == v.X - 1) g += 2.0 * (cv - 0.5) * (cv - 0.5);
    } else {
        double sum = 0.0;
        int n = 0;
        double cu[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            cu[i] = 0.0;
            if ((i == 0 && x == v.X - 1) || (i == 1 && x == 0)) continue;
            if (kat(o + offs[i]) > 0.0) {
                cu[i] = V[o + offs[i]];
                sum += cu[i];
                ++n;
            }
        }
        if (x == 0) { sum += -0.5; ++n; }
        if (x == v.X - 1) { sum += 0.5; ++n; }
        if (n == 0) return 0.0;
        const double cs = sum / n;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            if ((i == 0 && x == v.X - 1) || (i == 1 && x == 0)) continue;
            if (kat(o + offs[i]) > 0.0) g += 2.0 * (cu[i] - cs) * (cu[i] - cs);
        }
        if (x == 0) g += 2.0 * (cs + 0.5) * (cs + 0.5);
        if (x == v.X - 1) g += 2.0 * (cs - 0.5) * (cs - 0.5);
    }
    return g;
}

__global__ void k_grad_field(const float* __restrict__ F, const float* __restrict__ K, Vol v, int B,
                             double scale, float* __restrict__ grad) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int bx = blockIdx.z;
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
    if ((long long)B * X > 65535 || Y > 65535) return fail(-1, "grid too large for gradient");
    Vol v = make_vol(X, Y, Z);
    const double scale = (double)X / ((double)Y * (double)Z);
    hipLaunchKernelGGL(k_grad_field, dim3((Z + 255) / 256, Y, B * X), dim3(256), 0,
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

extern "C" int tf_grad_classes(const float* field, const uint8_t* codes, const float* d_cls,
                               int n_classes, int B, int X, int Y, int Z, double* grad_layers,
                               void* stream) {
    if (int rc = check_dims(B, X, Y, Z)) return rc;
    if (!field || !codes || !d_cls || !grad_layers) return fail(-1, "null pointer argument");
    if (n_classes < 1 || n_classes > kMaxGradClasses) return fail(-1, "n_classes=%d out of range", n_classes);
    if (X > 65535 || B > 65535) return fail(-1, "grid too large for gradient");
    Vol v = make_vol(X, Y, Z);
    const double scale = (double)X / ((double)Y * (double)Z);
    hipLaunchKernelGGL(k_grad_classes, dim3(X, B), dim3(256), 0, (hipStream_t)stream, field, codes,
                       d_cls, n_classes, v, scale, grad_layers);
    TF_LAUNCH_CHECK("k_grad_classes");
    return 0;
}

// ---------------------------------------------------------------------------
// uneven batch: many two-phase volumes, each with its own convergence
// ---------------------------------------------------------------------------
// Volumes live back to back in two packed field buffers and one packed code
// buffer.  A batch pass covers the tiles of every volume in one launch;
// tiles of volumes that have stopped return at once.  Every active volume has
// done the same number of iterations (they all start at 0 and advance
// together), so one colour and one source buffer hold for the whole launch.
// The per-volume convergence decision runs on the device (k_batch_check), so
// a sequence of passes and checks needs no host round trip and can be
// captured in a graph.
struct BatchArgs {
    float* fa;
    float* fb;
    const uint8_t* codes;
    const tf_volume* vols;
    const int* tile_vol;
    const tf_volume_state* state;
    int n_tiles;
    int src_b;
    int colour0;
    int TX;
};

template <int H, int NW, int RPW>
__global__ __launch_bounds__(NW * 64) void k_sweep_batch(BatchArgs a) {
    using Sh = RegShape<H, NW, RPW>;
    const int per = gridDim.x >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= a.n_tiles) return;
    const int vi = a.tile_vol[logical];
    if (!a.state[vi].active) return;
    const tf_volume vol = a.vols[vi];
    const Vol v = make_vol(vol.X, vol.Y, vol.Z);
    const int nTz = (vol.Z + Sh::TZ - 1) / Sh::TZ;
    const int nTy = (vol.Y + Sh::TY - 1) / Sh::TY;
    int t = logical - vol.tile_first;
    const int tz = t % nTz; t /= nTz;
    const int ty = t % nTy;
    const int tx = t / nTy;
    const float* src = (a.src_b ? a.fb : a.fa) + vol.offset;
    float* dst = (a.src_b ? a.fa : a.fb) + vol.offset;
    rsweep_tile<H, TF_KIND_SOLVER, NW, RPW>(src, dst, a.codes + vol.code_offset, nullptr, v,
                                            vol.omega, a.colour0, a.TX, 1 + tz * Sh::TZ, ty, tx,
                                            false);
}

template <int H, int NW, int RPW>
static int launch_sweep_batch(BatchArgs& a, hipStream_t st) {
    using Sh = RegShape<H, NW, RPW>;
    if constexpr (Sh::TY < 2) {
        return fail(-1, "workgroup shape %dx%d has no interior rows for n_half=%d", NW, RPW, H);
    } else {
        const size_t lds = sweep_lds_bytes<H, NW, RPW>(TF_KIND_SOLVER, 0);
        if (lds > 160 * 1024)
            return fail(-1, "sweep shape %dx%d with n_half=%d needs %zu B of LDS", NW, RPW, H, lds);
        const int grid = ((a.n_tiles + 7) / 8) * 8;
        static bool attr_set = false;
        if (!attr_set && lds > 64 * 1024) {
            if (hipFuncSetAttribute((const void*)k_sweep_batch<H, NW, RPW>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
                (void)hipGetLastError();
            attr_set = true;
        }
        hipLaunchKernelGGL((k_sweep_batch<H, NW, RPW>), dim3(grid), dim3(NW * 64), lds, st, a);
        TF_LAUNCH_CHECK("k_sweep_batch");
        return 0;
    }
}

template <int H>
static int dispatch_batch_shape(BatchArgs& a, int nw, int rpw, hipStream_t st) {
    if (nw == 8 && rpw == 2) return launch_sweep_batch<H, 8, 2>(a, st);
    if (nw == 12 && rpw == 2) return launch_sweep_batch<H, 12, 2>(a, st);
    if (nw == 16 && rpw == 2) return launch_sweep_batch<H, 16, 2>(a, st);
    if constexpr (H <= 4) {
        if (nw == 4 && rpw == 4) return launch_sweep_batch<H, 4, 4>(a, st);
        if (nw == 8 && rpw == 4) return launch_sweep_batch<H, 8, 4>(a, st);
    }
    return fail(-1, "workgroup shape %d waves x %d rows is not compiled for n_half %d", nw, rpw, H);
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
    const long long ty = (long long)t.waves * t.rows_per_wave - 2 * n_half;
    const long long tz = reg_tile_tz(n_half);
    long long n = ((Z + tz - 1) / tz) * ((Y + ty - 1) / ty) * ((X + t.tx - 1) / t.tx);
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
    BatchArgs a;
    a.fa = field_a; a.fb = field_b; a.codes = codes; a.vols = vols; a.tile_vol = tile_vol;
    a.state = state; a.n_tiles = n_tiles; a.src_b = src_buffer; a.colour0 = first_colour;
    a.TX = t.tx;
    hipStream_t st = (hipStream_t)stream;
    switch (n_half) {
        case 1: return dispatch_batch_shape<1>(a, t.waves, t.rows_per_wave, st);
        case 2: return dispatch_batch_shape<2>(a, t.waves, t.rows_per_wave, st);
        case 3: return dispatch_batch_shape<3>(a, t.waves, t.rows_per_wave, st);
        case 4: return dispatch_batch_shape<4>(a, t.waves, t.rows_per_wave, st);
        case 5: return dispatch_batch_shape<5>(a, t.waves, t.rows_per_wave, st);
        default: return dispatch_batch_shape<6>(a, t.waves, t.rows_per_wave, st);
    }
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
