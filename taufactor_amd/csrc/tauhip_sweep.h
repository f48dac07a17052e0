#pragma once
// tauhip_sweep.h — device code shared by the translation units of libtauhip:
// volume geometry, raw buffer access, the exact division, the voxel updates,
// and the two temporally blocked sweep kernels with their launch helpers.
// The sweep kernels are instantiated only in sweep_<kind>.hip (one file per
// kind, compiled in parallel); tauhip.hip holds the C-ABI and calls them
// through the sweep_dispatch_* functions declared at the end.
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
// tf_last_error's message (defined in tauhip.hip)
int fail(int code, const char* fmt, ...);
int hip_status(hipError_t e, const char* what);

#define TF_LAUNCH_CHECK(what)                                   \
    do {                                                        \
        hipError_t e_ = hipGetLastError();                      \
        if (e_ != hipSuccess) return hip_status(e_, what);      \
    } while (0)


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

static inline int check_dims(int B, int X, int Y, int Z) {
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
// gives, a zero possibly of the other sign, which gives the same update).
// One reciprocal product plus one residual correction with the correctly
// rounded reciprocal r = RN(1/d) (r = 0 and a zero divisor stand for inf);
// the correction is exact unless the residual is subnormal, and those
// quotients take the IEEE division.  Exhaustively checked against s/d over
// all float32 s for each d (tests/csrc/divcheck.c, run by
// tests/test_division.py).
__device__ __forceinline__ float div_by_count(float s, float r, float d, int code) {
    const float dd = code != 0 ? d : 0.0f;
    float q = s * r;
    const float e = __builtin_fmaf(-q, dd, s);
    q = __builtin_fmaf(e, r, q);
    if (__builtin_amdgcn_classf(e, 0x90)) q = s / d;
    return q;
}

// table of (1/d, d) per Solver code; code 0 means factor = inf
static __constant__ float c_rcp[9] = {0.0f, 1.0f, 0.5f, 1.0f / 3.0f, 0.25f,
                               0.2f, 1.0f / 6.0f, 1.0f / 7.0f, 0.125f};

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
// temporally blocked sweep: H half-sweeps per HBM pass
// ---------------------------------------------------------------------------
// A workgroup owns an interior tile [y0,y1) x [z0,z1) x [x0,x1) (padded
// coordinates) and streams it plane by plane along x, each plane extended by
// a halo in y and z.  Half-sweep h (h = 0..H-1, colour first_colour^(h&1))
// runs on plane s-h at step s, on the tile grown by H-1-h voxels, so that
// after the last half-sweep the owned tile holds exactly the values of H
// successive reference iterations.  Red-black colouring makes each
// half-sweep a Jacobi step within one colour, so reading src and writing dst
// (no in-place hazard between workgroups that recompute each other's halos)
// gives the in-place result.  The kernel is rsweep_tile below.
//
// Global memory conventions.  Loads: one descriptor per array for the whole
// tile.  The field descriptor starts kSlackBytes before the volume and
// spans it plus the slack on both sides, so the window's first pair (z = ew,
// at least 1 - 6) is at a non-negative offset whenever its row exists; the
// lane offset holds row and z, the wave-uniform offset the plane.  Rows
// outside [0, Y+1] and planes outside the tile's stream use a lane offset
// past any record count: loads read 0, no memory traffic.  The range check
// is per dword, so pairs running past the volume read 0 there.  Stores: one
// descriptor per row and plane, based at z = 0 of that row with record
// count 4*z1 (0 outside the owned rows and planes), so a pair straddling
// z1 = Z+1 stores only its first voxel; pairs before z0 use the
// out-of-range lane offset.  No branches around memory instructions, so the
// compiler's waitcnt pass keeps the prefetched planes in flight.
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
    long long fsec;    // field kind: floats between the gx, gy, gz face sections
    int resolve_only;  // set TX and the tile counts, launch nothing
    int periodic;      // y/z neighbours wrap round (Y and Z even)
    int whole_z;       // one z window per row covers z = 1..Z (Z <= 128, closed)
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

constexpr int kRowPairs = 64;           // lanes per row
constexpr int kRowLen = 2 * kRowPairs;  // voxels per LDS row
constexpr int kSlackBytes = TF_FIELD_SLACK_BYTES;



// 2ab/(a+b), 0 when a+b == 0, with the reference's float32 operation order
// ((2*a)*b)/(a+b) (MultiPhaseSolver._harmonic_mean, taufactor/taufactor.py:577-583)
__device__ __forceinline__ float hmean(float a, float b) {
    const float den = a + b;
    return den > 0.0f ? ((2.0f * a) * b) / den : 0.0f;
}

// per-voxel auxiliary data the layer sums read: a one-byte code (Solver,
// MultiPhase, anisotropic, electrode) or the voxel diffusivity (field)
template <int KIND>
using code_t = std::conditional_t<KIND == TF_KIND_FIELD, float, uint8_t>;

// (1/d, d) per Solver code for the sweep.  Code 0 (factor inf) uses (0, 0):
// the sequence of div_by_count then yields a zero whose sign can differ from
// s/inf only when s is -0, and c + w*(q - c) is the same float for either
// zero q whatever c is (c = +-0 gives +0, c != 0 gives c - w*c).
static __constant__ float c_rd[18] = {0.0f, 0.0f, 1.0f, 1.0f, 0.5f, 2.0f, 1.0f / 3.0f, 3.0f, 0.25f, 4.0f,
                               0.2f, 5.0f, 1.0f / 6.0f, 6.0f, 1.0f / 7.0f, 7.0f, 0.125f, 8.0f};

// ---------------------------------------------------------------------------
// the register-blocked sweep (every kind)
// ---------------------------------------------------------------------------
// Compute mapping: one lane per pair of voxels in z (a 128-voxel window of z
// starting at ew per row), so the red/black pattern is uniform across a
// wave.  A wave owns RPW consecutive rows and keeps every plane of the ring
// in registers:
// x neighbours and the y neighbours inside the wave are registers, the z
// neighbour across lanes comes by a DPP wave shift, and only the first and
// last row of each wave go through LDS.  Half-sweep h of step s reads the
// edge rows that the neighbouring waves published at step s-1 (the values
// of plane s-h after half-sweep h-1, or as loaded for h = 0), double
// buffered by step parity, so one barrier per step suffices.
__host__ __device__ constexpr bool own_code_kind(int kind) {
    return kind == TF_KIND_SOLVER || kind == TF_KIND_ANISOTROPIC || kind == TF_KIND_ELECTRODE;
}


// z tiling of the register-blocked sweep: wide tiles of TZ while a full one
// fits, then one narrow strip (z0 = 1 + wide * TZ) if the rest fits TZN,
// else one more wide tile
struct ZTiling {
    int wide;     // wide tiles
    bool narrow;  // a narrow strip follows
};
__host__ __device__ inline ZTiling z_tiling(int Z, int TZ, int TZN) {
    const int w = Z / TZ, rest = Z - w * TZ;
    return {w + (rest > TZN ? 1 : 0), rest > 0 && rest <= TZN};
}
// whole-z windows (rsweep_tile's wz): closed y/z boundaries and a row that
// one window of 128 z covers while the strips would need more than one
__host__ __device__ inline bool whole_z_window(int Z, int TZ, bool periodic) {
    return !periodic && Z > TZ && Z <= 128;
}
// owned z of a wide and of a narrow tile at depth H (RegShape::TZ, ::TZN)
__host__ __device__ constexpr int reg_tile_tz(int H) { return ((128 - (H + (H & 1)) - H) / 2) * 2; }
__host__ __device__ constexpr int reg_tile_tzn(int H) { return ((64 - (H + (H & 1)) - H) / 2) * 2; }

template <int H, int NW, int RPW>
struct RegShape {
    static constexpr int EY = NW * RPW;
    static constexpr int TY = EY - 2 * H;
    static constexpr int ML = H + (H & 1);                    // z0 - ew: even, so ew is odd
    static constexpr int TZ = reg_tile_tz(H);    // owned voxels of z (even): 128 - ML - H
    static constexpr int TZN = reg_tile_tzn(H);  // the same for a narrow tile: 64 - ML - H
    // planes in registers (even): H + 2 in use and D in flight; D = 1 at
    // H = 5 keeps a 16-wave workgroup within 128 VGPRs per lane
    static constexpr int NR = H == 5 ? 8 : ((H + 5) / 2) * 2;
    static constexpr int D = NR - H - 2;
    static constexpr int TBL = 640;                           // LDS bytes for the table
    static constexpr int EB = (NW + 2) * 2 * kRowPairs;       // floats per edge buffer
    // field kind: the y faces of each wave's last row, per ring plane
    static constexpr int GEB = (NW + 2) * 2 * kRowPairs;     // floats per ring plane
    // MultiPhase kind: the class codes (16 bits: the lane's pair) of each
    // wave's first and last row, per ring plane
    static constexpr int KEB = (NW + 2) * 2 * kRowPairs;     // uint16 per ring plane
    // bytes before the edge buffers: the (1/d, d) or prefactor table, or the
    // ncls x ncls face-conductance table of the MultiPhase kind
    static constexpr int table_bytes(int kind, int ncls) {
        return kind == TF_KIND_MULTIPHASE ? std::max(TBL, (ncls * ncls * 4 + 15) / 16 * 16)
               : kind == TF_KIND_MULTIPHASE_FACES ? 16 : TBL;
    }
    static constexpr size_t lds_bytes(int kind, int ncls = 0) {
        return table_bytes(kind, ncls) + (size_t)H * 2 * EB * sizeof(float) +
               (kind == TF_KIND_FIELD ? (size_t)NR * GEB * sizeof(float) : 0) +
               (kind == TF_KIND_MULTIPHASE || kind == TF_KIND_MULTIPHASE_FACES
                    ? (size_t)NR * KEB * sizeof(uint16_t) : 0);
    }
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
// A whole-z tile (wz; closed y/z boundaries, Z <= 128) has its window start
// at z = 1: the 64 lanes hold z = 1..128, so no z halo is recomputed and one
// wave covers the row.  The z ghosts then lie outside the window (z = 0) or,
// for Z = 128, past its end (z = 129), and the DPP shifts put 0 there: the
// ghost's value (closed boundaries keep the y/z ghosts at 0) and, read as a
// class code or a face-code offset, the zero class and the zero face
// conductance (the field kind's z- face of z = 1 is the face stored at the
// ghost, 0 as well).  Used where the strips below would cover Z with more
// than one wave per row (TZ < Z <= 128).
//
// A narrow tile (nar) covers the last strip of z when it is at most TZN wide:
// each half-wave holds a 64-voxel window of z, half 0 for the rows of y tile
// 2*ty and half 1 for those of y tile 2*ty + 1 (TY rows further, an even
// count, so the active element stays wave-uniform).  The DPP shifts mix the
// halves only at lanes 31 and 32, which are halo lanes of both windows, and
// the edge rows go through LDS per lane as before; only the addresses and
// the stores (one per half) differ.
// The per-code table of the Solver ((1/d, d) pairs), anisotropic and
// electrode kinds, in static LDS: its address is known at compile time, so
// a lookup is one LDS read with no base add (the dynamic area's base is
// not).  The other kinds allocate nothing here.
template <int KIND>
__device__ __forceinline__ float* code_table() {
    if constexpr (KIND == TF_KIND_SOLVER) {
        __shared__ __align__(16) float t[18];
        return t;
    } else if constexpr (KIND == TF_KIND_ANISOTROPIC) {
        __shared__ __align__(16) float t[TF_ANISO_TABLE];
        return t;
    } else if constexpr (KIND == TF_KIND_ELECTRODE) {
        __shared__ __align__(16) float t[TF_ELECTRODE_TABLE];
        return t;
    } else if constexpr (KIND == TF_KIND_MULTIPHASE_FACES) {
        __shared__ __align__(16) float t[TF_MPF_FACE_SLOTS + 2 * TF_MPF_MAX_FACTORS + 2];
        return t;
    } else {
        return nullptr;
    }
}

// AC: the active element of row 0 at step 0 of a block, (colour0 & 1) ^ 1 for
// the launch's first colour (AA in the loop body); each value is its own
// kernel, so that the register allocation of one colour's loop is not the
// other's (the 512^3 Solver sweep ran 8 % slower in one of the two bodies of
// a kernel holding both)
template <int H, int KIND, int NW, int RPW, int AC>
__device__ __forceinline__ void rsweep_tile(const float* __restrict__ S, float* __restrict__ D,
                                            const uint8_t* __restrict__ Cg,
                                            const float* __restrict__ hm, const Vol v, float w,
                                            int TX, int z0, int ty, int tx, bool nar,
                                            long long fsec = 0, int ncls = 0, bool wrap = false,
                                            bool wz = false) {
    using Sh = RegShape<H, NW, RPW>;
    // field kind: Cg holds three sections of face conductances in the padded
    // field layout, fsec floats apart: gx, gy, gz, each stored at the voxel
    // on the face's minus side (tf_field_faces)
    constexpr bool kFaces = KIND == TF_KIND_FIELD;
    // MultiPhase kind: Cg holds the class codes (sweep-code layout), hm the
    // ncls x ncls face-conductance table
    constexpr bool kClasses = KIND == TF_KIND_MULTIPHASE;
    // MultiPhase face codes: Cg holds one uint32 per voxel in the field
    // layout (byte offsets of the x+, y+, z+ face conductances and of the
    // (1/f, f) prefactor pair), hm the face table and the pairs
    constexpr bool kFaceCodes = KIND == TF_KIND_MULTIPHASE_FACES;
    constexpr int NG = kFaces ? Sh::NR : 1;
    constexpr int NKF = kFaceCodes ? Sh::NR : 1;
    constexpr int TY = Sh::TY, ML = Sh::ML, NR = Sh::NR, DP = Sh::D, EB = Sh::EB;
    static_assert(TY >= 2 && TY % 2 == 0, "tile rows must be even and positive");
    static_assert(RPW % 2 == 0, "rows per wave must be even");
    extern __shared__ __align__(16) unsigned char smem[];
    float* const sT = reinterpret_cast<float*>(smem);
    float* const sE = reinterpret_cast<float*>(smem + Sh::table_bytes(KIND, ncls));
    float* const sGE = sE + (size_t)H * 2 * EB;  // field kind: y faces of edge rows
    uint16_t* const sKE = reinterpret_cast<uint16_t*>(sGE);  // MultiPhase: classes of edge rows

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // y tile of each half (for a wide tile half 1 owns nothing)
    const int y0h[2] = {1 + (nar ? 2 * ty : ty) * TY, 1 + (nar ? 2 * ty + 1 : ty) * TY};
    const int x0 = 1 + tx * TX;
    const int x1 = min(x0 + TX, v.X + 1);
    const int z1 = wz ? v.Z + 1 : min(z0 + (nar ? Sh::TZN : Sh::TZ), v.Z + 1);
    const int half = nar ? lane >> 5 : 0;                 // this lane's half
    const int y0 = half ? y0h[1] : y0h[0];                  // this lane's y tile
    const int ey0 = y0 - H;
    const int ew = wz ? 1 : z0 - ML;                        // odd
    const int zl = ew + 2 * (nar ? (lane & 31) : lane);     // z of the lane's first voxel
    const int lx1 = min(x1 + H, v.X + 2);

    float* const s_code = code_table<KIND>();
    if (KIND == TF_KIND_SOLVER) {
        if (threadIdx.x < 18) s_code[threadIdx.x] = c_rd[threadIdx.x];
    } else if (KIND == TF_KIND_ANISOTROPIC) {
        if (threadIdx.x < TF_ANISO_TABLE) s_code[threadIdx.x] = hm[threadIdx.x];
    } else if (KIND == TF_KIND_ELECTRODE) {
        if (threadIdx.x < TF_ELECTRODE_TABLE) s_code[threadIdx.x] = hm[threadIdx.x];
    } else if (KIND == TF_KIND_MULTIPHASE_FACES) {
        // ncls = the table's length in floats
        for (int i = threadIdx.x; i < ncls; i += NW * 64) s_code[i] = hm[i];
    }
    if (KIND == TF_KIND_MULTIPHASE) {
        for (int i = threadIdx.x; i < ncls * ncls; i += NW * 64) sT[i] = hm[i];
    }
    if (KIND == TF_KIND_MULTIPHASE || KIND == TF_KIND_MULTIPHASE_FACES) {
        // the class / face ring's pad slots (no wave above the first / below
        // the last) are read as table indices or offsets: keep them 0
        for (int i = threadIdx.x; i < Sh::NR * 2 * 2 * kRowPairs; i += NW * 64) {
            const int pl = i / (4 * kRowPairs), r = i % (4 * kRowPairs);
            const int slot = r < 2 * kRowPairs ? 0 : NW + 1;
            sKE[pl * Sh::KEB + slot * 2 * kRowPairs + r % (2 * kRowPairs)] = 0;
        }
    }
    const float ky = KIND == TF_KIND_ANISOTROPIC ? hm[TF_ANISO_TABLE - 2] : 0.0f;
    const float kz = KIND == TF_KIND_ANISOTROPIC ? hm[TF_ANISO_TABLE - 1] : 0.0f;

    // ---- global memory (descriptor conventions: see the sweep section head) ----
    // loads: one 8-byte field load (the lane's pair) and one 2-byte code
    // load per row and plane, straight into the register ring
    constexpr int kOff = 0x7fffff00;
    const int plane_f = v.sX * 4;
    // the descriptors start at the tile's first streamed plane (pbase), so
    // offsets stay within the tile's planes (< 2 GiB, checked by the
    // launch) whatever the size of the volume
    const int pbase = max(x0 - H, 0);
    const uint32_t span_f = (uint32_t)(lx1 - pbase) * (uint32_t)plane_f;
    const rsrc_t rF = make_rsrc(reinterpret_cast<const char*>(S + (size_t)pbase * v.sX) - kSlackBytes,
                                span_f + 2u * kSlackBytes);
    const int row_c = sweep_code_row(v.Z);
    const int plane_c = (v.Y + 2) * row_c;
    const rsrc_t rC = make_rsrc(Cg + (size_t)pbase * plane_c, (uint32_t)(lx1 - pbase) * (uint32_t)plane_c);
    // face codes: one descriptor over the uint32 array, laid out (and
    // offset) like the field
    const rsrc_t rK = make_rsrc(reinterpret_cast<const char*>(reinterpret_cast<const uint32_t*>(Cg) +
                                                             (size_t)pbase * v.sX) -
                                    kSlackBytes,
                                span_f + 2u * kSlackBytes);
    rsrc_t rG[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        rG[k] = make_rsrc(reinterpret_cast<const char*>(reinterpret_cast<const float*>(Cg) + k * fsec +
                                                        (size_t)pbase * v.sX) -
                              kSlackBytes,
                          span_f + 2u * kSlackBytes);
    int vf[RPW], vc[RPW];
    bool own_row[2][RPW];   // wave-uniform, per half
    int ys[2][RPW];
    // periodic y/z (wrap): every row and z pair of the tile, ghosts and halo
    // beyond them included, is loaded from the interior voxels it repeats
    // ((p - 1) mod N + 1), codes included, so it is updated exactly as they
    // are; with Y and Z even the red/black colouring repeats as well.  The
    // lane's pair starts at an odd z, and Z is even, so the pair never
    // straddles the wrap.
    auto wrap_at = [](int p, int n) {
        int r = (p - 1) % n;
        return (r < 0 ? r + n : r) + 1;
    };
    const int zr = wrap ? wrap_at(zl, v.Z) : zl;
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
        const int y = ey0 + wave * RPW + j;
        const bool rok = wrap || (y >= 0 && y <= v.Y + 1);
        const int yr = wrap ? wrap_at(y, v.Y) : y;
        vf[j] = rok ? kSlackBytes + (yr * v.sY + zr) * 4 : kOff;
        vc[j] = rok ? yr * row_c + zr + 3 : kOff;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int yk = y0h[k] - H + wave * RPW + j;
            own_row[k][j] = (k == 0 || nar) && yk >= y0h[k] && yk < min(y0h[k] + TY, v.Y + 1);
            ys[k][j] = own_row[k][j] ? yk : 0;
        }
    }
    u32x2_t F[NR][RPW];
    uint32_t K[NR][RPW];
    u32x2_t G[NG][RPW][3];  // field kind: gx, gy, gz pairs
    u32x2_t KF[NKF][RPW];   // face codes of the lane's pair
    // slots the prologue does not load hold planes before the stream start
    // in the first steps: defined values, read by no active half-sweep
#pragma unroll
    for (int i = 0; i < NR; ++i)
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
            F[i][j] = u32x2_t{0u, 0u};
            K[i][j] = 0u;
        }
#pragma unroll
    for (int i = 0; i < NG; ++i)
#pragma unroll
        for (int j = 0; j < RPW; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) G[i][j][k] = u32x2_t{0u, 0u};
#pragma unroll
    for (int i = 0; i < NKF; ++i)
#pragma unroll
        for (int j = 0; j < RPW; ++j) KF[i][j] = u32x2_t{0u, 0u};
    // a plane outside the tile's stream [pbase, lx1) is loaded with the
    // scalar offset at the descriptor's record count, so every lane's offset
    // is out of range and the load returns zeros (no per-lane select); the
    // 32-bit sums stay below 2^32: lane offsets are below 2^31 - 128 and the
    // record counts below 2^31 (a tile's planes span less than 2 GiB)
    const int rec_f = (int)(span_f + 2u * kSlackBytes);
    const int rec_c = (int)((uint32_t)(lx1 - pbase) * (uint32_t)plane_c);
    auto issue_load = [&](int p, auto slot_c) {
        constexpr int sl = decltype(slot_c)::value;
        const bool pok = p >= pbase && p < lx1;
        const int sf = pok ? (p - pbase) * plane_f : rec_f, sc = pok ? (p - pbase) * plane_c : rec_c;
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
            F[sl][j] = __builtin_amdgcn_raw_buffer_load_b64(rF, vf[j], sf, 0);
            if constexpr (kFaces) {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    G[sl][j][k] = __builtin_amdgcn_raw_buffer_load_b64(rG[k], vf[j], sf, 0);
            } else if constexpr (kFaceCodes) {
                KF[sl][j] = __builtin_amdgcn_raw_buffer_load_b64(rK, vf[j], sf, 0);
            } else {
                K[sl][j] = __builtin_amdgcn_raw_buffer_load_b16(rC, vc[j], sc, 0);
            }
        }
    };
    // stores: one descriptor per plane step (based at the plane, record
    // count the plane's bytes) and one 8-byte store per row; a lane writes
    // its pair of row j when it owns the pair's first voxel (its half's row j
    // is owned and z0 <= zl < z1), else its offset lies past the plane.  With
    // Z odd the last pair's second voxel is the z = Z + 1 ghost, written too:
    // its register holds what the ghost holds (+0 with closed boundaries,
    // code 0 keeping it; a periodic volume with Z odd re-wraps its ghosts
    // before every half-sweep).  The stores are nontemporal (cache policy
    // nt): the next read of the plane is the next launch's, so the written
    // lines need not stay in L2, where the loaded planes' halo lines are
    // reused by neighbouring tiles (512^3 Solver at H = 6: 0.3854 -> 0.3813
    // ms per launch, H = 4: 0.3476 -> 0.3375)
    constexpr int kStoreNT = 2;
    int so[RPW];
#pragma unroll
    for (int j = 0; j < RPW; ++j) {
        const bool own = (half ? own_row[1][j] : own_row[0][j]) && zl >= z0 && zl < z1;
        so[j] = own ? ((half ? ys[1][j] : ys[0][j]) * v.sY + zl) * 4 : kOff;
    }
    auto store_plane = [&](int q, auto slot_c) {
        constexpr int sl = decltype(slot_c)::value;
        const bool qok = q >= x0 && q < x1;
        const rsrc_t r = make_rsrc(D + (size_t)(qok ? q : 0) * v.sX, qok ? (uint32_t)(4 * v.sX) : 0u);
#pragma unroll
        for (int j = 0; j < RPW; ++j) __builtin_amdgcn_raw_buffer_store_b64(F[sl][j], r, so[j], 0, kStoreNT);
    };

    // ---- per-tile bounds of each half-sweep (planes only: rows and z are
    // updated unconditionally; outside a half-sweep's region the result is
    // never read by a later half-sweep of this pass nor stored, and ghosts
    // (value 0 with code 0, whose table entry gives a zero quotient) keep
    // their value) ----
    // A wave whose rows all lie outside half-sweep h's rows, [h + 1, EY - 1 - h)
    // of the EY rows from ey0 (the tile grown by H - 1 - h), skips that
    // half-sweep: its plane count is 0.  The rows it leaves stale are read
    // only by rows outside the next half-sweep's region (one row narrower at
    // each end), whose results no owned row reads and no store writes.
    int plo[H], pn[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int grow = H - 1 - h;
        const int r0 = wave * RPW;
        const bool live = r0 + RPW > h + 1 && r0 < Sh::EY - 1 - h;
        plo[h] = max(x0 - grow, 1);
        pn[h] = live ? min(x1 + grow, v.X + 1) - plo[h] : 0;
    }
    // unclipped stream start: with TX and TY even, S0 + ey0 is odd, so the
    // active element of row j at step u of an unrolled block,
    // (s + y + colour0) & 1, is AA ^ ((u + j) & 1) with AA = AC = colour0 ^ 1 for
    // every tile of the launch; the parity of S0 is that of H
    const int S0 = x0 - (H - 1);
    const int S1 = (x1 - 1) + (H - 1);

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
        // field kind: the y faces of the last row of a freshly loaded plane,
        // read by the first row of the wave below in every half-sweep of
        // that plane (LDS ring slot = register ring slot)
        auto publish_faces = [&](auto slot_c) {
            constexpr int sl = decltype(slot_c)::value;
            if constexpr (kFaces) {
                reinterpret_cast<u32x2_t*>(sGE + sl * Sh::GEB)[(wave + 1) * kRowPairs + lane] =
                    G[sl][RPW - 1][1];
            } else if constexpr (kClasses) {
                // classes of the first (top) and last (bottom) row of a freshly
                // loaded plane, read by the neighbouring waves in every
                // half-sweep of that plane
                uint16_t* ke = sKE + sl * Sh::KEB + (wave + 1) * 2 * kRowPairs + lane;
                ke[0] = (uint16_t)K[sl][0];
                ke[kRowPairs] = (uint16_t)K[sl][RPW - 1];
            } else if constexpr (kFaceCodes) {
                // the y+ face offsets of the last row's pair, read as the y-
                // faces of the first row of the wave below
                uint16_t* ke = sKE + sl * Sh::KEB + (wave + 1) * 2 * kRowPairs + lane;
                const u32x2_t k = KF[sl][RPW - 1];
                ke[kRowPairs] = (uint16_t)(((k.x >> 6) & 63u) | (((k.y >> 6) & 63u) << 8));
            }
        };

        // chk: test the plane bounds of the half-sweep (the steady-state
        // blocks, where every plane of every half-sweep is in bounds, skip it)
        auto stage = [&](auto h_c, auto u_c, int p, auto chk_c) {
            constexpr int h = decltype(h_c)::value, u = decltype(u_c)::value;
            constexpr int sl = ((u - h + 1) % NR + NR) % NR;
            constexpr int slm = (sl + NR - 1) % NR, slp = (sl + 1) % NR;
            if constexpr (decltype(chk_c)::value) {
                if ((unsigned)(p - plo[h]) >= (unsigned)pn[h]) return;
            }
            const float* e = ebuf(h, par(u + 1));
            const float ym0 = e[-kRowPairs];        // bottom row of the wave above
            const float ypL = e[2 * kRowPairs];     // top row of the wave below
            u32x2_t gy0{0u, 0u};                    // y faces of the wave above's last row
            if constexpr (kFaces)
                gy0 = reinterpret_cast<const u32x2_t*>(sGE + sl * Sh::GEB)[wave * kRowPairs + lane];
            uint32_t ka = 0, kb = 0;                // classes: last row above, first row below
            if constexpr (kClasses) {
                const uint16_t* ke = sKE + sl * Sh::KEB + lane;
                ka = ke[wave * 2 * kRowPairs + kRowPairs];
                kb = ke[(wave + 2) * 2 * kRowPairs];
            } else if constexpr (kFaceCodes) {
                // face codes: the y+ face offsets of the last row above
                ka = sKE[sl * Sh::KEB + lane + wave * 2 * kRowPairs + kRowPairs];
            }
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
                // Solver: the byte is 8 x the count, the offset of its table entry
                const uint32_t code = a ? (K[sl][j] >> 8) : (K[sl][j] & 0xffu);
                float out;
                if constexpr (KIND == TF_KIND_SOLVER) {
                    const float2 rdv =
                        *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(s_code) + code);
                    float s = xp + xm;
                    s = s + yp;
                    s = s + ym;
                    s = s + zp;
                    s = s + zm;
                    float qt = s * rdv.x;
                    const float e2 = __builtin_fmaf(-qt, rdv.y, s);
                    qt = __builtin_fmaf(e2, rdv.x, qt);
                    // rare quotients: a wave-uniform test first, so that the
                    // common path never writes exec (an exec write costs wait
                    // states before the next DPP add).  The correction is
                    // exact unless the residual e2 is subnormal (then some of
                    // its bits are lost; tests/csrc/divcheck.c checks every
                    // float32 s): one class test of e2
                    const bool rare = __builtin_amdgcn_classf(e2, 0x90);
                    if (__builtin_expect(__builtin_amdgcn_ballot_w64(rare) != 0, 0)) {
                        if (rare) qt = s / (code ? (float)(code >> 3) : __builtin_inff());
                    }
                    s = qt - c;
                    s = s * w;
                    out = c + s;
                } else if constexpr (KIND == TF_KIND_ELECTRODE) {
                    float s = xp + xm;
                    s = s + yp;
                    s = s + ym;
                    s = s + zp;
                    s = s + zm;
                    s = s / s_code[code];
                    s = s - c;
                    s = s * w;
                    out = c + s;
                } else if constexpr (KIND == TF_KIND_MULTIPHASE) {
                    // face conductances from the classes of the voxel and of
                    // its six neighbours (MultiPhaseSolver, taufactor/taufactor.py:598-613)
                    auto cls = [a](uint32_t k) { return a ? (k >> 8) & 0xffu : k & 0xffu; };
                    const uint32_t kc = K[sl][j];
                    const float* row = sT + cls(kc) * ncls;
                    const uint32_t kyp = j < RPW - 1 ? K[sl][j + 1] : kb;
                    const uint32_t kym = j > 0 ? K[sl][j - 1] : ka;
                    const uint32_t czp =
                        a ? (uint32_t)__builtin_amdgcn_update_dpp(0, (int)kc, 0x130, 0xf, 0xf, false) & 0xffu
                          : (kc >> 8) & 0xffu;
                    const uint32_t czm =
                        a ? kc & 0xffu
                          : ((uint32_t)__builtin_amdgcn_update_dpp(0, (int)kc, 0x138, 0xf, 0xf, false) >> 8) & 0xffu;
                    const Nbr n{c, xp, xm, yp, ym, zp, zm};
                    out = update_multiphase(n, row[cls(K[slp][j])], row[cls(K[slm][j])], row[cls(kyp)],
                                            row[cls(kym)], row[czp], row[czm], p == 1, p == v.X, w);
                } else if constexpr (kFaceCodes) {
                    // face conductances at the byte offsets of the codes: the
                    // voxel's x+, y+, z+ faces and the x-, y-, z- faces as the
                    // plus faces of the neighbours on the minus side; the
                    // prefactor f and its correctly rounded reciprocal from
                    // the pair table (MultiPhaseSolver, taufactor/taufactor.py:
                    // 598-613, with the sums of :598-604 precomputed)
                    const char* tb = reinterpret_cast<const char*>(s_code);
                    auto face = [tb](uint32_t off) { return *reinterpret_cast<const float*>(tb + off); };
                    const u32x2_t kc = KF[sl][j];
                    const uint32_t own = a ? kc.y : kc.x;
                    const uint32_t kxm = a ? KF[slm][j].y : KF[slm][j].x;
                    const float gxp = face(own & 63u), gyp = face((own >> 6) & 63u);
                    const float gzp = face((own >> 12) & 63u);
                    const float gxm = face(kxm & 63u);
                    const float gym = j > 0 ? face(((a ? KF[sl][j - 1].y : KF[sl][j - 1].x) >> 6) & 63u)
                                            : face(a ? (ka >> 8) & 63u : ka & 63u);
                    const uint32_t kzm =
                        a ? kc.x : (uint32_t)__builtin_amdgcn_update_dpp(0, (int)kc.y, 0x138, 0xf, 0xf, false);
                    const float gzm = face((kzm >> 12) & 63u);
                    const float2 rf = *reinterpret_cast<const float2*>(tb + 4 * TF_MPF_FACE_SLOTS + (own >> 18));
                    float s = xp * gxp + xm * gxm;
                    s = s + yp * gyp;
                    s = s + ym * gym;
                    s = s + zp * gzp;
                    s = s + zm * gzm;
                    // s / f: reciprocal product and one residual correction
                    // (exact for 2^-100 <= |s| with the correctly rounded 1/f
                    // and f in [2^-20, 2^20], tests/csrc/divcheck_pairs.c;
                    // unlike the integer counts of the Solver kind, a
                    // subnormal-range s can leave a residual that rounds to
                    // zero, so the guard is on s); pair (0, 0) is a zero
                    // (infinite) prefactor, whose quotient 0 gives the same
                    // update as s/inf; smaller |s| take the IEEE division
                    float qt = s * rf.x;
                    const float e2 = __builtin_fmaf(-qt, rf.y, s);
                    qt = __builtin_fmaf(e2, rf.x, qt);
                    const bool rare = __builtin_amdgcn_frexp_expf(s) < -99;
                    if (__builtin_expect(__builtin_amdgcn_ballot_w64(rare) != 0, 0)) {
                        if (rare) qt = rf.y != 0.0f ? s / rf.y : 0.0f;
                    }
                    s = qt - c;
                    s = s * w;
                    out = c + s;
                } else if constexpr (KIND == TF_KIND_FIELD) {
                    // faces of the voxel: its own gx, gy, gz and those of its
                    // x-1 (previous plane), y-1 (row above) and z-1 (other
                    // element or previous lane) neighbours
                    auto sel = [a](u32x2_t q) { return a ? hi_f(q) : lo_f(q); };
                    const u32x2_t gz = G[sl][j][2];
                    const float gxp = sel(G[sl][j][0]), gxm = sel(G[slm][j][0]);
                    const float gyp = sel(G[sl][j][1]);
                    const float gym = sel(j > 0 ? G[sl][j - 1][1] : gy0);
                    const float gzp = sel(gz);
                    const float gzm = a ? lo_f(gz) : from_prev_lane(gz.y);
                    const Nbr n{c, xp, xm, yp, ym, zp, zm};
                    out = update_multiphase(n, gxp, gxm, gyp, gym, gzp, gzm, p == 1, p == v.X, w);
                } else {
                    float s = xp + xm;
                    float t = yp + ym;
                    t = ky * t;
                    s = s + t;
                    t = zp + zm;
                    t = kz * t;
                    s = s + t;
                    s = s / s_code[code];
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
        publish_faces(ic<1>{});
        __syncthreads();

        // one block of NR steps from s0 (ring slots are compile-time in u)
        auto block = [&](int s0, auto chk_c) {
            static_for<0, NR>([&](auto u_c) {
                constexpr int u = decltype(u_c)::value;
                const int s = s0 + u;
                static_for<0, H>([&](auto h_c) {
                    constexpr int h = decltype(h_c)::value;
                    stage(h_c, u_c, s - h, chk_c);
                    if constexpr (h + 1 < H) publish(h + 1, u, ic<((u - h + 1) % NR + NR) % NR>{});
                });
                publish(0, u, ic<(u + 2) % NR>{});
                publish_faces(ic<(u + 2) % NR>{});
                store_plane(s - (H - 1), ic<((u - H + 2) % NR + NR) % NR>{});
                issue_load(s + 2 + DP, ic<((u - H + 1) % NR + NR) % NR>{});
                __syncthreads();
            });
        };
        // blocks whose every stage is in bounds for this wave, [s_lo, s_hi]:
        // stage h of step s works on plane s - h, inside [plo[h], plo[h] + pn[h])
        // for s in [s0, s0 + NR); a wave that skips some half-sweep has none
        int s_lo = S0, s_hi = S1;
        bool skips = false;
#pragma unroll
        for (int h = 0; h < H; ++h) {
            s_lo = max(s_lo, plo[h] + h);
            s_hi = min(s_hi, plo[h] + pn[h] + h - NR);
            skips = skips || pn[h] == 0;
        }
        if (skips) s_hi = s_lo - 1;
        int s0 = S0;
        if constexpr (KIND == TF_KIND_SOLVER) {
            // the Solver kind (plain and batched) runs the steady-state blocks
            // without the bounds tests (512^3: 0.402 -> 0.385 ms per launch);
            // the kinds with more live registers per update spill in that
            // body and keep the tested one throughout (MultiPhase face codes
            // and the per-voxel field were 3 % slower with it).  Every wave
            // runs the same number of blocks, hence of barriers, whichever
            // bodies it takes.
            for (; s0 <= S1 && !(s0 >= s_lo && s0 <= s_hi); s0 += NR) block(s0, std::true_type{});
            for (; s0 >= s_lo && s0 <= s_hi; s0 += NR) block(s0, std::false_type{});
        }
        for (; s0 <= S1; s0 += NR) block(s0, std::true_type{});
    };
    body(ic<AC>{});
}

template <int H, int KIND, int NW, int RPW, int AA>
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
    // per-volume codes: the sweep-code layout, or for the field kind the
    // face sections (floats in the padded field layout)
    const uint8_t* codes =
        KIND == TF_KIND_FIELD || KIND == TF_KIND_MULTIPHASE_FACES
            ? reinterpret_cast<const uint8_t*>(static_cast<const float*>(a.codes) + off)
            : static_cast<const uint8_t*>(a.codes) + (size_t)b * sweep_code_volume(a.v.X, a.v.Y, a.v.Z);
    const float* hm = KIND == TF_KIND_ELECTRODE ? a.hm + (size_t)b * TF_ELECTRODE_TABLE : a.hm;
    rsweep_tile<H, KIND, NW, RPW, AA>(a.src + off, a.dst + off, codes, hm, a.v, a.w, a.TX,
                                  nar ? a.z0N : 1 + tz * RegShape<H, NW, RPW>::TZ, ty, tx, nar,
                                  a.fsec, a.ncls, a.periodic != 0, a.whole_z != 0);
}

// Tile geometry of the kernel k_sweep<H, KIND, NW, RPW> runs: owned rows
// and z voxels per tile, steps per unrolled block, LDS bytes.
template <int H, int KIND, int NW, int RPW>
struct SweepGeom {
    static constexpr int TY = RegShape<H, NW, RPW>::TY;
    static constexpr int TZ = RegShape<H, NW, RPW>::TZ;
    static constexpr int NP = RegShape<H, NW, RPW>::NR;
};

template <int H, int NW, int RPW>
static size_t sweep_lds_bytes(int kind, int ncls) {
    return RegShape<H, NW, RPW>::lds_bytes(kind, ncls);
}

// LDS of the kind's static table (code_table), with room for its alignment;
// a launch needs it on top of the dynamic bytes
static constexpr size_t static_table_bytes(int kind) {
    return 64 + sizeof(float) * (kind == TF_KIND_SOLVER ? 18
                                 : kind == TF_KIND_ANISOTROPIC ? TF_ANISO_TABLE
                                 : kind == TF_KIND_ELECTRODE ? TF_ELECTRODE_TABLE
                                 : kind == TF_KIND_MULTIPHASE_FACES ? TF_MPF_FACE_SLOTS + 2 * TF_MPF_MAX_FACTORS + 2
                                 : 0);
}

// Workgroups of one sweep kernel resident on the current device at once
// (CUs x workgroups per CU); 0 if the runtime cannot say.  `cache` holds the
// answer + 1 per device (0: not asked yet).
constexpr int kMaxDevices = 64;
static inline int resident_workgroups(const void* fn, int threads, size_t lds, int* cache) {
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
static inline int balanced_tx(int X, long long yz_tiles, int H, int NP, int slots, bool even) {
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
    if constexpr (Sh::TY < 2) {
        return fail(-1, "workgroup shape %dx%d has no interior rows for n_half=%d", NW, RPW, H);
    } else {
        const size_t lds = sweep_lds_bytes<H, NW, RPW>(KIND, a.ncls);
        if (lds + static_table_bytes(KIND) > 160 * 1024)
            return fail(-1, "sweep shape %dx%d with n_half=%d needs %zu B of LDS", NW, RPW, H,
                        lds + static_table_bytes(KIND));
        // one kernel per starting colour (rsweep_tile's AA)
        const void* const fn[2] = {(const void*)k_sweep<H, KIND, NW, RPW, 0>,
                                   (const void*)k_sweep<H, KIND, NW, RPW, 1>};
        static bool attr_set = false;
        if (!attr_set && lds > 64 * 1024) {
            // opt in to more than 64 KiB of dynamic LDS; if the runtime
            // refuses, the launch below reports the error
            for (const void* f : fn)
                if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)(160 * 1024 - static_table_bytes(KIND))) != hipSuccess)
                    (void)hipGetLastError();
            attr_set = true;
        }
        a.nTy = (a.v.Y + Sh::TY - 1) / Sh::TY;
        a.nTz = (a.v.Z + Sh::TZ - 1) / Sh::TZ;
        a.nTyN = 0;
        a.whole_z = whole_z_window(a.v.Z, Sh::TZ, a.periodic != 0);
        if (a.whole_z) {
            a.nTz = 1;
        } else {
            const ZTiling zt = z_tiling(a.v.Z, Sh::TZ, RegShape<H, NW, RPW>::TZN);
            a.nTz = zt.wide;
            if (zt.narrow) a.nTyN = (a.nTy + 1) / 2;
            a.z0N = 1 + (a.v.Z / Sh::TZ) * Sh::TZ;
        }
        const long long yz = (long long)a.nTz * a.nTy + a.nTyN;
        if (a.TX <= 0) {
            // the fewer resident workgroups of the two colours' kernels, so
            // that both colours of a solve tile alike
            static int slots_cache[2][kMaxDevices] = {};
            const int slots = std::min(resident_workgroups(fn[0], NW * 64, lds, slots_cache[0]),
                                       resident_workgroups(fn[1], NW * 64, lds, slots_cache[1]));
            a.TX = balanced_tx(a.v.X, yz * B, H, Sh::NP, slots, true);
        }
        // the register-blocked kernel needs an even x tile (rsweep_tile),
        // and a tile's planes must span less than 2 GiB (32-bit offsets)
        a.TX = (a.TX + 1) & ~1;
        const long long tx_max = (((1LL << 31) - 64) / ((long long)a.v.sX * 4) - 2 * H) & ~1LL;
        if (tx_max < 2) return fail(-1, "a plane of %d voxels is too large for the sweep", a.v.sX);
        if (a.TX > tx_max) a.TX = (int)tx_max;
        a.nTx = (a.v.X + a.TX - 1) / a.TX;
        const long long wide_tiles = (long long)a.nTz * a.nTy * a.nTx * B;
        const long long total = yz * a.nTx * B;
        if (total * 8 >= (1LL << 31)) return fail(-1, "too many tiles (%lld)", total);
        a.nWide = (int)wide_tiles;
        a.total = (int)total;
        if (a.resolve_only) return 0;
        const int grid = (int)(((total + 7) / 8) * 8);
        if ((a.colour0 & 1) ^ 1)
            hipLaunchKernelGGL((k_sweep<H, KIND, NW, RPW, 1>), dim3(grid), dim3(NW * 64), lds, st, a);
        else
            hipLaunchKernelGGL((k_sweep<H, KIND, NW, RPW, 0>), dim3(grid), dim3(NW * 64), lds, st, a);
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
    if constexpr (KIND == TF_KIND_SOLVER || KIND == TF_KIND_FIELD || KIND == TF_KIND_MULTIPHASE ||
                  KIND == TF_KIND_MULTIPHASE_FACES) {
        if (nw == 12 && rpw == 2) return launch_sweep<H, KIND, 12, 2>(a, B, st);
    }
    if constexpr (KIND == TF_KIND_SOLVER && H <= 4) {
        if (nw == 8 && rpw == 4) return launch_sweep<H, KIND, 8, 4>(a, B, st);
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
    // work queue (tf_batch_sweep_queue): live tile list and counters, or null
    const int* live;
    int* queue;
    int pass;
};

// Tile `logical` of the batch's tile list (every tile of every volume, in
// volume order) for the active volumes; returns at once for the others.
template <int H, int NW, int RPW, int AA>
__device__ __forceinline__ void batch_tile(const BatchArgs& a, int logical) {
    using Sh = RegShape<H, NW, RPW>;
    const int vi = a.tile_vol[logical];
    if (!a.state[vi].active) return;
    const tf_volume vol = a.vols[vi];
    const Vol v = make_vol(vol.X, vol.Y, vol.Z);
    // this volume's tiles, numbered as in k_sweep: wide tiles (tz fastest,
    // then ty, then tx), then the narrow strip (ty, then tx); a volume of
    // whole-z windows has one wide tile per (ty, tx)
    const bool wz = whole_z_window(vol.Z, Sh::TZ, false);
    const ZTiling zt = wz ? ZTiling{1, false} : z_tiling(vol.Z, Sh::TZ, Sh::TZN);
    const int nTy = (vol.Y + Sh::TY - 1) / Sh::TY;
    const int nTx = (vol.X + a.TX - 1) / a.TX;
    int t = logical - vol.tile_first;
    const int n_wide = zt.wide * nTy * nTx;
    const bool nar = t >= n_wide;
    int z0, ty, tx;
    if (!nar) {  // (zt.wide >= 1 here)
        z0 = 1 + (t % zt.wide) * Sh::TZ;
        t /= zt.wide;
        ty = t % nTy;
        tx = t / nTy;
    } else {
        t -= n_wide;
        const int nTyN = (nTy + 1) / 2;
        z0 = 1 + (vol.Z / Sh::TZ) * Sh::TZ;
        ty = t % nTyN;
        tx = t / nTyN;
    }
    const float* src = (a.src_b ? a.fb : a.fa) + vol.offset;
    float* dst = (a.src_b ? a.fa : a.fb) + vol.offset;
    rsweep_tile<H, TF_KIND_SOLVER, NW, RPW, AA>(src, dst, a.codes + vol.code_offset, nullptr, v,
                                                vol.omega, a.TX, z0, ty, tx, nar, 0, 0, false, wz);
}

template <int H, int NW, int RPW, int AA>
__global__ __launch_bounds__(NW * 64) void k_sweep_batch(BatchArgs a) {
    const int per = gridDim.x >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= a.n_tiles) return;
    batch_tile<H, NW, RPW, AA>(a, logical);
}

// The same pass as a work queue: a grid of at most the resident workgroups
// takes the live tiles (a.live[0 .. a.queue[0]), listed by k_batch_compact
// after the last check) one at a time from the pass's counter
// a.queue[1 + a.pass], which k_batch_compact zeroed.  Every workgroup leaves
// once the counter has passed the live count, so tiles of stopped volumes
// are never dispatched and no workgroup waits for another.
template <int H, int NW, int RPW, int AA>
__global__ __launch_bounds__(NW * 64) void k_sweep_batch_queue(BatchArgs a) {
    __shared__ int s_next;
    const int n_live = a.queue[0];
    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(&a.queue[1 + a.pass], 1);
        __syncthreads();
        const int t = s_next;
        __syncthreads();  // every wave has read s_next before it is written again
        if (t >= n_live) break;
        batch_tile<H, NW, RPW, AA>(a, a.live[t]);
    }
}

template <int H, int NW, int RPW>
static int launch_sweep_batch(BatchArgs& a, hipStream_t st) {
    using Sh = RegShape<H, NW, RPW>;
    if constexpr (Sh::TY < 2) {
        return fail(-1, "workgroup shape %dx%d has no interior rows for n_half=%d", NW, RPW, H);
    } else {
        const size_t lds = sweep_lds_bytes<H, NW, RPW>(TF_KIND_SOLVER, 0);
        if (lds + static_table_bytes(TF_KIND_SOLVER) > 160 * 1024)
            return fail(-1, "sweep shape %dx%d with n_half=%d needs %zu B of LDS", NW, RPW, H, lds);
        const void* const fq[2] = {(const void*)k_sweep_batch_queue<H, NW, RPW, 0>,
                                   (const void*)k_sweep_batch_queue<H, NW, RPW, 1>};
        const void* const fb[2] = {(const void*)k_sweep_batch<H, NW, RPW, 0>,
                                   (const void*)k_sweep_batch<H, NW, RPW, 1>};
        int grid = ((a.n_tiles + 7) / 8) * 8;
        if (a.queue) {
            // the queue form: at most the resident workgroups (the fewer of
            // the two colours' kernels)
            static int slots_cache[2][kMaxDevices] = {};
            const int slots = std::min(resident_workgroups(fq[0], NW * 64, lds, slots_cache[0]),
                                       resident_workgroups(fq[1], NW * 64, lds, slots_cache[1]));
            grid = slots > 0 ? std::min(a.n_tiles, slots) : std::min(a.n_tiles, 1024);
        }
        static bool attr_set = false;
        if (!attr_set && lds > 64 * 1024) {
            for (const void* f : {fq[0], fq[1], fb[0], fb[1]})
                if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)(160 * 1024 - static_table_bytes(TF_KIND_SOLVER))) != hipSuccess)
                    (void)hipGetLastError();
            attr_set = true;
        }
        const bool aa = ((a.colour0 & 1) ^ 1) != 0;
        if (a.queue && aa)
            hipLaunchKernelGGL((k_sweep_batch_queue<H, NW, RPW, 1>), dim3(grid), dim3(NW * 64), lds, st, a);
        else if (a.queue)
            hipLaunchKernelGGL((k_sweep_batch_queue<H, NW, RPW, 0>), dim3(grid), dim3(NW * 64), lds, st, a);
        else if (aa)
            hipLaunchKernelGGL((k_sweep_batch<H, NW, RPW, 1>), dim3(grid), dim3(NW * 64), lds, st, a);
        else
            hipLaunchKernelGGL((k_sweep_batch<H, NW, RPW, 0>), dim3(grid), dim3(NW * 64), lds, st, a);
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

// ---------------------------------------------------------------------------
// per-kind entry points (one translation unit each)
// ---------------------------------------------------------------------------
// Launch (or, with a.resolve_only, only size) the sweep of one kind.
int sweep_dispatch_solver(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st);
int sweep_dispatch_multiphase(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st);
int sweep_dispatch_mpfaces(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st);
int sweep_dispatch_field(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st);
int sweep_dispatch_anisotropic(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st);
int sweep_dispatch_electrode(int H, SweepArgs& a, int B, int nw, int rpw, hipStream_t st);
int sweep_dispatch_batch(int H, BatchArgs& a, int nw, int rpw, hipStream_t st);
