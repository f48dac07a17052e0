/*
 * tauhip.h — C-ABI of the MI355X (gfx950) steady-state diffusion solve behind
 * taufactor_amd.Solver / taufactor_amd.MultiPhaseSolver.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (0 = ok, <0 = library error, >0 = hipError_t) and never throws.
 * The message of the last failure on the calling thread is tf_last_error().
 * Device pointers are caller-owned (PyTorch-ROCm tensors); the library does
 * not allocate device memory (its one device-resident array is the static
 * scratch of tf_layer_sums_noalloc). `stream` is a hipStream_t passed as void*; all
 * device work is enqueued asynchronously on it.
 *
 * Volume layout (shared by every call):
 *   field  float32 [B][X+2][Y+2][Z+2], z fastest.  Padded exactly like the
 *          reference's SORSolver.field: x ghosts hold 2*bc (-1 at x=0, +1 at
 *          x=X+1), y/z ghosts hold 0.
 *   codes  uint8   [B][X+2][Y+2][Z+2], same indexing as field.
 *          kind TF_KIND_SOLVER:     code = number of conductive neighbours
 *                                   (x ghost counts 2), 0 = factor inf
 *                                   (non-conductive voxel, or conductive
 *                                   voxel with no conductive neighbour).
 *          kind TF_KIND_MULTIPHASE: code = phase class index into the
 *                                   n_classes x n_classes face-conductance
 *                                   table; x ghosts copy the adjacent voxel's
 *                                   class, y/z ghosts hold class 0, which is
 *                                   always a zero-diffusivity class.
 *   Interior voxel (x,y,z) (0-based) has colour (x+y+z)&1; reference
 *   iteration k updates colour k&1 (colour 0 first).
 */
#ifndef TAUHIP_H
#define TAUHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TF_ABI_VERSION 12

#define TF_KIND_SOLVER 0
#define TF_KIND_MULTIPHASE 1
/* per-voxel diffusivity: `codes` is a float32 [B][X+2][Y+2][Z+2] array of
 * voxel diffusivities (x ghosts copy the adjacent voxel, y/z ghosts 0);
 * faces use the MultiPhase harmonic mean of the two voxels. */
#define TF_KIND_FIELD 2
/* voxel-spacing weights (AnisotropicSolver): y and z neighbours weighted by
 * Ky = (dx/dy)^2 and Kz = (dx/dz)^2.  `codes` as TF_KIND_SOLVER but set up
 * with TF_SETUP_ANISOTROPIC (code = 1 + 48*xm + 16*xp + 8*ym + 4*yp + 2*zm + zp
 * from the neighbour states, x states 0/1/2, 0 = inactive); `hm_lut` holds
 * TF_ANISO_TABLE floats: the prefactor nn of codes 0..144 (inf for 0 and for
 * a zero sum), then Ky and Kz rounded to float32. */
#define TF_KIND_ANISOTROPIC 3
#define TF_ANISO_TABLE 147
/* transport-reaction electrode (ElectrodeSolver): `codes` from
 * tf_setup_electrode (code = 1 + 7*conductive neighbours + reactive
 * neighbours, 0 = not conductive); `hm_lut` holds TF_ELECTRODE_TABLE floats
 * per batch element: the prefactor cond + k0*reac of each code (inf for 0 and
 * for a zero sum). */
#define TF_KIND_ELECTRODE 4
#define TF_ELECTRODE_TABLE 64
/* MultiPhase stencil read through per-voxel face codes (tf_multiphase_face_codes):
 * `codes` is a uint32 array in the field layout [B][X+2][Y+2][Z+2] with
 * TF_FIELD_SLACK_BYTES of slack before and after it; bits 0-5, 6-11 and 12-17
 * hold the byte offsets (4 x table index) of the voxel's x+, y+ and z+ face
 * conductances in the face table, bits 18-31 the byte offset of its
 * (1/f, f) prefactor pair in the pair table (1/f correctly rounded).
 * `hm_lut` holds TF_MPF_FACE_SLOTS face conductances (4 x 4 face classes)
 * followed by the pairs, pair 0 = (0, 0); n_classes is the table's length in
 * floats.  Prefactors must lie in [2^-20, 2^20].  Same iterates as
 * TF_KIND_MULTIPHASE. */
#define TF_KIND_MULTIPHASE_FACES 5
#define TF_MPF_MAX_CLASSES 4
#define TF_MPF_FACE_SLOTS 16
#define TF_MPF_MAX_FACTORS 2047

/* Workgroup shape of the temporally blocked sweep.  A workgroup streams a
 * tile of (waves*rows_per_wave - 2*n_half) rows in y by TZ voxels in z
 * (TZ = 120 for n_half <= 4, 112 for 5 and 6) along tx planes of x; one
 * wavefront per row.  Compiled shapes (waves x rows_per_wave): 8x2, 12x2
 * and 16x2 at every n_half, 4x4 and 8x4 at n_half <= 4 (12x2 and 8x4 for
 * the Solver kind and batches only).  Zero fields select the defaults: 16x2 for one-byte
 * codes with n_half <= 4 and Y >= 64, else 8x2 (8x2 for batches).  tx 0
 * (tf_sweep) lets the launch choose the x-tile length: the one whose tiles
 * fill the device's resident workgroup slots in the fewest, shortest
 * rounds; batches default to tx 64. */
typedef struct tf_tile {
    int waves, rows_per_wave, tx;
} tf_tile;

/* Readable bytes tf_sweep and tf_batch_sweep require before the first and
 * after the last element of every field buffer (and of the per-voxel
 * diffusivity array of TF_KIND_FIELD); they must hold finite float32
 * values (zeros).  The sweep reads the field in 16-byte quads that start up
 * to 3 voxels before z = 0. */
#define TF_FIELD_SLACK_BYTES 16

/* ABI version of the loaded library (TF_ABI_VERSION). No reference
 * counterpart: lets the Python loader refuse a stale build. */
int tf_abi_version(void);

/* Thread-local message for the last non-zero status. */
const char* tf_last_error(void);

/* set-up flags: y and z periodic (PeriodicSolver / PeriodicMultiPhaseSolver) */
#define TF_SETUP_PERIODIC_YZ 1
/* set-up flag of tf_setup_solver: codes of TF_KIND_ANISOTROPIC */
#define TF_SETUP_ANISOTROPIC 2

/* Set-up of a two-phase volume.  Replaces SORSolver.__init__ field/prefactor
 * construction: ThroughTransportSolver.init_field (taufactor/taufactor.py:282-291),
 * SORSolver._pad (:227-238), Solver.init_conductive_neighbours (:402-410)
 * and vol_x = mean(mask,(2,3)) (:42).
 *   mask    device uint8 [B][X][Y][Z], 1 = conductive, 0 = not
 *   xinit   device float [X], the linear initial profile (linspace of
 *           -0.5+1/(2X) .. 0.5-1/(2X) in float32)
 *   codes   device uint8 padded, written (ghosts 0)
 *   field_a, field_b  device float padded, both written with the initial field
 *   layer_count device uint32 [B][X], written with the number of
 *           conductive voxels per x layer.
 *   flags   TF_SETUP_PERIODIC_YZ: neighbour counts wrap round in y and z
 *           (PeriodicSolver.init_conductive_neighbours,
 *           taufactor/taufactor.py:493-499)
 *           TF_SETUP_ANISOTROPIC: codes of TF_KIND_ANISOTROPIC, the state of
 *           the six neighbours (AnisotropicSolver.init_conductive_neighbours,
 *           taufactor/taufactor.py:459-471); not combined with periodic */
int tf_setup_solver(const uint8_t* mask, const float* xinit, uint8_t* codes,
                    float* field_a, float* field_b, uint32_t* layer_count,
                    int B, int X, int Y, int Z, int flags, void* stream);

/* Set-up of a labelled volume for the face-conductance stencil.  Replaces
 * MultiPhaseSolver.return_mask / init_conductive_neighbours
 * (taufactor/taufactor.py:571-604) and the shared field set-up above.
 *   cls      device uint8 [B][X][Y][Z], class index per voxel (< n_classes)
 *   cls_conductive device uint8 [n_classes], 1 if the class diffusivity > 0
 *   flags    TF_SETUP_PERIODIC_YZ: y/z ghost codes copy the opposite
 *            interior layer (PeriodicMultiPhaseSolver.init_conductive_neighbours,
 *            taufactor/taufactor.py:626-650)
 *   Other arguments as tf_setup_solver.                                    */
int tf_setup_multiphase(const uint8_t* cls, const uint8_t* cls_conductive,
                        int n_classes, const float* xinit, uint8_t* codes,
                        float* field_a, float* field_b, uint32_t* layer_count,
                        int B, int X, int Y, int Z, int flags, void* stream);

/* Set-up of a transport-reaction electrode volume.  Replaces
 * ElectrodeSolver.return_mask / init_field / init_conductive_neighbours /
 * init_reactive_neighbours (taufactor/electrode.py:36-70; the periodic forms
 * :133-151) and the reactive-surface sum of SORSolver.__init__
 * (taufactor/taufactor.py:47-48).
 *   cls      device uint8 [B][X][Y][Z]: 1 conductive, 2 reactive, 0 inert
 *   xinit    device float [X], the initial profile cosh(1 - (x+0.5)/X)/cosh(1)
 *            rounded to float32
 *   codes    device uint8 padded: 1 + 7*cond + reac for conductive voxels
 *            (cond counts the x = 0 ghost twice and the x = X+1 ghost not at
 *            all; reac counts reactive neighbours, ghosts excluded), 0 else
 *   field_a, field_b  padded initial field; x ghosts 2 (x = 0) and 0 (x = X+1)
 *   layer_count     device uint32 [B][X], conductive voxels per layer
 *   reactive_count  device uint32 [B][X], sum of reac over the layer's
 *            conductive voxels
 *   flags    TF_SETUP_PERIODIC_YZ: neighbour counts wrap round in y and z */
int tf_setup_electrode(const uint8_t* cls, const float* xinit, uint8_t* codes,
                       float* field_a, float* field_b, uint32_t* layer_count,
                       uint32_t* reactive_count, int B, int X, int Y, int Z, int flags,
                       void* stream);

/* Periodic boundaries: every y/z ghost of every x plane takes the value of
 * the opposite interior layer (corners both wraps).  Replaces
 * apply_boundary_conditions of PeriodicSolver / PeriodicMultiPhaseSolver
 * (taufactor/taufactor.py:501-505, 652-656), called before every
 * iteration.  With odd Y or Z a periodic solve alternates tf_wrap_ghosts on
 * the source buffer and a one-half-sweep tf_sweep; with Y and Z even it
 * sweeps with TF_SWEEP_PERIODIC_YZ instead (no ghost is read).            */
int tf_wrap_ghosts(float* field, int B, int X, int Y, int Z, void* stream);

/* The same copy from the interior of src into the ghosts of dst (src ==
 * dst is tf_wrap_ghosts).  After a TF_SWEEP_PERIODIC_YZ sweep whose last
 * launch is one half-sweep src -> dst, it leaves in dst the ghosts the
 * reference holds after that iteration (the wrap made before it).  No
 * reference counterpart as an entry point.                                */
int tf_wrap_ghosts_from(const float* src, float* dst, int B, int X, int Y, int Z, void* stream);

/* Size in bytes of the sweep-code layout of one volume: (X+2)*(Y+2) rows of
 * tf_sweep_code_row(Z) = roundup(Z+5, 4) bytes; byte z+3 of row (x, y)
 * holds the code of padded voxel (x, y, z) for z in [0, Z+1], every other
 * byte is 0.  Rows are 4-byte aligned so that the sweep reads four codes
 * per lane with one aligned load.  No reference counterpart (the reference
 * keeps a float32 prefactor array, taufactor/taufactor.py:402-410).      */
long long tf_sweep_codes_bytes(int X, int Y, int Z);

/* Repack codes (layout above, B volumes) into the sweep-code layout
 * (B consecutive volumes of tf_sweep_codes_bytes each; 4-byte aligned) for
 * tf_sweep / tf_batch_sweep of the given kind: TF_KIND_SOLVER stores each
 * count times 8 (the byte offset of its (1/d, d) table entry, at most 64),
 * the other kinds the code itself.  */
int tf_pack_sweep_codes(const uint8_t* codes, uint8_t* sweep_codes, int kind, int B, int X, int Y,
                        int Z, void* stream);

/* Face codes of a labelled volume for tf_sweep with TF_KIND_MULTIPHASE_FACES
 * (layout at TF_KIND_MULTIPHASE_FACES).  No reference counterpart as an
 * entry point: it precomputes what MultiPhaseSolver.init_conductive_neighbours
 * (taufactor/taufactor.py:585-604) derives per voxel - the face
 * conductances hm(D_a, D_b) and their sum, the prefactor - as table offsets.
 *   codes       the padded class codes of tf_setup_multiphase
 *   face_class  device uint8 [n_classes]: the face class (< TF_MPF_MAX_CLASSES,
 *               one per distinct diffusivity, 0 for diffusivity 0) of each
 *               class; the face table is 4 x 4, face (a, b) at slot 4a + b
 *   key_table   device uint16 [4^7 * 4]: the byte offset in the pair table of
 *               the prefactor of every neighbourhood, indexed by
 *               (((((((c*4 + xm)*4 + xp)*4 + ym)*4 + yp)*4 + zm)*4 + zp) * 4 + b)
 *               with c the voxel's face class, xm..zp its neighbours' and
 *               b = (x == 0) + 2 (x == X-1) the Dirichlet-doubling flags
 *               (interior 0-based x); offset 0 is the pair (0, 0) of a zero
 *               (infinite) prefactor
 *   face_codes  device uint32 padded, written (ghost voxels get pair 0)     */
int tf_multiphase_face_codes(const uint8_t* codes, const uint8_t* face_class, int n_classes,
                             const uint16_t* key_table, uint32_t* face_codes, int B, int X, int Y,
                             int Z, void* stream);

/* n_half over-relaxed red-black half-sweeps (reference iterations), the
 * first one on colour first_colour, fused in one HBM pass: reads src,
 * writes every interior voxel of dst (ghosts of dst are left untouched and
 * must already hold the ghost values).  src and dst must not alias.
 * Replaces the loop body of SORSolver.solve (taufactor/taufactor.py:175-182)
 * with SORSolver.sum_weighted_neighbours (:95-103) or
 * MultiPhaseSolver.sum_weighted_neighbours (:606-613).  Arithmetic is the
 * reference's float32 operation order, without contraction, so the result
 * equals n_half reference iterations value for value.
 *   src, dst field buffers with TF_FIELD_SLACK_BYTES of slack
 *   codes    TF_KIND_SOLVER / TF_KIND_MULTIPHASE: the sweep-code layout
 *            (tf_pack_sweep_codes); TF_KIND_FIELD: the face conductances
 *            of tf_field_faces (three sections, TF_FIELD_SLACK_BYTES of
 *            slack before the first and after the last)
 *   hm_lut   device float [n_classes*n_classes] face conductances
 *            (TF_KIND_MULTIPHASE), ignored for TF_KIND_SOLVER
 *   omega    the over-relaxation factor rounded to float32
 *   n_half   1..TF_MAX_FUSED_HALF_SWEEPS
 *   tile     NULL for the default tile shape
 * Sizes: a padded volume holds fewer than 2^31 voxels (every entry point);
 * the sweep addresses each tile from its first plane, so a volume may exceed
 * 2 GiB per field as long as (tx + 2 n_half) planes stay below 2 GiB (the x
 * tile is capped to that).                                                  */
#define TF_MAX_FUSED_HALF_SWEEPS 6
/* Flag OR-ed into the kind of tf_sweep / tf_sweep_launch_config: periodic y
 * and z.  Every y/z neighbour across the boundary is read from the opposite
 * interior layer (the value tf_wrap_ghosts would copy into the ghost before
 * each iteration, taufactor/taufactor.py:501-505, 652-656), so n_half fused
 * half-sweeps equal n_half periodic reference iterations; the ghosts of src
 * are not read and those of dst not written.  TF_KIND_SOLVER,
 * TF_KIND_MULTIPHASE and TF_KIND_ELECTRODE with codes set up with
 * TF_SETUP_PERIODIC_YZ; Y and Z must be even (the red/black colouring then
 * repeats across the boundary).                                            */
#define TF_SWEEP_PERIODIC_YZ 0x100
int tf_sweep(const float* src, float* dst, const void* codes,
             const float* hm_lut, int n_classes, int kind,
             int B, int X, int Y, int Z, float omega, int first_colour,
             int n_half, const tf_tile* tile, void* stream);

/* The launch configuration tf_sweep would use for these arguments: the
 * workgroup shape and the x-tile length after defaults (tx 0 resolved on
 * the current device), and the number of tiles.  Launches nothing; no
 * counterpart in the reference (the launch shape is this library's). */
int tf_sweep_launch_config(int kind, int n_classes, int B, int X, int Y, int Z,
                           int n_half, const tf_tile* tile, tf_tile* out, int* n_tiles);

/* One half-sweep of colour `colour` in place, one thread per voxel.  The
 * unblocked form of tf_sweep, kept as the simplest statement of the update
 * for tests; same arithmetic.                                            */
int tf_sweep_half_inplace(float* field, const uint8_t* codes,
                          const float* hm_lut, int n_classes, int kind,
                          int B, int X, int Y, int Z, float omega, int colour,
                          void* stream);

/* Per-layer float64 sums for the convergence check.  Replaces
 * Solver.vertical_flux (taufactor/taufactor.py:412-419) /
 * MultiPhaseSolver.vertical_flux (:615-620) and the two torch.mean(...,(2,3))
 * reductions of ThroughTransportSolver.compute_metrics (:296, :307).
 *   flux_sum device double [B][X-1]: sum over (y,z) of the float32 face flux
 *            (c[x+1]-c[x], zeroed when either voxel has code 0, for
 *            TF_KIND_SOLVER; hm*(c[x+1]-c[x]) for TF_KIND_MULTIPHASE;
 *            zeroed where the factor exceeds 8 for TF_KIND_ANISOTROPIC).
 *            TF_KIND_ELECTRODE: [B][X], ElectrodeSolver.compute_metrics'
 *            fluxes (taufactor/electrode.py:80-83): c[x-1]-c[x] for x >= 1,
 *            (1 - c[0])/0.5 for x = 0, zeroed where either value is 0
 *   conc_sum device double [B][X]:   sum over (y,z) of c
 * The summation order is fixed (chunks of 2048 positions y*Z + z, each
 * summed in a fixed tree, then added in chunk order), so repeated calls,
 * tf_layer_sums_ws and tf_batch_layer_sums give identical bits.          */
int tf_layer_sums(const float* field, const void* codes,
                  const float* hm_lut, int n_classes, int kind,
                  int B, int X, int Y, int Z,
                  double* flux_sum, double* conc_sum, void* stream);

/* Bytes of device workspace tf_layer_sums_ws needs: 8*B*X*ceil(Y*Z/2048)
 * doubles (the partial sums of every chunk and wave); negative for invalid
 * sizes.                                                                 */
long long tf_layer_sums_workspace_bytes(int B, int X, int Y, int Z);

/* tf_layer_sums with the same result bits, faster on large volumes: each
 * block walks a slab of x layers of one chunk and reads every plane once
 * (tf_layer_sums reads the field twice per face), writing the chunk partials
 * to `workspace`; a second kernel adds them in chunk order.  No reference
 * counterpart beyond tf_layer_sums'.
 *   workspace device memory of at least tf_layer_sums_workspace_bytes,
 *             8-byte aligned; overwritten                                  */
int tf_layer_sums_ws(const float* field, const void* codes,
                     const float* hm_lut, int n_classes, int kind,
                     int B, int X, int Y, int Z,
                     double* flux_sum, double* conc_sum, void* workspace,
                     long long workspace_bytes, void* stream);

/* tf_layer_sums_ws without a workspace argument and without allocating:
 * the partial sums go to a static device array of TF_NOALLOC_SCRATCH_BYTES
 * that the library holds on every device it runs on until the process ends
 * (its address is looked up once per device, on the first call there).  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same result bits as tf_layer_sums_ws.
 * Sizes whose tf_layer_sums_workspace_bytes exceeds TF_NOALLOC_SCRATCH_BYTES
 * are rejected (-1).  The calling thread's current device must be the device
 * of `stream`.  Every call on one device shares that device's array: calls
 * that overlap on one device (different streams, or host threads) race; calls
 * on one stream are ordered and safe.                                       */
#define TF_NOALLOC_SCRATCH_BYTES (16LL << 20)
int tf_layer_sums_noalloc(const float* field, const void* codes,
                          const float* hm_lut, int n_classes, int kind,
                          int B, int X, int Y, int Z,
                          double* flux_sum, double* conc_sum, void* stream);

/* Set-up of a per-voxel diffusivity volume (TF_KIND_FIELD).  No reference
 * counterpart: it is MultiPhaseSolver (taufactor/taufactor.py:585-604) with a
 * diffusivity map given per voxel instead of per label.
 *   k     device float [B][X][Y][Z], voxel diffusivity >= 0 (0 = inert)
 *   kpad  device float padded, written (x ghosts copy, y/z ghosts 0)
 *   other arguments as tf_setup_solver; conductive means k > 0.          */
int tf_setup_field(const float* k, const float* xinit, float* kpad,
                   float* field_a, float* field_b, uint32_t* layer_count,
                   int B, int X, int Y, int Z, void* stream);

/* Face conductances of the per-voxel field for tf_sweep (TF_KIND_FIELD):
 * faces = 3 consecutive sections of B*(X+2)*(Y+2)*(Z+2) floats (gx, gy, gz),
 * each face stored at the padded voxel on its minus side:
 * gx[o] = hm(kpad[o], kpad[o + x]) and likewise for y and z, 0 where that
 * neighbour is outside the padded volume; hm(a, b) = ((2a) b)/(a + b), 0 for
 * a + b = 0.  Computing each face once replaces the six harmonic means of
 * every update (MultiPhaseSolver._harmonic_mean, taufactor/taufactor.py:
 * 577-583, per voxel pair) with the same float32 values.
 *   kpad   the padded diffusivity of tf_setup_field                         */
int tf_field_faces(const float* kpad, float* faces, int B, int X, int Y, int Z, void* stream);

/* Gradient of D_rel with respect to every voxel diffusivity, from a
 * converged field (no reference counterpart: the reference solves under
 * torch.no_grad, taufactor/taufactor.py:172).  D_rel = (X/(Y*Z)) * W, W the
 * dissipation sum_faces g (dc)^2 + sum_x-faces 2k (c - bc)^2, which the field
 * minimises, so dD_rel/dk_v is the explicit derivative of W (one pass, no
 * adjoint solve).  Inert voxels (k = 0) get the one-sided derivative.
 *   grad  device float [B][X][Y][Z]                                        */
int tf_grad_field(const float* field, const float* kpad, int B, int X, int Y, int Z,
                  float* grad, void* stream);

/* The same derivative summed per class of a labelled volume, layer by layer:
 *   d_cls       device float [n_classes], diffusivity of each class (class 0 = 0)
 *   grad_layers device double [B][X][n_classes]; summing over x gives
 *               dD_rel/dD_class (fixed summation order)                    */
int tf_grad_classes(const float* field, const uint8_t* codes, const float* d_cls,
                    int n_classes, int B, int X, int Y, int Z, double* grad_layers,
                    void* stream);

/* ---- uneven batch (no reference counterpart: the reference batches only
 * volumes of one shape and stops when all have converged,
 * taufactor/taufactor.py:30-31,143; here every volume keeps the solo
 * Solver semantics: its own omega, check points, old_tau and stopping
 * iteration) ---- */

/* One two-phase volume of a batch.  Offsets are element offsets into the
 * packed buffers; layout of each volume as above (B = 1). */
typedef struct tf_volume {
    long long offset;   /* first element of the padded volume in field / codes */
    long long code_offset;  /* first byte of the volume in the packed sweep codes */
    int X, Y, Z;
    float omega;        /* over-relaxation factor rounded to float32 */
    int sum_offset;     /* first entry in the packed layer-sum arrays (X per volume) */
    int tile_first;     /* first tile of this volume in the batch tile list */
    int spans;          /* 1 if a 6-connected conductive cluster joins both x faces */
    int reserved;
    double D_mean;      /* tau = D_mean / D_rel (float32 mean porosity, widened) */
} tf_volume;

/* Device-resident convergence state of one volume; initialise with
 * active = 1 and everything else 0. */
typedef struct tf_volume_state {
    int active;         /* still iterating */
    int converged;      /* stopped by the convergence rule (not the limit) */
    int iters;          /* reference iterations done at the last check */
    int cur;            /* 0: field_a holds the final iterate, 1: field_b */
    double tau, old_tau, D_rel, rel_err;
} tf_volume_state;

/* Number of workgroup tiles tf_batch_sweep uses for one volume (tile
 * defaults of batches: 8x2, tx 64). */
int tf_sweep_tile_count(int X, int Y, int Z, int n_half, const tf_tile* tile, int* count);

/* n_half half-sweeps on every active volume of a batch, one launch.  All
 * active volumes read field_a (src_buffer 0) or field_b (1) and write the
 * other.  tile_vol[t] = volume of tile t (tiles of volume v are
 * tile_first .. tile_first + tf_sweep_tile_count - 1).  sweep_codes holds
 * each volume's sweep-code layout at its code_offset; field_a and field_b
 * need TF_FIELD_SLACK_BYTES of slack.  Same arithmetic as tf_sweep.      */
int tf_batch_sweep(float* field_a, float* field_b, const uint8_t* sweep_codes,
                   const tf_volume* vols, const int* tile_vol, int n_tiles,
                   const tf_volume_state* state, int src_buffer, int first_colour,
                   int n_half, const tf_tile* tile, void* stream);

/* Live tile list of a batch pass depth for tf_batch_sweep_queue: writes
 * live_tiles[0 .. n) = the tiles (indices into tile_vol) of the volumes
 * still active, in tile order, queue[0] = n, and clears the pass counters
 * queue[1 .. n_counters].  Run after every convergence check (and once
 * before the first pass); live_tiles holds n_tiles ints, queue
 * 1 + n_counters.  No reference counterpart. */
int tf_batch_compact(const int* tile_vol, int n_tiles, const tf_volume_state* state,
                     int* live_tiles, int* queue, int n_counters, void* stream);

/* tf_batch_sweep as a work queue: a grid of at most the device's resident
 * workgroups takes the listed live tiles one at a time from the counter
 * queue[1 + pass] (each pass of a period its own counter), so tiles of
 * stopped volumes are not dispatched.  Same results as tf_batch_sweep. */
int tf_batch_sweep_queue(float* field_a, float* field_b, const uint8_t* sweep_codes,
                         const tf_volume* vols, const int* tile_vol, const int* live_tiles,
                         int* queue, int pass, int n_tiles, const tf_volume_state* state,
                         int src_buffer, int first_colour, int n_half, const tf_tile* tile,
                         void* stream);

/* Per-layer float64 sums of every active volume (layer_vol[l] = volume of
 * packed layer l; layer x of volume v is l = sum_offset + x).  flux_sum[l]
 * is 0 for the last layer of a volume.  Same order as tf_layer_sums.      */
int tf_batch_layer_sums(const float* field_a, const float* field_b, const uint8_t* codes,
                        const tf_volume* vols, const int* layer_vol, int n_layers,
                        const tf_volume_state* state, int src_buffer,
                        double* flux_sum, double* conc_sum, void* stream);

/* The convergence check of every active volume, on the device: adds
 * iters_per_check to iters, evaluates rel. error and tau from the layer
 * sums and applies the reference's stopping rules
 * (taufactor/taufactor.py:109-153, 293-331); a volume stops when it
 * converges or reaches iter_limit, recording which buffer holds its field. */
int tf_batch_check(const tf_volume* vols, int nvol, const double* flux_sum,
                   tf_volume_state* state, int src_buffer, int iters_per_check,
                   int iter_limit, double conv_crit, void* stream);

/* Host-side test: does any 6-connected cluster of `mask` (host uint8
 * [X][Y][Z]) touch both x faces?  Returns 1 or 0, or <0 on error.
 * Replaces the spanning test of extract_through_feature
 * (taufactor/metrics/connectivity.py:138-213, find_spanning_labels :83-115)
 * as called from compute_metrics (taufactor/taufactor.py:318-327).        */
int tf_spans_x(const uint8_t* mask, int X, int Y, int Z);

/* Device-to-device copy of `bytes` bytes (a multiple of 16; both pointers
 * 16-byte aligned) by a streaming kernel: nontemporal 16-byte loads and
 * stores, four per lane in flight, one pass over the buffer.  No reference counterpart: it
 * calibrates the HBM traffic counters and gives the device's copy rate
 * against which the sweep's measured traffic is read.                    */
int tf_hbm_copy(const void* src, void* dst, long long bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TAUHIP_H */
