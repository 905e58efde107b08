/*
 * aerolbm.h — C ABI of the MI355X (gfx950) D3Q19 wind-tunnel solver.
 *
 * This is the drop-in boundary for the solver path of necrolingus/aerosim: the reference's TypeScript
 * host drives `LBMEngine` (src/solver/lbmEngine.ts:39-710) and `voxelizeMeshToLattice`
 * (src/geometry/voxelizer.ts:86); a host binding (N-API addon, ctypes) calls these entry points instead.
 * Every call returns 0 on success or a negative AERO_E* code; nothing throws across the ABI, and the
 * message of the last failure is available from aero_last_error(). Buffers passed in are caller-owned
 * host memory; the context owns all device memory. Calls on one context must be serialized (the
 * reference serializes with its `isBusy` flag, lbmEngine.ts:60,420,478).
 *
 * Lattice conventions (lbmConstants.ts:10-66, lbmCpuReference.ts:169-171):
 *   cell index idx = ix + nx*(iy + ny*iz), x fastest; +x is the wind direction, y is vertical
 *   (floor at iy = 0), z is spanwise. Populations are reported structure-of-arrays, f[d*N + idx],
 *   with the reference's D3Q19 velocity order (0 rest, 1..6 faces, 7..18 edges in +/- pairs).
 *   Cell flags: 0 fluid, 1 solid, 2 boundary (fluid with a solid D3Q19 neighbour).
 */
#ifndef AEROLBM_H
#define AEROLBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AERO_OK 0
#define AERO_E_ARG (-1)     /* invalid argument or size */
#define AERO_E_HIP (-2)     /* HIP runtime error (message holds the HIP error string) */
#define AERO_E_STATE (-3)   /* call not valid in the current state (e.g. no geometry yet) */
#define AERO_E_NOMEM (-4)   /* device or host allocation failed */
#define AERO_E_FORMAT (-5)  /* input bytes are not in the expected file format */

#define AERO_STORE_FP32 0       /* populations stored as float (152 B per cell update) */
#define AERO_STORE_FP16 1       /* populations stored as shifted half (f - w)*2^k (76 B per cell update) */

#define AERO_PROFILE_UNIFORM 0  /* uniform inflow; with ground_noslip the reference's power law */
#define AERO_PROFILE_LOGLAW 2   /* u_in * ln((iy+0.5)/z0) / ln(z_ref/z0) */

#define AERO_CLOSURE_SMAGORINSKY 0
#define AERO_CLOSURE_MLP 1

typedef struct aero_ctx aero_ctx;

/* Grid and storage (the reference sizes its grid in gpuContext.ts:17-58). */
typedef struct {
  int32_t nx, ny, nz;
  int32_t storage; /* AERO_STORE_* */
  int32_t device;  /* HIP device ordinal */
} aero_grid;

/* TunnelDimensions (src/core/types.ts:73-77): length along x, width along z, height along y. */
typedef struct {
  double length, width, height;
} aero_tunnel;

/* Solver parameters: LBMSolverConfig's numerical fields (lbmEngine.ts:17-32) plus the extensions of
 * this solver (inflow profile, ramp, closure, FP16 scale). */
typedef struct {
  double u_inlet;        /* target inlet velocity, lattice units */
  double nu;             /* lattice kinematic viscosity */
  double smagorinsky_cs; /* Smagorinsky constant */
  int32_t ground_noslip; /* 1: no-slip floor and the reference's ground power-law inflow */
  int32_t inflow_profile;/* AERO_PROFILE_* */
  double z0, z_ref;      /* log-law roughness length and reference height, in cells */
  int32_t ramp_steps;    /* 0: constant u_inlet; n: u = u_inlet*sin^2(pi/2*(s+1)/n) for step s < n */
  int32_t closure;       /* AERO_CLOSURE_* */
  int32_t fp16_shift;    /* k of the FP16 codec (storage = AERO_STORE_FP16): populations are stored as
                            half((f - w_k) * 2^k). 0 selects 15; otherwise 10..16. A power-of-two scale
                            changes no significand bit, and below FP16's normal range f - w_k is a
                            multiple of FP32 f's ulp that half subnormals hold exactly, so every
                            allowed k stores the same bits (overflow needs |f - w_k| >= 1) */
  int32_t reserved;
  double u_init;         /* inlet velocity used by aero_reset; < 0 selects u_inlet (the reference
                            test protocol starts at 0.001 and ramps, testRunner.ts:76,89-96) */
} aero_params;

/* Reference quantities for the coefficients (computeNormalizedForceCoefficients inputs,
 * forceIntegrator.ts:240-275). */
typedef struct {
  double frontal_area;       /* A_ref, m^2 */
  double char_length;        /* L_ref, m */
  double tunnel_width, tunnel_height; /* m, for the Maskell blockage ratio */
  double dx;                 /* cell size h, m */
  double u_ref;              /* lattice velocity used for q_lat = 0.5*u^2 */
} aero_refs;

/* Momentum-exchange force and torque about the domain centre (forces.wgsl:95-102,
 * forceIntegrator.ts:106-234), accumulated in FP64, and the coefficients of
 * forceIntegrator.ts:240-315. */
typedef struct {
  double fx, fy, fz, mx, my, mz;
  int64_t cut_face_links;
  double cd_raw, cd, cl, cm, xcp;
} aero_forces;

/* ---- lifecycle -------------------------------------------------------------------------------- */
/* Replaces `new LBMEngine(device)` (lbmEngine.ts:92-97). */
int aero_create(const aero_grid* grid, aero_ctx** out);
int aero_destroy(aero_ctx* ctx);
/* Message of the last failed call on ctx (or of the last failed aero_create when ctx is NULL). */
const char* aero_last_error(const aero_ctx* ctx);

/* ---- geometry --------------------------------------------------------------------------------- */
/* Loads a VoxelGridResult (voxelizer.ts:19-59) produced on the host: flags[N], bmap[N], blist[nb],
 * q[nb*18] (q[b*18 + d-1], -1 for links without a wall). Replaces the geometry upload of
 * LBMEngine.setVoxelGrid (lbmEngine.ts:142-287). */
int aero_set_voxels(aero_ctx* ctx, const uint8_t* flags, const int32_t* bmap, const int32_t* blist,
                    const float* q, int32_t nb);
/* Voxelizes a triangle soup on the GPU (replaces voxelizeMeshToLattice, voxelizer.ts:86-281).
 * tri_pos: ntri*9 floats (A, B, C per triangle, a BufferGeometry position attribute); world: 16
 * doubles, column-major like THREE.Matrix4.elements, or NULL for identity. Cell size is
 * tunnel->height / ny and the lattice is centred on the origin (voxelizer.ts:101-109). */
int aero_set_mesh(aero_ctx* ctx, const float* tri_pos, int32_t ntri, const double* world,
                  const aero_tunnel* tunnel);
/* Reads back the current VoxelGridResult. counts[0..2] = solid, boundary, fluid cell counts.
 * Any output pointer may be NULL; call with all arrays NULL first to size blist/q/minq. */
int aero_read_voxels(aero_ctx* ctx, uint8_t* flags, int32_t* bmap, int32_t* blist, float* q,
                     float* minq, int32_t* counts);
/* Voxelization time of the last aero_set_mesh in milliseconds (voxelizer.ts:92,259). */
int aero_voxelize_ms(aero_ctx* ctx, double* ms);

/* ---- parameters and state --------------------------------------------------------------------- */
/* Replaces LBMEngine.updateConfig (lbmEngine.ts:289) for the solver fields. */
int aero_set_params(aero_ctx* ctx, const aero_params* p);
/* Equilibrium initialisation (lbmCpuReference.ts:114-143, init.wgsl:24-61) with optional seeded
 * velocity noise of amplitude noise_amp on non-solid cells; step index back to 0. Replaces resetFlow
 * (lbmEngine.ts:677). */
int aero_reset(aero_ctx* ctx, uint64_t seed, double noise_amp);

/* ---- stepping --------------------------------------------------------------------------------- */
/* Advances n steps (collide+stream, domain boundaries, mass anchor: lbmCpuReference.ts:145-456).
 * Stream-ordered; returns once the work is queued. With AERO_STEP_READOUT the last step also
 * writes the macroscopic fields and a population snapshot for aero_read_macro /
 * aero_read_populations; with AERO_STEP_MACROS only the macroscopic fields (the reference's
 * readMacroField of stepAndExtractSlice, lbmEngine.ts:404, which the slice, Q-criterion, probe and
 * mass-statistics readouts read). A readout step runs the same split kernels as a plain step, so
 * AERO_STEP_MACROS adds 16 B per cell of writes, AERO_STEP_READOUT 92 B. */
#define AERO_STEP_READOUT 1
#define AERO_STEP_MACROS 2
int aero_step(aero_ctx* ctx, int32_t n, int32_t flags);
int aero_sync(aero_ctx* ctx);
int aero_step_index(aero_ctx* ctx, int64_t* out);   /* getStepCount, lbmEngine.ts:690 */
/* Sets the step counter (inlet perturbation and ramp depend on it; lbmCpuReference.ts:340), for restarts
 * from populations written with aero_write_populations. */
int aero_set_step_index(aero_ctx* ctx, int64_t step);
/* Times n steps with HIP events on the context's stream. out[0] = total ms, out[1] = mean ms of the
 * stream-collide launches alone (events around them), out[2] = mean ms of the whole step, out[3] = mean ms
 * of the bulk kernel, out[4] = mean ms from the bulk kernel's end until the stream has joined the listed-cell
 * kernels run beside it: their remaining work, if any, plus the join itself (out[3] = out[1], out[4] = 0
 * for steps that run one kernel). out must hold 5 values. */
int aero_time_steps(aero_ctx* ctx, int32_t n, double* out);
/* Cells per thread of the bulk step kernel (1, 2 or 4, dividing nx); the default is 1 for FP32 and 2 for
 * FP16 storage. 0 runs no bulk kernel: every cell goes through the one-cell-per-thread kernel that the split
 * step is checked against. Results do not depend on it. */
int aero_set_cells_per_thread(aero_ctx* ctx, int32_t v);
/* How a plain step splits the grid at the current cells-per-thread: out[0] = cells the bulk kernel updates
 * (plain fluid cells off the domain faces, each reading and writing its 19 populations), out[1] = cells the
 * listed-cell kernels visit, out[2] = cells per thread, out[3] = how many of the out[1] are plain fluid
 * cells (face rows), the rest being boundary cells and cells with non-fluid inlet or outlet neighbours.
 * out must hold 4 values. */
int aero_split_info(aero_ctx* ctx, int64_t* out);

/* ---- readouts --------------------------------------------------------------------------------- */
/* Force readout (forces.wgsl + host sum lbmEngine.ts:627-642 + coefficients forceIntegrator.ts:240).
 * refs may be NULL (coefficients then 0). */
int aero_read_forces(aero_ctx* ctx, const aero_refs* refs, aero_forces* out);
/* The force readout of aero_read_forces with nothing allocated, on the first call or any other: no counterpart in
 * the reference; it is for callers that cannot allocate device memory (code under stream capture, serving loops,
 * harnesses that forbid hipMalloc). Its kernels keep their block partials and sums in one fixed 256 KiB array per
 * GPU (the library's own, for the process's lifetime), so it serves geometries of up to 2^20 boundary cells and
 * fails with AERO_E_ARG beyond. Calls on one GPU share that array: calls that overlap on a GPU race and must be
 * serialized by the caller. Results equal aero_read_forces's bit for bit. */
int aero_read_forces_noalloc(aero_ctx* ctx, const aero_refs* refs, aero_forces* out);
/* Macroscopic fields after the last AERO_STEP_READOUT step: ux, uy, uz, rho (N floats each; NULL skips). */
int aero_read_macro(aero_ctx* ctx, float* ux, float* uy, float* uz, float* rho);
/* Post-step populations (the reference's fSrc after the swap) after the last AERO_STEP_READOUT step. */
int aero_read_populations(aero_ctx* ctx, float* f);
/* Replaces the state by post-step populations f[19*N] (as aero_read_populations returns them), for
 * checkpoint/restart and single-step parity checks. */
int aero_write_populations(aero_ctx* ctx, const float* f);
/* computeFluidMassStats (lbmCpuReference.ts:461-495) over the last readout fields, reduced on the GPU: over the
 * non-solid cells before the first cell (in index order) whose density or velocity is NaN, as the reference's
 * loop stops there. out: total mass, mean rho, max speed, has-NaN flag, counted cells. */
int aero_mass_stats(aero_ctx* ctx, double* out);
/* Number of steps on which the global mass anchor fired since the last reset. */
int aero_anchor_count(aero_ctx* ctx, int64_t* out);

/* ---- many tunnels in one launch ------------------------------------------------------------- */
/* A batch steps n independent contexts (same device and storage; each keeps its own geometry, parameters
 * and readouts through the entry points above) over tile lists shared by all tunnels (bulk cells and each
 * tunnel's listed cells), pairs of steps captured in a HIP graph. A step is either one persistent launch over
 * every tile, split by XCD, drawn from per-XCD work counters and finalized by its last block, or per-kind
 * launches (bulk, lists, finalize). No counterpart in the reference (it runs one grid, lbmEngine.ts:483-527);
 * it serves angle-of-attack and specimen sweeps. The contexts must outlive the batch and be reset together before batched stepping;
 * batched stepping sets every context's cells per thread to the batch's. */
typedef struct aero_batch aero_batch;
int aero_batch_create(aero_ctx* const* tunnels, int32_t n, aero_batch** out);
int aero_batch_destroy(aero_batch* b);
const char* aero_batch_last_error(const aero_batch* b);  /* NULL: why the last aero_batch_create failed */
/* Launch shape of a batched step: threads per block (128/256/512), cells per thread (1/2/4, dividing every
 * tunnel's nx), and grid: a positive multiple of 8 selects the one-launch persistent form with that many
 * blocks, 0 the per-kind launches with one block per tile. AERO_E_ARG for any other value. A new batch starts
 * at 512 threads (FP32 storage) or 256 (FP16), 1 cell per thread, grid 0. */
int aero_batch_set_shape(aero_batch* b, int32_t block, int32_t cells_per_thread, int32_t grid);
/* out3 = block, cells per thread, grid. */
int aero_batch_get_shape(aero_batch* b, int32_t* out3);
/* Times every candidate shape for `steps` steps and keeps the fastest (out3 = shape as above, ms = ms per
 * step). The flow advances while tuning: reset the tunnels afterwards. */
int aero_batch_tune(aero_batch* b, int32_t steps, int32_t* out3, double* ms);
int aero_batch_step(aero_batch* b, int32_t n);
/* out[0] = total ms of n steps (HIP events), out[1] = ms per step. */
int aero_batch_time_steps(aero_batch* b, int32_t n, double* out);

/* ---- learned closure ------------------------------------------------------------------------- */
/* A 6 -> 32 -> 32 -> 1 MLP mapping invariants of the non-equilibrium stress to a factor on the Smagorinsky
 * eddy viscosity for that stress (nu_t = nu_S * softplus(y + ln(e - 1)), tau = 3 (nu + nu_t) + 1/2; the zero
 * network is exactly the Smagorinsky closure). No counterpart in the reference, whose only closure
 * is Smagorinsky (collideStream.wgsl:154-158, lbmCpuReference.ts:280-282); it serves callers that train a
 * closure on filtered fine-grid runs. Forward and backward run on the matrix cores (BF16 operands, FP32
 * accumulation); parameters are 1313 FP32 values: W1[32][6], b1[32], W2[32][32], b2[32], w3[32], b3. */
typedef struct aero_mlp aero_mlp;
int aero_mlp_create(int32_t device, aero_mlp** out);
int aero_mlp_destroy(aero_mlp* m);
const char* aero_mlp_last_error(const aero_mlp* m);
int aero_mlp_set_params(aero_mlp* m, const float* params);
int aero_mlp_get_params(aero_mlp* m, float* params);
/* Training samples: 8 columns (features 0-5, column 6 zero, column 7 the sample's loss weight with 0
 * standing for 1; the network reads only the features) and one target factor on the Smagorinsky viscosity
 * each; the loss is w (softplus(y + ln(e - 1)) - t)^2. */
int aero_mlp_add_samples(aero_mlp* m, const float* x, const float* t, int64_t n);
int aero_mlp_clear_samples(aero_mlp* m);
int aero_mlp_num_samples(aero_mlp* m, int64_t* n);
int aero_mlp_get_samples(aero_mlp* m, float* x, float* t);
/* Appends samples filtered from the last readout fields of a fine-grid context onto the coarse grid of half
 * its size (2x2x2 blocks), on a sampling lattice of the given stride (offset chosen by seed). */
int aero_mlp_sample_fine(aero_mlp* m, aero_ctx* fine, int32_t stride, uint64_t seed, double nu_coarse, double cs,
                         int64_t* added);
/* Gradient of the mean loss over n host samples (grad: 1313 floats) and the loss. */
int aero_mlp_grad(aero_mlp* m, const float* x, const float* t, int32_t n, float* grad, double* loss);
/* Adam over the stored samples in contiguous batches; loss (may be NULL) receives each iteration's loss. */
int aero_mlp_train(aero_mlp* m, int32_t iters, int32_t batch, double lr, float* loss);
int aero_mlp_predict(aero_mlp* m, const float* x, int32_t n, float* y);
/* Uses m as the solver's closure (m must outlive ctx's use of it); NULL restores Smagorinsky. */
int aero_set_closure(aero_ctx* ctx, const aero_mlp* m);

/* ---- derived-field readouts (GPU, on the macro fields of the last AERO_STEP_READOUT step) -------- */
#define AERO_SLICE_Z 0          /* x-y plane at iz; width nx, height ny */
#define AERO_SLICE_Y 1          /* x-z plane at iy; width nx, height nz */
#define AERO_SLICE_X 2          /* z-y plane at ix; width nz, height ny */
#define AERO_FIELD_VELOCITY 0   /* |u| / (1.55 u_ref) */
#define AERO_FIELD_VORTICITY 1  /* |curl u| / (0.42 u_ref), alpha from it */
#define AERO_FIELD_CP 2         /* Cp = ((rho - 1)/3) / (u_ref^2 / 2), shown over [-1.1, 1.0] */
#define AERO_SLICE_RULES_CPU 0  /* extractSliceRGBA (lbmCpuReference.ts:500-596), FP64 */
#define AERO_SLICE_RULES_WGSL 1 /* the GPU post-process shader (postProcess.wgsl:52-146), FP32: its
                                   constants 1.6 and 0.45, alpha 0.82 / 0.25 + 0.70 |curl|, truncation */

/* Replaces extractSliceRGBA (lbmCpuReference.ts:500-596) and the post-process dispatch of
 * stepAndExtractSlice (lbmEngine.ts:529-553): the cross-section at index
 * min(n - 1, max(0, floor(fraction * n))) along the axis, one RGBA8 pixel per cell (row v, column u
 * at byte offset 4*(v*width + u)), solid cells transparent black. u_ref < 0 selects the context's
 * u_inlet; the reference takes max(1e-4, u_ref). wh (2 ints) receives width and height; rgba may be
 * NULL to get only those. */
int aero_read_slice_rgba(aero_ctx* ctx, int32_t axis, double fraction, int32_t field, int32_t rules, double u_ref,
                         uint8_t* rgba, int32_t* wh);
/* Replaces computeQCriterion (vortexIso.ts:22-119) on the context's velocity field and cell flags:
 * Q = (|Omega|^2 - |S|^2)/2 from central differences (one-sided at the domain faces) with cell size dx,
 * the vorticity magnitude and the vorticity vector (3 floats per cell, x y z), solid cells 0. Any output
 * may be NULL. maxima (may be NULL) receives maxQ and maxOmega (both at least 1e-6). */
int aero_read_qcriterion(aero_ctx* ctx, double dx, float* q, float* omega_mag, float* omega_vec, double* maxima);
/* computeQCriterion on caller-supplied fields (no solver context): ux, uy, uz of nx*ny*nz cells,
 * cell_flags may be NULL (no cell skipped). Runs on HIP device `device`. */
int aero_qcriterion(int32_t device, int32_t nx, int32_t ny, int32_t nz, double dx, const float* ux, const float* uy,
                    const float* uz, const uint8_t* cell_flags, float* q, float* omega_mag, float* omega_vec,
                    double* maxima);

/* Probe inputs: the voxel grid's placement (VoxelGridResult originX/Y/Z and dx, voxelizer.ts:19-59) and
 * the physical scaling of computeProbePhysics (pointProbe.ts:23-34). */
typedef struct {
  double origin[3];
  double dx;
  double wind_speed_mps;
  double u_inlet_lattice;
  double fluid_density;
} aero_probe_params;

/* Replaces the point probe's readout (pointProbe.ts:284-355 with computeProbePhysics, :23-92, and
 * sampleVelocityTrilinear, streamlines.ts:20-81) for n world points xyz (3 doubles each). on_surface
 * (n bytes, may be NULL) marks points whose ray hit the mesh. Per point, 13 doubles: ix, iy, iz,
 * speed m/s, speed mph, ux uy uz m/s, gauge pressure Pa, Cp, rho, dynamic pressure Pa, inside-solid
 * (0/1). Points must be finite. */
int aero_probe(aero_ctx* ctx, const aero_probe_params* params, const double* xyz, const uint8_t* on_surface,
               int32_t n, double* out);

/* ---- host geometry helpers (no GPU needed) ---------------------------------------------------- */
/* Mesh generators. Each writes up to cap triangles (9 floats each) into out and the count into
 * ntri; call with out = NULL to size. Lengths are in the caller's units (cells for the lattice
 * workloads). */
int aero_mesh_icosphere(int32_t level, double radius, const double centre[3], float* out, int32_t cap,
                        int32_t* ntri);
/* Finned rocket along +x: nose tip at x0, tangent-ogive nose of nose_len, cylinder body_len x diam,
 * 4 trapezoidal fins; jitter > 0 perturbs the dimensions by a seeded uniform +-jitter fraction. */
int aero_mesh_rocket(const double tip[3], double body_len, double diam, double nose_len, double jitter,
                     uint64_t seed, int32_t segments, float* out, int32_t cap, int32_t* ntri);
/* Symmetric NACA 00tt section (thickness ratio t), chord c, leading edge at le[], pitched nose-up by
 * alpha_deg about the quarter chord, extruded along z over [zmin, zmax]. */
int aero_mesh_naca(double t, double chord, const double le[3], double alpha_deg, double zmin, double zmax,
                   int32_t npts, float* out, int32_t cap, int32_t* ntri);
/* Axis-aligned box [lo, hi]. */
int aero_mesh_box(const double lo[3], const double hi[3], float* out, int32_t cap, int32_t* ntri);
/* Mesh statistics (stlLoader.ts:133-244): out[0] = projected frontal area on the y-z plane
 * (frontalArea.ts:22-179, 512^2 x 16 sub-pixels), out[1..3] = bounding box size x, y, z,
 * out[4] = volume, out[5] = open edge count. */
int aero_mesh_stats(const float* tri_pos, int32_t ntri, const double* world, double* out);

/* ---- STL ingestion (host, no GPU needed) ------------------------------------------------------ */
/* Replaces parseSTL (stlLoader.ts:9-95): binary or ASCII STL bytes to triangle positions (9 floats
 * per triangle, the vertices in file order). The format is decided as the reference decides it;
 * AERO_E_FORMAT with aero_last_error(NULL) = "Invalid or corrupted STL file." or "ASCII STL file
 * contains no valid triangles." when it would throw. Call with out = NULL to get the count; an
 * 84-byte binary file with count 0 yields 0 triangles. Normals are not produced (the solver path
 * never reads them). */
int aero_stl_parse(const uint8_t* bytes, int64_t len, float* out, int64_t cap, int64_t* ntri);
/* Replaces normalizeAndCenterGeometry (stlLoader.ts:101-127), in place: centre the bounding box
 * on the origin, then, if its largest side exceeds 10, scale by 0.001 (side > 50) or 1/side.
 * scaled_from_mm (may be NULL) receives autoScaledFromMm. */
int aero_stl_normalize(float* tri_pos, int64_t ntri, int32_t* scaled_from_mm);

#ifdef __cplusplus
}
#endif

#endif /* AEROLBM_H */
