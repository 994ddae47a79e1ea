/*
 * gocean.h -- C ABI of the MI355X (gfx950) CDKLM16 shallow-water step.
 *
 * Every entry point is extern "C", takes plain device pointers, element
 * pitches and a hipStream_t passed as void*, returns 0 on success or a
 * non-zero code whose message gocean_last_error() returns (thread-local).
 * No C++ exception crosses this boundary.  All arrays are float32 device
 * memory owned by the caller; the library owns only its small scratch.
 *
 * Layout (the reference's, gpu_ocean/SWESimulators/Common.py:503-546):
 *   state arrays eta, hu, hv : (ny+4) rows x pitch floats, two ghost cells on
 *                              every side, interior cells [2,ny+2) x [2,nx+2)
 *   corner depth Hi          : (ny+5) rows x hi_pitch floats, Hi[j][i] is the
 *                              south-west corner of cell (j,i)
 *   cell depth Hm            : (ny+4) rows x pitch floats
 * Land is marked in Hi by land_value (1e20, Common.py:1331).
 */
#ifndef GOCEAN_H
#define GOCEAN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GOCEAN_BC_WALL 1
#define GOCEAN_BC_PERIODIC 2

/* Return codes: 0 success, > 0 a hipError_t passed through, < 0 below. */
#define GOCEAN_OK 0
#define GOCEAN_EINVAL (-1)   /* invalid argument (message says which) */
#define GOCEAN_ENOMEM (-2)   /* device allocation failed */
#define GOCEAN_ESTATE (-3)   /* call out of order (e.g. replay before capture) */

/* Scheme constants of one domain.  The reference bakes these into the kernel
 * with #defines at JIT time (gpu_ocean/SWESimulators/CDKLM16.py:179-191);
 * here they are run-time arguments so that one launch can serve domains of
 * different sizes. */
typedef struct gocean_domain {
    int32_t nx, ny;          /* interior cells */
    int32_t pitch;           /* floats per row of state / Hm arrays, >= nx+4 */
    int32_t hi_pitch;        /* floats per row of Hi, >= nx+5 */
    float dx, dy, g, r;      /* spacing, gravity, quadratic bottom friction */
    float theta;             /* minmod theta (1.3) */
    float desing_eps;        /* KPSIMULATOR_DESING_EPS (0.1) */
    float flux_slope_eps;    /* KPSIMULATOR_FLUX_SLOPE_EPS (0.1) */
    float depth_cutoff;      /* KPSIMULATOR_DEPTH_CUTOFF (1e-5) */
    float land_value;        /* 1e20 */
    float f;                 /* constant Coriolis parameter */
    float north_x, north_y;  /* constant north vector (sin a, cos a) */
    int32_t bc_north, bc_east, bc_south, bc_west;   /* GOCEAN_BC_* */
    int32_t rk_order;        /* 1, 2 or 3 */
} gocean_domain_t;

/* Wind stress (tau/rho_water, m^2/s^2) at the two bracketing times of its
 * time series and the interpolation weight t in [0,1] (Simulator.py:145-223,
 * gpu_kernels/common.cu:616-650).
 *   grid == NULL: uniform wind x0, x1 (current, next) and y0, y1 -- the
 *                 reference's 1x1 textures;
 *   grid != NULL: gridded wind, four device planes of grid_ny x grid_nx
 *                 floats back to back: X current, X next, Y current, Y next,
 *                 read at each cell like the reference's bilinear textures
 *                 (normalised coordinates (ti+0.5)/nx with the
 *                 ghost-inclusive cell index ti, clamped, linear filter;
 *                 CDKLM16_kernel.cu:827-828).  Only the plain substep and the
 *                 step loop take a grid; ensembles, drifters and the adjoint
 *                 reject it. */
typedef struct gocean_wind {
    float t, x0, x1, y0, y1;
    const float* grid;
    int32_t grid_nx, grid_ny;
} gocean_wind_t;

/* Optional per-cell fields, (ny+4) x pitch floats each; NULL selects the
 * constant f / north vector of the domain.  Replace the angle_tex and
 * coriolis_f_tex textures (gpu_kernels/CDKLM16_kernel.cu:30-77). */
typedef struct gocean_fields {
    const float* f;
    const float* north_x;
    const float* north_y;
} gocean_fields_t;

/* Library version string. */
const char* gocean_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* gocean_last_error(void);

/* Number of visible HIP devices (0 when there is none). */
int gocean_device_count(void);

/* Fill the two-deep halo of Hi (walls mirror, periodic wraps) and compute
 * Hm = mean of the wet corners.  Replaces Common.Bathymetry's device work:
 * _boundaryConditions + initBm (gpu_ocean/SWESimulators/Common.py:1376-1389,
 * 1436-1459 -> gpu_kernels/boundary_kernels.cu:121-246,
 * gpu_kernels/initBm_kernel.cu:38-107). */
int gocean_bathymetry_init(const gocean_domain_t* d, float* Hi, float* Hm,
                           void* stream);

/* One Runge-Kutta substep for all interior cells (cdklm_swe_2D,
 * gpu_ocean/SWESimulators/gpu_kernels/CDKLM16_kernel.cu:429-1006, launched by
 * CDKLM16.callKernel, CDKLM16.py:613-639).  Reads the stencil from *_in,
 * reads Q^n from *_qn when rk_step >= 1 (pass the output arrays, as the
 * reference does, or NULL for "same as output"), writes *_out.  Wall and
 * periodic boundaries are applied while loading the tile, so no separate
 * ghost-cell launch is needed; for periodic sides the kernel also refreshes
 * the output's ghost cells.  dt_dev, when not NULL, is a device float that
 * overrides dt (device-resident time step). */
int gocean_cdklm_substep(const gocean_domain_t* d, float dt, const float* dt_dev,
                         int rk_step,
                         const float* eta_in, const float* hu_in, const float* hv_in,
                         const float* eta_qn, const float* hu_qn, const float* hv_qn,
                         float* eta_out, float* hu_out, float* hv_out,
                         const float* Hi, const gocean_fields_t* fields,
                         const gocean_wind_t* wind, int tile_config,
                         void* stream);

/* n full RK steps at fixed dt (rk_order substeps each), h0 -> h0 with h1 as
 * the intermediate buffer; the sequence of CDKLM16.step's loop body
 * (CDKLM16.py:508-593) for a run of equal local_dt. */
int gocean_rk_steps(const gocean_domain_t* d, int n_steps, float dt,
                    const float* dt_dev,
                    float* eta0, float* hu0, float* hv0,
                    float* eta1, float* hu1, float* hv1,
                    float* eta2, float* hu2, float* hv2,   /* RK3 only, else NULL */
                    const float* Hi, const gocean_fields_t* fields,
                    const gocean_wind_t* wind, int tile_config, void* stream);

/* Periodic ghost-cell exchange, north/south then east/west
 * (BoundaryConditionsArakawaA.boundaryCondition, Common.py:1219-1239 ->
 * gpu_kernels/boundary_kernels.cu:35-110).  Wall sides are left untouched. */
int gocean_periodic_bc(const gocean_domain_t* d, float* eta, float* hu, float* hv,
                       void* stream);

/* CFL time step (per_block_max_dt + max_dt_reduction,
 * gpu_kernels/max_dt.cu:42-194, CDKLM16.updateDt, CDKLM16.py:712-739).
 * Writes out_dev[0] = min over wet interior cells of
 * min(0.25dx/|u+-c|, 0.25dy/|v+-c|) (1e5 if none), and, when courant > 0,
 * out_dev[1] = courant*out_dev[0] rounded to float.  out_dev holds
 * 2 + min(ny, 1024) floats: the partial minima of the reduction follow the
 * two results.  No host sync. */
int gocean_max_dt(const gocean_domain_t* d, const float* eta, const float* hu,
                  const float* hv, const float* Hm, float courant, float* out_dev,
                  void* stream);

/* gocean_max_dt without the caller's partial-minima buffer: out_dev holds
 * just the two results, and the partial minima go to a static device array of
 * the library (1024 floats, one copy per GPU, kept until the process ends).
 * No counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc): it calls no allocator.  Calls that overlap on one GPU
 * share the array and race; calls on one stream, or ordered by events, are
 * safe.  Same results as gocean_max_dt, bit for bit. */
int gocean_max_dt_static(const gocean_domain_t* d, const float* eta, const float* hu,
                         const float* hv, const float* Hm, float courant, float* out_dev,
                         void* stream);

/* Tile configuration (index into gocean_tile_configs; -1, the default: the
 * column-marching kernel from a million cells on, 32x16/512 below) of the
 * recording forward substep the adjoint entry points run
 * (gocean_adjoint_record / _segment).  Process-wide; every configuration
 * writes the same record.  An ensemble records with its own configuration
 * (gocean_ensemble_set_tile_config): through its strips when that is a
 * column-marching one, else with 32x16/512 tiles.  No counterpart in the
 * reference (the adjoint is new); it lets a tuner pick the fastest recording
 * tile, as for the forward. */
int gocean_adjoint_set_record_tile_config(int cfg);
int gocean_adjoint_record_tile_config(void);

/* Kernel of the substep adjoint in the reverse entry points: 0 = one 16x16
 * tile per workgroup through LDS, 1 = column-marching wavefronts, -1 (the
 * default) = by size, the marching kernel from 2 M cells on (a domain's
 * cells, or all members' cells together for an ensemble).  Both give the
 * same bits.  An ensemble whose member planes exceed the marching kernel's
 * 32-bit offsets takes the tile kernel.  Process-wide; no counterpart in the
 * reference. */
int gocean_adjoint_set_reverse_kernel(int kind);
int gocean_adjoint_reverse_kernel(void);

/* Names of the tile configurations the library is built with; returns the
 * count and writes up to max entries of "BWxBH/threads" into names. */
int gocean_tile_configs(const char** names, int max);
/* The configuration a tile_config of -1 selects for an nx x ny domain (or an
 * ensemble whose members average nx*ny cells): the column-marching kernel
 * from a million cells on, the 32x16/512 tile below. */
int gocean_default_tile_config(int nx, int ny);

/* Device self-check of the square root the substep kernels take for wave
 * speeds when the host enables their fast arithmetic: counts the float bit
 * patterns in [lo_bits, hi_bits] (positive floats, lo_bits <= hi_bits <=
 * 0x7F7FFFFF) for which the reciprocal-square-root-plus-Newton sequence
 * differs from the correctly rounded square root, and writes the count to
 * *mismatches.  Synchronous.  No counterpart in the reference: it backs the
 * claim that the shortcut is exact for every argument >= 2^-60. */
int gocean_sqrt_newton_check(uint32_t lo_bits, uint32_t hi_bits, unsigned long long* mismatches);

/* ------------------------------------------------------------------------
 * Batched ensembles.  The reference steps ensemble members one at a time
 * from Python (OceanModelEnsemble.modelStep,
 * gpu_ocean/SWESimulators/OceanModelEnsemble.py:105-117;
 * EnsembleFromFiles.stepToObservation, EnsembleFromFiles.py:302-325;
 * CDKLM16.dataAssimilationStep, CDKLM16.py:650-704).  Here all members, of
 * any sizes, advance together: one launch per RK substep covers every
 * member's tiles.  The caller owns flat device buffers:
 *   state buffers (two, three for RK3): member m's eta at state_offset,
 *       hu at state_offset + plane, hv at state_offset + 2*plane,
 *       plane = (ny+4)*pitch;
 *   Hi buffer (member m at hi_offset), Hm buffer (hm_offset);
 *   optional field buffer: f at field_offset, north_x at +plane,
 *       north_y at +2*plane (field_offset -1 = constant f / north);
 *   dt: one device float per member.
 * The handle owns the member table, the tile -> member map and an optional
 * captured graph.  Calls on one handle must not overlap.
 * ------------------------------------------------------------------------ */
typedef struct gocean_member {
    gocean_domain_t domain;
    int64_t state_offset;
    int64_t hi_offset;
    int64_t hm_offset;
    int64_t field_offset;
    gocean_wind_t wind;
} gocean_member_t;

typedef struct gocean_ensemble gocean_ensemble_t;

/* Build the member table (all members share rk_order). */
int gocean_ensemble_create(const gocean_member_t* members, int n, int tile_config,
                           gocean_ensemble_t** out);
int gocean_ensemble_destroy(gocean_ensemble_t* ens);
/* Switch the tile configuration of the batched substep (index into
 * gocean_tile_configs; -1 = default) and rebuild the tile -> member map, for
 * an autotuner.  Synchronises the device first and drops a captured graph
 * (capture again).  Members with per-cell fields or a rotated grid always use
 * the 32x8 tile. */
int gocean_ensemble_set_tile_config(gocean_ensemble_t* ens, int tile_config);
/* n_members, total tiles and kernel mode (0 zero angle, 1 constant north,
 * 2 per-cell fields). */
int gocean_ensemble_info(const gocean_ensemble_t* ens, int* n_members, int* n_tiles, int* mode);
/* Hi halo + Hm for every member (gocean_bathymetry_init per member). */
int gocean_ensemble_bathymetry_init(gocean_ensemble_t* ens, float* Hi, float* Hm, void* stream);
/* Periodic ghost exchange for every member (gocean_periodic_bc per member). */
int gocean_ensemble_periodic_bc(gocean_ensemble_t* ens, float* state, void* stream);
/* dt_out[m] = courant * CFL minimum of member m (raw minimum if courant <= 0),
 * on the device; the per-member form of gocean_max_dt. */
int gocean_ensemble_max_dt(gocean_ensemble_t* ens, const float* state, const float* Hm,
                           float courant, float* dt_out, void* stream);
/* One substep of every member (gocean_cdklm_substep per member), dt per
 * member from the device array.  A member whose dt is 0 is left untouched
 * (its output buffer is not written): members with different dt can then
 * share the launches of a step to a common end time.  The same holds for
 * gocean_ensemble_rk_steps and the captured graphs. */
int gocean_ensemble_substep(gocean_ensemble_t* ens, int rk_step, const float* state_in,
                            const float* state_qn, float* state_out, const float* Hi,
                            const float* fields, const float* dt, void* stream);
/* n full RK steps of every member (gocean_rk_steps per member). */
int gocean_ensemble_rk_steps(gocean_ensemble_t* ens, int n_steps, float* state0, float* state1,
                             float* state2, const float* Hi, const float* fields,
                             const float* dt, void* stream);
/* Capture n full RK steps of every member into a HIP graph held by the
 * handle (no host synchronisation inside; dt is read from the device), and
 * replay it on a stream. */
int gocean_ensemble_capture(gocean_ensemble_t* ens, int n_steps, float* state0, float* state1,
                            float* state2, const float* Hi, const float* fields, const float* dt);
int gocean_ensemble_replay(gocean_ensemble_t* ens, void* stream);

/* Reverse mode of the batched ensemble (no counterpart in the reference; the
 * batched form of gocean_adjoint_segment for data assimilation over an
 * ensemble of domains).  Members must use constant f, north = (0, 1), r == 0
 * and rk_order 2.  Setup builds the reverse-mode tile maps for segments of
 * up to max_steps steps and returns the floats of work buffer the segments
 * need (member m's block holds its (16 max_steps + 9) planes), or a negative
 * error code. */
long long gocean_ensemble_adjoint_setup(gocean_ensemble_t* ens, int max_steps);
/* n_steps (<= max_steps) RK2 steps of every member from the checkpoint
 * state0 with the per-member device dt: recompute with branch records, then
 * map the members' adjoints at the segment end (lam_in, laid out like the
 * state) to the segment start (lam_out).  Fused launches: each recording
 * substep, adjoint tile pass and ghost fold covers all members at once. */
int gocean_ensemble_adjoint_segment(gocean_ensemble_t* ens, int n_steps, const float* dt,
                                    const float* state0, const float* lam_in, float* lam_out,
                                    const float* Hi, float* work, long long work_floats,
                                    void* stream);
/* The two halves of gocean_ensemble_adjoint_segment, for callers that keep
 * one work buffer per segment from the forward run instead of recomputing
 * it (as gocean_adjoint_record / gocean_adjoint_reverse for one domain).
 * record: n_steps recording RK2 steps of every member from state0 into
 * `work`; when state_out is not NULL (and not state0) it receives the state
 * after the segment, laid out like state0, with the cells a step does not
 * write (wall ghosts, pitch padding) taken from state0.  reverse: lambda at
 * the segment's end to lambda at its start from a work buffer that record
 * filled for the same n_steps, dt, state0 and Hi; it may be repeated. */
int gocean_ensemble_adjoint_record(gocean_ensemble_t* ens, int n_steps, const float* dt,
                                   const float* state0, const float* Hi, float* work,
                                   long long work_floats, float* state_out, void* stream);
int gocean_ensemble_adjoint_reverse(gocean_ensemble_t* ens, int n_steps, const float* dt,
                                    const float* state0, const float* lam_in, float* lam_out,
                                    const float* Hi, float* work, long long work_floats,
                                    void* stream);

/* ---- adjoint (reverse mode) of the fixed-dt RK2 step ----
 * No counterpart in the reference, which has no gradients: this serves
 * variational data assimilation / parameter estimation callers that need
 * d(misfit)/d(initial state) of gocean_rk_steps with a fixed dt.  Scope:
 * north = (0, 1), constant f, r == 0, rk_order == 2; other domains return
 * GOCEAN_EINVAL.  All state / adjoint arrays are padded (ny+4) x pitch
 * planes; ghost entries of adjoint inputs are ignored and those of adjoint
 * outputs are zero (their sensitivity is folded into the interior cells the
 * ghosts mirror or wrap).  Branches (minmod candidate, wave-speed min/max,
 * upwinding, positivity adjustment, wet / desingularised / dry cells) follow
 * the forward evaluation: each entry point first re-runs the forward
 * substeps it differentiates in a recording variant that stores every such
 * choice, then applies the hand-written Jacobian-transpose kernels.  `wind`
 * (nullable = none) is the uniform wind of the forward run being
 * differentiated; the recomputed trajectory must match it.
 * Scratch / work buffers are caller-owned device memory. */

/* Scratch of gocean_adjoint_substep / gocean_adjoint_rk2_step, in floats. */
long long gocean_adjoint_scratch_floats(const gocean_domain_t* d);

/* o = alpha * J^T w + gamma * x, with J the Jacobian of the substep map
 * Q -> base(Q) + dt * L(Q) (the rk_step 0 update, CDKLM16_kernel.cu:905-1001)
 * at Q = (eta, hu, hv).  Cells that the forward update clamps to the dry state
 * (eta = -Hm + cutoff, CDKLM16_kernel.cu:986-1001) pass no w: the mask is
 * taken from w_clamp (a forward eta output) or, when w_clamp is NULL, from
 * this substep's own output.  x_clamp (nullable) masks x the same way; x may
 * be NULL. */
int gocean_adjoint_substep(const gocean_domain_t* d, const gocean_wind_t* wind, float dt,
                           const float* dt_dev, const float* eta, const float* hu, const float* hv,
                           const float* w_eta, const float* w_hu, const float* w_hv,
                           const float* w_clamp, const float* x_eta, const float* x_hu,
                           const float* x_hv, const float* x_clamp, float alpha, float gamma,
                           float* o_eta, float* o_hu, float* o_hv, const float* Hi,
                           float* scratch, void* stream);

/* lambda^n = (dQ^{n+1}/dQ^n)^T lambda^{n+1} for one SSP-RK2 step
 * Q* = P(Q^n), Q^{n+1} = (Q^n + P(Q*)) / 2 (CDKLM16.py:533-544).  Needs the
 * step's Q^n and Q* (and, nullable, eta of Q^{n+1} for the dry-clamp mask;
 * NULL recomputes it); o must not alias l. */
int gocean_adjoint_rk2_step(const gocean_domain_t* d, const gocean_wind_t* wind, float dt,
                            const float* dt_dev, const float* eta_n, const float* hu_n,
                            const float* hv_n,
                            const float* eta_s, const float* hu_s, const float* hv_s,
                            const float* eta_next, const float* l_eta, const float* l_hu,
                            const float* l_hv, float* o_eta, float* o_hu, float* o_hv,
                            const float* Hi, float* scratch, void* stream);

/* Work buffer of gocean_adjoint_segment for n_steps steps, in floats
 * ((16 n + 9) planes). */
long long gocean_adjoint_segment_work_floats(const gocean_domain_t* d, int n_steps);

/* Reverse mode of n_steps fixed-dt RK2 steps started from the checkpoint
 * (eta0, hu0, hv0): recomputes the segment (with branch records) into `work`,
 * then maps lambda at the segment's end (l_*) to lambda at its start (o_*).
 * One call per checkpoint segment; everything runs on `stream`. */
int gocean_adjoint_segment(const gocean_domain_t* d, const gocean_wind_t* wind, int n_steps,
                           float dt, const float* dt_dev, const float* eta0, const float* hu0,
                           const float* hv0,
                           const float* l_eta, const float* l_hu, const float* l_hv,
                           float* o_eta, float* o_hu, float* o_hv, const float* Hi,
                           float* work, long long work_floats, void* stream);

/* The two halves of gocean_adjoint_segment, for callers that keep the
 * records of the forward run instead of recomputing them in the reverse
 * run.  No reference counterpart (the reference has no reverse mode).
 *
 * gocean_adjoint_record: n_steps recording RK2 steps from (eta0, hu0, hv0)
 * into `work`.  Work layout, plane = (ny + 4) * pitch floats: Q_1..Q_n
 * (eta, hu, hv planes each; Q_n is the state after the segment, the same
 * bits as gocean_rk_steps on the cells a step writes: interior and periodic
 * ghost cells; wall ghost cells and pitch padding are left unwritten),
 * Q*_0..Q*_{n-1} (3 planes each), the branch
 * records of the 2n substeps (5 planes each), 9 planes of lambda scratch.
 *
 * gocean_adjoint_reverse: lambda at the segment's end (l_*) to lambda at its
 * start (o_*) from a work buffer that gocean_adjoint_record filled for the
 * same n_steps, dt, start state (eta0 is read) and bathymetry.  The records
 * stay valid, so the reverse run may be repeated for other l_*. */
int gocean_adjoint_record(const gocean_domain_t* d, const gocean_wind_t* wind, int n_steps, float dt,
                          const float* dt_dev, const float* eta0, const float* hu0, const float* hv0,
                          const float* Hi, float* work, long long work_floats, void* stream);
int gocean_adjoint_reverse(const gocean_domain_t* d, int n_steps, float dt, const float* dt_dev,
                           const float* eta0, const float* l_eta, const float* l_hu, const float* l_hv,
                           float* o_eta, float* o_hu, float* o_hv, const float* Hi, float* work,
                           long long work_floats, void* stream);

/* ------------------------------------------------------------------------
 * Model error: SOAR-correlated eta perturbation with geostrophically
 * balanced (hu, hv), added to the state (OceanStateNoise,
 * gpu_ocean/SWESimulators/OceanStateNoise.py:399-496; kernels
 * gpu_kernels/ocean_noise.cu and gpu_kernels/reductions.cu).
 *
 * Buffers (device, row-major, no padding):
 *   seed    seed_ny x seed_nx uint64 (LCG generator state)
 *   random  rand_ny x rand_nx float  (xi; nu for the perpendicular field)
 *   coarse  (coarse_ny+4) x (coarse_nx+4) float
 * eta/hu/hv are (ny+4) x pitch, Hi (ny+5) x hi_pitch, as everywhere else;
 * only interior cells of eta/hu/hv are written (the reference's kernels
 * leave ghost cells to the boundary-condition launch that follows).
 * ------------------------------------------------------------------------ */
typedef struct gocean_noise {
    int32_t nx, ny;                 /* interior cells of the simulator grid */
    int32_t pitch, hi_pitch;        /* floats per row of eta/hu/hv (and fields), of Hi */
    float dx, dy;
    int32_t interpolation_factor;   /* positive odd, divides nx and ny */
    int32_t periodic_ns, periodic_ew;
    float soar_q0, soar_L;          /* SOAR amplitude and length scale */
    float g, f, beta;               /* coriolis = f + beta ((i+1/2)dx n_x + (j+1/2)dy n_y),
                                       ghosted (i, j); f == 0 reads fields->f */
    float north_x, north_y;         /* north vector when fields->north_* are NULL */
    float land_value;               /* Hi value marking land (1e20) */
} gocean_noise_t;

#define GOCEAN_RNG_LCG_NORMAL 0     /* normalDistribution, ocean_noise.cu:163 */
#define GOCEAN_RNG_LCG_UNIFORM 1    /* uniformDistribution, ocean_noise.cu:121 */
#define GOCEAN_RNG_PHILOX_NORMAL 2  /* in place of XORWOW fill_normal, OceanStateNoise.py:324 */
#define GOCEAN_RNG_PHILOX_UNIFORM 3 /* in place of XORWOW fill_uniform, OceanStateNoise.py:346 */

/* sizes = coarse_nx, coarse_ny, rand_nx, rand_ny, seed_nx, seed_ny,
 * rand_ghost_x, rand_ghost_y (OceanStateNoise.py:97-135). */
int gocean_noise_sizes(const gocean_noise_t* nz, int32_t sizes[8]);

/* Fill `random` with N(0,1) or U(0,1) numbers (generateNormalDistribution /
 * generateUniformDistribution, OceanStateNoise.py:320-352).  LCG kinds
 * advance `seed` twice per pair, bit for bit as ocean_noise.cu:80-195
 * (including its column rule: with odd rand_nx the last column is not
 * written).  Philox kinds (Philox4x32-10, counter = (element/4, *counter,
 * member 0), key) read the device counter and advance it by one. */
int gocean_noise_generate(const gocean_noise_t* nz, int kind, uint64_t* seed, float* random,
                          uint64_t key, uint64_t* counter, void* stream);

/* coarse = [coarse +] scale * q0 * SOAR(random): the SOAR kernel,
 * ocean_noise.cu:264-370 (additive != 0 adds, as for the perpendicular
 * field, OceanStateNoise.py:449-458).  q0 = soar_q0 * q0_scale. */
int gocean_noise_soar(const gocean_noise_t* nz, float q0, float scale, const float* random,
                      float* coarse, int additive, void* stream);

/* Add d_eta and the balanced (d_hu, d_hv) to the interior of eta, hu, hv:
 * geostrophicBalance (ocean_noise.cu:379-505) for interpolation_factor 1,
 * bicubicInterpolation (ocean_noise.cu:522-772, alignment offsets from
 * OceanStateNoise._obtain_coarse_grid_offset) for odd factors > 1.
 * fields may be NULL (constant f and north). */
int gocean_noise_balance(const gocean_noise_t* nz, const float* coarse, float* eta, float* hu,
                         float* hv, const float* Hi, const gocean_fields_t* fields, int offset_i,
                         int offset_j, void* stream);

/* The SOAR kernel followed by geostrophicBalance in one launch, for
 * interpolation factor 1 (ocean_noise.cu:264-505): the same coarse buffer
 * and the same state as gocean_noise_soar (not additive) then
 * gocean_noise_balance with zero offsets, without the coarse buffer's round
 * trip through memory between the two. */
int gocean_noise_soar_balance(const gocean_noise_t* nz, float q0, float scale, const float* random,
                              float* coarse, float* eta, float* hu, float* hv, const float* Hi,
                              const gocean_fields_t* fields, void* stream);

/* out = [sum xi^2, sum nu^2, sum xi nu] over the random field (squareSum /
 * squareSumDouble, reductions.cu:27-185); nu == NULL writes out[0] only. */
int gocean_noise_norms(const gocean_noise_t* nz, const float* xi, const float* nu, float* out,
                       void* stream);

/* nu <- sqrt(|nu|^2/|nu_perp|^2) (nu - (xi.nu/|xi|^2) xi) with `norms` from
 * gocean_noise_norms (makePerpendicular, ocean_noise.cu:206-243). */
int gocean_noise_make_perpendicular(const gocean_noise_t* nz, const float* xi, float* nu,
                                    const float* norms, void* stream);

/* Batched model error: every member of an ensemble perturbed by one launch
 * per kernel (the reference calls perturbSim member by member,
 * CDKLM16.py:581-582 inside OceanModelEnsemble.modelStep).  Member m's
 * state lives in one flat buffer as for gocean_ensemble_*: eta at
 * state_offset, hu at +plane, hv at +2*plane, plane = (ny+4)*pitch; Hi at
 * hi_offset; fields (f, north_x, north_y planes) at field_offset or -1.
 * The batch lays out flat seed / random / coarse buffers (offsets from
 * gocean_noise_batch_layout, in elements).  Calls on one batch must not
 * overlap. */
typedef struct gocean_noise_member {
    gocean_noise_t noise;
    int64_t state_offset;
    int64_t hi_offset;
    int64_t field_offset;
} gocean_noise_member_t;

typedef struct gocean_noise_batch gocean_noise_batch_t;

int gocean_noise_batch_create(const gocean_noise_member_t* members, int n, int kind,
                              gocean_noise_batch_t** out);
int gocean_noise_batch_destroy(gocean_noise_batch_t* batch);
/* offsets[3m+0..2] = member m's seed, random and coarse offsets; totals the
 * buffer lengths (either may be NULL). */
int gocean_noise_batch_layout(const gocean_noise_batch_t* batch, int64_t* offsets, int64_t totals[3]);
/* New random fields for every member (Philox: counter (element/4, *counter,
 * member index)). */
int gocean_noise_batch_generate(gocean_noise_batch_t* batch, uint64_t* seed, float* random,
                                uint64_t key, uint64_t* counter, void* stream);
/* perturbOceanState for every member: [generate,] SOAR with amplitude
 * soar_q0 * q0_scale, balance with zero alignment offsets (one fused launch
 * when every member has interpolation factor 1). */
int gocean_noise_batch_perturb(gocean_noise_batch_t* batch, float* state, const float* Hi,
                               const float* fields, float q0_scale, float perturbation_scale,
                               int update_random, uint64_t* seed, float* random, float* coarse,
                               uint64_t key, uint64_t* counter, void* stream);

/* ------------------------------------------------------------------------
 * Drifters and the observation operator.
 * positions: (num_drifters + 1) x 2 floats (x, y in metres from the
 * south-west interior corner): the drifters, then the observation position
 * (GPUDrifterCollection.driftersDevice, GPUDrifterCollection.py:61-64).
 * x_zero / y_zero: the ghosted cell column / row holding x = 0 / y = 0 (2).
 * ------------------------------------------------------------------------ */

/* One drift step of every position, observation included
 * (passiveDrifterKernel, gpu_kernels/driftKernels.cu:108-210, launched by
 * GPUDrifterCollection.drift, GPUDrifterCollection.py:212-232): bilinear
 * velocity hu/(Hm+eta), hv/(Hm+eta) with the reference's ceil-based cell
 * index, plus wind_drift_factor times the uniform wind lerped by wind->t,
 * times sensitivity * dt; periodic sides wrap.  wind may be NULL when
 * wind_drift_factor is 0. */
int gocean_drift(const gocean_domain_t* d, int num_drifters, float* positions, const float* eta,
                 const float* hu, const float* hv, const float* Hm, float dt, float sensitivity,
                 const gocean_wind_t* wind, float wind_drift_factor, int x_zero, int y_zero,
                 void* stream);

/* Periodic wrap of every position (enforceBoundaryConditions,
 * driftKernels.cu:214-250). */
int gocean_drifters_enforce_bc(float domain_size_x, float domain_size_y, int periodic_ns,
                               int periodic_ew, int num_drifters, float* positions, void* stream);

/* obs[k] = (hu, hv) of the cell holding position k, k < num_drifters
 * (observeUnderlyingFlow, gpu_kernels/observationKernels.cu:29-82). */
int gocean_observe_flow(const gocean_domain_t* d, int num_drifters, const float* positions,
                        const float* hu, const float* hv, float* obs, int x_zero, int y_zero,
                        void* stream);

/* Batched drifters over the members of an ensemble (the reference moves and
 * observes them particle by particle, BaseOceanStateEnsemble.py:105-136).
 * Member m's state is in one flat buffer as for gocean_ensemble_* (eta at
 * state_offset, hu at +plane, hv at +2*plane), Hm at hm_offset, its
 * positions at positions_offset ((num+1) x 2 floats) and its observations at
 * observations_offset (num x 2 floats); x_zero = y_zero = 2. */
typedef struct gocean_drift_member {
    gocean_domain_t domain;
    int64_t state_offset;
    int64_t hm_offset;
    int64_t positions_offset;
    int64_t observations_offset;
    int32_t num_drifters;
    float sensitivity;
    float wind_drift_factor;
    gocean_wind_t wind;
} gocean_drift_member_t;

typedef struct gocean_drift_batch gocean_drift_batch_t;

int gocean_drift_batch_create(const gocean_drift_member_t* members, int n, gocean_drift_batch_t** out);
int gocean_drift_batch_destroy(gocean_drift_batch_t* batch);
/* One drift step of every member's positions; dt: one device float per member. */
int gocean_drift_batch_step(gocean_drift_batch_t* batch, const float* state, const float* Hm,
                            float* positions, const float* dt, void* stream);
/* Every member observes hu, hv at its num_drifters positions into its
 * observation block: shared != 0 reads one positions array for all members
 * (the true state's drifters, as BaseOceanStateEnsemble.py:105-136), else
 * each member's own block of the drift positions buffer. */
int gocean_observe_flow_batch(gocean_drift_batch_t* batch, const float* state, const float* positions,
                              int shared, float* observations, void* stream);

/* ---- Two-stage IEWPF (IEWPFOcean.iewpf_2stage, gpu_ocean/SWESimulators/IEWPFOcean.py:195-329) ----
 * Coarse random fields of every member of an ensemble: member m's field is n
 * floats (coarse_ny rows of coarse_nx) at offsets[m] (device int64) of a flat
 * buffer, the layout of gocean_noise_batch_layout for doubly periodic
 * domains.  One launch covers every member where the reference launches per
 * particle. */

/* out[3m..3m+2] = |xi|^2, |nu|^2, xi . nu of member m, float64
 * (squareSumDouble, reductions.cu; OceanStateNoise.findDoubleNormAndDot). */
int gocean_iewpf_dots(const float* xi, const float* nu, const int64_t* offsets, int32_t n_members, int32_t n,
                      double* out, void* stream);
/* nu <- sqrt(|nu|^2 / |nu_perp|^2) (nu - (xi.nu / |xi|^2) xi) with dots from
 * gocean_iewpf_dots (makePerpendicular, ocean_noise.cu:206-243). */
int gocean_iewpf_make_perpendicular(const float* xi, float* nu, const int64_t* offsets, int32_t n_members,
                                    int32_t n, const double* dots, void* stream);
/* xi <- a[m] xi + b nu, a per member on the device (blas_xaxpby,
 * iewpf_kernels.cu:47-66, as addBetaNuIntoAlphaXi uses it, IEWPFOcean.py:522-546). */
int gocean_iewpf_axpby(float* xi, const float* nu, const int64_t* offsets, int32_t n_members, int32_t n,
                       const float* a, float b, void* stream);
/* For each drifter in order, the 7 x 7 window of xi centred on coarse cell
 * cells[2d], cells[2d+1] (periodic) <- svd (49 x 49, row-major) x window
 * (localSVDOnGlobalXi, iewpf_kernels.cu:213-258; applySVDtoPerpendicular,
 * IEWPFOcean.py:693-711). */
int gocean_iewpf_local_svd(float* xi, const int64_t* offsets, int32_t n_members, int32_t coarse_nx,
                           int32_t coarse_ny, const float* svd, const int32_t* cells, int32_t n_drifters,
                           void* stream);
/* Member m's field <- zero except the 7 x 7 window centred on coarse cell
 * (cx, cy) (periodic), which gets e[m][drifter][0] patch[0] + e[m][drifter][1]
 * patch[1] (patch: 2 x 49 floats, e: n_members x n_drifters x 2 floats)
 * (setNoiseBufferToZero + halfTheKalmanGain, IEWPFOcean.py:509-520, 548-595;
 * iewpf_kernels.cu:34-45, 115-176). */
int gocean_iewpf_kalman_patch(float* xi, const int64_t* offsets, int32_t n_members, int32_t coarse_nx,
                              int32_t coarse_ny, const float* patch, const float* e, int32_t n_drifters,
                              int32_t drifter, int32_t cx, int32_t cy, void* stream);
/* Every member's state += the Kalman-gain term of every drifter for
 * interpolated model error (addKalmanGain, IEWPFOcean.py:548-595, whose term
 * is the model-error perturbation of the patch above): state is n_members
 * blocks of three (ny+4) x (nx+4) planes (eta, hu, hv), member_stride floats
 * apart.  Drifter d's term is sum_k coef[d][k][m] basis[d][k] (coef:
 * n_drifters x 5 x coef_stride floats, coef_stride at least n_members
 * rounded up to a multiple of 16, the padding read but not used; basis:
 * n_drifters x 5 x 3 x win_h x win_w floats), a win_h x win_w window of interior cells starting at row
 * origins[2d], column origins[2d+1], periodic.  tiles: n_tiles (row,
 * segment) pairs of 256-cell interior row segments; the drifters meeting tile
 * t are list[list_off[t] .. list_off[t+1]), ascending.  Ghost cells are not
 * written. */
int gocean_iewpf_kalman_window(float* state, int32_t n_members, int32_t nx, int32_t ny, int64_t member_stride,
                               const float* basis, int32_t n_drifters, int32_t win_h, int32_t win_w,
                               const int32_t* origins, const float* coef, int32_t coef_stride,
                               const int32_t* tiles, int32_t n_tiles, const int32_t* list_off, const int32_t* list,
                               void* stream);

/* ---- Localised ETKF (ETKFOcean.LETKF, gpu_ocean/SWESimulators/ETKFOcean.py:583-744) ----
 * The ensemble state is n_members blocks of three (ny+4) x pitch float planes
 * (eta, hu, hv), plane floats apart, member_stride floats between members;
 * interior cell (x, y) sits at (y+2) * pitch + x+2 of a plane.  One group of
 * observations (ETKFOcean.initializeGroups, :478-514) is analysed by
 * gocean_letkf_coefficients followed by gocean_letkf_apply on one stream;
 * groups run one after the other, and the caller refreshes periodic ghost
 * cells afterwards (uploadAnalysis, :796-812). */
typedef struct gocean_letkf_layout {
    int32_t n_members, nx, ny, pitch;
    int64_t plane;
    int64_t member_stride;
} gocean_letkf_layout_t;

/* Per observation, written by gocean_letkf_coefficients (device memory). */
typedef struct gocean_letkf_coef {
    float a_eta, a_uv;   /* anomaly scale: 1, and ff^-1/2 when hu, hv are inflated */
    double sv0, sv1;     /* member sums of the two V columns */
} gocean_letkf_coef_t;

/* Cells [x0, x0 + tile_w) x [y0, y0 + tile_h) (clipped to the domain) and the
 * observations tile_obs[begin .. begin+count) whose nonzero weights reach them,
 * in group order. */
typedef struct gocean_letkf_tile {
    int32_t x0, y0, begin, count;
} gocean_letkf_tile_t;

/* The tile shape gocean_letkf_apply expects, the most observations of one group
 * whose nonzero weights may overlap on a cell, and the largest ensemble. */
int gocean_letkf_tile_shape(int32_t* tile_w, int32_t* tile_h, int32_t* max_cover, int32_t* max_members);

/* For each observation o = group[i], i < n_group: observe hu, hv of every
 * member at interior cell cells[2o], cells[2o+1] (observe_particles_from_X_f,
 * :777-793), innovation D = y[2o..2o+1] - mean, and the local transform of
 * :651-720 (R^-1 from the 2x2 host array obs_cov, inflation: 1 none, 0 the
 * adaptive factor of :667-671, else the factor for hu and hv) in rank-2 form:
 * coef[o][k][0..5] = V_k0, V_k1, B^eta_k0, B^eta_k1, B^uv_k0, B^uv_k1 (float,
 * n_obs x n_members x 6) and oc[o].  All other pointers are device memory. */
int gocean_letkf_coefficients(const gocean_letkf_layout_t* layout, const float* state, const int32_t* group,
                              int32_t n_group, const int32_t* cells, const double* y, int32_t n_obs,
                              const double* obs_cov, double inflation, float* coef, gocean_letkf_coef_t* oc,
                              void* stream);

/* In place, for every cell of the tiles: x_e <- W_f x_e + sum over covering
 * observations o of w_loc(o) (mean + a (x_e - mean) + p . B_e), with w_loc the
 * local_n x local_n weight stencil (float64) placed at the box origin
 * boxes[2o], boxes[2o+1] (periodic), W_f = 1 - W / max(W, 1), W the sum of the
 * covering weights (getCombinedWeights and the blend, :421-443, 464-468,
 * 722-740). */
int gocean_letkf_apply(const gocean_letkf_layout_t* layout, float* state, const double* w_loc, int32_t local_n,
                       const int32_t* boxes, int32_t n_obs, const float* coef, const gocean_letkf_coef_t* oc,
                       const gocean_letkf_tile_t* tiles, int32_t n_tiles, const int32_t* tile_obs,
                       int32_t n_tile_obs, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GOCEAN_H */
