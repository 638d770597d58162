/*
 * atm.h -- C ABI of the MI355X (gfx950) HIP library for the atmosphere
 * model's per-timestep hot path.
 *
 * Drop-in boundary for felix-goddard/atmosphere-model.  The reference has no
 * FFI: its hot path is a set of Fortran module procedures that share module
 * arrays (src/mod_fields.f90:15-37).  Each entry point below replaces one
 * of those procedures; a Fortran host binds them through ISO_C_BINDING
 * (the modules under atmosphere-model_amd/fortran).  All functions return
 * an atm_status (0 = success);
 * atm_last_error() gives the message of the last failure on a context.
 *
 * Data layout.  Every field crossing this boundary is FP64 in the
 * reference's halo'd Fortran layout: (is:ie, js:je[, k]) column-major with
 * is = 1-7, ie = nx+7 (halo width 7, src/mod_tiles.f90:12), i.e. a
 * contiguous array of (nx+14)*(ny+14)*nk doubles, i fastest.  Layer fields
 * have nk = nlay, level fields nk = nlay+1, surface fields nk = 1.
 *
 * Threading.  Calls on one context are stream-ordered and asynchronous;
 * atm_get_field*, atm_sync, atm_check_stability and atm_radiation block.
 * A context is not thread-safe; separate contexts may be used from
 * separate threads.
 */
#ifndef ATM_H
#define ATM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct atm_ctx atm_ctx;

typedef enum atm_status {
    ATM_OK = 0,
    ATM_ERR_ARGUMENT = 1,      /* bad pointer, size, id or sequence */
    ATM_ERR_HIP = 2,           /* HIP runtime error or missing device */
    ATM_ERR_TABLE_RANGE = 3,   /* gas-optics lookup outside its table
                                  (reference aborts: mod_gas_optics.f90:4562-4592) */
    ATM_ERR_NONFINITE = 4,     /* state failed the stability check */
    ATM_ERR_STATE = 5          /* call out of order (e.g. step before init) */
} atm_status;

typedef enum atm_field {
    ATM_DP = 0,       /* layer pressure thickness dp (nlay)        mod_fields.f90:17 */
    ATM_PT = 1,       /* potential temperature pt (nlay)           mod_fields.f90:18 */
    ATM_UD = 2,       /* D-grid u (nlay)                           mod_fields.f90:19 */
    ATM_VD = 3,       /* D-grid v (nlay)                           mod_fields.f90:20 */
    ATM_TS = 4,       /* surface temperature ts (1)                mod_fields.f90:21 */
    ATM_GZ = 5,       /* geopotential on interfaces gz (nlay+1)    mod_fields.f90:29 */
    ATM_NET_FLUX = 6, /* net radiative flux on interfaces (nlay+1) mod_fields.f90:32 */
    ATM_PLEV = 7,     /* interface pressure plev (nlay+1)          mod_fields.f90:27 */
    ATM_PKAP = 8,     /* plev**kappa (nlay+1)                      mod_fields.f90:28 */
    ATM_RADIUS = 9,   /* distance from the domain centre (1)       mod_fields.f90:37 */
    /* mod_sw_dyn module-private C-grid state (mod_sw_dyn.f90:23-27); the
       reference host never sees these, they are exposed for testing. */
    ATM_DPC = 10, ATM_PTC = 11, ATM_UC = 12, ATM_VC = 13,
    /* tangent and adjoint of dp, pt, ud, vd for the linearised step
       (atm_lin_init); only their interior is read and meaningful */
    ATM_TL_DP = 14, ATM_TL_PT = 15, ATM_TL_UD = 16, ATM_TL_VD = 17,
    ATM_AD_DP = 18, ATM_AD_PT = 19, ATM_AD_UD = 20, ATM_AD_VD = 21,
    ATM_NFIELDS = 22
} atm_field;

typedef struct atm_grid {
    int32_t nx, ny, nlay;      /* interior sizes (mod_config.f90:22) */
    double dx, dy;             /* grid spacing, m */
} atm_grid;

typedef struct atm_params {
    double gravity;                  /* physics_parameters namelist, config.nml:16-21 */
    double coriolis_parameter;
    double divergence_damping_coefficient;
    double vorticity_damping_coefficient;
} atm_params;

/* Arithmetic.  FP64: the drop-in, the reference's operations in double.
   MIXED: the C- and D-step tile kernels form the PPM reconstructions,
   fluxes, kinetic energy, vorticity and damping terms in float and update
   dp, pt and the winds in double; radiation forms its g-point optics in
   float and runs the adding method in double; the radiative heating rate
   (a tendency of pt) and the playkap it divides by are formed in float;
   the column scans' pressures, pkap and geopotential, the pressure
   gradient, remap and physics stay FP64.
   Each face flux is one value shared by the two cells it separates, so mass
   is conserved to FP64 rounding.  MIXED runs on the fused path (any grid);
   it has no counterpart in the reference. */
enum { ATM_PRECISION_FP64 = 0, ATM_PRECISION_MIXED = 1 };

/* Context lifetime.  Allocates every device buffer for one model instance.
   Replaces allocate_fields (mod_fields.f90:41-67) and the module
   allocations of init_sw_dyn / allocate_remapping_arrays /
   allocate_radiation_arrays. */
int atm_create(const atm_grid *grid, const atm_params *params, int precision, atm_ctx **out);
int atm_destroy(atm_ctx *ctx);
const char *atm_last_error(const atm_ctx *ctx);
/* Run on an existing HIP stream (hipStream_t passed as void*); NULL = the
   context's own stream. */
int atm_set_stream(atm_ctx *ctx, void *hip_stream);
/* A HIP stream of default priority from the runtime this library uses, for
   hosts that run several contexts on a few streams (an ensemble), and its
   release.  No reference counterpart (plumbing of this library). */
int atm_stream_create(void **hip_stream);
int atm_stream_destroy(void *hip_stream);
/* Wait for the work enqueued on a stream of this library's runtime; time
   the work enqueued between timer_begin and timer_end with HIP events on
   that stream (ms; timer_end waits for the end event).  One timer per
   calling thread.  For hosts that drive the library from another runtime
   (torch bundles its own HIP runtime, whose streams and events this one
   does not see).  No reference counterpart. */
int atm_stream_synchronize(void *hip_stream);
int atm_stream_timer_begin(void *hip_stream);
int atm_stream_timer_end(void *hip_stream, double *ms);

/* Hybrid coordinate coefficients Ak, Bk (nlay+1 each) -- the reference's
   coord_Ak/coord_Bk read by init_prognostic_fields (mod_input.f90:161-168). */
int atm_set_coordinate(atm_ctx *ctx, const double *ak, const double *bk);

/* Field transfer, host arrays in the halo'd layout described above.  A put
   replaces the whole array (halo included); a get returns exactly what the
   reference's module array would hold at that point. */
int atm_put_field(atm_ctx *ctx, int field, const double *host);
int atm_get_field(atm_ctx *ctx, int field, double *host);
/* Same, device pointers (same layout), stream-ordered. */
int atm_put_field_device(atm_ctx *ctx, int field, const double *dev);
int atm_get_field_device(atm_ctx *ctx, int field, double *dev);

/* init_prognostic_fields tail (mod_input.f90:229-231): plev/pkap top level. */
int atm_init_prognostic_fields(atm_ctx *ctx);
/* initial_halo_exchange (mod_input.f90:247-252): dp, pt, gz(1:nlay), ud, vd, ts. */
int atm_initial_halo_exchange(atm_ctx *ctx);

/* mod_sw_dyn public procedures (mod_sw_dyn.f90:17-19) */
int atm_init_sw_dyn(atm_ctx *ctx);                 /* init_sw_dyn          958-1023 */
int atm_cgrid_step(atm_ctx *ctx, double dt);       /* cgrid_dynamics_step   80-416  */
int atm_cgrid_halo(atm_ctx *ctx);                  /* cgrid_halo_exchange  418-422  */
int atm_dgrid_step(atm_ctx *ctx, double dt);       /* dgrid_dynamics_step  424-773  */
int atm_dgrid_halo(atm_ctx *ctx);                  /* dgrid_halo_exchange  775-779  */

/* mod_remap public procedures (mod_remap.f90:13) */
int atm_remap_init(atm_ctx *ctx);                  /* allocate_remapping_arrays 358-388 */
int atm_remap(atm_ctx *ctx);                       /* perform_vertical_remapping 27-163 */

/* mod_radiation public procedures (mod_radiation.f90:20-21) */
int atm_radiation_init(atm_ctx *ctx);              /* allocate_radiation_arrays 406-446 */
int atm_radiation(atm_ctx *ctx, double time);      /* calculate_radiative_heating 41-156;
                                                      blocks; ATM_ERR_TABLE_RANGE on a
                                                      lookup outside the tables */
int atm_radiation_halo(atm_ctx *ctx);              /* radiation_halo_exchange 400-404 */

/* mod_physics public procedures (mod_physics.f90:13-14) on the device.
   Tendencies are computed when the host's physics interval is due, before
   the C-grid step; they are applied every step after the D-grid step
   (mod_model.f90:137-159). */
int atm_physics_init(atm_ctx *ctx);                /* allocate_physics_arrays 193-214 */
int atm_physics_tendencies(atm_ctx *ctx);          /* calculate_physics_tendencies 27-78
                                                      (dry_convective_adjustment 80-150,
                                                      surface_heating 152-160) */
int atm_apply_physics(atm_ctx *ctx, double dt);    /* apply_physics_tendencies 162-185 */
int atm_physics_halo(atm_ctx *ctx);                /* physics_halo_exchange 187-191 (ts) */

/* mod_output's time-mean accumulation on the device (mod_output.f90:37-160,
   without the netCDF writes).  atm_output_mean returns what write_output
   writes -- output_field / accumulation_time as a (nx, ny, nlay+1, 6)
   Fortran-order array of dp, pt, u, v (A grid), ts, net_flux -- and resets
   the accumulators.  Blocks. */
int atm_output_init(atm_ctx *ctx);                 /* init_output 37-80 (accumulators) */
int atm_accumulate_output(atm_ctx *ctx, double dt); /* accumulate_output 94-130 */
int atm_output_mean(atm_ctx *ctx, double *host);   /* write_output 144-154 */

/* mod_sync generic halo_exchange (mod_sync.f90:14-17) for one field: a
   periodic fill of the 7-wide halo, edges and corners.  For ATM_GZ only
   levels 1..nlay are filled (the 5-argument specific's view of gz). */
int atm_halo_exchange(atm_ctx *ctx, int field);

/* check_stability (mod_fields.f90:69-106): *stable = 1 when dp, pt, ud, vd
   are finite over the interior.  Blocks. */
int atm_check_stability(atm_ctx *ctx, int *stable);

/* The same check, deferred (no counterpart in the reference; for drivers
   of many contexts that keep the GPU fed): _begin queues the check behind
   the calls issued so far and returns at once; _end waits for that check
   only, not for calls issued after _begin, and sets *stable for the state
   at _begin.  One deferred check per context at a time; atm_check_stability
   and atm_graph_begin are refused while one is pending.  Where the fields
   are tracked (every write since the last scan came from a kernel that
   reports non-finite values), a reported non-finite write reads as
   unstable without atm_check_stability's confirming scan. */
int atm_check_stability_begin(atm_ctx *ctx);
int atm_check_stability_end(atm_ctx *ctx, int *stable);

/* Execution path of the dynamics (no counterpart in the reference; both
   give identical results).  ATM_PATH_FUSED (the default, any nx and ny):
   LDS-tiled kernels, four per step; tiles of 32 x 8 and 32 x 16 cells, the
   last of a row or column overhanging the domain when nx or ny is not a
   multiple.  ATM_PATH_REFERENCE: one kernel per reference loop nest, in
   place.  Settable between steps; the environment variable
   ATM_PATH=reference selects it at creation (the context then holds no
   fused buffers and cannot switch to the fused path). */
enum { ATM_PATH_FUSED = 0, ATM_PATH_REFERENCE = 1 };
int atm_set_path(atm_ctx *ctx, int path);
int atm_get_path(atm_ctx *ctx, int *path);

/* Instrumentation (no counterpart in the reference): HIP-event timing of
   the library's kernels on the context stream.  Enabling resets the
   counters; atm_timer_read blocks and returns the number of timed launches
   and their summed duration for one kernel group:
   0 C1, 1 C2, 2 D1, 3 D2, 4 column scans, 5 halo fills, 6 remap, 7 radiation,
   8 physics. */
int atm_timer_enable(atm_ctx *ctx, int on);
int atm_timer_read(atm_ctx *ctx, int kernel, int64_t *launches, double *total_ms);
/* The kernel groups timed while timing is enabled: bit g for group g
   above (default all); fewer groups record fewer events in the stream.
   Resets the counters. */
int atm_timer_select(atm_ctx *ctx, unsigned groups);

/* Read-extent diagnostics (no counterpart in the reference; for tests).
   atm_poison writes `value` (a NaN, say) into device memory that no kernel
   may read for a result inside the domain, in every field buffer of the
   context: ATM_POISON_PADDING, the row padding past ie (the device's row
   stride is nx+14 rounded up to 16) and the tail past the last plane (what
   the overhanging tiles of the fused path read); ATM_POISON_OUTER_HALO, the
   halo cells more than 4 from the interior (no kernel of the fused path
   reads farther for a cell inside; the halo fills rewrite them; the
   reference-structured path evaluates the first phase of a step on the
   ring, as the reference does, from cells farther out).  Stream-ordered; refused
   while recording a graph.  Fields read back through atm_get_field show the
   value in the poisoned outer halo until the next halo fill. */
enum { ATM_POISON_PADDING = 1, ATM_POISON_OUTER_HALO = 2 };
int atm_poison(atm_ctx *ctx, int what, double value);

/* Launch-shape tuning (no counterpart in the reference): atm_autotune times
   each candidate launch shape of the fused path's tunable kernels inside
   whole trial steps (C step, C halo fill, D step) from the context's
   current state into the other copy of its fields, on its grid and stream
   (one warm-up step, then `reps` timed ones), and keeps the fastest; it
   blocks.  The trial steps leave the model where it was: the next real
   step gives the bits it gives without them.  Candidates differ only in
   how the work is divided, never in the values computed.  The knobs, tuned
   one at a time in this order: the rows of the dp/pt tiles (8 or 16,
   default 8) and of the wind tiles (16 or 8, default 16), each 32 cells
   wide, which atm_get_tiles reads and atm_set_tiles sets directly; the
   corner kernel's tile depth, 1, 2, 4 or 8 levels per workgroup (default
   4), which atm_get_tuning reads and atm_set_tuning sets.  Needs
   atm_init_sw_dyn and a step boundary (ATM_ERR_STATE otherwise).  On the
   reference-structured path these return ATM_OK and change nothing.  The
   batched ensemble calls use the default shapes. */
int atm_autotune(atm_ctx *ctx, int reps);
int atm_get_tuning(atm_ctx *ctx, int *corner_levels);
int atm_set_tuning(atm_ctx *ctx, int corner_levels);
int atm_get_tiles(atm_ctx *ctx, int *pair_rows, int *wind_rows);
int atm_set_tiles(atm_ctx *ctx, int pair_rows, int wind_rows);

/* Wait for all work queued on the context. */
int atm_sync(atm_ctx *ctx);

/* Graphs of whole steps (no reference counterpart; for hosts whose steps
   are short enough that launch overhead shows).  Between atm_graph_begin
   and atm_graph_end the calls atm_cgrid_step, atm_cgrid_halo,
   atm_dgrid_step, atm_dgrid_halo, atm_remap, atm_physics_tendencies,
   atm_apply_physics, atm_physics_halo, atm_halo_exchange,
   atm_radiation_halo and atm_accumulate_output are recorded instead of run;
   every call that waits on the device or copies host memory returns
   ATM_ERR_STATE meanwhile.  The recording must end where it started: at a
   step boundary and, on the fused path, with the same current copy of the
   double-buffered fields (an even number of D steps + remaps), else
   atm_graph_end refuses it.  atm_graph_launch runs the recording `times`
   times on the context's stream; the dt values are those recorded, and
   each launch advances the output accumulation time by the recorded dt
   values.  A launch must find the copy of the fields that was current when
   the recording began (an even number of eager D steps + remaps since),
   else it returns ATM_ERR_STATE.  One recording per context; a new
   atm_graph_begin replaces it. */
int atm_graph_begin(atm_ctx *ctx);
int atm_graph_end(atm_ctx *ctx);
int atm_graph_launch(atm_ctx *ctx, int times);
int atm_graph_free(atm_ctx *ctx);

/* Batched ensemble steps (no reference counterpart; for ensembles of
   members with their own grids).  An ensemble binds up to 64 contexts that
   run the fused path in one precision on one device; each call below does
   for every member what the single-context call does (atm_cgrid_step,
   atm_cgrid_halo, atm_dgrid_step, atm_dgrid_halo: same checks, same values
   bit for bit, same host-side state), with one kernel launch per stage for
   all members together (a member table indexed by workgroup) on the first
   member's stream.  Members on other streams are joined by events: each
   call waits for the work already issued on every member's stream, and
   every member's later work waits for the call.  Members may differ in nx,
   ny and nlay (nlay <= 64); they share dt.  Physics stays per member.  Errors are reported on the member concerned
   (atm_last_error of that context), else on the first.  The ensemble holds
   pointers to the members' device arrays: destroy it before any member. */
typedef struct atm_ensemble atm_ensemble;
int atm_ensemble_create(atm_ctx *const *members, int n, atm_ensemble **out);
int atm_ensemble_destroy(atm_ensemble *ens);
int atm_ensemble_cgrid_step(atm_ensemble *ens, double dt);
int atm_ensemble_cgrid_halo(atm_ensemble *ens);
int atm_ensemble_dgrid_step(atm_ensemble *ens, double dt);
int atm_ensemble_dgrid_halo(atm_ensemble *ens);
/* perform_vertical_remapping (mod_remap.f90:27-163) of every member after
   atm_ensemble_dgrid_step, as atm_remap does for one context, with one
   launch per remap stage for all members; needs every member's
   atm_remap_init before atm_ensemble_create. */
int atm_ensemble_remap(atm_ensemble *ens);
/* calculate_radiative_heating (mod_radiation.f90:41-156) of every member,
   as atm_radiation does for one context, with one launch per stage for all
   members and one wait for their lookups' range flags (ATM_ERR_TABLE_RANGE
   on the first member concerned); needs every member's atm_radiation_init
   before atm_ensemble_create.  The halo fill of the net flux stays per
   member (atm_radiation_halo). */
int atm_ensemble_radiation(atm_ensemble *ens, double time);

/* Graphs of batched steps (no reference counterpart).  Between
   atm_ensemble_graph_begin and atm_ensemble_graph_end the batched step
   calls above (atm_ensemble_radiation excepted: it waits on the device)
   and the members' own step calls are recorded into one HIP graph instead
   of run: the first member's stream starts the
   recording and the other member streams join it.  The rules of
   atm_graph_begin / end / launch hold for every member (whole steps, an
   even number of buffer switches, no deferred check pending; each launch
   applies every member's recorded host-side effects).  Each
   atm_ensemble_graph_launch follows the work already issued on every
   member's stream and precedes what is issued there afterwards.  One
   recording per ensemble; a new atm_ensemble_graph_begin replaces it, and
   atm_ensemble_destroy frees it.  Members may not record graphs of their
   own meanwhile. */
int atm_ensemble_graph_begin(atm_ensemble *ens);
int atm_ensemble_graph_end(atm_ensemble *ens);
int atm_ensemble_graph_launch(atm_ensemble *ens, int times);

/* Tangent-linear and adjoint of one model step (no reference counterpart;
   for gradient computations).  The step is the reference's sequence
   cgrid_dynamics_step(dt/2), cgrid_halo_exchange, dgrid_dynamics_step(dt),
   perform_vertical_remapping when `remap` is nonzero, dgrid_halo_exchange
   (src/mod_model.f90:144-173) with radiation and physics off (the net flux
   is a fixed forcing), linearised about the current state with every PPM
   limiter branch and upwind switch frozen as the state takes them.  The
   state is the interior of dp, pt, ud, vd at a step boundary; the
   linearised kernels form the forward values themselves (the FP64 path).
     atm_lin_init  allocates the linearised step's arrays (needs
                   atm_init_sw_dyn).
     atm_tl_step   advances the state by one step and the tangent fields
                   (ATM_TL_*) by its tangent-linear model about the state
                   at the step's start.
     atm_ad_step   replaces the adjoint fields (ATM_AD_*: the adjoint of the
                   step's end state) with the adjoint of the start state,
                   linearised about the current state, which it leaves
                   unchanged.
   Both refuse while recording a graph or between atm_dgrid_step and
   atm_dgrid_halo. */
int atm_lin_init(atm_ctx *ctx);
int atm_tl_step(atm_ctx *ctx, double dt, int remap);
int atm_ad_step(atm_ctx *ctx, double dt, int remap);

/* General two-stream columns.
   Replaces the column loop of calculate_radiative_heating
   (mod_radiation.f90:117-154) with solve_shortwave_column (158-268),
   solve_longwave_column (270-347), calculate_fluxes (349-381) and
   delta_eddington_rescaling (383-398) over inputs the reference fixes:
   per-layer cloud optical depth, single-scattering albedo and asymmetry
   mixed with the gas and Rayleigh optical depths, per-column cosine of the
   solar zenith angle, surface albedo, surface temperature and irradiance,
   and gas optics from an analytic k-distribution, per g-point g:
     ln k = c0 + c1 x + c2 x^2 + c3 y + c4 y^2,  x = ln(p / p_ref),
     y = T - t_ref,  tau_gas = k q dm  (k in m^2 per kg of water vapour,
     q the layer's mixing ratio, dm = (p_lower - p_upper) / gravity);
   shortwave Rayleigh tau = k_ray dm; Planck source f_g sigma T^4 / pi
   (sigma 5.670374419e-8, pi the binary32 value); surface emissivity 0.99.
   The reference's column formulas are kept as written, with its own
   properties: alpha2 = gamma2 gamma3 + gamma2 gamma4 (214), the diffuse R/T
   with the direct beam's e^(-lambda tau) (254-260), the top's diffuse
   downward flux equal to the incident beam (264), no surface reflection of
   the direct beam.  guard: where |1 - (lambda mu0)^2| < guard, lambda is
   scaled by (1 - 2 guard) for that layer (every term), moving lambda mu0
   off the removable singularity of the direct-beam R/T.
   Layout: every array is FP64 device memory, [column][level] with the level
   index contiguous; level 0 is the surface, layer k lies between levels k
   and k + 1.  optics: n_sw x 5 shortwave coefficients (c0..c4), n_lw x 5
   longwave, n_sw Rayleigh coefficients (m^2 per kg of air), n_sw solar
   fractions, n_lw Planck fractions (6 n_sw + 6 n_lw + n_sw values). */
typedef struct atm_ts_columns {
    int64_t ncol;
    int32_t nlay, n_sw, n_lw;          /* 1 <= nlay <= 256; n_sw, n_lw >= 0 */
    double p_ref, t_ref, gravity, guard;
    const double *optics;
    const double *play, *tlay, *h2o;                 /* [ncol][nlay] */
    const double *cloud_tau, *cloud_ssa, *cloud_asy; /* [ncol][nlay] */
    const double *plev, *tlev;                       /* [ncol][nlay + 1] */
    const double *mu0, *sfc_alb, *tsfc, *irradiance; /* [ncol] */
} atm_ts_columns;

/* Column diagnostics: shortwave layers (x g-points) where the guard moved
   lambda; columns with an input outside the solver's domain (non-positive
   or non-finite optical depth, no absorption in a layer (omega = 1, where
   the reference's R/T divide 0 by 0), mu0 outside (0, 1], non-positive
   pressure thickness), whose fluxes are then not meaningful. */
typedef struct atm_ts_stats {
    int64_t guarded_layers, bad_columns;
} atm_ts_stats;

/* net_flux [ncol][nlay + 1] = (sw down + lw down) - (sw up + lw up), each
   summed over its g-points in order; fluxes (optional, NULL for none):
   4 x [ncol][nlay + 1] shortwave up, shortwave down (direct beam
   included), longwave up, longwave down.  precision: ATM_PRECISION_FP64
   (the reference's arithmetic) or ATM_PRECISION_MIXED (g-point optics in
   float, the longwave sources in a rearranged form that avoids the
   cancellation of 1 + R - T for thin layers; beam, adding method and sums
   in double).  Runs on hip_stream (NULL: the default stream; a stream of
   this library's runtime, see atm_stream_create).  stats NULL:
   asynchronous, allocates nothing.  stats non-NULL: allocates a per-column
   status buffer (4 ncol + 64 bytes), waits for the stream and fills stats.
   Arrays must be device memory: pointers this runtime identifies as host
   memory are refused, others are taken as device addresses (memory of
   another runtime in the process, such as torch's, is not known to this
   one).  Errors are reported by atm_twostream_last_error (per thread). */
int atm_twostream(const atm_ts_columns *cols, int precision, double *net_flux, double *fluxes,
                  atm_ts_stats *stats, void *hip_stream);
/* atm_twostream without any allocation, for callers that cannot allocate
   (code under HIP graph capture, serving loops, hosts that forbid
   hipMalloc).  No counterpart in the reference.  Same kernels and results;
   the per-column status words and the totals live in a static device array
   of the library (4 MiB + 256 B per GPU, held until the process ends), so
   ncol is limited to 1048576 (larger: ATM_ERR_ARGUMENT).  Every call on a
   GPU shares that array: calls that overlap on one GPU race.  stats NULL:
   asynchronous (no copy, no wait); non-NULL: waits for the stream and
   copies the totals. */
int atm_twostream_noalloc(const atm_ts_columns *cols, int precision, double *net_flux, double *fluxes,
                          atm_ts_stats *stats, void *hip_stream);
const char *atm_twostream_last_error(void);

/* Library build information (kernel path, compile flags), for logs. */
const char *atm_build_info(void);

#ifdef __cplusplus
}
#endif
#endif
