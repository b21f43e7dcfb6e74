/*
 * ddyn.h -- C ABI of the MI355X-native DiskDyn dynamical hot path.
 *
 * The reference drives this path through three Grid member calls made once
 * per time step by Grid::evolve (src/grid.cu:387-389):
 *     massive_RK4(dt)   src/grid.cu:693   four massive-body RK4 stages
 *     small_RK4(dt)     src/grid.cu:729   fused four-stage grain RK4
 *     massive_final(dt) src/grid.cu:742   massive-body RK4 combination
 * on the device SoA `coord gpu` (src/ddyn.cuh:44-58) that Grid allocates and
 * fills in allocgpu/sync_to_gpu/sync_dust_to_gpu (src/init.cu:543-667) and
 * reads back in sync_from_gpu (src/init.cu:670-682).  The reference has no
 * plugin or FFI layer; the functions below replace that member-call boundary
 * one for one, with plain pointers and sizes, so a C, C++, Python (ctypes) or
 * any other FFI caller can drive the path.
 *
 * Units follow the reference: AU, yr, Msun, and every mass is stored as
 * mu = G*M with G = 4*pi^2 (src/ddyn.cuh:30, src/read_in.cu:76).  Particle
 * order in every array is [star (0) | planets | swarm | dust], exactly as in
 * the reference (n_grav = star + planets + swarm, src/grid.cu:696).
 *
 * Conventions: functions return DDYN_OK (0) or a negative DDYN_ERR_* code and
 * never exit the process; ddyn_last_error() gives the message.  After a
 * DDYN_ERR_HIP from ddyn_step or ddyn_adjoint the device state is undefined:
 * upload a state again before stepping.  The caller owns every host array
 * (copied in upload/download); the context owns all device memory and two HIP
 * streams (the main stream, and a grain stream used by the default schedule
 * of ddyn_step).  A context is not thread safe.  ddyn_step() is asynchronous
 * on the main stream (which waits for the grain stream before the call
 * returns); download, accel and energy synchronise.
 */
#ifndef DDYN_H
#define DDYN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDYN_OK 0
#define DDYN_ERR_ARG (-1)     /* bad argument or size */
#define DDYN_ERR_HIP (-2)     /* HIP runtime failure */
#define DDYN_ERR_STATE (-3)   /* call out of order (e.g. step before upload) */
#define DDYN_ERR_NOMEM (-4)   /* device allocation failed */

/* Force precision forms. */
#define DDYN_FP64 0           /* every term in FP64 (the plain form) */
#define DDYN_FP32_FORCE 1     /* swarm-sourced pairs in FP32 with FP64 per-tile flushes;
                                 star/planet terms, close pairs and RK4 state in FP64 */
#define DDYN_FP64_RSQ 2       /* every term in FP64; every pair law takes 1/R^3 (and the grain
                                 star term 1/R^2) from the v_rsq_f64 estimate refined by one
                                 third-order step (relative error at most 1.4e-16) instead of
                                 an IEEE square root and division */

typedef struct ddyn_ctx ddyn_ctx;

/* Physics switches and constants.  Replaces the cconfig globals and the
 * scalars that small_RK4 passes to RK4calc_dust (src/grid.cu:737). */
typedef struct {
    int32_t n_planets;   /* star + planets (cfg->n_planets counts the star, src/read_in.cu:63-79) */
    int32_t SWQ;         /* 1: stellar-wind drag on (.param record 10) */
    int32_t FLQ;         /* 1: Lorentz force on (.param record 10) */
    int32_t precision;   /* DDYN_FP64, DDYN_FP32_FORCE or DDYN_FP64_RSQ */
    double  eps;         /* Plummer softening [AU] for swarm-sourced pairs; 0 = reference law (none) */
    double  Rmin;        /* grain sublimation radius [AU] (Grid::Rmin, src/grid.cu:737) */
    double  vsw;         /* stellar-wind speed [AU/yr] (cfg->vsw*year/AU, src/grid.cu:737) */
    double  Wz;          /* stellar rotation rate [1/yr] (cfg->Wz) */
    double  Bstar;       /* stellar surface field in the reference's units (cfg->Bstar) */
    double  Rstar;       /* stellar radius [AU] (cfg->rstar) */
    double  light;       /* speed of light [AU/yr]; 0 selects lightAU = 63239.6717367 */
    double  fp64_radius; /* DDYN_FP32_FORCE: swarm pairs closer than this [AU] are evaluated in
                            FP64; <= 0 selects 10*eps.  With a swarm and eps = 0 (the
                            reference law) it must be > 0, else ddyn_step, ddyn_accel and
                            the adjoint return DDYN_ERR_ARG: the FP64 redo is what evaluates
                            close pairs and detects exact hits */
    int32_t schedule;    /* 0: fuse many steps per launch where the system allows it (star +
                            planets only, <= 32 massive bodies), and with grains run the grain
                            kernel of step k beside the massive stages of step k+1 on a second
                            stream; 1: one launch per kernel, in order, on one stream.
                            Both give identical results.  In the adjoint with grains and
                            more than 64 massive bodies, 0 also sums the grains' pulls on
                            the massive bodies as a pair kernel, 1 per grain wave: those
                            gradients agree to rounding, not bit for bit. */
    int32_t reserved;    /* must be 0 */
} ddyn_params;

/* Create a context on HIP device `device` (replaces checkcuda, src/cuda.cu:92-98). */
int  ddyn_create(ddyn_ctx **out, int device);
/* Free all device memory (replaces the Grid destructor frees, src/init.cu:111-215). */
void ddyn_destroy(ddyn_ctx *ctx);
/* Message of the last failure on this context ("" if none). */
const char *ddyn_last_error(const ddyn_ctx *ctx);

/* Set physics and precision (replaces the cconfig fields read by the kernels). */
int ddyn_set_params(ddyn_ctx *ctx, const ddyn_params *p);

/* Upload the particle state (replaces allocgpu + null_gpu + sync_to_gpu,
 * src/init.cu:543-637).  Arrays have n_grav + n_dust entries in reference
 * order; mu = G*M; N is the tracer count (0 = removed, skipped forever,
 * src/gpu_functions.cu:643); id is the grain size-bin index (-1 = none;
 * may be NULL for all -1). */
int ddyn_upload_state(ddyn_ctx *ctx, int64_t n_grav, int64_t n_dust,
                      const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz,
                      const double *mu, const double *N, const int32_t *id);

/* Per-size-bin beta, beta_sw and charge tables, float as in the reference's
 * __constant__ arrays (src/ddyn.cu:30-32, filled by sync_dust_to_gpu,
 * src/init.cu:639-667).  n_bins <= 2700 (nsize, src/ddyn.cuh:35). */
int ddyn_upload_bin_tables(ddyn_ctx *ctx, int32_t n_bins,
                           const float *beta, const float *betasw, const float *q);

/* Per-grain FP64 beta and beta_sw (n_dust each), overriding the bin tables.
 * No reference counterpart: the reference only has per-bin float tables.
 * betasw may be NULL (zeros).  beta == NULL removes the override. */
int ddyn_upload_grain_beta(ddyn_ctx *ctx, const double *beta, const double *betasw);

/* Advance nsteps RK4 steps of size dt [yr]; one step is
 * massive_RK4(dt) + small_RK4(dt) + massive_final(dt) (src/grid.cu:387-389). */
int ddyn_step(ddyn_ctx *ctx, double dt, int64_t nsteps);

/* One force evaluation at the current state (stage-1 forces of the next step):
 * ax/ay/az get n_grav + n_dust accelerations [AU/yr^2]; removed (may be NULL)
 * gets n_dust flags (1 = the grain would be removed: sublimation or exact hit);
 * a removed grain's acceleration is reported as zero.
 * No reference counterpart (the reference never exposes accelerations). */
int ddyn_accel(ddyn_ctx *ctx, double *ax, double *ay, double *az, int32_t *removed);

/* ddyn_accel without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No reference counterpart.  The inputs and outputs are the
 * caller's DEVICE arrays, not the context's state: x..vz, mu (n_grav + n_dust
 * each, reference order), id (n_grav + n_dust, may be NULL: every grain without
 * a bin), beta_p and betasw_p (n_dust, may be NULL: the context's bin tables),
 * ax, ay, az (n_grav + n_dust) and removed (n_dust, may be NULL).  Uses the
 * context's parameters, bin tables and main stream; the stage-0 positions and
 * the pair kernels' slice partials live in a static 32 MiB device array of the
 * library (per device, held until the process ends), which limits n_grav to
 * about 466,000 (DDYN_ERR_ARG beyond).  Asynchronous: no allocation, no copy
 * and no synchronisation; call ddyn_synchronize (or wait on the stream) before
 * reading the outputs.  Every call on a device shares that device's array, so
 * calls that overlap on one device (two contexts, two host threads) race.
 * Results equal ddyn_accel's for the same state up to the order of the slice
 * sums (bit for bit when the pair split is the same). */
int ddyn_accel_noalloc(ddyn_ctx *ctx, int64_t n_grav, int64_t n_dust, const double *x, const double *y,
                       const double *z, const double *vx, const double *vy, const double *vz,
                       const double *mu, const int32_t *id, const double *beta_p, const double *betasw_p,
                       double *ax, double *ay, double *az, int32_t *removed);

/* Download the state (replaces sync_from_gpu, src/init.cu:670-682).  Unlike
 * the reference, N is downloaded too so removals are visible.  Any pointer may
 * be NULL to skip that array. */
int ddyn_download_state(ddyn_ctx *ctx, double *x, double *y, double *z,
                        double *vx, double *vy, double *vz, double *N);

/* Invariants of the massive set in mu units: out7 = {E, Lx, Ly, Lz, Px, Py, Pz}
 * with E = sum 0.5 mu v^2 - sum_{i<j} mu_i mu_j / sqrt(r^2 + eps_ij^2).
 * No reference counterpart. */
int ddyn_energy(ddyn_ctx *ctx, double *out7);

/* Discrete adjoint of nsteps RK4 steps of size dt from the current state.
 * No reference counterpart: the reference has no differentiation.
 * Runs the forward pass (the context state ends at the state after nsteps, as
 * ddyn_step would leave it) keeping a checkpoint every ckpt_every steps
 * (<= 0: one checkpoint), then reverses every step, recomputing each segment
 * from its checkpoint.  Arrays are SoA blocks: lam_r = [x | y | z] of n_all
 * each = dL/d(final positions), lam_v likewise for velocities.  Outputs:
 * g_r, g_v (3*n_all) = dL/d(initial positions, velocities); g_mu (n_grav) =
 * dL/d mu of every massive body (direct force dependence); g_beta (n_dust,
 * may be NULL) = dL/d beta per grain (meaningful with ddyn_upload_grain_beta).
 * Grains removed during the run are treated as frozen.  The Lorentz force
 * (FLQ) is not supported.  Stage forces are recomputed in the context's
 * precision for massive bodies and in FP64 for grains; the Jacobian products
 * are FP64. */
int ddyn_adjoint(ddyn_ctx *ctx, double dt, int64_t nsteps, int64_t ckpt_every,
                 const double *lam_r, const double *lam_v, double *g_r, double *g_v,
                 double *g_mu, double *g_beta);

/* ddyn_adjoint in two calls, for a loss that needs the final state before its
 * gradient lam_r, lam_v is known.  No reference counterpart.
 * ddyn_adjoint_record runs the forward pass from the current state (the
 * context ends at the state after nsteps, as ddyn_step would leave it) and
 * keeps a tape on the device: a checkpoint every ckpt_every steps (<= 0: one
 * segment), and every step's state and stage positions of the last segment,
 * 56 B per particle plus 96 B per massive body per step of it.
 * ddyn_adjoint_reverse then gives what ddyn_adjoint gives for the recorded run,
 * bit for bit, without a second forward pass (earlier segments are recomputed
 * from their checkpoints, the last one is on the tape); it ends at the final
 * state and keeps the tape, so it may be called again with another lam.
 * ddyn_step, the uploads and ddyn_set_params drop the tape; reversing without
 * one returns DDYN_ERR_STATE. */
int ddyn_adjoint_record(ddyn_ctx *ctx, double dt, int64_t nsteps, int64_t ckpt_every);
int ddyn_adjoint_reverse(ddyn_ctx *ctx, const double *lam_r, const double *lam_v,
                         double *g_r, double *g_v, double *g_mu, double *g_beta);

/* Wait for all work of the context (the main stream, which every call joins). */
int ddyn_synchronize(ddyn_ctx *ctx);

/* Kernel timing with HIP events around each launch (on the stream it runs on).  Classes: */
#define DDYN_KCLASS_PAIRS 0   /* massive swarm-pair kernel */
#define DDYN_KCLASS_STAGE 1   /* massive stage epilogue (direct terms + RK4 update) */
#define DDYN_KCLASS_DUST 2    /* fused four-stage grain kernel; also the star-and-planets
                                 multi-step kernel (whole steps per launch) */
#define DDYN_KCLASS_COUNT 3
int ddyn_set_timing(ddyn_ctx *ctx, int enable);
/* Totals since timing was enabled (synchronises). */
int ddyn_read_timing(ddyn_ctx *ctx, int32_t kclass, int64_t *launches, double *total_ms);

/* The source-slice split of the massive pair kernels for the uploaded state:
 * nslice slices of slice_len swarm sources (0 and 0 without a swarm).  Each
 * slice adds 3 FP64 partials per massive body per stage.  Needs parameters
 * and a state.  No reference counterpart (measurement support). */
int ddyn_pair_plan(ddyn_ctx *ctx, int32_t *nslice, int32_t *slice_len);

/* ---------------------------------------------------------------------------
 * Images and SEDs of the grains' thermal emission and scattered starlight
 * (src/gpu_functions.cu:945-1156, driven by Grid::null_flux, calc_flux,
 * calc_flux_sed and copy_sed, src/grid.cu:12-65).
 * ------------------------------------------------------------------------- */

/* The dust grid's optical tables (the reference's Qabs, Qsca, Qpfc, Td, w, in
 * device arrays and the size table, uploaded in sync_to_gpu,
 * src/init.cu:595-664; ddyn_grid_optics_tables in ddyn_host.h fills one from a
 * .param run's optical data).  Bin b is the grains' size-bin id. */
typedef struct {
    int32_t n_bins;        /* size bins */
    int32_t n_wav;         /* wavelengths */
    int32_t n_theta;       /* scattering-angle points, cos from -1 to 1 */
    int32_t n_r;           /* radii of the temperature table (nR = 100 in the reference) */
    const double *w;       /* n_wav wavelengths [m], increasing */
    const double *in;      /* n_wav stellar flux at the observer [W m^-3] */
    const double *size;    /* n_bins grain radii [m] */
    const double *Qabs;    /* n_bins * n_wav absorption efficiencies, [bin][wav] */
    const double *Qsca;    /* n_bins * n_wav scattering efficiencies, [bin][wav] */
    const double *Qpfunc;  /* n_bins * n_theta * n_wav phase function, [bin][theta][wav] */
    const double *Td;      /* n_bins * n_r grain temperature [K] on n_r radii log-spaced
                              from Rmin to Rmax, [bin][radius] */
    double Rmin, Rmax;     /* radial range of Td [AU]; grains within Rmin do not emit */
} ddyn_optics_tables;

/* Observer and detector (the reference's cfg->dist, longitude, inc, PA, xsize,
 * ysize and fov, .param record 7). */
typedef struct {
    double dist;           /* distance [m] */
    double rot;            /* rotation of the system about its z axis [rad] (longitude) */
    double incl;           /* inclination [rad] */
    double PA;             /* position angle [rad] */
    double fov;            /* field of view [AU] across the larger image side */
    int32_t xsize, ysize;  /* pixels */
} ddyn_view;

/* Copy the optical tables to the device (replaces their cudaMemcpy's in
 * sync_to_gpu, src/init.cu:639-664).  n_bins <= 2700, n_wav and n_theta >= 2,
 * n_r >= 2, n_bins * n_theta * n_wav <= 2^31, 0 < Rmin < Rmax. */
int ddyn_upload_optics(ddyn_ctx *ctx, const ddyn_optics_tables *t);

/* Image at wavelength index wj of the current state [Jy per pixel] (replaces
 * null_flux + calc_flux and the download in write_image, src/grid.cu:12-54,
 * src/fits.cu:21): F gets xsize * ysize values in the reference's layout,
 * pixel (x, y) at F[ysize * x + y].  A grain with a size bin in the tables,
 * outside Rmin and projected inside the field adds N (thermal + scattered)
 * flux to its pixel.  Unlike the reference's atomic adds, each pixel sums its
 * grains in grain order, so images are deterministic.  xsize * ysize <= 2^30.
 * Synchronises. */
int ddyn_calc_flux(ddyn_ctx *ctx, const ddyn_view *v, int32_t wj, double *F);

/* SED of the grains [Jy] at every wavelength of the tables (replaces
 * null_flux_part + calc_flux_sed + copy_sed per wavelength, src/grid.cu:21-41,
 * 56-65, 366-373): Fexc gets n_wav totals over all grains, summed in a fixed
 * order (deterministic).  As in the reference's SED kernel the system is
 * turned about z with the grain's 3-D distance, not its cylindrical one
 * (src/gpu_functions.cu:1104-1105); xsize, ysize and fov are not used.
 * Synchronises. */
int ddyn_calc_sed(ddyn_ctx *ctx, const ddyn_view *v, double *Fexc);

/* Mie efficiencies and phase functions of homogeneous spheres: the tables
 * the reference's tools/runmie computes with mie_single (tools/mie_single.c,
 * tools/runmie.c), in FP64 on the device.  For every radius s_um[i] (n_sizes,
 * microns) and wavelength lam_um[j] (n_wav, microns) with complex refractive
 * index n_re[j] - i k_im[j]: Qabs, Qsca [i * n_wav + j] and the phase function
 * Qpfunc [(i * n_theta + t) * n_wav + j] at the n_theta cosines cost[t] of the
 * scattering angle, normalised to 1 over the sphere (the layout of a
 * composition table, ddyn_optics_read).  n_theta <= 256, size parameters
 * 2 pi s / lambda <= 7.2e6 (the reference's limit), n_sizes * n_wav <= 2^30,
 * finite n_re > 0 and k_im >= 0.
 * Uses the context's device and main stream only (no state needed);
 * synchronises.  Negative efficiencies are returned as computed (the tool
 * clamps them to 0 when writing). */
int ddyn_mie(ddyn_ctx *ctx, int32_t n_sizes, const double *s_um, int32_t n_wav, const double *lam_um,
             const double *n_re, const double *k_im, int32_t n_theta, const double *cost, double *Qabs,
             double *Qsca, double *Qpfunc);

/* y[i] = the library's FP64 reciprocal square root of x[i] > 0, evaluated on the
 * device (the v_rsq_f64 estimate refined by one third-order step, used by every
 * pair law of DDYN_FP64_RSQ and by the adjoint's Jacobian products).  n <= 2^30;
 * host arrays; synchronises.  No reference counterpart (accuracy check). */
int ddyn_eval_rsqrt_fp64(ddyn_ctx *ctx, int64_t n, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
