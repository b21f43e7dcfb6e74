/*
 * ddyn_oracle.h -- FP64 CPU oracle for DiskDyn's dynamical hot path.
 * TEST INFRASTRUCTURE ONLY (the oracle of the tests, never the product).
 */
#ifndef DDYN_ORACLE_H
#define DDYN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physics switches and constants of one run (units AU, yr, Msun; mu = G*M). */
typedef struct {
    int32_t n_planets;   /* star + planets, as cfg->n_planets (src/read_in.cu:63-79) */
    int32_t SWQ;         /* stellar-wind drag on (record 10) */
    int32_t FLQ;         /* Lorentz force on (record 10) */
    int32_t pad_;
    double  eps;         /* Plummer softening [AU] on swarm-sourced pairs; 0 = reference law */
    double  Rmin;        /* sublimation radius [AU] */
    double  vsw;         /* stellar-wind speed [AU/yr] (grid.cu:737 passes vsw*year/AU) */
    double  Wz;          /* stellar rotation [1/yr] */
    double  Bstar;       /* stellar field, reference units (read_in.cu:597) */
    double  Rstar;       /* stellar radius [AU] */
    double  light;       /* speed of light [AU/yr] (lightAU, ddyn.cuh:22) */
} orc_params;

/* Particle state in the reference order [star | planets | swarm | dust]. */
typedef struct {
    int64_t n_grav, n_dust;
    double *x, *y, *z, *vx, *vy, *vz, *mu, *N;   /* n_grav + n_dust each */
    const int32_t *id;                            /* size-bin id (may be NULL: all -1) */
    int32_t n_bins;
    int32_t pad_;
    const float *beta_tab, *betasw_tab, *q_tab;   /* per-bin tables (may be NULL) */
    const double *beta_p, *betasw_p;              /* per-grain overrides (may be NULL) */
} orc_state;

void orc_massive_accel(const orc_params *p, int64_t n_grav,
                       const double *x, const double *y, const double *z, const double *mu,
                       double *ax, double *ay, double *az);
void orc_massive_accel_targets(const orc_params *p, int64_t n_grav,
                               const double *x, const double *y, const double *z, const double *mu,
                               int64_t nt, const int64_t *idx, double *ax, double *ay, double *az);
int  orc_dust_accel(const orc_params *p, const orc_state *s, double *ax, double *ay, double *az,
                    int32_t *removed);
int  orc_step(const orc_params *p, orc_state *s, double dt, int64_t nsteps);
/* Discrete adjoint of nsteps RK4 steps from the state s (not modified).
 * lam_r, lam_v: dL/d(final r, v), 3*n_all interleaved (x,y,z per particle).
 * Outputs: g_r, g_v (3*n_all) = dL/d(initial r, v); g_mu (n_grav) = dL/dmu;
 * g_beta (n_dust, may be NULL) = dL/dbeta.  Returns -2 if FLQ is on. */
int  orc_adjoint(const orc_params *p, const orc_state *s, double dt, int64_t nsteps,
                 const double *lam_r, const double *lam_v, double *g_r, double *g_v,
                 double *g_mu, double *g_beta);
void orc_energy(const orc_params *p, int64_t n, const double *x, const double *y, const double *z,
                const double *vx, const double *vy, const double *vz, const double *mu,
                double *out7);
void orc_coe2rv(double M, double m, double a, double e, double inc, double w, double o,
                double nu, double beta, double *out6);
void orc_recalc_com(int64_t n_planets, double *x, double *y, double *z,
                    double *vx, double *vy, double *vz, const double *mu);
/* Images and SEDs (src/gpu_functions.cu:983-1156): the dust grid's optical
 * tables, in the layouts of the reference's device arrays. */
typedef struct {
    int32_t n_bins, n_wav, n_theta, n_r;
    const double *w, *in, *size, *Qabs, *Qsca, *Qpfunc, *Td;   /* [wav], [wav], [bin], [bin][wav] x2,
                                                                 [bin][theta][wav], [bin][radius] */
    double Rmin, Rmax;                                          /* [AU] */
} orc_optics;
typedef struct {
    double dist, rot, incl, PA, fov;   /* [m], [rad] x3, [AU] */
    int32_t xsize, ysize;
} orc_view;
/* Image at wavelength index wj: F (xsize*ysize, pixel (x, y) at F[ysize*x + y])
 * is zeroed, then every grain's flux is added to its pixel in grain order. */
void orc_image(const orc_optics *o, const orc_view *v, int64_t n_grav, int64_t n_dust, const double *x,
               const double *y, const double *z, const double *N, const int32_t *id, int32_t wj, double *F);
/* SED: Fexc (n_wav) = the sum over grains, in grain order, of every grain's flux. */
void orc_sed(const orc_optics *o, const orc_view *v, int64_t n_grav, int64_t n_dust, const double *x,
             const double *y, const double *z, const double *N, const int32_t *id, double *Fexc);
/* Mie efficiencies and phase function of a homogeneous sphere of radius s [um]
 * at nl wavelengths lam [um] with complex index n - i k (tools/mie_single.c):
 * Qabs, Qsca (nl) and Qpfunc[nl * t + i] at the nt cosines cost (phase function
 * normalised to 1 over the sphere).  Returns 0, or -1 if a size parameter
 * exceeds 7.2e6. */
int  orc_mie(double s, const double *lam, const double *nre, const double *kim, int nl, const double *cost, int nt,
             double *Qabs, double *Qsca, double *Qpfunc);
int  orc_max_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
