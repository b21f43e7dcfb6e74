/*
 * ddyn_host.h -- host-side envelope of the dynamical hot path (libddyn_host.so).
 *
 * Pure C++ (no GPU): the .param reader, the time-step rule, the data-file
 * writer/reader, with the record layout of the reference so files move both
 * ways between the two codes, and the set-up of a .param run (optical data,
 * dust grid, grain tables, particle generation).
 *   ddyn_read_param      replaces read_in_param      (src/read_in.cu:508-733)
 *                        with add_star, add_planets, add_disks, add_blobs and
 *                        recalc_COM (src/read_in.cu:63-234, 344-375)
 *   ddyn_det_step        replaces Grid::det_step     (src/grid.cu:312-342)
 *   ddyn_write_datafile  replaces Grid::writegrid    (src/grid.cu:441-673)
 *   ddyn_read_datafile   replaces read_in_model + Grid::read_in
 *                        (src/read_in.cu:736-1014, src/grid.cu:205-310)
 *   ddyn_write_image     replaces Grid::write_image  (src/fits.cu:9-85)
 *   ddyn_write_sed       replaces Grid::write_sed    (src/grid.cu:675-691)
 * Functions return 0 on success or the reference's error number (the index of
 * its message table, src/error.cu:5-64); ddyn_host_error_text() gives a text.
 * Units as in the reference's configdata after parsing.
 */
#ifndef DDYN_HOST_H
#define DDYN_HOST_H
#include <stdint.h>

#include "ddyn.h"   /* ddyn_optics_tables */

#ifdef __cplusplus
extern "C" {
#endif

#define DDYN_MAX_PLANETS 64
#define DDYN_MAX_DISKS 16
#define DDYN_MAX_BLOBS 16
#define DDYN_MAX_WAVS 64
#define DDYN_STR 256

typedef struct {
    double Rin, Rout, dhdr, sigmae, amax, amin, eta, Mtot, mmax, mmin, gammas;
    double dmax, dmmax, C;           /* optics-dependent; 0 until computed */
    int64_t Ngrav, Nsmall;
} ddyn_disk;

typedef struct {
    double a, ecc, inc, w, o, nu, sigma, amax, amin, eta, Mtot, mmax, mmin, vkepp;
    double dmax, dmmax, C;
    int64_t Ngrav, Nsmall;
    int32_t collcloud, pad_;
} ddyn_blob;

typedef struct {
    int32_t verb, gpuid, gl, n_step;
    double t_gl, t_write, t_fits, t_sed, t_end;
    char sptype[DDYN_STR];
    double teff, rstar, Bstar, Wz, mloss, vsw;     /* rstar [AU], Bstar ref units, Wz [1/yr], vsw [m/s] */
    double dist, fov, PA, inc, longitude;          /* dist [m], angles [rad] */
    int32_t xsize, ysize, nwavs, radQ;
    double wavs[DDYN_MAX_WAVS];                    /* [m] */
    char specfile[DDYN_STR];
    int32_t SWQ, FLQ, n_planets, n_disks;          /* n_planets counts the star */
    /* star (index 0) and planets after recalc_COM; m = G*M */
    double px[DDYN_MAX_PLANETS], py[DDYN_MAX_PLANETS], pz[DDYN_MAX_PLANETS];
    double pvx[DDYN_MAX_PLANETS], pvy[DDYN_MAX_PLANETS], pvz[DDYN_MAX_PLANETS], pm[DDYN_MAX_PLANETS];
    double rho, kqe;                               /* rho [kg/m^3], kqe per yr */
    char comp[DDYN_STR];
    int32_t n_blobs, n_grid, n_wav, n_theta;
    ddyn_disk disks[DDYN_MAX_DISKS];
    ddyn_blob blobs[DDYN_MAX_BLOBS];
    double amin, amax, mmin, mmax, f;              /* dust grid (optics; 0 if unknown) */
    int64_t seed;
    char name_stub[DDYN_STR];
} ddyn_config;

/* Bookkeeping of a run, as Grid's time, dt, tlastd/f/s and counts. */
typedef struct {
    double time, dt, tlastd, tlastf, tlasts;
    double Rmax;                      /* outer radius for the "## " marker */
    int64_t n_mass, n_dust, n_all;    /* n_mass = swarm bodies (excludes star + planets) */
} ddyn_run_info;

const char *ddyn_host_error_text(int code);

/* Parse a .param file (positional records 0-15). */
int ddyn_read_param(const char *path, ddyn_config *cfg);

/* Time step of Grid::det_step: min over massive bodies of 2 pi sqrt(r^3/(mu0+mui))/n_step
 * and the disk/blob terms 2 pi sqrt(R/mu0)/n_step, as the reference writes them. */
double ddyn_det_step(const ddyn_config *cfg, int64_t n_grav, const double *x, const double *y,
                     const double *z, const double *mu);

/* Write the data file.  Arrays are n_all long in the reference order; mu = G*M;
 * beta_bin (n_bins floats, may be NULL) gives the beta column by size bin;
 * beta_p (n_dust doubles, may be NULL) a per-grain beta column instead. */
int ddyn_write_datafile(const char *path, const char *version, const ddyn_config *cfg,
                        const ddyn_run_info *run, const double *x, const double *y, const double *z,
                        const double *vx, const double *vy, const double *vz, const double *N,
                        const double *mu, const int32_t *id, const float *beta_bin, int32_t n_bins,
                        const double *beta_p);

/* Read a data file.  Call once with x == NULL to fill cfg and run (counts);
 * then again with arrays of run->n_all entries (beta may be NULL).  The
 * configuration's n_planets is taken from the file.  Returns 55 on a version
 * mismatch and 56 on an nR mismatch, as the reference does. */
int ddyn_read_datafile(const char *path, const char *version, ddyn_config *cfg, ddyn_run_info *run,
                       double *x, double *y, double *z, double *vx, double *vy, double *vz,
                       double *N, double *mu, int32_t *id, double *beta);

/* ---------------------------------------------------------------------------
 * Set-up of a .param run (ddyn_setup.cpp): optical data, the dust size grid,
 * the grain coefficient tables and the particles of the disks and blobs.
 * ------------------------------------------------------------------------- */

/* The grain composition table and the stellar spectrum (replaces
 * add_optical_data and add_specfile, src/read_in.cu:377-506).  comp_path: the
 * .param record 12 file (first line skipped; then n_wav, n_theta, the number of
 * sizes and the size range in um as tokens 1-5 of the second line; per size a
 * "# <size um>" line, n_wav lines "<wavelength um> <Qabs> <Qsca>" and the phase
 * function, theta-major, ten values per line).  spec_path: lines "<wavelength m>
 * <surface flux>", scaled to the observer by (rstar / dist)^2 (cfg's rstar and
 * dist).  Returns 39 if a file cannot be opened, 42 if it is not well formed. */
typedef struct ddyn_optics ddyn_optics;
int  ddyn_optics_read(const char *comp_path, const char *spec_path, const ddyn_config *cfg, ddyn_optics **out);
void ddyn_optics_free(ddyn_optics *o);
int  ddyn_optics_info(const ddyn_optics *o, int32_t *n_wav, int32_t *n_theta, int32_t *n_sizes, double *dmin,
                      double *dmax);

/* The dust size grid of a .param run (replaces detminmax, confCdisks and
 * confCblobs, src/read_in.cu:236-342): each disk's and blob's grain sizes
 * clipped to the optical range (amin, mmin, dmax, dmmax), the global grid amin,
 * amax, mmin, mmax, n_grid (mass ratio 1.1 between bins) and f, the
 * normalisations C, n_wav and n_theta.  Returns 54 if n_grid is outside
 * [2, 2700] (the device tables hold 2,700 bins). */
int ddyn_set_dust_grid(ddyn_config *cfg, const ddyn_optics *o);

/* Outer radius [AU] of the system, twice the largest of the disks' Rout, the
 * blobs' a (1 + e) + sigma and the planets' distances (Grid::Rmax,
 * src/grid.cu:757-764): the data file marks grains beyond it. */
double ddyn_outer_radius(const ddyn_config *cfg);

/* The grain coefficient tables of the grid's n_grid size bins (replaces
 * interpolate_dust_to_grid, interpolate_stellar_to_grid and calculate_beta,
 * src/functions.cu:243-388) and the sublimation radius (calculate_dust_temp,
 * src/grid.cu:754-792): beta (radiation pressure over gravity), beta_sw
 * (stellar-wind drag), q (charge factor of the Lorentz force) as the float
 * tables ddyn_upload_bin_tables takes, size [m] (each may be NULL; n_max >=
 * n_grid entries), Rmin [AU] and the luminosity [Lsun] (may be NULL). */
int ddyn_grain_tables(const ddyn_config *cfg, const ddyn_optics *o, int32_t n_max, float *beta, float *betasw,
                      float *q, double *size, double *Rmin_au, double *lum);

/* The dust grid's optical tables for images and SEDs: the optical data
 * interpolated to the grid's sizes and the stellar flux to its wavelengths (as
 * ddyn_grain_tables does), and the grain temperature table of
 * calculate_dust_temp (src/grid.cu:754-834): per bin, the distance at which a
 * grain is in equilibrium at each temperature from 2 T_sub = 4000 K down to
 * 4 K in steps of 2 K, 0.5 sqrt(absorbed / emitted) with trapezoid sums over
 * the wavelengths; then the temperature on 100 radii log-spaced from the
 * sublimation radius to twice the outer radius, log-log between those points.
 * Where a radius lies inside the 4000 K distance or beyond the 4 K one (the
 * reference then reads outside its table) the nearest interval extrapolates.
 * Needs the dust grid (ddyn_set_dust_grid).  Returns 54 on a grid error. */
typedef struct ddyn_grid_optics ddyn_grid_optics;
int  ddyn_grid_optics_make(const ddyn_config *cfg, const ddyn_optics *o, ddyn_grid_optics **out);
void ddyn_grid_optics_free(ddyn_grid_optics *t);
/* Views of the tables for ddyn_upload_optics (valid while t lives). */
int  ddyn_grid_optics_tables(const ddyn_grid_optics *t, ddyn_optics_tables *v);

/* Wavelength indices of the .param file's image wavelengths (nwavs entries;
 * src/functions.cu:329-335): with i the first optical wavelength at or above
 * the requested one (at most the second last), the image is made at index
 * i - 1, the wavelength below it (0 where the reference would use -1).
 * Returns 43 for a wavelength outside the optical data. */
int ddyn_image_wavelengths(const ddyn_config *cfg, const ddyn_grid_optics *t, int32_t *wj);

/* Write an image as a FITS file <stub>_<wvl [um] %.2f>_<time %.3f>.fits
 * (replaces Grid::write_image, src/fits.cu:9-85): a primary array of
 * xsize x ysize FP64 pixels, F in ddyn_calc_flux's layout (pixel (x, y) at
 * F[ysize * x + y]), with the keys WAVELENG [um], FLUX (the sum of all pixels
 * [Jy]), PIXELSCA [arcsec/pixel], VERSION, a HISTORY line and DATE.  An
 * existing file is replaced.  name (may be NULL, name_len bytes) gets the
 * file name.  Returns 0, or 39 if the file cannot be written. */
int ddyn_write_image(const char *version, const ddyn_config *cfg, double time, double wvl, const double *F,
                     char *name, int32_t name_len);

/* Write the SED <stub>_SED_<time %.2f>.dat (replaces Grid::write_sed,
 * src/grid.cu:675-691): per wavelength "<w [um] %.3f> <stellar flux [Jy]
 * %.6e> <grain flux [Jy] %.6e>".  in is the stellar flux at the observer per
 * wavelength [W m^-3] (ddyn_optics_tables' in), Fexc ddyn_calc_sed's totals.
 * Returns 0, or 39 if the file cannot be written. */
int ddyn_write_sed(const char *version, const ddyn_config *cfg, double time, int32_t n_wav, const double *w,
                   const double *in, const double *Fexc, char *name, int32_t name_len);

/* Particle counts of the disks and blobs: gravitating bodies (sum of Ngrav) and
 * grain tracers (per disk or blob, bins x round(Nsmall / bins)); needs the dust
 * grid (ddyn_set_dust_grid) when grains are asked for. */
int ddyn_generate_counts(const ddyn_config *cfg, int64_t *n_mass, int64_t *n_dust);

/* The particles of a .param run (replaces Grid::setup's loops with getbod,
 * genorb_disk and genorb_blob, src/grid.cu:99-189, src/functions.cu:157-240),
 * n_all = n_planets + n_mass + n_dust entries in reference order and units:
 * star and planets; each disk's, then each blob's gravitating bodies down the
 * mass ladder of dN/dm = C m^-eta; then each disk's and blob's grains, Npg
 * tracers per mass bin with N the bin's number per tracer and id the bin.
 * beta (n_bins floats, from ddyn_grain_tables) reduces the grains' orbital
 * speeds as coe2rv does; o gives the optical size range (grains outside it get
 * id -1; may be NULL).  Orbits follow the reference's distributions (surface
 * density ~ r^gamma, |normal| eccentricities and inclinations, uniform angles;
 * Gaussian blobs with an optional collisional velocity kick) from the
 * reference's random stream (Knuth's subtractive generator in the parameters
 * of src/ran3.cu:12-48, seeded with cfg->seed, -2 as the reference), drawn in
 * the reference's order: the reference's particles, bit for bit, with the
 * same C library.
 * Returns 54 when n_all does not match the counts, or when a law's rejection
 * draw (e < 1 for a disk, a > 0 for a blob) fails 2^20 times in a row. */
int ddyn_generate(const ddyn_config *cfg, const ddyn_optics *o, const float *beta, int32_t n_bins, int64_t n_all,
                  double *x, double *y, double *z, double *vx, double *vy, double *vz, double *N, double *mu,
                  int32_t *id);

/* Size-bin ids of a restart from the masses (Grid::read_in, src/grid.cu:285-287):
 * round(log(m / mmin) / f), -1 beyond the grid, for massless bodies and for the
 * star and planets. */
int ddyn_restart_ids(const ddyn_config *cfg, int64_t n_all, const double *mu, int32_t *id);

#ifdef __cplusplus
}
#endif
#endif
