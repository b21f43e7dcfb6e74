/*
 * ddyn_oracle.c -- FP64 CPU restatement of DiskDyn's dynamical hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle for the HIP library in
 * diskdyn_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and never as the thing measured or shipped.
 *
 * What it restates (reference paths relative to the DiskDyn tree):
 *   - massive pair law           src/gpu_functions.cu:101-217 (RK4calc_massive2, same law in 4/6/8)
 *   - massive RK4 stage algebra  src/gpu_functions.cu:90-99, 219-230, 350-362, 482-493, 613-625
 *   - fused grain RK4            src/gpu_functions.cu:627-943 (RK4calc_dust)
 *   - step ordering              src/grid.cu:385-389 and :693-752
 *   - orbital elements           src/functions.cu:6-46 (coe2rv)
 *   - barycentre correction      src/read_in.cu:344-375 (recalc_COM)
 *   - constants                  src/ddyn.cuh:16-38
 *
 * Arithmetic conventions.  The reference is compiled by nvcc, which contracts
 * a*b+c into one fused multiply-add.  This restatement writes every such
 * contraction as an explicit fma() and is compiled with -ffp-contract=off,
 * so its rounding is fully specified.  Sums over sources run in ascending
 * source index (the reference sums massive pairs through CAS atomics in an
 * unspecified order; ascending order is one admissible order).
 *
 * Deliberate differences from the reference (each is a reference quirk):
 *   - optional Plummer softening eps on swarm-sourced pairs (eps = 0 gives the
 *     reference law exactly);
 *   - stage-4 sublimation flags the grain itself (the reference writes N[i],
 *     another particle's slot, src/gpu_functions.cu:878);
 *   - the charge q is 0 for grains with id < 0 (the reference reads q[id]
 *     out of bounds, src/gpu_functions.cu:665);
 *   - a lone star (n_grav == 1) is still advanced with its own velocity (the
 *     reference skips massive RK4 then, src/grid.cu:698).
 *
 * Pinning: tests/test_oracle.py checks every function here against closed
 * forms (two-body acceleration, Kepler orbits, the RK4 stability polynomial,
 * Wyatt-Whipple PR-drag rates), invariants (Newton's third law, momentum,
 * energy, angular momentum, Jacobi constant), brute force in long double and
 * central finite differences (adjoint).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "ddyn_oracle.h"

/* ------------------------------------------------------------------------- */
/* Pair laws                                                                  */
/* ------------------------------------------------------------------------- */

/* Softening applied to the pair (target i, source j): swarm bodies are the
 * massive bodies with index >= n_planets; grains count as swarm targets. */
static inline double pair_eps2(const orc_params *p, int64_t i_is_swarm, int64_t j)
{
    return (i_is_swarm && j >= p->n_planets) ? p->eps * p->eps : 0.0;
}

/* Massive acceleration on body i from all massive bodies, ascending j.
 * Reference per pair: f = RR*sqrt(RR); kv = dX/f (0 if f == 0); K += kv*m
 * (src/gpu_functions.cu:157-172, 188-193). */
static void massive_accel_one(const orc_params *p, int64_t n, int64_t i,
                              const double *x, const double *y, const double *z,
                              const double *mu, double *ax, double *ay, double *az)
{
    const double xi = x[i], yi = y[i], zi = z[i];
    const int64_t i_swarm = (i >= p->n_planets);
    double KX = 0.0, KY = 0.0, KZ = 0.0;
    for (int64_t j = 0; j < n; j++) {
        if (j == i) continue;
        double dX = x[j] - xi, dY = y[j] - yi, dZ = z[j] - zi;
        double RR = fma(dZ, dZ, fma(dY, dY, dX * dX)) + pair_eps2(p, i_swarm, j);
        double f = RR * sqrt(RR);
        double kx = 0.0, ky = 0.0, kz = 0.0;
        if (f != 0.0) { kx = dX / f; ky = dY / f; kz = dZ / f; }
        KX = fma(kx, mu[j], KX);
        KY = fma(ky, mu[j], KY);
        KZ = fma(kz, mu[j], KZ);
    }
    *ax = KX; *ay = KY; *az = KZ;
}

void orc_massive_accel(const orc_params *p, int64_t n_grav,
                       const double *x, const double *y, const double *z, const double *mu,
                       double *ax, double *ay, double *az)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_grav; i++)
        massive_accel_one(p, n_grav, i, x, y, z, mu, &ax[i], &ay[i], &az[i]);
}

/* Accelerations of the targets idx[0..nt) only, each summed over all n_grav
 * sources (sampled parity checks and the bounded CPU baseline). */
void orc_massive_accel_targets(const orc_params *p, int64_t n_grav,
                               const double *x, const double *y, const double *z, const double *mu,
                               int64_t nt, const int64_t *idx, double *ax, double *ay, double *az)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nt; k++)
        massive_accel_one(p, n_grav, idx[k], x, y, z, mu, &ax[k], &ay[k], &az[k]);
}

/* Grain acceleration at one RK stage (src/gpu_functions.cu:679-735 for stage
 * 1; stages 2-4 repeat it with the stage position and velocity).
 * (sx,sy,sz) are the massive bodies' positions at this stage, body 0 = star.
 * Returns 1 if the grain is removed (sublimation or exact hit), else 0. */
static int grain_accel(const orc_params *p, int64_t n_grav,
                       const double *sx, const double *sy, const double *sz, const double *mu,
                       double x, double y, double z, double vx, double vy, double vz,
                       double bet, double betsw, double qe, double m,
                       double *kx, double *ky, double *kz)
{
    double dx = x - sx[0], dy = y - sy[0], dz = z - sz[0];
    double RR = fma(dz, dz, fma(dy, dy, dx * dx));
    double RR3 = RR * sqrt(RR);
    if (RR <= p->Rmin * p->Rmin) return 1;                        /* :686-689 */
    double m0 = mu[0];
    double Fg = (1.0 - bet - betsw) * m0 / RR3;                    /* :691 */
    double Fb1;
    if (p->SWQ == 1) Fb1 = m0 / RR * (bet / p->light + betsw / p->vsw);   /* :694-696 */
    else             Fb1 = bet * m0 / p->light / RR;                       /* :697-699 */
    double Fb2 = fma(dz, vz, fma(dy, vy, dx * vx)) / RR;          /* :700 */
    double Fdx = Fb1 * fma(Fb2, dx, vx);                           /* :702-704 */
    double Fdy = Fb1 * fma(Fb2, dy, vy);
    double Fdz = Fb1 * fma(Fb2, dz, vz);
    double Flx = 0.0, Fly = 0.0;
    if (p->FLQ == 1) {                                             /* :706-714 */
        double Bz = p->Bstar * p->Rstar * p->Rstar * p->Rstar / RR3;
        Flx =  qe * fma(-p->Wz, dx, vy) * Bz / m;
        Fly = -qe * fma(p->Wz, dy, vx) * Bz / m;
    }
    double ax = fma(-Fg, dx, -Fdx) + Flx;                          /* :716-718 */
    double ay = fma(-Fg, dy, -Fdy) + Fly;
    double az = fma(-Fg, dz, -Fdz);
    const double e2 = p->eps * p->eps;
    for (int64_t j = 1; j < n_grav; j++) {                         /* :722-735 */
        double ex = x - sx[j], ey = y - sy[j], ez = z - sz[j];
        double RRu = fma(ez, ez, fma(ey, ey, ex * ex));
        if (RRu == 0.0) return 1;                                  /* :727-730 */
        double R2 = (j >= p->n_planets) ? RRu + e2 : RRu;
        double f = mu[j] / R2 / sqrt(R2);
        ax = fma(-f, ex, ax);
        ay = fma(-f, ey, ay);
        az = fma(-f, ez, az);
    }
    *kx = ax; *ky = ay; *kz = az;
    return 0;
}

static inline void grain_coeffs(const orc_state *s, int64_t t,
                                double *bet, double *betsw, double *qe)
{
    int32_t id = s->id ? s->id[t] : -1;
    int64_t g = t - s->n_grav;
    if (s->beta_p) {                       /* per-grain FP64 coefficients */
        *bet = s->beta_p[g];
        *betsw = s->betasw_p ? s->betasw_p[g] : 0.0;
    } else if (id >= 0 && id < s->n_bins) {  /* per-size-bin float tables (ddyn.cu:30-32) */
        *bet = (double)s->beta_tab[id];
        *betsw = (double)s->betasw_tab[id];
    } else {
        *bet = 0.0; *betsw = 0.0;
    }
    *qe = (s->q_tab && id >= 0 && id < s->n_bins) ? (double)s->q_tab[id] : 0.0;
}

int orc_dust_accel(const orc_params *p, const orc_state *s, double *ax, double *ay, double *az,
                   int32_t *removed)
{
    const int64_t ng = s->n_grav;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < s->n_dust; g++) {
        int64_t t = ng + g;
        double bet, betsw, qe;
        grain_coeffs(s, t, &bet, &betsw, &qe);
        double kx = 0, ky = 0, kz = 0;
        int r = grain_accel(p, ng, s->x, s->y, s->z, s->mu, s->x[t], s->y[t], s->z[t],
                            s->vx[t], s->vy[t], s->vz[t], bet, betsw, qe, s->mu[t], &kx, &ky, &kz);
        ax[g] = kx; ay[g] = ky; az[g] = kz;
        removed[g] = r;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* One RK4 step                                                               */
/* ------------------------------------------------------------------------- */

/* Workspace for the massive stages: stage positions k1..k4 (x,y,z), stage
 * velocities k2r..k4r and stage accelerations k1v..k4v, as in coord
 * (src/ddyn.cuh:44-58). */
typedef struct {
    double *kx[4], *ky[4], *kz[4];
    double *rx[4], *ry[4], *rz[4];     /* rx[0] unused: stage-1 velocity is v */
    double *vx[4], *vy[4], *vz[4];     /* stage accelerations */
} massive_ws;

static int ws_alloc(massive_ws *w, int64_t n)
{
    double **all[] = {w->kx, w->ky, w->kz, w->rx, w->ry, w->rz, w->vx, w->vy, w->vz};
    for (int a = 0; a < 9; a++)
        for (int s = 0; s < 4; s++) {
            all[a][s] = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
            if (!all[a][s]) return -1;
        }
    return 0;
}

static void ws_free(massive_ws *w)
{
    double **all[] = {w->kx, w->ky, w->kz, w->rx, w->ry, w->rz, w->vx, w->vy, w->vz};
    for (int a = 0; a < 9; a++)
        for (int s = 0; s < 4; s++) free(all[a][s]);
}

/* Grid::massive_RK4 (src/grid.cu:693-727): four massive stages. */
static void massive_stages(const orc_params *p, const orc_state *s, massive_ws *w, double dt)
{
    const int64_t n = s->n_grav;
    const double h[3] = {0.5 * dt, 0.5 * dt, dt};
    for (int64_t i = 0; i < n; i++) {              /* massive1: k1x = x (:90-99) */
        w->kx[0][i] = s->x[i]; w->ky[0][i] = s->y[i]; w->kz[0][i] = s->z[i];
    }
    for (int st = 0; st < 4; st++) {
        orc_massive_accel(p, n, w->kx[st], w->ky[st], w->kz[st], s->mu,
                          w->vx[st], w->vy[st], w->vz[st]);
        if (st == 3) break;
        /* massive3/5/7 (:219-230, 350-362, 482-493):
         *   k{s+1}x = x + k{s}r*h',  k{s+1}r = v + k{s}v*h'  with k1r = v. */
        const double *vrx = st ? w->rx[st] : s->vx, *vry = st ? w->ry[st] : s->vy,
                     *vrz = st ? w->rz[st] : s->vz;
        for (int64_t i = 0; i < n; i++) {
            w->kx[st + 1][i] = fma(vrx[i], h[st], s->x[i]);
            w->ky[st + 1][i] = fma(vry[i], h[st], s->y[i]);
            w->kz[st + 1][i] = fma(vrz[i], h[st], s->z[i]);
            w->rx[st + 1][i] = fma(w->vx[st][i], h[st], s->vx[i]);
            w->ry[st + 1][i] = fma(w->vy[st][i], h[st], s->vy[i]);
            w->rz[st + 1][i] = fma(w->vz[st][i], h[st], s->vz[i]);
        }
    }
}

/* RK4calc_massive_final (src/gpu_functions.cu:613-625), dt6 = dt/6 (grid.cu:750):
 *   x += dt6*(vx + 2*k2r + 2*k3r + k4r), evaluated left to right. */
static void massive_final(const orc_state *s, massive_ws *w, double dt)
{
    const double dt6 = dt / 6.0;
    for (int64_t i = 0; i < s->n_grav; i++) {
        double Sx = fma(2.0, w->rx[1][i], s->vx[i]);
        double Sy = fma(2.0, w->ry[1][i], s->vy[i]);
        double Sz = fma(2.0, w->rz[1][i], s->vz[i]);
        Sx = fma(2.0, w->rx[2][i], Sx) + w->rx[3][i];
        Sy = fma(2.0, w->ry[2][i], Sy) + w->ry[3][i];
        Sz = fma(2.0, w->rz[2][i], Sz) + w->rz[3][i];
        double Ax = fma(2.0, w->vx[1][i], w->vx[0][i]);
        double Ay = fma(2.0, w->vy[1][i], w->vy[0][i]);
        double Az = fma(2.0, w->vz[1][i], w->vz[0][i]);
        Ax = fma(2.0, w->vx[2][i], Ax) + w->vx[3][i];
        Ay = fma(2.0, w->vy[2][i], Ay) + w->vy[3][i];
        Az = fma(2.0, w->vz[2][i], Az) + w->vz[3][i];
        s->x[i] = fma(dt6, Sx, s->x[i]);
        s->y[i] = fma(dt6, Sy, s->y[i]);
        s->z[i] = fma(dt6, Sz, s->z[i]);
        s->vx[i] = fma(dt6, Ax, s->vx[i]);
        s->vy[i] = fma(dt6, Ay, s->vy[i]);
        s->vz[i] = fma(dt6, Az, s->vz[i]);
    }
}

/* RK4calc_dust (src/gpu_functions.cu:627-943) for one grain; massive stage
 * positions come from the workspace. */
static void grain_step(const orc_params *p, const orc_state *s, const massive_ws *w,
                       int64_t t, double dt)
{
    if (s->N[t] == 0.0) return;                                    /* :643 */
    double bet, betsw, qe;
    grain_coeffs(s, t, &bet, &betsw, &qe);
    const double dt2 = dt / 2.0, dt6 = dt / 6.0;
    const double x = s->x[t], y = s->y[t], z = s->z[t];
    const double vx = s->vx[t], vy = s->vy[t], vz = s->vz[t];
    const double m = s->mu[t];
    const int64_t ng = s->n_grav;
    double k1x, k1y, k1z, k2x, k2y, k2z, k3x, k3y, k3z, k4x, k4y, k4z;
    /* K1 at (x, v) */
    if (grain_accel(p, ng, w->kx[0], w->ky[0], w->kz[0], s->mu, x, y, z, vx, vy, vz,
                    bet, betsw, qe, m, &k1x, &k1y, &k1z)) { s->N[t] = 0.0; return; }
    /* K2 (:737-743) */
    double r2x = fma(k1x, dt2, vx), r2y = fma(k1y, dt2, vy), r2z = fma(k1z, dt2, vz);
    if (grain_accel(p, ng, w->kx[1], w->ky[1], w->kz[1], s->mu,
                    fma(vx, dt2, x), fma(vy, dt2, y), fma(vz, dt2, z), r2x, r2y, r2z,
                    bet, betsw, qe, m, &k2x, &k2y, &k2z)) { s->N[t] = 0.0; return; }
    /* K3 (:800-806) */
    double r3x = fma(k2x, dt2, vx), r3y = fma(k2y, dt2, vy), r3z = fma(k2z, dt2, vz);
    if (grain_accel(p, ng, w->kx[2], w->ky[2], w->kz[2], s->mu,
                    fma(r2x, dt2, x), fma(r2y, dt2, y), fma(r2z, dt2, z), r3x, r3y, r3z,
                    bet, betsw, qe, m, &k3x, &k3y, &k3z)) { s->N[t] = 0.0; return; }
    /* K4 (:863-869) */
    double r4x = fma(k3x, dt, vx), r4y = fma(k3y, dt, vy), r4z = fma(k3z, dt, vz);
    if (grain_accel(p, ng, w->kx[3], w->ky[3], w->kz[3], s->mu,
                    fma(r3x, dt, x), fma(r3y, dt, y), fma(r3z, dt, z), r4x, r4y, r4z,
                    bet, betsw, qe, m, &k4x, &k4y, &k4z)) { s->N[t] = 0.0; return; }
    /* Evolve (:927-933): x += dt6*(vx + 2*(k2r + k3r) + k4r) */
    s->x[t] = fma(dt6, fma(2.0, r2x + r3x, vx) + r4x, x);
    s->y[t] = fma(dt6, fma(2.0, r2y + r3y, vy) + r4y, y);
    s->z[t] = fma(dt6, fma(2.0, r2z + r3z, vz) + r4z, z);
    s->vx[t] = fma(dt6, fma(2.0, k2x + k3x, k1x) + k4x, vx);
    s->vy[t] = fma(dt6, fma(2.0, k2y + k3y, k1y) + k4y, vy);
    s->vz[t] = fma(dt6, fma(2.0, k2z + k3z, k1z) + k4z, vz);
}

int orc_step(const orc_params *p, orc_state *s, double dt, int64_t nsteps)
{
    massive_ws w;
    memset(&w, 0, sizeof w);
    if (ws_alloc(&w, s->n_grav)) { ws_free(&w); return -1; }
    for (int64_t k = 0; k < nsteps; k++) {
        massive_stages(p, s, &w, dt);                  /* grid.cu:387 */
#pragma omp parallel for schedule(static)
        for (int64_t g = 0; g < s->n_dust; g++)        /* grid.cu:388 */
            grain_step(p, s, &w, s->n_grav + g, dt);
        massive_final(s, &w, dt);                      /* grid.cu:389 */
    }
    ws_free(&w);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Discrete adjoint of the RK4 step (no reference counterpart)               */
/* ------------------------------------------------------------------------- */
/* The step map is y -> y' with stages R1 = r, V1 = v; R2 = r + h/2 V1,
 * V2 = v + h/2 A1; R3 = r + h/2 V2, V3 = v + h/2 A2; R4 = r + h V3,
 * V4 = v + h A3; r' = r + h/6 (V1 + 2V2 + 2V3 + V4), v' = v + h/6 (A1 + ...),
 * A_s = f(R_s, V_s).  Its reverse is written out stage by stage; the force
 * Jacobians are, for K(d) = d / s^3 with s^2 = |d|^2 + eps^2,
 *   dK/dd = J(d) = (I s^2 - 3 d d^T) / s^5   (symmetric, J(-d) = J(d)).
 * Massive bodies: a_i = sum_j mu_j K(R_j - R_i), so
 *   Rbar_k += sum_{j != k} J(R_j - R_k) (mu_k Abar_j - mu_j Abar_k),
 *   mubar_k += sum_{i != k} Abar_i . K(R_k - R_i).
 * Grains: the star term (gravity, radiation pressure, PR + wind drag) is
 * differentiated in closed form; the Lorentz force is not supported here. */

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Rbar[k] += ..., mubar[k] += ... for all massive k at stage positions R. */
static void massive_vjp(const orc_params *p, int64_t n, const double *Rx, const double *Ry,
                        const double *Rz, const double *mu, const double *Ab /* 3n */,
                        double *Rb /* 3n */, double *mub)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; k++) {
        const int64_t k_swarm = (k >= p->n_planets);
        double acc[3] = {0, 0, 0}, am = 0.0;
        const double *Ak = &Ab[3 * k];
        for (int64_t j = 0; j < n; j++) {
            if (j == k) continue;
            double d[3] = {Rx[j] - Rx[k], Ry[j] - Ry[k], Rz[j] - Rz[k]};
            double s2 = dot3(d, d) + pair_eps2(p, k_swarm, j);
            if (s2 == 0.0) continue;
            double inv = 1.0 / sqrt(s2), inv3 = inv * inv * inv, inv5 = inv3 * inv * inv;
            const double *Aj = &Ab[3 * j];
            double w[3] = {mu[k] * Aj[0] - mu[j] * Ak[0], mu[k] * Aj[1] - mu[j] * Ak[1],
                           mu[k] * Aj[2] - mu[j] * Ak[2]};
            double dw = dot3(d, w);
            for (int c = 0; c < 3; c++) acc[c] += w[c] * inv3 - 3.0 * d[c] * dw * inv5;
            am -= dot3(Aj, d) * inv3;
        }
        for (int c = 0; c < 3; c++) Rb[3 * k + c] += acc[c];
        mub[k] += am;
    }
}

/* Grain state at the four stages of one step (forward, as grain_step). */
typedef struct { double R[4][3], V[4][3]; int gone; } grain_stages;

static void grain_forward_stages(const orc_params *p, const orc_state *s, const massive_ws *w,
                                 int64_t t, double dt, double bet, double betsw, grain_stages *G)
{
    const double dt2 = dt / 2.0;
    const double x0[3] = {s->x[t], s->y[t], s->z[t]}, v0[3] = {s->vx[t], s->vy[t], s->vz[t]};
    const double hh[3] = {dt2, dt2, dt};
    double A[3];
    G->gone = 0;
    for (int c = 0; c < 3; c++) { G->R[0][c] = x0[c]; G->V[0][c] = v0[c]; }
    for (int st = 0; st < 4; st++) {
        if (grain_accel(p, s->n_grav, w->kx[st], w->ky[st], w->kz[st], s->mu, G->R[st][0], G->R[st][1],
                        G->R[st][2], G->V[st][0], G->V[st][1], G->V[st][2], bet, betsw, 0.0,
                        s->mu[t], &A[0], &A[1], &A[2])) { G->gone = 1; return; }
        if (st == 3) break;
        for (int c = 0; c < 3; c++) {
            G->R[st + 1][c] = fma(G->V[st][c], hh[st], x0[c]);
            G->V[st + 1][c] = fma(A[c], hh[st], v0[c]);
        }
    }
}

/* Reverse of one grain's step.  rb, vb: in = dL/d(state after), out = dL/d(state
 * before).  Accumulates the grain's pull on the massive stage positions into
 * Rbm[st][3*j] and on the masses into mub[j]; betab gets dL/dbeta. */
static void grain_reverse(const orc_params *p, const orc_state *s, const massive_ws *w, int64_t t,
                          double dt, double bet, double betsw, const grain_stages *G,
                          double *rb, double *vb, double *betab, double *Rbm[4], double *mub)
{
    const double h = dt;
    const double wts[4] = {h / 6.0, h / 3.0, h / 3.0, h / 6.0};
    double Vb[4][3], Ab[4][3];
    for (int st = 0; st < 4; st++)
        for (int c = 0; c < 3; c++) { Vb[st][c] = wts[st] * rb[c]; Ab[st][c] = wts[st] * vb[c]; }
    double r_bar[3] = {rb[0], rb[1], rb[2]}, v_bar[3] = {vb[0], vb[1], vb[2]};
    const double m0 = s->mu[0];
    const double kap = (p->SWQ == 1) ? bet / p->light + betsw / p->vsw : bet / p->light;
    const double dkap = 1.0 / p->light;            /* d kappa / d beta */
    const double e2 = p->eps * p->eps;
    const int64_t ng = s->n_grav;
    const double hh[3] = {h / 2.0, h / 2.0, h};
    for (int st = 3; st >= 0; st--) {
        const double *Rs = G->R[st], *Vs = G->V[st], *A = Ab[st];
        double Rb[3] = {0, 0, 0};
        /* star term */
        double d[3] = {Rs[0] - w->kx[st][0], Rs[1] - w->ky[st][0], Rs[2] - w->kz[st][0]};
        double R2 = dot3(d, d), R = sqrt(R2), R3 = R2 * R, R4 = R2 * R2, R5 = R3 * R2, R6 = R4 * R2;
        double Gc = (1.0 - bet - betsw) * m0;
        double dv = dot3(d, Vs), dA = dot3(d, A), vA = dot3(Vs, A);
        double db[3];
        for (int c = 0; c < 3; c++) {
            double J0A = A[c] / R3 - 3.0 * d[c] * dA / R5;
            db[c] = -Gc * J0A - m0 * kap * ((Vs[c] * dA + dv * A[c]) / R4 - 4.0 * dv * dA * d[c] / R6
                                              - 2.0 * d[c] * vA / R4);
            Vb[st][c] += -m0 * kap * (d[c] * dA / R4 + A[c] / R2);
        }
        double term2[3];
        for (int c = 0; c < 3; c++) term2[c] = dv * d[c] / R4 + Vs[c] / R2;
        *betab += m0 * dA / R3 - m0 * dkap * dot3(term2, A);
        mub[0] += -(1.0 - bet - betsw) * dA / R3 - kap * dot3(term2, A);
        for (int c = 0; c < 3; c++) { Rb[c] += db[c]; Rbm[st][c] -= db[c]; }
        /* massive pulls, j >= 1: a += mu_j K(R_j - R_t) */
        for (int64_t j = 1; j < ng; j++) {
            double q[3] = {w->kx[st][j] - Rs[0], w->ky[st][j] - Rs[1], w->kz[st][j] - Rs[2]};
            double s2 = dot3(q, q) + ((j >= p->n_planets) ? e2 : 0.0);
            double inv = 1.0 / sqrt(s2), inv3 = inv * inv * inv, inv5 = inv3 * inv * inv;
            double qA = dot3(q, A);
            for (int c = 0; c < 3; c++) {
                double JA = A[c] * inv3 - 3.0 * q[c] * qA * inv5;
                Rbm[st][3 * j + c] += s->mu[j] * JA;
                Rb[c] -= s->mu[j] * JA;
            }
            mub[j] += qA * inv3;
        }
        /* chain through the stage definitions */
        for (int c = 0; c < 3; c++) {
            r_bar[c] += Rb[c];
            if (st == 0) { v_bar[c] += Vb[0][c]; continue; }
            v_bar[c] += Vb[st][c];
            Vb[st - 1][c] += hh[st - 1] * Rb[c];
            Ab[st - 1][c] += hh[st - 1] * Vb[st][c];
        }
    }
    for (int c = 0; c < 3; c++) { rb[c] = r_bar[c]; vb[c] = v_bar[c]; }
}

int orc_adjoint(const orc_params *p, const orc_state *s0, double dt, int64_t nsteps,
                const double *lam_r, const double *lam_v, double *g_r, double *g_v,
                double *g_mu, double *g_beta)
{
    if (p->FLQ == 1) return -2;
    const int64_t ng = s0->n_grav, nd = s0->n_dust, na = ng + nd;
    const size_t per = (size_t)na * 7;
    double *hist = (double *)malloc(per * (size_t)(nsteps + 1) * sizeof(double));
    double *Rbm[4];
    double *mub = (double *)calloc((size_t)ng, sizeof(double));
    double *Abm = (double *)calloc(3 * 4 * (size_t)ng, sizeof(double));
    double *Vbm = (double *)calloc(3 * 4 * (size_t)ng, sizeof(double));
    for (int st = 0; st < 4; st++) Rbm[st] = (double *)calloc(3 * (size_t)ng, sizeof(double));
    massive_ws w;
    memset(&w, 0, sizeof w);
    if (!hist || !mub || !Abm || !Vbm || !Rbm[0] || !Rbm[1] || !Rbm[2] || !Rbm[3] || ws_alloc(&w, ng)) {
        free(hist); free(mub); free(Abm); free(Vbm);
        for (int st = 0; st < 4; st++) free(Rbm[st]);
        ws_free(&w);
        return -1;
    }
    /* forward, keeping every step's state (x,y,z,vx,vy,vz,N) */
    orc_state s = *s0;
    double *cur = (double *)malloc(per * sizeof(double));
    double *arr[7];
    for (int a = 0; a < 7; a++) arr[a] = cur + (size_t)a * na;
    const double *src0[7] = {s0->x, s0->y, s0->z, s0->vx, s0->vy, s0->vz, s0->N};
    for (int a = 0; a < 7; a++) memcpy(arr[a], src0[a], (size_t)na * sizeof(double));
    s.x = arr[0]; s.y = arr[1]; s.z = arr[2]; s.vx = arr[3]; s.vy = arr[4]; s.vz = arr[5]; s.N = arr[6];
    memcpy(hist, cur, per * sizeof(double));
    for (int64_t k = 0; k < nsteps; k++) {
        orc_step(p, &s, dt, 1);
        memcpy(hist + per * (size_t)(k + 1), cur, per * sizeof(double));
    }
    /* adjoint state */
    double *rb = g_r, *vb = g_v;
    memcpy(rb, lam_r, 3 * (size_t)na * sizeof(double));
    memcpy(vb, lam_v, 3 * (size_t)na * sizeof(double));
    memset(g_mu, 0, (size_t)ng * sizeof(double));
    if (g_beta) memset(g_beta, 0, (size_t)(nd > 0 ? nd : 0) * sizeof(double));
    const double h = dt, wts[4] = {h / 6.0, h / 3.0, h / 3.0, h / 6.0}, hh[3] = {h / 2.0, h / 2.0, h};
    for (int64_t k = nsteps - 1; k >= 0; k--) {
        memcpy(cur, hist + per * (size_t)k, per * sizeof(double));
        massive_stages(p, &s, &w, dt);     /* stage positions of step k */
        for (int st = 0; st < 4; st++) memset(Rbm[st], 0, 3 * (size_t)ng * sizeof(double));
        memset(mub, 0, (size_t)ng * sizeof(double));
        /* grains (serial: they scatter into the massive adjoints) */
        for (int64_t g = 0; g < nd; g++) {
            int64_t t = ng + g;
            if (s.N[t] == 0.0) continue;
            double bet, betsw, qe;
            grain_coeffs(&s, t, &bet, &betsw, &qe);
            grain_stages G;
            grain_forward_stages(p, &s, &w, t, dt, bet, betsw, &G);
            if (G.gone) continue;                        /* removed this step: identity map */
            double betab = 0.0;
            grain_reverse(p, &s, &w, t, dt, bet, betsw, &G, &rb[3 * t], &vb[3 * t], &betab, Rbm, mub);
            if (g_beta) g_beta[g] += betab;
        }
        /* massive bodies: Vbar_s = w_s rbar', Abar_s = w_s vbar' (before any update) */
        for (int st = 0; st < 4; st++)
            for (int64_t i = 0; i < 3 * ng; i++) {
                Vbm[st * 3 * ng + i] = wts[st] * rb[i];
                Abm[st * 3 * ng + i] = wts[st] * vb[i];
            }
        for (int st = 3; st >= 0; st--) {
            massive_vjp(p, ng, w.kx[st], w.ky[st], w.kz[st], s.mu, &Abm[st * 3 * ng], Rbm[st], mub);
            for (int64_t i = 0; i < 3 * ng; i++) {
                const double Rb = Rbm[st][i], Vb = Vbm[st * 3 * ng + i];
                rb[i] += Rb;                       /* R_s = r + ... */
                vb[i] += Vb;                       /* V_s = v + ... (V_1 = v) */
                if (st > 0) {
                    Vbm[(st - 1) * 3 * ng + i] += hh[st - 1] * Rb;   /* R_s = r + h' V_{s-1} */
                    Abm[(st - 1) * 3 * ng + i] += hh[st - 1] * Vb;   /* V_s = v + h' A_{s-1} */
                }
            }
        }
        for (int64_t j = 0; j < ng; j++) g_mu[j] += mub[j];
    }
    free(cur); free(hist); free(mub); free(Abm); free(Vbm);
    for (int st = 0; st < 4; st++) free(Rbm[st]);
    ws_free(&w);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Invariants of the massive set (units: mu = G*M, so these are G*E, G*L, G*P) */
/* ------------------------------------------------------------------------- */

void orc_energy(const orc_params *p, int64_t n, const double *x, const double *y, const double *z,
                const double *vx, const double *vy, const double *vz, const double *mu,
                double *out7)
{
    double K = 0.0, U = 0.0;
    double L[3] = {0, 0, 0}, P[3] = {0, 0, 0};
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : K, U)
    for (int64_t i = 0; i < n; i++) {
        K += 0.5 * mu[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
        const int64_t i_swarm = (i >= p->n_planets);
        double u = 0.0;
        for (int64_t j = i + 1; j < n; j++) {
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double R2 = dx * dx + dy * dy + dz * dz + pair_eps2(p, i_swarm, j);
            if (R2 > 0.0) u -= mu[j] / sqrt(R2);
        }
        U += mu[i] * u;
    }
    for (int64_t i = 0; i < n; i++) {
        L[0] += mu[i] * (y[i] * vz[i] - z[i] * vy[i]);
        L[1] += mu[i] * (z[i] * vx[i] - x[i] * vz[i]);
        L[2] += mu[i] * (x[i] * vy[i] - y[i] * vx[i]);
        P[0] += mu[i] * vx[i]; P[1] += mu[i] * vy[i]; P[2] += mu[i] * vz[i];
    }
    out7[0] = K + U;
    out7[1] = L[0]; out7[2] = L[1]; out7[3] = L[2];
    out7[4] = P[0]; out7[5] = P[1]; out7[6] = P[2];
}

/* ------------------------------------------------------------------------- */
/* Initial conditions                                                         */
/* ------------------------------------------------------------------------- */

/* coe2rv (src/functions.cu:6-46): orbital elements -> Cartesian, mu = (M+m)(1-beta).
 * Argument order as in the reference: (M, m, a, e, i, w, o, nu, ..., beta). */
void orc_coe2rv(double M, double m, double a, double e, double inc, double w, double o,
                double nu, double beta, double *out6)
{
    double p = a * (1.0 - e * e);
    double mu = (M + m) * (1.0 - beta);
    double xe = p * cos(nu) / (1.0 + e * cos(nu));
    double ye = p * sin(nu) / (1.0 + e * cos(nu));
    double vxe = -sqrt(mu / p) * sin(nu);
    double vye = sqrt(mu / p) * (e + cos(nu));
    double co = cos(o), so = sin(o), cw = cos(w), sw = sin(w), ci = cos(inc), si = sin(inc);
    double r11 = co * cw - so * sw * ci, r12 = -co * sw - so * cw * ci;
    double r21 = so * cw + co * sw * ci, r22 = -so * sw + co * cw * ci;
    double r31 = sw * si, r32 = cw * si;
    out6[0] = xe * r11 + ye * r12;
    out6[1] = xe * r21 + ye * r22;
    out6[2] = xe * r31 + ye * r32;
    out6[3] = vxe * r11 + vye * r12;
    out6[4] = vxe * r21 + vye * r22;
    out6[5] = vxe * r31 + vye * r32;
}

/* recalc_COM (src/read_in.cu:344-375): shift star+planets to the barycentre,
 * then give the star the velocity that zeroes the total momentum.  The
 * planets' velocities are left unchanged, as in the reference. */
void orc_recalc_com(int64_t n_planets, double *x, double *y, double *z,
                    double *vx, double *vy, double *vz, const double *mu)
{
    if (n_planets == 1) return;
    double Xc = 0, Yc = 0, Zc = 0, M = 0;
    for (int64_t j = 0; j < n_planets; j++) {
        Xc += mu[j] * x[j]; Yc += mu[j] * y[j]; Zc += mu[j] * z[j]; M += mu[j];
    }
    Xc /= M; Yc /= M; Zc /= M;
    for (int64_t j = 0; j < n_planets; j++) { x[j] -= Xc; y[j] -= Yc; z[j] -= Zc; }
    double Vx = 0, Vy = 0, Vz = 0;
    for (int64_t j = 1; j < n_planets; j++) {
        Vx -= mu[j] * vx[j]; Vy -= mu[j] * vy[j]; Vz -= mu[j] * vz[j];
    }
    vx[0] = Vx / mu[0]; vy[0] = Vy / mu[0]; vz[0] = Vz / mu[0];
}

/* ------------------------------------------------------------------------
 * Images and SEDs: src/gpu_functions.cu:983-1156, one grain at a time in
 * grain order, with the reference's expressions and its operation order.
 * ---------------------------------------------------------------------- */
#define ORC_AU 1.49598e11           /* src/ddyn.cuh:18 */
#define ORC_CSPEED 299792458.0      /* :23 */
#define ORC_KBOLTZ 1.3806503E-23    /* :24 */
#define ORC_HPLANK 6.626068E-34     /* :25 */
#define ORC_PI 3.14159265358979323846

/* Grain temperature at R [AU] (src/gpu_functions.cu:1050-1064); the index is
 * kept inside the table where rounding at R just below Rmax could step past it. */
static double orc_temp(const orc_optics *o, int id, double R)
{
    const double f = log(o->Rmax / o->Rmin) / ((double)o->n_r - 1.0);
    const double *Td = o->Td + (int64_t)id * o->n_r;
    if (R <= o->Rmin) return Td[0];
    if (R >= o->Rmax) return Td[o->n_r - 1];
    int Rid = (int)(log(R / o->Rmin) / f) + 1;
    if (Rid > o->n_r - 1) Rid = o->n_r - 1;
    const double Rin = o->Rmin * exp((Rid - 1) * f);
    const double Rout = o->Rmin * exp(Rid * f);
    return exp(log(Td[Rid - 1]) + log(Td[Rid] / Td[Rid - 1]) * log(R / Rin) / log(Rout / Rin));
}

/* Flux [Jy] at wavelength j (src/gpu_functions.cu:1065-1080). */
static double orc_flux(const orc_optics *o, double dist, int id, double n, double R, double cossc, double T, int j)
{
    const double wvl = o->w[j];
    const double I = (2 * ORC_HPLANK * ORC_CSPEED * ORC_CSPEED) /
                     (pow(wvl, 5) * (exp(ORC_HPLANK * ORC_CSPEED / (wvl * ORC_KBOLTZ * T)) - 1.0)) * ORC_PI;
    int cosinm = (int)floor(0.5 * (cossc + 1.0) * (o->n_theta - 1));
    if (cosinm < 0) cosinm = 0;
    if (cosinm > o->n_theta - 2) cosinm = o->n_theta - 2;
    const int cosinp = cosinm + 1;
    const double cosm = (double)(cosinm)*2.0 / (o->n_theta - 1) - 1.0;
    const double cosp = (double)(cosinp)*2.0 / (o->n_theta - 1) - 1.0;
    const double Pm = o->Qpfunc[(int64_t)o->n_wav * ((int64_t)o->n_theta * id + cosinm) + j];
    const double Pp = o->Qpfunc[(int64_t)o->n_wav * ((int64_t)o->n_theta * id + cosinp) + j];
    const double P = exp(log(Pm) + (log(Pp) - log(Pm)) * (cossc - cosm) / (cosp - cosm));
    const double s = o->size[id];
    const double Fscat = o->in[j] * ORC_PI * s * s * o->Qsca[(int64_t)o->n_wav * id + j] * P / R / R / ORC_AU / ORC_AU;
    const double Fther = I * s * s * o->Qabs[(int64_t)o->n_wav * id + j] / dist / dist;
    return n * (Fther + Fscat) * 1e26 * wvl * wvl / ORC_CSPEED;
}

/* Scattering cosine (src/gpu_functions.cu:1010-1023). */
static double orc_cossc(const orc_optics *o, const orc_view *v, double x, double y, double z, double R)
{
    const double ly = -v->dist * sin(v->incl) - y;
    const double lz = v->dist * cos(v->incl) - z;
    double c = (x * x + (-y) * ly + (-z) * lz) / R / sqrt(x * x + ly * ly + lz * lz);
    if (c < -1) c = -1.0;
    if (c >= 1.0 - 2.0 / o->n_theta) c = 1.0 - 2.0 / o->n_theta;
    return c;
}

void orc_image(const orc_optics *o, const orc_view *v, int64_t n_grav, int64_t n_dust, const double *x,
               const double *y, const double *z, const double *N, const int32_t *id, int32_t wj, double *F)
{
    const int64_t npix = (int64_t)v->xsize * v->ysize;
    for (int64_t k = 0; k < npix; k++) F[k] = 0.0;
    for (int64_t i = 0; i < n_dust; i++) {
        const int64_t t = n_grav + i;
        const int b = id ? id[t] : -1;
        if (b < 0 || b >= o->n_bins) continue;   /* (:994) */
        double X = x[t], Y = y[t];
        const double Z = z[t], n = N[t];
        double phi = atan2(Y, X);
        const double R = sqrt(X * X + Y * Y + Z * Z);
        double Rd = sqrt(X * X + Y * Y);
        X = Rd * cos(phi + v->rot);   /* (:1005-1006) */
        Y = Rd * sin(phi + v->rot);
        const double cossc = orc_cossc(o, v, X, Y, Z, R);
        Y = sqrt(Y * Y + Z * Z) * cos(atan2(Z, Y) - v->incl);   /* (:1027) */
        phi = atan2(Y, X);   /* (:1030-1034) */
        Rd = sqrt(X * X + Y * Y);
        X = Rd * cos(phi + v->PA);
        Y = Rd * sin(phi + v->PA);
        const double psca = v->xsize >= v->ysize ? v->fov / v->xsize : v->fov / v->ysize;   /* (:1036-1041) */
        X = X + 0.5 * psca * v->xsize;
        Y = Y + 0.5 * psca * v->ysize;
        const double fx = floor(X / psca), fy = floor(Y / psca);
        if (!(fx >= 0.0 && fx <= v->xsize - 1 && fy >= 0.0 && fy <= v->ysize - 1)) continue;   /* (:1047) */
        if (!(R > o->Rmin)) continue;   /* (:1078) */
        const double T = orc_temp(o, b, R);
        F[(int64_t)v->ysize * (int64_t)fx + (int64_t)fy] += orc_flux(o, v->dist, b, n, R, cossc, T, wj);
    }
}

void orc_sed(const orc_optics *o, const orc_view *v, int64_t n_grav, int64_t n_dust, const double *x,
             const double *y, const double *z, const double *N, const int32_t *id, double *Fexc)
{
    for (int j = 0; j < o->n_wav; j++) Fexc[j] = 0.0;
    for (int64_t i = 0; i < n_dust; i++) {
        const int64_t t = n_grav + i;
        const int b = id ? id[t] : -1;
        if (b < 0 || b >= o->n_bins) continue;   /* (:1092) */
        double X = x[t], Y = y[t];
        const double Z = z[t], n = N[t];
        const double phi = atan2(Y, X);
        const double R = sqrt(X * X + Y * Y + Z * Z);
        X = R * cos(phi + v->rot);   /* the 3-D distance (:1104-1105) */
        Y = R * sin(phi + v->rot);
        const double cossc = orc_cossc(o, v, X, Y, Z, R);
        if (!(R > o->Rmin)) continue;   /* Fdust = 0 (:1151-1154) */
        const double T = orc_temp(o, b, R);
        for (int j = 0; j < o->n_wav; j++) Fexc[j] += orc_flux(o, v->dist, b, n, R, cossc, T, j);
    }
}

/* ------------------------------------------------------------------------
 * Mie scattering (tools/mie_single.c:54-224): the logarithmic derivative
 * D_n(m x) by downward recurrence from n_max, the Riccati-Bessel functions
 * psi_n, chi_n upward, the coefficients a_n, b_n, the efficiency sums and
 * the amplitude sums S+ = S1 + S2, S- = S1 - S2 with the angular functions
 * pi_n, tau_n, in the reference's order.  Complex numbers are written out as
 * real and imaginary parts, as there.
 * ---------------------------------------------------------------------- */
#define ORC_MIE_PI 3.141592653589793238
#define ORC_MIE_MAXX 7200000.0

int orc_mie(double s, const double *lam, const double *nre, const double *kim, int nl, const double *cost, int nt,
            double *Qabs, double *Qsca, double *Qpfunc)
{
    double *Spr = calloc((size_t)nt, sizeof(double)), *Spi = calloc((size_t)nt, sizeof(double));
    double *Smr = calloc((size_t)nt, sizeof(double)), *Smi = calloc((size_t)nt, sizeof(double));
    double *P0r = calloc((size_t)nt, sizeof(double)), *P0i = calloc((size_t)nt, sizeof(double));
    double *P1r = calloc((size_t)nt, sizeof(double)), *P1i = calloc((size_t)nt, sizeof(double));
    int rc = 0;
    for (int i = 0; i < nl && rc == 0; i++) {   /* (:93) */
        const double x = 2.0 * ORC_MIE_PI * s / lam[i];
        if (x > ORC_MIE_MAXX) { rc = -1; break; }
        const double cr = nre[i], ci = -kim[i];
        const double Irr = cr / (cr * cr + ci * ci), Iri = -ci / (cr * cr + ci * ci);
        const double Yr = x * cr, Yi = x * ci;
        double nstop;   /* (:105-112) */
        if (x < 0.02) nstop = 2.0;
        else if (x <= 8.0) nstop = x + 4.0 * pow(x, 0.3333333333333333) + 2.0;
        else if (x < 4200.0) nstop = x + 4.05 * pow(x, 0.3333333333333333) + 2.0;
        else nstop = x + 4.0 * pow(x, 0.3333333333333333) + 2.0;
        double t = sqrt(Yr * Yr + Yi * Yi);
        t = (nstop > t) ? nstop : t;
        t += 15.0;
        const long nmax = (long)floor(t);
        double *dr = calloc((size_t)nmax + 1, sizeof(double)), *di = calloc((size_t)nmax + 1, sizeof(double));
        for (long n = nmax - 1; n > 0; n--) {   /* (:124-129) */
            const double tr = (((double)(n + 1)) * Yr) / (Yr * Yr + Yi * Yi);
            const double ti = (-((double)(n + 1)) * Yi) / (Yr * Yr + Yi * Yi);
            const double ur = tr + dr[n + 1], ui = ti + di[n + 1];
            dr[n] = tr - (ur / (ur * ur + ui * ui));
            di[n] = ti - (-ui / (ur * ur + ui * ui));
        }
        for (int k = 0; k < nt; k++) { Spr[k] = Spi[k] = Smr[k] = Smi[k] = P0r[k] = P0i[k] = P1i[k] = 0.0; P1r[k] = 1.0; }
        double Psi0 = cos(x), Psi1 = sin(x), Chi0 = -Psi1, Chi1 = Psi0;
        double Dg = 0.0, Dqsc = 0.0, Dqxt = 0.0, Tnp1 = 1.0;
        double Ar1 = 0.0, Ai1 = 0.0, Br1 = 0.0, Bi1 = 0.0;   /* a_{n-1}, b_{n-1} */
        for (long n = 1; n <= nstop; n++) {   /* (:147) */
            Tnp1 += 2.0;
            const double Tnm1 = Tnp1 - 2.0;
            const double DN = (double)n;
            const double A2 = Tnp1 / (DN * (DN + 1.0));
            const double Turbo = (DN + 1.0) / DN;
            const double Rnx = n / x;
            const double Psi = Tnm1 * (Psi1 / x) - Psi0;
            const double Chi = Tnm1 * (Chi1 / x) - Chi0;
            double nr, ni, qr, qi, q2r, q2i, den, Ar, Ai, Br, Bi;
            /* a_n with D_n / m (:162-173) */
            nr = ((dr[n] * Irr - di[n] * Iri) + Rnx) * Psi - Psi1;
            ni = (dr[n] * Iri + di[n] * Irr) * Psi;
            qr = ((dr[n] * Irr - di[n] * Iri) + Rnx);
            qi = (dr[n] * Iri + di[n] * Irr);
            q2r = qr * Psi - qi * Chi;
            q2i = qr * Chi + qi * Psi;
            qr = q2r - Psi1;
            qi = q2i - Chi1;
            den = 1. / (qr * qr + qi * qi);
            Ar = (nr * qr + ni * qi) * den;
            Ai = (ni * qr - nr * qi) * den;
            /* b_n with m D_n (:176-187) */
            nr = ((dr[n] * cr - di[n] * ci) + Rnx) * Psi - Psi1;
            ni = (dr[n] * ci + di[n] * cr) * Psi;
            qr = ((dr[n] * cr - di[n] * ci) + Rnx);
            qi = (dr[n] * ci + di[n] * cr);
            q2r = qr * Psi - qi * Chi;
            q2i = qr * Chi + qi * Psi;
            qr = q2r - Psi1;
            qi = q2i - Chi1;
            den = 1. / (qr * qr + qi * qi);
            Br = (nr * qr + ni * qi) * den;
            Bi = (ni * qr - nr * qi) * den;
            Dqxt = Tnp1 * (Ar + Br) + Dqxt;
            Dqsc = Tnp1 * (Ar * Ar + Ai * Ai + Br * Br + Bi * Bi) + Dqsc;
            if (n > 1)
                Dg = Dg + ((DN * DN - 1.0) / DN) * ((Ar1 * Ar + Ai1 * Ai) + (Br1 * Br + Bi1 * Bi)) +
                     Tnm1 * (Ar1 * Br1 + Ai1 * Bi1) / (DN * DN - DN);
            Ar1 = Ar; Ai1 = Ai; Br1 = Br; Bi1 = Bi;
            for (int k = 0; k < nt; k++) {   /* (:201-219) */
                const double Sr = cost[k] * P1r[k], Si = cost[k] * P1i[k];
                const double Tr = Sr - P0r[k], Ti = Si - P0i[k];
                const double Tar = n * Tr - P0r[k], Tai = n * Ti - P0i[k];
                const double aAr = A2 * Ar, aAi = A2 * Ai, aBr = A2 * Br, aBi = A2 * Bi;
                Spr[k] = (aAr + aBr) * (P1r[k] + Tar) - (aAi + aBi) * (P1i[k] + Tai) + Spr[k];
                Spi[k] = (aAi + aBi) * (P1r[k] + Tar) + (aAr + aBr) * (P1i[k] + Tai) + Spi[k];
                Smr[k] = (aAr - aBr) * (P1r[k] - Tar) - (aAi - aBi) * (P1i[k] - Tai) + Smr[k];
                Smi[k] = (aAi - aBi) * (P1r[k] - Tar) + (aAr - aBr) * (P1i[k] - Tai) + Smi[k];
                P0r[k] = P1r[k];
                P0i[k] = P1i[k];
                P1r[k] = Sr + Turbo * Tr;
                P1i[k] = Si + Turbo * Ti;
            }
            Psi0 = Psi1; Psi1 = Psi; Chi0 = Chi1; Chi1 = Chi;
        }
        free(dr);
        free(di);
        for (int k = 0; k < nt; k++) {   /* (:232-241) */
            const double X1r = (Spr[k] + Smr[k]) * 0.5, X1i = (Spi[k] + Smi[k]) * 0.5;
            const double X2r = (Spr[k] - Smr[k]) * 0.5, X2i = (Spi[k] - Smi[k]) * 0.5;
            Qpfunc[i + (long)nl * k] = ((X1r * X1r) + (X1i * X1i) + (X2r * X2r) + (X2i * X2i)) / (4 * ORC_MIE_PI * Dqsc);
        }
        Dqsc = 2.0 * Dqsc / (x * x);
        Dqxt = 2.0 * Dqxt / (x * x);
        Qabs[i] = Dqxt - Dqsc;
        Qsca[i] = Dqsc;
        (void)Dg;
    }
    free(Spr); free(Spi); free(Smr); free(Smi); free(P0r); free(P0i); free(P1r); free(P1i);
    return rc;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
