// ddyn_hip.hip -- MI355X (gfx950) kernels and C ABI for DiskDyn's dynamical hot path.
//
// One RK4 step runs, in this order (the reference order of src/grid.cu:387-389):
//   for stage s = 0..3 of the massive set:
//     k_pairs64 / k_pairs32   swarm-sourced pair forces on every massive body,
//                             i-parallel over targets and split over source slices,
//                             LDS-tiled sources, one FP64 partial per (slice, target)
//     k_stage<s>              (stages 0-2) FP64 star/planet direct terms + slice
//                             partials, then the RK4 stage update that writes the
//                             next stage's positions (FP64 and an FP32 shadow)
//   k_dust                    all four grain stages (RK4calc_dust,
//                             src/gpu_functions.cu:627-943) against the massive
//                             stage positions; four waves per 64 grains, each over
//                             a quarter of the swarm (scalar-loaded sources), star
//                             and planet terms in FP64, RK4 bookkeeping in LDS
//   k_stage<3>                massive RK4 combination (RK4calc_massive_final,
//                             src/gpu_functions.cu:613-625) fused into stage 4.
//
// With grains, ddyn_step's default schedule runs k_dust of step k on a second
// stream beside the massive stages of step k+1 (overlapped_step).  Without a
// swarm and with at most 32 massive bodies (star and planets only), the default
// schedule is k_small_steps instead: up to 1,024 whole steps per launch, each
// workgroup integrating the massive set redundantly beside its 64 grains.
//
// The reference evaluates massive pairs on 32x32 tile pairs with Newton's third
// law and CAS atomics (src/gpu_functions.cu:5-13, 101-217); here every target sums
// its own sources in a fixed order, so results are deterministic and no atomics
// are needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "ddyn.h"
#include "ddyn_ctx.h"

namespace {

constexpr int kWG = 256;            // threads per workgroup (4 waves of 64)
constexpr int kTile = 256;          // sources per LDS tile in the massive pair kernels
#ifndef DDYN_DUST_SPLIT
#define DDYN_DUST_SPLIT 4
#endif
#ifndef DDYN_DUST_TILE
#define DDYN_DUST_TILE 128
#endif
constexpr int kDustSplit = DDYN_DUST_SPLIT;   // grain kernel: waves per workgroup, each over a share of the swarm
constexpr int kDustWG = 64 * kDustSplit;      // 64 grains per workgroup, one per lane
constexpr int kDustTile = DDYN_DUST_TILE;     // grain kernel: sources per FP32 partial and close-pair test
#ifndef DDYN_PAIR_TARGET_WGS
#define DDYN_PAIR_TARGET_WGS 65536
#endif
constexpr int kTargetWGs = DDYN_PAIR_TARGET_WGS;   // pair-kernel workgroups wanted per launch: short workgroups keep
                                    // the launch's tail short (measured on the 65,537- and
                                    // 1,048,577-body swarms, 16,384 to 65,536)
constexpr int kMinSlice = 1024;     // fewest sources per slice (bounds the partials' traffic)
constexpr int kAdjPullMin = 64;     // more massive bodies than this: grain pulls of the adjoint as a pair kernel
constexpr double kLightAU = 63239.6717367;   // src/ddyn.cuh:22
constexpr double kFar = 1.0e15;     // position of padding sources/targets [AU]
constexpr float kTiny2 = 1.0e-20f;  // FP32 r^2 floor for unsoftened pairs (keeps self pairs finite)
constexpr int kMaxBins = 2700;      // nsize, src/ddyn.cuh:35

// ---------------------------------------------------------------------------
// Massive swarm-pair kernels
// 1/sqrt(x) for x > 0 in FP64 (the adjoint kernels and every pair law of the
// DDYN_FP64_RSQ form): the hardware estimate v_rsq_f64 (relative error up to
// 5.2e-8) refined by one third-order step y <- y + y e (1/2 + 3e/8) with
// e = 1 - x y^2, the series of (1 - e)^(-1/2) to e^2.  The estimate's error
// cubed is far below FP64 rounding: relative error at most 1.4e-16 (1.24 x
// 2^-53) over 2^22 inputs spread over 24 decades, as two Newton steps give, in
// five operations instead of eight.  Cheaper than an IEEE square root and division.
__device__ inline double rsqrt_fp64(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = __builtin_fma(-x * y, y, 1.0);
    return __builtin_fma(y * e, __builtin_fma(0.375, e, 0.5), y);
}

// mu / R^3 of an unsoftened (or softened) pair with R^2 = r2, 0 when R^3 = 0:
// an IEEE square root and division (the plain form), or rsqrt_fp64 with R^2
// clamped at 1e-200 (DDYN_FP64_RSQ), where a coincident pair (d = 0) still
// contributes d * q = 0 without a select; only separations below 1e-100 AU
// are affected.
template <bool RSQ>
__device__ inline double pair_q(double r2, double mu)
{
    if (RSQ) {
        const double y = rsqrt_fp64(__builtin_fmax(r2, 1e-200));
        return mu * (y * y * y);
    }
    const double f = r2 * __builtin_sqrt(r2);
    return (f != 0.0) ? mu / f : 0.0;
}

// ---------------------------------------------------------------------------
// Targets: every massive body i in [0, n_grav).  Sources: swarm bodies j in the
// slice [n_planets + y*slice_len, ...).  Lane l of wave w handles targets
// wbase + l + 64k, k < TPT, so each wave covers one contiguous index range.
// Pair law (src/gpu_functions.cu:157-172): a_i += mu_j (r_j - r_i) / R^3 with
// R^2 = |r_j - r_i|^2 + eps^2 for swarm targets (eps = 0 for star/planet targets),
// and a zero contribution when R^3 == 0 (pair_q; RSQ: the DDYN_FP64_RSQ form).

// The massive-body kernels (k_pairs64, k_pairs32, k_stage) raise their waves'
// issue priority.  With grains they run beside the grain kernel of the
// previous step (overlapped schedule), one or two waves per SIMD among seven
// grain waves: at equal priority they get about an eighth of the issue slots,
// each of the eight launches of a step takes 40 us instead of 5, and the
// chain becomes the critical path of the step.  Alone on the GPU (no grains)
// every wave has the same priority and nothing changes.
__device__ __forceinline__ void massive_priority()
{
    __builtin_amdgcn_s_setprio(3);
}

template <int TPT, bool RSQ>
__global__ __launch_bounds__(kWG) void k_pairs64(const double4* __restrict__ pos, int n_grav,
                                                 int n_planets, int slice_len, double eps2,
                                                 double* __restrict__ part)
{
    __shared__ double4 tile[kTile];
    massive_priority();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wbase = blockIdx.x * (kWG * TPT) + wave * (64 * TPT);
    const int jbeg = n_planets + blockIdx.y * slice_len;
    const int jend = min(jbeg + slice_len, n_grav);

    double xi[TPT], yi[TPT], zi[TPT], e2[TPT], ax[TPT], ay[TPT], az[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int i = wbase + lane + 64 * k;
        const double4 p = (i < n_grav) ? pos[i] : make_double4(kFar, kFar, kFar, 0.0);
        xi[k] = p.x; yi[k] = p.y; zi[k] = p.z;
        e2[k] = (i >= n_planets) ? eps2 : 0.0;
        ax[k] = 0.0; ay[k] = 0.0; az[k] = 0.0;
    }
    for (int j0 = jbeg; j0 < jend; j0 += kTile) {
        const int tl = min(kTile, jend - j0);
        __syncthreads();
        tile[threadIdx.x] = ((int)threadIdx.x < tl) ? pos[j0 + threadIdx.x]
                                               : make_double4(kFar, kFar, kFar, 0.0);
        __syncthreads();
#pragma unroll 2
        for (int jj = 0; jj < kTile; ++jj) {
            const double4 s = tile[jj];
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                const double dx = s.x - xi[k], dy = s.y - yi[k], dz = s.z - zi[k];
                const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx)) + e2[k];
                const double q = pair_q<RSQ>(r2, s.w);
                ax[k] = __builtin_fma(dx, q, ax[k]);
                ay[k] = __builtin_fma(dy, q, ay[k]);
                az[k] = __builtin_fma(dz, q, az[k]);
            }
        }
    }
    const size_t n = (size_t)n_grav;
    double* px = part + (size_t)blockIdx.y * 3 * n;
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int i = wbase + lane + 64 * k;
        if (i < n_grav) { px[i] = ax[k]; px[n + i] = ay[k]; px[2 * n + i] = az[k]; }
    }
}

// FP64 evaluation of one source tile for TPT targets (close pairs and the
// diagonal tiles of the FP32 kernel).  Sources come from global memory.
template <int TPT>
__device__ inline void tile_fp64(const double4* __restrict__ pos, int j0, int j1,
                                 const double* xi, const double* yi, const double* zi,
                                 const double* e2, double* ax, double* ay, double* az)
{
    for (int j = j0; j < j1; ++j) {
        const double4 s = pos[j];
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            const double dx = s.x - xi[k], dy = s.y - yi[k], dz = s.z - zi[k];
            const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx)) + e2[k];
            const double f = r2 * __builtin_sqrt(r2);
            const double q = (f != 0.0) ? s.w / f : 0.0;
            ax[k] = __builtin_fma(dx, q, ax[k]);
            ay[k] = __builtin_fma(dy, q, ay[k]);
            az[k] = __builtin_fma(dz, q, az[k]);
        }
    }
}

// FP32 sum of one LDS tile of kTile sources for TPT targets.  With SELF, the
// tile holds the wave's own targets: the pair with source slot self0 + 64k
// (the target itself) is masked to r^2 = 3e38, so it adds exactly zero
// (inv^3 underflows) and is not seen by the close-pair minimum.
template <int TPT, bool SELF>
__device__ inline void tile_sum32(const float4* __restrict__ tile, int self0, const float* xf, const float* yf,
                                  const float* zf, const float* e2f, float* fx, float* fy, float* fz, float* rmin)
{
    auto one = [&](int jj) {
        const float4 s = tile[jj];
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            const float dx = s.x - xf[k], dy = s.y - yf[k], dz = s.z - zf[k];
            float r2 = __builtin_fmaf(dx, dx, e2f[k]);
            r2 = __builtin_fmaf(dy, dy, r2);
            r2 = __builtin_fmaf(dz, dz, r2);
            if (SELF) r2 = (jj == self0 + 64 * k) ? 3.0e38f : r2;
            rmin[k] = fminf(rmin[k], r2);
            const float inv = __builtin_amdgcn_rsqf(r2);
            const float inv2 = inv * inv;
            const float q = (s.w * inv) * inv2;
            fx[k] = __builtin_fmaf(dx, q, fx[k]);
            fy[k] = __builtin_fmaf(dy, q, fy[k]);
            fz[k] = __builtin_fmaf(dz, q, fz[k]);
        }
    };
    // Eight sources per loop iteration in two groups of four that the
    // scheduler may not merge: the pair law's transcendental and accumulation
    // chains issue fastest in groups of four, and a longer loop body issues
    // slower again.
#pragma unroll 1
    for (int jj = 0; jj < kTile; jj += 8) {
#pragma unroll
        for (int u = 0; u < 4; ++u) one(jj + u);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 4; u < 8; ++u) one(jj + u);
    }
}

// FP32-force form.  Each source tile is summed in FP32 (v_rsq_f32) into a
// per-tile partial that is flushed into the FP64 accumulators.  A tile is
// instead evaluated in FP64 for the whole wave when any pair of the wave came
// closer than the FP64 radius (r^2 < r64^2 + eps^2, tracked as a running
// minimum).  In a tile that holds the wave's own targets the self pairs are
// masked out of both the sum and the minimum.
template <int TPT>
__global__ __launch_bounds__(kWG) void k_pairs32(const float4* __restrict__ pos32,
                                                 const double4* __restrict__ pos64, int n_grav,
                                                 int n_planets, int slice_len, float eps2f,
                                                 float r64sq, double eps2, double* __restrict__ part)
{
    __shared__ float4 tile[kTile];
    massive_priority();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wbase = blockIdx.x * (kWG * TPT) + wave * (64 * TPT);
    const int wend = wbase + 64 * TPT;
    const int jbeg = n_planets + blockIdx.y * slice_len;
    const int jend = min(jbeg + slice_len, n_grav);

    // FP64 target positions are read only when a tile is redone in FP64
    float xf[TPT], yf[TPT], zf[TPT], e2f[TPT], thr[TPT];
    double ax[TPT], ay[TPT], az[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int i = wbase + lane + 64 * k;
        const bool ok = i < n_grav;
        const float4 q = ok ? pos32[i] : make_float4((float)kFar, (float)kFar, (float)kFar, 0.f);
        xf[k] = q.x; yf[k] = q.y; zf[k] = q.z;
        const bool sw = i >= n_planets;
        e2f[k] = sw ? fmaxf(eps2f, kTiny2) : kTiny2;
        thr[k] = r64sq + e2f[k];
        ax[k] = 0.0; ay[k] = 0.0; az[k] = 0.0;
    }
    for (int j0 = jbeg; j0 < jend; j0 += kTile) {
        const int tl = min(kTile, jend - j0);
        __syncthreads();
        tile[threadIdx.x] = ((int)threadIdx.x < tl)
                                ? pos32[j0 + threadIdx.x]
                                : make_float4((float)kFar, (float)kFar, (float)kFar, 0.f);
        __syncthreads();
        // wave-uniform: does the tile hold one of this wave's own targets?
        const bool own = (j0 < wend) && (j0 + tl > wbase);
        bool redo = false;
        {
            float fx[TPT], fy[TPT], fz[TPT], rmin[TPT];
#pragma unroll
            for (int k = 0; k < TPT; ++k) { fx[k] = 0.f; fy[k] = 0.f; fz[k] = 0.f; rmin[k] = 3.0e38f; }
            if (own)
                tile_sum32<TPT, true>(tile, wbase + lane - j0, xf, yf, zf, e2f, fx, fy, fz, rmin);
            else
                tile_sum32<TPT, false>(tile, 0, xf, yf, zf, e2f, fx, fy, fz, rmin);
            bool close = false;
#pragma unroll
            for (int k = 0; k < TPT; ++k) close = close || (rmin[k] < thr[k]);
            redo = __any(close);
            if (!redo) {
#pragma unroll
                for (int k = 0; k < TPT; ++k) {
                    ax[k] += (double)fx[k]; ay[k] += (double)fy[k]; az[k] += (double)fz[k];
                }
            }
        }
        if (redo) {
            double xi[TPT], yi[TPT], zi[TPT], e2d[TPT];
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                const int i = wbase + lane + 64 * k;
                const double4 p = (i < n_grav) ? pos64[i] : make_double4(kFar, kFar, kFar, 0.0);
                xi[k] = p.x; yi[k] = p.y; zi[k] = p.z;
                e2d[k] = (i >= n_planets) ? eps2 : 0.0;
            }
            tile_fp64<TPT>(pos64, j0, j0 + tl, xi, yi, zi, e2d, ax, ay, az);
        }
    }
    const size_t n = (size_t)n_grav;
    double* px = part + (size_t)blockIdx.y * 3 * n;
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int i = wbase + lane + 64 * k;
        if (i < n_grav) { px[i] = ax[k]; px[n + i] = ay[k]; px[2 * n + i] = az[k]; }
    }
}

// ---------------------------------------------------------------------------
// Massive stage epilogue: direct terms + slice partials + RK4 update
// ---------------------------------------------------------------------------
struct StageArgs {
    int n_grav, n_planets, nslice;
    const double* part;
    double *x, *y, *z, *vx, *vy, *vz;
    const double* mu;
    double4* p64[4];
    float4* p32[4];
    double *vrx, *vry, *vrz;     // stage velocity k{s}r
    double *Srx, *Sry, *Srz;     // v + 2 v2 + 2 v3 + v4, built left to right
    double *Sax, *Say, *Saz;     // a1 + 2 a2 + 2 a3 + a4, built left to right
    double *ox, *oy, *oz;        // accel output (STAGE == 4)
    double4* n64;                // stage 3: where the next step's stage-0 positions go
    float4* n32;
    double h;                    // dt/2, dt/2, dt for stages 0-2; dt/6 for stage 3
};

__device__ inline void store_pos(double4* p64, float4* p32, int i, double x, double y, double z,
                                 double mu)
{
    p64[i] = make_double4(x, y, z, mu);
    p32[i] = make_float4((float)x, (float)y, (float)z, (float)mu);
}

// STAGE 0..3: RK4 stage s; STAGE 4: write accelerations only.  RSQ: the
// DDYN_FP64_RSQ form of the direct terms (pair_q).
template <int STAGE, bool RSQ>
__global__ __launch_bounds__(kWG) void k_stage(StageArgs A)
{
    massive_priority();
    const int i = blockIdx.x * kWG + threadIdx.x;
    if (i >= A.n_grav) return;
    const double4* P = A.p64[STAGE == 4 ? 0 : STAGE];
    const double4 pi = P[i];
    double ax = 0.0, ay = 0.0, az = 0.0;
    // Star and planets, FP64 and unsoftened (src/gpu_functions.cu:157-172).
    for (int j = 0; j < A.n_planets; ++j) {
        const double4 pj = P[j];
        const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
        const double q = pair_q<RSQ>(r2, pj.w);
        ax = __builtin_fma(dx, q, ax);
        ay = __builtin_fma(dy, q, ay);
        az = __builtin_fma(dz, q, az);
    }
    const size_t n = (size_t)A.n_grav;
    // slice partials in slice order; unrolled so that many loads are in flight
#pragma unroll 8
    for (int k = 0; k < A.nslice; ++k) {
        const double* pk = A.part + (size_t)k * 3 * n;
        ax += pk[i]; ay += pk[n + i]; az += pk[2 * n + i];
    }
    if (STAGE == 4) { A.ox[i] = ax; A.oy[i] = ay; A.oz[i] = az; return; }

    const double x = A.x[i], y = A.y[i], z = A.z[i];
    const double vx = A.vx[i], vy = A.vy[i], vz = A.vz[i];
    const double h = A.h;
    if (STAGE == 0) {   // massive3 (src/gpu_functions.cu:219-230) with k1r = v
        const double rx = __builtin_fma(ax, h, vx), ry = __builtin_fma(ay, h, vy),
                     rz = __builtin_fma(az, h, vz);
        store_pos(A.p64[1], A.p32[1], i, __builtin_fma(vx, h, x), __builtin_fma(vy, h, y),
                  __builtin_fma(vz, h, z), pi.w);
        A.vrx[i] = rx; A.vry[i] = ry; A.vrz[i] = rz;
        A.Srx[i] = __builtin_fma(2.0, rx, vx);
        A.Sry[i] = __builtin_fma(2.0, ry, vy);
        A.Srz[i] = __builtin_fma(2.0, rz, vz);
        A.Sax[i] = ax; A.Say[i] = ay; A.Saz[i] = az;
    } else if (STAGE == 1 || STAGE == 2) {   // massive5 / massive7 (:350-362, :482-493)
        const double ux = A.vrx[i], uy = A.vry[i], uz = A.vrz[i];
        const double rx = __builtin_fma(ax, h, vx), ry = __builtin_fma(ay, h, vy),
                     rz = __builtin_fma(az, h, vz);
        store_pos(A.p64[STAGE + 1], A.p32[STAGE + 1], i, __builtin_fma(ux, h, x),
                  __builtin_fma(uy, h, y), __builtin_fma(uz, h, z), pi.w);
        if (STAGE == 1) {
            A.vrx[i] = rx; A.vry[i] = ry; A.vrz[i] = rz;
            A.Srx[i] = __builtin_fma(2.0, rx, A.Srx[i]);
            A.Sry[i] = __builtin_fma(2.0, ry, A.Sry[i]);
            A.Srz[i] = __builtin_fma(2.0, rz, A.Srz[i]);
        } else {
            A.Srx[i] = A.Srx[i] + rx; A.Sry[i] = A.Sry[i] + ry; A.Srz[i] = A.Srz[i] + rz;
        }
        A.Sax[i] = __builtin_fma(2.0, ax, A.Sax[i]);
        A.Say[i] = __builtin_fma(2.0, ay, A.Say[i]);
        A.Saz[i] = __builtin_fma(2.0, az, A.Saz[i]);
    } else {   // STAGE 3: massive_final (:613-625), h = dt/6
        const double nx = __builtin_fma(h, A.Srx[i], x), ny = __builtin_fma(h, A.Sry[i], y),
                     nz = __builtin_fma(h, A.Srz[i], z);
        A.x[i] = nx; A.y[i] = ny; A.z[i] = nz;
        A.vx[i] = __builtin_fma(h, A.Sax[i] + ax, vx);
        A.vy[i] = __builtin_fma(h, A.Say[i] + ay, vy);
        A.vz[i] = __builtin_fma(h, A.Saz[i] + az, vz);
        store_pos(A.n64, A.n32, i, nx, ny, nz, pi.w);
    }
}

// Stage-0 positions from the base state (after upload).
__global__ void k_init_pos(int n_grav, const double* x, const double* y, const double* z,
                           const double* mu, double4* p64, float4* p32)
{
    const int i = blockIdx.x * kWG + threadIdx.x;
    if (i < n_grav) store_pos(p64, p32, i, x[i], y[i], z[i], mu[i]);
}

// ---------------------------------------------------------------------------
// Grain kernel: RK4calc_dust (src/gpu_functions.cu:627-943), one grain per lane
// ---------------------------------------------------------------------------
struct DustArgs {
    int n_grav, n_planets, n_dust, n_bins;
    int blk0;               // k_dust: first 64-grain block of this launch
    const double4* p64[4];
    const float4* p32[4];
    double *x, *y, *z, *vx, *vy, *vz;
    const double* mu;
    double* N;
    const int32_t* id;
    const double *beta_p, *betasw_p;
    const float *beta_t, *betasw_t, *q_t;
    double dt, eps2, Rmin2, vsw, Wz, Bstar, Rstar, light;
    float eps2f, thr2f;
    int SWQ, FLQ;
    double *ox, *oy, *oz;   // accel mode outputs (n_dust)
    int32_t* removed;       // accel mode outputs (n_dust)
    double* mfin;           // k_small_steps: final massive x, y, z, vx, vy, vz (n_grav each)
};

// A grain's step-invariant factors of the star term, each computed with the
// same operations as inside the term, so hoisting them changes no bit:
// gfac = 1 - beta - beta_sw; drag = beta/c + beta_sw/v_sw with wind drag, else
// beta mu0 / c (src/gpu_functions.cu:691-699).
struct GrainCoef {
    double gfac, drag, qe, m;
};

__device__ inline GrainCoef grain_coef(const DustArgs& A, double bet, double betsw, double qe, double m,
                                       double m0)
{
    GrainCoef k;
    k.gfac = 1.0 - bet - betsw;
    k.drag = (A.SWQ == 1) ? bet / A.light + betsw / A.vsw : bet * m0 / A.light;
    k.qe = qe;
    k.m = m;
    return k;
}

// Star term with radiation pressure, PR and stellar-wind drag and the Lorentz
// force (src/gpu_functions.cu:679-718).  Returns true if the grain sublimates.
// P(c) and U(c) give coordinate c of the grain's stage position and velocity.
// LEAN (the grain kernel, whose VGPR peak this is) keeps the scheduler from
// overlapping the three divisions, which holds fewer registers live; the
// operations and results are the same.
template <bool RSQ, bool LEAN, class Pos, class Vel>
__device__ inline bool star_term_at(const DustArgs& A, const double4 s0, Pos P, Vel U, const GrainCoef& k,
                                    double& ax, double& ay, double& az)
{
    const double dx = P(0) - s0.x, dy = P(1) - s0.y, dz = P(2) - s0.z;
    const double RR = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
    if (RR <= A.Rmin2) return true;
    const double m0 = s0.w;
    const double du = __builtin_fma(dz, U(2), __builtin_fma(dy, U(1), dx * U(0)));
    double RR3 = 0.0, iR2 = 0.0, iR3 = 0.0, Fg, Fb1, Fb2;
    if (RSQ) {   // DDYN_FP64_RSQ: powers of one refined 1/R
        const double i1 = rsqrt_fp64(RR);
        iR2 = i1 * i1;
        iR3 = iR2 * i1;
        Fg = k.gfac * m0 * iR3;
        Fb1 = (A.SWQ == 1) ? m0 * iR2 * k.drag : k.drag * iR2;
        Fb2 = du * iR2;
    } else {
        RR3 = RR * __builtin_sqrt(RR);
        if (LEAN) __builtin_amdgcn_sched_barrier(0);
        Fg = k.gfac * m0 / RR3;
        if (LEAN) __builtin_amdgcn_sched_barrier(0);
        Fb1 = (A.SWQ == 1) ? m0 / RR * k.drag : k.drag / RR;
        if (LEAN) __builtin_amdgcn_sched_barrier(0);
        Fb2 = du / RR;
        if (LEAN) __builtin_amdgcn_sched_barrier(0);
    }
    ax = __builtin_fma(-Fg, dx, -(Fb1 * __builtin_fma(Fb2, dx, U(0))));
    ay = __builtin_fma(-Fg, dy, -(Fb1 * __builtin_fma(Fb2, dy, U(1))));
    az = __builtin_fma(-Fg, dz, -(Fb1 * __builtin_fma(Fb2, dz, U(2))));
    if (A.FLQ == 1) {
        // the Lorentz force is added last, so the gravity and drag terms are no
        // longer live while it is formed
        const double B3 = A.Bstar * A.Rstar * A.Rstar * A.Rstar;
        const double Bz = RSQ ? B3 * iR3 : B3 / RR3;
        ax = ax + k.qe * __builtin_fma(-A.Wz, dx, U(1)) * Bz / k.m;
        ay = ay + -k.qe * __builtin_fma(A.Wz, dy, U(0)) * Bz / k.m;
    }
    return false;
}

template <bool RSQ>
__device__ inline bool star_term(const DustArgs& A, const double4 s0, double px, double py,
                                 double pz, double ux, double uy, double uz, const GrainCoef& k,
                                 double& ax, double& ay, double& az)
{
    const double p[3] = {px, py, pz}, u[3] = {ux, uy, uz};
    return star_term_at<RSQ, false>(A, s0, [&](int c) { return p[c]; }, [&](int c) { return u[c]; }, k, ax, ay, az);
}

// FP64 pull of sources [j0, j1) on the grain; `soft` adds eps^2.  Sets `hit`
// on an exact coincidence (src/gpu_functions.cu:727-730).
template <bool RSQ>
__device__ inline void grain_fp64(const double4* __restrict__ P, int j0, int j1, double px,
                                  double py, double pz, double e2, double& ax, double& ay,
                                  double& az, bool& hit)
{
    for (int j = j0; j < j1; ++j) {
        const double4 s = P[j];
        const double ex = px - s.x, ey = py - s.y, ez = pz - s.z;
        const double RRu = __builtin_fma(ez, ez, __builtin_fma(ey, ey, ex * ex));
        hit = hit || (RRu == 0.0);
        const double R2 = RRu + e2;
        double f;
        if (RSQ) {
            const double y = rsqrt_fp64(R2);
            f = s.w * (y * y * y);
        } else {
            f = s.w / (R2 * __builtin_sqrt(R2));
        }
        ax = __builtin_fma(-f, ex, ax);
        ay = __builtin_fma(-f, ey, ay);
        az = __builtin_fma(-f, ez, az);
    }
}

// Sources are read through scalar loads: every lane of a wave sweeps the same
// source index, so a constant-address-space view of the stage positions lets
// the compiler fetch them with s_load and feed the VALU from SGPRs (no LDS
// staging and no barriers in the source loop).
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(4))) const f4v cf4v;
typedef __attribute__((address_space(4))) const d4v cd4v;

// Per-workgroup grain records in LDS: the RK4 bookkeeping of the 64 grains,
// the direct terms (star and planets) of the current stage and every wave's
// partial swarm acceleration.  Keeping this state in LDS instead of registers
// keeps the source loop lean (VGPRs for eight waves per SIMD).
#ifndef DDYN_DUST_WAVES_PER_EU
#define DDYN_DUST_WAVES_PER_EU 7
#endif
#ifndef DDYN_DUST_UNROLL
#define DDYN_DUST_UNROLL 8
#endif
#ifndef DDYN_DUST_F32_RSQ_DIRECT
#define DDYN_DUST_F32_RSQ_DIRECT 0
#endif
#ifndef DDYN_DUST_PREFETCH
#define DDYN_DUST_PREFETCH 1
#endif
#ifndef DDYN_DUST_GROUP
#define DDYN_DUST_GROUP 4
#endif
#ifndef DDYN_DUST_LEAD_CUT_F32
#define DDYN_DUST_LEAD_CUT_F32 3
#endif
#ifndef DDYN_DUST_LEAD_CUT_RSQ
#define DDYN_DUST_LEAD_CUT_RSQ 1
#endif
#ifndef DDYN_DUST_LEAD_CUT_F64
#define DDYN_DUST_LEAD_CUT_F64 0
#endif
constexpr int kDustGroup = DDYN_DUST_GROUP;   // grain kernel: sources per scalar-load group
static_assert(DDYN_DUST_TILE % (2 * DDYN_DUST_GROUP) == 0, "tile must hold whole pairs of groups");
static_assert(kDustSplit >= 2, "the grain kernel needs a lead wave and at least one other");
constexpr int kDustUnit = 2 * kDustGroup;    // sources per share unit: two scalar-load groups

// Sources [b, e) of wave w among ns swarm sources, in units of kDustUnit: the
// lead (last) wave takes ceil(units / waves) - cut units, the other waves
// split the rest as evenly as units allow, the larger shares first.  The split
// depends on the sizes only, so a grain's sums do not depend on its workgroup.
__device__ inline void dust_share(int ns, int cut, int w, int& b, int& e)
{
    const int nu = (ns + kDustUnit - 1) / kDustUnit;
    const int ul = max(0, min(nu, (nu + kDustSplit - 1) / kDustSplit - cut));
    const int nr = nu - ul;
    const int q = nr / (kDustSplit - 1), r = nr % (kDustSplit - 1);
    int ub = nr, ue = nu;
    if (w < kDustSplit - 1) {
        ub = w * q + min(w, r);
        ue = ub + q + (w < r ? 1 : 0);
    }
    b = min(ns, kDustUnit * ub);
    e = min(ns, kDustUnit * ue);
}
constexpr int kDustWavesPerEU = DDYN_DUST_WAVES_PER_EU;
struct DustShared {
    double x0[3][64], v0[3][64];     // state at the start of the step
    double p[3][64], u[3][64];       // current stage position / velocity
    double k0[3][64], k12[3][64];    // k1 and k2 + k3 (accelerations)
    double r12[3][64];               // r2 + r3 (stage velocities)
    double coef[3][64];              // GrainCoef: 1 - beta - beta_sw, drag factor, charge factor
    double d[3][64];                 // direct terms of the current stage (star and planets)
    double a[kDustSplit][3][64];     // per-wave partial swarm accelerations
    int32_t hit[64];                 // set (never cleared) when a term removes the grain
};

// A workgroup owns 64 grains, one per lane, in kDustSplit waves; each wave
// sweeps its share of the swarm sources for all 64 grains (so a launch has
// many short waves and its drain at the end is short), in FP32 partials over
// tiles of up to kDustTile sources.  The lead (last) wave also evaluates the
// FP64 star and planet terms, with a smaller share of the swarm.  After each
// stage waves 0-2 update one coordinate each: the total acceleration is the
// direct terms plus the wave partials in wave order.
template <bool F32, bool ACCEL, bool RSQ = false>
__global__ __launch_bounds__(kDustWG) __attribute__((amdgpu_waves_per_eu(kDustWavesPerEU, 8))) void k_dust(
    DustArgs A)
{
    __shared__ DustShared S;
    // the star and planet terms' inverse powers of R: refined hardware rsqrt in FP64_RSQ
    constexpr bool kDirectRsq = RSQ || (F32 && DDYN_DUST_F32_RSQ_DIRECT);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = (A.blk0 + (int)blockIdx.x) * 64 + lane;
    const int64_t t = (int64_t)A.n_grav + g;
    const bool valid = g < A.n_dust;
    const bool live = valid && (ACCEL || A.N[t] != 0.0);
    // This wave's share of the swarm sources (dust_share).  The last wave leads:
    // it also evaluates the star and planet terms, and gives up as many
    // sources as they cost, so the waves meet at the barriers together.
    constexpr int kLead = kDustSplit - 1;
    constexpr int kCut = F32 ? DDYN_DUST_LEAD_CUT_F32 : (RSQ ? DDYN_DUST_LEAD_CUT_RSQ : DDYN_DUST_LEAD_CUT_F64);
    const int ns = A.n_grav - A.n_planets;
    int jb, je;
    dust_share(ns, kCut, wave, jb, je);
    jb += A.n_planets;
    je += A.n_planets;
    unsigned parts = 0;   // waves holding sources
    for (int w = 0; w < kDustSplit; ++w) {
        int b, e;
        dust_share(ns, kCut, w, b, e);
        parts |= (b < e ? 1u : 0u) << w;
    }
    const double dt = A.dt, dt2 = dt / 2.0;
    if (wave == 0) {
        const double xs[3] = {valid ? A.x[t] : kFar, valid ? A.y[t] : kFar, valid ? A.z[t] : kFar};
        const double vs[3] = {valid ? A.vx[t] : 0.0, valid ? A.vy[t] : 0.0, valid ? A.vz[t] : 0.0};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            S.x0[c][lane] = xs[c]; S.v0[c][lane] = vs[c];
            S.p[c][lane] = xs[c]; S.u[c][lane] = vs[c];
        }
        double m = 1.0, bet = 0.0, betsw = 0.0, qe = 0.0;
        if (valid) {
            m = A.mu[t];
            const int32_t id = A.id ? A.id[t] : -1;
            if (A.beta_p) {
                bet = A.beta_p[g];
                betsw = A.betasw_p ? A.betasw_p[g] : 0.0;
            } else if (id >= 0 && id < A.n_bins) {
                bet = (double)A.beta_t[id];
                betsw = (double)A.betasw_t[id];
            }
            qe = (id >= 0 && id < A.n_bins) ? (double)A.q_t[id] : 0.0;
        }
        const GrainCoef k = grain_coef(A, bet, betsw, qe, m, A.p64[0][0].w);
        S.coef[0][lane] = k.gfac; S.coef[1][lane] = k.drag; S.coef[2][lane] = k.qe;
        S.hit[lane] = 0;
    }
    __syncthreads();
    bool gone = false;
    constexpr int kStages = ACCEL ? 1 : 4;
#pragma unroll 1
    for (int s = 0; s < kStages; ++s) {
        const bool act = live && !gone;
        if (wave == kLead) {
            // star and planets (src/gpu_functions.cu:679-730), FP64
            double ax = 0.0, ay = 0.0, az = 0.0;
            bool hit = false;
            if (act) {
                // the grain's mass enters only the Lorentz force; it is read from
                // the state rather than kept in the LDS record
                const GrainCoef k = {S.coef[0][lane], S.coef[1][lane], S.coef[2][lane],
                                     A.FLQ == 1 ? A.mu[t] : 1.0};
                hit = star_term_at<kDirectRsq, true>(
                    A, A.p64[s][0], [&](int c) { return S.p[c][lane]; }, [&](int c) { return S.u[c][lane]; }, k,
                    ax, ay, az);
                // the stage position is read again for the planets, so it is not
                // held in registers through the star term (the kernel's VGPR peak)
                __asm__ volatile("" ::: "memory");
                grain_fp64<kDirectRsq>(A.p64[s], 1, A.n_planets, S.p[0][lane], S.p[1][lane], S.p[2][lane], 0.0,
                                       ax, ay, az, hit);
            }
            S.d[0][lane] = ax; S.d[1][lane] = ay; S.d[2][lane] = az;
            if (hit) S.hit[lane] = 1;
        }
        bool hit = false;
        S.a[wave][0][lane] = 0.0; S.a[wave][1][lane] = 0.0; S.a[wave][2][lane] = 0.0;
        if (__builtin_amdgcn_readfirstlane(__any(act)) && jb < je) {
            if (F32) {
                const cf4v* src = (const cf4v*)A.p32[s];
                const float xf = (float)S.p[0][lane], yf = (float)S.p[1][lane], zf = (float)S.p[2][lane];
                const float e2f = fmaxf(A.eps2f, kTiny2);
                for (int j0 = jb; j0 < je; j0 += kDustTile) {
                    const int tl = min(kDustTile, je - j0);
                    float fx = 0.f, fy = 0.f, fz = 0.f, rmin = 3.0e38f;
                    auto pull = [&](const f4v sj) {
                        const float dx = sj.x - xf, dy = sj.y - yf, dz = sj.z - zf;
                        float r2 = __builtin_fmaf(dx, dx, e2f);
                        r2 = __builtin_fmaf(dy, dy, r2);
                        r2 = __builtin_fmaf(dz, dz, r2);
                        rmin = fminf(rmin, r2);
                        const float inv = __builtin_amdgcn_rsqf(r2);
                        const float inv2 = inv * inv;
                        const float q = (sj.w * inv) * inv2;
                        fx = __builtin_fmaf(dx, q, fx);
                        fy = __builtin_fmaf(dy, q, fy);
                        fz = __builtin_fmaf(dz, q, fz);
                    };
                    if (tl % kDustUnit == 0) {
#if DDYN_DUST_PREFETCH
                        // kDustGroup sources per group in scalar registers, two groups
                        // in turn.  Scalar loads return out of order, so a wait covers
                        // every load in flight: the first pull of a group comes
                        // before the next group's loads are issued (its wait then
                        // covers only its own group), and the group's other pulls
                        // run while those loads are in flight.
                        constexpr int G = kDustGroup;
                        const cf4v* sp = src + j0;
                        f4v ga[G], gb[G];
#pragma unroll
                        for (int u = 0; u < G; ++u) ga[u] = sp[u];
#pragma unroll 1
                        for (int jj = 0; jj < tl; jj += 2 * G) {
                            pull(ga[0]);
                            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                            for (int u = 0; u < G; ++u) gb[u] = sp[jj + G + u];
                            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                            for (int u = 1; u < G; ++u) pull(ga[u]);
                            pull(gb[0]);
                            __builtin_amdgcn_sched_barrier(0);
                            // past the tile's end the last group is loaded again and not used
                            const int jn = min(jj + 2 * G, tl - G);
#pragma unroll
                            for (int u = 0; u < G; ++u) ga[u] = sp[jn + u];
                            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                            for (int u = 1; u < G; ++u) pull(gb[u]);
                        }
#else
#pragma unroll DDYN_DUST_UNROLL
                        for (int jj = 0; jj < tl; ++jj) pull(src[j0 + jj]);
#endif
                    } else {
                        for (int jj = 0; jj < tl; ++jj) pull(src[j0 + jj]);
                    }
                    const bool redo = __any(act && rmin < A.thr2f);
                    if (act) {
                        // the FP64 accumulation of the wave's tiles lives in LDS
                        if (!redo) {
                            S.a[wave][0][lane] += (double)fx;
                            S.a[wave][1][lane] += (double)fy;
                            S.a[wave][2][lane] += (double)fz;
                        } else {
                            double ax = S.a[wave][0][lane], ay = S.a[wave][1][lane], az = S.a[wave][2][lane];
                            grain_fp64<false>(A.p64[s], j0, j0 + tl, S.p[0][lane], S.p[1][lane], S.p[2][lane],
                                              A.eps2, ax, ay, az, hit);
                            S.a[wave][0][lane] = ax; S.a[wave][1][lane] = ay; S.a[wave][2][lane] = az;
                        }
                    }
                }
            } else {
                const cd4v* src = (const cd4v*)A.p64[s];
                const double e2 = A.eps2;
                const double px = S.p[0][lane], py = S.p[1][lane], pz = S.p[2][lane];
                double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll 4
                for (int j = jb; j < je; ++j) {
                    const d4v sj = src[j];
                    const double ex = px - sj.x, ey = py - sj.y, ez = pz - sj.z;
                    const double RRu = __builtin_fma(ez, ez, __builtin_fma(ey, ey, ex * ex));
                    hit = hit || (RRu == 0.0);
                    const double R2 = RRu + e2;
                    double f;
                    if (RSQ) {
                        const double y = rsqrt_fp64(R2);
                        f = sj.w * (y * y * y);
                    } else {
                        f = sj.w / (R2 * __builtin_sqrt(R2));
                    }
                    ax = __builtin_fma(-f, ex, ax);
                    ay = __builtin_fma(-f, ey, ay);
                    az = __builtin_fma(-f, ez, az);
                }
                if (act) { S.a[wave][0][lane] = ax; S.a[wave][1][lane] = ay; S.a[wave][2][lane] = az; }
            }
        }
        // Removal is permanent, so one flag per grain serves every stage: a lane
        // that is still active at this stage found no hit at an earlier one.
        // Waves only ever store 1 into it.
        if (hit) S.hit[lane] = 1;
        __syncthreads();
        gone = gone || (act && S.hit[lane] != 0);
        if (wave < 3 && valid) {
            // coordinate c = wave: the direct terms, then the wave partials in wave order
            const int c = wave;
            double tot = S.d[c][lane];
            for (int w = 0; w < kDustSplit; ++w)
                if ((parts >> w) & 1u) tot += S.a[w][c][lane];
            if (ACCEL) {
                double* const O[3] = {A.ox, A.oy, A.oz};
                O[c][g] = gone ? 0.0 : tot;
                if (c == 0 && A.removed) A.removed[g] = gone ? 1 : 0;
            } else if (s < 3) {
                // next stage (src/gpu_functions.cu:737-743, 800-806, 863-869)
                const double x0 = S.x0[c][lane], v0 = S.v0[c][lane], uo = S.u[c][lane];
                if (s == 0) {
                    S.k0[c][lane] = tot;
                    S.p[c][lane] = __builtin_fma(v0, dt2, x0);
                    S.u[c][lane] = __builtin_fma(tot, dt2, v0);
                } else if (s == 1) {
                    const double r2v = __builtin_fma(tot, dt2, v0);
                    S.k12[c][lane] = tot;
                    S.p[c][lane] = __builtin_fma(uo, dt2, x0);
                    S.r12[c][lane] = uo + r2v;
                    S.u[c][lane] = r2v;
                } else {
                    S.k12[c][lane] = S.k12[c][lane] + tot;
                    S.p[c][lane] = __builtin_fma(uo, dt, x0);
                    S.u[c][lane] = __builtin_fma(tot, dt, v0);
                }
            } else if (live) {
                // evolve (src/gpu_functions.cu:927-933); removal sets N = 0
                if (gone) {
                    if (c == 0) A.N[t] = 0.0;
                } else {
                    const double dt6 = dt / 6.0;
                    const double x0 = S.x0[c][lane], v0 = S.v0[c][lane];
                    double* const X[3] = {A.x, A.y, A.z};
                    double* const V[3] = {A.vx, A.vy, A.vz};
                    X[c][t] = __builtin_fma(dt6, __builtin_fma(2.0, S.r12[c][lane], v0) + S.u[c][lane], x0);
                    V[c][t] = __builtin_fma(dt6, __builtin_fma(2.0, S.k12[c][lane], S.k0[c][lane]) + tot, v0);
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Many steps in one launch for small massive sets (no swarm, n_grav <= 32)
// ---------------------------------------------------------------------------
// With only a star and a few planets the massive integration is a few hundred
// FP64 operations per step, so every workgroup integrates it redundantly and
// advances its own 64 grains; nothing is exchanged between workgroups and one
// launch covers many steps.  The launch is latency-bound (a few workgroups,
// long FP64 square-root and division chains), so the workgroup is a two-wave
// pipeline: wave 0 (one massive body per lane) runs one step ahead and, while
// wave 1 advances the grains through step n against the four stage positions
// in one LDS buffer, finishes step n's massive update and writes step n+1's
// stage positions into the other buffer.  One barrier per step.  The
// arithmetic is that of k_stage and k_dust (same expressions, same order), so
// results equal the one-launch-per-kernel schedule bit for bit.
//
// With many grains the launch is throughput-bound instead. The massive wave,
// redundant in every workgroup, then takes wave slots from the grains (the
// kernel needs about 160 VGPRs, 3 waves per SIMD: a CU holds 6 grain waves
// and 6 massive waves), and every grain wave waits at the per-step barrier.
// From kSmallSplitMin grains the split form runs instead: k_small_massive
// integrates the star and planets once per launch and records their stage
// positions, then k_small_grains advances the grains alone (below).
constexpr int kSmallMax = 32;
constexpr int kSmallChunk = 1024;   // steps per launch of k_small_steps
constexpr int kSmallWG = 2 * 64;    // wave 0 massive, wave 1 grains
constexpr int64_t kSmallSplitMin = 131072;
constexpr int kSmallGrainWG = 256;  // k_small_grains: four waves of 64 grains

// Pair sum of the star and planets on massive body i at one stage
// (src/gpu_functions.cu:157-172, as in k_stage).
template <bool RSQ>
__device__ inline void small_massive_accel(const double4* __restrict__ P, int i, int np, double& ax,
                                           double& ay, double& az)
{
    const double4 pi = P[i];
    ax = 0.0; ay = 0.0; az = 0.0;
    for (int j = 0; j < np; ++j) {
        const double4 pj = P[j];
        const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
        const double q = pair_q<RSQ>(r2, pj.w);
        ax = __builtin_fma(dx, q, ax);
        ay = __builtin_fma(dy, q, ay);
        az = __builtin_fma(dz, q, az);
    }
}

// Within one wave, LDS stores are seen by the wave's later loads (its LDS
// operations run in order); this keeps the compiler from moving them across.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One RK4 step of a grain against the four stage positions Ps of the star and
// planets (k_dust in the FP64 form).  k_small_steps and k_small_grains both
// call it, so their results agree bit for bit.  A grain removed during the
// step (sublimation or an exact hit) sets removed and keeps its state.
template <bool RSQ>
__device__ __forceinline__ void small_grain_step(const DustArgs& A, const double4 (*Ps)[kSmallMax], int np,
                                                 double dt, double dt2, double dt6, const GrainCoef& kg,
                                                 double& x0, double& y0, double& z0, double& vx0, double& vy0,
                                                 double& vz0, bool& removed)
{
    double px = x0, py = y0, pz = z0, ux = vx0, uy = vy0, uz = vz0;
    double kx[4], ky[4], kz[4], rx[4], ry[4], rz[4];
    bool gone = false;
    for (int s = 0; s < 4 && !gone; ++s) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        gone = star_term<RSQ>(A, Ps[s][0], px, py, pz, ux, uy, uz, kg, ax, ay, az);
        bool hit = false;
        grain_fp64<RSQ>(Ps[s], 1, np, px, py, pz, 0.0, ax, ay, az, hit);
        gone = gone || hit;
        kx[s] = ax; ky[s] = ay; kz[s] = az;
        if (s == 0) {
            rx[1] = __builtin_fma(ax, dt2, vx0); ry[1] = __builtin_fma(ay, dt2, vy0);
            rz[1] = __builtin_fma(az, dt2, vz0);
            px = __builtin_fma(vx0, dt2, x0); py = __builtin_fma(vy0, dt2, y0);
            pz = __builtin_fma(vz0, dt2, z0);
            ux = rx[1]; uy = ry[1]; uz = rz[1];
        } else if (s == 1) {
            rx[2] = __builtin_fma(ax, dt2, vx0); ry[2] = __builtin_fma(ay, dt2, vy0);
            rz[2] = __builtin_fma(az, dt2, vz0);
            px = __builtin_fma(rx[1], dt2, x0); py = __builtin_fma(ry[1], dt2, y0);
            pz = __builtin_fma(rz[1], dt2, z0);
            ux = rx[2]; uy = ry[2]; uz = rz[2];
        } else if (s == 2) {
            rx[3] = __builtin_fma(ax, dt, vx0); ry[3] = __builtin_fma(ay, dt, vy0);
            rz[3] = __builtin_fma(az, dt, vz0);
            px = __builtin_fma(rx[2], dt, x0); py = __builtin_fma(ry[2], dt, y0);
            pz = __builtin_fma(rz[2], dt, z0);
            ux = rx[3]; uy = ry[3]; uz = rz[3];
        }
    }
    if (gone) {
        removed = true;
    } else {
        const double nx0 = __builtin_fma(dt6, __builtin_fma(2.0, rx[1] + rx[2], vx0) + rx[3], x0);
        const double ny0 = __builtin_fma(dt6, __builtin_fma(2.0, ry[1] + ry[2], vy0) + ry[3], y0);
        const double nz0 = __builtin_fma(dt6, __builtin_fma(2.0, rz[1] + rz[2], vz0) + rz[3], z0);
        vx0 = __builtin_fma(dt6, __builtin_fma(2.0, kx[1] + kx[2], kx[0]) + kx[3], vx0);
        vy0 = __builtin_fma(dt6, __builtin_fma(2.0, ky[1] + ky[2], ky[0]) + ky[3], vy0);
        vz0 = __builtin_fma(dt6, __builtin_fma(2.0, kz[1] + kz[2], kz[0]) + kz[3], vz0);
        x0 = nx0; y0 = ny0; z0 = nz0;
    }
}

// NP > 0 fixes the number of star + planets at compile time (1 to 9, the star
// and up to eight planets), so the pair loops unroll and their independent
// square-root and division chains overlap (the sums keep their order, so every
// NP gives the bits of NP = 0); NP = 0 takes it from A.
template <int NP, bool RSQ>
__global__ __launch_bounds__(kSmallWG) void k_small_steps(DustArgs A, int nsteps)
{
    __shared__ double4 P[2][4][kSmallMax];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const bool mwave = wave == 0;
    const int ng = A.n_grav, np = NP > 0 ? NP : A.n_planets;
    const double dt = A.dt, dt2 = dt / 2.0, dt6 = dt / 6.0;
    const double hs[3] = {0.5 * dt, 0.5 * dt, dt};

    // wave 0, lane i < ng: massive body i (base state, stage velocity and sums)
    const bool body = mwave && lane < ng;
    double mx = 0, my = 0, mz = 0, mvx = 0, mvy = 0, mvz = 0, mmu = 0;
    double vrx = 0, vry = 0, vrz = 0, Srx = 0, Sry = 0, Srz = 0, Sax = 0, Say = 0, Saz = 0;
    if (body) {
        mx = A.x[lane]; my = A.y[lane]; mz = A.z[lane];
        mvx = A.vx[lane]; mvy = A.vy[lane]; mvz = A.vz[lane]; mmu = A.mu[lane];
    }
    // massive stages 0-2 of a step (k_stage<0..2>) into P[buf]
    auto massive_stages = [&](int buf) {
        if (body) P[buf][0][lane] = make_double4(mx, my, mz, mmu);
        wave_lds_sync();
        for (int s = 0; s < 3; ++s) {
            if (body) {
                double ax, ay, az;
                small_massive_accel<RSQ>(P[buf][s], lane, np, ax, ay, az);
                const double h = hs[s];
                const double ux = s ? vrx : mvx, uy = s ? vry : mvy, uz = s ? vrz : mvz;
                const double rx = __builtin_fma(ax, h, mvx), ry = __builtin_fma(ay, h, mvy),
                             rz = __builtin_fma(az, h, mvz);
                P[buf][s + 1][lane] = make_double4(__builtin_fma(ux, h, mx), __builtin_fma(uy, h, my),
                                                   __builtin_fma(uz, h, mz), mmu);
                if (s == 0) {
                    Srx = __builtin_fma(2.0, rx, mvx); Sry = __builtin_fma(2.0, ry, mvy);
                    Srz = __builtin_fma(2.0, rz, mvz);
                    Sax = ax; Say = ay; Saz = az;
                } else if (s == 1) {
                    Srx = __builtin_fma(2.0, rx, Srx); Sry = __builtin_fma(2.0, ry, Sry);
                    Srz = __builtin_fma(2.0, rz, Srz);
                    Sax = __builtin_fma(2.0, ax, Sax); Say = __builtin_fma(2.0, ay, Say);
                    Saz = __builtin_fma(2.0, az, Saz);
                } else {
                    Srx = Srx + rx; Sry = Sry + ry; Srz = Srz + rz;
                    Sax = __builtin_fma(2.0, ax, Sax); Say = __builtin_fma(2.0, ay, Say);
                    Saz = __builtin_fma(2.0, az, Saz);
                }
                vrx = rx; vry = ry; vrz = rz;
            }
            wave_lds_sync();
        }
    };

    // wave 1, lane l: grain blockIdx.x * 64 + l
    const int g = blockIdx.x * 64 + lane;
    const int64_t t = (int64_t)ng + g;
    const bool mine = !mwave && g < A.n_dust;
    bool live = false;
    double x0 = 0, y0 = 0, z0 = 0, vx0 = 0, vy0 = 0, vz0 = 0, m = 1, bet = 0, betsw = 0, qe = 0;
    if (mine) {
        live = A.N[t] != 0.0;
        x0 = A.x[t]; y0 = A.y[t]; z0 = A.z[t];
        vx0 = A.vx[t]; vy0 = A.vy[t]; vz0 = A.vz[t];
        m = A.mu[t];
        const int32_t id = A.id[t];
        if (A.beta_p) {
            bet = A.beta_p[g];
            betsw = A.betasw_p ? A.betasw_p[g] : 0.0;
        } else if (id >= 0 && id < A.n_bins) {
            bet = (double)A.beta_t[id];
            betsw = (double)A.betasw_t[id];
        }
        qe = (id >= 0 && id < A.n_bins) ? (double)A.q_t[id] : 0.0;
    }
    const GrainCoef kg = grain_coef(A, bet, betsw, qe, m, A.mu[0]);
    bool removed = false;

    if (mwave) massive_stages(0);
    __syncthreads();
    for (int step = 0; step < nsteps; ++step) {
        const int buf = step & 1;
        if (mwave) {
            // massive stage 3 = final combination of this step (k_stage<3>) ...
            if (body) {
                double ax, ay, az;
                small_massive_accel<RSQ>(P[buf][3], lane, np, ax, ay, az);
                mx = __builtin_fma(dt6, Srx, mx); my = __builtin_fma(dt6, Sry, my);
                mz = __builtin_fma(dt6, Srz, mz);
                mvx = __builtin_fma(dt6, Sax + ax, mvx);
                mvy = __builtin_fma(dt6, Say + ay, mvy);
                mvz = __builtin_fma(dt6, Saz + az, mvz);
            }
            // ... and the stage positions of the next step into the other buffer
            if (step + 1 < nsteps) massive_stages(buf ^ 1);
        } else if (live && !removed) {
            // grains against this step's four stage positions (k_dust in the FP64 form)
            small_grain_step<RSQ>(A, P[buf], np, dt, dt2, dt6, kg, x0, y0, z0, vx0, vy0, vz0, removed);
        }
        __syncthreads();
    }
    if (mine && live) {
        if (removed) {
            A.N[t] = 0.0;
        } else {
            A.x[t] = x0; A.y[t] = y0; A.z[t] = z0;
            A.vx[t] = vx0; A.vy[t] = vy0; A.vz[t] = vz0;
        }
    }
    // Every workgroup read the massive state at its start, and a workgroup may
    // start after workgroup 0 has finished (more workgroups than fit on the GPU
    // at once), so the final massive state goes to mfin; k_small_commit, the
    // next launch on the stream, moves it into the state.
    if (blockIdx.x == 0 && body) {
        const int n = ng;
        A.mfin[lane] = mx; A.mfin[n + lane] = my; A.mfin[2 * n + lane] = mz;
        A.mfin[3 * n + lane] = mvx; A.mfin[4 * n + lane] = mvy; A.mfin[5 * n + lane] = mvz;
    }
}

// The split form, for many grains: the star and planets alone first, then the
// grains against their recorded stage positions, so that no grain workgroup
// carries a massive wave or a per-step barrier.  k_small_massive (one wave,
// one body per lane) runs the massive wave's code of k_small_steps for nsteps
// steps and writes the four stage positions of every step to stg, and the
// final state to mfin (k_small_commit moves it into the state);
// k_small_grains advances every grain through the same steps against them,
// one grain per lane, reading them with wave-uniform loads.  Same operations
// in the same order as k_small_steps, so the same bits.
template <int NP, bool RSQ>
__global__ __launch_bounds__(64) void k_small_massive(DustArgs A, int nsteps, double4* __restrict__ stg)
{
    __shared__ double4 P[4][kSmallMax];
    const int lane = threadIdx.x;
    const int ng = A.n_grav, np = NP > 0 ? NP : A.n_planets;
    const double dt = A.dt, dt6 = dt / 6.0;
    const double hs[3] = {0.5 * dt, 0.5 * dt, dt};
    const bool body = lane < ng;
    double mx = 0, my = 0, mz = 0, mvx = 0, mvy = 0, mvz = 0, mmu = 0;
    double vrx = 0, vry = 0, vrz = 0, Srx = 0, Sry = 0, Srz = 0, Sax = 0, Say = 0, Saz = 0;
    if (body) {
        mx = A.x[lane]; my = A.y[lane]; mz = A.z[lane];
        mvx = A.vx[lane]; mvy = A.vy[lane]; mvz = A.vz[lane]; mmu = A.mu[lane];
    }
    for (int step = 0; step < nsteps; ++step) {
        double4* out = stg + (int64_t)step * 4 * kSmallMax;
        // stages 0-2 (k_stage<0..2>)
        if (body) {
            P[0][lane] = make_double4(mx, my, mz, mmu);
            out[lane] = P[0][lane];
        }
        wave_lds_sync();
        for (int s = 0; s < 3; ++s) {
            if (body) {
                double ax, ay, az;
                small_massive_accel<RSQ>(P[s], lane, np, ax, ay, az);
                const double h = hs[s];
                const double ux = s ? vrx : mvx, uy = s ? vry : mvy, uz = s ? vrz : mvz;
                const double rx = __builtin_fma(ax, h, mvx), ry = __builtin_fma(ay, h, mvy),
                             rz = __builtin_fma(az, h, mvz);
                P[s + 1][lane] = make_double4(__builtin_fma(ux, h, mx), __builtin_fma(uy, h, my),
                                              __builtin_fma(uz, h, mz), mmu);
                out[(s + 1) * kSmallMax + lane] = P[s + 1][lane];
                if (s == 0) {
                    Srx = __builtin_fma(2.0, rx, mvx); Sry = __builtin_fma(2.0, ry, mvy);
                    Srz = __builtin_fma(2.0, rz, mvz);
                    Sax = ax; Say = ay; Saz = az;
                } else if (s == 1) {
                    Srx = __builtin_fma(2.0, rx, Srx); Sry = __builtin_fma(2.0, ry, Sry);
                    Srz = __builtin_fma(2.0, rz, Srz);
                    Sax = __builtin_fma(2.0, ax, Sax); Say = __builtin_fma(2.0, ay, Say);
                    Saz = __builtin_fma(2.0, az, Saz);
                } else {
                    Srx = Srx + rx; Sry = Sry + ry; Srz = Srz + rz;
                    Sax = __builtin_fma(2.0, ax, Sax); Say = __builtin_fma(2.0, ay, Say);
                    Saz = __builtin_fma(2.0, az, Saz);
                }
                vrx = rx; vry = ry; vrz = rz;
            }
            wave_lds_sync();
        }
        // stage 3 = final combination (k_stage<3>)
        if (body) {
            double ax, ay, az;
            small_massive_accel<RSQ>(P[3], lane, np, ax, ay, az);
            mx = __builtin_fma(dt6, Srx, mx); my = __builtin_fma(dt6, Sry, my);
            mz = __builtin_fma(dt6, Srz, mz);
            mvx = __builtin_fma(dt6, Sax + ax, mvx);
            mvy = __builtin_fma(dt6, Say + ay, mvy);
            mvz = __builtin_fma(dt6, Saz + az, mvz);
        }
        wave_lds_sync();
    }
    if (body) {
        const int n = ng;
        A.mfin[lane] = mx; A.mfin[n + lane] = my; A.mfin[2 * n + lane] = mz;
        A.mfin[3 * n + lane] = mvx; A.mfin[4 * n + lane] = mvy; A.mfin[5 * n + lane] = mvz;
    }
}

template <int NP, bool RSQ>
__global__ __launch_bounds__(kSmallGrainWG) void k_small_grains(DustArgs A, int nsteps,
                                                               const double4* __restrict__ stg)
{
    const int np = NP > 0 ? NP : A.n_planets;
    const double dt = A.dt, dt2 = dt / 2.0, dt6 = dt / 6.0;
    const int64_t g = (int64_t)blockIdx.x * kSmallGrainWG + threadIdx.x;
    if (g >= A.n_dust) return;
    const int64_t t = (int64_t)A.n_grav + g;
    if (A.N[t] == 0.0) return;
    double x0 = A.x[t], y0 = A.y[t], z0 = A.z[t], vx0 = A.vx[t], vy0 = A.vy[t], vz0 = A.vz[t];
    const double m = A.mu[t];
    double bet = 0, betsw = 0;
    const int32_t id = A.id[t];
    if (A.beta_p) {
        bet = A.beta_p[g];
        betsw = A.betasw_p ? A.betasw_p[g] : 0.0;
    } else if (id >= 0 && id < A.n_bins) {
        bet = (double)A.beta_t[id];
        betsw = (double)A.betasw_t[id];
    }
    const double qe = (id >= 0 && id < A.n_bins) ? (double)A.q_t[id] : 0.0;
    const GrainCoef kg = grain_coef(A, bet, betsw, qe, m, A.mu[0]);
    bool removed = false;
    for (int step = 0; step < nsteps && !removed; ++step) {
        const double4(*Ps)[kSmallMax] = reinterpret_cast<const double4(*)[kSmallMax]>(stg + (int64_t)step * 4 * kSmallMax);
        small_grain_step<RSQ>(A, Ps, np, dt, dt2, dt6, kg, x0, y0, z0, vx0, vy0, vz0, removed);
    }
    if (removed) {
        A.N[t] = 0.0;
    } else {
        A.x[t] = x0; A.y[t] = y0; A.z[t] = z0;
        A.vx[t] = vx0; A.vy[t] = vy0; A.vz[t] = vz0;
    }
}

// One launch of k_small_steps, specialised on the number of star + planets.
template <bool RSQ>
void launch_small(const DustArgs& A, unsigned blocks, hipStream_t st, int n)
{
    switch (A.n_planets) {
    case 1: k_small_steps<1, RSQ><<<blocks, kSmallWG, 0, st>>>(A, n); break;
    case 2: k_small_steps<2, RSQ><<<blocks, kSmallWG, 0, st>>>(A, n); break;
    case 3: k_small_steps<3, RSQ><<<blocks, kSmallWG, 0, st>>>(A, n); break;
    case 4: k_small_steps<4, RSQ><<<blocks, kSmallWG, 0, st>>>(A, n); break;
    case 5: k_small_steps<5, RSQ><<<blocks, kSmallWG, 0, st>>>(A, n); break;
    case 6: k_small_steps<6, RSQ><<<blocks, kSmallWG, 0, st>>>(A, n); break;
    case 7: k_small_steps<7, RSQ><<<blocks, kSmallWG, 0, st>>>(A, n); break;
    case 8: k_small_steps<8, RSQ><<<blocks, kSmallWG, 0, st>>>(A, n); break;
    case 9: k_small_steps<9, RSQ><<<blocks, kSmallWG, 0, st>>>(A, n); break;
    default: k_small_steps<0, RSQ><<<blocks, kSmallWG, 0, st>>>(A, n); break;
    }
}

// The split form: k_small_massive, then k_small_grains, specialised on the
// number of star + planets.
template <bool RSQ>
void launch_small_split(const DustArgs& A, double4* stg, int64_t n_dust, hipStream_t st, int n)
{
    const unsigned gb = (unsigned)((n_dust + kSmallGrainWG - 1) / kSmallGrainWG);
    switch (A.n_planets) {
#define DDYN_SPLIT_CASE(K)                                                       \
    case K:                                                                      \
        k_small_massive<K, RSQ><<<1, 64, 0, st>>>(A, n, stg);                    \
        k_small_grains<K, RSQ><<<gb, kSmallGrainWG, 0, st>>>(A, n, stg);          \
        break;
    DDYN_SPLIT_CASE(1) DDYN_SPLIT_CASE(2) DDYN_SPLIT_CASE(3) DDYN_SPLIT_CASE(4) DDYN_SPLIT_CASE(5)
    DDYN_SPLIT_CASE(6) DDYN_SPLIT_CASE(7) DDYN_SPLIT_CASE(8) DDYN_SPLIT_CASE(9)
#undef DDYN_SPLIT_CASE
    default:
        k_small_massive<0, RSQ><<<1, 64, 0, st>>>(A, n, stg);
        k_small_grains<0, RSQ><<<gb, kSmallGrainWG, 0, st>>>(A, n, stg);
        break;
    }
}

// Moves the massive state k_small_steps left in mfin into the state arrays and
// the stage-0 positions (one lane per body, n_grav <= kSmallMax).
__global__ __launch_bounds__(64) void k_small_commit(DustArgs A)
{
    const int i = threadIdx.x;
    const int n = A.n_grav;
    if (i >= n) return;
    const double x = A.mfin[i], y = A.mfin[n + i], z = A.mfin[2 * n + i];
    A.x[i] = x; A.y[i] = y; A.z[i] = z;
    A.vx[i] = A.mfin[3 * n + i]; A.vy[i] = A.mfin[4 * n + i]; A.vz[i] = A.mfin[5 * n + i];
    store_pos(const_cast<double4*>(A.p64[0]), const_cast<float4*>(A.p32[0]), i, x, y, z, A.mu[i]);
}

// ---------------------------------------------------------------------------
// Discrete adjoint of the RK4 step
// ---------------------------------------------------------------------------
// Stage map of one step: R1 = r, V1 = v; R2 = r + h/2 V1, V2 = v + h/2 A1;
// R3 = r + h/2 V2, V3 = v + h/2 A2; R4 = r + h V3, V4 = v + h A3;
// r' = r + h/6 (V1 + 2V2 + 2V3 + V4), v' = v + h/6 (A1 + 2A2 + 2A3 + A4).
// Reverse: Vbar_s = w_s rbar', Abar_s = w_s vbar' (w = h/6, h/3, h/3, h/6), then
// for s = 4..1: Rbar_s = J_R^T Abar_s, Vbar_s += J_V^T Abar_s, rbar += Rbar_s,
// vbar += Vbar_s, Vbar_{s-1} += h'_s Rbar_s, Abar_{s-1} += h'_s Vbar_s.
// With K(d) = d/s^3, s^2 = |d|^2 + eps^2, J(d) = dK/dd = (I s^2 - 3 d d^T)/s^5:
//   massive: Rbar_k = sum_j J(R_j - R_k)(mu_k Abar_j - mu_j Abar_k),
//            mubar_k = sum_i Abar_i . K(R_k - R_i);
//   grains:  star term differentiated in closed form (gravity, radiation
//            pressure, PR and wind drag), pulls mu_j K(R_j - R_t).


// One FP64 pair of the massive VJP: target k (position and mass pk, Abar_k),
// source j (pj, Abar_j), softening e2 of this pair.  Adds
// J(d) (mu_k Abar_j - mu_j Abar_k) to g and -Abar_j . K(d) to gm, with
// d = r_j - r_k, s^2 = |d|^2 + e2, K(d) = d / s^3 and J(d) w = w / s^3 - 3 d (d.w) / s^5.
// w is formed from plain products, so a softened self pair (j == k, d = 0)
// has w = 0 exactly and contributes nothing; s^2 = 0 (an unsoftened self or
// coincident pair) contributes nothing either.
__device__ inline void vjp_pair_fp64(const double4 pk, double akx, double aky, double akz, const double4 pj,
                                     double ajx, double ajy, double ajz, double e2, double& gx, double& gy,
                                     double& gz, double& gm)
{
    const double dx = pj.x - pk.x, dy = pj.y - pk.y, dz = pj.z - pk.z;
    const double s2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, e2)));
    if (s2 > 0.0) {
        const double inv = rsqrt_fp64(s2);
        const double inv2 = inv * inv, inv3 = inv2 * inv, inv5 = inv3 * inv2;
        const double wx = pk.w * ajx - pj.w * akx, wy = pk.w * ajy - pj.w * aky, wz = pk.w * ajz - pj.w * akz;
        const double dw = __builtin_fma(dz, wz, __builtin_fma(dy, wy, dx * wx));
        const double t3 = 3.0 * dw * inv5;
        gx = __builtin_fma(wx, inv3, __builtin_fma(-dx, t3, gx));
        gy = __builtin_fma(wy, inv3, __builtin_fma(-dy, t3, gy));
        gz = __builtin_fma(wz, inv3, __builtin_fma(-dz, t3, gz));
        const double da = __builtin_fma(ajz, dz, __builtin_fma(ajy, dy, ajx * dx));
        gm = __builtin_fma(-da, inv3, gm);
    }
}

// Massive VJP, all pairs: targets k (one per lane), sources j in slice y of
// [0, n_grav).  Writes part[(y*4 + c) * n + k], c = x, y, z, mu.  The LDS tile
// carries each source's softening flag (1 for swarm sources, 0 for the star
// and planets) so that a pair's softening is e2k * flag.
__global__ __launch_bounds__(kWG) void k_vjp_pairs(const double4* __restrict__ pos,
                                                   const double* __restrict__ Ab, int n_grav,
                                                   int n_planets, int slice_len, double eps2,
                                                   double* __restrict__ part)
{
    __shared__ double4 tp[kTile];
    __shared__ double4 ta[kTile];
    const size_t n = (size_t)n_grav;
    const int k = blockIdx.x * kWG + threadIdx.x;
    const bool ok = k < n_grav;
    const double4 pk = ok ? pos[k] : make_double4(kFar, kFar, kFar, 0.0);
    const double akx = ok ? Ab[k] : 0.0, aky = ok ? Ab[n + k] : 0.0, akz = ok ? Ab[2 * n + k] : 0.0;
    const double e2k = (k >= n_planets) ? eps2 : 0.0;
    const int jbeg = blockIdx.y * slice_len;
    const int jend = min(jbeg + slice_len, n_grav);
    double gx = 0.0, gy = 0.0, gz = 0.0, gm = 0.0;
    for (int j0 = jbeg; j0 < jend; j0 += kTile) {
        const int tl = min(kTile, jend - j0);
        __syncthreads();
        if ((int)threadIdx.x < tl) {
            const int j = j0 + threadIdx.x;
            tp[threadIdx.x] = pos[j];
            ta[threadIdx.x] = make_double4(Ab[j], Ab[n + j], Ab[2 * n + j], (j >= n_planets) ? 1.0 : 0.0);
        }
        __syncthreads();
        for (int jj = 0; jj < tl; ++jj) {
            const double4 aj = ta[jj];
            vjp_pair_fp64(pk, akx, aky, akz, tp[jj], aj.x, aj.y, aj.z, e2k * aj.w, gx, gy, gz, gm);
        }
    }
    if (ok) {
        double* p = part + (size_t)blockIdx.y * 4 * n;
        p[k] = gx; p[n + k] = gy; p[2 * n + k] = gz; p[3 * n + k] = gm;
    }
}

// FP64 VJP of one source tile [j0, j1) for target k (self tiles and close pairs
// of the FP32 form).  Sources come from global memory.
__device__ inline void vjp_tile_fp64(const double4* __restrict__ pos, const double* __restrict__ Ab, size_t n,
                                     int j0, int j1, int n_planets, double e2k, const double4 pk, double akx,
                                     double aky, double akz, double& gx, double& gy, double& gz, double& gm)
{
    for (int j = j0; j < j1; ++j)
        vjp_pair_fp64(pk, akx, aky, akz, pos[j], Ab[j], Ab[n + j], Ab[2 * n + j], (j >= n_planets) ? e2k : 0.0,
                      gx, gy, gz, gm);
}

// FP32 VJP sum over one LDS tile for one target; with SELF the source slot
// self0 is the target itself and is masked (s^2 = 3e38: every term underflows
// to zero and the close-pair minimum ignores it).
template <bool SELF>
__device__ inline void vjp_tile32(const float4* __restrict__ tp, const float4* __restrict__ ta, int self0,
                                  float xf, float yf, float zf, float mk, float axf, float ayf, float azf,
                                  float e2f, float& fx, float& fy, float& fz, float& fm, float& rmin)
{
#pragma unroll 2
    for (int jj = 0; jj < kTile; ++jj) {
        const float4 pj = tp[jj];
        const float4 aj = ta[jj];
        const float dx = pj.x - xf, dy = pj.y - yf, dz = pj.z - zf;
        float s2 = __builtin_fmaf(dx, dx, e2f);
        s2 = __builtin_fmaf(dy, dy, s2);
        s2 = __builtin_fmaf(dz, dz, s2);
        if (SELF) s2 = (jj == self0) ? 3.0e38f : s2;
        rmin = fminf(rmin, s2);
        const float inv = __builtin_amdgcn_rsqf(s2);
        const float inv2 = inv * inv;
        const float inv3 = inv2 * inv;
        const float inv5 = inv3 * inv2;
        const float wx = __builtin_fmaf(mk, aj.x, -pj.w * axf);
        const float wy = __builtin_fmaf(mk, aj.y, -pj.w * ayf);
        const float wz = __builtin_fmaf(mk, aj.z, -pj.w * azf);
        const float dw = __builtin_fmaf(dz, wz, __builtin_fmaf(dy, wy, dx * wx));
        const float t3 = 3.0f * dw * inv5;
        fx = __builtin_fmaf(wx, inv3, __builtin_fmaf(-dx, t3, fx));
        fy = __builtin_fmaf(wy, inv3, __builtin_fmaf(-dy, t3, fy));
        fz = __builtin_fmaf(wz, inv3, __builtin_fmaf(-dz, t3, fz));
        const float da = __builtin_fmaf(aj.z, dz, __builtin_fmaf(aj.y, dy, aj.x * dx));
        fm = __builtin_fmaf(-da, inv3, fm);
    }
}

// FP32-tile form of k_vjp_pairs: per 256-source tile an FP32 partial flushed
// to FP64; the tile of star/planet sources and tiles where a pair of the wave
// came closer than the FP64 radius are redone in FP64 (as k_pairs32).
__global__ __launch_bounds__(kWG) void k_vjp_pairs32(const double4* __restrict__ pos,
                                                     const double* __restrict__ Ab, int n_grav,
                                                     int n_planets, int slice_len, double eps2,
                                                     float r64sq, double* __restrict__ part)
{
    __shared__ float4 tp[kTile];
    __shared__ float4 ta[kTile];
    const size_t n = (size_t)n_grav;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wbase = blockIdx.x * kWG + wave * 64;
    const int k = wbase + lane;
    const bool ok = k < n_grav;
    const double4 pk = ok ? pos[k] : make_double4(kFar, kFar, kFar, 0.0);
    const double akx = ok ? Ab[k] : 0.0, aky = ok ? Ab[n + k] : 0.0, akz = ok ? Ab[2 * n + k] : 0.0;
    const bool sw = k >= n_planets;
    const double e2k = sw ? eps2 : 0.0;
    const float xf = (float)pk.x, yf = (float)pk.y, zf = (float)pk.z, mk = (float)pk.w;
    const float axf = (float)akx, ayf = (float)aky, azf = (float)akz;
    const float e2f = sw ? fmaxf((float)eps2, kTiny2) : kTiny2;
    const float thr = r64sq + e2f;
    const int jbeg = blockIdx.y * slice_len;
    const int jend = min(jbeg + slice_len, n_grav);
    double gx = 0.0, gy = 0.0, gz = 0.0, gm = 0.0;
    for (int j0 = jbeg; j0 < jend; j0 += kTile) {
        const int tl = min(kTile, jend - j0);
        __syncthreads();
        if ((int)threadIdx.x < tl) {
            const int j = j0 + threadIdx.x;
            const double4 pj = pos[j];
            tp[threadIdx.x] = make_float4((float)pj.x, (float)pj.y, (float)pj.z, (float)pj.w);
            ta[threadIdx.x] = make_float4((float)Ab[j], (float)Ab[n + j], (float)Ab[2 * n + j], 0.f);
        } else {
            tp[threadIdx.x] = make_float4((float)kFar, (float)kFar, (float)kFar, 0.f);
            ta[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();
        // star/planet sources (unsoftened for swarm targets too): FP64
        bool redo = j0 < n_planets;
        if (!redo) {
            float fx = 0.f, fy = 0.f, fz = 0.f, fm = 0.f, rmin = 3.0e38f;
            const bool own = (j0 < wbase + 64) && (j0 + tl > wbase);   // wave-uniform
            if (own)
                vjp_tile32<true>(tp, ta, k - j0, xf, yf, zf, mk, axf, ayf, azf, e2f, fx, fy, fz, fm, rmin);
            else
                vjp_tile32<false>(tp, ta, 0, xf, yf, zf, mk, axf, ayf, azf, e2f, fx, fy, fz, fm, rmin);
            redo = __any(ok && rmin < thr);
            if (!redo) { gx += (double)fx; gy += (double)fy; gz += (double)fz; gm += (double)fm; }
        }
        if (redo && ok) vjp_tile_fp64(pos, Ab, n, j0, j0 + tl, n_planets, e2k, pk, akx, aky, akz, gx, gy, gz, gm);
    }
    if (ok) {
        double* p = part + (size_t)blockIdx.y * 4 * n;
        p[k] = gx; p[n + k] = gy; p[2 * n + k] = gz; p[3 * n + k] = gm;
    }
}

struct AdjArgs {
    int n_grav, n_planets, n_dust, nslice, nwaves;
    int n_wsum;          // massive bodies whose grain pulls k_adj_dust sums per wave: n_grav, or 1 (the star)
    size_t n_all;
    const double4* p64[4];
    const double* x; const double* y; const double* z;
    const double* vx; const double* vy; const double* vz;
    const double* mu;
    const double* N;
    const int32_t* id;
    const double *beta_p, *betasw_p;
    const float *beta_t, *betasw_t;
    int n_bins, SWQ;
    double dt, eps2, Rmin2, vsw, light;
    double* rb;          // 3 * n_all, SoA blocks: dL/dr
    double* vb;          // 3 * n_all: dL/dv
    double* Vbm;         // 4 stages * 3 * n_grav
    double* Abm;         // 4 stages * 3 * n_grav
    double* Rbg;         // 4 stages * 3 * n_grav: grain pulls on massive stage positions
    double* mug;         // n_grav: grain pulls on massive masses (this step)
    double* wpart;       // per-wave partials: nwaves * 4 stages * n_wsum * 4
    double* gR;          // k_adj_pull path (else null): grain stage positions, 4 stages * 3 * n_dust (kFar if inactive)
    double* gA;          // k_adj_pull path: grain acceleration adjoints, 4 stages * 3 * n_dust (0 if inactive)
    double* ppart;       // k_adj_pull path: grain-pull slice partials, nslice_g * 4 stages * 4 * n_grav
    const double* part;  // VJP pair partials
    double* g_mu;        // n_grav accumulated
    double* g_beta;      // n_dust accumulated
};

__global__ void k_adj_init(AdjArgs A)
{
    const int i = blockIdx.x * kWG + threadIdx.x;
    if (i >= A.n_grav) return;
    const double h = A.dt;
    const double w[4] = {h / 6.0, h / 3.0, h / 3.0, h / 6.0};
    const size_t n = (size_t)A.n_grav, na = A.n_all;
    for (int st = 0; st < 4; ++st)
        for (int c = 0; c < 3; ++c) {
            A.Vbm[(st * 3 + c) * n + i] = w[st] * A.rb[c * na + i];
            A.Abm[(st * 3 + c) * n + i] = w[st] * A.vb[c * na + i];
        }
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Grain reverse: one grain per lane; recomputes its four FP64 stages (as k_dust
// does in the FP64 form), then runs the stage reverse in registers.  The pull
// of the grain on the first n_wsum massive bodies is summed per wave and
// written to wpart.  With n_wsum = 1 (many massive bodies) only the star's
// pull is, since it carries the drag terms; the grain's stage positions and
// acceleration adjoints go to gR, gA instead, and k_adj_pull sums the pulls
// on the other massive bodies as a pair kernel over LDS tiles of grains.
__global__ __launch_bounds__(kWG) void k_adj_dust(AdjArgs A)
{
    const int g = blockIdx.x * kWG + threadIdx.x;
    const int wave = (blockIdx.x * kWG + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)A.n_grav + g;
    const size_t na = A.n_all;
    bool live = g < A.n_dust && A.N[t] != 0.0;
    double x0[3] = {0, 0, 0}, v0[3] = {0, 0, 0}, bet = 0.0, betsw = 0.0, m = 1.0;
    if (g < A.n_dust) {
        x0[0] = A.x[t]; x0[1] = A.y[t]; x0[2] = A.z[t];
        v0[0] = A.vx[t]; v0[1] = A.vy[t]; v0[2] = A.vz[t];
        m = A.mu[t];
        const int32_t id = A.id[t];
        if (A.beta_p) {
            bet = A.beta_p[g];
            betsw = A.betasw_p ? A.betasw_p[g] : 0.0;
        } else if (id >= 0 && id < A.n_bins) {
            bet = (double)A.beta_t[id];
            betsw = (double)A.betasw_t[id];
        }
    }
    const double dt = A.dt, dt2 = dt / 2.0;
    const double hh[3] = {dt2, dt2, dt};
    double R[4][3], V[4][3];
    bool gone = false;
    DustArgs D{};
    D.Rmin2 = A.Rmin2; D.SWQ = A.SWQ; D.FLQ = 0; D.vsw = A.vsw; D.light = A.light;
    for (int c = 0; c < 3; ++c) { R[0][c] = x0[c]; V[0][c] = v0[c]; }
    // forward stages (src/gpu_functions.cu:679-869 in the FP64 form of k_dust)
    for (int st = 0; st < 4; ++st) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        if (live && !gone) {
            const double4* P = A.p64[st];
            gone = star_term<false>(D, P[0], R[st][0], R[st][1], R[st][2], V[st][0], V[st][1], V[st][2],
                             grain_coef(D, bet, betsw, 0.0, m, P[0].w), ax, ay, az);
            bool hit = false;
            grain_fp64<false>(P, 1, A.n_planets, R[st][0], R[st][1], R[st][2], 0.0, ax, ay, az, hit);
            grain_fp64<false>(P, A.n_planets, A.n_grav, R[st][0], R[st][1], R[st][2], A.eps2, ax, ay, az, hit);
            gone = gone || hit;
        }
        if (st == 3) break;
        const double a[3] = {ax, ay, az};
        for (int c = 0; c < 3; ++c) {
            R[st + 1][c] = __builtin_fma(V[st][c], hh[st], x0[c]);
            V[st + 1][c] = __builtin_fma(a[c], hh[st], v0[c]);
        }
    }
    const bool act = live && !gone;   // removed grains: identity map, nothing to reverse
    double rbar[3] = {0, 0, 0}, vbar[3] = {0, 0, 0};
    if (act)
        for (int c = 0; c < 3; ++c) { rbar[c] = A.rb[c * na + t]; vbar[c] = A.vb[c * na + t]; }
    const double w4[4] = {dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0};
    double Vb[4][3], Ab[4][3];
    for (int st = 0; st < 4; ++st)
        for (int c = 0; c < 3; ++c) { Vb[st][c] = w4[st] * rbar[c]; Ab[st][c] = w4[st] * vbar[c]; }
    double r_out[3] = {rbar[0], rbar[1], rbar[2]}, v_out[3] = {vbar[0], vbar[1], vbar[2]};
    double betab = 0.0;
    const double kap = (A.SWQ == 1) ? bet / A.light + betsw / A.vsw : bet / A.light;
    const double dkap = 1.0 / A.light;
    const size_t n = (size_t)A.n_grav;
    for (int st = 3; st >= 0; --st) {
        const double4* P = A.p64[st];
        const double* Rs = R[st];
        const double* Vs = V[st];
        const double* Av = Ab[st];
        double Rb[3] = {0, 0, 0};
        const size_t nw = (size_t)A.n_wsum;
        double* wp = A.wpart + ((size_t)wave * 4 + st) * nw * 4;
        if (A.gR && g < A.n_dust) {
            const size_t nd = (size_t)A.n_dust;
            for (int c = 0; c < 3; ++c) {
                A.gR[(st * 3 + c) * nd + g] = act ? Rs[c] : kFar;
                A.gA[(st * 3 + c) * nd + g] = act ? Av[c] : 0.0;
            }
        }
        {   // star term
            const double4 s0 = P[0];
            const double m0 = s0.w;
            const double d[3] = {Rs[0] - s0.x, Rs[1] - s0.y, Rs[2] - s0.z};
            const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            // powers of 1/R from one refined reciprocal square root: the
            // Jacobian below needs 1/R^2 .. 1/R^6 about thirty times per stage
            const double i1 = rsqrt_fp64(R2), i2 = i1 * i1, i3 = i2 * i1, i4 = i2 * i2, i5 = i4 * i1,
                         i6 = i4 * i2;
            const double Gc = (1.0 - bet - betsw) * m0;
            const double dv = d[0] * Vs[0] + d[1] * Vs[1] + d[2] * Vs[2];
            const double dA = d[0] * Av[0] + d[1] * Av[1] + d[2] * Av[2];
            const double vA = Vs[0] * Av[0] + Vs[1] * Av[1] + Vs[2] * Av[2];
            double db[3], t2A = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double J0A = Av[c] * i3 - 3.0 * d[c] * dA * i5;
                db[c] = -Gc * J0A - m0 * kap * ((Vs[c] * dA + dv * Av[c]) * i4 - 4.0 * dv * dA * d[c] * i6 -
                                                 2.0 * d[c] * vA * i4);
                Vb[st][c] += act ? -m0 * kap * (d[c] * dA * i4 + Av[c] * i2) : 0.0;
                t2A += (dv * d[c] * i4 + Vs[c] * i2) * Av[c];
            }
            const double bb = m0 * dA * i3 - m0 * dkap * t2A;
            const double mb = -(1.0 - bet - betsw) * dA * i3 - kap * t2A;
            double c4[4];
            for (int c = 0; c < 3; ++c) {
                if (act) Rb[c] += db[c];
                c4[c] = act ? -db[c] : 0.0;
            }
            c4[3] = act ? mb : 0.0;
            if (act) betab += bb;
            for (int c = 0; c < 4; ++c) {
                const double s = wave_sum(c4[c]);
                if (lane == 0) wp[c] = s;
            }
        }
        for (int j = 1; j < A.n_grav; ++j) {   // massive pulls
            const double4 pj = P[j];
            const double q[3] = {pj.x - Rs[0], pj.y - Rs[1], pj.z - Rs[2]};
            const double s2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + ((j >= A.n_planets) ? A.eps2 : 0.0);
            const double inv = rsqrt_fp64(s2), inv3 = inv * inv * inv, inv5 = inv3 * inv * inv;
            const double qA = q[0] * Av[0] + q[1] * Av[1] + q[2] * Av[2];
            double c4[4];
            for (int c = 0; c < 3; ++c) {
                const double JA = Av[c] * inv3 - 3.0 * q[c] * qA * inv5;
                c4[c] = act ? pj.w * JA : 0.0;
                if (act) Rb[c] -= pj.w * JA;
            }
            c4[3] = act ? qA * inv3 : 0.0;
            if (nw > 1)
                for (int c = 0; c < 4; ++c) {
                    const double s = wave_sum(c4[c]);
                    if (lane == 0) wp[(size_t)j * 4 + c] = s;
                }
        }
        for (int c = 0; c < 3; ++c) {
            r_out[c] += Rb[c];
            v_out[c] += Vb[st][c];
            if (st > 0) {
                Vb[st - 1][c] += hh[st - 1] * Rb[c];
                Ab[st - 1][c] += hh[st - 1] * Vb[st][c];
            }
        }
    }
    if (act) {
        for (int c = 0; c < 3; ++c) { A.rb[c * na + t] = r_out[c]; A.vb[c * na + t] = v_out[c]; }
        A.g_beta[g] += betab;
    }
}

// Sums of the per-wave grain partials: Rbg[st][c][j], mug[j].
// One wave per (stage, massive body): lane l sums the per-wave pulls of grain
// waves l, l + 64, ... in order, and a fixed shuffle tree adds the 64 lane
// sums, so the result is deterministic.  (The number of items, 4 n_grav, is
// small next to the number of grain waves, so a thread per item would wait
// through one load latency per grain wave.)
__global__ __launch_bounds__(kWG) void k_adj_reduce(AdjArgs A)
{
    const int item = blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);   // (st, j)
    const int lane = threadIdx.x & 63;
    const int n = A.n_grav, nw = A.n_wsum;
    if (item >= 4 * nw) return;   // whole waves
    const int st = item / nw, j = item % nw;
    double s[4] = {0, 0, 0, 0};
    for (int w = lane; w < A.nwaves; w += 64) {
        const double* p = A.wpart + (((size_t)w * 4 + st) * nw + j) * 4;
        for (int c = 0; c < 4; ++c) s[c] += p[c];
    }
    for (int c = 0; c < 4; ++c) s[c] = wave_sum(s[c]);
    if (lane == 0)
        for (int c = 0; c < 3; ++c) A.Rbg[((size_t)st * 3 + c) * n + j] = s[c];
    if (st == 0) {
        double mt = s[3];
        for (int s2 = 1; s2 < 4; ++s2) {
            double acc = 0.0;
            for (int w = lane; w < A.nwaves; w += 64) acc += A.wpart[(((size_t)w * 4 + s2) * nw + j) * 4 + 3];
            mt += wave_sum(acc);
        }
        if (lane == 0) A.mug[j] = mt;
    }
}

// Grain pulls on massive bodies j >= 1 (reverse sweep, many massive bodies):
// targets j, one per lane; sources the grains of slice blockIdx.y at stage
// blockIdx.z, their stage positions and acceleration adjoints staged through
// LDS in tiles of kTile.  Per pair (q = r_j - r_g, s2 = |q|^2 + eps^2 for
// swarm bodies, A the grain's acceleration adjoint) it adds
// J(q) A = A / s^3 - 3 q (q.A) / s^5, times mu_j at the end, and (q.A) / s^3
// for mu_j, the same terms that k_adj_dust sums per wave.  Writes
// ppart[((y * 4 + st) * 4 + c) * n_grav + j].  Inactive grains sit at kFar
// with A = 0 and add exactly zero.
__global__ __launch_bounds__(kWG) void k_adj_pull(AdjArgs A, int slice_len)
{
    __shared__ double tr[3][kTile];
    __shared__ double ta[3][kTile];
    const int st = blockIdx.z;
    const int j = 1 + blockIdx.x * kWG + threadIdx.x;
    const bool ok = j < A.n_grav;
    const double4 pj = ok ? A.p64[st][j] : make_double4(-kFar, -kFar, -kFar, 0.0);
    const double e2 = (j >= A.n_planets) ? A.eps2 : 0.0;
    const size_t nd = (size_t)A.n_dust;
    const int gb = blockIdx.y * slice_len;
    const int ge = min(gb + slice_len, A.n_dust);
    const double* R = A.gR + (size_t)st * 3 * nd;
    const double* Av = A.gA + (size_t)st * 3 * nd;
    double sx = 0.0, sy = 0.0, sz = 0.0, sm = 0.0;
    for (int g0 = gb; g0 < ge; g0 += kTile) {
        const int tl = min(kTile, ge - g0);
        __syncthreads();
        if ((int)threadIdx.x < tl) {
            const size_t g = (size_t)g0 + threadIdx.x;
            for (int c = 0; c < 3; ++c) {
                tr[c][threadIdx.x] = R[c * nd + g];
                ta[c][threadIdx.x] = Av[c * nd + g];
            }
        }
        __syncthreads();
        for (int gg = 0; gg < tl; ++gg) {
            const double qx = pj.x - tr[0][gg], qy = pj.y - tr[1][gg], qz = pj.z - tr[2][gg];
            const double ax = ta[0][gg], ay = ta[1][gg], az = ta[2][gg];
            const double s2 = __builtin_fma(qz, qz, __builtin_fma(qy, qy, qx * qx)) + e2;
            const double inv = rsqrt_fp64(s2), inv3 = inv * inv * inv, inv5 = inv3 * inv * inv;
            const double qA = __builtin_fma(qz, az, __builtin_fma(qy, ay, qx * ax));
            const double t3 = 3.0 * qA * inv5;
            sx = __builtin_fma(ax, inv3, __builtin_fma(-qx, t3, sx));
            sy = __builtin_fma(ay, inv3, __builtin_fma(-qy, t3, sy));
            sz = __builtin_fma(az, inv3, __builtin_fma(-qz, t3, sz));
            sm = __builtin_fma(qA, inv3, sm);
        }
    }
    if (ok) {
        const size_t n = (size_t)A.n_grav;
        double* p = A.ppart + ((size_t)blockIdx.y * 4 + st) * 4 * n;
        p[j] = pj.w * sx; p[n + j] = pj.w * sy; p[2 * n + j] = pj.w * sz; p[3 * n + j] = sm;
    }
}

// Sums of k_adj_pull's slice partials in slice order: Rbg[st][c][j] and
// mug[j] (stages in order) for j >= 1.  One wave per massive body: lane l
// sums slices l, l + 64, ..., and a fixed shuffle tree adds the lane sums.
__global__ __launch_bounds__(kWG) void k_adj_pull_reduce(AdjArgs A, int nslice_g)
{
    const int j = 1 + blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= A.n_grav) return;   // whole waves
    const size_t n = (size_t)A.n_grav;
    double mt = 0.0;
    for (int st = 0; st < 4; ++st) {
        double s[4] = {0, 0, 0, 0};
        for (int y = lane; y < nslice_g; y += 64) {
            const double* p = A.ppart + ((size_t)y * 4 + st) * 4 * n;
            for (int c = 0; c < 4; ++c) s[c] += p[c * n + j];
        }
        for (int c = 0; c < 4; ++c) s[c] = wave_sum(s[c]);
        if (lane == 0)
            for (int c = 0; c < 3; ++c) A.Rbg[((size_t)st * 3 + c) * n + j] = s[c];
        mt += s[3];
    }
    if (lane == 0) A.mug[j] = mt;
}

// Massive stage reverse (stage ST): combine VJP pair partials with the grain
// pulls, update rbar, vbar and the previous stage's Vbar and Abar.
template <int ST>
__global__ void k_adj_mstage(AdjArgs A, int with_grains)
{
    const int i = blockIdx.x * kWG + threadIdx.x;
    if (i >= A.n_grav) return;
    const size_t n = (size_t)A.n_grav, na = A.n_all;
    const double h = A.dt;
    const double hh[3] = {h / 2.0, h / 2.0, h};
    double Rb[3] = {0, 0, 0}, mb = 0.0;
#pragma unroll 8
    for (int y = 0; y < A.nslice; ++y) {
        const double* p = A.part + (size_t)y * 4 * n;
        Rb[0] += p[i]; Rb[1] += p[n + i]; Rb[2] += p[2 * n + i]; mb += p[3 * n + i];
    }
    if (with_grains) {
        for (int c = 0; c < 3; ++c) Rb[c] += A.Rbg[((size_t)ST * 3 + c) * n + i];
        if (ST == 3) mb += A.mug[i];
    }
    A.g_mu[i] += mb;
    for (int c = 0; c < 3; ++c) {
        const double Vb = A.Vbm[(ST * 3 + c) * n + i];
        A.rb[c * na + i] += Rb[c];
        A.vb[c * na + i] += Vb;
        if (ST > 0) {
            A.Vbm[((ST - 1) * 3 + c) * n + i] += hh[ST - 1] * Rb[c];
            A.Abm[((ST - 1) * 3 + c) * n + i] += hh[ST - 1] * Vb;
        }
    }
}

// The massive part of one reverse step when n_grav <= kWG, in one workgroup
// with one body per thread: k_adj_init, then for stages 3..0 the VJP pair sums
// and the stage chain rule of k_adj_mstage.  Each stage's Abar of every body
// goes through LDS to the next stage's pair sums.  The operations and their
// order are those of the separate launches with one source slice (where the
// FP32-force form redoes its only tile, which holds the star, in FP64), so the
// result is the same bit for bit; nine launches per reverse step become one.
__global__ __launch_bounds__(kWG) void k_adj_massive_small(AdjArgs A, int with_grains)
{
    __shared__ double4 sp[kWG];   // stage positions and masses
    __shared__ double4 sa[kWG];   // stage Abar and the softening flag
    const int i = threadIdx.x;
    const int ng = A.n_grav;
    const bool ok = i < ng;
    const size_t n = (size_t)ng, na = A.n_all;
    const double h = A.dt;
    const double w[4] = {h / 6.0, h / 3.0, h / 3.0, h / 6.0};
    const double hh[3] = {h / 2.0, h / 2.0, h};
    const double e2i = (i >= A.n_planets) ? A.eps2 : 0.0;
    double rb[3] = {0, 0, 0}, vb[3] = {0, 0, 0}, Vb[4][3], Ab[4][3], gmu = 0.0;
    if (ok) {
        for (int c = 0; c < 3; ++c) { rb[c] = A.rb[c * na + i]; vb[c] = A.vb[c * na + i]; }
        gmu = A.g_mu[i];
    }
    for (int st = 0; st < 4; ++st)
        for (int c = 0; c < 3; ++c) { Vb[st][c] = w[st] * rb[c]; Ab[st][c] = w[st] * vb[c]; }
#pragma unroll
    for (int ST = 3; ST >= 0; --ST) {
        __syncthreads();   // the previous stage's readers of sp and sa are done
        if (ok) {
            sp[i] = A.p64[ST][i];
            sa[i] = make_double4(Ab[ST][0], Ab[ST][1], Ab[ST][2], (i >= A.n_planets) ? 1.0 : 0.0);
        }
        __syncthreads();
        if (!ok) continue;
        const double4 pk = sp[i];
        double gx = 0.0, gy = 0.0, gz = 0.0, gm = 0.0;
        for (int j = 0; j < ng; ++j) {
            const double4 aj = sa[j];
            vjp_pair_fp64(pk, Ab[ST][0], Ab[ST][1], Ab[ST][2], sp[j], aj.x, aj.y, aj.z, e2i * aj.w, gx, gy, gz,
                          gm);
        }
        double Rb[3] = {0.0, 0.0, 0.0}, mb = 0.0;
        Rb[0] += gx; Rb[1] += gy; Rb[2] += gz; mb += gm;
        if (with_grains) {
            for (int c = 0; c < 3; ++c) Rb[c] += A.Rbg[((size_t)ST * 3 + c) * n + i];
            if (ST == 3) mb += A.mug[i];
        }
        gmu += mb;
        for (int c = 0; c < 3; ++c) {
            const double V = Vb[ST][c];
            rb[c] += Rb[c];
            vb[c] += V;
            if (ST > 0) {
                Vb[ST - 1][c] += hh[ST - 1] * Rb[c];
                Ab[ST - 1][c] += hh[ST - 1] * V;
            }
        }
    }
    if (ok) {
        for (int c = 0; c < 3; ++c) { A.rb[c * na + i] = rb[c]; A.vb[c * na + i] = vb[c]; }
        A.g_mu[i] = gmu;
    }
}

// Copy of the full state (x, y, z, vx, vy, vz, N: 7 blocks of n_all).
__global__ void k_copy_state(size_t n_all, const double* x, const double* y, const double* z,
                             const double* vx, const double* vy, const double* vz, const double* N,
                             double* dst)
{
    const size_t i = (size_t)blockIdx.x * kWG + threadIdx.x;
    if (i >= n_all) return;
    dst[i] = x[i]; dst[n_all + i] = y[i]; dst[2 * n_all + i] = z[i];
    dst[3 * n_all + i] = vx[i]; dst[4 * n_all + i] = vy[i]; dst[5 * n_all + i] = vz[i];
    dst[6 * n_all + i] = N[i];
}

__global__ void k_restore_state(size_t n_all, const double* src, double* x, double* y, double* z,
                                double* vx, double* vy, double* vz, double* N)
{
    const size_t i = (size_t)blockIdx.x * kWG + threadIdx.x;
    if (i >= n_all) return;
    x[i] = src[i]; y[i] = src[n_all + i]; z[i] = src[2 * n_all + i];
    vx[i] = src[3 * n_all + i]; vy[i] = src[4 * n_all + i]; vz[i] = src[5 * n_all + i];
    N[i] = src[6 * n_all + i];
}

// ---------------------------------------------------------------------------
// Potential of the massive set (diagnostics): phi_i = -sum_{j != i} mu_j / R_ij
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWG) void k_potential(const double4* __restrict__ pos, int n_grav,
                                                   int n_planets, double eps2,
                                                   double* __restrict__ phi)
{
    __shared__ double4 tile[kTile];
    const int i = blockIdx.x * kWG + threadIdx.x;
    const double4 pi = (i < n_grav) ? pos[i] : make_double4(kFar, kFar, kFar, 0.0);
    const bool sw = i >= n_planets;
    double u = 0.0;
    for (int j0 = 0; j0 < n_grav; j0 += kTile) {
        const int tl = min(kTile, n_grav - j0);
        __syncthreads();
        if ((int)threadIdx.x < tl) tile[threadIdx.x] = pos[j0 + threadIdx.x];
        __syncthreads();
        for (int jj = 0; jj < tl; ++jj) {
            const int j = j0 + jj;
            const double4 s = tile[jj];
            const double dx = s.x - pi.x, dy = s.y - pi.y, dz = s.z - pi.z;
            const double R2 = dx * dx + dy * dy + dz * dz + ((sw && j >= n_planets) ? eps2 : 0.0);
            if (j != i && R2 > 0.0) u -= s.w / __builtin_sqrt(R2);
        }
    }
    if (i < n_grav) phi[i] = u;
}

// The library's FP64 reciprocal square root on n inputs (accuracy check).
__global__ void k_rsqrt_fp64(const double* __restrict__ x, double* __restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (i < n) y[i] = rsqrt_fp64(x[i]);
}

}  // namespace

// ---------------------------------------------------------------------------
// Static scratch of ddyn_accel_noalloc: the stage-0 positions (FP64 and FP32
// shadow) and the pair kernels' slice partials live here instead of in
// library allocations.  32 MiB per device for the life of the process; every
// call on a device shares it, so calls that overlap on one device race.
// Host code never passes `scratch` itself to a kernel: in host code the name
// is a stand-in, not a device address.  ddyn_create looks up each device's
// copy once with hipGetSymbolAddress and every use goes through that address.
constexpr size_t kScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

// ---------------------------------------------------------------------------
// Context and C ABI
// ---------------------------------------------------------------------------
// (struct ddyn_ctx is in ddyn_ctx.h, shared with ddyn_image.hip)

namespace {

// Device address of `scratch` on each device, looked up once per device.
constexpr int kMaxDevices = 64;
void* g_scratch_addr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// With `device` current: the device's scratch address, looked up on first use.
hipError_t scratch_addr(int device, void** out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[device] = p;
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}

int fail(ddyn_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

int hipfail(ddyn_ctx* c, hipError_t e, const char* where)
{
    return fail(c, DDYN_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define DDYN_HIP(c, call)                                          \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return hipfail((c), e_, #call);      \
    } while (0)

template <class T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Frees the adjoint tape, after the work queued on the main stream that may
// still write it; every call that changes the state or the physics makes the
// tape stale.
void drop_tape(ddyn_ctx* c)
{
    if (c->tape_ck || c->tape_seg || c->tape_fin || c->tape_stg) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
    }
    dfree(c->tape_ck); dfree(c->tape_seg); dfree(c->tape_fin); dfree(c->tape_stg);
    c->have_tape = false;
}

void free_state(ddyn_ctx* c)
{
    dfree(c->x); dfree(c->y); dfree(c->z); dfree(c->vx); dfree(c->vy); dfree(c->vz);
    dfree(c->mu); dfree(c->N); dfree(c->id);
    for (int s = 0; s < 4; ++s) { dfree(c->p64[s]); dfree(c->p32[s]); dfree(c->q64[s]); dfree(c->q32[s]); }
    dfree(c->tape_ck); dfree(c->tape_seg); dfree(c->tape_fin); dfree(c->tape_stg);
    c->have_tape = false;
    dfree(c->stage); dfree(c->part); dfree(c->acc); dfree(c->rem); dfree(c->small_stg);
    dfree(c->beta_p); dfree(c->betasw_p);
    c->use_beta_p = false; c->have_betasw_p = false;
    c->n_grav = c->n_dust = 0;
    c->nslice = 0;
}

template <class T>
int dalloc(ddyn_ctx* c, T*& p, size_t count)
{
    p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) { p = nullptr; return fail(c, DDYN_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)); }
    return DDYN_OK;
}

// HIP event pair around one launch on stream `st` (kernel timing, when on).
struct TimedLaunch {
    ddyn_ctx* c;
    int cls;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(ddyn_ctx* c_, int cls_, hipStream_t st_ = nullptr) : c(c_), cls(cls_), st(st_ ? st_ : c_->stream)
    {
        if (!c->timing) return;
        if (!c->event_pool.empty()) {
            a = c->event_pool.back().a;
            b = c->event_pool.back().b;
            c->event_pool.pop_back();
        } else if (hipEventCreate(&a) != hipSuccess) {
            a = nullptr;
            return;
        } else if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            a = nullptr;
            return;
        }
        (void)hipEventRecord(a, st);
    }
    ~TimedLaunch()
    {
        if (!a) return;
        (void)hipEventRecord(b, st);
        c->events.push_back({cls, a, b});
    }
};

void drain_events(ddyn_ctx* c)
{
    for (auto& e : c->events) {
        float ms = 0.f;
        if (hipEventSynchronize(e.b) == hipSuccess && hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            c->t_ms[e.cls] += ms;
            c->t_launch[e.cls] += 1;
        }
        c->event_pool.push_back(e);
    }
    c->events.clear();
}

void destroy_events(ddyn_ctx* c)
{
    drain_events(c);
    for (auto& e : c->event_pool) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    c->event_pool.clear();
}

int n_swarm(const ddyn_ctx* c) { return (int)(c->n_grav - c->prm.n_planets); }

// The source-slice split of the pair kernels for n_grav targets and ns swarm
// sources, so that a launch has about kTargetWGs workgroups; at most
// max_nslice slices (when > 0).
void pair_split(int64_t n_grav, int ns, int max_nslice, int* nslice, int* slice_len)
{
    if (ns <= 0) { *nslice = 0; *slice_len = 0; return; }
    const int tblocks = (int)((n_grav + kWG - 1) / kWG);
    int want = (kTargetWGs + tblocks - 1) / tblocks;
    // few targets (a small swarm): slices down to one tile, the launch is latency-bound
    const int min_slice = (tblocks < 64) ? kTile : kMinSlice;
    const int max_slices = std::max(1, ns / min_slice);
    if (want > max_slices) want = max_slices;
    if (max_nslice > 0 && want > max_nslice) want = max_nslice;
    if (want < 1) want = 1;
    int len = (ns + want - 1) / want;
    len = ((len + kTile - 1) / kTile) * kTile;
    *slice_len = len;
    *nslice = (ns + len - 1) / len;
}

// Choose the source-slice split of the uploaded state (one target per lane).
void plan_pairs(ddyn_ctx* c)
{
    c->tpt = 1;
    pair_split(c->n_grav, n_swarm(c), 0, &c->nslice, &c->slice_len);
}

// Without softening the default FP64 radius 10 eps is 0: no swarm pair would
// be redone in FP64, unsoftened close pairs would come from FP32 positions and
// an exact hit on a swarm body would never be seen.  So a swarm in the
// FP32-force form with eps = 0 needs an FP64 radius of its own.
int check_fp64_radius(ddyn_ctx* c, int64_t n_swarm_bodies)
{
    if (c->prm.precision == DDYN_FP32_FORCE && n_swarm_bodies > 0 && c->prm.eps == 0.0 &&
        !(c->prm.fp64_radius > 0.0))
        return fail(c, DDYN_ERR_ARG, "DDYN_FP32_FORCE with a swarm and eps = 0 needs fp64_radius > 0");
    return DDYN_OK;
}

// Parameters and a state are present (enough for diagnostics).
int check_state(ddyn_ctx* c)
{
    if (!c) return DDYN_ERR_ARG;
    if (!c->have_params) return fail(c, DDYN_ERR_STATE, "ddyn_set_params has not been called");
    if (!c->x) return fail(c, DDYN_ERR_STATE, "ddyn_upload_state has not been called");
    if (c->prm.n_planets > c->n_grav)
        return fail(c, DDYN_ERR_ARG, "n_planets exceeds n_grav");
    return DDYN_OK;
}

// ... and the forces can be evaluated (steps, accelerations, the adjoint).
int check_ready(ddyn_ctx* c)
{
    const int rc = check_state(c);
    return rc ? rc : check_fp64_radius(c, n_swarm(c));
}

double eps2_of(const ddyn_ctx* c) { return c->prm.eps * c->prm.eps; }

double r64_of(const ddyn_ctx* c)
{
    return c->prm.fp64_radius > 0.0 ? c->prm.fp64_radius : 10.0 * c->prm.eps;
}

// Swarm-pair forces of the stage positions p32/p64 (ng bodies) into part, in
// nslice slices of slice_len sources.
int launch_pairs_on(ddyn_ctx* c, const float4* p32, const double4* p64, int ng, int nslice, int slice_len,
                    double* part)
{
    if (nslice == 0) return DDYN_OK;
    TimedLaunch tl(c, DDYN_KCLASS_PAIRS);
    const int np = c->prm.n_planets;
    dim3 grid((unsigned)((ng + kWG - 1) / kWG), (unsigned)nslice);
    if (c->prm.precision == DDYN_FP32_FORCE) {
        const double r64 = r64_of(c);
        k_pairs32<1><<<grid, kWG, 0, c->stream>>>(p32, p64, ng, np, slice_len, (float)eps2_of(c),
                                                   (float)(r64 * r64), eps2_of(c), part);
    } else if (c->prm.precision == DDYN_FP64_RSQ) {
        k_pairs64<1, true><<<grid, kWG, 0, c->stream>>>(p64, ng, np, slice_len, eps2_of(c), part);
    } else {
        k_pairs64<1, false><<<grid, kWG, 0, c->stream>>>(p64, ng, np, slice_len, eps2_of(c), part);
    }
    DDYN_HIP(c, hipGetLastError());
    return DDYN_OK;
}

// Swarm-pair forces at stage s into c->part.
int launch_pairs(ddyn_ctx* c, int s)
{
    return launch_pairs_on(c, c->p32[s], c->p64[s], (int)c->n_grav, c->nslice, c->slice_len, c->part);
}

StageArgs stage_args(ddyn_ctx* c, double h)
{
    StageArgs A{};
    const size_t n = (size_t)c->n_grav;
    A.n_grav = (int)c->n_grav; A.n_planets = c->prm.n_planets; A.nslice = c->nslice;
    A.part = c->part;
    A.x = c->x; A.y = c->y; A.z = c->z; A.vx = c->vx; A.vy = c->vy; A.vz = c->vz; A.mu = c->mu;
    for (int s = 0; s < 4; ++s) { A.p64[s] = c->p64[s]; A.p32[s] = c->p32[s]; }
    A.vrx = c->stage; A.vry = c->stage + n; A.vrz = c->stage + 2 * n;
    A.Srx = c->stage + 3 * n; A.Sry = c->stage + 4 * n; A.Srz = c->stage + 5 * n;
    A.Sax = c->stage + 6 * n; A.Say = c->stage + 7 * n; A.Saz = c->stage + 8 * n;
    A.n64 = c->p64[0]; A.n32 = c->p32[0];
    A.h = h;
    return A;
}

int launch_stage(ddyn_ctx* c, int s, const StageArgs& A)
{
    TimedLaunch tl(c, DDYN_KCLASS_STAGE);
    const unsigned blocks = (unsigned)((A.n_grav + kWG - 1) / kWG);
    if (c->prm.precision == DDYN_FP64_RSQ) {
        switch (s) {
            case 0: k_stage<0, true><<<blocks, kWG, 0, c->stream>>>(A); break;
            case 1: k_stage<1, true><<<blocks, kWG, 0, c->stream>>>(A); break;
            case 2: k_stage<2, true><<<blocks, kWG, 0, c->stream>>>(A); break;
            case 3: k_stage<3, true><<<blocks, kWG, 0, c->stream>>>(A); break;
            default: k_stage<4, true><<<blocks, kWG, 0, c->stream>>>(A); break;
        }
    } else {
        switch (s) {
            case 0: k_stage<0, false><<<blocks, kWG, 0, c->stream>>>(A); break;
            case 1: k_stage<1, false><<<blocks, kWG, 0, c->stream>>>(A); break;
            case 2: k_stage<2, false><<<blocks, kWG, 0, c->stream>>>(A); break;
            case 3: k_stage<3, false><<<blocks, kWG, 0, c->stream>>>(A); break;
            default: k_stage<4, false><<<blocks, kWG, 0, c->stream>>>(A); break;
        }
    }
    DDYN_HIP(c, hipGetLastError());
    return DDYN_OK;
}

DustArgs dust_args(ddyn_ctx* c, double dt)
{
    DustArgs A{};
    A.n_grav = (int)c->n_grav; A.n_planets = c->prm.n_planets; A.n_dust = (int)c->n_dust;
    A.n_bins = c->n_bins;
    for (int s = 0; s < 4; ++s) { A.p64[s] = c->p64[s]; A.p32[s] = c->p32[s]; }
    A.x = c->x; A.y = c->y; A.z = c->z; A.vx = c->vx; A.vy = c->vy; A.vz = c->vz;
    A.mu = c->mu; A.N = c->N; A.id = c->id;
    A.beta_p = c->use_beta_p ? c->beta_p : nullptr;
    A.betasw_p = (c->use_beta_p && c->have_betasw_p) ? c->betasw_p : nullptr;
    A.beta_t = c->beta_t; A.betasw_t = c->betasw_t; A.q_t = c->q_t;
    A.dt = dt;
    A.eps2 = eps2_of(c);
    A.Rmin2 = c->prm.Rmin * c->prm.Rmin;
    A.vsw = c->prm.vsw; A.Wz = c->prm.Wz; A.Bstar = c->prm.Bstar; A.Rstar = c->prm.Rstar;
    A.light = c->prm.light > 0.0 ? c->prm.light : kLightAU;
    A.eps2f = (float)A.eps2;
    const double r64 = r64_of(c);
    A.thr2f = (float)(r64 * r64 + A.eps2);
    A.SWQ = c->prm.SWQ; A.FLQ = c->prm.FLQ;
    return A;
}

// Grain kernel over 64-grain blocks [blk0, blk0 + nblk) (all blocks when nblk < 0).
int launch_dust(ddyn_ctx* c, DustArgs A, bool accel, hipStream_t st = nullptr, int blk0 = 0, int nblk = -1)
{
    if (A.n_dust == 0) return DDYN_OK;
    if (!st) st = c->stream;
    TimedLaunch tl(c, DDYN_KCLASS_DUST, st);
    const int all = (int)((A.n_dust + 63) / 64);
    if (nblk < 0) nblk = all - blk0;
    if (blk0 < 0 || nblk <= 0 || blk0 + nblk > all) return fail(c, DDYN_ERR_ARG, "grain block range");
    A.blk0 = blk0;
    const unsigned blocks = (unsigned)nblk;
    const int form = c->prm.precision;
    if (accel) {
        if (form == DDYN_FP32_FORCE) k_dust<true, true><<<blocks, kDustWG, 0, st>>>(A);
        else if (form == DDYN_FP64_RSQ) k_dust<false, true, true><<<blocks, kDustWG, 0, st>>>(A);
        else k_dust<false, true><<<blocks, kDustWG, 0, st>>>(A);
    } else {
        if (form == DDYN_FP32_FORCE) k_dust<true, false><<<blocks, kDustWG, 0, st>>>(A);
        else if (form == DDYN_FP64_RSQ) k_dust<false, false, true><<<blocks, kDustWG, 0, st>>>(A);
        else k_dust<false, false><<<blocks, kDustWG, 0, st>>>(A);
    }
    DDYN_HIP(c, hipGetLastError());
    return DDYN_OK;
}

int one_step(ddyn_ctx* c, double dt)
{
    int rc;
    const double hs[3] = {0.5 * dt, 0.5 * dt, dt};
    for (int s = 0; s < 3; ++s) {
        if ((rc = launch_pairs(c, s))) return rc;
        if ((rc = launch_stage(c, s, stage_args(c, hs[s])))) return rc;
    }
    if ((rc = launch_pairs(c, 3))) return rc;
    if ((rc = launch_dust(c, dust_args(c, dt), false))) return rc;   // before the final update
    return launch_stage(c, 3, stage_args(c, dt / 6.0));
}

#ifndef DDYN_GRAIN_STREAMS
#define DDYN_GRAIN_STREAMS 2
#endif
// With many grains the overlapped schedule splits the grain blocks over two
// streams, so that the next step's launch of one half can start while the
// other half's launch drains (results unchanged: each grain's arithmetic does
// not depend on the launch it is in).  Measured: 3% off the step with
// 1,048,576 grains, no change with 262,144.
bool grain_split(const ddyn_ctx* c) { return DDYN_GRAIN_STREAMS == 2 && c->n_dust >= 2 * 64 * 1024; }

// Step k of a run of steps with the grain kernel on its own stream, so that it
// overlaps the massive stages of step k+1 (same kernels and arguments as
// one_step, so the results are identical).  The massive stage positions are
// double buffered by step parity: the final massive stage writes the next
// step's stage-0 positions into the other set, after the grain kernel of step
// k-1 (the last reader of that set) has finished.  The caller joins the grain
// stream after the last step.
int overlapped_step(ddyn_ctx* c, double dt, int64_t k)
{
    int rc;
    const double hs[3] = {0.5 * dt, 0.5 * dt, dt};
    for (int s = 0; s < 3; ++s) {
        if ((rc = launch_pairs(c, s))) return rc;
        if ((rc = launch_stage(c, s, stage_args(c, hs[s])))) return rc;
    }
    DDYN_HIP(c, hipEventRecord(c->ev_stage[k & 1], c->stream));
    if ((rc = launch_pairs(c, 3))) return rc;
    const int blocks = (int)((c->n_dust + 63) / 64);
    const int half = grain_split(c) ? blocks / 2 : blocks;
    DDYN_HIP(c, hipStreamWaitEvent(c->gstream, c->ev_stage[k & 1], 0));
    if ((rc = launch_dust(c, dust_args(c, dt), false, c->gstream, 0, half))) return rc;
    DDYN_HIP(c, hipEventRecord(c->ev_dust[k & 1], c->gstream));
    if (half < blocks) {
        DDYN_HIP(c, hipStreamWaitEvent(c->gstream2, c->ev_stage[k & 1], 0));
        if ((rc = launch_dust(c, dust_args(c, dt), false, c->gstream2, half, blocks - half))) return rc;
        DDYN_HIP(c, hipEventRecord(c->ev_dust2[k & 1], c->gstream2));
    }
    if (k > 0) {
        DDYN_HIP(c, hipStreamWaitEvent(c->stream, c->ev_dust[(k - 1) & 1], 0));
        if (half < blocks) DDYN_HIP(c, hipStreamWaitEvent(c->stream, c->ev_dust2[(k - 1) & 1], 0));
    }
    StageArgs S = stage_args(c, dt / 6.0);
    S.n64 = c->q64[0];
    S.n32 = c->q32[0];
    if ((rc = launch_stage(c, 3, S))) return rc;
    for (int s = 0; s < 4; ++s) {
        std::swap(c->p64[s], c->q64[s]);
        std::swap(c->p32[s], c->q32[s]);
    }
    return DDYN_OK;
}

}  // namespace

extern "C" {

int ddyn_create(ddyn_ctx** out, int device)
{
    if (!out) return DDYN_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return DDYN_ERR_HIP;
    if (device < 0 || device >= ndev) return DDYN_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return DDYN_ERR_HIP;
    ddyn_ctx* c = new ddyn_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return DDYN_ERR_HIP;
    }
    bool ok = hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&c->gstream2, hipStreamNonBlocking) == hipSuccess;
    for (int k = 0; k < 2 && ok; ++k)
        ok = hipEventCreateWithFlags(&c->ev_stage[k], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_dust[k], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_dust2[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        ddyn_destroy(c);
        return DDYN_ERR_HIP;
    }
    if (dalloc(c, c->beta_t, kMaxBins) || dalloc(c, c->betasw_t, kMaxBins) || dalloc(c, c->q_t, kMaxBins)) {
        ddyn_destroy(c);
        return DDYN_ERR_NOMEM;
    }
    // the device's static scratch of ddyn_accel_noalloc (looked up once per device)
    void* sp = nullptr;
    const hipError_t se = scratch_addr(device, &sp);
    if (se != hipSuccess) c->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
    *out = c;
    return DDYN_OK;
}

void ddyn_destroy(ddyn_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gstream) (void)hipStreamSynchronize(c->gstream);
    if (c->gstream2) (void)hipStreamSynchronize(c->gstream2);
    destroy_events(c);
    free_state(c);
    ddyn_image_release(c);
    dfree(c->beta_t); dfree(c->betasw_t); dfree(c->q_t);
    for (int k = 0; k < 2; ++k) {
        if (c->ev_stage[k]) (void)hipEventDestroy(c->ev_stage[k]);
        if (c->ev_dust[k]) (void)hipEventDestroy(c->ev_dust[k]);
        if (c->ev_dust2[k]) (void)hipEventDestroy(c->ev_dust2[k]);
    }
    if (c->gstream) (void)hipStreamDestroy(c->gstream);
    if (c->gstream2) (void)hipStreamDestroy(c->gstream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* ddyn_last_error(const ddyn_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ddyn_set_params(ddyn_ctx* c, const ddyn_params* p)
{
    if (!c || !p) return DDYN_ERR_ARG;
    if (p->n_planets < 1) return fail(c, DDYN_ERR_ARG, "n_planets must count the star (>= 1)");
    if (p->precision != DDYN_FP64 && p->precision != DDYN_FP32_FORCE && p->precision != DDYN_FP64_RSQ)
        return fail(c, DDYN_ERR_ARG, "unknown precision form");
    if (!(p->eps >= 0.0) || !(p->Rmin >= 0.0)) return fail(c, DDYN_ERR_ARG, "eps and Rmin must be >= 0");
    if (p->SWQ == 1 && !(p->vsw > 0.0)) return fail(c, DDYN_ERR_ARG, "SWQ needs vsw > 0");
    if (c->x && p->n_planets > c->n_grav) return fail(c, DDYN_ERR_ARG, "n_planets exceeds n_grav");
    drop_tape(c);
    c->prm = *p;
    c->have_params = true;
    c->fuse_small = p->schedule == 0;
    if (c->x) {
        (void)hipSetDevice(c->device);
        const int old = c->nslice;
        plan_pairs(c);
        if (c->nslice > old || (c->nslice && !c->part)) {
            dfree(c->part);
            if (dalloc(c, c->part, (size_t)c->nslice * 3 * (size_t)c->n_grav)) return DDYN_ERR_NOMEM;
        }
    }
    return DDYN_OK;
}

int ddyn_upload_state(ddyn_ctx* c, int64_t n_grav, int64_t n_dust, const double* x, const double* y,
                      const double* z, const double* vx, const double* vy, const double* vz,
                      const double* mu, const double* N, const int32_t* id)
{
    if (!c) return DDYN_ERR_ARG;
    if (n_grav < 1 || n_dust < 0 || n_grav > (int64_t)1 << 30 || n_dust > (int64_t)1 << 30)
        return fail(c, DDYN_ERR_ARG, "need 1 <= n_grav and 0 <= n_dust, each below 2^30");
    if (!x || !y || !z || !vx || !vy || !vz || !mu || !N)
        return fail(c, DDYN_ERR_ARG, "state arrays must not be NULL");
    if (c->have_params && c->prm.n_planets > n_grav)
        return fail(c, DDYN_ERR_ARG, "n_planets exceeds n_grav");
    DDYN_HIP(c, hipSetDevice(c->device));
    // nothing queued on either of the context's streams may still use the buffers
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    DDYN_HIP(c, hipStreamSynchronize(c->gstream));
    DDYN_HIP(c, hipStreamSynchronize(c->gstream2));
    free_state(c);
    const size_t na = (size_t)(n_grav + n_dust), ng = (size_t)n_grav;
    double** arr[] = {&c->x, &c->y, &c->z, &c->vx, &c->vy, &c->vz, &c->mu, &c->N};
    const double* src[] = {x, y, z, vx, vy, vz, mu, N};
    for (int a = 0; a < 8; ++a)
        if (dalloc(c, *arr[a], na)) { free_state(c); return DDYN_ERR_NOMEM; }
    if (dalloc(c, c->id, na)) { free_state(c); return DDYN_ERR_NOMEM; }
    for (int s = 0; s < 4; ++s)
        if (dalloc(c, c->p64[s], ng) || dalloc(c, c->p32[s], ng)) { free_state(c); return DDYN_ERR_NOMEM; }
    if (n_dust > 0)
        for (int s = 0; s < 4; ++s)
            if (dalloc(c, c->q64[s], ng) || dalloc(c, c->q32[s], ng)) { free_state(c); return DDYN_ERR_NOMEM; }
    if (dalloc(c, c->stage, 9 * ng) || dalloc(c, c->acc, 3 * na) ||
        dalloc(c, c->rem, (size_t)(n_dust > 0 ? n_dust : 1)) || dalloc(c, c->beta_p, (size_t)(n_dust > 0 ? n_dust : 1)) ||
        dalloc(c, c->betasw_p, (size_t)(n_dust > 0 ? n_dust : 1))) {
        free_state(c);
        return DDYN_ERR_NOMEM;
    }
    c->n_grav = n_grav; c->n_dust = n_dust;
    for (int a = 0; a < 8; ++a) DDYN_HIP(c, hipMemcpy(*arr[a], src[a], na * sizeof(double), hipMemcpyHostToDevice));
    if (id) {
        DDYN_HIP(c, hipMemcpy(c->id, id, na * sizeof(int32_t), hipMemcpyHostToDevice));
    } else {
        std::vector<int32_t> m1(na, -1);
        DDYN_HIP(c, hipMemcpy(c->id, m1.data(), na * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    k_init_pos<<<(unsigned)((ng + kWG - 1) / kWG), kWG, 0, c->stream>>>((int)ng, c->x, c->y, c->z, c->mu,
                                                                       c->p64[0], c->p32[0]);
    DDYN_HIP(c, hipGetLastError());
    if (c->have_params) {
        plan_pairs(c);
        if (c->nslice && dalloc(c, c->part, (size_t)c->nslice * 3 * ng)) { free_state(c); return DDYN_ERR_NOMEM; }
    }
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    return DDYN_OK;
}

int ddyn_upload_bin_tables(ddyn_ctx* c, int32_t n_bins, const float* beta, const float* betasw, const float* q)
{
    if (!c) return DDYN_ERR_ARG;
    if (n_bins < 0 || n_bins > kMaxBins)
        return fail(c, DDYN_ERR_ARG, "n_bins must be in [0, 2700] (nsize)");
    if (n_bins > 0 && (!beta || !betasw || !q)) return fail(c, DDYN_ERR_ARG, "tables must not be NULL");
    DDYN_HIP(c, hipSetDevice(c->device));
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    drop_tape(c);
    if (n_bins > 0) {
        DDYN_HIP(c, hipMemcpy(c->beta_t, beta, n_bins * sizeof(float), hipMemcpyHostToDevice));
        DDYN_HIP(c, hipMemcpy(c->betasw_t, betasw, n_bins * sizeof(float), hipMemcpyHostToDevice));
        DDYN_HIP(c, hipMemcpy(c->q_t, q, n_bins * sizeof(float), hipMemcpyHostToDevice));
    }
    c->n_bins = n_bins;
    return DDYN_OK;
}

int ddyn_upload_grain_beta(ddyn_ctx* c, const double* beta, const double* betasw)
{
    if (!c) return DDYN_ERR_ARG;
    if (!c->x) return fail(c, DDYN_ERR_STATE, "ddyn_upload_state has not been called");
    DDYN_HIP(c, hipSetDevice(c->device));
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    drop_tape(c);
    if (!beta) { c->use_beta_p = false; c->have_betasw_p = false; return DDYN_OK; }
    const size_t nd = (size_t)c->n_dust;
    if (nd) {
        DDYN_HIP(c, hipMemcpy(c->beta_p, beta, nd * sizeof(double), hipMemcpyHostToDevice));
        if (betasw) DDYN_HIP(c, hipMemcpy(c->betasw_p, betasw, nd * sizeof(double), hipMemcpyHostToDevice));
    }
    c->use_beta_p = true;
    c->have_betasw_p = betasw != nullptr;
    return DDYN_OK;
}

int ddyn_step(ddyn_ctx* c, double dt, int64_t nsteps)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!(dt > 0.0) || nsteps < 0) return fail(c, DDYN_ERR_ARG, "need dt > 0 and nsteps >= 0");
    DDYN_HIP(c, hipSetDevice(c->device));
    drop_tape(c);
    if (n_swarm(c) == 0 && c->n_grav <= kSmallMax && c->fuse_small) {
        // star + planets only: many steps per launch (bounded launch length)
        DustArgs A = dust_args(c, dt);
        A.mfin = c->stage;   // 9 * n_grav of stage scratch, unused by this path
        const bool split = c->n_dust >= kSmallSplitMin;
        const unsigned blocks = (unsigned)std::max<int64_t>(1, (c->n_dust + 63) / 64);
        if (split && !c->small_stg && (rc = dalloc(c, c->small_stg, (size_t)kSmallChunk * 4 * kSmallMax)))
            return rc;
        for (int64_t k = 0; k < nsteps; k += kSmallChunk) {
            const int n = (int)std::min<int64_t>(kSmallChunk, nsteps - k);
            TimedLaunch tl(c, DDYN_KCLASS_DUST);
            const bool rsq = c->prm.precision == DDYN_FP64_RSQ;
            if (split) {
                if (rsq) launch_small_split<true>(A, c->small_stg, c->n_dust, c->stream, n);
                else launch_small_split<false>(A, c->small_stg, c->n_dust, c->stream, n);
            } else {
                if (rsq) launch_small<true>(A, blocks, c->stream, n);
                else launch_small<false>(A, blocks, c->stream, n);
            }
            DDYN_HIP(c, hipGetLastError());
            k_small_commit<<<1, 64, 0, c->stream>>>(A);
            DDYN_HIP(c, hipGetLastError());
        }
        return DDYN_OK;
    }
    if (c->n_dust > 0 && c->fuse_small) {
        for (int64_t k = 0; k < nsteps; ++k)
            if ((rc = overlapped_step(c, dt, k))) return rc;
        if (nsteps > 0) {
            DDYN_HIP(c, hipStreamWaitEvent(c->stream, c->ev_dust[(nsteps - 1) & 1], 0));
            if (grain_split(c)) DDYN_HIP(c, hipStreamWaitEvent(c->stream, c->ev_dust2[(nsteps - 1) & 1], 0));
        }
        return DDYN_OK;
    }
    for (int64_t k = 0; k < nsteps; ++k)
        if ((rc = one_step(c, dt))) return rc;
    return DDYN_OK;
}

int ddyn_accel(ddyn_ctx* c, double* ax, double* ay, double* az, int32_t* removed)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!ax || !ay || !az) return fail(c, DDYN_ERR_ARG, "output arrays must not be NULL");
    DDYN_HIP(c, hipSetDevice(c->device));
    const size_t na = (size_t)(c->n_grav + c->n_dust), ng = (size_t)c->n_grav;
    if ((rc = launch_pairs(c, 0))) return rc;
    StageArgs S = stage_args(c, 0.0);
    S.ox = c->acc; S.oy = c->acc + na; S.oz = c->acc + 2 * na;
    if ((rc = launch_stage(c, 4, S))) return rc;
    DustArgs D = dust_args(c, 0.0);
    D.ox = c->acc + ng; D.oy = c->acc + na + ng; D.oz = c->acc + 2 * na + ng;
    D.removed = c->rem;
    if ((rc = launch_dust(c, D, true))) return rc;
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    DDYN_HIP(c, hipMemcpy(ax, c->acc, na * sizeof(double), hipMemcpyDeviceToHost));
    DDYN_HIP(c, hipMemcpy(ay, c->acc + na, na * sizeof(double), hipMemcpyDeviceToHost));
    DDYN_HIP(c, hipMemcpy(az, c->acc + 2 * na, na * sizeof(double), hipMemcpyDeviceToHost));
    if (removed && c->n_dust)
        DDYN_HIP(c, hipMemcpy(removed, c->rem, (size_t)c->n_dust * sizeof(int32_t), hipMemcpyDeviceToHost));
    return DDYN_OK;
}

int ddyn_accel_noalloc(ddyn_ctx* c, int64_t n_grav, int64_t n_dust, const double* x, const double* y,
                       const double* z, const double* vx, const double* vy, const double* vz, const double* mu,
                       const int32_t* id, const double* beta_p, const double* betasw_p, double* ax, double* ay,
                       double* az, int32_t* removed)
{
    if (!c) return DDYN_ERR_ARG;
    if (!c->have_params) return fail(c, DDYN_ERR_STATE, "ddyn_set_params has not been called");
    if (n_grav < 1 || n_dust < 0 || n_grav > (int64_t)1 << 30 || n_dust > (int64_t)1 << 30)
        return fail(c, DDYN_ERR_ARG, "need 1 <= n_grav and 0 <= n_dust, each below 2^30");
    if (!x || !y || !z || !vx || !vy || !vz || !mu || !ax || !ay || !az)
        return fail(c, DDYN_ERR_ARG, "state and output arrays must not be NULL");
    if (c->prm.n_planets > n_grav) return fail(c, DDYN_ERR_ARG, "n_planets exceeds n_grav");
    int rc = check_fp64_radius(c, n_grav - c->prm.n_planets);
    if (rc) return rc;
    // scratch layout: stage-0 positions (double4, float4) then the slice partials
    const size_t ng = (size_t)n_grav;
    const size_t pos_bytes = ((ng * (sizeof(double4) + sizeof(float4)) + 255) / 256) * 256;
    if (pos_bytes + 3 * ng * sizeof(double) > kScratchBytes)
        return fail(c, DDYN_ERR_ARG, "n_grav too large for the static scratch of ddyn_accel_noalloc (" +
                                         std::to_string(kScratchBytes >> 20) + " MiB: at most " +
                                         std::to_string((kScratchBytes - 256) / 72) + " massive bodies)");
    const int max_nslice = (int)std::min<size_t>((kScratchBytes - pos_bytes) / (3 * ng * sizeof(double)), 1 << 20);
    int nslice = 0, slice_len = 0;
    pair_split(n_grav, (int)(n_grav - c->prm.n_planets), max_nslice, &nslice, &slice_len);
    DDYN_HIP(c, hipSetDevice(c->device));
    void* base = nullptr;
    DDYN_HIP(c, scratch_addr(c->device, &base));
    double4* p64 = (double4*)base;
    float4* p32 = (float4*)(p64 + ng);
    double* part = (double*)((char*)base + pos_bytes);
    k_init_pos<<<(unsigned)((ng + kWG - 1) / kWG), kWG, 0, c->stream>>>((int)ng, x, y, z, mu, p64, p32);
    DDYN_HIP(c, hipGetLastError());
    if ((rc = launch_pairs_on(c, p32, p64, (int)ng, nslice, slice_len, part))) return rc;
    StageArgs S{};
    S.n_grav = (int)ng; S.n_planets = c->prm.n_planets; S.nslice = nslice; S.part = part;
    S.p64[0] = p64; S.p32[0] = p32;
    S.ox = ax; S.oy = ay; S.oz = az;
    if ((rc = launch_stage(c, 4, S))) return rc;
    DustArgs D = dust_args(c, 0.0);
    D.n_grav = (int)ng; D.n_dust = (int)n_dust;
    for (int s = 0; s < 4; ++s) { D.p64[s] = p64; D.p32[s] = p32; }
    D.x = const_cast<double*>(x); D.y = const_cast<double*>(y); D.z = const_cast<double*>(z);
    D.vx = const_cast<double*>(vx); D.vy = const_cast<double*>(vy); D.vz = const_cast<double*>(vz);
    D.mu = mu; D.N = nullptr; D.id = id;
    D.beta_p = beta_p; D.betasw_p = beta_p ? betasw_p : nullptr;
    D.ox = ax + ng; D.oy = ay + ng; D.oz = az + ng;
    D.removed = removed;
    return launch_dust(c, D, true);
}

int ddyn_download_state(ddyn_ctx* c, double* x, double* y, double* z, double* vx, double* vy,
                        double* vz, double* N)
{
    if (!c) return DDYN_ERR_ARG;
    if (!c->x) return fail(c, DDYN_ERR_STATE, "ddyn_upload_state has not been called");
    DDYN_HIP(c, hipSetDevice(c->device));
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    const size_t bytes = (size_t)(c->n_grav + c->n_dust) * sizeof(double);
    double* dst[] = {x, y, z, vx, vy, vz, N};
    const double* src[] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->N};
    for (int a = 0; a < 7; ++a)
        if (dst[a]) DDYN_HIP(c, hipMemcpy(dst[a], src[a], bytes, hipMemcpyDeviceToHost));
    return DDYN_OK;
}

namespace {

// Device buffer freed when it goes out of scope.
extern "C++" {
template <class T>
struct DevBufT {
    T* p = nullptr;
    ~DevBufT() { if (p) (void)hipFree(p); }
};
}
using DevBuf = DevBufT<double>;
using DevBuf4 = DevBufT<double4>;

}  // namespace

namespace {

// Segment length and count of an adjoint run (ckpt_every <= 0 or >= nsteps:
// one segment).
int64_t seg_len(int64_t nsteps, int64_t ckpt_every)
{
    return (ckpt_every > 0 && ckpt_every < nsteps) ? ckpt_every : (nsteps > 0 ? nsteps : 1);
}
int64_t seg_count(int64_t nsteps, int64_t ck) { return nsteps > 0 ? (nsteps + ck - 1) / ck : 0; }

void save_state(ddyn_ctx* c, double* dst)
{
    const size_t na = (size_t)(c->n_grav + c->n_dust);
    k_copy_state<<<(unsigned)((na + kWG - 1) / kWG), kWG, 0, c->stream>>>(na, c->x, c->y, c->z, c->vx, c->vy,
                                                                          c->vz, c->N, dst);
}

void restore_state(ddyn_ctx* c, const double* src)
{
    const size_t na = (size_t)(c->n_grav + c->n_dust), ng = (size_t)c->n_grav;
    k_restore_state<<<(unsigned)((na + kWG - 1) / kWG), kWG, 0, c->stream>>>(na, src, c->x, c->y, c->z, c->vx,
                                                                             c->vy, c->vz, c->N);
    k_init_pos<<<(unsigned)((ng + kWG - 1) / kWG), kWG, 0, c->stream>>>((int)ng, c->x, c->y, c->z, c->mu,
                                                                        c->p64[0], c->p32[0]);
}

// Steps k0..k1-1 of a segment from the current state, keeping each step's state
// in tape_seg and its massive positions of stages 1-3 in tape_stg (the stage
// kernels write them there directly).  The last step is completed only when
// `finish` is set: the reverse sweep needs the states at the steps' starts, and
// the step after a segment is the next segment's checkpoint.
int tape_segment(ddyn_ctx* c, double dt, int64_t k0, int64_t k1, bool finish)
{
    const size_t na = (size_t)(c->n_grav + c->n_dust), ng = (size_t)c->n_grav;
    const double hs[3] = {0.5 * dt, 0.5 * dt, dt};
    double4* const own64[3] = {c->p64[1], c->p64[2], c->p64[3]};
    int rc = DDYN_OK;
    for (int64_t k = k0; k < k1 && !rc; ++k) {
        save_state(c, c->tape_seg + (size_t)(k - k0) * 7 * na);
        for (int st = 1; st <= 3; ++st) c->p64[st] = c->tape_stg + ((size_t)(k - k0) * 3 + (size_t)(st - 1)) * ng;
        for (int st = 0; st < 3 && !rc; ++st) {   // massive stages 0-2 of step k
            if (!(rc = launch_pairs(c, st))) rc = launch_stage(c, st, stage_args(c, hs[st]));
        }
        if (!rc && (k + 1 < k1 || finish)) {      // the rest of the step
            if (!(rc = launch_pairs(c, 3)) && !(rc = launch_dust(c, dust_args(c, dt), false)))
                rc = launch_stage(c, 3, stage_args(c, dt / 6.0));
        }
        for (int st = 1; st <= 3; ++st) c->p64[st] = own64[st - 1];
    }
    return rc;
}

// Forward pass of the adjoint: a checkpoint at the start of every segment, the
// last segment taped (tape_segment), the final state kept.  The context ends at
// the state after nsteps.
int record_tape(ddyn_ctx* c, double dt, int64_t nsteps, int64_t ckpt_every)
{
    drop_tape(c);
    const size_t na = (size_t)(c->n_grav + c->n_dust), ng = (size_t)c->n_grav;
    const int64_t ck = seg_len(nsteps, ckpt_every), nseg = seg_count(nsteps, ck);
    if (dalloc(c, c->tape_ck, (size_t)std::max<int64_t>(nseg, 1) * 7 * na) ||
        dalloc(c, c->tape_seg, (size_t)ck * 7 * na) || dalloc(c, c->tape_stg, (size_t)ck * 3 * ng) ||
        dalloc(c, c->tape_fin, 7 * na)) {
        drop_tape(c);
        return DDYN_ERR_NOMEM;
    }
    int rc;
    const int64_t klast = nseg > 0 ? (nseg - 1) * ck : 0;
    for (int64_t k = 0; k < klast; ++k) {
        if (k % ck == 0) save_state(c, c->tape_ck + (size_t)(k / ck) * 7 * na);
        if ((rc = one_step(c, dt))) { drop_tape(c); return rc; }
    }
    if (nseg > 0) {
        save_state(c, c->tape_ck + (size_t)(nseg - 1) * 7 * na);
        if ((rc = tape_segment(c, dt, klast, nsteps, true))) { drop_tape(c); return rc; }
    }
    save_state(c, c->tape_fin);
    DDYN_HIP(c, hipGetLastError());
    c->tape_dt = dt;
    c->tape_n = nsteps;
    c->tape_ck_every = ck;
    c->have_tape = true;
    c->tape_last = nseg > 0;
    return DDYN_OK;
}

// Reverse sweep over a recorded tape: the last segment from the tape, every
// earlier one recomputed from its checkpoint into the same buffers (after which
// the last segment is no longer on the tape, and a later sweep recomputes it
// too).  Ends at the recorded final state, with the tape still valid.
int reverse_tape(ddyn_ctx* c, const double* lam_r, const double* lam_v, double* g_r, double* g_v,
                 double* g_mu, double* g_beta)
{
    const double dt = c->tape_dt;
    const int64_t nsteps = c->tape_n, ck = c->tape_ck_every, nseg = seg_count(nsteps, ck);
    const size_t na = (size_t)(c->n_grav + c->n_dust), ng = (size_t)c->n_grav, nd = (size_t)c->n_dust;
    const int nwaves = (int)((nd + kWG - 1) / kWG) * (kWG / 64);
    // Grain pulls on the massive bodies: summed per grain wave when there are
    // few massive bodies (or in the one-launch-per-kernel schedule), else as a
    // pair kernel with massive targets and grain sources (k_adj_pull)
    const bool pull_pairs = nd > 0 && ng > (size_t)kAdjPullMin && c->fuse_small;
    const int n_wsum = pull_pairs ? 1 : (int)ng;
    const int pull_tb = pull_pairs ? (int)((ng - 1 + kWG - 1) / kWG) : 0;
    int pull_len = 0, pull_nsl = 0;
    if (pull_pairs) {
        const int want = std::max(1, kTargetWGs / (4 * pull_tb));
        pull_len = (int)std::max<size_t>(kMinSlice, (nd + want - 1) / want);
        pull_len = (pull_len + kTile - 1) / kTile * kTile;
        pull_nsl = (int)((nd + pull_len - 1) / pull_len);
    }
    // VJP source split over [0, n_grav)
    const int tblocks = (int)((ng + kWG - 1) / kWG);
    int vsl = std::max(1, std::min((kTargetWGs + tblocks - 1) / tblocks, (int)(ng / kMinSlice)));
    int vlen = (int)(((ng + vsl - 1) / vsl + kTile - 1) / kTile) * kTile;
    vsl = (int)((ng + vlen - 1) / vlen);
    DevBuf rbb, vbb, Vbm, Abm, Rbg, mug, gmu, gbeta, wpart, vpart, gR, gA, ppart;
    if (dalloc(c, rbb.p, 3 * na) || dalloc(c, vbb.p, 3 * na) ||
        dalloc(c, Vbm.p, 12 * ng) || dalloc(c, Abm.p, 12 * ng) || dalloc(c, Rbg.p, 12 * ng) ||
        dalloc(c, mug.p, ng) || dalloc(c, gmu.p, ng) || dalloc(c, gbeta.p, std::max<size_t>(nd, 1)) ||
        dalloc(c, wpart.p, (size_t)std::max(nwaves, 1) * 16 * (size_t)n_wsum) ||
        dalloc(c, vpart.p, (size_t)vsl * 4 * ng))
        return DDYN_ERR_NOMEM;
    if (pull_pairs && (dalloc(c, gR.p, 12 * nd) || dalloc(c, gA.p, 12 * nd) ||
                       dalloc(c, ppart.p, (size_t)pull_nsl * 16 * ng)))
        return DDYN_ERR_NOMEM;
    const unsigned bg = (unsigned)((ng + kWG - 1) / kWG);
    DDYN_HIP(c, hipMemcpyAsync(rbb.p, lam_r, 3 * na * sizeof(double), hipMemcpyHostToDevice, c->stream));
    DDYN_HIP(c, hipMemcpyAsync(vbb.p, lam_v, 3 * na * sizeof(double), hipMemcpyHostToDevice, c->stream));
    DDYN_HIP(c, hipMemsetAsync(gmu.p, 0, ng * sizeof(double), c->stream));
    DDYN_HIP(c, hipMemsetAsync(gbeta.p, 0, std::max<size_t>(nd, 1) * sizeof(double), c->stream));

    AdjArgs A{};
    A.n_grav = (int)ng; A.n_planets = c->prm.n_planets; A.n_dust = (int)nd; A.nslice = vsl;
    A.nwaves = nwaves; A.n_wsum = n_wsum; A.n_all = na;
    A.gR = gR.p; A.gA = gA.p; A.ppart = ppart.p;
    for (int s = 0; s < 4; ++s) A.p64[s] = c->p64[s];
    A.x = c->x; A.y = c->y; A.z = c->z; A.vx = c->vx; A.vy = c->vy; A.vz = c->vz; A.mu = c->mu;
    A.N = c->N; A.id = c->id;
    A.beta_p = c->use_beta_p ? c->beta_p : nullptr;
    A.betasw_p = (c->use_beta_p && c->have_betasw_p) ? c->betasw_p : nullptr;
    A.beta_t = c->beta_t; A.betasw_t = c->betasw_t; A.n_bins = c->n_bins; A.SWQ = c->prm.SWQ;
    A.dt = dt; A.eps2 = eps2_of(c); A.Rmin2 = c->prm.Rmin * c->prm.Rmin; A.vsw = c->prm.vsw;
    A.light = c->prm.light > 0.0 ? c->prm.light : kLightAU;
    A.rb = rbb.p; A.vb = vbb.p; A.Vbm = Vbm.p; A.Abm = Abm.p; A.Rbg = Rbg.p; A.mug = mug.p;
    A.wpart = wpart.p; A.part = vpart.p; A.g_mu = gmu.p; A.g_beta = gbeta.p;

    int rc;
    for (int64_t s = nseg - 1; s >= 0; --s) {
        const int64_t k0 = s * ck, k1 = std::min(k0 + ck, nsteps);
        if (s < nseg - 1 || !c->tape_last) {   // recompute the segment
            restore_state(c, c->tape_ck + (size_t)s * 7 * na);
            if ((rc = tape_segment(c, dt, k0, k1, s == nseg - 1))) return rc;
            c->tape_last = false;
        }
        for (int64_t k = k1 - 1; k >= k0; --k) {
            restore_state(c, c->tape_seg + (size_t)(k - k0) * 7 * na);
            for (int st = 1; st <= 3; ++st) A.p64[st] = c->tape_stg + ((size_t)(k - k0) * 3 + (size_t)(st - 1)) * ng;
            // star + planets and few other massive bodies: one launch for the massive part
            const bool small = ng <= (size_t)kWG && c->fuse_small;
            if (!small) k_adj_init<<<bg, kWG, 0, c->stream>>>(A);
            if (nd) {
                k_adj_dust<<<(unsigned)((nd + kWG - 1) / kWG), kWG, 0, c->stream>>>(A);
                k_adj_reduce<<<(unsigned)((4 * (size_t)n_wsum + kWG / 64 - 1) / (kWG / 64)), kWG, 0, c->stream>>>(A);
                if (pull_pairs) {
                    k_adj_pull<<<dim3((unsigned)pull_tb, (unsigned)pull_nsl, 4), kWG, 0, c->stream>>>(A, pull_len);
                    k_adj_pull_reduce<<<(unsigned)((ng - 1 + kWG / 64 - 1) / (kWG / 64)), kWG, 0, c->stream>>>(
                        A, pull_nsl);
                }
            }
            const int wg = nd ? 1 : 0;
            if (small) {
                k_adj_massive_small<<<1, kWG, 0, c->stream>>>(A, wg);
                DDYN_HIP(c, hipGetLastError());
                continue;
            }
            for (int st = 3; st >= 0; --st) {
                if (c->prm.precision == DDYN_FP32_FORCE) {
                    const double r64 = r64_of(c);
                    k_vjp_pairs32<<<dim3(bg, (unsigned)vsl), kWG, 0, c->stream>>>(
                        A.p64[st], Abm.p + (size_t)st * 3 * ng, (int)ng, c->prm.n_planets, vlen, eps2_of(c),
                        (float)(r64 * r64), vpart.p);
                } else {
                    k_vjp_pairs<<<dim3(bg, (unsigned)vsl), kWG, 0, c->stream>>>(
                        A.p64[st], Abm.p + (size_t)st * 3 * ng, (int)ng, c->prm.n_planets, vlen, eps2_of(c),
                        vpart.p);
                }
                switch (st) {
                    case 3: k_adj_mstage<3><<<bg, kWG, 0, c->stream>>>(A, wg); break;
                    case 2: k_adj_mstage<2><<<bg, kWG, 0, c->stream>>>(A, wg); break;
                    case 1: k_adj_mstage<1><<<bg, kWG, 0, c->stream>>>(A, wg); break;
                    default: k_adj_mstage<0><<<bg, kWG, 0, c->stream>>>(A, wg); break;
                }
            }
            DDYN_HIP(c, hipGetLastError());
        }
    }
    restore_state(c, c->tape_fin);
    DDYN_HIP(c, hipGetLastError());
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    DDYN_HIP(c, hipMemcpy(g_r, rbb.p, 3 * na * sizeof(double), hipMemcpyDeviceToHost));
    DDYN_HIP(c, hipMemcpy(g_v, vbb.p, 3 * na * sizeof(double), hipMemcpyDeviceToHost));
    DDYN_HIP(c, hipMemcpy(g_mu, gmu.p, ng * sizeof(double), hipMemcpyDeviceToHost));
    if (g_beta && nd) DDYN_HIP(c, hipMemcpy(g_beta, gbeta.p, nd * sizeof(double), hipMemcpyDeviceToHost));
    return DDYN_OK;
}

int check_adjoint_args(ddyn_ctx* c, double dt, int64_t nsteps)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!(dt > 0.0) || nsteps < 0) return fail(c, DDYN_ERR_ARG, "need dt > 0 and nsteps >= 0");
    if (c->prm.FLQ == 1) return fail(c, DDYN_ERR_ARG, "the adjoint does not support the Lorentz force (FLQ)");
    return DDYN_OK;
}

}  // namespace

int ddyn_adjoint(ddyn_ctx* c, double dt, int64_t nsteps, int64_t ckpt_every, const double* lam_r,
                 const double* lam_v, double* g_r, double* g_v, double* g_mu, double* g_beta)
{
    int rc = check_adjoint_args(c, dt, nsteps);
    if (rc) return rc;
    if (!lam_r || !lam_v || !g_r || !g_v || !g_mu) return fail(c, DDYN_ERR_ARG, "adjoint arrays must not be NULL");
    DDYN_HIP(c, hipSetDevice(c->device));
    if ((rc = record_tape(c, dt, nsteps, ckpt_every))) return rc;
    rc = reverse_tape(c, lam_r, lam_v, g_r, g_v, g_mu, g_beta);
    drop_tape(c);
    return rc;
}

int ddyn_adjoint_record(ddyn_ctx* c, double dt, int64_t nsteps, int64_t ckpt_every)
{
    int rc = check_adjoint_args(c, dt, nsteps);
    if (rc) return rc;
    DDYN_HIP(c, hipSetDevice(c->device));
    return record_tape(c, dt, nsteps, ckpt_every);
}

int ddyn_adjoint_reverse(ddyn_ctx* c, const double* lam_r, const double* lam_v, double* g_r, double* g_v,
                         double* g_mu, double* g_beta)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!c->have_tape)
        return fail(c, DDYN_ERR_STATE, "no adjoint tape: call ddyn_adjoint_record after the last change of state");
    if (!lam_r || !lam_v || !g_r || !g_v || !g_mu) return fail(c, DDYN_ERR_ARG, "adjoint arrays must not be NULL");
    DDYN_HIP(c, hipSetDevice(c->device));
    return reverse_tape(c, lam_r, lam_v, g_r, g_v, g_mu, g_beta);
}

int ddyn_energy(ddyn_ctx* c, double* out7)
{
    int rc = check_state(c);
    if (rc) return rc;
    if (!out7) return fail(c, DDYN_ERR_ARG, "out7 must not be NULL");
    DDYN_HIP(c, hipSetDevice(c->device));
    const size_t ng = (size_t)c->n_grav;
    k_potential<<<(unsigned)((ng + kWG - 1) / kWG), kWG, 0, c->stream>>>(c->p64[0], (int)ng,
                                                                        c->prm.n_planets, eps2_of(c), c->acc);
    DDYN_HIP(c, hipGetLastError());
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<double> h(7 * ng);
    DDYN_HIP(c, hipMemcpy(h.data(), c->acc, ng * sizeof(double), hipMemcpyDeviceToHost));
    const double* src[] = {c->x, c->y, c->z, c->vx, c->vy, c->vz};
    for (int a = 0; a < 6; ++a)
        DDYN_HIP(c, hipMemcpy(h.data() + (a + 1) * ng, src[a], ng * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> m(ng);
    DDYN_HIP(c, hipMemcpy(m.data(), c->mu, ng * sizeof(double), hipMemcpyDeviceToHost));
    long double K = 0, U = 0, L[3] = {0, 0, 0}, P[3] = {0, 0, 0};
    const double *phi = h.data(), *X = phi + ng, *Y = X + ng, *Z = Y + ng, *VX = Z + ng, *VY = VX + ng,
                 *VZ = VY + ng;
    for (size_t i = 0; i < ng; ++i) {
        K += 0.5L * m[i] * ((long double)VX[i] * VX[i] + (long double)VY[i] * VY[i] + (long double)VZ[i] * VZ[i]);
        U += 0.5L * m[i] * phi[i];
        L[0] += (long double)m[i] * (Y[i] * VZ[i] - Z[i] * VY[i]);
        L[1] += (long double)m[i] * (Z[i] * VX[i] - X[i] * VZ[i]);
        L[2] += (long double)m[i] * (X[i] * VY[i] - Y[i] * VX[i]);
        P[0] += (long double)m[i] * VX[i]; P[1] += (long double)m[i] * VY[i]; P[2] += (long double)m[i] * VZ[i];
    }
    out7[0] = (double)(K + U);
    for (int a = 0; a < 3; ++a) { out7[1 + a] = (double)L[a]; out7[4 + a] = (double)P[a]; }
    return DDYN_OK;
}

int ddyn_synchronize(ddyn_ctx* c)
{
    if (!c) return DDYN_ERR_ARG;
    DDYN_HIP(c, hipSetDevice(c->device));
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    // the grain streams are joined by the main stream at the end of ddyn_step;
    // a step that stopped on an error may have left work on them
    DDYN_HIP(c, hipStreamSynchronize(c->gstream));
    DDYN_HIP(c, hipStreamSynchronize(c->gstream2));
    return DDYN_OK;
}

int ddyn_set_timing(ddyn_ctx* c, int enable)
{
    if (!c) return DDYN_ERR_ARG;
    DDYN_HIP(c, hipSetDevice(c->device));
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    drain_events(c);
    for (int k = 0; k < DDYN_KCLASS_COUNT; ++k) { c->t_launch[k] = 0; c->t_ms[k] = 0.0; }
    c->timing = enable != 0;
    return DDYN_OK;
}

int ddyn_pair_plan(ddyn_ctx* c, int32_t* nslice, int32_t* slice_len)
{
    const int rc = check_state(c);
    if (rc) return rc;
    if (nslice) *nslice = c->nslice;
    if (slice_len) *slice_len = c->slice_len;
    return DDYN_OK;
}

int ddyn_eval_rsqrt_fp64(ddyn_ctx* c, int64_t n, const double* x, double* y)
{
    if (!c) return DDYN_ERR_ARG;
    if (n < 0 || n > ((int64_t)1 << 30) || (n > 0 && (!x || !y)))
        return fail(c, DDYN_ERR_ARG, "need 0 <= n <= 2^30 and non-NULL arrays");
    if (n == 0) return DDYN_OK;
    DDYN_HIP(c, hipSetDevice(c->device));
    DevBuf dx, dy;
    if (dalloc(c, dx.p, (size_t)n) || dalloc(c, dy.p, (size_t)n)) return DDYN_ERR_NOMEM;
    DDYN_HIP(c, hipMemcpyAsync(dx.p, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    k_rsqrt_fp64<<<(unsigned)((n + kWG - 1) / kWG), kWG, 0, c->stream>>>(dx.p, dy.p, n);
    DDYN_HIP(c, hipGetLastError());
    DDYN_HIP(c, hipMemcpyAsync(y, dy.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    return DDYN_OK;
}

int ddyn_read_timing(ddyn_ctx* c, int32_t kclass, int64_t* launches, double* total_ms)
{
    if (!c) return DDYN_ERR_ARG;
    if (kclass < 0 || kclass >= DDYN_KCLASS_COUNT) return fail(c, DDYN_ERR_ARG, "unknown kernel class");
    DDYN_HIP(c, hipSetDevice(c->device));
    DDYN_HIP(c, hipStreamSynchronize(c->stream));
    drain_events(c);
    if (launches) *launches = c->t_launch[kclass];
    if (total_ms) *total_ms = c->t_ms[kclass];
    return DDYN_OK;
}

}  // extern "C"
