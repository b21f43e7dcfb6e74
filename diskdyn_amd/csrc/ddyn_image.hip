// ddyn_image.hip -- images and SEDs of the grains on gfx950: thermal emission
// and scattered starlight of every grain (src/gpu_functions.cu:945-1156,
// driven by Grid::null_flux, calc_flux, calc_flux_sed and copy_sed,
// src/grid.cu:12-65).
//
// Per grain (one lane each) the flux is the reference's: the grain's
// temperature from its bin's table at its distance (log-log between the
// table's radii), pi B_lambda(T) s^2 Qabs / d^2 of thermal emission, and the
// stellar flux times pi s^2 Qsca P(cos) / R^2 of scattered light, with the
// phase function P interpolated log-linearly in the scattering cosine; times
// the grain's tracer count N, in Jy.
//
// The reference adds each grain into its pixel (and each block's SED sum into
// the total) with CAS atomics, so its images depend on scheduling.  Here:
//   image  k_img_flux writes one (pixel, flux) pair per grain; a stable radix
//          sort by pixel (rocPRIM) keeps each pixel's grains in grain order;
//          k_img_sum adds every pixel's run left to right.  The image is the
//          sum in grain order, pixel by pixel, deterministic.
//   SED    k_sed computes every wavelength of a grain in one lane (geometry
//          and temperature once), sums each wave's 64 grains with a fixed
//          butterfly and writes one partial per (wave, wavelength);
//          k_sed_reduce adds the partials of each wavelength in a fixed order.
// All three passes read each grain's x, y, z, N and id once (36 B per grain)
// plus the optical tables, which the L2 and the last-level cache hold.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "ddyn.h"
#include "ddyn_ctx.h"

struct ImageData {
    int32_t n_bins = 0, n_wav = 0, n_theta = 0, n_r = 0;
    double Rmin = 0.0, Rmax = 0.0;
    double *w = nullptr, *in = nullptr, *size = nullptr, *Qabs = nullptr, *Qsca = nullptr, *Qpf = nullptr,
           *Td = nullptr;
    // derived on the device at upload (k_optics_prep): per wavelength lambda^5 and
    // lambda k_B, and the logarithm of every phase-function entry
    double *w5 = nullptr, *wk = nullptr, *lQp = nullptr;
    // scratch, grown as needed
    size_t cap_dust = 0;
    uint32_t *key = nullptr, *key_s = nullptr;
    double *val = nullptr, *val_s = nullptr;
    void* sort_tmp = nullptr;
    size_t sort_bytes = 0;
    size_t cap_pix = 0;
    double* pix = nullptr;
    size_t cap_part = 0;
    double* part = nullptr;
    size_t cap_sed = 0;
    double* sed = nullptr;
};

namespace {

// Constants of src/ddyn.cuh:16-38.
constexpr double kAU = 1.49598e11;
constexpr double kCspeed = 299792458.0;
constexpr double kBoltz = 1.3806503E-23;
constexpr double kPlanck = 6.626068E-34;
constexpr double kPI = 3.14159265358979323846;
constexpr int kMaxBins = 2700;
constexpr int kImgWG = 256;

struct ImgArgs {
    int64_t n_grav, n_dust;
    const double *x, *y, *z, *N;
    const int32_t* id;
    const double *w, *in, *size, *Qabs, *Qsca, *lQp, *Td;
    const double *w5, *wk;
    int32_t n_bins, n_wav, n_theta, n_r;
    double Rmin, Rmax;
    double dist, rot, incl, PA, fov;
    int32_t xsize, ysize;
};

// Temperature of a grain of bin id at R [AU] (src/gpu_functions.cu:1050-1064):
// the table's end values outside [Rmin, Rmax], else log-log between the two
// radii around R.  Rounding at R just below Rmax cannot step past the table.
__device__ double grain_temp(const ImgArgs& A, int id, double R)
{
    const double f = log(A.Rmax / A.Rmin) / ((double)A.n_r - 1.0);
    const double* Td = A.Td + (int64_t)id * A.n_r;
    if (R <= A.Rmin) return Td[0];
    if (R >= A.Rmax) return Td[A.n_r - 1];
    const int Rid = min((int)(log(R / A.Rmin) / f) + 1, A.n_r - 1);
    const double Rin = A.Rmin * exp((Rid - 1) * f);
    const double Rout = A.Rmin * exp(Rid * f);
    return exp(log(Td[Rid - 1]) + log(Td[Rid] / Td[Rid - 1]) * log(R / Rin) / log(Rout / Rin));
}

// Flux [Jy] of one grain (tracer count n, bin id, distance R [AU], scattering
// cosine cs, temperature T) at wavelength j (src/gpu_functions.cu:1065-1080).
// The wavelength's lambda^5 and lambda k_B and the logarithms of the phase
// function come from tables made at upload with the same operations, so the
// results are those of forming them here.
__device__ double grain_flux(const ImgArgs& A, int id, double n, double R, double cs, double T, int j)
{
    const double wvl = A.w[j];
    const double I = (2 * kPlanck * kCspeed * kCspeed) /
                     (A.w5[j] * (exp(kPlanck * kCspeed / (A.wk[j] * T)) - 1.0)) * kPI;   // pi B_lambda
    // (the clamp only matters for a non-finite cosine, from a non-finite position)
    const int cosinm = (int)fmin(fmax(floor(0.5 * (cs + 1.0) * (A.n_theta - 1)), 0.0), (double)(A.n_theta - 2));
    const int cosinp = cosinm + 1;
    const double cosm = (double)(cosinm)*2.0 / (A.n_theta - 1) - 1.0;
    const double cosp = (double)(cosinp)*2.0 / (A.n_theta - 1) - 1.0;
    const double* lQp = A.lQp + (int64_t)A.n_wav * ((int64_t)A.n_theta * id + cosinm) + j;
    const double lm = lQp[0], lp = lQp[A.n_wav];
    const double P = exp(lm + (lp - lm) * (cs - cosm) / (cosp - cosm));
    const double s = A.size[id];
    const int64_t q = (int64_t)A.n_wav * id + j;
    const double Fscat = A.in[j] * kPI * s * s * A.Qsca[q] * P / R / R / kAU / kAU;
    const double Fther = I * s * s * A.Qabs[q] / A.dist / A.dist;
    return n * (Fther + Fscat) * 1e26 * wvl * wvl / kCspeed;
}

// Cosine the phase function is read at (src/gpu_functions.cu:1010-1023): the
// direction to the star (-x, -y, -z) against the direction to the observer,
// who sits at (0, -dist sin i, dist cos i) about the star, clamped to
// [-1, 1 - 2/n_theta].  (dist is in m and the position in AU, as in the
// reference: the observer is in effect at infinity.)
__device__ double scatter_cos(const ImgArgs& A, double x, double y, double z, double R)
{
    const double ly = -A.dist * sin(A.incl) - y;
    const double lz = A.dist * cos(A.incl) - z;
    double c = (x * x + (-y) * ly + (-z) * lz) / R / sqrt(x * x + ly * ly + lz * lz);
    if (c < -1) c = -1.0;
    if (c >= 1.0 - 2.0 / A.n_theta) c = 1.0 - 2.0 / A.n_theta;
    return c;
}

// One (pixel, flux) pair per grain at wavelength wj (calc_flux_driver,
// src/gpu_functions.cu:983-1081); grains without a bin in the tables, within
// Rmin or outside the field get the key npix, which sorts after every pixel.
__global__ __launch_bounds__(kImgWG) void k_img_flux(ImgArgs A, int wj, uint32_t npix, uint32_t* key, double* val)
{
    const int64_t i = (int64_t)blockIdx.x * kImgWG + threadIdx.x;
    if (i >= A.n_dust) return;
    uint32_t k = npix;
    double v = 0.0;
    const int64_t t = A.n_grav + i;
    const int id = A.id ? A.id[t] : -1;
    if (id >= 0 && id < A.n_bins) {
        double x = A.x[t], y = A.y[t];
        const double z = A.z[t], n = A.N[t];
        double phi = atan2(y, x);
        const double R = sqrt(x * x + y * y + z * z);
        double Rd = sqrt(x * x + y * y);
        // turn the system about z by the longitude
        x = Rd * cos(phi + A.rot);
        y = Rd * sin(phi + A.rot);
        const double cs = scatter_cos(A, x, y, z, R);
        // incline it, then turn the sky plane by the position angle
        y = sqrt(y * y + z * z) * cos(atan2(z, y) - A.incl);
        phi = atan2(y, x);
        Rd = sqrt(x * x + y * y);
        x = Rd * cos(phi + A.PA);
        y = Rd * sin(phi + A.PA);
        const double psca = A.xsize >= A.ysize ? A.fov / A.xsize : A.fov / A.ysize;
        x = x + 0.5 * psca * A.xsize;
        y = y + 0.5 * psca * A.ysize;
        const double fx = floor(x / psca), fy = floor(y / psca);
        const bool inside = fx >= 0.0 && fx <= (double)(A.xsize - 1) && fy >= 0.0 && fy <= (double)(A.ysize - 1);
        if (inside && R > A.Rmin) {
            const double T = grain_temp(A, id, R);
            v = grain_flux(A, id, n, R, cs, T, wj);
            k = (uint32_t)((int64_t)A.ysize * (int64_t)fx + (int64_t)fy);
        }
    }
    key[i] = k;
    val[i] = v;
}

__global__ void k_img_zero(double* F, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kImgWG + threadIdx.x;
    if (i < n) F[i] = 0.0;
}

// Each pixel's run of the sorted pairs, added left to right (grain order) by
// the thread at its head.
__global__ __launch_bounds__(kImgWG) void k_img_sum(const uint32_t* key, const double* val, int64_t n,
                                                   uint32_t npix, double* F)
{
    const int64_t i = (int64_t)blockIdx.x * kImgWG + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = key[i];
    if (k >= npix || (i > 0 && key[i - 1] == k)) return;
    double s = val[i];
    for (int64_t j = i + 1; j < n && key[j] == k; ++j) s += val[j];
    F[k] = s;
}

// Every wavelength of one grain per lane (calc_flux_sed_driver,
// src/gpu_functions.cu:1083-1156: the system is turned about z with the 3-D
// distance R, as there); each wave's 64 fluxes summed by a fixed butterfly,
// one partial per (wave, wavelength).
__global__ __launch_bounds__(kImgWG) void k_sed(ImgArgs A, double* part)
{
    const int64_t i = (int64_t)blockIdx.x * kImgWG + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool emit = false;
    double R = 0.0, cs = 0.0, T = 0.0, n = 0.0;
    int id = -1;
    if (i < A.n_dust) {
        const int64_t t = A.n_grav + i;
        id = A.id ? A.id[t] : -1;
        if (id >= 0 && id < A.n_bins) {
            double x = A.x[t], y = A.y[t];
            const double z = A.z[t];
            n = A.N[t];
            const double phi = atan2(y, x);
            R = sqrt(x * x + y * y + z * z);
            x = R * cos(phi + A.rot);
            y = R * sin(phi + A.rot);
            cs = scatter_cos(A, x, y, z, R);
            emit = R > A.Rmin;
            if (emit) T = grain_temp(A, id, R);
        }
    }
    double* out = part + ((int64_t)blockIdx.x * (kImgWG / 64) + wave) * A.n_wav;
    for (int j = 0; j < A.n_wav; ++j) {
        double f = emit ? grain_flux(A, id, n, R, cs, T, j) : 0.0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) f += __shfl_xor(f, o);
        if (lane == 0) out[j] = f;
    }
}

// Total of wavelength blockIdx.x over the nq wave partials: each thread adds
// every 256th partial in order, then a fixed tree over the threads.
__global__ __launch_bounds__(kImgWG) void k_sed_reduce(const double* part, int64_t nq, int n_wav, double* sed)
{
    __shared__ double sh[kImgWG];
    const int j = blockIdx.x, t = threadIdx.x;
    double s = 0.0;
    for (int64_t q = t; q < nq; q += kImgWG) s += part[q * n_wav + j];
    sh[t] = s;
    __syncthreads();
    for (int o = kImgWG / 2; o > 0; o >>= 1) {
        if (t < o) sh[t] += sh[t + o];
        __syncthreads();
    }
    if (t == 0) sed[j] = sh[0];
}

// Tables derived from the uploaded ones: pow(lambda, 5) and lambda k_B per
// wavelength (the first factors of the Planck function as the reference forms
// it, (lambda k_B) T), and log of every phase-function entry.
__global__ void k_optics_prep(const double* w, int n_wav, const double* Qpf, int64_t n_pf, double* w5, double* wk,
                              double* lQp)
{
    const int64_t i = (int64_t)blockIdx.x * kImgWG + threadIdx.x;
    if (i < n_wav) {
        w5[i] = pow(w[i], 5);
        wk[i] = w[i] * kBoltz;
    }
    if (i < n_pf) lQp[i] = log(Qpf[i]);
}

int fail(ddyn_ctx* c, int code, const std::string& msg)
{
    c->err = msg;
    return code;
}

int hipfail(ddyn_ctx* c, hipError_t e, const char* where)
{
    return fail(c, DDYN_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define IMG_HIP(c, call)                                           \
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

void free_tables(ImageData* d)
{
    dfree(d->w); dfree(d->in); dfree(d->size); dfree(d->Qabs); dfree(d->Qsca); dfree(d->Qpf); dfree(d->Td);
    dfree(d->w5); dfree(d->wk); dfree(d->lQp);
}

// Grows a scratch array to at least n elements (contents not kept).
template <class T>
hipError_t grow(T*& p, size_t& cap, size_t n)
{
    if (n <= cap && p) return hipSuccess;
    dfree(p);
    cap = 0;
    const hipError_t e = hipMalloc((void**)&p, (n ? n : 1) * sizeof(T));
    if (e == hipSuccess) cap = n;
    return e;
}

// Checks common to both entry points and the kernel arguments of the current state.
int prepare(ddyn_ctx* c, const ddyn_view* v, bool image, ImgArgs& A)
{
    if (!c->x) return fail(c, DDYN_ERR_STATE, "ddyn_upload_state has not been called");
    if (!c->img || !c->img->w) return fail(c, DDYN_ERR_STATE, "ddyn_upload_optics has not been called");
    if (!v) return fail(c, DDYN_ERR_ARG, "view must not be NULL");
    if (!(v->dist > 0.0)) return fail(c, DDYN_ERR_ARG, "dist must be > 0");
    if (image) {
        if (!(v->fov > 0.0)) return fail(c, DDYN_ERR_ARG, "fov must be > 0");
        if (v->xsize < 1 || v->ysize < 1 || (int64_t)v->xsize * v->ysize > ((int64_t)1 << 30))
            return fail(c, DDYN_ERR_ARG, "xsize and ysize must be >= 1 with xsize * ysize <= 2^30");
    }
    const ImageData* d = c->img;
    A.n_grav = c->n_grav;
    A.n_dust = c->n_dust;
    A.x = c->x; A.y = c->y; A.z = c->z; A.N = c->N; A.id = c->id;
    A.w = d->w; A.in = d->in; A.size = d->size; A.Qabs = d->Qabs; A.Qsca = d->Qsca; A.lQp = d->lQp; A.Td = d->Td;
    A.w5 = d->w5; A.wk = d->wk;
    A.n_bins = d->n_bins; A.n_wav = d->n_wav; A.n_theta = d->n_theta; A.n_r = d->n_r;
    A.Rmin = d->Rmin; A.Rmax = d->Rmax;
    A.dist = v->dist; A.rot = v->rot; A.incl = v->incl; A.PA = v->PA; A.fov = v->fov;
    A.xsize = v->xsize; A.ysize = v->ysize;
    return DDYN_OK;
}

}  // namespace

void ddyn_image_release(ddyn_ctx* c)
{
    if (!c || !c->img) return;
    ImageData* d = c->img;
    free_tables(d);
    dfree(d->key); dfree(d->key_s); dfree(d->val); dfree(d->val_s);
    if (d->sort_tmp) (void)hipFree(d->sort_tmp);
    dfree(d->pix); dfree(d->part); dfree(d->sed);
    delete d;
    c->img = nullptr;
}

extern "C" {

int ddyn_upload_optics(ddyn_ctx* c, const ddyn_optics_tables* t)
{
    if (!c) return DDYN_ERR_ARG;
    if (!t) return fail(c, DDYN_ERR_ARG, "tables must not be NULL");
    if (t->n_bins < 1 || t->n_bins > kMaxBins) return fail(c, DDYN_ERR_ARG, "n_bins must be in [1, 2700] (nsize)");
    if (t->n_wav < 2 || t->n_theta < 2 || t->n_r < 2)
        return fail(c, DDYN_ERR_ARG, "n_wav, n_theta and n_r must be >= 2");
    if ((int64_t)t->n_bins * t->n_theta * t->n_wav > ((int64_t)1 << 31))
        return fail(c, DDYN_ERR_ARG, "n_bins * n_theta * n_wav must be <= 2^31");
    if (!t->w || !t->in || !t->size || !t->Qabs || !t->Qsca || !t->Qpfunc || !t->Td)
        return fail(c, DDYN_ERR_ARG, "table pointers must not be NULL");
    if (!(t->Rmin > 0.0) || !(t->Rmax > t->Rmin)) return fail(c, DDYN_ERR_ARG, "need 0 < Rmin < Rmax");
    IMG_HIP(c, hipSetDevice(c->device));
    IMG_HIP(c, hipStreamSynchronize(c->stream));
    if (!c->img) c->img = new ImageData();
    ImageData* d = c->img;
    free_tables(d);
    const size_t nb = (size_t)t->n_bins, nw = (size_t)t->n_wav, nt = (size_t)t->n_theta, nr = (size_t)t->n_r;
    struct { double** dst; const double* src; size_t n; } cp[] = {
        {&d->w, t->w, nw}, {&d->in, t->in, nw}, {&d->size, t->size, nb}, {&d->Qabs, t->Qabs, nb * nw},
        {&d->Qsca, t->Qsca, nb * nw}, {&d->Qpf, t->Qpfunc, nb * nt * nw}, {&d->Td, t->Td, nb * nr},
        {&d->w5, nullptr, nw}, {&d->wk, nullptr, nw}, {&d->lQp, nullptr, nb * nt * nw}};
    for (auto& e : cp) {
        if (hipMalloc((void**)e.dst, e.n * sizeof(double)) != hipSuccess) {
            *e.dst = nullptr;
            free_tables(d);
            return fail(c, DDYN_ERR_NOMEM, "optical tables: device allocation failed");
        }
        if (e.src) IMG_HIP(c, hipMemcpy(*e.dst, e.src, e.n * sizeof(double), hipMemcpyHostToDevice));
    }
    const int64_t npf = (int64_t)(nb * nt * nw);
    const int64_t nprep = npf > (int64_t)nw ? npf : (int64_t)nw;
    k_optics_prep<<<(unsigned)((nprep + kImgWG - 1) / kImgWG), kImgWG, 0, c->stream>>>(d->w, t->n_wav, d->Qpf, npf,
                                                                                     d->w5, d->wk, d->lQp);
    IMG_HIP(c, hipGetLastError());
    IMG_HIP(c, hipStreamSynchronize(c->stream));
    d->n_bins = t->n_bins; d->n_wav = t->n_wav; d->n_theta = t->n_theta; d->n_r = t->n_r;
    d->Rmin = t->Rmin; d->Rmax = t->Rmax;
    return DDYN_OK;
}

int ddyn_calc_flux(ddyn_ctx* c, const ddyn_view* v, int32_t wj, double* F)
{
    if (!c) return DDYN_ERR_ARG;
    ImgArgs A;
    if (const int rc = prepare(c, v, true, A)) return rc;
    if (wj < 0 || wj >= A.n_wav) return fail(c, DDYN_ERR_ARG, "wj must be in [0, n_wav)");
    if (!F) return fail(c, DDYN_ERR_ARG, "F must not be NULL");
    IMG_HIP(c, hipSetDevice(c->device));
    ImageData* d = c->img;
    const hipStream_t st = c->stream;
    const int64_t nd = A.n_dust;
    const uint32_t npix = (uint32_t)((int64_t)A.xsize * A.ysize);
    IMG_HIP(c, grow(d->pix, d->cap_pix, npix));
    k_img_zero<<<(unsigned)((npix + kImgWG - 1) / kImgWG), kImgWG, 0, st>>>(d->pix, npix);
    IMG_HIP(c, hipGetLastError());
    if (nd > 0) {
        if ((size_t)nd > d->cap_dust) {
            dfree(d->key); dfree(d->key_s); dfree(d->val); dfree(d->val_s);
            d->cap_dust = 0;
            IMG_HIP(c, hipMalloc((void**)&d->key, nd * sizeof(uint32_t)));
            IMG_HIP(c, hipMalloc((void**)&d->key_s, nd * sizeof(uint32_t)));
            IMG_HIP(c, hipMalloc((void**)&d->val, nd * sizeof(double)));
            IMG_HIP(c, hipMalloc((void**)&d->val_s, nd * sizeof(double)));
            d->cap_dust = (size_t)nd;
        }
        const unsigned blocks = (unsigned)((nd + kImgWG - 1) / kImgWG);
        k_img_flux<<<blocks, kImgWG, 0, st>>>(A, wj, npix, d->key, d->val);
        IMG_HIP(c, hipGetLastError());
        // keys run from 0 to npix (the mark of grains that add nothing)
        const unsigned end_bit = 32u - (unsigned)__builtin_clz(npix);
        size_t bytes = 0;
        IMG_HIP(c, rocprim::radix_sort_pairs(nullptr, bytes, d->key, d->key_s, d->val, d->val_s, (size_t)nd, 0u,
                                             end_bit, st));
        if (bytes > d->sort_bytes) {
            if (d->sort_tmp) (void)hipFree(d->sort_tmp);
            d->sort_tmp = nullptr;
            d->sort_bytes = 0;
            IMG_HIP(c, hipMalloc(&d->sort_tmp, bytes));
            d->sort_bytes = bytes;
        }
        IMG_HIP(c, rocprim::radix_sort_pairs(d->sort_tmp, bytes, d->key, d->key_s, d->val, d->val_s, (size_t)nd, 0u,
                                             end_bit, st));
        k_img_sum<<<blocks, kImgWG, 0, st>>>(d->key_s, d->val_s, nd, npix, d->pix);
        IMG_HIP(c, hipGetLastError());
    }
    IMG_HIP(c, hipMemcpyAsync(F, d->pix, (size_t)npix * sizeof(double), hipMemcpyDeviceToHost, st));
    IMG_HIP(c, hipStreamSynchronize(st));
    return DDYN_OK;
}

int ddyn_calc_sed(ddyn_ctx* c, const ddyn_view* v, double* Fexc)
{
    if (!c) return DDYN_ERR_ARG;
    ImgArgs A;
    if (const int rc = prepare(c, v, false, A)) return rc;
    if (!Fexc) return fail(c, DDYN_ERR_ARG, "Fexc must not be NULL");
    IMG_HIP(c, hipSetDevice(c->device));
    ImageData* d = c->img;
    const hipStream_t st = c->stream;
    const int64_t nd = A.n_dust;
    const int nw = A.n_wav;
    IMG_HIP(c, grow(d->sed, d->cap_sed, (size_t)nw));
    const unsigned blocks = (unsigned)((nd + kImgWG - 1) / kImgWG);
    const int64_t nq = (int64_t)blocks * (kImgWG / 64);   // wave partials
    if (nd > 0) {
        IMG_HIP(c, grow(d->part, d->cap_part, (size_t)nq * nw));
        k_sed<<<blocks, kImgWG, 0, st>>>(A, d->part);
        IMG_HIP(c, hipGetLastError());
    }
    k_sed_reduce<<<nw, kImgWG, 0, st>>>(d->part, nq, nw, d->sed);
    IMG_HIP(c, hipGetLastError());
    IMG_HIP(c, hipMemcpyAsync(Fexc, d->sed, (size_t)nw * sizeof(double), hipMemcpyDeviceToHost, st));
    IMG_HIP(c, hipStreamSynchronize(st));
    return DDYN_OK;
}

}  // extern "C"
