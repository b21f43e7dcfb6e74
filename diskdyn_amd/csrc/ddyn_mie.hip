// ddyn_mie.hip -- Mie efficiencies and phase functions of homogeneous spheres
// on gfx950: the tables the reference's tools/runmie writes for every grain
// size and wavelength of a refractive-index (.lnk) table
// (tools/mie_single.c:54-241).
//
// One wave per (size, wavelength) problem.  The series over n is sequential:
// every lane runs the logarithmic derivative D_n(m x) down from n_max (lane
// n mod 64 keeps D_n in a per-problem scratch area) and then the upward
// recurrences of psi_n, chi_n, the coefficients a_n, b_n and the efficiency
// sums, all wave-uniform; the lanes hold the scattering angles (up to AP per
// lane) and update the angular functions pi_n, tau_n and the amplitude sums
// S+ = S1 + S2, S- = S1 - S2 of their own angles.  D_n comes back to every
// lane from the lane that stored it (each lane reads only what it wrote, then
// a cross-lane broadcast).  The arithmetic is FP64 with the reference's
// expressions in its order; the angular functions are real (the reference
// carries imaginary parts that stay exactly zero and add only zeros).
//
// Problems are sorted by their number of terms, largest first, and launched
// in chunks whose D_n scratch fits a fixed budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <string>
#include <vector>

#include "ddyn.h"
#include "ddyn_ctx.h"

namespace {

constexpr double kMiePI = 3.141592653589793238;
constexpr double kMaxX = 7200000.0;                  // largest size parameter (tools/mie_single.c:86)
constexpr size_t kScratchBudget = (size_t)2 << 30;   // bytes of D_n scratch per launch

struct MieArgs {
    int32_t n_wav, n_theta;
    const double *s, *lam, *nre, *kim, *cost;
    const int32_t* prob;     // problem p = size * n_wav + wavelength, one per wave
    const int64_t* doff;     // start of each problem's D_n scratch (entries)
    double2* D;
    double *Qabs, *Qsca, *Qpf;
};

// Number of terms of the series and the start of the D_n recurrence
// (tools/mie_single.c:105-118).
__host__ __device__ inline double mie_nstop(double x)
{
    if (x < 0.02) return 2.0;
    if (x <= 8.0) return x + 4.0 * pow(x, 0.3333333333333333) + 2.0;
    if (x < 4200.0) return x + 4.05 * pow(x, 0.3333333333333333) + 2.0;
    return x + 4.0 * pow(x, 0.3333333333333333) + 2.0;
}

__host__ __device__ inline int64_t mie_nmax(double x, double nstop, double Yr, double Yi)
{
    double t = sqrt(Yr * Yr + Yi * Yi);
    t = (nstop > t) ? nstop : t;
    t += 15.0;
    return (int64_t)floor(t);
}

__device__ inline double bcast(double v, int src)
{
    return __shfl(v, src, 64);
}

template <int AP>
__global__ __launch_bounds__(64) void k_mie(MieArgs A)
{
    const int lane = threadIdx.x;
    const int p = A.prob[blockIdx.x];
    const int is = p / A.n_wav, iw = p % A.n_wav;
    const double x = 2.0 * kMiePI * A.s[is] / A.lam[iw];
    const double cr = A.nre[iw], ci = -A.kim[iw];
    const double Irr = cr / (cr * cr + ci * ci), Iri = -ci / (cr * cr + ci * ci);
    const double Yr = x * cr, Yi = x * ci;
    const double nstop = mie_nstop(x);
    const int64_t nmax = mie_nmax(x, nstop, Yr, Yi);
    double2* D = A.D + A.doff[blockIdx.x];

    // D_n, n = nmax-1 .. 1, from D_nmax = 0 (tools/mie_single.c:124-129)
    {
        double dr = 0.0, di = 0.0;
        for (int64_t n = nmax - 1; n > 0; --n) {
            const double tr = (((double)(n + 1)) * Yr) / (Yr * Yr + Yi * Yi);
            const double ti = (-((double)(n + 1)) * Yi) / (Yr * Yr + Yi * Yi);
            const double ur = tr + dr, ui = ti + di;
            dr = tr - (ur / (ur * ur + ui * ui));
            di = ti - (-ui / (ur * ur + ui * ui));
            if (lane == (int)(n & 63)) D[n] = make_double2(dr, di);
        }
    }

    double c[AP], Spr[AP], Spi[AP], Smr[AP], Smi[AP], P0[AP], P1[AP];
#pragma unroll
    for (int a = 0; a < AP; ++a) {
        const int k = lane + 64 * a;
        c[a] = k < A.n_theta ? A.cost[k] : 0.0;
        Spr[a] = Spi[a] = Smr[a] = Smi[a] = P0[a] = 0.0;
        P1[a] = 1.0;
    }
    double Psi0 = cos(x), Psi1 = sin(x), Chi0 = -Psi1, Chi1 = Psi0;
    double Dqsc = 0.0, Dqxt = 0.0, Tnp1 = 1.0;
    double2 dv = make_double2(0.0, 0.0);
    for (int64_t n = 1; n <= nstop; ++n) {
        // every 64 terms each lane fetches the D_n it stored itself
        if ((n & 63) == 0 || n == 1) {
            const int64_t m = (n & ~(int64_t)63) + lane;
            dv = (m > 0 && m < nmax) ? D[m] : make_double2(0.0, 0.0);
        }
        const double dnr = bcast(dv.x, (int)(n & 63)), dni = bcast(dv.y, (int)(n & 63));
        Tnp1 += 2.0;
        const double Tnm1 = Tnp1 - 2.0;
        const double DN = (double)n;
        const double A2 = Tnp1 / (DN * (DN + 1.0));
        const double Turbo = (DN + 1.0) / DN;
        const double Rnx = n / x;
        const double Psi = Tnm1 * (Psi1 / x) - Psi0;
        const double Chi = Tnm1 * (Chi1 / x) - Chi0;
        double nr, ni, qr, qi, q2r, q2i, den;
        // a_n with D_n / m (tools/mie_single.c:162-173)
        nr = ((dnr * Irr - dni * Iri) + Rnx) * Psi - Psi1;
        ni = (dnr * Iri + dni * Irr) * Psi;
        qr = ((dnr * Irr - dni * Iri) + Rnx);
        qi = (dnr * Iri + dni * Irr);
        q2r = qr * Psi - qi * Chi;
        q2i = qr * Chi + qi * Psi;
        qr = q2r - Psi1;
        qi = q2i - Chi1;
        den = 1. / (qr * qr + qi * qi);
        const double Ar = (nr * qr + ni * qi) * den;
        const double Ai = (ni * qr - nr * qi) * den;
        // b_n with m D_n (:176-187)
        nr = ((dnr * cr - dni * ci) + Rnx) * Psi - Psi1;
        ni = (dnr * ci + dni * cr) * Psi;
        qr = ((dnr * cr - dni * ci) + Rnx);
        qi = (dnr * ci + dni * cr);
        q2r = qr * Psi - qi * Chi;
        q2i = qr * Chi + qi * Psi;
        qr = q2r - Psi1;
        qi = q2i - Chi1;
        den = 1. / (qr * qr + qi * qi);
        const double Br = (nr * qr + ni * qi) * den;
        const double Bi = (ni * qr - nr * qi) * den;
        Dqxt = Tnp1 * (Ar + Br) + Dqxt;
        Dqsc = Tnp1 * (Ar * Ar + Ai * Ai + Br * Br + Bi * Bi) + Dqsc;
        const double aAr = A2 * Ar, aAi = A2 * Ai, aBr = A2 * Br, aBi = A2 * Bi;
#pragma unroll
        for (int a = 0; a < AP; ++a) {   // (:201-219), the angular functions real
            const double S = c[a] * P1[a];
            const double T = S - P0[a];
            const double Tau = n * T - P0[a];
            Spr[a] = (aAr + aBr) * (P1[a] + Tau) + Spr[a];
            Spi[a] = (aAi + aBi) * (P1[a] + Tau) + Spi[a];
            Smr[a] = (aAr - aBr) * (P1[a] - Tau) + Smr[a];
            Smi[a] = (aAi - aBi) * (P1[a] - Tau) + Smi[a];
            P0[a] = P1[a];
            P1[a] = S + Turbo * T;
        }
        Psi0 = Psi1;
        Psi1 = Psi;
        Chi0 = Chi1;
        Chi1 = Chi;
    }
    // phase function normalised to 1 over the sphere, the efficiencies (:232-247)
#pragma unroll
    for (int a = 0; a < AP; ++a) {
        const int k = lane + 64 * a;
        if (k < A.n_theta) {
            const double X1r = (Spr[a] + Smr[a]) * 0.5, X1i = (Spi[a] + Smi[a]) * 0.5;
            const double X2r = (Spr[a] - Smr[a]) * 0.5, X2i = (Spi[a] - Smi[a]) * 0.5;
            A.Qpf[((int64_t)is * A.n_theta + k) * A.n_wav + iw] =
                ((X1r * X1r) + (X1i * X1i) + (X2r * X2r) + (X2i * X2i)) / (4 * kMiePI * Dqsc);
        }
    }
    if (lane == 0) {
        Dqsc = 2.0 * Dqsc / (x * x);
        Dqxt = 2.0 * Dqxt / (x * x);
        A.Qabs[(int64_t)is * A.n_wav + iw] = Dqxt - Dqsc;
        A.Qsca[(int64_t)is * A.n_wav + iw] = Dqsc;
    }
}

int fail(ddyn_ctx* c, int code, const std::string& msg)
{
    c->err = msg;
    return code;
}

// Device buffers of one call, freed on every return.
struct Bufs {
    std::vector<void*> p;
    template <class T>
    hipError_t get(T** out, size_t n)
    {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(T));
        if (e == hipSuccess) p.push_back(q);
        *out = (T*)q;
        return e;
    }
    ~Bufs()
    {
        for (void* q : p) (void)hipFree(q);
    }
};

}  // namespace

extern "C" int ddyn_mie(ddyn_ctx* c, int32_t n_sizes, const double* s_um, int32_t n_wav, const double* lam_um,
                        const double* n_re, const double* k_im, int32_t n_theta, const double* cost, double* Qabs,
                        double* Qsca, double* Qpfunc)
{
    if (!c) return DDYN_ERR_ARG;
    if (n_sizes < 1 || n_wav < 1 || n_theta < 1 || n_theta > 256 || (int64_t)n_sizes * n_wav > (1 << 30))
        return fail(c, DDYN_ERR_ARG, "need n_sizes, n_wav >= 1, 1 <= n_theta <= 256, n_sizes * n_wav <= 2^30");
    if (!s_um || !lam_um || !n_re || !k_im || !cost || !Qabs || !Qsca || !Qpfunc)
        return fail(c, DDYN_ERR_ARG, "arrays must not be NULL");
    // each problem's terms and scratch, with the formulas the kernel uses (one
    // entry of margin where host and device rounding could differ)
    const int64_t np = (int64_t)n_sizes * n_wav;
    std::vector<int64_t> need((size_t)np);
    std::vector<double> work((size_t)np);
    for (int64_t p = 0; p < np; ++p) {
        const int is = (int)(p / n_wav), iw = (int)(p % n_wav);
        if (!(s_um[is] > 0.0) || !(lam_um[iw] > 0.0))
            return fail(c, DDYN_ERR_ARG, "sizes and wavelengths must be > 0");
        if (!(n_re[iw] > 0.0) || !(k_im[iw] >= 0.0) || !std::isfinite(n_re[iw]) || !std::isfinite(k_im[iw]))
            return fail(c, DDYN_ERR_ARG, "refractive indices need a finite n > 0 and k >= 0");
        const double x = 2.0 * kMiePI * s_um[is] / lam_um[iw];
        if (!(x <= kMaxX)) return fail(c, DDYN_ERR_ARG, "size parameter 2 pi s / lambda beyond 7.2e6");
        const double nstop = mie_nstop(x);
        const int64_t nmax = mie_nmax(x, nstop, x * n_re[iw], -x * k_im[iw]);
        need[(size_t)p] = nmax + 2;
        work[(size_t)p] = nstop + (double)nmax;
    }
    std::vector<int32_t> order((size_t)np);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return work[(size_t)a] > work[(size_t)b]; });
    const size_t nt = (size_t)n_theta;
    Bufs B;
    double *ds, *dl, *dn, *dk, *dc, *dqa, *dqs, *dqp;
    int32_t* dprob;
    int64_t* doff;
    double2* dD;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, DDYN_ERR_HIP, "hipSetDevice failed");
    size_t max_chunk = 0, max_need = 0;
    for (int64_t p = 0; p < np; ++p) max_need = std::max(max_need, (size_t)need[(size_t)p]);
    const size_t budget = std::max(kScratchBudget / sizeof(double2), max_need);
    // chunks of consecutive problems (in the sorted order) within the budget
    std::vector<size_t> cut{0};
    {
        size_t acc = 0;
        for (size_t q = 0; q < (size_t)np; ++q) {
            const size_t e = (size_t)need[(size_t)order[q]];
            if (acc + e > budget && q > cut.back()) {
                cut.push_back(q);
                max_chunk = std::max(max_chunk, acc);
                acc = 0;
            }
            acc += e;
        }
        max_chunk = std::max(max_chunk, acc);
        cut.push_back((size_t)np);
    }
    if (B.get(&ds, (size_t)n_sizes) || B.get(&dl, (size_t)n_wav) || B.get(&dn, (size_t)n_wav) ||
        B.get(&dk, (size_t)n_wav) || B.get(&dc, nt) || B.get(&dqa, (size_t)np) || B.get(&dqs, (size_t)np) ||
        B.get(&dqp, (size_t)np * nt) || B.get(&dprob, (size_t)np) || B.get(&doff, (size_t)np) ||
        B.get(&dD, max_chunk))
        return fail(c, DDYN_ERR_NOMEM, "ddyn_mie: device allocation failed");
    std::vector<int64_t> off((size_t)np);
    for (size_t k = 0; k + 1 < cut.size(); ++k) {
        int64_t a = 0;
        for (size_t q = cut[k]; q < cut[k + 1]; ++q) {
            off[q] = a;
            a += need[(size_t)order[q]];
        }
    }
    const hipStream_t st = c->stream;
    struct { void* d; const void* h; size_t bytes; } up[] = {
        {ds, s_um, (size_t)n_sizes * 8}, {dl, lam_um, (size_t)n_wav * 8}, {dn, n_re, (size_t)n_wav * 8},
        {dk, k_im, (size_t)n_wav * 8}, {dc, cost, nt * 8}, {dprob, order.data(), (size_t)np * 4},
        {doff, off.data(), (size_t)np * 8}};
    for (auto& u : up)
        if (hipMemcpyAsync(u.d, u.h, u.bytes, hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(c, DDYN_ERR_HIP, "ddyn_mie: copy to the device failed");
    for (size_t k = 0; k + 1 < cut.size(); ++k) {
        MieArgs A;
        A.n_wav = n_wav; A.n_theta = n_theta;
        A.s = ds; A.lam = dl; A.nre = dn; A.kim = dk; A.cost = dc;
        A.prob = dprob + cut[k];
        A.doff = doff + cut[k];
        A.D = dD;
        A.Qabs = dqa; A.Qsca = dqs; A.Qpf = dqp;
        const unsigned blocks = (unsigned)(cut[k + 1] - cut[k]);
        if (n_theta <= 64) k_mie<1><<<blocks, 64, 0, st>>>(A);
        else if (n_theta <= 128) k_mie<2><<<blocks, 64, 0, st>>>(A);
        else k_mie<4><<<blocks, 64, 0, st>>>(A);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(c, DDYN_ERR_HIP, std::string("k_mie: ") + hipGetErrorString(e));
    }
    if (hipMemcpyAsync(Qabs, dqa, (size_t)np * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(Qsca, dqs, (size_t)np * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(Qpfunc, dqp, (size_t)np * nt * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
        return fail(c, DDYN_ERR_HIP, "ddyn_mie: copy to the host failed");
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(c, DDYN_ERR_HIP, std::string("ddyn_mie: ") + hipGetErrorString(e));
    return DDYN_OK;
}
