// ddyn_setup.cpp -- what a .param run needs before its first step: the optical
// data of the grains and the star, the dust size grid, the per-bin beta,
// beta_sw and charge tables with the sublimation radius, and the particles of
// the disks and blobs (see include/ddyn_host.h).  Host C++ only.
//
// The reference does this in main and Grid::setup (src/ddyn.cu:78-96,
// src/grid.cu:62-189) with add_optical_data and add_specfile
// (src/read_in.cu:377-506), detminmax, confCdisks and confCblobs
// (src/read_in.cu:236-342), interpolate_dust_to_grid,
// interpolate_stellar_to_grid and calculate_beta (src/functions.cu:243-388),
// calculate_dust_temp (src/grid.cu:754-792) and getbod, genorb_disk and
// genorb_blob (src/functions.cu:157-240).  The file formats, units and the
// formulas are the reference's, and so is the random stream (Subtractive55).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "ddyn_host.h"
#include "ddyn_host_impl.h"

struct ddyn_optics {
    int n_wav = 0, n_theta = 0, n_sizes = 0;
    double dmin = 0.0, dmax = 0.0;           // optical size range [m] (the table header)
    std::vector<double> w;                   // wavelengths [m], n_wav
    std::vector<double> s;                   // grain sizes [m], n_sizes
    std::vector<double> Qabs, Qsca;          // [size][wav]
    std::vector<double> Qpf;                 // phase function [size][theta][wav]
    std::vector<double> spec_w, spec_F;      // stellar spectrum: wavelength [m], flux at the observer
};

// The dust grid's optical tables for images and SEDs (ddyn_grid_optics_make).
struct ddyn_grid_optics {
    int n_bins = 0, n_wav = 0, n_theta = 0;
    std::vector<double> w, in, size, Qabs, Qsca, Qpf, Td;   // as ddyn_optics_tables describes them
    double Rmin = 0.0, Rmax = 0.0;                          // radial grid of Td [AU]
};

namespace {

// Constants of src/ddyn.cuh:16-38.
constexpr double kAU = 1.49598e11;
constexpr double kMSun = 1.98892e30;
constexpr double kYear = 3.1556926e7;
constexpr double kPI = 3.14159265358979323846;
constexpr double kPI2 = 6.28318530717958623200;
constexpr double kGrav = 4.0 * kPI * kPI;
constexpr double kC = 299792458.0;
constexpr double kBoltz = 1.3806503E-23;
constexpr double kPlanck = 6.626068E-34;
constexpr double kTsub = 2000.0;         // sublimation temperature [K]
constexpr double kMgrid = 1.1;           // mass ratio of neighbouring dust bins
constexpr double kCd = 2.0;              // drag coefficient factor
constexpr double kLsunW = 3.827e26;      // solar luminosity [W] (src/functions.cu:355)
constexpr int kMaxBins = 2700;           // size bins the device tables hold (nsize)

// Error numbers of the reference's message table (src/error.cu:5-64).
constexpr int kErrFile = 39;       // a file could not be opened
constexpr int kErrOptics = 42;     // optical data not well formed
constexpr int kErrInterp = 54;     // interpolation / grid error
constexpr int kErrWav = 43;        // an image wavelength outside the optical data
constexpr int kNR = 100;           // radii of the dust temperature table (nR, src/ddyn.cuh:34)

bool read_lines(const char* path, std::vector<std::string>& lines)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::string cur;
    int ch;
    while ((ch = std::fgetc(f)) != EOF) {
        if (ch == '\n') { lines.push_back(cur); cur.clear(); }
        else cur.push_back((char)ch);
    }
    if (!cur.empty()) lines.push_back(cur);
    std::fclose(f);
    return true;
}

bool to_double(const std::string& t, double& v)
{
    char* end = nullptr;
    v = std::strtod(t.c_str(), &end);
    return end != t.c_str();
}

// The composition table (src/read_in.cu:419-506): a first line that is
// skipped; a definition line whose tokens 1-5 are n_wav, n_theta, the number
// of sizes and the size range [um]; then per size a line "# <size [um]>",
// n_wav lines "<wavelength [um]> <Qabs> <Qsca>" (wavelengths taken from the
// first size) and the phase function, n_theta x n_wav values ten per line,
// theta-major.
int read_composition(const char* path, ddyn_optics& o)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) return kErrFile;
    if (lines.size() < 2) return kErrOptics;
    {
        const auto d = ddyn_impl::tokenize(lines[1].c_str(), 12);
        double v[5];
        if (d.size() < 6) return kErrOptics;
        for (int k = 0; k < 5; ++k)
            if (!to_double(d[k + 1], v[k])) return kErrOptics;
        o.n_wav = (int)v[0]; o.n_theta = (int)v[1]; o.n_sizes = (int)v[2];
        o.dmin = v[3] * 1e-6; o.dmax = v[4] * 1e-6;
    }
    if (o.n_wav < 2 || o.n_theta < 2 || o.n_sizes < 2 || o.n_wav > (1 << 20) || o.n_theta > (1 << 16) ||
        o.n_sizes > (1 << 20) || (size_t)o.n_sizes * o.n_wav * o.n_theta > ((size_t)1 << 31))
        return kErrOptics;
    const size_t nw = (size_t)o.n_wav, nt = (size_t)o.n_theta, ns = (size_t)o.n_sizes;
    o.w.assign(nw, 0.0); o.s.assign(ns, 0.0);
    o.Qabs.assign(ns * nw, 0.0); o.Qsca.assign(ns * nw, 0.0); o.Qpf.assign(ns * nt * nw, 0.0);
    size_t j = 0, iw = 0, ip = 0;   // sizes seen, wavelength rows and phase values of the current size
    std::vector<size_t> rows(ns, 0), pfs(ns, 0);
    for (size_t i = 2; i < lines.size(); ++i) {
        const std::string& ln = lines[i];
        const auto d = ddyn_impl::tokenize(ln.c_str(), 12);
        if (!ln.empty() && ln[0] == '#') {
            if (j >= ns || d.size() < 2) return kErrOptics;
            double v;
            if (!to_double(d[1], v)) return kErrOptics;
            o.s[j++] = v * 1e-6;
            iw = ip = 0;
            continue;
        }
        if (j == 0) continue;
        if (d.size() == 3) {
            double v[3];
            for (int k = 0; k < 3; ++k)
                if (!to_double(d[k], v[k])) return kErrOptics;
            if (iw >= nw) return kErrOptics;
            if (j == 1) o.w[iw] = v[0] * 1e-6;
            o.Qabs[(j - 1) * nw + iw] = v[1];
            o.Qsca[(j - 1) * nw + iw] = v[2];
            rows[j - 1] = ++iw;
        } else if (d.size() == 10) {
            if (ip + 10 > nt * nw) return kErrOptics;
            for (int k = 0; k < 10; ++k) {
                double v;
                if (!to_double(d[k], v)) return kErrOptics;
                o.Qpf[(j - 1) * nt * nw + ip + k] = v;
            }
            ip += 10;
            pfs[j - 1] = ip;
        }
    }
    if (j != ns) return kErrOptics;
    for (size_t k = 0; k < ns; ++k)
        if (rows[k] != nw || pfs[k] != nt * nw) return kErrOptics;
    for (size_t k = 1; k < ns; ++k)
        if (!(o.s[k] > o.s[k - 1])) return kErrOptics;
    for (size_t k = 1; k < nw; ++k)
        if (!(o.w[k] > o.w[k - 1])) return kErrOptics;
    return 0;
}

// The stellar spectrum (src/read_in.cu:377-417): lines "<wavelength [m]>
// <surface flux>" ('#' lines skipped), the flux scaled to the observer at
// distance cfg->dist by (R_star / dist)^2.
int read_spectrum(const char* path, const ddyn_config& c, ddyn_optics& o)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) return kErrFile;
    const double scale = std::pow(c.rstar * kAU, 2) / std::pow(c.dist, 2);
    for (const auto& ln : lines) {
        if (!ln.empty() && ln[0] == '#') continue;
        const auto d = ddyn_impl::tokenize(ln.c_str(), 2);
        double w, F;
        if (d.size() < 2 || !to_double(d[0], w) || !to_double(d[1], F)) continue;
        o.spec_w.push_back(w);
        o.spec_F.push_back(F * scale);
    }
    if (o.spec_w.size() < 2) return kErrOptics;
    for (size_t k = 1; k < o.spec_w.size(); ++k)
        if (!(o.spec_w[k] > o.spec_w[k - 1])) return kErrOptics;
    return 0;
}

// Index ip of the first table size >= a, clamped so that [ip-1, ip] is an
// interval of the table (the reference indexes below the table when a equals
// its smallest size, src/functions.cu:285-286; here the first interval is
// used, which gives the first row exactly there).
size_t upper_index(const std::vector<double>& s, double a)
{
    size_t ip = 0;
    while (ip < s.size() && s[ip] < a) ++ip;
    if (ip == 0) ip = 1;
    if (ip >= s.size()) ip = s.size() - 1;
    return ip;
}

// log-log interpolation between (x0, y0) and (x1, y1) at x, as the reference
// writes it: exp(log y0 + log(x/x0) log(y1/y0) / log(x1/x0)).  That form is
// undefined where an efficiency is zero (a grain that does not scatter): equal
// values are returned as they are, and a zero at one end is interpolated
// linearly in log x.
double loglog(double x, double x0, double x1, double y0, double y1)
{
    if (y0 == y1) return y0;
    const double t = std::log(x / x0) / std::log(x1 / x0);
    if (!(y0 > 0.0) || !(y1 > 0.0)) return y0 + t * (y1 - y0);
    return std::exp(std::log(y0) + t * std::log(y1 / y0));
}

// Planck B_lambda (src/functions.cu:390-397).
double planck(double T, double wvl)
{
    return (2 * kPlanck * kC * kC) / (std::pow(wvl, 5) * (std::exp(kPlanck * kC / (wvl * kBoltz * T)) - 1));
}

// The optical data on the dust grid's sizes and the stellar flux on the optical
// wavelengths (interpolate_dust_to_grid and interpolate_stellar_to_grid,
// src/functions.cu:243-327): sizes a_i = amin exp(i f_a), efficiencies and the
// phase function log-log in size between the table's sizes; the spectrum
// log-log in wavelength, zero below its first wavelength and a Rayleigh-Jeans
// tail beyond its last.
struct GridTables {
    std::vector<double> a, Qa, Qs, Qp, in;   // [bin], [bin][wav], [bin][wav], [bin][theta][wav], [wav]
};

int grid_tables(const ddyn_config* c, const ddyn_optics* o, GridTables& g)
{
    const int n = c->n_grid;
    if (n < 2 || n > kMaxBins || !(c->amax > c->amin) || !(c->amin > 0.0)) return kErrInterp;
    const size_t nw = (size_t)o->n_wav, nt = (size_t)o->n_theta;
    const double fa = std::log(c->amax / c->amin) / ((double)n - 1.0);
    g.a.assign((size_t)n, 0.0);
    g.Qa.assign((size_t)n * nw, 0.0);
    g.Qs.assign((size_t)n * nw, 0.0);
    g.Qp.assign((size_t)n * nt * nw, 0.0);
    for (int i = 0; i < n; ++i) {
        double ai = c->amin * std::exp((double)i * fa);
        if (i == n - 1 && (ai - o->dmax) > 1e-12) return kErrInterp;
        if (i == n - 1 && (ai - o->dmax) < 1e-12 && (ai - o->dmax) > 0) ai = o->dmax;
        g.a[(size_t)i] = ai;
        const size_t ip = upper_index(o->s, ai), im = ip - 1;
        const double s0 = o->s[im], s1 = o->s[ip];
        for (size_t j = 0; j < nw; ++j) {
            g.Qa[i * nw + j] = loglog(ai, s0, s1, o->Qabs[im * nw + j], o->Qabs[ip * nw + j]);
            g.Qs[i * nw + j] = loglog(ai, s0, s1, o->Qsca[im * nw + j], o->Qsca[ip * nw + j]);
            for (size_t k = 0; k < nt; ++k)
                g.Qp[nw * (nt * i + k) + j] =
                    loglog(ai, s0, s1, o->Qpf[nw * (nt * im + k) + j], o->Qpf[nw * (nt * ip + k) + j]);
        }
    }
    g.in.assign(nw, 0.0);
    const size_t nsp = o->spec_w.size();
    for (size_t j = 0; j < nw; ++j) {
        const double wj = o->w[j];
        if (wj < o->spec_w[0]) {
            g.in[j] = 0.0;
        } else if (wj > o->spec_w[nsp - 1]) {
            g.in[j] = o->spec_F[nsp - 1] * std::pow(o->spec_w[nsp - 1] / wj, 4.0);   // Rayleigh-Jeans tail
        } else {
            size_t ip = 0;
            while (ip < nsp - 1 && o->spec_w[ip] <= wj) ++ip;
            if (ip == 0) ip = 1;
            const size_t im = ip - 1;
            const double F0 = o->spec_F[im], F1 = o->spec_F[ip];
            g.in[j] = (F0 != 0 && F1 != 0) ? loglog(wj, o->spec_w[im], o->spec_w[ip], F0, F1) : 0.0;
        }
    }
    return 0;
}

// The sublimation radius [m] (calculate_dust_temp, src/grid.cu:766-783): the
// smallest over the bins of the distance at which a grain reaches T_sub,
// 0.5 sqrt(absorbed / emitted), starting from Rmax_m.
double sublimation_radius(const ddyn_config* c, const ddyn_optics* o, const GridTables& g, double Rmax_m)
{
    const size_t nw = (size_t)o->n_wav;
    double Rmin = Rmax_m;
    for (int k = 0; k < c->n_grid; ++k) {
        double sk = 0.0, gf = 0.0;
        for (size_t i = 1; i < nw; ++i) {
            sk += 0.5 * (o->w[i] - o->w[i - 1]) * (g.Qa[k * nw + i - 1] * g.in[i - 1] + g.Qa[k * nw + i] * g.in[i]) /
                  kPI;
            gf += 0.5 * (o->w[i] - o->w[i - 1]) *
                  (g.Qa[k * nw + i - 1] * planck(kTsub, o->w[i - 1]) + g.Qa[k * nw + i] * planck(kTsub, o->w[i]));
        }
        sk *= std::pow(c->dist, 2.0);
        const double Rt = 0.5 * std::sqrt(sk / gf);
        if (Rt < Rmin) Rmin = Rt;
    }
    return Rmin;
}

// The random stream of particle generation: Knuth's subtractive generator
// (The Art of Computer Programming vol. 2, 2nd ed., 3.2.2 and the IN55/IRN55
// routines of 3.6), X[n] = X[n-55] - X[n-24] mod 10^9, in the parameters the
// reference draws with (ran3, src/ran3.cu:12-48), so that a .param file gives
// the reference's particles:
//  - the 55 starting words: t0 = (161803398 - |seed|) mod 10^9 (C's
//    truncating remainder), t1 = 1, t[n+1] = t[n-1] - t[n] (+10^9 if
//    negative), word t[n] in slot 21 n mod 55 (slot 55 for n = 0);
//  - the recurrence updates one slot per draw, slot i (1..55, cyclic) minus
//    slot i + 31, which holds X[n-24];
//  - 220 draws (four passes over the slots) are discarded, and a deviate is
//    X times the double nearest 10^-9, in [0, 1).
// The reference seeds with -2 (src/grid.cu:80) and keeps one stream for the
// whole setup, in the order of the draws below.
class Subtractive55 {
public:
    explicit Subtractive55(int64_t seed)
    {
        const int64_t mag = seed < 0 ? -seed : seed;
        int64_t prev = (kSeedWord - mag) % kM, cur = 1;
        slot_[55] = prev;
        for (int n = 1; n <= 54; ++n) {
            slot_[(21 * n) % 55] = cur;
            int64_t nxt = prev - cur;
            if (nxt < 0) nxt += kM;
            prev = cur;
            cur = nxt;
        }
        for (int k = 0; k < 4 * 55; ++k) next();
    }
    int64_t next()
    {
        i_ = i_ == 55 ? 1 : i_ + 1;
        const int p = i_ + 31 > 55 ? i_ - 24 : i_ + 31;
        int64_t v = slot_[i_] - slot_[p];
        if (v < 0) v += kM;
        slot_[i_] = v;
        return v;
    }
    double uniform() { return (double)next() * (1.0 / (double)kM); }
    // Box-Muller, sqrt(-2 ln u1) cos(2 pi u2), u1 drawn first: the order in
    // which the reference's host compiler evaluates the two draws of
    // sqrt(-2.0*log(ran3(&seed)))*cos(PI2*ran3(&seed)) (left operand first).
    double normal()
    {
        const double u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(kPI2 * u2);
    }

private:
    static constexpr int64_t kM = 1000000000;
    static constexpr int64_t kSeedWord = 161803398;
    int64_t slot_[56] = {};   // slots 1..55
    int i_ = 0;
};
struct Sink {
    int64_t n = 0, cap;
    double *x, *y, *z, *vx, *vy, *vz, *N, *mu;
    int32_t* id;
    bool add(const double* r, double Nv, double m, int32_t idv)
    {
        if (n >= cap) return false;
        x[n] = r[0]; y[n] = r[1]; z[n] = r[2]; vx[n] = r[3]; vy[n] = r[4]; vz[n] = r[5];
        N[n] = Nv; mu[n] = m; id[n] = idv;
        ++n;
        return true;
    }
};

double bin_beta(const float* beta, int32_t n_bins, int32_t id)
{
    double b = (id >= 0 && id < n_bins && beta) ? (double)beta[id] : 0.0;
    return b >= 0.5 ? 0.0 : b;   // such grains leave at once: launched on a Kepler orbit (functions.cu:185)
}

// Draws of a rejection loop before generation gives up (a law whose accepted
// region has a probability below about 1e-5 per draw is a bad input).
constexpr int kMaxDraws = 1 << 20;

// genorb_disk (src/functions.cu:157-191): e = |normal| sigma_e (< 1), a with
// a^(2+gamma) uniform between Rin/(1-e) and Rout/(1+e) (surface density ~
// r^gamma), i = |normal| atan(dh/dr), node, periapsis and true anomaly uniform.
// Returns false if no e below 1 was drawn.
bool orbit_disk(const ddyn_config& c, const ddyn_disk& D, double m, double beta, Subtractive55& g, double* r)
{
    double e = 1.0;
    for (int k = 0; k < kMaxDraws && !(e < 1.0); ++k) e = std::fabs(g.normal()) * D.sigmae;
    if (!(e < 1.0)) return false;
    const double amin = D.Rin / (1.0 - e), amax = D.Rout / (1.0 + e);
    const double u = g.uniform();
    double a;
    if (D.gammas == -2.0) {   // the power law's limit: a log-uniform (the reference divides by zero)
        a = amin * std::exp(u * std::log(amax / amin));
    } else {
        const double p = 2.0 + D.gammas;
        a = std::pow(u * (std::pow(amax, p) - std::pow(amin, p)) + std::pow(amin, p), 1.0 / p);
    }
    const double i = std::fabs(g.normal()) * D.dhdr;
    const double nu = kPI2 * g.uniform(), w = kPI2 * g.uniform(), o = kPI2 * g.uniform();
    ddyn_impl::coe2rv(c.pm[0], m, a, e, i, w, o, nu, beta, r);
    return true;
}

// genorb_blob (src/functions.cu:194-232): a = a_blob + normal sigma (> 0), the
// blob's e, i, w, Omega, true anomaly nu_blob + normal sigma / a; for a
// collision cloud a velocity kick of vkepp % of the orbital speed in an
// isotropic direction.  Returns false if no positive a was drawn.
bool orbit_blob(const ddyn_config& c, const ddyn_blob& B, double m, double beta, Subtractive55& g, double* r)
{
    double a = 0.0;
    for (int k = 0; k < kMaxDraws && !(a > 0.0); ++k) a = B.a + g.normal() * B.sigma;
    if (!(a > 0.0)) return false;
    const double nu = B.nu + g.normal() * B.sigma / a;
    ddyn_impl::coe2rv(c.pm[0], m, a, B.ecc, B.inc, B.w, B.o, nu, beta, r);
    if (B.collcloud == 0) {
        const double voff = B.vkepp * 0.01 * std::sqrt(r[3] * r[3] + r[4] * r[4] + r[5] * r[5]);
        const double ct = -1.0 + 2.0 * g.uniform();
        const double st = std::sqrt(1.0 - ct * ct);
        const double gam = kPI2 * g.uniform();
        double sg, cg;
        ::sincos(gam, &sg, &cg);   // one call, as coe2rv's angles
        r[3] += voff * st * cg;
        r[4] += voff * st * sg;
        r[5] += voff * ct;
    }
    return true;
}

// getbod (src/functions.cu:235-240): the next body of the mass ladder.  The
// interval [mmin, mmax] below the current mmax holds one body of
// dN/dm = C m^-eta (ome = 1 - eta); its mass is the interval's mass.  eta = 1
// and eta = 2 take the logarithmic forms of the integrals (the reference
// divides by zero there).  Returns mu = G M of the body.  The precisions are
// the reference's: ome, ome / C, 1 / ome and 1 + ome in double, the powers and
// the mass in long double, mu rounded to double at the end.
double next_body(long double& mmax, double C, double eta)
{
    using LD = long double;
    const double ome = 1.0 - eta;
    LD mmin, mass;
    if (eta == 1.0) mmin = mmax * std::exp(-1.0L / (LD)C);
    else mmin = std::pow(std::pow(mmax, (LD)ome) - (LD)(ome / C), (LD)(1.0 / ome));
    if (eta == 2.0) mass = (LD)C * std::log(mmax / mmin);
    else mass = (LD)C * (std::pow(mmax, (LD)(1.0 + ome)) - std::pow(mmin, (LD)(1.0 + ome))) / (LD)(1.0 + ome);
    mmax = mmin;
    return (double)(mass / (LD)kMSun * (LD)kGrav);
}

// The dust mass bins of one disk or blob (src/grid.cu:136-189): its mass
// range snapped to the global grid, Ngrid bins of ratio exp(f), Npg tracers
// per bin.
struct DustBins {
    double mmin = 0.0;
    int ngrid = 0, npg = 0;
};

DustBins dust_bins(const ddyn_config& c, double dmmax, double mmin_body, int64_t nsmall)
{
    DustBins b;
    if (nsmall <= 0 || !(c.mmin > 0.0) || !(c.f > 0.0)) return b;
    const double mx = c.mmin * std::exp(std::round(std::log(dmmax / c.mmin) / c.f) * c.f);
    b.mmin = c.mmin * std::exp(std::round(std::log(mmin_body / c.mmin) / c.f) * c.f);
    b.ngrid = (int)std::round(std::log(mx / b.mmin) / std::log(kMgrid) + 1);
    if (b.ngrid < 1) { b.ngrid = 0; return b; }
    b.npg = (int)std::round((double)nsmall / (double)b.ngrid);
    return b;
}

// Tracer weight of mass bin [in, ax] of dN/dm = C m^-eta (src/grid.cu:156-158).
double bin_number(double C, double eta, double in, double ax)
{
    if (eta == 0) return C * (ax - in);
    if (eta == 1) return C * std::log(ax / in);
    return C * (std::pow(ax, 1.0 - eta) - std::pow(in, 1.0 - eta)) / (1.0 - eta);
}

template <class F>
int dust_of(const ddyn_config& c, const ddyn_optics* o, const DustBins& b, double C, double eta, const float* beta,
            int32_t n_bins, Subtractive55& g, Sink& out, F orbit)
{
    const double dmin = o ? o->dmin : 0.0, dmax = o ? o->dmax : std::numeric_limits<double>::infinity();
    for (int l = 0; l < b.ngrid; ++l) {
        double m = b.mmin * std::exp((double)l * c.f);
        const double size = std::pow(m * 3.0 / 4.0 / kPI / c.rho, 1.0 / 3.0);
        int32_t id = (int32_t)std::round(std::log(m / c.mmin) / c.f);
        if (size < dmin) id = -1;
        if (size - dmax > 1e-9) id = -1;
        const double in = b.mmin * std::exp(((double)l - 0.5) * c.f);
        const double ax = b.mmin * std::exp(((double)l + 0.5) * c.f);
        m *= kGrav / kMSun;
        const double N = bin_number(C, eta, in, ax);
        const double bt = bin_beta(beta, n_bins, id);
        for (int k = 0; k < b.npg; ++k) {
            double r[6];
            if (!orbit(m, bt, g, r) || !out.add(r, N / b.npg, m, id)) return kErrInterp;
        }
    }
    return 0;
}

double norm_c(double Mtot, double eta, double mmax, double mmin)
{
    if (eta == 1) return Mtot / (mmax - mmin);
    if (eta == 2) return Mtot / std::log(mmax / mmin);
    return Mtot / ((std::pow(mmax, 2.0 - eta) - std::pow(mmin, 2.0 - eta)) / (2.0 - eta));
}

}  // namespace

extern "C" {

int ddyn_optics_read(const char* comp_path, const char* spec_path, const ddyn_config* cfg, ddyn_optics** out)
{
    if (!out) return kErrOptics;
    *out = nullptr;
    if (!comp_path || !spec_path || !cfg) return kErrFile;
    ddyn_optics* o = new ddyn_optics();
    int rc = read_composition(comp_path, *o);
    if (!rc) rc = read_spectrum(spec_path, *cfg, *o);
    if (rc) { delete o; return rc; }
    *out = o;
    return 0;
}

void ddyn_optics_free(ddyn_optics* o) { delete o; }

int ddyn_optics_info(const ddyn_optics* o, int32_t* n_wav, int32_t* n_theta, int32_t* n_sizes, double* dmin,
                     double* dmax)
{
    if (!o) return kErrOptics;
    if (n_wav) *n_wav = o->n_wav;
    if (n_theta) *n_theta = o->n_theta;
    if (n_sizes) *n_sizes = o->n_sizes;
    if (dmin) *dmin = o->dmin;
    if (dmax) *dmax = o->dmax;
    return 0;
}

int ddyn_set_dust_grid(ddyn_config* c, const ddyn_optics* o)
{
    if (!c || !o) return kErrOptics;
    const double kappa = 4.0 * kPI * c->rho / 3.0;
    for (int i = 0; i < c->n_disks; ++i) {
        ddyn_disk& D = c->disks[i];
        if (D.amin < o->dmin) { D.amin = o->dmin; D.mmin = kappa * std::pow(D.amin, 3.0); }
        D.dmax = D.amax > o->dmax ? o->dmax : D.amax;
        D.dmmax = kappa * std::pow(D.dmax, 3.0);
    }
    for (int i = 0; i < c->n_blobs; ++i) {
        ddyn_blob& B = c->blobs[i];
        if (B.amin < o->dmin) { B.amin = o->dmin; B.mmin = kappa * std::pow(B.amin, 3.0); }
        B.dmax = B.amax > o->dmax ? o->dmax : B.amax;
        B.dmmax = kappa * std::pow(B.dmax, 3.0);
    }
    double amin, amax;
    if (c->n_disks > 0) { amin = c->disks[0].amin; amax = c->disks[0].dmax; }
    else if (c->n_blobs > 0) { amin = c->blobs[0].amin; amax = c->blobs[0].dmax; }
    else { amin = o->dmin; amax = o->dmax; }
    for (int i = 1; i < c->n_disks; ++i) {
        amin = std::fmin(amin, c->disks[i].amin);
        amax = std::fmax(amax, c->disks[i].dmax);
    }
    for (int i = 0; i < c->n_blobs; ++i) {
        amin = std::fmin(amin, c->blobs[i].amin);
        amax = std::fmax(amax, c->blobs[i].dmax);
    }
    c->amin = amin;
    c->amax = amax;
    c->mmin = kappa * std::pow(amin, 3.0);
    c->mmax = kappa * std::pow(amax, 3.0);
    c->n_grid = (int)(std::log(c->mmax / c->mmin) / std::log(kMgrid) + 1.0);
    if (c->n_grid < 2 || c->n_grid > kMaxBins) return kErrInterp;
    c->f = std::log(c->mmax / c->mmin) / ((double)c->n_grid - 1.0);
    // normalisations after the size ranges are final (confCdisks/confCblobs)
    for (int i = 0; i < c->n_disks; ++i)
        c->disks[i].C = norm_c(c->disks[i].Mtot, c->disks[i].eta, c->disks[i].mmax, c->disks[i].mmin);
    for (int i = 0; i < c->n_blobs; ++i)
        c->blobs[i].C = norm_c(c->blobs[i].Mtot, c->blobs[i].eta, c->blobs[i].mmax, c->blobs[i].mmin);
    c->n_wav = o->n_wav;
    c->n_theta = o->n_theta;
    return 0;
}

double ddyn_outer_radius(const ddyn_config* c)
{
    double R = 0.0;
    for (int i = 0; i < c->n_disks; ++i) R = std::fmax(R, c->disks[i].Rout);
    for (int i = 0; i < c->n_blobs; ++i)
        R = std::fmax(R, c->blobs[i].a * (1.0 + c->blobs[i].ecc) + c->blobs[i].sigma);
    for (int i = 0; i < c->n_planets; ++i)
        R = std::fmax(R, std::sqrt(c->px[i] * c->px[i] + c->py[i] * c->py[i] + c->pz[i] * c->pz[i]));
    return 2.0 * R;
}

int ddyn_grain_tables(const ddyn_config* c, const ddyn_optics* o, int32_t n_max, float* beta, float* betasw,
                      float* q, double* size, double* Rmin_au, double* lum)
{
    if (!c || !o) return kErrOptics;
    const int n = c->n_grid;
    if (n > n_max) return kErrInterp;
    GridTables g;
    if (const int rc = grid_tables(c, o, g)) return rc;
    const size_t nw = (size_t)o->n_wav, nt = (size_t)o->n_theta;
    const std::vector<double>& a = g.a;
    // calculate_beta: radiation pressure efficiency averaged over the spectrum
    double star = 0.0;
    for (size_t j = 0; j + 1 < nw; ++j) star += 0.5 * (o->w[j + 1] - o->w[j]) * (g.in[j + 1] + g.in[j]);
    const double L = star * 4.0 * kPI * std::pow(c->dist, 2.0) / kLsunW;
    const double m0 = c->pm[0];
    const double swconst = 3.0 * c->mloss * kCd * kMSun * kYear * c->vsw / 32.0 / kPI / m0 / c->rho / kAU / kAU / kAU;
    std::vector<double> Qpr(nw);
    for (int i = 0; i < n; ++i) {
        double b = 0.0;
        if (c->radQ == 1 && star > 0.0) {
            for (size_t j = 0; j < nw; ++j) {
                double gq = 0.0, S = 0.0, cm = -1.0;
                for (size_t k = 1; k < nt; ++k) {
                    const double cp = (double)k * 2.0 / ((double)nt - 1.0) - 1.0;
                    const double p0 = g.Qp[nw * (nt * i + k - 1) + j], p1 = g.Qp[nw * (nt * i + k) + j];
                    S += 0.5 * (p0 + p1) * (cp - cm);
                    gq += 0.5 * (cm * p0 + cp * p1) * (cp - cm);
                    cm = cp;
                }
                gq = S != 0.0 ? gq / S : 0.0;
                Qpr[j] = (g.Qa[i * nw + j] + g.Qs[i * nw + j] * (1.0 - gq)) * g.in[j];
            }
            double Qsum = 0.0;
            for (size_t j = 0; j + 1 < nw; ++j) Qsum += 0.5 * (o->w[j + 1] - o->w[j]) * (Qpr[j + 1] + Qpr[j]);
            b = 0.57 * Qsum / star * (L / m0 * kGrav) / a[i] / 1e6 / c->rho / 1e-3;
        }
        if (beta) beta[i] = (float)b;
        if (betasw) betasw[i] = (float)(c->SWQ == 1 ? swconst / a[i] : 0.0);
        if (q) q[i] = (float)(c->kqe * a[i] * 1e9);
        if (size) size[i] = a[i];
    }
    if (Rmin_au) *Rmin_au = sublimation_radius(c, o, g, ddyn_outer_radius(c) * kAU) / kAU;
    if (lum) *lum = L;
    return 0;
}

int ddyn_grid_optics_make(const ddyn_config* c, const ddyn_optics* o, ddyn_grid_optics** out)
{
    if (!out) return kErrOptics;
    *out = nullptr;
    if (!c || !o) return kErrOptics;
    GridTables g;
    if (const int rc = grid_tables(c, o, g)) return rc;
    const int n = c->n_grid;
    const size_t nw = (size_t)o->n_wav;
    ddyn_grid_optics* t = new ddyn_grid_optics();
    t->n_bins = n;
    t->n_wav = o->n_wav;
    t->n_theta = o->n_theta;
    t->w = o->w;
    t->in = g.in;
    t->size = g.a;
    t->Qabs = g.Qa;
    t->Qsca = g.Qs;
    t->Qpf = g.Qp;
    // calculate_dust_temp (src/grid.cu:754-834), radii in m until the end:
    // R_max twice the outer radius, R_min the sublimation radius of the most
    // easily heated bin; per bin the equilibrium distance of every temperature
    // from 2 T_sub down to 4 K in steps of 2 K, then the temperature on kNR
    // radii log-spaced from R_min to R_max, log-log between those points.
    const double Rmax = ddyn_outer_radius(c) * kAU;
    const double Rmin = sublimation_radius(c, o, g, Rmax);
    const int nT = (int)std::floor((2.0 * kTsub - 3.0) * 0.5) + 1;
    std::vector<double> star((size_t)n), Tarr((size_t)nT), Rtab((size_t)n * nT), B(nw);
    for (int k = 0; k < n; ++k) {
        const double* Qa = g.Qa.data() + (size_t)k * nw;
        double sk = 0.0;
        for (size_t i = 1; i < nw; ++i)
            sk += 0.5 * (o->w[i] - o->w[i - 1]) * (Qa[i - 1] * g.in[i - 1] + Qa[i] * g.in[i]) / kPI;
        star[(size_t)k] = sk * std::pow(c->dist, 2.0);
    }
    // the Planck function of each temperature once for every bin (the same
    // values and sums as bin by bin)
    int j = 0;
    for (double T = 2.0 * kTsub; T >= 3.0; T -= 2.0) {
        for (size_t i = 0; i < nw; ++i) B[i] = planck(T, o->w[i]);
        for (int k = 0; k < n; ++k) {
            const double* Qa = g.Qa.data() + (size_t)k * nw;
            double gf = 0.0;
            for (size_t i = 1; i < nw; ++i) gf += 0.5 * (o->w[i] - o->w[i - 1]) * (Qa[i - 1] * B[i - 1] + Qa[i] * B[i]);
            Rtab[(size_t)k * nT + j] = 0.5 * std::sqrt(star[(size_t)k] / gf);
        }
        Tarr[(size_t)j] = T;
        ++j;
    }
    t->Td.assign((size_t)n * kNR, 0.0);
    for (int k = 0; k < n; ++k) {
        const double* Rarr = Rtab.data() + (size_t)k * nT;
        // The reference searches upwards from the previous radius's interval
        // and indexes outside the table when a radius lies inside the hottest
        // (4000 K) distance or beyond the coldest; here the first or last
        // interval extrapolates instead.
        int ip = 0;
        for (int i = 0; i < kNR; ++i) {
            const double R = Rmin * std::exp(std::log(Rmax / Rmin) * ((double)i) / ((double)kNR - 1.0));
            while (ip < nT - 1 && Rarr[ip] <= R) ++ip;
            if (ip == 0) ip = 1;
            const int im = ip - 1;
            t->Td[(size_t)k * kNR + i] =
                std::exp(std::log(Tarr[(size_t)im]) + std::log(Tarr[(size_t)ip] / Tarr[(size_t)im]) *
                                                          std::log(R / Rarr[im]) / std::log(Rarr[ip] / Rarr[im]));
            --ip;
        }
    }
    t->Rmin = Rmin / kAU;
    t->Rmax = Rmax / kAU;
    *out = t;
    return 0;
}

void ddyn_grid_optics_free(ddyn_grid_optics* t) { delete t; }

int ddyn_grid_optics_tables(const ddyn_grid_optics* t, ddyn_optics_tables* v)
{
    if (!t || !v) return kErrOptics;
    v->n_bins = t->n_bins;
    v->n_wav = t->n_wav;
    v->n_theta = t->n_theta;
    v->n_r = kNR;
    v->w = t->w.data();
    v->in = t->in.data();
    v->size = t->size.data();
    v->Qabs = t->Qabs.data();
    v->Qsca = t->Qsca.data();
    v->Qpfunc = t->Qpf.data();
    v->Td = t->Td.data();
    v->Rmin = t->Rmin;
    v->Rmax = t->Rmax;
    return 0;
}

int ddyn_image_wavelengths(const ddyn_config* c, const ddyn_grid_optics* t, int32_t* wj)
{
    if (!c || !t || (c->nwavs > 0 && !wj)) return kErrOptics;
    const int nw = t->n_wav;
    for (int j = 0; j < c->nwavs; ++j) {
        const double v = c->wavs[j];
        if (v < t->w[0] || v > t->w[(size_t)nw - 1]) return kErrWav;
        // the first wavelength at or above the requested one, at most the
        // second last, and the image is made one below it
        // (src/functions.cu:329-335; where the reference would index -1, for a
        // request equal to the first wavelength, the first is used)
        int i = 0;
        while (t->w[(size_t)i] < v && i < nw - 2) ++i;
        wj[j] = i > 0 ? i - 1 : 0;
    }
    return 0;
}

int ddyn_generate_counts(const ddyn_config* c, int64_t* n_mass, int64_t* n_dust)
{
    if (!c) return kErrInterp;
    int64_t nm = 0, nd = 0;
    for (int j = 0; j < c->n_disks; ++j) {
        nm += c->disks[j].Ngrav;
        const DustBins b = dust_bins(*c, c->disks[j].dmmax, c->disks[j].mmin, c->disks[j].Nsmall);
        nd += (int64_t)b.ngrid * b.npg;
    }
    for (int j = 0; j < c->n_blobs; ++j) {
        nm += c->blobs[j].Ngrav;
        const DustBins b = dust_bins(*c, c->blobs[j].dmmax, c->blobs[j].mmin, c->blobs[j].Nsmall);
        nd += (int64_t)b.ngrid * b.npg;
    }
    if (n_mass) *n_mass = nm;
    if (n_dust) *n_dust = nd;
    return 0;
}

int ddyn_generate(const ddyn_config* c, const ddyn_optics* o, const float* beta, int32_t n_bins, int64_t n_all,
                  double* x, double* y, double* z, double* vx, double* vy, double* vz, double* N, double* mu,
                  int32_t* id)
{
    if (!c || !x || !y || !z || !vx || !vy || !vz || !N || !mu || !id) return kErrInterp;
    Sink out{0, n_all, x, y, z, vx, vy, vz, N, mu, id};
    // star and planets first (src/grid.cu:102-106)
    for (int j = 0; j < c->n_planets; ++j) {
        const double r[6] = {c->px[j], c->py[j], c->pz[j], c->pvx[j], c->pvy[j], c->pvz[j]};
        if (!out.add(r, 1.0, c->pm[j], -1)) return kErrInterp;
    }
    Subtractive55 g(c->seed);
    // gravitating bodies of the disks, then of the blobs: the mass ladder from each one's mmax down
    for (int j = 0; j < c->n_disks; ++j) {
        const ddyn_disk& D = c->disks[j];
        long double mmax = D.mmax;
        for (int64_t l = 0; l < D.Ngrav; ++l) {
            const double m = next_body(mmax, D.C, D.eta);
            double r[6];
            if (!orbit_disk(*c, D, m, 0.0, g, r) || !out.add(r, 1.0, m, -1)) return kErrInterp;
        }
    }
    for (int j = 0; j < c->n_blobs; ++j) {
        const ddyn_blob& B = c->blobs[j];
        long double mmax = B.mmax;
        for (int64_t l = 0; l < B.Ngrav; ++l) {
            const double m = next_body(mmax, B.C, B.eta);
            double r[6];
            if (!orbit_blob(*c, B, m, 0.0, g, r)) return kErrInterp;
            if (!out.add(r, 1.0, m, -1)) return kErrInterp;
        }
    }
    // massless grains: every disk's mass bins, then every blob's
    int rc;
    for (int j = 0; j < c->n_disks; ++j) {
        const ddyn_disk& D = c->disks[j];
        const DustBins b = dust_bins(*c, D.dmmax, D.mmin, D.Nsmall);
        auto orb = [&](double m, double bt, Subtractive55& gg, double* r) { return orbit_disk(*c, D, m, bt, gg, r); };
        if ((rc = dust_of(*c, o, b, D.C, D.eta, beta, n_bins, g, out, orb))) return rc;
    }
    for (int j = 0; j < c->n_blobs; ++j) {
        const ddyn_blob& B = c->blobs[j];
        const DustBins b = dust_bins(*c, B.dmmax, B.mmin, B.Nsmall);
        auto orb = [&](double m, double bt, Subtractive55& gg, double* r) { return orbit_blob(*c, B, m, bt, gg, r); };
        if ((rc = dust_of(*c, o, b, B.C, B.eta, beta, n_bins, g, out, orb))) return rc;
    }
    return out.n == n_all ? 0 : kErrInterp;
}

int ddyn_restart_ids(const ddyn_config* c, int64_t n_all, const double* mu, int32_t* id)
{
    if (!c || !mu || !id) return kErrInterp;
    for (int64_t i = 0; i < n_all; ++i) {
        if (i < c->n_planets) { id[i] = -1; continue; }
        const double m = mu[i] * kMSun / kGrav;
        if (!(m > 0.0) || !(c->mmin > 0.0) || !(c->f > 0.0)) { id[i] = -1; continue; }
        const double r = std::round(std::log(m / c->mmin) / c->f);
        id[i] = (r < 0.0 || r > (double)(c->n_grid - 1)) ? -1 : (int32_t)r;
    }
    return 0;
}

}  // extern "C"
