// ddyn_host.cpp -- .param reader, time-step rule and data-file writer/reader
// of the dynamical hot path (see include/ddyn_host.h).  Host C++ only.
#include "ddyn_host.h"
#include "ddyn_host_impl.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

namespace {

// Constants of src/ddyn.cuh:16-38.
constexpr double kMEarth = 5.97e24;
constexpr double kAU = 1.49598e11;
constexpr double kMSun = 1.98892e30;
constexpr double kRSun = 695700000.0;
constexpr double kYear = 3.1556926e7;
constexpr double kPI = 3.14159265358979323846;
constexpr double kPI2 = 6.28318530717958623200;
constexpr double kGrav = 4.0 * kPI * kPI;
constexpr double kPc = 3.08567758149137e16;
constexpr int kNR = 100;

// Error texts, numbered as the reference's message table (src/error.cu:5-64).
const char* error_text(int code)
{
    switch (code) {
        case 0: return "ok";
        case 3: return "missing or empty input/model file";
        case 4: return "invalid verbose parameter";
        case 6: return "GPU id must be >= 0";
        case 7: return "graphics step size must be > 0";
        case 8: return "graphics parameter must be yes or no";
        case 9: return "write-out step sizes must be > 0";
        case 10: return "simulation end time must be > 0";
        case 11: return "steps per orbit must be > 0";
        case 12: return "stellar parameters must be > 0";
        case 13: return "all eight stellar parameters are required";
        case 14: return "bulk density must be > 0";
        case 15: return "density, composition file and charge coefficient are required";
        case 17: return "parameter file has too few records";
        case 18: return "planet orbit needs seven values";
        case 30: return "all three physics switches are required";
        case 31: return "physics switches must be yes or no";
        case 32: return "charge coefficient must be > 0 with the Lorentz force on";
        case 36: return "system geometry needs seven values";
        case 37: return "distance must be > 0";
        case 38: return "field of view must be > 0";
        case 39: return "spectral or optical data file missing";
        case 40: return "Teff outside the spectral library range";
        case 41: return "image wavelengths are required";
        case 42: return "optical data not well formed";
        case 44: return "too many planets";
        case 45: return "too many disks";
        case 46: return "too many blobs";
        case 54: return "interpolation or particle generation error";
        case 55: return "data file version differs from this version";
        case 56: return "data file nR differs from the compiled radial grid size";
        case 57: return "blob needs 'cloud' or 'collision'";
        case 58: return "blob needs fourteen values";
        default: return "error";
    }
}

}  // namespace

namespace ddyn_impl {

// Whitespace tokenizer with double-quoted tokens (quotes removed), as the
// reference's tokenize_spaces/remove_quotes do; at most `maxtok` tokens.
std::vector<std::string> tokenize(const char* line, size_t maxtok)
{
    std::vector<std::string> out;
    const char* p = line;
    while (*p && out.size() < maxtok) {
        while (*p && std::isspace((unsigned char)*p)) ++p;
        if (!*p) break;
        std::string tok;
        bool quoted = false;
        while (*p && (quoted || !std::isspace((unsigned char)*p))) {
            if (*p == '"') quoted = !quoted;
            else tok.push_back(*p);
            ++p;
        }
        out.push_back(tok);
    }
    return out;
}

}  // namespace ddyn_impl

namespace {

using ddyn_impl::coe2rv;
using ddyn_impl::tokenize;

double num(const std::string& s) { return std::strtod(s.c_str(), nullptr); }
long lnum(const std::string& s) { return std::strtol(s.c_str(), nullptr, 10); }

int yesno(const std::string& s)
{
    if (s == "yes" || s == "Y" || s == "YES" || s == "y" || s == "true" || s == "TRUE") return 1;
    if (s == "no" || s == "N" || s == "NO" || s == "n" || s == "false" || s == "FALSE") return 0;
    return -1;
}

void copy_str(char* dst, const std::string& s)
{
    std::snprintf(dst, DDYN_STR, "%s", s.c_str());
}

}  // namespace

namespace ddyn_impl {

// coe2rv (src/functions.cu:6-46), mu = (M + m)(1 - beta).  The sine and
// cosine of each angle come from one sincos call, as in the reference built
// by g++ at -O3 (its sin/cos pairs of one argument become sincos calls, and
// the C library's sincos differs from sin and cos in the last bit for about
// one argument in 700).
void coe2rv(double M, double m, double a, double e, double i, double w, double o, double nu,
            double beta, double* r)
{
    double snu, cnu, so, co, sw, cw, si, ci;
    ::sincos(nu, &snu, &cnu);
    ::sincos(o, &so, &co);
    ::sincos(w, &sw, &cw);
    ::sincos(i, &si, &ci);
    const double p = a * (1.0 - e * e);
    const double mu = (M + m) * (1.0 - beta);
    const double xe = p * cnu / (1.0 + e * cnu);
    const double ye = p * snu / (1.0 + e * cnu);
    const double vxe = -std::sqrt(mu / p) * snu;
    const double vye = std::sqrt(mu / p) * (e + cnu);
    // The rotation with the reference's products and sums, including the
    // terms of the zero perifocal z and v_z (they only turn -0 sums into +0).
    const double ze = 0.0, vze = 0.0;
    const double r11 = co * cw - so * sw * ci, r12 = -co * sw - so * cw * ci, r13 = so * si;
    const double r21 = so * cw + co * sw * ci, r22 = -so * sw + co * cw * ci, r23 = -co * si;
    r[0] = xe * r11 + ye * r12 + ze * r13;
    r[1] = xe * r21 + ye * r22 + ze * r23;
    r[2] = xe * sw * si + ye * cw * si + ze * ci;
    r[3] = vxe * r11 + vye * r12 + vze * r13;
    r[4] = vxe * r21 + vye * r22 + vze * r23;
    r[5] = vxe * sw * si + vye * cw * si + vze * ci;
}

}  // namespace ddyn_impl

namespace {

// recalc_COM (src/read_in.cu:344-375).
void recalc_com(ddyn_config* c)
{
    if (c->n_planets == 1) return;
    double X = 0, Y = 0, Z = 0, M = 0;
    for (int j = 0; j < c->n_planets; ++j) {
        X += c->pm[j] * c->px[j]; Y += c->pm[j] * c->py[j]; Z += c->pm[j] * c->pz[j]; M += c->pm[j];
    }
    X /= M; Y /= M; Z /= M;
    for (int j = 0; j < c->n_planets; ++j) { c->px[j] -= X; c->py[j] -= Y; c->pz[j] -= Z; }
    double vx = 0, vy = 0, vz = 0;
    for (int j = 1; j < c->n_planets; ++j) {
        vx -= c->pm[j] * c->pvx[j]; vy -= c->pm[j] * c->pvy[j]; vz -= c->pm[j] * c->pvz[j];
    }
    c->pvx[0] = vx / c->pm[0]; c->pvy[0] = vy / c->pm[0]; c->pvz[0] = vz / c->pm[0];
}

int record(ddyn_config* c, int i, const std::vector<std::string>& d, int& stay)
{
    const size_t t = d.size();
    stay = 0;
    switch (i) {
        case 0: {   // verbose
            const long v = lnum(d[0]);
            if (v < 0) return 4;
            c->verb = (int)v;
            return 0;
        }
        case 1: {   // GPU id
            const long v = lnum(d[0]);
            if (v < 0) return 6;
            c->gpuid = (int)v;
            return 0;
        }
        case 2: {   // graphics
            const int yn = yesno(d[0]);
            if (yn == 1) {
                if (t < 2 || num(d[1]) <= 0) return 7;
                c->gl = 1; c->t_gl = num(d[1]);
            } else if (yn == 0) {
                c->gl = 0; c->t_gl = 0;
            } else {
                return 8;
            }
            return 0;
        }
        case 3:     // data, FITS and SED output times
            if (t < 3 || num(d[0]) <= 0 || num(d[1]) <= 0 || num(d[2]) <= 0) return 9;
            c->t_write = num(d[0]); c->t_fits = num(d[1]); c->t_sed = num(d[2]);
            return 0;
        case 4:
            if (num(d[0]) <= 0) return 10;
            c->t_end = num(d[0]);
            return 0;
        case 5:
            if (lnum(d[0]) <= 0) return 11;
            c->n_step = (int)lnum(d[0]);
            return 0;
        case 6: {   // star: M, sptype, Teff, R, B, P_rot, mass loss, wind speed
            if (t < 8) return 13;
            for (int k : {0, 2, 3, 4, 5, 6, 7})
                if (num(d[k]) <= 0) return 12;
            c->n_planets = 1;
            c->px[0] = c->py[0] = c->pz[0] = c->pvx[0] = c->pvy[0] = c->pvz[0] = 0.0;
            c->pm[0] = num(d[0]) * kGrav;
            copy_str(c->sptype, d[1]);
            c->teff = num(d[2]);
            c->rstar = num(d[3]) * kRSun / kAU;
            c->Bstar = num(d[4]) * 1.0e-4 * kYear * kYear / kMSun;
            c->Wz = 365.2422 / num(d[5]);
            c->mloss = num(d[6]);
            c->vsw = num(d[7]) * 1000.0;
            return 0;
        }
        case 7:     // geometry
            if (t < 7) return 36;
            if (num(d[0]) <= 0) return 37;
            if (num(d[1]) <= 0) return 38;
            c->dist = num(d[0]) * kPc; c->fov = num(d[1]);
            c->PA = num(d[2]) * kPI / 180.0; c->inc = num(d[3]) * kPI / 180.0;
            c->longitude = num(d[4]) * kPI / 180.0;
            c->xsize = (int)lnum(d[5]); c->ysize = (int)lnum(d[6]);
            return 0;
        case 8:     // image wavelengths [um]
            if (t < 1) return 41;
            c->nwavs = (int)std::min<size_t>(t, DDYN_MAX_WAVS);
            for (int k = 0; k < c->nwavs; ++k) c->wavs[k] = num(d[k]) * 1e-6;
            return 0;
        case 9:     // spectrum: Library (Kurucz grid) or a file
            if (d[0] == "Library" || d[0] == "library" || d[0] == "LIBRARY") {
                double te = c->teff;
                if (te >= 3375 && te < 10000) te = std::round(te / 250.0) * 250;
                else if (te >= 10000 && te < 13000) te = std::round(te / 500.0) * 500;
                else if (te >= 13000 && te < 35000) te = std::round(te / 1000.0) * 1000;
                else if (te >= 35000 && te < 50000) te = std::round(te / 2500.0) * 2500;
                else return 40;
                c->teff = te;
                std::snprintf(c->specfile, DDYN_STR, "Kurucz/Kurucz_%.0f.dat", te);
            } else {
                copy_str(c->specfile, d[0]);
            }
            return 0;
        case 10: {  // physics switches
            if (t < 3) return 30;
            const int r = yesno(d[0]), s = yesno(d[1]), f = yesno(d[2]);
            if (r < 0 || s < 0 || f < 0) return 31;
            c->radQ = r; c->SWQ = s; c->FLQ = f;
            return 0;
        }
        case 11:    // planets block
            if (d[0] == "begin_planets") { stay = 1; return 0; }
            if (d[0] == "end_planets") { recalc_com(c); return 0; }
            if (t < 7) return 18;
            if (c->n_planets >= DDYN_MAX_PLANETS) return 44;
            {
                const int j = c->n_planets++;
                // degrees to radians as the reference's i *= M_PI/180.0 (src/read_in.cu:96-100)
                const double inc = num(d[0]) * (M_PI / 180.0), a = num(d[1]), e = num(d[2]);
                const double nu = num(d[3]) * (M_PI / 180.0), o = num(d[4]) * (M_PI / 180.0);
                const double w = num(d[5]) * (M_PI / 180.0), m = num(d[6]) * kGrav;
                double r[6];
                coe2rv(c->pm[0], m, a, e, inc, w, o, nu, 0.0, r);
                c->px[j] = r[0]; c->py[j] = r[1]; c->pz[j] = r[2];
                c->pvx[j] = r[3]; c->pvy[j] = r[4]; c->pvz[j] = r[5]; c->pm[j] = m;
            }
            stay = 1;
            return 0;
        case 12:    // density, composition, charge coefficient
            if (t < 3) return 15;
            if (num(d[0]) <= 0) return 14;
            c->rho = num(d[0]) * 1000.0;
            copy_str(c->comp, d[1]);
            if (num(d[2]) <= 0 && c->FLQ == 1) return 32;
            c->kqe = num(d[2]) / kYear;
            return 0;
        case 13:    // disks block
            if (d[0] == "begin_disks") { stay = 1; return 0; }
            if (d[0] == "end_disks") return 0;
            if (c->n_disks >= DDYN_MAX_DISKS) return 45;
            {
                ddyn_disk& k = c->disks[c->n_disks++];
                std::memset(&k, 0, sizeof k);
                const double amax = t > 4 ? num(d[4]) : 0, amin = t > 5 ? num(d[5]) : 0;
                k.Rin = num(d[0]); k.Rout = t > 1 ? num(d[1]) : 0;
                k.dhdr = std::atan(t > 2 ? num(d[2]) : 0); k.sigmae = t > 3 ? num(d[3]) : 0;
                k.amax = amax * 1000.0; k.amin = amin / 1.0e6;
                k.eta = ((t > 6 ? num(d[6]) : 0) + 2.0) / 3.0;
                k.Mtot = (t > 7 ? num(d[7]) : 0) * kMEarth;
                k.mmax = 4.0 * std::pow(amax * 1000.0, 3.0) * kPI * c->rho / 3.0;
                k.mmin = 4.0 * std::pow(amin / 1.0e6, 3.0) * kPI * c->rho / 3.0;
                k.gammas = t > 8 ? num(d[8]) : 0;
                k.Ngrav = t > 9 ? lnum(d[9]) : 0; k.Nsmall = t > 10 ? lnum(d[10]) : 0;
            }
            stay = 1;
            return 0;
        case 14:    // blobs block
            if (d[0] == "begin_blobs") { stay = 1; return 0; }
            if (d[0] == "end_blobs") return 0;
            if (t < 14) return 58;
            if (c->n_blobs >= DDYN_MAX_BLOBS) return 46;
            {
                ddyn_blob& b = c->blobs[c->n_blobs];
                std::memset(&b, 0, sizeof b);
                if (d[12] == "cloud" || d[12] == "Cloud") { b.collcloud = 1; b.sigma = num(d[13]); }
                else if (d[12] == "collision" || d[12] == "Collision") { b.collcloud = 0; b.vkepp = num(d[13]); }
                else return 57;
                b.a = num(d[0]); b.ecc = num(d[1]);
                b.inc = num(d[2]) * M_PI / 180.0; b.w = num(d[3]) * M_PI / 180.0;
                b.o = num(d[4]) * M_PI / 180.0; b.nu = num(d[5]) * M_PI / 180.0;
                b.amax = num(d[6]) * 1000.0; b.amin = num(d[7]) / 1.0e6;
                b.eta = (num(d[8]) + 2.0) / 3.0; b.Mtot = num(d[9]) * kMEarth;
                b.mmax = 4.0 * std::pow(num(d[6]) * 1000.0, 3.0) * kPI * c->rho / 3.0;
                b.mmin = 4.0 * std::pow(num(d[7]) / 1.0e6, 3.0) * kPI * c->rho / 3.0;
                b.Ngrav = lnum(d[10]); b.Nsmall = lnum(d[11]);
                c->n_blobs++;
            }
            stay = 1;
            return 0;
        case 15:
            copy_str(c->name_stub, d[0]);
            return 0;
        default:
            return 0;
    }
}

// Disk/blob normalisation (confCdisks/confCblobs, src/read_in.cu:236-266).
double norm_c(double Mtot, double eta, double mmax, double mmin)
{
    if (eta == 1) return Mtot / (mmax - mmin);
    if (eta == 2) return Mtot / std::log(mmax / mmin);
    return Mtot / ((std::pow(mmax, 2.0 - eta) - std::pow(mmin, 2.0 - eta)) / (2.0 - eta));
}

}  // namespace

extern "C" {

const char* ddyn_host_error_text(int code) { return error_text(code); }

int ddyn_read_param(const char* path, ddyn_config* cfg)
{
    if (!path || !cfg) return 3;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->seed = -2;   // the reference seeds its generator with -2 (src/grid.cu:80)
    FILE* f = std::fopen(path, "rb");
    if (!f) return 3;
    char buff[512];
    int i = 0, rc = 0;
    while (std::fgets(buff, sizeof buff, f)) {
        if (buff[0] == '#') continue;
        const auto d = tokenize(buff, 16);
        if (d.empty()) continue;
        int stay = 0;
        if ((rc = record(cfg, i, d, stay))) break;
        if (!stay) ++i;
    }
    std::fclose(f);
    if (rc) return rc;
    if (i < 15) return 17;
    for (int k = 0; k < cfg->n_disks; ++k)
        cfg->disks[k].C = norm_c(cfg->disks[k].Mtot, cfg->disks[k].eta, cfg->disks[k].mmax, cfg->disks[k].mmin);
    for (int k = 0; k < cfg->n_blobs; ++k)
        cfg->blobs[k].C = norm_c(cfg->blobs[k].Mtot, cfg->blobs[k].eta, cfg->blobs[k].mmax, cfg->blobs[k].mmin);
    return 0;
}

double ddyn_det_step(const ddyn_config* c, int64_t n_grav, const double* x, const double* y,
                     const double* z, const double* mu)
{
    double dtmin = 1e137;
    if (n_grav > 1) {
        for (int64_t i = 1; i < n_grav; ++i) {
            const double dx = x[i] - x[0], dy = y[i] - y[0], dz = z[i] - z[0];
            const double R = std::pow(dx * dx + dy * dy + dz * dz, 1.5);
            const double dtt = kPI2 * std::sqrt(R / (mu[0] + mu[i])) / c->n_step;
            if (dtt < dtmin) dtmin = dtt;
        }
    }
    // disk and blob terms exactly as the reference writes them (sqrt(R/mu0))
    for (int k = 0; k < c->n_disks; ++k) {
        const double dtt = kPI2 * std::sqrt(c->disks[k].Rin / mu[0]) / c->n_step;
        if (dtt < dtmin) dtmin = dtt;
    }
    for (int k = 0; k < c->n_blobs; ++k) {
        const ddyn_blob& b = c->blobs[k];
        const double dtt = kPI2 * std::sqrt((b.a * (1.0 - b.ecc) - b.sigma) / mu[0]) / c->n_step;
        if (dtt < dtmin) dtmin = dtt;
    }
    return dtmin;
}

int ddyn_write_datafile(const char* path, const char* version, const ddyn_config* c,
                        const ddyn_run_info* run, const double* x, const double* y, const double* z,
                        const double* vx, const double* vy, const double* vz, const double* N,
                        const double* mu, const int32_t* id, const float* beta_bin, int32_t n_bins,
                        const double* beta_p)
{
    FILE* fw = std::fopen(path, "wb");
    if (!fw) return 3;
    std::fprintf(fw, "#\n#Calculated_with_DiskDyn_Version %s\n#\n#Config file parameters\n", version);
    std::fprintf(fw, "#Verbose\t\t\t%d\n#GPUID\t\t\t\t%d\n#Graphical_Output\t\t%d\n", c->verb, c->gpuid, c->gl);
    std::fprintf(fw, "#Time_Out_for_Graphics\t\t%.3f\n#Time_Out_for_Data\t\t%.3f\n", c->t_gl, c->t_write);
    std::fprintf(fw, "#Time_Out_for_SED\t\t%.3f\n#Time_Out_for_FITS\t\t%.3f\n", c->t_sed, c->t_fits);
    std::fprintf(fw, "#Time_End_for_model\t\t%.3f\n#N_step_per_orbit\t\t%d\n", c->t_end, c->n_step);
    std::fprintf(fw, "#Spectral-type_of_star\t\t%s\n#Temperature_of_star\t\t%.2f\n", c->sptype, c->teff);
    std::fprintf(fw, "#Radius_of_star\t\t\t%.5f\n", c->rstar * kAU / kRSun);
    std::fprintf(fw, "#Magnetic_field_of_star\t\t%.5e\n", c->Bstar * kMSun * 1.0e4 / kYear / kYear);
    std::fprintf(fw, "#Stellar_rotation_of_star\t%.5f\n", c->Wz != 0.0 ? 365.2422 / c->Wz : 0.0);
    std::fprintf(fw, "#Stellar_mass_loss\t\t%.5e\n#Stellar_wind_speed\t\t%.5e\n", c->mloss, c->vsw);
    std::fprintf(fw, "#Include_rad_press\t\t%d\n#Include_SWD\t\t\t%d\n#Include_Lorentz\t\t%d\n", c->radQ, c->SWQ, c->FLQ);
    std::fprintf(fw, "#Bulk_density\t\t\t%.3f\n#Electric_charge_coeff\t\t%.3f\n", c->rho, c->kqe * kYear);
    std::fprintf(fw, "#Name_Stub\t\t\t%s\n#Distance\t\t\t%.5f\n", c->name_stub, c->dist / kPc);
    std::fprintf(fw, "#Field_of_View\t\t\t%.6f\n#Position_Angle\t\t\t%.5f\n", c->fov, c->PA * 180.0 / kPI);
    std::fprintf(fw, "#Inclination_Angle\t\t%.5f\n#Longitude_Angle\t\t%.5f\n", c->inc * 180.0 / kPI,
                 c->longitude * 180.0 / kPI);
    std::fprintf(fw, "#Image_x_size\t\t\t%d\n#Image_y_size\t\t\t%d\n#Spectral-file\t\t\t%s\n", c->xsize,
                 c->ysize, c->specfile);
    std::fprintf(fw, "#Minimum_particle_size\t\t%.8e\n#Maximum_particle_size\t\t%.8e\n", c->amin, c->amax);
    std::fprintf(fw, "#Minimum_particle_mass\t\t%.8e\n#Maximum_particle_mass\t\t%.8e\n", c->mmin, c->mmax);
    std::fprintf(fw, "#Number_of_dust_grids\t\t%d\n#Mass_grid_ratio\t\t%.8e\n", c->n_grid, c->f);
    std::fprintf(fw, "#Optical_composition_file\t%s\n#Wavelengths_in_interpolation\t%d\n", c->comp, c->n_wav);
    std::fprintf(fw, "#Number_of_thetas\t\t%d\n#Number_of_dust_grids\t\t%d\n", c->n_theta, c->n_grid);
    std::fprintf(fw, "#Number_of_planets\t\t%d\n#Number_of_disks\t\t%d\n#Number_of_blobs\t\t%d\n",
                 c->n_planets, c->n_disks, c->n_blobs);
    std::fprintf(fw, "#Image_wavelengths\t");
    for (int i = 0; i < c->nwavs; ++i) std::fprintf(fw, "\t%.3f", c->wavs[i] * 1e6);
    std::fprintf(fw, "\n#Seed_of_model\t\t\t%ld\n", (long)c->seed);
    const ddyn_disk* D = c->disks;
    const int nd = c->n_disks, nb = c->n_blobs;
#define ROW(label, fmt, expr, n)                                  \
    std::fprintf(fw, label);                                      \
    for (int i = 0; i < (n); ++i) std::fprintf(fw, fmt, expr);    \
    std::fprintf(fw, "\n");
    ROW("#Disk_Rin\t", "\t\t%.3f", D[i].Rin, nd)
    ROW("#Disk_Rout\t", "\t\t%.3f", D[i].Rout, nd)
    ROW("#Disk_dhdr\t", "\t\t%.4f", D[i].dhdr, nd)
    ROW("#Disk_sigmae\t", "\t\t%.4f", D[i].sigmae, nd)
    ROW("#Disk_amax\t\t", "\t%.6e", D[i].amax, nd)
    ROW("#Disk_mmax\t\t", "\t%.6e", D[i].mmax, nd)
    ROW("#Disk_amin\t\t", "\t%.6e", D[i].amin, nd)
    ROW("#Disk_mmin\t\t", "\t%.6e", D[i].mmin, nd)
    ROW("#Disk_dmax\t\t", "\t%.6e", D[i].dmax, nd)
    ROW("#Disk_dmmax\t\t", "\t%.6e", D[i].dmmax, nd)
    ROW("#Disk_eta\t", "\t\t%.4f", D[i].eta, nd)
    ROW("#Disk_gammas\t", "\t\t%.3f", D[i].gammas, nd)
    ROW("#Disk_C\t\t\t", "\t%.8e", D[i].C, nd)
    ROW("#Disk_Mtot\t\t", "\t%.8e", D[i].Mtot, nd)
    ROW("#Disk_Ngrav\t", "\t\t%ld", (long)D[i].Ngrav, nd)
    ROW("#Disk_Nsmall\t", "\t\t%ld", (long)D[i].Nsmall, nd)
    const ddyn_blob* B = c->blobs;
    ROW("#Blob_a\t\t", "\t\t%.5f", B[i].a, nb)
    ROW("#Blob_ecc\t", "\t\t%.5f", B[i].ecc, nb)
    ROW("#Blob_inc\t", "\t\t%.5f", B[i].inc, nb)
    ROW("#Blob_w\t\t", "\t\t%.5f", B[i].w, nb)
    ROW("#Blob_o\t", "\t\t%.5f", B[i].o, nb)
    ROW("#Blob_nu\t", "\t\t%.5f", B[i].nu, nb)
    ROW("#Blob_sigma\t", "\t\t%.3f", B[i].sigma, nb)
    ROW("#Blob_amax\t\t", "\t%.6e", B[i].amax, nb)
    ROW("#Blob_mmax\t\t", "\t%.6e", B[i].mmax, nb)
    ROW("#Blob_amin\t\t", "\t%.6e", B[i].amin, nb)
    ROW("#Blob_mmin\t\t", "\t%.6e", B[i].mmin, nb)
    ROW("#Blob_dmax\t\t", "\t%.6e", B[i].dmax, nb)
    ROW("#Blob_dmmax\t\t", "\t%.6e", B[i].dmmax, nb)
    ROW("#Blob_eta\t", "\t\t%.3f", B[i].eta, nb)
    ROW("#Blob_C\t\t\t", "\t%.8e", B[i].C, nb)
    ROW("#Blob_Mtot\t\t", "\t%.8e", B[i].Mtot, nb)
    ROW("#Blob_Ngrav\t", "\t\t%ld", (long)B[i].Ngrav, nb)
    ROW("#Blob_Nsmall\t", "\t\t%ld", (long)B[i].Nsmall, nb)
    ROW("#Blob_collcloud\t", "\t\t%d", B[i].collcloud, nb)
    ROW("#Blob_vkepp\t", "\t\t%.5f", B[i].vkepp, nb)
    std::fprintf(fw, "#\n# t=0 positions and velocities of planets and Star\n");
    const int np = c->n_planets;
    ROW("#Planet_x\t", "\t\t%.4f", c->px[i], np)
    ROW("#Planet_y\t", "\t\t%.4f", c->py[i], np)
    ROW("#Planet_z\t", "\t\t%.4f", c->pz[i], np)
    ROW("#Planet_vx\t", "\t\t%.4f", c->pvx[i], np)
    ROW("#Planet_vy\t", "\t\t%.4f", c->pvy[i], np)
    ROW("#Planet_vz\t", "\t\t%.4f", c->pvz[i], np)
    ROW("#Planet_m\t\t", "\t%.6e", c->pm[i], np)
#undef ROW
    std::fprintf(fw, "#NR_used\t\t\t%d\n#\n# Remaining parameters and data are for Grid class read-in\n#\n", kNR);
    std::fprintf(fw, "#Time_of_model\t\t\t%.8f\n#dt_of_model\t\t\t%.8f\n", run->time, run->dt);
    std::fprintf(fw, "#Last_t_of_data\t\t\t%.8f\n#Last_t_of_fits\t\t\t%.8f\n#Last_t_of_SED\t\t\t%.8f\n",
                 run->tlastd, run->tlastf, run->tlasts);
    std::fprintf(fw, "#Number_of_grav_bodies\t\t%d\n#Number_of_dust_particles\t%d\n#Number_of_all_particles\t%d\n",
                 (int)run->n_mass, (int)run->n_dust, (int)run->n_all);
    std::fprintf(fw, "#Data: x[AU] y[AU] z[AU] vx[AU/yr] vy[AU/yr] vz[AU/yr] N m[kg] r[m] beta\n#Planets:\n");
    for (int i = 0; i < np; ++i)
        std::fprintf(fw, "% .8e % .8e % .8e % .8e % .8e % .8e %.12e %.8e %d\n", x[i], y[i], z[i], vx[i], vy[i],
                     vz[i], 1.0, mu[i] * kMSun / kGrav, id ? id[i] : -1);
    std::fprintf(fw, "#Bodies:\n");
    const int64_t n_grav = run->n_all - run->n_dust;
    for (int64_t i = np; i < run->n_all; ++i) {
        const double R = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
        if (R > run->Rmax * run->Rmax) std::fprintf(fw, "## ");
        if (R < c->rstar * c->rstar) std::fprintf(fw, "## ");
        const int32_t idi = id ? id[i] : -1;
        double beta = 0.0;   // grains without a size bin print 0 (the reference reads beta[-1])
        if (beta_p && i >= n_grav) beta = beta_p[i - n_grav];
        else if (beta_bin && idi >= 0 && idi < n_bins) beta = (double)beta_bin[idi];
        const double mkg = mu[i] * kMSun / kGrav;
        std::fprintf(fw, "% .8e % .8e % .8e % .8e % .8e % .8e %.12e %.8e %.8e %.8e %d\n", x[i], y[i], z[i],
                     vx[i], vy[i], vz[i], N[i], mkg,
                     c->rho > 0 ? std::pow(3.0 * mkg / 4.0 / kPI / c->rho, 1.0 / 3.0) : 0.0, beta, idi);
    }
    std::fclose(fw);
    return 0;
}

int ddyn_read_datafile(const char* path, const char* version, ddyn_config* c, ddyn_run_info* run,
                       double* x, double* y, double* z, double* vx, double* vy, double* vz, double* N,
                       double* mu, int32_t* id, double* beta)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return 3;
    std::memset(c, 0, sizeof *c);
    std::memset(run, 0, sizeof *run);
    std::vector<char> buf(1 << 16);
    int i = 0;
    int64_t k = 0;
    bool bodies = false;
    int rc = 0;
    while (std::fgets(buf.data(), (int)buf.size(), f)) {
        const auto d = tokenize(buf.data(), 64);
        if (d.empty()) continue;
        const size_t t = d.size();
        auto v = [&](size_t j) { return j < t ? num(d[j]) : 0.0; };
        auto s = [&](size_t j) { return j < t ? d[j] : std::string(); };
        switch (i) {
            case 1: if (s(1) != version) rc = 55; break;
            case 4: c->verb = (int)v(1); break;
            case 5: c->gpuid = (int)v(1); break;
            case 6: c->gl = (int)v(1); break;
            case 7: c->t_gl = v(1); break;
            case 8: c->t_write = v(1); break;
            case 9: c->t_sed = v(1); break;
            case 10: c->t_fits = v(1); break;
            case 11: c->t_end = v(1); break;
            case 12: c->n_step = (int)v(1); break;
            case 13: copy_str(c->sptype, s(1)); break;
            case 14: c->teff = v(1); break;
            case 15: c->rstar = v(1) * kRSun / kAU; break;
            case 16: c->Bstar = v(1) * 1.0e-4 * kYear * kYear / kMSun; break;
            case 17: c->Wz = v(1) != 0.0 ? 365.2422 / v(1) : 0.0; break;
            case 18: c->mloss = v(1); break;
            case 19: c->vsw = v(1); break;
            case 20: c->radQ = (int)v(1); break;
            case 21: c->SWQ = (int)v(1); break;
            case 22: c->FLQ = (int)v(1); break;
            case 23: c->rho = v(1); break;
            case 24: c->kqe = v(1) / kYear; break;
            case 25: copy_str(c->name_stub, s(1)); break;
            case 26: c->dist = v(1) * kPc; break;
            case 27: c->fov = v(1); break;
            case 28: c->PA = v(1) * kPI / 180.0; break;
            case 29: c->inc = v(1) * kPI / 180.0; break;
            case 30: c->longitude = v(1) * kPI / 180.0; break;
            case 31: c->xsize = (int)v(1); break;
            case 32: c->ysize = (int)v(1); break;
            case 33: copy_str(c->specfile, s(1)); break;
            case 34: c->amin = v(1); break;
            case 35: c->amax = v(1); break;
            case 36: c->mmin = v(1); break;
            case 37: c->mmax = v(1); break;
            case 38: c->n_grid = (int)v(1); break;
            case 39: c->f = v(1); break;
            case 40: copy_str(c->comp, s(1)); break;
            case 41: c->n_wav = (int)v(1); break;
            case 42: c->n_theta = (int)v(1); break;
            case 44: c->n_planets = (int)std::min<double>(v(1), DDYN_MAX_PLANETS); break;
            case 45: c->n_disks = (int)std::min<double>(v(1), DDYN_MAX_DISKS); break;
            case 46: c->n_blobs = (int)std::min<double>(v(1), DDYN_MAX_BLOBS); break;
            case 47:
                c->nwavs = (int)std::min<size_t>(t - 1, DDYN_MAX_WAVS);
                for (int q = 0; q < c->nwavs; ++q) c->wavs[q] = v(q + 1) * 1e-6;
                break;
            case 48: c->seed = (int64_t)v(1); break;
            case 94: if ((int)v(1) != kNR) rc = 56; break;
            case 98: run->time = v(1); break;
            case 99: run->dt = v(1); break;
            case 100: run->tlastd = v(1); break;
            case 101: run->tlastf = v(1); break;
            case 102: run->tlasts = v(1); break;
            case 103: run->n_mass = (int64_t)v(1); break;
            case 104: run->n_dust = (int64_t)v(1); break;
            case 105: run->n_all = (int64_t)v(1); break;
            default: break;
        }
        if (i >= 49 && i <= 64) {
            for (int q = 0; q < c->n_disks && (size_t)(q + 1) < t; ++q) {
                ddyn_disk& D = c->disks[q];
                double* fields[] = {&D.Rin, &D.Rout, &D.dhdr, &D.sigmae, &D.amax, &D.mmax, &D.amin, &D.mmin,
                                    &D.dmax, &D.dmmax, &D.eta, &D.gammas, &D.C, &D.Mtot};
                if (i - 49 < 14) *fields[i - 49] = v(q + 1);
                else if (i == 63) D.Ngrav = (int64_t)v(q + 1);
                else D.Nsmall = (int64_t)v(q + 1);
            }
        }
        if (i >= 65 && i <= 84) {
            for (int q = 0; q < c->n_blobs && (size_t)(q + 1) < t; ++q) {
                ddyn_blob& Bq = c->blobs[q];
                double* fields[] = {&Bq.a, &Bq.ecc, &Bq.inc, &Bq.w, &Bq.o, &Bq.nu, &Bq.sigma, &Bq.amax, &Bq.mmax,
                                    &Bq.amin, &Bq.mmin, &Bq.dmax, &Bq.dmmax, &Bq.eta, &Bq.C, &Bq.Mtot};
                const int r = i - 65;
                if (r < 16) *fields[r] = v(q + 1);
                else if (r == 16) Bq.Ngrav = (int64_t)v(q + 1);
                else if (r == 17) Bq.Nsmall = (int64_t)v(q + 1);
                else if (r == 18) Bq.collcloud = (int)v(q + 1);
                else Bq.vkepp = v(q + 1);
            }
        }
        if (i >= 87 && i <= 93) {
            double* arr[] = {c->px, c->py, c->pz, c->pvx, c->pvy, c->pvz, c->pm};
            for (int q = 0; q < c->n_planets && (size_t)(q + 1) < t; ++q) arr[i - 87][q] = v(q + 1);
        }
        if (rc) break;
        if (i >= 106 && d[0] == "#Bodies:") bodies = true;
        const bool planet_row = i >= 108 && i < 108 + c->n_planets;
        const bool body_row = bodies && d[0] != "#Bodies:";
        if (x && (planet_row || body_row) && k < run->n_all) {
            const size_t o = (d[0] == "##") ? 1 : 0;
            x[k] = v(o); y[k] = v(o + 1); z[k] = v(o + 2);
            vx[k] = v(o + 3); vy[k] = v(o + 4); vz[k] = v(o + 5);
            N[k] = v(o + 6);
            mu[k] = v(o + 7) * kGrav / kMSun;
            if (planet_row) {
                id[k] = (int32_t)v(o + 8);
                if (beta) beta[k] = 0.0;
            } else {
                id[k] = (int32_t)v(o + 10);
                if (beta) beta[k] = v(o + 9);
            }
            ++k;
        }
        ++i;
    }
    std::fclose(f);
    if (rc) return rc;
    if (x && k != run->n_all) return 17;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Images (FITS) and SEDs
// ---------------------------------------------------------------------------
namespace {

constexpr double kCspeed = 299792458.0;

// FITS header cards (the FITS standard, NASA/IAU 4.0: 80-character cards in
// 2880-byte blocks; a fixed-format value right-justified in columns 11-30, a
// string value quoted from column 11, the comment after " / ").
void card(std::string& h, const char* key, const std::string& value, bool quoted, const char* comment)
{
    char buf[96];
    if (quoted) {
        std::string q = "'" + value;
        while (q.size() < 9) q += ' ';   // at least eight characters inside the quotes
        q += "'";
        std::snprintf(buf, sizeof buf, "%-8.8s= %-20s / %s", key, q.c_str(), comment);
    } else {
        std::snprintf(buf, sizeof buf, "%-8.8s= %20s / %s", key, value.c_str(), comment);
    }
    std::string c(buf);
    c.resize(80, ' ');
    h += c;
}

void text_card(std::string& h, const char* key, const char* text)
{
    char buf[96];
    std::snprintf(buf, sizeof buf, "%-8.8s%s", key, text);
    std::string c(buf);
    c.resize(80, ' ');
    h += c;
}

// A real value with `digits` significant digits, always with a decimal point
// or an exponent (a FITS real, not an integer).
std::string real(double v, int digits)
{
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.*G", digits, v);
    std::string s(buf);
    if (s.find_first_of(".EN") == std::string::npos) s += ".";
    return s;
}

void copy_name(char* name, int32_t name_len, const char* s)
{
    if (name && name_len > 0) {
        std::strncpy(name, s, (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
}

}  // namespace

extern "C" {

int ddyn_write_image(const char* version, const ddyn_config* c, double time, double wvl, const double* F,
                     char* name, int32_t name_len)
{
    if (!c || !F || c->xsize < 1 || c->ysize < 1) return 39;
    char fname[DDYN_STR + 64];
    std::snprintf(fname, sizeof fname, "%s_%.2f_%.3f.fits", c->name_stub, wvl * 1e6, time);
    copy_name(name, name_len, fname);
    const int64_t nx = c->xsize, ny = c->ysize;
    // row y of the FITS array holds pixels (0..nx-1, y); the total in the
    // reference's order (x outer, y inner, src/fits.cu:52-58)
    double Ptot = 0.0;
    for (int64_t x = 0; x < nx; ++x)
        for (int64_t y = 0; y < ny; ++y) Ptot += F[ny * x + y];
    const int size = c->xsize >= c->ysize ? c->xsize : c->ysize;
    const float pxscale = (float)(c->fov / c->dist * kPc / size);
    const float lambda = (float)(wvl * 1e6);
    std::string h;
    card(h, "SIMPLE", "T", false, "file does conform to FITS standard");
    card(h, "BITPIX", "-64", false, "number of bits per data pixel");
    card(h, "NAXIS", "2", false, "number of data axes");
    card(h, "NAXIS1", std::to_string(nx), false, "length of data axis 1");
    card(h, "NAXIS2", std::to_string(ny), false, "length of data axis 2");
    card(h, "EXTEND", "T", false, "FITS dataset may contain extensions");
    card(h, "WAVELENG", real(lambda, 7), false, "[micron] Wavelength of dataset");
    card(h, "FLUX", real(Ptot, 15), false, "[Jy] Total flux detected");
    card(h, "PIXELSCA", real(pxscale, 7), false, "[''/px] pixel scale of instrument setup");
    card(h, "VERSION", version ? version : "", true, "DiskDyn version");
    text_card(h, "HISTORY", "Created with DiskDyn");
    {
        char date[32];
        const std::time_t now = std::time(nullptr);
        std::tm tmv{};
        gmtime_r(&now, &tmv);
        std::strftime(date, sizeof date, "%Y-%m-%dT%H:%M:%S", &tmv);
        card(h, "DATE", date, true, "file creation date (YYYY-MM-DDThh:mm:ss UT)");
    }
    text_card(h, "END", "");
    h.resize((h.size() + 2879) / 2880 * 2880, ' ');
    // data: FP64 big-endian, NAXIS1 (x) fastest, zero-padded to whole blocks
    std::vector<unsigned char> data((size_t)(nx * ny * 8 + 2879) / 2880 * 2880, 0);
    for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x) {
            uint64_t u;
            std::memcpy(&u, &F[ny * x + y], 8);
            unsigned char* d = &data[(size_t)(y * nx + x) * 8];
            for (int b = 0; b < 8; ++b) d[b] = (unsigned char)(u >> (56 - 8 * b));
        }
    std::remove(fname);
    FILE* fw = std::fopen(fname, "wb");
    if (!fw) return 39;
    const bool ok = std::fwrite(h.data(), 1, h.size(), fw) == h.size() &&
                    std::fwrite(data.data(), 1, data.size(), fw) == data.size();
    return (std::fclose(fw) == 0 && ok) ? 0 : 39;
}

int ddyn_write_sed(const char* version, const ddyn_config* c, double time, int32_t n_wav, const double* w,
                   const double* in, const double* Fexc, char* name, int32_t name_len)
{
    if (!c || n_wav < 0 || (n_wav > 0 && (!w || !in || !Fexc))) return 39;
    char fname[DDYN_STR + 64];
    std::snprintf(fname, sizeof fname, "%s_SED_%.2f.dat", c->name_stub, time);
    copy_name(name, name_len, fname);
    FILE* fw = std::fopen(fname, "wb");
    if (!fw) return 39;
    std::fprintf(fw, "#Calculated with DiskDyn Version %s\n", version ? version : "");
    std::fprintf(fw, "#Wav[mu]\tFlux_star[Jy]\tFlux_exc[Jy]\n");
    for (int32_t j = 0; j < n_wav; ++j)
        std::fprintf(fw, "%.3f %.6e %.6e\n", w[j] * 1e6, in[j] * 1e26 * w[j] * w[j] / kCspeed, Fexc[j]);
    return std::fclose(fw) == 0 ? 0 : 39;
}

}  // extern "C"
