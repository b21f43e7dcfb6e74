// ddyn_main.cpp -- command-line driver of the dynamical hot path on one GPU.
//
//   ddyn -i <file.param> [options]   start from a .param file: star, planets and the
//                                    particles its disks and blobs generate
//   ddyn -m <file.dat>   [options]   continue from a data file (every particle)
//
// It replaces the dynamics of the reference's main + Grid::setup / read_in +
// Grid::evolve (src/ddyn.cu:37-111, src/grid.cu:62-310, 344-439): the optical
// data and the grain tables of the dust grid (with grains), the particles of
// the disks and blobs, the same time step (det_step),
// the same clipping of steps to the data/FITS/SED output times, data files
// <stub>_<time>.dat written at the data cadence and at the end, and with the
// optical data the images <stub>_<wavelength>_<time>.fits at the FITS cadence
// and the SEDs <stub>_SED_<time>.dat at the SED cadence (src/grid.cu:344-439:
// images, SED and data file of the initial state of a .param run; then at each
// output time the SED, the images and the data file, in that order).
//
// Options: --fp32 (FP32-force form), --fp64-rsq (FP64 with refined hardware
//          reciprocal square roots in every pair law), --eps <AU> (softening of
//          swarm pairs), --r64 <AU> (FP32-force form: swarm pairs closer than
//          this are evaluated in FP64; default 10 eps, or 0.01 AU without
//          softening), --rmin <AU> (sublimation radius), --no-images (data
//          files only), --device <n>, --quiet.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ddyn.h"
#include "ddyn_host.h"

namespace {

constexpr const char* kVersion = "1.01";
constexpr double kAU = 1.49598e11;
constexpr double kYear = 3.1556926e7;

void usage()
{
    std::fprintf(stderr,
                 "Usage: ddyn -i <init.param> | -m <model.dat> [--fp32 | --fp64-rsq] [--eps AU] [--r64 AU] [--rmin AU] "
                 "[--no-images] [--device n] [--quiet]\n");
}

struct State {
    int64_t n_grav = 0, n_dust = 0;
    std::vector<double> x, y, z, vx, vy, vz, N, mu, beta;
    std::vector<int32_t> id;
    void resize(int64_t n)
    {
        for (auto* v : {&x, &y, &z, &vx, &vy, &vz, &N, &mu, &beta}) v->assign((size_t)n, 0.0);
        id.assign((size_t)n, -1);
    }
};

int check(ddyn_ctx* ctx, int rc, const char* what)
{
    if (rc) std::fprintf(stderr, "ddyn: %s failed (%d): %s\n", what, rc, ddyn_last_error(ctx));
    return rc;
}

}  // namespace

int main(int argc, char** argv)
{
    const char* param = nullptr;
    const char* model = nullptr;
    int precision = DDYN_FP64, device = 0, quiet = 0, images = 1;
    double eps = 0.0, rmin = 0.0, r64 = 0.0;
    for (int a = 1; a < argc; ++a) {
        const std::string s = argv[a];
        if (s == "-i" && a + 1 < argc) param = argv[++a];
        else if (s == "-m" && a + 1 < argc) model = argv[++a];
        else if (s == "--fp32") precision = DDYN_FP32_FORCE;
        else if (s == "--fp64-rsq") precision = DDYN_FP64_RSQ;
        else if (s == "--eps" && a + 1 < argc) eps = std::atof(argv[++a]);
        else if (s == "--rmin" && a + 1 < argc) rmin = std::atof(argv[++a]);
        else if (s == "--r64" && a + 1 < argc) r64 = std::atof(argv[++a]);
        else if (s == "--device" && a + 1 < argc) device = std::atoi(argv[++a]);
        else if (s == "--quiet") quiet = 1;
        else if (s == "--no-images") images = 0;
        else { usage(); return s == "-h" ? 0 : 1; }
    }
    if (!param == !model) { usage(); return 1; }

    static ddyn_config cfg;
    ddyn_run_info run{};
    State st;
    int rc;
    bool param_mode = param != nullptr;
    // grain coefficient tables of the dust grid (from the optical data) and the
    // sublimation radius they give
    std::vector<float> tb, tbsw, tq;
    double rmin_optics = 0.0;
    bool have_tables = false;
    ddyn_optics* optics = nullptr;
    // the dust grid's optical tables for images and SEDs, and the image wavelengths
    ddyn_grid_optics* gopt = nullptr;
    std::vector<int32_t> wj;
    auto load_imaging = [&]() {
        if (!images || !optics) return;
        int r = ddyn_grid_optics_make(&cfg, optics, &gopt);
        if (!r) {
            wj.assign((size_t)std::max(cfg.nwavs, 1), 0);
            r = ddyn_image_wavelengths(&cfg, gopt, wj.data());
            wj.resize((size_t)std::max(cfg.nwavs, 0));
        }
        if (r) {
            std::fprintf(stderr, "ddyn: images and SEDs: %s; not written\n", ddyn_host_error_text(r));
            ddyn_grid_optics_free(gopt);
            gopt = nullptr;
        }
    };
    auto load_tables = [&]() -> int {
        const size_t n = (size_t)std::max(cfg.n_grid, 0);
        tb.assign(n, 0.f); tbsw.assign(n, 0.f); tq.assign(n, 0.f);
        double lum = 0.0;
        const int r = ddyn_grain_tables(&cfg, optics, cfg.n_grid, tb.data(), tbsw.data(), tq.data(), nullptr,
                                        &rmin_optics, &lum);
        if (r) return r;
        if (!quiet)
            std::printf("ddyn: %d size bins, L = %.4g Lsun, sublimation radius %.5g AU\n", cfg.n_grid, lum,
                        rmin_optics);
        have_tables = true;
        return 0;
    };
    if (param_mode) {
        if ((rc = ddyn_read_param(param, &cfg))) {
            std::fprintf(stderr, "ddyn: %s: %s\n", param, ddyn_host_error_text(rc));
            return 1;
        }
        bool generate = false;
        for (int k = 0; k < cfg.n_disks; ++k) generate = generate || cfg.disks[k].Ngrav || cfg.disks[k].Nsmall;
        for (int k = 0; k < cfg.n_blobs; ++k) generate = generate || cfg.blobs[k].Ngrav || cfg.blobs[k].Nsmall;
        if (generate) {
            // optical data, dust grid and grain tables (src/ddyn.cu:78-96, src/grid.cu:66-70)
            if ((rc = ddyn_optics_read(cfg.comp, cfg.specfile, &cfg, &optics))) {
                std::fprintf(stderr, "ddyn: optical data %s / spectrum %s: %s\n", cfg.comp, cfg.specfile,
                             ddyn_host_error_text(rc));
                return 1;
            }
            if ((rc = ddyn_set_dust_grid(&cfg, optics)) || (rc = load_tables())) {
                std::fprintf(stderr, "ddyn: dust grid: %s\n", ddyn_host_error_text(rc));
                ddyn_optics_free(optics);
                return 1;
            }
            load_imaging();
        }
        int64_t n_mass = 0, n_dust = 0;
        ddyn_generate_counts(&cfg, &n_mass, &n_dust);
        st.n_grav = cfg.n_planets + n_mass;
        st.n_dust = n_dust;
        st.resize(st.n_grav + st.n_dust);
        rc = ddyn_generate(&cfg, optics, have_tables ? tb.data() : nullptr, have_tables ? cfg.n_grid : 0,
                           st.n_grav + st.n_dust, st.x.data(), st.y.data(), st.z.data(), st.vx.data(), st.vy.data(),
                           st.vz.data(), st.N.data(), st.mu.data(), st.id.data());
        ddyn_optics_free(optics);
        optics = nullptr;
        if (rc) {
            std::fprintf(stderr, "ddyn: particle generation: %s\n", ddyn_host_error_text(rc));
            return 1;
        }
        if (!quiet && generate)
            std::printf("ddyn: generated %lld gravitating bodies and %lld grain tracers\n", (long long)n_mass,
                        (long long)n_dust);
        run.n_mass = n_mass; run.n_dust = n_dust; run.n_all = st.n_grav + st.n_dust;
    } else {
        if ((rc = ddyn_read_datafile(model, kVersion, &cfg, &run, nullptr, nullptr, nullptr, nullptr, nullptr,
                                     nullptr, nullptr, nullptr, nullptr, nullptr))) {
            std::fprintf(stderr, "ddyn: %s: %s\n", model, ddyn_host_error_text(rc));
            return 1;
        }
        st.resize(run.n_all);
        if ((rc = ddyn_read_datafile(model, kVersion, &cfg, &run, st.x.data(), st.y.data(), st.z.data(),
                                     st.vx.data(), st.vy.data(), st.vz.data(), st.N.data(), st.mu.data(),
                                     st.id.data(), st.beta.data()))) {
            std::fprintf(stderr, "ddyn: %s: %s\n", model, ddyn_host_error_text(rc));
            return 1;
        }
        st.n_dust = run.n_dust;
        st.n_grav = run.n_all - run.n_dust;
        if (st.n_dust > 0) {
            // as the reference's restart (src/grid.cu:285-296): the grain tables
            // from the optical data on the file's dust grid, the bins from the masses
            if (ddyn_optics_read(cfg.comp, cfg.specfile, &cfg, &optics) == 0 && load_tables() == 0) {
                ddyn_restart_ids(&cfg, run.n_all, st.mu.data(), st.id.data());
                load_imaging();
            } else if (cfg.SWQ == 1 || cfg.FLQ == 1) {
                // the data file carries beta per grain only: without the optical
                // data the wind drag and the Lorentz force would silently be zero
                std::fprintf(stderr, "ddyn: %s has stellar-wind drag or the Lorentz force on, which need "
                                     "beta_sw and charge tables that a data file does not carry, and the optical "
                                     "data %s / spectrum %s cannot be read\n", model, cfg.comp, cfg.specfile);
                ddyn_optics_free(optics);
                return 2;
            } else {
                std::fprintf(stderr, "ddyn: warning: optical data %s / spectrum %s cannot be read: per-grain beta "
                                     "from the data file%s\n", cfg.comp, cfg.specfile,
                             rmin > 0.0 ? "" : ", and no sublimation radius (--rmin)");
            }
            ddyn_optics_free(optics);
            optics = nullptr;
        }
    }
    run.Rmax = ddyn_outer_radius(&cfg);
    if (run.Rmax <= 0.0) run.Rmax = 1e300;

    ddyn_ctx* ctx = nullptr;
    if ((rc = ddyn_create(&ctx, device))) {
        std::fprintf(stderr, "ddyn: no usable GPU (device %d, code %d)\n", device, rc);
        return 1;
    }
    ddyn_params prm{};
    prm.n_planets = cfg.n_planets;
    prm.SWQ = cfg.SWQ;
    prm.FLQ = cfg.FLQ;
    prm.precision = precision;
    prm.eps = eps;
    prm.Rmin = rmin > 0.0 ? rmin : (have_tables ? rmin_optics : 0.0);
    prm.vsw = cfg.vsw * kYear / kAU;
    prm.Wz = cfg.Wz;
    prm.Bstar = cfg.Bstar;
    prm.Rstar = cfg.rstar;
    prm.light = 0.0;
    prm.fp64_radius = r64;
    if (precision == DDYN_FP32_FORCE && eps == 0.0 && !(r64 > 0.0)) {
        // unsoftened pairs: an FP64 radius of its own (10 eps would be 0)
        prm.fp64_radius = 0.01;
        if (!quiet) std::printf("ddyn: FP32-force form without softening: pairs within 0.01 AU in FP64\n");
    }
    if (check(ctx, ddyn_set_params(ctx, &prm), "set_params") ||
        check(ctx, ddyn_upload_state(ctx, st.n_grav, st.n_dust, st.x.data(), st.y.data(), st.z.data(), st.vx.data(),
                                     st.vy.data(), st.vz.data(), st.mu.data(), st.N.data(), st.id.data()),
              "upload_state") ||
        check(ctx, ddyn_set_params(ctx, &prm), "set_params")) {
        ddyn_destroy(ctx);
        return 1;
    }
    if (have_tables) {   // the reference's per-bin float tables (src/ddyn.cu:29-32)
        if (check(ctx, ddyn_upload_bin_tables(ctx, cfg.n_grid, tb.data(), tbsw.data(), tq.data()),
                  "upload_bin_tables")) {
            ddyn_destroy(ctx);
            return 1;
        }
    } else if (st.n_dust) {   // beta per grain from the data file's beta column
        std::vector<double> b(st.beta.begin() + st.n_grav, st.beta.end());
        if (check(ctx, ddyn_upload_grain_beta(ctx, b.data(), nullptr), "upload_grain_beta")) {
            ddyn_destroy(ctx);
            return 1;
        }
    }
    ddyn_optics_tables otab{};
    if (gopt) {
        ddyn_grid_optics_tables(gopt, &otab);
        if (check(ctx, ddyn_upload_optics(ctx, &otab), "upload_optics")) {
            ddyn_destroy(ctx);
            ddyn_grid_optics_free(gopt);
            return 1;
        }
    }
    ddyn_view view{};
    view.dist = cfg.dist;
    view.rot = cfg.longitude;
    view.incl = cfg.inc;
    view.PA = cfg.PA;
    view.fov = cfg.fov;
    view.xsize = cfg.xsize;
    view.ysize = cfg.ysize;
    std::vector<double> img, sed;
    // Grid::calc_flux + write_image per image wavelength (src/grid.cu:43-54, src/fits.cu)
    auto write_images = [&](double time) -> int {
        if (!gopt) return 0;
        img.resize((size_t)std::max<int64_t>((int64_t)view.xsize * view.ysize, 1));
        for (const int32_t j : wj) {
            if (check(ctx, ddyn_calc_flux(ctx, &view, j, img.data()), "calc_flux")) return 1;
            char name[600];
            const int r = ddyn_write_image(kVersion, &cfg, time, otab.w[j], img.data(), name, (int32_t)sizeof name);
            if (r) { std::fprintf(stderr, "ddyn: %s: %s\n", name, ddyn_host_error_text(r)); return 1; }
            if (!quiet) std::printf("ddyn: wrote %s\n", name);
        }
        return 0;
    };
    // Grid::calc_flux_sed + copy_sed per wavelength, write_sed (src/grid.cu:56-65, 675-691)
    auto write_sed = [&](double time) -> int {
        if (!gopt) return 0;
        sed.resize((size_t)otab.n_wav);
        if (check(ctx, ddyn_calc_sed(ctx, &view, sed.data()), "calc_sed")) return 1;
        char name[600];
        const int r = ddyn_write_sed(kVersion, &cfg, time, otab.n_wav, otab.w, otab.in, sed.data(), name,
                                     (int32_t)sizeof name);
        if (r) { std::fprintf(stderr, "ddyn: %s: %s\n", name, ddyn_host_error_text(r)); return 1; }
        if (!quiet) std::printf("ddyn: wrote %s\n", name);
        return 0;
    };
    auto finish = [&](int code) {
        ddyn_destroy(ctx);
        ddyn_grid_optics_free(gopt);
        return code;
    };
    auto write = [&](double time) -> int {
        if (check(ctx, ddyn_download_state(ctx, st.x.data(), st.y.data(), st.z.data(), st.vx.data(), st.vy.data(),
                                           st.vz.data(), st.N.data()), "download_state"))
            return 1;
        char name[600];
        std::snprintf(name, sizeof name, "%s_%.2f.dat", cfg.name_stub[0] ? cfg.name_stub : "ddyn", time);
        const double* bp = (st.n_dust && !have_tables) ? st.beta.data() + st.n_grav : nullptr;
        int r = ddyn_write_datafile(name, kVersion, &cfg, &run, st.x.data(), st.y.data(), st.z.data(),
                                    st.vx.data(), st.vy.data(), st.vz.data(), st.N.data(), st.mu.data(),
                                    st.id.data(), have_tables ? tb.data() : nullptr, have_tables ? cfg.n_grid : 0,
                                    bp);
        if (r) std::fprintf(stderr, "ddyn: %s: %s\n", name, ddyn_host_error_text(r));
        else if (!quiet) std::printf("ddyn: wrote %s\n", name);
        return r;
    };

    // Grid::evolve (src/grid.cu:344-439)
    const double dt = ddyn_det_step(&cfg, st.n_grav, st.x.data(), st.y.data(), st.z.data(), st.mu.data());
    run.dt = dt;
    double dttmp = dt;
    if (!quiet) std::printf("ddyn: dt = %.8e yr, t_end = %.3f yr\n", dt, cfg.t_end);
    if (param_mode && (write_images(run.time) || write_sed(run.time) || write(run.time))) return finish(1);
    auto clip = [&]() {
        if (run.time + dttmp > run.tlastd + cfg.t_write) dttmp = cfg.t_write + run.tlastd - run.time;
        if (run.time + dttmp > run.tlastf + cfg.t_fits) dttmp = cfg.t_fits + run.tlastf - run.time;
        if (run.time + dttmp > run.tlasts + cfg.t_sed) dttmp = cfg.t_sed + run.tlasts - run.time;
        if (run.time + dttmp > cfg.t_end) dttmp = cfg.t_end - run.time;
    };
    clip();
    int64_t steps = 0;
    // The time bookkeeping is host arithmetic only, so consecutive steps of the
    // same size with no output between them go to the GPU as one ddyn_step
    // call (same steps, same results, and the library can overlap or fuse them).
    constexpr int64_t kMaxRun = 4096;
    do {
        const double h = dttmp;
        int64_t n = 0;
        bool out_sed = false, out_fits = false, out_data = false;
        do {
            ++n;
            run.time += h;
            if (run.time >= run.tlasts + cfg.t_sed || run.time >= cfg.t_end) {
                run.tlasts = run.time;
                out_sed = true;
                dttmp = dt;
            }
            if (run.time >= run.tlastf + cfg.t_fits || run.time >= cfg.t_end) {
                run.tlastf = run.time;
                out_fits = true;
                dttmp = dt;
            }
            if (run.time >= run.tlastd + cfg.t_write || run.time >= cfg.t_end) {
                run.tlastd = run.time;
                out_data = true;
                dttmp = dt;
            }
            clip();
        } while (!out_sed && !out_fits && !out_data && dttmp == h && run.time < cfg.t_end && n < kMaxRun);
        if (check(ctx, ddyn_step(ctx, h, n), "step")) return finish(1);
        steps += n;
        if (out_sed && write_sed(run.time)) return finish(1);
        if (out_fits && write_images(run.time)) return finish(1);
        if (out_data && write(run.time)) return finish(1);
    } while (run.time < cfg.t_end);
    if (!quiet) std::printf("ddyn: %lld steps to t = %.6f yr\n", (long long)steps, run.time);
    return finish(0);
}
