// ddyn_runmie.cpp -- the Mie-table tool: refractive indices (.lnk: lines
// "<wavelength um> <n> <k>") in, a composition table (the format
// ddyn_optics_read and the reference's read_in read) out.  Replaces the
// reference's tools/runmie (tools/runmie.c:134-224) with the same options,
// defaults and output layout; the Mie sums run on the GPU (ddyn_mie).
//
//   runmie -Ntheta <angles> -Ndust <sizes> -amin <um> -amax <um>
//          -lnkfile <file.lnk> -output <file.dat> [--device n] [--quiet]
//
// As in the reference, the wavelengths, indices, sizes and scattering cosines
// are single precision values (the sums are FP64), negative efficiencies are
// written as 0, and the phase-function block of each size holds
// floor(n_theta * n_lnk / 10) lines of ten values.  The absorption efficiency
// is the FP64 difference Qext - Qsca; the reference rounds Qext to single
// precision first (tools/mie_single.c:245), which shifts small Qabs by up to
// 6e-8 Qext.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ddyn.h"

namespace {

void help(FILE* f)
{
    std::fprintf(f,
                 "Usage:\n\trunmie -Ntheta <number of angles> -Ndust <number of dust sizes>\n"
                 "\t       -amin <min dust size in micron> -amax <max dust size in micron>\n"
                 "\t       -lnkfile <L,N,K input file> -output <output file name> [--device n] [--quiet]\n");
}

int fail(const char* msg)
{
    std::fprintf(stderr, "ERROR:\t%s\n\n", msg);
    help(stderr);
    return 1;
}

// "<wavelength> <n> <k>" lines, '#' lines and short lines skipped.
bool read_lnk(const char* path, std::vector<float>& l, std::vector<float>& n, std::vector<float>& k)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[512];
    while (std::fgets(buf, sizeof buf, f)) {
        if (buf[0] == '#') continue;
        double a, b, c;
        if (std::sscanf(buf, "%lf %lf %lf", &a, &b, &c) == 3) {
            l.push_back((float)a);
            n.push_back((float)b);
            k.push_back((float)c);
        }
    }
    std::fclose(f);
    return true;
}

}  // namespace

int main(int argc, char** argv)
{
    int ntheta = 90, ndust = 250, device = 0, quiet = 0;
    float amin = 0.001f, amax = 1e5f;
    const char* lnk = nullptr;
    const char* out = nullptr;
    if (argc == 1) return fail("Missing input parameter(s)");
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> const char* {
            if (++i == argc) { fail(what); std::exit(1); }
            return argv[i];
        };
        if (a == "-h") { help(stdout); return 0; }
        else if (a == "-Ntheta") ntheta = std::atoi(next("Missing number of angles"));
        else if (a == "-Ndust") ndust = std::atoi(next("Missing number of dust particles"));
        else if (a == "-amin") amin = (float)std::atof(next("Missing minimum dust size"));
        else if (a == "-amax") amax = (float)std::atof(next("Missing maximum dust size"));
        else if (a == "-lnkfile") lnk = next("Missing LNK file");
        else if (a == "-output") out = next("Missing output filename");
        else if (a == "--device") device = std::atoi(next("Missing device"));
        else if (a == "--quiet") quiet = 1;
        else return fail("Missing input parameter(s)");
    }
    if (!lnk) return fail("Missing LNK file");
    if (!out) return fail("Missing output filename");
    if (ntheta < 2 || ntheta > 256 || ndust < 2 || !(amin > 0.0f) || !(amax > amin))
        return fail("need 2 <= Ntheta <= 256, Ndust >= 2 and 0 < amin < amax");
    std::vector<float> l, n, k;
    if (!read_lnk(lnk, l, n, k) || l.empty()) {
        std::fprintf(stderr, "runmie: cannot read %s\n", lnk);
        return 1;
    }
    const int nl = (int)l.size();
    // sizes and cosines as the reference forms them (tools/runmie.c:195-206)
    std::vector<float> s((size_t)ndust), cst((size_t)ntheta);
    // (the logarithm of the single-precision ratio in double precision, as C's log)
    for (int z = 0; z < ndust; ++z)
        s[(size_t)z] = (float)(amin * std::exp(std::log((double)(amax / amin)) * z / ((float)ndust - 1.0)));
    for (int i = 0; i < ntheta; ++i) cst[(size_t)i] = (float)((double)i * 2.0 / (ntheta - 1) - 1.0);
    std::vector<double> sd(s.begin(), s.end()), ld(l.begin(), l.end()), nd(n.begin(), n.end()),
        kd(k.begin(), k.end()), cd(cst.begin(), cst.end());
    const size_t np = (size_t)ndust * nl;
    std::vector<double> qa(np), qs(np), qp(np * (size_t)ntheta);
    ddyn_ctx* ctx = nullptr;
    if (ddyn_create(&ctx, device)) {
        std::fprintf(stderr, "runmie: no usable GPU (device %d)\n", device);
        return 1;
    }
    const int rc = ddyn_mie(ctx, ndust, sd.data(), nl, ld.data(), nd.data(), kd.data(), ntheta, cd.data(), qa.data(),
                            qs.data(), qp.data());
    if (rc) {
        std::fprintf(stderr, "runmie: %s\n", ddyn_last_error(ctx));
        ddyn_destroy(ctx);
        return 1;
    }
    ddyn_destroy(ctx);
    FILE* fw = std::fopen(out, "wb");
    if (!fw) {
        std::fprintf(stderr, "runmie: cannot write %s\n", out);
        return 1;
    }
    auto f32 = [](double v) { const float x = (float)v; return x < 0.0f ? 0.0f : x; };
    std::fprintf(fw, "#\tn_lnk\tnthetas\tndust\tamin\t\tamax\n");
    std::fprintf(fw, "#\t%d\t%d\t%d\t%.6e\t%.6e\n", nl, ntheta, ndust, amin, amax);
    for (int z = 0; z < ndust; ++z) {
        if (!quiet) std::printf("Working on dust size %.6f micron\n", s[(size_t)z]);
        std::fprintf(fw, "#s %.6e\n", s[(size_t)z]);
        const size_t b = (size_t)z * nl;
        for (int i = 0; i < nl; ++i)
            std::fprintf(fw, "%.6e %.6e %.6e\n", l[(size_t)i], f32(qa[b + i]), f32(qs[b + i]));
        // theta-major, ten values per line
        const double* pf = qp.data() + (size_t)z * ntheta * nl;
        for (long i = 0; i < (long)ntheta * nl / 10; ++i) {
            for (int j = 0; j < 10; ++j) std::fprintf(fw, "%.6e ", f32(pf[i * 10 + j]));
            std::fprintf(fw, "\n");
        }
    }
    if (std::fclose(fw) != 0) {
        std::fprintf(stderr, "runmie: cannot write %s\n", out);
        return 1;
    }
    return 0;
}
