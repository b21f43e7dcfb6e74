// ddyn_host_impl.h -- helpers shared by the host library's sources
// (ddyn_host.cpp, ddyn_setup.cpp); not part of the C ABI.
#ifndef DDYN_HOST_IMPL_H
#define DDYN_HOST_IMPL_H
#include <cstddef>
#include <string>
#include <vector>

namespace ddyn_impl {

// Whitespace tokens of a line (double-quoted tokens keep their spaces), at most maxtok.
std::vector<std::string> tokenize(const char* line, size_t maxtok);

// Orbital elements about a central mass M (with the orbiting mass m and the
// radiation-pressure ratio beta: mu = (M + m)(1 - beta)) to x, y, z, vx, vy, vz.
void coe2rv(double M, double m, double a, double e, double i, double w, double o, double nu, double beta,
            double* r);

}  // namespace ddyn_impl
#endif
