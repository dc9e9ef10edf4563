// Host-side math of the engine that has to reproduce the reference CPU
// path's libm values bit for bit. Compiled by the host C++ compiler
// (host_math.cpp), not in HIP mode: HIP's host headers replace
// std::exp(std::complex<double>) with an implementation that differs from
// the C library's cexp in the last bit for some arguments.
#pragma once

#include <cstddef>

namespace tbtkx {

// Factors of the direct contraction (generateGreensFunctionCPU,
// Lib/src/Solver/ChebyshevExpander.cpp:386-430) for n = 1 .. M-1 and the
// energies e0 + e*de (e < E), scale s: kind 0 writes exp(-i n acos(E~)) as
// (re, im) pairs at fac[2((n-1)E + e)], kind 1 sin(n acos(E~)) and kind 2
// cos(n acos(E~)) at fac[(n-1)E + e]; den[e] = sqrt(1 - E~^2), E~ =
// (e0 + e*de)/s. Runs on up to 16 host threads.
void direct_factor_table(int kind, int M, int E, double e0, double de, double s, double *fac,
                         double *den);

}  // namespace tbtkx
