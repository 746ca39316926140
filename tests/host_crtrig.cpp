// Host build of the device sin/cos source (gargantua_amd/csrc/gg_crtrig.hpp)
// for the CPU test that checks it against correctly rounded values.
#include <stdint.h>

#include "gg_crtrig.hpp"

extern "C" void host_sincos_cr(uint64_t n, const double *x, double *s, double *c) {
  for (uint64_t i = 0; i < n; ++i) gg::sincos_cr(x[i], s + i, c + i);
}

extern "C" void host_libm_sincos(uint64_t n, const double *x, double *s, double *c) {
  for (uint64_t i = 0; i < n; ++i) { s[i] = sin(x[i]); c[i] = cos(x[i]); }
}
