// Host build of the device VJP source (gargantua_amd/csrc/gg_vjp.hpp) for
// the CPU test that checks it against the oracle's dual-number Jacobian.
#include <stdint.h>

#include "gg_vjp.hpp"

extern "C" void host_kerr_rhs_vjp(double m, double a, uint64_t n, const double *pos, const double *vel,
                                  const double *cot, double *out7) {
  for (uint64_t i = 0; i < n; ++i) {
    gg::d4 v{vel[4 * i], vel[4 * i + 1], vel[4 * i + 2], vel[4 * i + 3]};
    gg::d4 w{cot[4 * i], cot[4 * i + 1], cot[4 * i + 2], cot[4 * i + 3]};
    gg::RhsBar b = gg::kerr_rhs_vjp(m, a, pos[4 * i + 1], pos[4 * i + 2], v, w);
    double *o = out7 + 7 * i;
    o[0] = b.r; o[1] = b.th; o[2] = b.v.x; o[3] = b.v.y; o[4] = b.v.z; o[5] = b.v.w; o[6] = b.a;
  }
}

extern "C" int host_sincos_near(double s0, double c0, double d, double *out2) {
  return gg::sincos_near(s0, c0, d, out2[0], out2[1]) ? 1 : 0;
}
