// Accuracy of the Box-Muller log and sin/cos of the device noise generator
// (sopg_device.h bm_log, bm_sincos2pi) against long-double libm, over 4M
// arguments of the generator's form.  Prints max ulp errors of log, sin, cos
// and the largest absolute sin/cos error.
#include "sopg_device.h"
#include <cstdio>
#include <cmath>
#include <random>
using namespace sopg;
static double ulp_err(double got, long double ref) {
  const double r = (double)ref;
  const double u = std::fabs(std::nextafter(r, INFINITY) - r);
  return (double)(std::fabs((long double)got - ref) / u);
}
int main() {
  std::mt19937_64 g(7);
  double mlog = 0, ms = 0, mc = 0, mabs = 0;
  for (long i = 0; i < 4000000; i++) {
    uint64_t a = g(), b = g();
    double u1 = (double)((a >> 11) + 1) * (1.0 / 9007199254740992.0);
    if (i < 64) u1 = std::ldexp(1.0, -(int)i);  // powers of two down to 2^-63 (u1 >= 2^-53 in use)
    if (u1 < 1.1102230246251565e-16) u1 = 1.1102230246251565e-16;
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    if (i < 16) u2 = i / 16.0;
    double l = bm_log(u1);
    if (u1 != 1.0) mlog = std::max(mlog, ulp_err(l, logl((long double)u1)));
    double s, c;
    bm_sincos2pi(u2, s, c);
    long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)u2;
    long double rs = sinl(ang), rc = cosl(ang);
    mabs = std::max(mabs, (double)std::max(std::fabs((long double)s - rs), std::fabs((long double)c - rc)));
    if (std::fabs(rs) > 1e-3) ms = std::max(ms, ulp_err(s, rs));
    if (std::fabs(rc) > 1e-3) mc = std::max(mc, ulp_err(c, rc));
  }
  std::printf("%.3f %.3f %.3f %.3e\n", mlog, ms, mc, mabs);
}
