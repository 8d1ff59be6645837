// sopg_device.h -- FP64 helpers shared by the gfx950 kernels (and by host
// checks of the same routines).  Compiled with -ffp-contract=off: every
// expression below is evaluated exactly as written, one IEEE rounding per
// operation, which is what makes the kernels reproduce SOP-CPU's arithmetic.
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define SOPG_HD __host__ __device__ __forceinline__
#else
#define SOPG_HD inline
#endif

namespace sopg {

// rnd(x): round half away from zero through floor/ceil (global.cpp:223-227).
SOPG_HD double rnd(double x) { return (x > 0) ? floor(x + 0.5) : ceil(x - 0.5); }

// Min-image component d - boxl*rnd(d/boxl) (e.g. energy.cpp:287-289).
// Exact fast path: for |d| <= boxl/4 the quotient lies in [-1/4, 1/4]
// (division is monotone and boxl/4 is exact), rnd returns +0 or -0, and
// d - boxl*(+-0) equals d + 0.0 (which maps -0 to +0, as the full form does).
// The division is skipped only there; everything else takes the full form.
SOPG_HD double mimg(double d, double boxl, double qbox) {
  if (fabs(d) <= qbox) return d + 0.0;
  return d - boxl * rnd(d / boxl);
}

// Correctly rounded d^8 and d^6 by double-double squaring.  The reference
// calls glibc pow(d,8.0) / pow(d,6.0) (energy.cpp:382, 180); glibc's pow is
// accurate to ~0.52 ulp, so it differs from the correctly rounded value by one
// ulp on ~0.08% of inputs.  fma gives the exact low part of each product.
SOPG_HD void two_prod(double a, double b, double& p, double& e) { p = a * b; e = fma(a, b, -p); }
SOPG_HD void dd_sq(double h, double l, double& oh, double& ol) {
  double p, e;
  two_prod(h, h, p, e);
  double c = h * l;
  e += c + c;
  double s = p + e;
  ol = e - (s - p);
  oh = s;
}
SOPG_HD double pow8_cr(double d) {
  double h2, l2, h4, l4, h8, l8;
  dd_sq(d, 0.0, h2, l2);
  dd_sq(h2, l2, h4, l4);
  dd_sq(h4, l4, h8, l8);
  return h8;
}
SOPG_HD double pow6_cr(double d) {
  double h2, l2, h4, l4, p, e;
  dd_sq(d, 0.0, h2, l2);
  dd_sq(h2, l2, h4, l4);
  two_prod(h4, h2, p, e);
  e += h4 * l2 + l4 * h2;
  return p + e;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based, so every
// (chain seed, force evaluation, bead, pair) has its own independent block.
SOPG_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += W0; k1 += W1;
  }
}

// The two transcendental functions of the Box-Muller transform, for its
// arguments only and to within a few ulp (noise needs accuracy, not correct
// rounding), in about a third of the instructions of the general library
// routines.  Textbook series, Horner form with fused multiply-adds; the same
// bits on the host and the device.
//
// log u for u in [2^-53, 1]: u = m 2^e with m in [sqrt(1/2), sqrt(2)), and
// log m = 2 atanh(s) = 2 s (1 + s^2/3 + s^4/5 + ...), s = (m - 1) / (m + 1),
// |s| <= 0.1716; the terms up to s^22 / 23 leave a truncation below 1e-18.
SOPG_HD double bm_log(double u) {
  int e;
  double m = frexp(u, &e);  // [0.5, 1)
  if (m < 0.70710678118654752) {
    m = 2.0 * m;
    e -= 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 1.0 / 23.0;
  p = fma(p, z, 1.0 / 21.0); p = fma(p, z, 1.0 / 19.0); p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0); p = fma(p, z, 1.0 / 13.0); p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0); p = fma(p, z, 1.0 / 7.0); p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  const double lm = fma(2.0 * s * z, p, 2.0 * s);  // 2 s + 2 s z p
  return fma((double)e, 0.69314718055994531, lm);
}
// sin and cos of 2 pi u for u in [0, 1): quarter turns q = rint(4u) taken out
// exactly, t = 2 pi (u - q/4) in [-pi/4, pi/4], Taylor series of sin t to
// t^17 and of cos t to t^18 (truncation below 1e-19).
SOPG_HD void bm_sincos2pi(double u, double& sn, double& cs) {
  const double q = rint(4.0 * u);
  const double t = fma(q, -0.25, u) * 6.283185307179586;
  const double z = t * t;
  double ps = -1.0 / 355687428096000.0;                         // -1/17!
  ps = fma(ps, z, 1.0 / 1307674368000.0); ps = fma(ps, z, -1.0 / 6227020800.0);
  ps = fma(ps, z, 1.0 / 39916800.0); ps = fma(ps, z, -1.0 / 362880.0); ps = fma(ps, z, 1.0 / 5040.0);
  ps = fma(ps, z, -1.0 / 120.0); ps = fma(ps, z, 1.0 / 6.0);
  const double st = fma(-t * z, ps, t);                         // t - t z (1/6 - ...)
  double pc = 1.0 / 6402373705728000.0;                         // 1/18!
  pc = fma(pc, z, -1.0 / 20922789888000.0); pc = fma(pc, z, 1.0 / 87178291200.0);
  pc = fma(pc, z, -1.0 / 479001600.0); pc = fma(pc, z, 1.0 / 3628800.0); pc = fma(pc, z, -1.0 / 40320.0);
  pc = fma(pc, z, 1.0 / 720.0); pc = fma(pc, z, -1.0 / 24.0); pc = fma(pc, z, 0.5);
  const double ct = fma(-z, pc, 1.0);                           // 1 - z (1/2 - ...)
  const int k = (int)q & 3;
  sn = (k == 0) ? st : (k == 1) ? ct : (k == 2) ? -st : -ct;
  cs = (k == 0) ? ct : (k == 1) ? -st : (k == 2) ? -ct : st;
}

// Two standard normal deviates (Box-Muller) from one Philox block.
SOPG_HD void philox_gauss2(uint64_t seed, uint64_t ctr, uint32_t bead, uint32_t pair, double& g0, double& g1) {
  uint32_t c[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), bead, pair};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  uint64_t a = ((uint64_t)c[0] << 32) | c[1], b = ((uint64_t)c[2] << 32) | c[3];
  double u1 = (double)((a >> 11) + 1) * (1.0 / 9007199254740992.0);  // (0, 1]
  double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);        // [0, 1)
  double r = sqrt(-2.0 * bm_log(u1));
  double s, co;
  bm_sincos2pi(u2, s, co);
  g0 = r * co;
  g1 = r * s;
}

// Slot word of the per-bead non-native neighbor list:
//   bits  0-24 partner bead (chain-local, 0-based)
//   bit     25 side: 0 = this bead is the pair's i (first) bead, 1 = its j bead
//   bits 26-27 itype, bits 28-29 jtype
//   bits 30-31 pair-list membership, one bit per parity (double-buffered)
constexpr uint32_t kPartnerMask = 0x1FFFFFFu;
// Native-contact constants from the pair's native distance r, computed as
// params.cpp:204-208 and sop.cpp:506-507 do (left to right, no contraction).
SOPG_HD void nat_powers(double r, double& r6, double& r12) {
  const double r2 = r * r;
  r6 = r2 * r2 * r2;
  r12 = r6 * r6;
}
SOPG_HD double nat_plrc2(double r) {
  const double rp = 2.5 * r;
  return rp * rp;
}
SOPG_HD double nat_nlrc2(double r) {  // neighbor_list.cpp:44-47
  const double rn = 3.2 * r;
  return rn * rn;
}
constexpr int kMaxChainBeads = 1 << 25;
// Beads in one context's pool: byte offsets into a bead array (8 B per
// bead) stay below 2^32, as the kernels' buffer loads take them.
constexpr size_t kMaxPoolBeads = (size_t)1 << 28;
SOPG_HD uint32_t slot_make(uint32_t partner, uint32_t side, uint32_t it, uint32_t jt) {
  return partner | (side << 25) | (it << 26) | (jt << 28);
}
// Layout of the slot words in HBM: tiles of kSlotTile consecutive beads x 4
// consecutive entries (cap words per bead, cap a multiple of 4; the bead pool
// a multiple of kSlotTile beads).  A wave that gives 16 consecutive beads 4
// lanes each reads one chunk of every bead's list as one contiguous 256-byte
// run, instead of 16 runs cap words apart.
constexpr int kSlotTile = 16;
SOPG_HD size_t slot_index(int g, int q, int cap) {
  return ((size_t)(g >> 4) * (size_t)cap + (size_t)(q & ~3)) * 16 + (size_t)(g & 15) * 4 + (size_t)(q & 3);
}

}  // namespace sopg
