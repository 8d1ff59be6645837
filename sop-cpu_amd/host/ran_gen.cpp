// ran_gen.cpp -- ran2 / gasdev of Numerical Recipes in C (2nd ed., sections
// 7.1 and 7.2) with the constants SOP-CPU uses (random_generator.h:15-36), and
// the reference's state-file layout (random_generator.cpp:24-70).
#include "ran_gen.h"

extern int sopg_restart_lossless;  // sop_shim.cpp

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>

namespace {
// L'Ecuyer's two multiplicative congruential generators, Schrage-factored.
constexpr int kM1 = 2147483563, kA1 = 40014, kQ1 = 53668, kR1 = 12211;
constexpr int kM2 = 2147483399, kA2 = 40692, kQ2 = 52774, kR2 = 3791;
constexpr int kTab = 32;
constexpr int kDiv = 1 + (kM1 - 1) / kTab;
constexpr double kAm = 1.0 / double(kM1);
constexpr double kMax = 1.0 - 3.0e-16;

inline int schrage(int s, int a, int q, int r, int m) {
  const int k = s / q;
  s = a * (s - k * q) - k * r;
  return s < 0 ? s + m : s;
}
}  // namespace

Ran_Gen::Ran_Gen() : mseed(0), iset(0), gset(0.0), idum2(123456789), iy(0) {
  std::memset(iv, 0, sizeof iv);
  std::strcpy(fname, "ran_gen.dat");
}

Ran_Gen::~Ran_Gen() {}

void Ran_Gen::set_fname(char* f) {
  std::strncpy(fname, f, sizeof fname - 1);
  fname[sizeof fname - 1] = 0;
}

void Ran_Gen::set_seed(int s) { mseed = s; }

double Ran_Gen::ran2() {
  if (mseed <= 0) {
    // (re)initialise: warm up 8 steps, then fill the shuffle table
    mseed = (mseed == 0) ? 1 : -mseed;
    idum2 = mseed;
    for (int j = kTab + 7; j >= 0; j--) {
      mseed = schrage(mseed, kA1, kQ1, kR1, kM1);
      if (j < kTab) iv[j] = mseed;
    }
    iy = iv[0];
  }
  mseed = schrage(mseed, kA1, kQ1, kR1, kM1);
  idum2 = schrage(idum2, kA2, kQ2, kR2, kM2);
  const int j = iy / kDiv;
  iy = iv[j] - idum2;
  iv[j] = mseed;
  if (iy < 1) iy += kM1 - 1;
  const double t = kAm * iy;
  return t > kMax ? kMax : t;
}

double Ran_Gen::gasdev() {
  if (iset) {
    iset = 0;
    return gset;
  }
  double v1, v2, rsq;
  do {
    v1 = 2.0 * ran2() - 1.0;
    v2 = 2.0 * ran2() - 1.0;
    rsq = v1 * v1 + v2 * v2;
  } while (rsq >= 1.0 || rsq == 0.0);
  const double fac = std::sqrt(-2.0 * std::log(rsq) / rsq);
  gset = v1 * fac;
  iset = 1;
  return v2 * fac;
}

// State file: mseed, iset, gset, idum2, iy, iv[0..31], one value per line.
// gset is written with %f as the reference writes it (which loses the cached
// deviate's low digits), or with %.17g under `set restart_format lossless;`;
// reading accepts either.
void Ran_Gen::save_state() {
  FILE* f = std::fopen(fname, "w");
  if (!f) return;
  std::fprintf(f, sopg_restart_lossless ? "%d\n%d\n%.17g\n%d\n%d\n" : "%d\n%d\n%f\n%d\n%d\n", mseed, iset, gset,
               idum2, iy);
  for (int i = 0; i < kTab; i++) std::fprintf(f, "%d\n", iv[i]);
  std::fclose(f);
}

void Ran_Gen::restart() {
  std::ifstream in(fname);
  in >> mseed >> iset >> gset >> idum2 >> iy;
  for (int i = 0; i < kTab; i++) in >> iv[i];
}
