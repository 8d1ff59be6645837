// sop_gen.cpp -- synthetic SOP chains and their input files.
//
// The reference ships no inputs (no topology, coordinates or scripts), so every
// chain used by the tests and the benchmarks is synthesised here:
//   * an off-lattice self-avoiding walk: 3.8 A bonds, every non-bonded pair
//     (|i-j| >= 2) at least 4.0 A apart;
//   * a Monte Carlo collapse into a sphere of radius 3.0*N^(1/3) A that keeps
//     both constraints (crankshaft, end and pivot moves, zero-temperature
//     acceptance on the excess radius);
//   * Maxwell velocities at kT (unit mass).
// Files are written in the reference's formats (params.cpp:128-259,
// io.cpp:180-242) with %.17g so they read back exactly:
//   bonds "bonds N-1" / "i i+1 r0", angles "angles N-2" / "i i+1 i+2 d(i,i+2)",
//   vdw   "natt NA nrep NR" / "i j r_ij 1 1" for every pair j-i>2 in (i,j)
//         lexicographic order (the loader splits natives at r_ij < 8 A),
//   init  "beads N" / "idx x y z", and "idx x y z" coordinate/velocity files.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <unordered_map>
#include <vector>

namespace {

// splitmix64: small, seedable, portable.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double gauss() {
    double u1 = 1.0 - uni(), u2 = uni();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  void unit(double* v) {
    double z = 2.0 * uni() - 1.0, phi = 6.283185307179586 * uni();
    double r = std::sqrt(1.0 - z * z);
    v[0] = r * std::cos(phi); v[1] = r * std::sin(phi); v[2] = z;
  }
};

const double kBond = 3.8, kExcl = 4.0, kExcl2 = kExcl * kExcl;

// Hash grid (cell = exclusion distance) for the non-bonded distance checks.
struct Grid {
  std::unordered_map<uint64_t, std::vector<int>> cells;
  static int64_t ci(double x) { return (int64_t)std::floor(x / kExcl); }
  static uint64_t key(int64_t a, int64_t b, int64_t c) {
    return ((uint64_t)(a & 0x1FFFFF) << 42) | ((uint64_t)(b & 0x1FFFFF) << 21) | (uint64_t)(c & 0x1FFFFF);
  }
  void add(int i, const double* p) { cells[key(ci(p[0]), ci(p[1]), ci(p[2]))].push_back(i); }
  void remove(int i, const double* p) {
    auto it = cells.find(key(ci(p[0]), ci(p[1]), ci(p[2])));
    if (it == cells.end()) return;
    auto& v = it->second;
    for (size_t k = 0; k < v.size(); k++)
      if (v[k] == i) { v[k] = v.back(); v.pop_back(); break; }
  }
  // true if some bead j with accept(j) lies within kExcl of p
  template <class F>
  bool clash(const double* p, const std::vector<double>& X, F accept) const {
    int64_t a = ci(p[0]), b = ci(p[1]), c = ci(p[2]);
    for (int64_t da = -1; da <= 1; da++)
      for (int64_t db = -1; db <= 1; db++)
        for (int64_t dc = -1; dc <= 1; dc++) {
          auto it = cells.find(key(a + da, b + db, c + dc));
          if (it == cells.end()) continue;
          for (int j : it->second) {
            if (!accept(j)) continue;
            double dx = X[3 * j] - p[0], dy = X[3 * j + 1] - p[1], dz = X[3 * j + 2] - p[2];
            if (dx * dx + dy * dy + dz * dz < kExcl2) return true;
          }
        }
    return false;
  }
};

inline double norm3(const double* p) { return std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }

// Rodrigues rotation of point p about the axis through c with unit direction k.
inline void rotate(const double* p, const double* c, const double* k, double ang, double* out) {
  double v[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
  double cs = std::cos(ang), sn = std::sin(ang);
  double kv = k[0] * v[0] + k[1] * v[1] + k[2] * v[2];
  double kx[3] = {k[1] * v[2] - k[2] * v[1], k[2] * v[0] - k[0] * v[2], k[0] * v[1] - k[1] * v[0]};
  for (int a = 0; a < 3; a++) out[a] = c[a] + v[a] * cs + kx[a] * sn + k[a] * kv * (1.0 - cs);
}

// Per bead i, its native partners j < i with their squared native distances.
using NativeFloor = std::vector<std::vector<std::pair<int, double>>>;

// Grow a SAW bead by bead inside radius rmax (<=0: unbounded) with backtracking.
// With floor, a bead is also placed no closer to an earlier native partner than
// their native distance.
bool grow_saw(int n, Rng& rng, double rmax, std::vector<double>& X, const NativeFloor* floor = nullptr) {
  X.assign(3 * (size_t)n, 0.0);
  Grid g;
  g.add(0, &X[0]);
  int i = 1;
  long budget = 400L * n + 100000;
  while (i < n) {
    if (--budget < 0) return false;
    bool ok = false;
    double p[3];
    for (int t = 0; t < 200 && !ok; t++) {
      double u[3];
      rng.unit(u);
      for (int a = 0; a < 3; a++) p[a] = X[3 * (i - 1) + a] + kBond * u[a];
      if (rmax > 0 && norm3(p) > rmax) continue;
      ok = !g.clash(p, X, [&](int j) { return j <= i - 2; });
      if (ok && floor)
        for (const auto& pr : (*floor)[i]) {
          const double* q = &X[3 * pr.first];
          const double dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
          if (dx * dx + dy * dy + dz * dz < pr.second) { ok = false; break; }
        }
    }
    if (ok) {
      for (int a = 0; a < 3; a++) X[3 * i + a] = p[a];
      g.add(i, &X[3 * i]);
      i++;
    } else {
      int back = 1 + (int)(rng.uni() * 8);
      if (back > i - 1) back = i - 1;
      for (int b = 0; b < back; b++) { i--; g.remove(i, &X[3 * i]); }
      if (i < 1) { i = 1; }
    }
  }
  return true;
}

// Zero-temperature MC collapse of X into radius R (center = origin).
bool collapse(int n, Rng& rng, double R, std::vector<double>& X, long max_sweeps) {
  Grid g;
  for (int i = 0; i < n; i++) g.add(i, &X[3 * i]);
  auto excess = [&](const double* p) { double r = norm3(p); return r > R ? (r - R) * (r - R) : 0.0; };
  double E = 0.0;
  for (int i = 0; i < n; i++) E += excess(&X[3 * i]);
  std::vector<double> Y;
  for (long sweep = 0; sweep < max_sweeps && E > 0.0; sweep++) {
    for (int m = 0; m < n; m++) {
      double mv = rng.uni();
      if (mv < 0.1 && n > 3) {
        // segment move (bond lengths kept): rotate beads k+1..m-1 about the
        // k--m axis (a long crankshaft), or a short end segment about bead k
        const int len = 2 + (int)(rng.uni() * 31);
        int k = (int)(rng.uni() * n), m = k + len + 1, lo, hi;
        double ax[3], ang;
        bool end = false;
        if (m < n) {
          lo = k + 1; hi = m - 1;
          for (int a = 0; a < 3; a++) ax[a] = X[3 * m + a] - X[3 * k + a];
          double l = norm3(ax);
          if (l < 1e-9) continue;
          for (int a = 0; a < 3; a++) ax[a] /= l;
          ang = (rng.uni() - 0.5) * 6.283185307179586;
        } else {
          end = true;
          if (rng.uni() < 0.5) { k = n - 1 - len; lo = k + 1; hi = n - 1; }
          else { k = len; lo = 0; hi = k - 1; }
          if (k < 1 || k > n - 2) continue;
          rng.unit(ax);
          ang = (rng.uni() - 0.5) * 1.2;
        }
        Y.resize(3 * (size_t)(hi - lo + 1));
        double dE = 0.0;
        for (int i = lo; i <= hi; i++) {
          rotate(&X[3 * i], &X[3 * k], ax, ang, &Y[3 * (i - lo)]);
          dE += excess(&Y[3 * (i - lo)]) - excess(&X[3 * i]);
        }
        if (dE > 0.0) continue;
        bool bad = false;
        for (int i = lo; i <= hi && !bad; i++)
          bad = g.clash(&Y[3 * (i - lo)], X, [&](int j) {
            bool fixed = j < lo || j > hi;
            return fixed && std::abs(j - i) >= 2;
          });
        (void)end;
        if (bad) continue;
        for (int i = lo; i <= hi; i++) {
          g.remove(i, &X[3 * i]);
          for (int a = 0; a < 3; a++) X[3 * i + a] = Y[3 * (i - lo) + a];
          g.add(i, &X[3 * i]);
        }
        E += dE;
      } else {
        int i = (int)(rng.uni() * n);
        double p[3];
        if (i == 0 || i == n - 1) {
          int nb = (i == 0) ? 1 : n - 2;
          double u[3];
          rng.unit(u);
          for (int a = 0; a < 3; a++) p[a] = X[3 * nb + a] + kBond * u[a];
        } else {
          double ax[3] = {X[3 * (i + 1)] - X[3 * (i - 1)], X[3 * (i + 1) + 1] - X[3 * (i - 1) + 1],
                          X[3 * (i + 1) + 2] - X[3 * (i - 1) + 2]};
          double l = norm3(ax);
          if (l < 1e-9) continue;
          for (int a = 0; a < 3; a++) ax[a] /= l;
          rotate(&X[3 * i], &X[3 * (i - 1)], ax, (rng.uni() - 0.5) * 6.283185307179586, p);
        }
        double dE = excess(p) - excess(&X[3 * i]);
        if (dE > 0.0) continue;
        if (g.clash(p, X, [&](int j) { return std::abs(j - i) >= 2; })) continue;
        g.remove(i, &X[3 * i]);
        for (int a = 0; a < 3; a++) X[3 * i + a] = p[a];
        g.add(i, &X[3 * i]);
        E += dE;
      }
    }
    if (sweep % 64 == 63) {  // refresh the running sum
      E = 0.0;
      for (int i = 0; i < n; i++) E += excess(&X[3 * i]);
    }
  }
  E = 0.0;
  for (int i = 0; i < n; i++) E += excess(&X[3 * i]);
  return E == 0.0;
}

double dist(const double* X, int i, int j) {
  double dx = X[3 * j] - X[3 * i], dy = X[3 * j + 1] - X[3 * i + 1], dz = X[3 * j + 2] - X[3 * i + 2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

}  // namespace

extern "C" {

// Self-avoiding walk without confinement (an unfolded start).  xyz: 3n doubles.
int sopgen_saw(int n, uint64_t seed, double* xyz) {
  if (n < 1 || !xyz) return -1;
  Rng rng(seed);
  std::vector<double> X;
  if (!grow_saw(n, rng, 0.0, X)) return -2;
  std::memcpy(xyz, X.data(), sizeof(double) * 3 * (size_t)n);
  return 0;
}

long sopgen_natives(int n, const double* xyz, double rcut, int* oi, int* oj, double* orr, long cap);

// Unfolded start of a folding run: a SAW inside radius rmax (<= 0: unbounded)
// that keeps every native pair of native_xyz (j - i > 2, closer than 8 A there)
// at least its native distance apart.  A plain SAW puts some native pairs at
// half their native distance, where the 12-6 native term's repulsion is
// thousands of times its well depth and breaks the FENE bonds in the first
// steps.
int sopgen_unfolded(int n, uint64_t seed, const double* native_xyz, double rmax, double* xyz) {
  if (n < 1 || !native_xyz || !xyz) return -1;
  const long cap = 64L * n + 64;
  std::vector<int> ni((size_t)cap), nj((size_t)cap);
  std::vector<double> nr((size_t)cap);
  const long nn = sopgen_natives(n, native_xyz, 8.0, ni.data(), nj.data(), nr.data(), cap);
  if (nn < 0 || nn > cap) return -1;
  NativeFloor fl((size_t)n);
  for (long k = 0; k < nn; k++) fl[nj[k] - 1].push_back({ni[k] - 1, nr[k] * nr[k]});
  Rng rng(seed);
  std::vector<double> X;
  for (int attempt = 0; attempt < 8; attempt++)
    if (grow_saw(n, rng, rmax, X, &fl)) {
      std::memcpy(xyz, X.data(), sizeof(double) * 3 * (size_t)n);
      return 0;
    }
  return -2;
}

// Native structure: SAW grown in a sphere of 1.3 R, collapsed into R = 3.0*n^(1/3).
int sopgen_native(int n, uint64_t seed, double* xyz) {
  if (n < 1 || !xyz) return -1;
  Rng rng(seed);
  double R = 3.0 * std::cbrt((double)n);
  std::vector<double> X;
  for (int attempt = 0; attempt < 8; attempt++) {
    if (!grow_saw(n, rng, 1.3 * R, X)) continue;
    if (collapse(n, rng, R, X, 200000)) {
      std::memcpy(xyz, X.data(), sizeof(double) * 3 * (size_t)n);
      return 0;
    }
  }
  return -2;
}

// Maxwell velocities at temperature kT (unit mass).
int sopgen_maxwell(int n, double kT, uint64_t seed, double* vel) {
  if (n < 1 || !vel || kT < 0) return -1;
  Rng rng(seed ^ 0xA5A5A5A5DEADBEEFull);
  double s = std::sqrt(kT);
  for (long i = 0; i < 3L * n; i++) vel[i] = s * rng.gauss();
  return 0;
}

// Smallest non-bonded distance (|i-j| >= 2) and largest bond deviation from 3.8.
int sopgen_check(int n, const double* xyz, double* min_nonbonded, double* max_bond_dev, double* max_radius) {
  double mn = 1e300, mb = 0.0, mr = 0.0;
  for (int i = 0; i < n; i++) {
    mr = std::fmax(mr, norm3(&xyz[3 * i]));
    if (i + 1 < n) mb = std::fmax(mb, std::fabs(dist(xyz, i, i + 1) - kBond));
    for (int j = i + 2; j < n; j++) mn = std::fmin(mn, dist(xyz, i, j));
  }
  *min_nonbonded = mn; *max_bond_dev = mb; *max_radius = mr;
  return 0;
}

// Write the topology (bonds, angles, vdw) and init files in the reference's
// formats.  Returns the number of native pairs (r_ij < rcut, default 8 A) or <0.
int sopgen_write_topology(int n, const double* xyz, const char* bonds_path, const char* angles_path,
                          const char* vdw_path, const char* init_path, double rcut_nat) {
  if (n < 4) return -1;
  FILE* f;
  if (bonds_path) {
    if (!(f = std::fopen(bonds_path, "w"))) return -3;
    std::fprintf(f, "bonds %d\n", n - 1);
    for (int i = 0; i + 1 < n; i++) std::fprintf(f, "%d %d %.17g\n", i + 1, i + 2, dist(xyz, i, i + 1));
    std::fclose(f);
  }
  if (angles_path) {
    if (!(f = std::fopen(angles_path, "w"))) return -3;
    std::fprintf(f, "angles %d\n", n - 2);
    for (int i = 0; i + 2 < n; i++)
      std::fprintf(f, "%d %d %d %.17g\n", i + 1, i + 2, i + 3, dist(xyz, i, i + 2));
    std::fclose(f);
  }
  int natt = 0;
  long nrep = 0;
  char buf[64];
  for (int i = 0; i < n; i++)
    for (int j = i + 3; j < n; j++) {
      // classify exactly as the loader will: atof(%.17g) reproduces the double
      std::snprintf(buf, sizeof buf, "%.17g", dist(xyz, i, j));
      if (std::atof(buf) < rcut_nat) natt++; else nrep++;
    }
  if (vdw_path) {
    if (!(f = std::fopen(vdw_path, "w"))) return -3;
    std::fprintf(f, "natt %d nrep %ld\n", natt, nrep);
    for (int i = 0; i < n; i++)
      for (int j = i + 3; j < n; j++) std::fprintf(f, "%d %d %.17g 1 1\n", i + 1, j + 1, dist(xyz, i, j));
    std::fclose(f);
  }
  if (init_path) {
    if (!(f = std::fopen(init_path, "w"))) return -3;
    std::fprintf(f, "beads %d\n", n);
    for (int i = 0; i < n; i++)
      std::fprintf(f, "%d %.17g %.17g %.17g\n", i + 1, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    std::fclose(f);
  }
  return natt;
}

// Native contacts of a structure: pairs j-i>2 with distance < rcut, in (i,j)
// lexicographic order, 1-based, with the distance exactly as the vdw file
// carries it.  O(n) with a cell grid.  Returns the count (entries beyond cap
// are counted, not written) or <0.
long sopgen_natives(int n, const double* xyz, double rcut, int* oi, int* oj, double* orr, long cap) {
  if (n < 1 || !xyz || !(rcut > 0)) return -1;
  const double cs = rcut;
  double lo[3] = {1e300, 1e300, 1e300};
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) lo[a] = std::fmin(lo[a], xyz[3 * i + a]);
  std::unordered_map<uint64_t, std::vector<int>> cells;
  auto key = [](int64_t a, int64_t b, int64_t c) {
    return ((uint64_t)(a & 0x1FFFFF) << 42) | ((uint64_t)(b & 0x1FFFFF) << 21) | (uint64_t)(c & 0x1FFFFF);
  };
  std::vector<int64_t> ci(3 * (size_t)n);
  for (int i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++) ci[3 * i + a] = (int64_t)std::floor((xyz[3 * i + a] - lo[a]) / cs);
    cells[key(ci[3 * i], ci[3 * i + 1], ci[3 * i + 2])].push_back(i);
  }
  long cnt = 0;
  std::vector<int> nb;
  for (int i = 0; i < n; i++) {
    nb.clear();
    for (int64_t da = -1; da <= 1; da++)
      for (int64_t db = -1; db <= 1; db++)
        for (int64_t dc = -1; dc <= 1; dc++) {
          auto it = cells.find(key(ci[3 * i] + da, ci[3 * i + 1] + db, ci[3 * i + 2] + dc));
          if (it == cells.end()) continue;
          for (int j : it->second)
            if (j > i + 2) nb.push_back(j);
        }
    std::sort(nb.begin(), nb.end());
    for (int j : nb) {
      const double d = dist(xyz, i, j);
      if (d < rcut) {
        if (cnt < cap) { oi[cnt] = i + 1; oj[cnt] = j + 1; orr[cnt] = d; }
        cnt++;
      }
    }
  }
  return cnt;
}

// "idx x y z" lines (coordinates or velocities, io.cpp:156-278 layout) at %.17g.
int sopgen_write_xyz(int n, const double* xyz, const char* path) {
  FILE* f = std::fopen(path, "w");
  if (!f) return -3;
  for (int i = 0; i < n; i++)
    std::fprintf(f, "%d %.17g %.17g %.17g\n", i + 1, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  std::fclose(f);
  return 0;
}

}  // extern "C"
