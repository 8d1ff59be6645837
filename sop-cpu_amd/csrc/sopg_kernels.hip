// sopg_kernels.hip -- gfx950 kernels of the SOP hot path (FP64, no MFMA).
//
// Work decomposition: a group of L lanes (1..64, inside one wave64) owns one
// bead.  Every per-bead list (bond, angle, native and non-native incidences)
// is walked L entries at a time; each lane evaluates one pair, and the group
// folds the pair forces into the bead's force strictly in list order with
// ballot + cross-lane shuffles.  That order -- noise, bonds, angles, native
// pairs, non-native pairs, each in the reference's list order -- is the order
// in which SOP-CPU's force_eval accumulates force[bead] (energy.cpp:21-33,
// 259-465), so the sums are bit-identical and independent of L, of the batch
// composition and of scheduling.  No floating-point atomics anywhere.
//
// Each pair is evaluated by both of its beads (the reference evaluates it once
// and applies +f / -f); both evaluations use the pair's own (i, j) orientation
// and therefore produce the same bits.
#include <hip/hip_runtime.h>
#include <mutex>
#include <type_traits>
#include "sopg_device.h"
#include "sopg_internal.h"
#include "../../include/sopg.h"

namespace sopg {

// ------------------------------------------------------------------ group ops

template <int L>
__device__ __forceinline__ uint64_t group_mask(bool c) {
  if constexpr (L == 1) {
    return c ? 1ull : 0ull;
  } else {
    uint64_t m = __ballot(c);
    const int gb = (threadIdx.x & 63) & ~(L - 1);
    if constexpr (L == 64) return m;
    else return (m >> gb) & ((1ull << L) - 1ull);
  }
}

// Lane K of each L-lane group (L = 2 or 4) to all lanes of the group: a DPP
// quad permutation (no LDS round trip), 32 bits at a time.
template <int K, int L>
__device__ __forceinline__ double group_bcast(double v) {
  constexpr int s01 = K, s23 = (L == 4) ? K : 2 + K;
  constexpr int ctrl = s01 | s01 << 2 | s23 << 4 | s23 << 6;
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, ctrl, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), ctrl, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (uint64_t)lo | ((uint64_t)hi << 32));
}

template <int K, int L>
__device__ __forceinline__ int group_bcast_i(int v) {
  constexpr int s01 = K, s23 = (L == 4) ? K : 2 + K;
  constexpr int ctrl = s01 | s01 << 2 | s23 << 4 | s23 << 6;
  return __builtin_amdgcn_mov_dpp(v, ctrl, 0xF, 0xF, false);
}

// F += contributions of the group's lanes, in lane order (= list order).
// For L = 2 and 4 every lane is added in turn: callers leave -0.0 in the
// values of a lane that does not contribute (c false), and x + (-0.0) is x
// for every x (both zeros included), so the sums are the same bits as adding
// the contributing lanes only.  There only lane 0's F is the bead's (lane 0
// adds its own value first, then lanes 1.. in turn); the other lanes' F is
// not used.
template <int L>
__device__ __forceinline__ void fold3(double& Fx, double& Fy, double& Fz, bool c, double fx, double fy, double fz) {
  if constexpr (L == 1) {
    if (c) { Fx += fx; Fy += fy; Fz += fz; }
  } else if constexpr (L == 2 || L == 4) {
    Fx += fx; Fy += fy; Fz += fz;
    Fx += group_bcast<1, L>(fx); Fy += group_bcast<1, L>(fy); Fz += group_bcast<1, L>(fz);
    if constexpr (L == 4) {
      Fx += group_bcast<2, L>(fx); Fy += group_bcast<2, L>(fy); Fz += group_bcast<2, L>(fz);
      Fx += group_bcast<3, L>(fx); Fy += group_bcast<3, L>(fy); Fz += group_bcast<3, L>(fz);
    }
  } else {
    uint64_t gm = group_mask<L>(c);
    // group-uniform loop over the contributing lanes, lowest lane (= earliest
    // list entry) first; every source lane belongs to this (active) group
    while (gm) {
      const int k = __builtin_ctzll(gm);
      gm &= gm - 1ull;
      Fx += __shfl(fx, k, L);
      Fy += __shfl(fy, k, L);
      Fz += __shfl(fz, k, L);
    }
  }
}

// fold3 for one value (energy terms): the same list order for every L, so
// a bead's energy sums do not depend on the lane count either.  Same
// conventions: e is -0.0 in a lane that does not contribute, and at L = 2, 4
// only lane 0's E is the bead's.
template <int L>
__device__ __forceinline__ void fold1(double& E, bool c, double e) {
  if constexpr (L == 1) {
    if (c) E += e;
  } else if constexpr (L == 2 || L == 4) {
    E += e;
    E += group_bcast<1, L>(e);
    if constexpr (L == 4) {
      E += group_bcast<2, L>(e);
      E += group_bcast<3, L>(e);
    }
  } else {
    uint64_t gm = group_mask<L>(c);
    while (gm) {
      const int k = __builtin_ctzll(gm);
      gm &= gm - 1ull;
      E += __shfl(e, k, L);
    }
  }
}

// Workgroups are dispatched round-robin over the chip's 8 XCDs, each with its
// own L2.  Logical block k of the returned numbering runs on one XCD together
// with its neighbours k-1, k+1, ..., so the contiguous beads of a chain -- and
// the partner positions they gather -- stay in one L2.  A bijection on
// [0, gridDim.x).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
  const int n = gridDim.x, b = blockIdx.x;
  const int q = n / kXcds, r = n % kXcds, x = b % kXcds;
  return x * q + (x < r ? x : r) + b / kXcds;
}

// ------------------------------------------------------------------ rare paths

// Native pair-list membership of master entry k at parity pr: the flag of its
// first-bead incidence.
__device__ __forceinline__ bool att_in_pl(const uint8_t* flag, const int* __restrict__ q0, int k, int pr) {
  return (flag[q0[k]] >> (1 + pr)) & 1;
}

// energy.cpp:293: the native skip test reads lj_nat_pdb_dist2[i] with i the
// PAIR-LIST position of the entry, not its master index.  Only reachable when
// d2 < tol * max(lj_nat_pdb_dist2); then the position is counted exactly.
__device__ __noinline__ bool att_tol_skip(const uint8_t* flag, const int* __restrict__ q0,
                                          const double* __restrict__ r2, int att_off, int k, int pr, double d2) {
  int p = 0;
  for (int kk = att_off; kk < k; kk++) p += att_in_pl(flag, q0, kk, pr);
  return d2 < kTol * r2[att_off + p];
}

// energy.cpp:292, 331: rep_tol comes from the LAST native pair-list entry.
// With an empty native pair list the reference reads it uninitialised; here no
// non-native pair is skipped in that case.
__device__ __noinline__ bool rep_tol_skip(const uint8_t* flag, const int* __restrict__ q0,
                                          const int8_t* __restrict__ am_t, const Tables* __restrict__ T, int att_off,
                                          int natt, int pr, double d2) {
  for (int kk = att_off + natt - 1; kk >= att_off; kk--) {
    if (att_in_pl(flag, q0, kk, pr)) {
      const int tt = am_t[kk];
      return d2 < T->s2[tt & 3][tt >> 2] * kTol;
    }
  }
  return false;
}

// Raw buffer view of a device array: loads through it take a 32-bit byte
// offset, so the three coordinate gathers of one partner share one offset
// register instead of three 64-bit addresses.  Dword 3 is the untyped
// 32-bit-data form of this family; bead arrays stay below 2^32 bytes (the
// bead pool is capped at kMaxPoolBeads).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_view(const double* p) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, -1, 0x00020000);
}
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
}

// Bead g's slot words (the neighbor list: tiled, slot_index) or candidate
// words (row-major, cap words per bead).  The bead's first word is located
// once; entry q of a tiled row is then (q & ~3) * 16 + (q & 3) words further,
// a 32-bit offset.
struct ListRow {
  uint32_t* row;
  bool tiled;
  __device__ __forceinline__ ListRow(uint32_t* base, int g, int cap, bool t)
      : row(base + (t ? slot_index(g, 0, cap) : (size_t)g * (size_t)cap)), tiled(t) {}
  __device__ __forceinline__ uint32_t& operator[](int q) const {
    return tiled ? row[(uint32_t)(((q & ~3) << 4) | (q & 3))] : row[(uint32_t)q];
  }
};

// A lane's force term cs * d, or -0.0 (which adds nothing in the folds) when
// the lane has none: the product's sign bit is set on such lanes, whose cs is
// 0 and whose product is therefore +0 or -0.
__device__ __forceinline__ void neutral3(bool c, double cs, double dx, double dy, double dz, double& fx, double& fy,
                                         double& fz) {
  const uint64_t nz = c ? 0ull : 0x8000000000000000ull;
  fx = __longlong_as_double(__double_as_longlong(cs * dx) | nz);
  fy = __longlong_as_double(__double_as_longlong(cs * dy) | nz);
  fz = __longlong_as_double(__double_as_longlong(cs * dz) | nz);
}
// The same term where cs and d are defined only on lanes with c: a select.
__device__ __forceinline__ void neutral3s(bool c, double cs, double dx, double dy, double dz, double& fx, double& fy,
                                          double& fz) {
  fx = c ? cs * dx : -0.0;
  fy = c ? cs * dy : -0.0;
  fz = c ? cs * dz : -0.0;
}

// Pair displacement u[j] - u[i] seen from bead g, min-imaged, all three
// components at once: c is g's own position, o the partner's.  The reference
// computes o - c when g is the pair's i bead (side 0) and c - o when it is j
// (side 1).  Side 1 here uses (o - c) * -1.0: rounding to nearest is
// symmetric, so the two are the same bits except for c == o, where they are
// -0 and +0; mimg maps both zeros to +0 (d + 0.0 on the fast path; d -
// boxl*rnd(d/boxl) = +0 on the full form), so the result is the reference's,
// with one subtraction per component.  The sign and the + 0.0 are one fma:
// (o - c) * +-1 is exact, so fma((o - c), +-1, 0.0) is that product with -0
// mapped to +0.  One branch covers the common case in which every component
// takes mimg's exact fast path (|d| <= boxl/4); otherwise all three take the
// full form, which equals the fast path's d + 0.0 on a component within it.
__device__ __forceinline__ void pair_d3(int side, double cx, double cy, double cz, double ox, double oy, double oz,
                                        double boxl, double qbox, double& dx, double& dy, double& dz) {
  const double sg = side ? -1.0 : 1.0;
  const double rx = fma(ox - cx, sg, 0.0);
  const double ry = fma(oy - cy, sg, 0.0);
  const double rz = fma(oz - cz, sg, 0.0);
  if (fabs(rx) <= qbox && fabs(ry) <= qbox && fabs(rz) <= qbox) {
    dx = rx; dy = ry; dz = rz;
  } else {
    dx = rx - boxl * rnd(rx / boxl);
    dy = ry - boxl * rnd(ry / boxl);
    dz = rz - boxl * rnd(rz / boxl);
  }
}

__device__ __forceinline__ bool cell_adjacent(const Dev& D, const ChainP& P, int i, int j) {
  // cell_list.cpp:24-44 with the cell indices precomputed per bead
  const double cx = mimg(D.cx[i] - D.cx[j], P.ncell, P.qncell);
  const double cy = mimg(D.cy[i] - D.cy[j], P.ncell, P.qncell);
  const double cz = mimg(D.cz[i] - D.cz[j], P.ncell, P.qncell);
  return cx >= -1.0 && cx <= 1.0 && cy >= -1.0 && cy <= 1.0 && cz >= -1.0 && cz <= 1.0;
}

// ------------------------------------------------------------------ kernels

// Cell index per bead: floor(unc_pos / lcell) (cell_list.cpp:24-31).
__device__ __forceinline__ void cellcoords_bead(const Dev& D, int cur, int g) {
  const ChainP& P = D.cp[D.chain[g]];
  D.cx[g] = floor((cur ? D.ux[1] : D.ux[0])[g] / P.lcell);
  D.cy[g] = floor((cur ? D.uy[1] : D.uy[0])[g] / P.lcell);
  D.cz[g] = floor((cur ? D.uz[1] : D.uz[0])[g] / P.lcell);
}
__global__ void k_cellcoords(Dev D, int cur, int g0, int g1) {
  const int g = g0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= g1 || D.chain[g] < 0) return;  // past the end, or a free bead
  if (D.cp[D.chain[g]].resident) return;  // the resident kernel keeps its own
  cellcoords_bead(D, cur, g);
}

// Neighbor-list (distance) or cell-list rebuild plus the pair-list filter at
// the same positions, scanning every candidate of bead g (the full scan; the
// binned kernel below covers compact chains).  mode < 0: each chain's own
// cutoff, only chains whose rebuild is due at relative step rel; mode >= 0:
// that criterion, all chains.
// RES: called from the resident kernel (which owns its chain and uses no grid).
template <int L, bool RES = false>
__device__ __forceinline__ void rebuild_scan_bead(const Dev& D, const Tables* __restrict__ T, int cur, int par,
                                                  int rel, int mode, int g, int l) {
  const int c = D.chain[g];
  if (c < 0) return;  // a free bead
  // the plan of k_grid (due chains only): the full scan, or another path
  if (!RES && !(D.grid[c].work & W_SCAN)) return;
  const ChainP& P = D.cp[c];
  int crit = mode;
  if (mode < 0) {
    if (P.cutoff == SOPG_CUTOFF_NONE || (P.phase + rel) % P.nnlup != 0) return;
    if (!RES && P.resident) return;
    crit = P.cutoff;
  }
  const double* ux = cur ? D.ux[1] : D.ux[0];
  const double* uy = cur ? D.uy[1] : D.uy[0];
  const double* uz = cur ? D.uz[1] : D.uz[0];
  const double boxl = P.boxl, qbox = P.qbox;
  // candidate-list build (a neighbor-list chain too wide for the grid): the
  // pairs within 3.2 sigma + skin, exactly, no pair-list bits; the natives are
  // left to the candidate filter, which runs next
  const bool cmode = !RES && (D.grid[c].work & W_CAND);

  // native contacts (neighbor_list.cpp:17-67 / cell_list.cpp:16-57)
  const int a0 = D.att_beg[g], a1 = D.att_end[g];
  for (int q = a0 + l; q < a1 && !cmode; q += L) {
    const int code = D.att_code[q], k = code >> 1, side = code & 1, o = D.att_partner[q];
    const int i = side ? o : g, j = side ? g : o;
    const double dx = mimg(ux[j] - ux[i], boxl, qbox);
    const double dy = mimg(uy[j] - uy[i], boxl, qbox);
    const double dz = mimg(uz[j] - uz[i], boxl, qbox);
    const double d2 = dx * dx + dy * dy + dz * dz;
    bool nl, pl;
    if (crit == SOPG_CUTOFF_NEIGHBOR) { nl = d2 < D.am_nlrc2[k]; pl = nl && d2 < nat_plrc2(D.att_r[q]); }
    else if (crit == SOPG_CUTOFF_CELL) { nl = cell_adjacent(D, P, i, j); pl = nl && d2 < nat_plrc2(D.att_r[q]); }
    else { nl = true; pl = true; }
    D.att_flag[q] = (uint8_t)((nl ? 1 : 0) | (pl ? (2 << par) : 0));
  }

  // non-native contacts: stable (ballot + prefix) compaction into the slots
  const int b = g - P.off;
  const ListRow S{cmode ? D.cand : D.slot, g, D.cap, !cmode};
  int* SK = D.slot_k ? D.slot_k + (size_t)g * D.cap : nullptr;
  int cnt = 0;
  if (!P.rep_explicit) {
    // implicit master list: all j-i>2 pairs in (i,j) order minus natives; for
    // bead b that is every partner c (|c-b|>2, not native) in ascending order.
    int np = a0;  // per-lane cursor into b's natives (ascending partners)
    for (int base = 0; base < P.n; base += L) {
      const int c = base + l;
      bool keep = false;
      uint32_t w = 0;
      if (c < P.n && (c - b > 2 || b - c > 2)) {
        const int gc = P.off + c;
        while (np < a1 && D.att_partner[np] < gc) np++;
        if (!(np < a1 && D.att_partner[np] == gc)) {
          const int side = (gc < g) ? 1 : 0;
          const int i = side ? gc : g, j = side ? g : gc;
          const int it = D.type[i], jt = D.type[j];
          const double dx = mimg(ux[j] - ux[i], boxl, qbox);
          const double dy = mimg(uy[j] - uy[i], boxl, qbox);
          const double dz = mimg(uz[j] - uz[i], boxl, qbox);
          const double d2 = dx * dx + dy * dy + dz * dz;
          bool pl;
          if (cmode) { keep = d2 < T->candc2[it][jt]; pl = false; }
          else if (crit == SOPG_CUTOFF_NEIGHBOR) { keep = d2 < T->nlrc2[it][jt]; pl = keep && d2 < T->plrc2[it][jt]; }
          else if (crit == SOPG_CUTOFF_CELL) { keep = cell_adjacent(D, P, i, j); pl = keep && d2 < T->plrc2[it][jt]; }
          else { keep = true; pl = true; }
          w = slot_make((uint32_t)c, (uint32_t)side, (uint32_t)it, (uint32_t)jt) | (pl ? (1u << (30 + par)) : 0u);
        }
      }
      const uint64_t m = group_mask<L>(keep);
      if (keep) {
        const int idx = cnt + __popcll(m & ((1ull << l) - 1ull));
        if (idx < D.cap) S[idx] = w;
      }
      cnt += __popcll(m);
    }
  } else {
    const int r0 = D.rep_beg[g], r1 = D.rep_end[g];
    for (int base = r0; base < r1; base += L) {
      const int q = base + l;
      bool keep = false;
      uint32_t w = 0;
      int kk = 0;
      if (q < r1) {
        const int code = D.rep_code[q], side = code & 1;
        kk = code >> 1;
        const int i = D.rm_i[kk], j = D.rm_j[kk];
        const int tt = D.rm_t[kk], it = tt & 3, jt = tt >> 2;
        const double dx = mimg(ux[j] - ux[i], boxl, qbox);
        const double dy = mimg(uy[j] - uy[i], boxl, qbox);
        const double dz = mimg(uz[j] - uz[i], boxl, qbox);
        const double d2 = dx * dx + dy * dy + dz * dz;
        bool pl;
        if (crit == SOPG_CUTOFF_NEIGHBOR) { keep = d2 < T->nlrc2[it][jt]; pl = keep && d2 < T->plrc2[it][jt]; }
        else if (crit == SOPG_CUTOFF_CELL) { keep = cell_adjacent(D, P, i, j); pl = keep && d2 < T->plrc2[it][jt]; }
        else { keep = true; pl = true; }
        const int o = side ? i : j;
        w = slot_make((uint32_t)(o - P.off), (uint32_t)side, (uint32_t)it, (uint32_t)jt) | (pl ? (1u << (30 + par)) : 0u);
      }
      const uint64_t m = group_mask<L>(keep);
      if (keep) {
        const int idx = cnt + __popcll(m & ((1ull << l) - 1ull));
        if (idx < D.cap) { S[idx] = w; if (SK) SK[idx] = kk; }
      }
      cnt += __popcll(m);
    }
  }
  if (l == 0) {
    (cmode ? D.ccnt : D.cnt)[g] = cnt < D.cap ? cnt : D.cap;
    if (cnt > D.cap) {
      atomicOr(D.err, 1);
      if (cmode) D.grid[c].cand_valid = 0;  // a truncated candidate list must not serve later rebuilds
    }
  }
}

// Grid-stride over beads: a bounded grid, so that a launch in which every due
// chain is binned (the usual case) costs little more than its dispatch.
template <int L>
__global__ void __launch_bounds__(256) k_rebuild(Dev D, const Tables* __restrict__ T, int cur, int par, int rel,
                                                 int mode, int g0, int g1) {
  rel += *D.relb;
  if (D.plan[2] != rel) return;  // no chain takes the full scan in this rebuild
  const int l = threadIdx.x & (L - 1);
  const int stride = (int)(gridDim.x * blockDim.x / L);
  for (int g = g0 + (xcd_block() * (int)blockDim.x + (int)threadIdx.x) / L; g < g1; g += stride)
    rebuild_scan_bead<L>(D, T, cur, par, rel, mode, g, l);
}

// ------------------------------------------------------- binned rebuild
//
// Same lists as k_rebuild, with O(neighbors) candidates per bead instead of
// O(N): beads are binned into a per-chain grid; a wave per bead tests the 27
// neighbouring cells, marks accepted partners in an LDS bitmap (one bit per
// chain bead) and emits the bitmap in ascending order -- the reference's
// lexicographic order -- with a wave prefix sum.  The grid covers the chain's
// extent, or, for a chain wider than the period less one grid edge, the period
// itself with wrapped coordinates (neighbouring cells modulo the cell count).

__device__ __forceinline__ bool rebuild_due(const ChainP& P, int rel, int mode, int& crit) {
  crit = mode;
  if (mode >= 0) return true;
  crit = P.cutoff;
  return P.cutoff != SOPG_CUTOFF_NONE && !P.resident && (P.phase + rel) % P.nnlup == 0;
}

// Per-chain plan of a list rebuild (GridP::work) and, for binned chains, the
// extent and grid; one block per chain, every chain of the launch.  Chains of
// more than maxn beads (whose LDS bitmap would not fit) keep the full scan.
// Neighbor-list chains in a run use candidate lists: while the chain's
// largest displacement from u0 stays within half the skin, the neighbor list
// is filtered from the candidate list; otherwise the candidate list is
// rebuilt (binned, cells >= 3.2 sigma + skin) at the current positions.
__device__ __forceinline__ int grid_coord(double v, double o, double inv, int n) {
  int k = (int)floor((v - o) * inv);
  return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

// One block per chain; a chain's beads (up to thousands) are strided over the
// block's threads, so a wide block keeps the per-chain loop short.
constexpr int kGridThreads = 1024;
__global__ void __launch_bounds__(kGridThreads) k_grid(Dev D, const Tables* __restrict__ T, int cur, int rel, int mode,
                                              int c0, int maxn) {
  rel += *D.relb;
  const int c = c0 + blockIdx.x;
  const ChainP& P = D.cp[c];
  GridP& G = D.grid[c];
  int crit;
  if (!rebuild_due(P, rel, mode, crit)) {
    if (threadIdx.x == 0) G.work = 0;
    return;
  }
  if (P.n > maxn || P.rep_explicit || crit == SOPG_CUTOFF_NONE) {
    if (threadIdx.x == 0) { G.work = W_SCAN; D.plan[2] = rel; }
    return;
  }
  // neighbor-list rebuilds, in a run or standalone, go through candidate lists
  const bool cand = crit == SOPG_CUTOFF_NEIGHBOR && T->cand_on;
  const double* px = cur ? D.ux[1] : D.ux[0];
  const double* py = cur ? D.uy[1] : D.uy[0];
  const double* pz = cur ? D.uz[1] : D.uz[0];
  // the reductions' scratch, then (binning) the per-cell counts and positions
  __shared__ union {
    double sh[6][kGridThreads];
    int cells[kGridCells];
  } u;
  auto& sh = u.sh;
  __shared__ int part[kGridThreads];
  __shared__ double gf[5];
  __shared__ int gi[5];
  if (cand && G.cand_valid) {
    // largest squared displacement from the candidate lists' positions
    double dmax = 0.0;
    for (int b = threadIdx.x; b < P.n; b += blockDim.x) {
      const int g = P.off + b;
      const double x = px[g], y = py[g], z = pz[g];
      const double ex = x - D.u0x[g], ey = y - D.u0y[g], ez = z - D.u0z[g];
      const double dd = ex * ex + ey * ey + ez * ez;
      dmax = (dd > dmax || dd != dd) ? dd : dmax;  // NaN propagates (a rebuild)
      D.fpos[g] = make_float4((float)x, (float)y, (float)z, (float)D.type[g]);  // for the filter's gathers
    }
    sh[0][threadIdx.x] = dmax;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        const double o = sh[0][threadIdx.x + w];
        if (o > sh[0][threadIdx.x] || o != o) sh[0][threadIdx.x] = o;
      }
      __syncthreads();
    }
    const bool keep = sh[0][0] <= T->cand_disp2;  // false for NaN
    __syncthreads();
    if (keep) {
      if (threadIdx.x == 0) { G.work = W_FILTER; D.plan[1] = rel; }
      return;
    }
  }
  const bool cell = crit == SOPG_CUTOFF_CELL;
  const double* ax = cell ? D.cx : px;
  const double* ay = cell ? D.cy : py;
  const double* az = cell ? D.cz : pz;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < P.n; b += blockDim.x) {
    const int g = P.off + b;
    const double v[3] = {ax[g], ay[g], az[g]};
    for (int a = 0; a < 3; a++) { lo[a] = fmin(lo[a], v[a]); hi[a] = fmax(hi[a], v[a]); }
    if (cand) {
      D.u0x[g] = v[0]; D.u0y[g] = v[1]; D.u0z[g] = v[2];
      D.fpos[g] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)D.type[g]);
    }
  }
  for (int a = 0; a < 3; a++) { sh[a][threadIdx.x] = lo[a]; sh[3 + a][threadIdx.x] = hi[a]; }
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; a++) {
        sh[a][threadIdx.x] = fmin(sh[a][threadIdx.x], sh[a][threadIdx.x + w]);
        sh[3 + a][threadIdx.x] = fmax(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // A chain whose extent on every axis is at most the period less the grid
    // edge has no pair that the criterion keeps across the boundary: a pair
    // whose plain difference exceeds half the period has the min image
    // period - |difference| >= the grid edge > the cutoff (the cell list: >= 2
    // cells).  So every kept pair lies in adjacent cells of a local grid over
    // the unwrapped coordinates; the distances themselves keep the reference's
    // min image.
    const double gedge = cell ? 1.0 : (cand ? P.gcand : P.gcell);
    const double limit = cell ? P.ncell - 2.0 : P.boxl - gedge;
    // cells of the grid edge, or wider where the extent would need more than
    // kGridDim of them (cubic cells of at least the edge keep the adjacency)
    double ext[3], emax = 0.0;
    bool fin = true;
    for (int a = 0; a < 3; a++) {
      ext[a] = sh[3 + a][0] - sh[a][0];
      fin = fin && isfinite(ext[a]);
      emax = fmax(emax, ext[a]);
    }
    const double edge = fmax(gedge, emax / (kGridDim - 1));
    double inv = 1.0 / edge;
    int dims[3];
    bool use = fin;
    for (int a = 0; a < 3; a++) {
      use = use && ext[a] <= limit;
      dims[a] = use ? (int)floor(ext[a] * inv) + 1 : 0;
      use = use && dims[a] <= kGridDim;
    }
    G.ox = sh[0][0]; G.oy = sh[1][0]; G.oz = sh[2][0];
    // A chain too wide for that grid is binned periodically: coordinates
    // wrapped into [0, period) (boxl; ncell for the cell list, whose
    // coordinates are the reference's cell indices), n cells of edge period / n
    // >= the grid edge per axis (3 <= n <= 32), neighbouring cells taken modulo
    // n.  A pair within the cutoff (under the reference's min image; for the
    // cell list, cells adjacent modulo ncell) has wrapped coordinates that
    // differ by less than one cell edge modulo the period, so its beads lie in
    // cells adjacent modulo n (n >= 3 keeps the 27 cells distinct).
    const double period = cell ? P.ncell : P.boxl;
    int per = 0;
    if (!use && fin && period > 0.0) {
      int n = (int)fmin(floor(period / gedge), (double)kGridDim);
      while (n >= 3 && period / n < gedge) n--;
      if (n >= 3) {
        per = 1;
        use = true;
        dims[0] = dims[1] = dims[2] = n;
        inv = (double)n / period;
        G.ox = G.oy = G.oz = 0.0;
      }
    }
    G.inv = inv;
    G.nx = dims[0]; G.ny = dims[1]; G.nz = dims[2];
    G.periodic = per;
    // a chain that neither grid takes (a period under three grid edges,
    // non-finite positions) builds its lists by the full scan
    G.work = (use ? W_BIN : W_SCAN) | (cand ? (W_CAND | W_FILTER) : 0);
    if (G.work & W_BIN) D.plan[0] = rel;
    if (G.work & W_FILTER) D.plan[1] = rel;
    if (G.work & W_SCAN) D.plan[2] = rel;
    if (cand) G.cand_valid = 1;
    gf[0] = G.ox; gf[1] = G.oy; gf[2] = G.oz; gf[3] = inv; gf[4] = period;
    gi[0] = dims[0]; gi[1] = dims[1]; gi[2] = dims[2]; gi[3] = G.work; gi[4] = per;
  }
  __syncthreads();
  if (!(gi[3] & W_BIN)) return;  // block-uniform
  // Counting sort of the chain's beads into the grid's cells (in LDS): the
  // cell of each bead, the cells' starts (exclusive scan of the counts), and
  // the beads of each cell, contiguous in cell order (any order inside a cell:
  // the rebuild's bitmaps make the lists independent of it).
  const int nx = gi[0], ny = gi[1], nz = gi[2], ncell = nx * ny * nz;
  const double gox = gf[0], goy = gf[1], goz = gf[2], ginv = gf[3], gbox = gf[4];
  const bool gper = gi[4] != 0;
  // periodic: the coordinate wrapped into [0, period) (the period itself, by
  // rounding, falls in the last cell, adjacent to the first)
  auto wrap = [&](double v) { return gper ? v - gbox * floor(v / gbox) : v; };
  int* cells = u.cells;
  const int t = (int)threadIdx.x, nt = (int)blockDim.x;
  for (int k = t; k < ncell; k += nt) cells[k] = 0;
  __syncthreads();
  for (int b = t; b < P.n; b += nt) {
    const int g = P.off + b;
    const int id = (grid_coord(wrap(az[g]), goz, ginv, nz) * ny + grid_coord(wrap(ay[g]), goy, ginv, ny)) * nx +
                   grid_coord(wrap(ax[g]), gox, ginv, nx);
    D.bead_cell[g] = id;
    atomicAdd(&cells[id], 1);
  }
  __syncthreads();
  // each thread scans a contiguous run of cells; the runs' sums are scanned
  // across the block
  const int per = (ncell + nt - 1) / nt, k0 = min(ncell, t * per), k1 = min(ncell, k0 + per);
  int sum = 0;
  for (int k = k0; k < k1; k++) sum += cells[k];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < nt; o <<= 1) {
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - sum;
  int* st = D.cell_start + (size_t)c * kGridCells;
  int* fl = D.cell_fill + (size_t)c * kGridCells;
  for (int k = k0; k < k1; k++) {
    const int nk = cells[k];
    st[k] = run;
    fl[k] = nk;
    cells[k] = run;  // the cell's next free position
    run += nk;
  }
  __syncthreads();
  for (int b = t; b < P.n; b += nt) {
    const int g = P.off + b;
    D.cell_beads[P.off + atomicAdd(&cells[D.bead_cell[g]], 1)] = b;
  }
}


// Counting sort of beads into cells: count, scan, fill.
// One group of LG lanes per bead, 256 / LG beads per block.  Each bead keeps
// two LDS bitmaps with one bit per bead of its chain: partners that pass the
// list criterion, and those of them inside the pair-list cutoff.  The 27
// neighbouring grid cells are walked as 9 x-runs: cells (x-1..x+1, y, z) are
// consecutive in the cell-sorted bead array.
template <int LG>
__global__ void __launch_bounds__(256) k_rebuild_grid(Dev D, const Tables* __restrict__ T, int cur, int par, int rel,
                                                      int mode, int g0, int g1, const int* __restrict__ tiles) {
  rel += *D.relb;
  if (D.plan[0] != rel) return;  // block-uniform: before any barrier
  constexpr int BPB = 256 / LG;
  static_assert(BPB == kBeadTile, "one block per bead tile");
  extern __shared__ uint32_t bitmaps[];
  // per type pair (it * 3 + jt): (3.2 sigma)^2, (2.5 sigma)^2, (3.2 sigma + skin)^2
  __shared__ double cut2[9][3];
  if (threadIdx.x < 9) {
    cut2[threadIdx.x][0] = T->nlrc2[threadIdx.x / 3][threadIdx.x % 3];
    cut2[threadIdx.x][1] = T->plrc2[threadIdx.x / 3][threadIdx.x % 3];
    cut2[threadIdx.x][2] = T->candc2[threadIdx.x / 3][threadIdx.x % 3];
  }
  const int grp = (int)threadIdx.x / LG, l = (int)threadIdx.x & (LG - 1);
  // block blk: beads g0 + [blk, blk + 1) * BPB, or with a tile table (the
  // tiles of kBeadTile beads that hold a bead of [g0, g1)) tile tiles[blk]
  const int blk = xcd_block();
  const int g = (tiles ? tiles[blk] * kBeadTile : g0 + blk * BPB) + grp;
  const int W = D.bitmap_words;
  uint32_t* nlb = bitmaps + (size_t)grp * 2 * W;
  uint32_t* plb = nlb + W;
  bool active = g >= g0 && g < g1;
  if (active) active = D.chain[g] >= 0;
  int crit = 0, work = 0;
  const ChainP* Pp = nullptr;
  if (active) {
    work = D.grid[D.chain[g]].work;
    active = (work & W_BIN) != 0;
  }
  if (active) {
    Pp = &D.cp[D.chain[g]];
    crit = mode >= 0 ? mode : Pp->cutoff;
  }
  // candidate-list build: the wider cutoff, no pair-list bits, natives left
  // to the filter
  const bool cmode = (work & W_CAND) != 0;
  const double* ux = cur ? D.ux[1] : D.ux[0];
  const double* uy = cur ? D.uy[1] : D.uy[0];
  const double* uz = cur ? D.uz[1] : D.uz[0];
  int words = 0, b = 0, a0 = 0, a1 = 0;
  if (active) {
    const ChainP& P = *Pp;
    b = g - P.off;
    words = (P.n + 31) >> 5;
    a0 = D.att_beg[g];
    a1 = D.att_end[g];
    const double boxl = P.boxl, qbox = P.qbox;
    // natives: NL / PL flags, exactly as the full scan
    for (int q = a0 + l; q < a1 && !cmode; q += LG) {
      const int code = D.att_code[q], k = code >> 1, side = code & 1, o = D.att_partner[q];
      const int i = side ? o : g, j = side ? g : o;
      const double dx = mimg(ux[j] - ux[i], boxl, qbox);
      const double dy = mimg(uy[j] - uy[i], boxl, qbox);
      const double dz = mimg(uz[j] - uz[i], boxl, qbox);
      const double d2 = dx * dx + dy * dy + dz * dz;
      bool nl, pl;
      if (crit == SOPG_CUTOFF_NEIGHBOR) { nl = d2 < D.am_nlrc2[k]; pl = nl && d2 < nat_plrc2(D.att_r[q]); }
      else { nl = cell_adjacent(D, P, i, j); pl = nl && d2 < nat_plrc2(D.att_r[q]); }
      D.att_flag[q] = (uint8_t)((nl ? 1 : 0) | (pl ? (2 << par) : 0));
    }
    for (int w = l; w < words; w += LG) { nlb[w] = 0u; plb[w] = 0u; }
  }
  __syncthreads();
  if (active) {
    const ChainP& P = *Pp;
    const int c = D.chain[g];
    const GridP& G = D.grid[c];
    const int id = D.bead_cell[g];
    const int cxi = id % G.nx, cyi = (id / G.nx) % G.ny, czi = id / (G.nx * G.ny);
    const size_t cbase = (size_t)c * kGridCells;
    const int* cb = D.cell_beads + P.off;
    const int off = P.off;
    const double boxl = P.boxl, qbox = P.qbox;
    const double ox = ux[g], oy = uy[g], oz = uz[g];
    const int gt = D.type[g];
    const bool nlcrit = crit == SOPG_CUTOFF_NEIGHBOR;
    // The 27 neighbouring cells as x-runs, contiguous in the sorted array,
    // walked as one candidate sequence (LG candidates per chunk, the next chunk
    // loaded before the current one is tested): 9 runs on the local grid (rows
    // outside it empty), and on the periodic grid 18, the two parts of each
    // row's x-range cx - 1 .. cx + 1 modulo n (the second empty unless it wraps).
    auto walk = [&](auto nruns) {
      constexpr int NR = decltype(nruns)::value;
      constexpr bool PER = NR == 18;
      int rs[NR], pre[NR + 1];
      pre[0] = 0;
#pragma unroll
      for (int k = 0; k < 9; k++) {
        int z = czi + k / 3 - 1, y = cyi + k % 3 - 1;
        if constexpr (!PER) {
          const int x0 = cxi > 0 ? cxi - 1 : 0, x1 = cxi + 1 < G.nx ? cxi + 1 : G.nx - 1;
          int s0 = 0, len = 0;
          if (z >= 0 && z < G.nz && y >= 0 && y < G.ny) {
            const int row = (z * G.ny + y) * G.nx;
            s0 = D.cell_start[cbase + row + x0];
            len = D.cell_start[cbase + row + x1] + D.cell_fill[cbase + row + x1] - s0;
          }
          rs[k] = s0;
          pre[k + 1] = pre[k] + len;
        } else {
          const int n = G.nx;  // nx = ny = nz
          z += z < 0 ? n : (z >= n ? -n : 0);
          y += y < 0 ? n : (y >= n ? -n : 0);
          const int row = (z * n + y) * n;
          const int a0 = cxi == 0 ? 0 : cxi - 1, a1 = cxi == n - 1 ? n - 1 : cxi + 1;
          const int bx = cxi == 0 ? n - 1 : 0;  // the wrapped cell, if any
          const bool wr = cxi == 0 || cxi == n - 1;
          const int s0 = D.cell_start[cbase + row + a0];
          rs[2 * k] = s0;
          pre[2 * k + 1] = pre[2 * k] + D.cell_start[cbase + row + a1] + D.cell_fill[cbase + row + a1] - s0;
          rs[2 * k + 1] = D.cell_start[cbase + row + bx];
          pre[2 * k + 2] = pre[2 * k + 1] + (wr ? D.cell_fill[cbase + row + bx] : 0);
        }
      }
      const int total = pre[NR];
      // position in the sorted array of candidate t: the last run starting at or before t
      auto locate = [&](int t) {
        int at = rs[0] + t;
#pragma unroll
        for (int k = 1; k < NR; k++) at = (pre[k] <= t) ? rs[k] + (t - pre[k]) : at;
        return at;
      };
      // A candidate-list build gathers the single-precision positions (and
      // types) of the planning step and keeps a superset of the pairs within
      // 3.2 sigma + skin: the cutoff is widened by the bound eps2 on the error of
      // the squared distance (as in the candidate filter), and the filter decides
      // every kept pair exactly.  A neighbor-list build, and a candidate build on
      // the periodic grid (whose partners' coordinates are not bounded by the
      // bead's own plus the reach, as eps2 assumes), use the double positions.
      constexpr bool FP = !PER;  // single-precision candidate gathers
      const double R = T->cand_reach;
      const double oxs = (double)(float)ox, oys = (double)(float)oy, ozs = (double)(float)oz;
      const double ec = 0x1p-24 * (2.0 * fmax(fabs(ox), fmax(fabs(oy), fabs(oz))) + R);
      const double eps2 = 2.0 * R * 1.7320508075688773 * ec + 3.0 * ec * ec + 1e-9;
      const double qlim = qbox - 1.0;
      // software pipeline: a chunk's candidate bead indices are loaded two
      // chunks ahead and their positions one chunk ahead, so neither load of
      // the dependent pair (index, then position) waits in the pass that issues it
      int ccn = -1, ccn2 = -1;
      double pxn = 0.0, pyn = 0.0, pzn = 0.0;
      float4 pfn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      auto fetch_pos = [&](int cc) {
        const int gc = off + cc;
        if (FP && cmode) pfn = D.fpos[gc];
        else { pxn = ux[gc]; pyn = uy[gc]; pzn = uz[gc]; }
      };
      if (l < total) {
        ccn = cb[locate(l)];
        fetch_pos(ccn);
      }
      if (l + LG < total) ccn2 = cb[locate(l + LG)];
      for (int base = 0; base < total; base += LG) {
        const int cc = ccn;
        const double px = pxn, py = pyn, pz = pzn;
        const float4 pf = pfn;
        if (base + LG + l < total) fetch_pos(ccn2);
        ccn = ccn2;
        if (base + 2 * LG + l < total) ccn2 = cb[locate(base + 2 * LG + l)];
        if (base + l >= total || !(cc - b > 2 || b - cc > 2)) continue;
        const int gc = off + cc;
        const int side = gc < g ? 1 : 0;  // 1: the partner is the pair's first bead
        bool keep;
        if (FP && cmode) {
          const int ct = (int)pf.w;
          const int tpi = side ? ct * 3 + gt : gt * 3 + ct;
          const double sx = (double)pf.x - oxs, sy = (double)pf.y - oys, sz = (double)pf.z - ozs;
          // beyond the min-image fast path: kept (the filter decides exactly)
          keep = !(fabs(sx) <= qlim && fabs(sy) <= qlim && fabs(sz) <= qlim) ||
                 sx * sx + sy * sy + sz * sz < cut2[tpi][2] + eps2;
        } else {
          const int ct = D.type[gc];
          const int tpi = side ? ct * 3 + gt : gt * 3 + ct;
          double ddx, ddy, ddz;
          pair_d3(side, ox, oy, oz, px, py, pz, boxl, qbox, ddx, ddy, ddz);
          const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
          if (cmode) {
            keep = d2 < cut2[tpi][2];  // periodic grid: the exact candidate test
          } else {
            keep = nlcrit ? d2 < cut2[tpi][0] : cell_adjacent(D, P, side ? gc : g, side ? g : gc);
            if (keep && d2 < cut2[tpi][1]) atomicOr(&plb[cc >> 5], 1u << (cc & 31));
          }
        }
        if (keep) atomicOr(&nlb[cc >> 5], 1u << (cc & 31));
      }
    };
    if (G.periodic) walk(std::integral_constant<int, 18>{});
    else walk(std::integral_constant<int, 9>{});
  }
  __syncthreads();
  if (active) {
    // natives are not in the non-native master list
    for (int q = a0 + l; q < a1; q += LG) {
      const int p = D.att_partner[q] - Pp->off;
      const uint32_t m = ~(1u << (p & 31));
      atomicAnd(&nlb[p >> 5], m);
      atomicAnd(&plb[p >> 5], m);
    }
  }
  __syncthreads();
  if (!active) return;
  // emit in ascending partner order (the master list's order for bead b)
  const ChainP& P = *Pp;
  const int off = P.off;
  const uint32_t gt = (uint32_t)D.type[g];
  const ListRow S{cmode ? D.cand : D.slot, g, D.cap, !cmode};
  const uint32_t pbit = cmode ? 0u : 1u << (30 + par);
  int cnt = 0;
  for (int base = 0; base < words; base += LG) {
    const int w = base + l;
    uint32_t bits = w < words ? nlb[w] : 0u;
    const uint32_t pls = w < words ? plb[w] : 0u;
    const int pc = __popc(bits);
    int incl = pc;
#pragma unroll
    for (int o = 1; o < LG; o <<= 1) {
      const int t = __shfl_up(incl, o, LG);
      if (l >= o) incl += t;
    }
    int idx = cnt + incl - pc;
    while (bits) {
      const int k = __builtin_ctz(bits);
      bits &= bits - 1u;
      const int c = w * 32 + k;
      const int gc = off + c;
      const int side = gc < g ? 1 : 0;
      const uint32_t ct = (uint32_t)D.type[gc];
      if (idx < D.cap)
        S[idx] = slot_make((uint32_t)c, (uint32_t)side, side ? ct : gt, side ? gt : ct) |
                 (((pls >> k) & 1u) ? pbit : 0u);
      idx++;
    }
    cnt += __shfl(incl, LG - 1, LG);
  }
  if (l == 0) {
    (cmode ? D.ccnt : D.cnt)[g] = cnt < D.cap ? cnt : D.cap;
    if (cnt > D.cap) {
      atomicOr(D.err, 1);
      // a truncated candidate list must not serve later rebuilds
      if (cmode) D.grid[D.chain[g]].cand_valid = 0;
    }
  }
}

// The chain fields the step kernel and the candidate filter read.  They are
// loaded through the scalar cache when every lane of the wave is on one chain
// (almost every wave: chains are longer than a wave's beads), so that they
// cost no vector memory (texture data) cycles, which bound both kernels;
// lanes on different chains load their own.
typedef __attribute__((address_space(4))) const ChainP ConstChainP;  // read through the scalar cache
struct ChainV {
  int off, att_off, natt, cutoff, noise_mode, iup0, nup, resident;
  double boxl, qbox, thr_att, var;
  uint64_t seed, noise_base;
  template <class P_>
  __device__ __forceinline__ void load(P_* __restrict__ P) {
    off = P->off; att_off = P->att_off; natt = P->natt; cutoff = P->cutoff; noise_mode = P->noise_mode;
    iup0 = P->iup0; nup = P->nup; resident = P->resident;
    boxl = P->boxl; qbox = P->qbox; thr_att = P->thr_att; var = P->var;
    seed = P->seed; noise_base = P->noise_base;
  }
};
// ... the candidate filter's
struct ChainF {
  int off;
  double boxl, qbox;
  template <class P_>
  __device__ __forceinline__ void load(P_* __restrict__ P) {
    off = P->off; boxl = P->boxl; qbox = P->qbox;
  }
};
// ... and the integration's, read in the epilogue
struct ChainI {
  int vel_on, sim_type;
  double a1, a2, a3, a4, a5;
  template <class P_>
  __device__ __forceinline__ void load(P_* __restrict__ P) {
    vel_on = P->vel_on; sim_type = P->sim_type;
    a1 = P->a1; a2 = P->a2; a3 = P->a3; a4 = P->a4; a5 = P->a5;
  }
};
template <class V>
__device__ __forceinline__ V wave_chain_fields(const ChainP* __restrict__ cp, int chain) {
  V v;
  const int c0 = __builtin_amdgcn_readfirstlane(chain);
  if (__builtin_amdgcn_ballot_w64(chain != c0) == 0)
    v.load((ConstChainP*)(cp + c0));
  else
    v.load(cp + chain);
  return v;
}

// Neighbor list from the candidate list (chains planned W_FILTER): the
// candidates within the neighbor-list cutoff at the current positions, in
// their (ascending, master) order, with the pair-list bit at the same d^2;
// natives as in the full scan.  One group of LF lanes per bead.
template <int LF>
__global__ void __launch_bounds__(256) k_cand_filter(Dev D, const Tables* __restrict__ T, int cur, int par, int rel,
                                                     int g0, int g1, const int* __restrict__ tiles, int ntiles) {
  rel += *D.relb;
  if (D.plan[1] != rel) return;  // block-uniform: before the barrier
  // thread t: bead g0 + t / LF, or with a tile table (the tiles of kBeadTile
  // beads that hold a bead of [g0, g1)) bead (t / LF) % kBeadTile of tile
  // tiles[t / (LF * kBeadTile)]
  const int t = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  int g = g0 + t / LF;
  bool act = true;
  if (tiles) {
    const int k = t / (LF * kBeadTile);
    act = k < ntiles;
    if (act) g = tiles[k] * kBeadTile + (t / LF) % kBeadTile;
  }
  act = act && g >= g0 && g < g1;
  const int l = t & (LF - 1);
  const double* ux = cur ? D.ux[1] : D.ux[0];
  const double* uy = cur ? D.uy[1] : D.uy[0];
  const double* uz = cur ? D.uz[1] : D.uz[0];
  // the bead's own loads go out with the cutoff table's, before the barrier
  int c = -1, n = 0;
  double cx = 0.0, cy = 0.0, cz = 0.0;
  if (act) {
    c = D.chain[g];
    n = D.ccnt[g];
    cx = ux[g]; cy = uy[g]; cz = uz[g];
  }
  __shared__ double cut2[9][2];  // per type pair (it * 3 + jt): (3.2 sigma)^2, (2.5 sigma)^2
  if (threadIdx.x < 9) {
    cut2[threadIdx.x][0] = T->nlrc2[threadIdx.x / 3][threadIdx.x % 3];
    cut2[threadIdx.x][1] = T->plrc2[threadIdx.x / 3][threadIdx.x % 3];
  }
  __syncthreads();
  if (!act || c < 0 || !(D.grid[c].work & W_FILTER)) return;
  const ChainF cf = wave_chain_fields<ChainF>(D.cp, c);
  const double boxl = cf.boxl, qbox = cf.qbox;
  const int off = cf.off;
  // non-natives (neighbor_list.cpp:97-103): stable compaction of the
  // candidates.  Software pipeline as in the step kernel (candidate words two
  // chunks ahead, partner coordinates one chunk ahead), started before the
  // natives so that their loads overlap.
  const uint32_t* C = D.cand + (size_t)g * D.cap;
  const ListRow S{D.slot, g, D.cap, true};
  // The candidates' positions are gathered in single precision (half the
  // data of the double positions).  A single-precision coordinate is within
  // 2^-24 |x| of the double one, so for a partner within R = cand_reach (every
  // candidate of a valid list is) each component of the displacement is within
  // ec of the exact one and the squared distance within eps2 (a bound on
  // 2 R sqrt(3) ec + 3 ec^2, with room for the double roundings).  A candidate
  // whose distance lies within eps2 of a cutoff, or near the min-image limit,
  // is decided in double precision from the double positions.
  const double R = T->cand_reach;
  const double cxs = (double)(float)cx, cys = (double)(float)cy, czs = (double)(float)cz;
  const double ec = 0x1p-24 * (2.0 * fmax(fabs(cx), fmax(fabs(cy), fabs(cz))) + R);
  const double eps2 = 2.0 * R * 1.7320508075688773 * ec + 3.0 * ec * ec + 1e-9;
  const double qlim = qbox - 1.0;
  uint32_t wn = 0u, wn2 = 0u;
  float4 pn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (l < n) {
    wn = C[l];
    pn = D.fpos[off + (int)(wn & kPartnerMask)];
  }
  if (l + LF < n) wn2 = C[l + LF];
  // natives (neighbor_list.cpp:17-67): the displacement from the bead's own
  // position as the step kernel forms it, the cutoffs from sigma
  const int a0 = D.att_beg[g], a1 = D.att_end[g];
  for (int q = a0 + l; q < a1; q += LF) {
    const uint32_t aw = D.att_w[q];
    const double r = D.att_r[q];
    const int o = off + (int)(aw & kPartnerMask), side = (int)(aw >> (kAttTShift + 4)) & 1;
    double dx, dy, dz;
    pair_d3(side, cx, cy, cz, ux[o], uy[o], uz[o], boxl, qbox, dx, dy, dz);
    const double d2 = dx * dx + dy * dy + dz * dz;
    const bool nl = d2 < nat_nlrc2(r), pl = nl && d2 < nat_plrc2(r);
    D.att_flag[q] = (uint8_t)((nl ? 1 : 0) | (pl ? (2 << par) : 0));
  }
  const uint32_t pbit = 1u << (30 + par);
  int cnt = 0;
  for (int base = 0; base < n; base += LF) {
    const int q = base + l;
    uint32_t w = wn;
    const float4 pf = pn;
    if (q + LF < n) pn = D.fpos[off + (int)(wn2 & kPartnerMask)];
    wn = wn2;
    if (q + 2 * LF < n) wn2 = C[q + 2 * LF];
    bool keep = false;
    if (q < n) {
      const int side = (w >> 25) & 1, tpi = (int)((w >> 26) & 3) * 3 + (int)((w >> 28) & 3);
      const double sx = (double)pf.x - cxs, sy = (double)pf.y - cys, sz = (double)pf.z - czs;
      const double s2 = sx * sx + sy * sy + sz * sz;
      const double cn = cut2[tpi][0], cp = cut2[tpi][1];
      bool pl;
      if (fabs(sx) <= qlim && fabs(sy) <= qlim && fabs(sz) <= qlim && fabs(s2 - cn) > eps2 &&
          fabs(s2 - cp) > eps2) {
        keep = s2 < cn;
        pl = s2 < cp;
      } else {
        // the exact decision, in double precision from the double positions
        const int o = off + (int)(w & kPartnerMask);
        double dx, dy, dz;
        pair_d3(side, cx, cy, cz, ux[o], uy[o], uz[o], boxl, qbox, dx, dy, dz);
        const double d2 = dx * dx + dy * dy + dz * dz;
        keep = d2 < cn;
        pl = d2 < cp;
      }
      w = (w & ~(3u << 30)) | ((keep && pl) ? pbit : 0u);
    }
    const uint64_t m = group_mask<LF>(keep);
    if (keep) S[cnt + __popcll(m & ((1ull << l) - 1ull))] = w;
    cnt += __popcll(m);
  }
  if (l == 0) D.cnt[g] = cnt;  // <= the candidate count <= cap
}

// Pair-list filter at the current positions (sop.cpp:473-553), parity par.
template <int L>
__device__ __forceinline__ void pairfilter_bead(const Dev& D, const Tables* __restrict__ T, int cur, int par, int g,
                                                int l) {
  const ChainP& P = D.cp[D.chain[g]];
  const double* ux = cur ? D.ux[1] : D.ux[0];
  const double* uy = cur ? D.uy[1] : D.uy[0];
  const double* uz = cur ? D.uz[1] : D.uz[0];
  const double boxl = P.boxl, qbox = P.qbox;
  const uint8_t pbit = (uint8_t)(2 << par);
  const int a0 = D.att_beg[g], a1 = D.att_end[g];
  for (int q = a0 + l; q < a1; q += L) {
    const int side = D.att_code[q] & 1, o = D.att_partner[q];
    const uint8_t f = D.att_flag[q];
    bool pl = false;
    if (f & 1) {
      const int i = side ? o : g, j = side ? g : o;
      const double dx = mimg(ux[j] - ux[i], boxl, qbox);
      const double dy = mimg(uy[j] - uy[i], boxl, qbox);
      const double dz = mimg(uz[j] - uz[i], boxl, qbox);
      const double d2 = dx * dx + dy * dy + dz * dz;
      pl = (P.cutoff == SOPG_CUTOFF_NONE) ? true : d2 < nat_plrc2(D.att_r[q]);
    }
    D.att_flag[q] = (uint8_t)((f & ~pbit) | (pl ? pbit : 0));
  }
  const int cnt = D.cnt[g];
  const ListRow S{D.slot, g, D.cap, true};
  const uint32_t sbit = 1u << (30 + par);
  for (int q = l; q < cnt; q += L) {
    const uint32_t w = S[q];
    const int o = P.off + (int)(w & kPartnerMask), side = (w >> 25) & 1, it = (w >> 26) & 3, jt = (w >> 28) & 3;
    const int i = side ? o : g, j = side ? g : o;
    const double dx = mimg(ux[j] - ux[i], boxl, qbox);
    const double dy = mimg(uy[j] - uy[i], boxl, qbox);
    const double dz = mimg(uz[j] - uz[i], boxl, qbox);
    const double d2 = dx * dx + dy * dy + dz * dz;
    const bool pl = (P.cutoff == SOPG_CUTOFF_NONE) ? true : d2 < T->plrc2[it][jt];
    S[q] = (w & ~sbit) | (pl ? sbit : 0u);
  }
}

template <int L>
__global__ void __launch_bounds__(256) k_pairfilter(Dev D, const Tables* __restrict__ T, int cur, int par, int g0, int g1) {
  const int t = xcd_block() * blockDim.x + threadIdx.x;
  const int g = g0 + t / L;
  if (g >= g1 || D.chain[g] < 0) return;  // past the end, or a free bead
  if (D.cp[D.chain[g]].resident) return;  // the resident kernel filters its own
  pairfilter_bead<L>(D, T, cur, par, g, t & (L - 1));
}

// Position half of underdamped_iteration (sop.cpp:288-310): incr, pos, unc_pos.
__device__ __forceinline__ void posupdate_bead(const Dev& D, int src, int dst, int g) {
  const ChainP& P = D.cp[D.chain[g]];
  // underdamped: incr = a1 v + a2 F (sop.cpp:292-294); overdamped: incr = a5 F (sop.cpp:339-341)
  const bool od = P.sim_type == 2;
  const double ix = od ? P.a5 * D.fx[g] : P.a1 * D.vx[g] + P.a2 * D.fx[g];
  const double iy = od ? P.a5 * D.fy[g] : P.a1 * D.vy[g] + P.a2 * D.fy[g];
  const double iz = od ? P.a5 * D.fz[g] : P.a1 * D.vz[g] + P.a2 * D.fz[g];
  (dst ? D.ix[1] : D.ix[0])[g] = ix; (dst ? D.iy[1] : D.iy[0])[g] = iy; (dst ? D.iz[1] : D.iz[0])[g] = iz;
  (dst ? D.px[1] : D.px[0])[g] = mimg((src ? D.px[1] : D.px[0])[g] + ix, P.boxl, P.qbox);
  (dst ? D.py[1] : D.py[0])[g] = mimg((src ? D.py[1] : D.py[0])[g] + iy, P.boxl, P.qbox);
  (dst ? D.pz[1] : D.pz[0])[g] = mimg((src ? D.pz[1] : D.pz[0])[g] + iz, P.boxl, P.qbox);
  (dst ? D.ux[1] : D.ux[0])[g] = (src ? D.ux[1] : D.ux[0])[g] + ix;
  (dst ? D.uy[1] : D.uy[0])[g] = (src ? D.uy[1] : D.uy[0])[g] + iy;
  (dst ? D.uz[1] : D.uz[0])[g] = (src ? D.uz[1] : D.uz[0])[g] + iz;
}

__global__ void k_posupdate(Dev D, int src, int dst, int g0, int g1) {
  const int g = g0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= g1 || D.chain[g] < 0) return;  // past the end, or a free bead
  if (D.cp[D.chain[g]].resident) return;  // the resident kernel moves its own
  posupdate_bead(D, src, dst, g);
}

// One force evaluation (force_eval, energy.cpp:21-33) at unc_pos[cur] with the
// pair list of parity pr, fused with what the step loop needs around it.
// flags: F_STORE   store forces (else energy only)
//        F_VEL     velocity update (sop.cpp:320-328) when the chain's T allows
//        F_ENERGY  energy partials (energy.cpp:115-257) on chains whose record
//                  falls on this step (all chains if F_ENERGY_ALL)
// NEXT:  also the next step's pair list (parity pr^1) from these positions,
//        and with F_ADVANCE the next position update into buffer cur^1
//        (a run's last step computes the pair list its next run starts with)
//        F_NOISE / F_BOND / F_ANGLE / F_VDW  which force terms to add (the
//                  reference's force_term[] table, energy.cpp:78-113)
//        F_ACCUM   add to the stored force instead of starting from zero
enum { F_STORE = 1, F_VEL = 2, F_ENERGY = 4, F_ENERGY_ALL = 8, F_NOISE = 16, F_BOND = 32, F_ANGLE = 64,
       F_VDW = 128, F_ACCUM = 256, F_ADVANCE = 512 };

// Type-pair tables the pair loops read, staged in LDS: row it + 4 * jt, the
// 4-bit type field of a slot word (bits 26-29) or of a native incidence word,
// so that a row's address is that field times 64 bytes.  Rows of 8 doubles:
// the non-native loop reads entries 0-3 of a row as two 16-byte LDS loads.
enum { TP_FREP = 0, TP_S6, TP_S12, TP_PLRC2, TP_CREP, TP_FATT, TP_CATT, TP_PAD, TP_N };
constexpr int kTypeRows = 16;

__device__ __forceinline__ void load_type_tables(const Tables* __restrict__ T, double (*tp)[TP_N]) {
  if (threadIdx.x < kTypeRows) {
    const int it = threadIdx.x & 3, jt = threadIdx.x >> 2;
    const bool ok = it < 3 && jt < 3;  // types are 0..2
    const int a = ok ? it : 0, b = ok ? jt : 0;
    tp[threadIdx.x][TP_FREP] = T->frep[a][b];
    tp[threadIdx.x][TP_S6] = T->s6[a][b];
    tp[threadIdx.x][TP_S12] = T->s12[a][b];
    tp[threadIdx.x][TP_PLRC2] = T->plrc2[a][b];
    tp[threadIdx.x][TP_CREP] = T->crep[a][b];
    tp[threadIdx.x][TP_FATT] = T->fatt[a][b];
    tp[threadIdx.x][TP_CATT] = T->catt[a][b];
  }
}

// The step for bead g, computed by the L lanes of its group (lane l); the
// body of k_force (SKIP_RES: chains the resident kernel advances are left
// alone) and of the resident kernel.

// EN false: the launch records no energy (no chain's record step), and the
// energy terms are compiled out.
// The bead's own fields.  bead_head_load issues the loads and bead_head_share
// completes them, so that a kernel can put the loads out before its block
// barrier.  L = 4: the group's four lanes share the fields: each lane loads
// two of the eight index fields and one coordinate (into chain, bond0 and cx)
// and the group exchanges them (a quarter of the vector-memory data of every
// lane loading every field).
struct BeadHead {
  int chain, bond0, bond1, ang0, ang1, att0, att1, nslot;
  double cx, cy, cz;
};
template <int L>
__device__ __forceinline__ void bead_head_load(const Dev& D, const double* __restrict__ ux,
                                               const double* __restrict__ uy, const double* __restrict__ uz, int g,
                                               int l, BeadHead& h) {
  if constexpr (L == 4) {
    const int* __restrict__ pa = l == 0 ? D.chain : l == 1 ? D.bond_end : l == 2 ? D.ang_end : D.att_end;
    const int* __restrict__ pb = l == 0 ? D.bond_beg : l == 1 ? D.ang_beg : l == 2 ? D.att_beg : D.cnt;
    const double* __restrict__ pc = l == 0 ? ux : l == 1 ? uy : uz;
    h.chain = pa[g];
    h.bond0 = pb[g];
    h.cx = pc[g];
  } else {
    h.chain = D.chain[g];
    h.cx = ux[g]; h.cy = uy[g]; h.cz = uz[g];
    h.bond0 = D.bond_beg[g]; h.bond1 = D.bond_end[g]; h.ang0 = D.ang_beg[g]; h.ang1 = D.ang_end[g];
    h.att0 = D.att_beg[g]; h.att1 = D.att_end[g]; h.nslot = D.cnt[g];
  }
}
template <int L>
__device__ __forceinline__ void bead_head_share(BeadHead& h) {
  if constexpr (L == 4) {
    const int va = h.chain, vb = h.bond0;
    const double vc = h.cx;
    h.chain = group_bcast_i<0, 4>(va); h.bond0 = group_bcast_i<0, 4>(vb);
    h.bond1 = group_bcast_i<1, 4>(va); h.ang0 = group_bcast_i<1, 4>(vb);
    h.ang1 = group_bcast_i<2, 4>(va); h.att0 = group_bcast_i<2, 4>(vb);
    h.att1 = group_bcast_i<3, 4>(va); h.nslot = group_bcast_i<3, 4>(vb);
    h.cx = group_bcast<0, 4>(vc); h.cy = group_bcast<1, 4>(vc); h.cz = group_bcast<2, 4>(vc);
  }
}

template <int L, bool NEXT, bool SKIP_RES = false, bool EN = true>
__device__ __forceinline__ void force_bead(const Dev& D, const Tables* __restrict__ T, const double (*tp)[TP_N],
                                           int cur, int pr, int rel, int flags, const double* __restrict__ noise,
                                           int g, int l, const BeadHead* pre = nullptr,
                                           const ChainV* pre_cv = nullptr) {
  // loads in two batches: bead-indexed ones (any bead below the high-water
  // mark is valid memory), then the chain's fields; pre, pre_cv: loaded (and
  // shared) by the caller
  const double* __restrict__ ux = cur ? D.ux[1] : D.ux[0];
  const double* __restrict__ uy = cur ? D.uy[1] : D.uy[0];
  const double* __restrict__ uz = cur ? D.uz[1] : D.uz[0];
  const __amdgpu_buffer_rsrc_t bux = buf_view(ux), buy = buf_view(uy), buz = buf_view(uz);
  BeadHead h;
  if (pre) {
    h = *pre;
  } else {
    bead_head_load<L>(D, ux, uy, uz, g, l, h);
    bead_head_share<L>(h);
  }
  const int chain = h.chain, bond0 = h.bond0, bond1 = h.bond1, ang0 = h.ang0, ang1 = h.ang1, att0 = h.att0,
            att1 = h.att1, nslot = h.nslot;
  const double cx = h.cx, cy = h.cy, cz = h.cz;
  if (chain < 0) return;  // a free bead
  // chain fields, in registers (the loops' stores could alias them)
  const ChainV cv = pre_cv ? *pre_cv : wave_chain_fields<ChainV>(D.cp, chain);
  const int off = cv.off, att_off = cv.att_off, natt = cv.natt;
  const int cutoff = cv.cutoff, noise_mode = cv.noise_mode, iup0 = cv.iup0, nup = cv.nup;
  const int resident = cv.resident;
  const double boxl = cv.boxl, qbox = cv.qbox, thr_att = cv.thr_att, var = cv.var;
  const uint64_t seed = cv.seed, noise_base = cv.noise_base;
  if (SKIP_RES && resident) return;
  const uint32_t uoff = (uint32_t)off;  // partner indices unsigned: no sign extension in addresses
  const bool no_cut = cutoff == SOPG_CUTOFF_NONE;
  const int pw = pr ^ 1;
  const bool energy = EN && ((flags & F_ENERGY_ALL) || ((flags & F_ENERGY) && ((iup0 + rel) % nup == 0)));
  double eb = 0.0, ea = 0.0, eatt = 0.0, erep = 0.0;

  // clear_forces, then random_force (energy.cpp:35-45, 449-465):
  // force = 0 + var * gasdev() for x, y, z
  double gx = 0.0, gy = 0.0, gz = 0.0;
  const bool with_noise = (flags & F_STORE) && (flags & F_NOISE);
  if (!with_noise) {
    // energy only, or the random term not requested
  } else if (noise_mode == SOPG_NOISE_REPLAY) {
    gx = noise[3 * (size_t)g];
    gy = noise[3 * (size_t)g + 1];
    gz = noise[3 * (size_t)g + 2];
  } else {
    const uint64_t ctr = noise_base + (uint64_t)rel;
    const uint32_t bead = (uint32_t)(g - off);
    if constexpr (L == 1) {
      double u;
      philox_gauss2(seed, ctr, bead, 0, gx, gy);
      philox_gauss2(seed, ctr, bead, 1, gz, u);
    } else {
      double a, c;
      philox_gauss2(seed, ctr, bead, (uint32_t)(l & 1), a, c);
      if constexpr (L == 2 || L == 4) {
        gx = group_bcast<0, L>(a);
        gy = group_bcast<0, L>(c);
        gz = group_bcast<1, L>(a);
      } else {
        gx = __shfl(a, 0, L);
        gy = __shfl(c, 0, L);
        gz = __shfl(a, 1, L);
      }
    }
  }
  double Fx = (flags & F_ACCUM) ? D.fx[g] : 0.0;
  double Fy = (flags & F_ACCUM) ? D.fy[g] : 0.0;
  double Fz = (flags & F_ACCUM) ? D.fz[g] : 0.0;
  if (with_noise) {
    Fx += var * gx;
    Fy += var * gy;
    Fz += var * gz;
  }

  // FENE bonds (energy.cpp:402-447): F_i -= f, F_j += f; then soft-sphere
  // (i, i+2) repulsion (energy.cpp:354-400): F_i -= f, F_k += f.  One loop
  // over the bead's bond incidences followed by its angle incidences, so a
  // chunk of L lanes covers both (about two of each per bead) and the fold
  // still adds every bond before every angle, each in list order.
  if (flags & (F_BOND | F_ANGLE)) {
    const int b0 = (flags & F_BOND) ? bond0 : 0, nbq = (flags & F_BOND) ? bond1 - b0 : 0;
    const int a0 = (flags & F_ANGLE) ? ang0 : 0, naq = (flags & F_ANGLE) ? ang1 - a0 : 0;
    const int ntot = nbq + naq;
    for (int base = 0; base < ntot; base += L) {
      const int q = base + l;
      const bool has = q < ntot, isb = q < nbq;
      // -0.0: what a lane without a term adds in the folds
      double cs = 0.0, dx = 0.0, dy = 0.0, dz = 0.0, e_b = -0.0, e_a = -0.0;
      bool ec = false;
      if (has) {
        const int pq = isb ? D.bond_pq[b0 + q] : D.ang_pq[a0 + q - nbq], o = pq >> 1, side = pq & 1;
        const double sg = side ? -1.0 : 1.0;
        pair_d3(side, cx, cy, cz, ux[o], uy[o], uz[o], boxl, qbox, dx, dy, dz);
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        double co1;
        if (isb) {
          const double dev = d - D.bond_r0[b0 + q];
          const double dev2 = dev * dev;
          co1 = -T->k_bnd * dev / d / (1.0 - dev2 / T->R0sq);
          if (energy && side == 0) { e_b = log1p(-dev * dev / T->R0sq); ec = true; }
        } else {
          co1 = T->f_ang_ss_coeff / pow8_cr(d);
          if (energy && side == 0) { e_a = T->e_ang_ss_coeff / pow6_cr(d); ec = true; }
        }
        // side 0 negates the force: (-co1 * 1) * d is -(co1 * d) bit for bit
        cs = -co1 * sg;
      }
      double fx, fy, fz;
      neutral3(has, cs, dx, dy, dz, fx, fy, fz);
      fold3<L>(Fx, Fy, Fz, has, fx, fy, fz);
      if (energy) {
        fold1<L>(eb, ec && isb, e_b);
        fold1<L>(ea, ec && !isb, e_a);
      }
    }
  }
  // native LJ 12-6 over the pair list (energy.cpp:274-311): F_i += f, F_j -= f
  if (flags & F_VDW) {
    const int p0 = att0, p1 = att1;
    // software pipeline as in the non-native loop below
    uint8_t fn = 0, fn2 = 0;
    uint32_t an = 0u, an2 = 0u;
    double pxn = 0.0, pyn = 0.0, pzn = 0.0;
    if (p0 + l < p1) {
      fn = D.att_flag[p0 + l];
      an = D.att_w[p0 + l];
      const uint32_t ob = (uoff + (an & kPartnerMask)) * 8u;
      pxn = buf_ld(bux, ob); pyn = buf_ld(buy, ob); pzn = buf_ld(buz, ob);
    }
    if (p0 + l + L < p1) { fn2 = D.att_flag[p0 + l + L]; an2 = D.att_w[p0 + l + L]; }
    const int p1a = p1 - L, p1b = p1 - 2 * L;  // q < p1a: entry q + L exists
    for (int base = p0; base < p1; base += L) {
      const int q = base + l;
      const uint8_t f = fn;
      const uint32_t aw = an;
      const double px = pxn, py = pyn, pz = pzn;
      if (q < p1a) {
        const uint32_t ob = (uoff + (an2 & kPartnerMask)) * 8u;
        pxn = buf_ld(bux, ob); pyn = buf_ld(buy, ob); pzn = buf_ld(buz, ob);
      }
      fn = fn2;
      an = an2;
      if (q < p1b) { fn2 = D.att_flag[q + 2 * L]; an2 = D.att_w[q + 2 * L]; }
      bool contrib = false, ec = false;
      // the displacement and cs are read (by neutral3s) only where contrib:
      // elsewhere they are left unset (no register is cleared for them)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wsometimes-uninitialized"
      double cs, dx, dy, dz, e = -0.0;
      {
        if (q < p1 && (f & 1)) {
          const double r = D.att_r[q];
          const int tt = (int)(aw >> kAttTShift), side = (tt >> 4) & 1, tpi = tt & 15;
          const double sg = side ? -1.0 : 1.0;
          pair_d3(side, cx, cy, cz, px, py, pz, boxl, qbox, dx, dy, dz);
          const double d2 = dx * dx + dy * dy + dz * dz;
          if ((f >> (1 + pr)) & 1) {
            double r6, r12;
            nat_powers(r, r6, r12);
            const double d6 = d2 * d2 * d2;
            const double d12 = d6 * d6;
            const bool skip = (d2 < thr_att) &&
                              att_tol_skip(D.att_flag, D.am_q0, D.am_r2, att_off, D.att_code[q] >> 1, pr, d2);
            if (!skip) {
              // side 1 negates the force: (co1 * -1) * d is -(co1 * d) bit for bit
              cs = (tp[tpi][TP_FATT] / d2 * ((r12 / d12) - (r6 / d6))) * sg;
              contrib = true;
            }
            if (energy && side == 0) { e = tp[tpi][TP_CATT] * (r12 / d12) - 2.0 * (r6 / d6); ec = true; }
          }
          if constexpr (NEXT) {
            // the next step's bit, stored only where it changes
            const bool npl = no_cut || d2 < nat_plrc2(r);
            if (npl != (((f >> (1 + pw)) & 1) != 0)) D.att_flag[q] = (uint8_t)(f ^ (2 << pw));
          }
        }
      }
      double fx, fy, fz;
      neutral3s(contrib, cs, dx, dy, dz, fx, fy, fz);
#pragma clang diagnostic pop
      fold3<L>(Fx, Fy, Fz, contrib, fx, fy, fz);
      if (energy) fold1<L>(eatt, ec, e);
    }
  }
  // non-native repulsion over the pair list (energy.cpp:313-350): F_i += f, F_j -= f
  if (flags & F_VDW) {
    const int cnt = nslot;
    const ListRow S{D.slot, g, D.cap, true};
    const uint32_t rbit = 1u << (30 + pr), wbit = 1u << (30 + pw);
    const double thr_rep = T->thr_rep;
    // software pipeline: the partner coordinates of the next chunk and the
    // slot words of the one after it are loaded before the current chunk is
    // evaluated and folded, so no load waits on one issued in the same pass
    uint32_t wn = 0u, wn2 = 0u;
    double pxn = 0.0, pyn = 0.0, pzn = 0.0;
    if (l < cnt) {
      wn = S[l];
      const uint32_t ob = (uoff + (wn & kPartnerMask)) * 8u;
      pxn = buf_ld(bux, ob); pyn = buf_ld(buy, ob); pzn = buf_ld(buz, ob);
    }
    if (l + L < cnt) wn2 = S[l + L];
    const int cnt1 = cnt - L, cnt2 = cnt - 2 * L;  // q < cnt1: chunk q + L exists
    for (int base = 0; base < cnt; base += L) {
      const int q = base + l;
      const uint32_t w = wn;
      const double px = pxn, py = pyn, pz = pzn;
      if (q < cnt1) {
        const uint32_t ob = (uoff + (wn2 & kPartnerMask)) * 8u;
        pxn = buf_ld(bux, ob); pyn = buf_ld(buy, ob); pzn = buf_ld(buz, ob);
      }
      wn = wn2;
      if (q < cnt2) wn2 = S[q + 2 * L];
      // every lane evaluates the displacement (a lane past the list end holds
      // finite stale coordinates); the branches cover only what needs the entry
      const bool in = q < cnt;
      const int side = (w >> 25) & 1;
      const double sg = side ? -1.0 : 1.0;
      const int tpi = (int)((w >> 26) & 15);
      double dx, dy, dz;
      pair_d3(side, cx, cy, cz, px, py, pz, boxl, qbox, dx, dy, dz);
      const double d2 = dx * dx + dy * dy + dz * dz;
      // the type pair's constants, read together (one LDS round trip)
      const double frep = tp[tpi][TP_FREP], s6 = tp[tpi][TP_S6], s12 = tp[tpi][TP_S12];
      const double plrc2 = tp[tpi][TP_PLRC2];
      bool contrib = false, ec = false;
      // cs is read (by neutral3s) only where contrib
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wsometimes-uninitialized"
      double cs, e = -0.0;
      if (in && (w & rbit)) {
        const double d6 = d2 * d2 * d2;
        const double d12 = d6 * d6;
        const bool skip = (d2 < thr_rep) && rep_tol_skip(D.att_flag, D.am_q0, D.am_t, T, att_off, natt, pr, d2);
        if (!skip) {
          // side 1 negates the force: (co1 * -1) * d is -(co1 * d) bit for bit
          cs = (frep / d2 * (2.0 * s12 / d12 + s6 / d6)) * sg;
          contrib = true;
        }
        if (energy && side == 0) { e = tp[tpi][TP_CREP] * (s12 / d12 + s6 / d6); ec = true; }
      }
      if constexpr (NEXT) {
        // the next step's bit, stored only where it changes
        const bool npl = no_cut || d2 < plrc2;
        if (in && npl != ((w & wbit) != 0u)) S[q] = w ^ wbit;
      }
      double fx, fy, fz;
      neutral3s(contrib, cs, dx, dy, dz, fx, fy, fz);
#pragma clang diagnostic pop
      fold3<L>(Fx, Fy, Fz, contrib, fx, fy, fz);
      if (energy) fold1<L>(erep, ec, e);
    }
  }

  if (l != 0) return;
  if (energy) {
    double* r = D.red + (size_t)g * kRed;
    r[R_EBND] = eb; r[R_EANG] = ea; r[R_EATT] = eatt; r[R_EREP] = erep;
  }
  if (!(flags & F_STORE)) return;
  // every load of the integration first (one memory round trip), then the
  // arithmetic and the stores
  const ChainI ci = wave_chain_fields<ChainI>(D.cp, chain);
  const int vel_on = ci.vel_on, sim = ci.sim_type;
  const double a1 = ci.a1, a2 = ci.a2, a3 = ci.a3, a4 = ci.a4, a5 = ci.a5;
  const double vx0 = D.vx[g], vy0 = D.vy[g], vz0 = D.vz[g];
  const double ix0 = (cur ? D.ix[1] : D.ix[0])[g], iy0 = (cur ? D.iy[1] : D.iy[0])[g];
  const double iz0 = (cur ? D.iz[1] : D.iz[0])[g];
  double psx = 0.0, psy = 0.0, psz = 0.0;
  const bool adv = NEXT && (flags & F_ADVANCE);
  if (adv) {
    psx = (cur ? D.px[1] : D.px[0])[g];
    psy = (cur ? D.py[1] : D.py[0])[g];
    psz = (cur ? D.pz[1] : D.pz[0])[g];
  }
  const bool vel = (flags & F_VEL) && vel_on && sim != 2;
  const bool od = sim == 2;
  D.fx[g] = Fx; D.fy[g] = Fy; D.fz[g] = Fz;
  double vx = vx0, vy = vy0, vz = vz0;
  if (vel) {
    vx = a3 * ix0 + a4 * Fx;
    vy = a3 * iy0 + a4 * Fy;
    vz = a3 * iz0 + a4 * Fz;
    D.vx[g] = vx; D.vy[g] = vy; D.vz[g] = vz;
  }
  if (adv) {
    // the next step's position half (sop.cpp:288-310) into buffer cur^1
    const double ix = od ? a5 * Fx : a1 * vx + a2 * Fx;
    const double iy = od ? a5 * Fy : a1 * vy + a2 * Fy;
    const double iz = od ? a5 * Fz : a1 * vz + a2 * Fz;
    (cur ? D.ix[0] : D.ix[1])[g] = ix; (cur ? D.iy[0] : D.iy[1])[g] = iy; (cur ? D.iz[0] : D.iz[1])[g] = iz;
    (cur ? D.px[0] : D.px[1])[g] = mimg(psx + ix, boxl, qbox);
    (cur ? D.py[0] : D.py[1])[g] = mimg(psy + iy, boxl, qbox);
    (cur ? D.pz[0] : D.pz[1])[g] = mimg(psz + iz, boxl, qbox);
    (cur ? D.ux[0] : D.ux[1])[g] = cx + ix;
    (cur ? D.uy[0] : D.uy[1])[g] = cy + iy;
    (cur ? D.uz[0] : D.uz[1])[g] = cz + iz;
  }
}

template <int L, bool NEXT, bool EN>
__global__ void __launch_bounds__(256, 5) k_force(Dev D, const Tables* __restrict__ T, int cur, int pr, int rel, int flags,
                                               const double* __restrict__ noise, int g0, int g1,
                                               const int* __restrict__ blk) {
  rel += *D.relb;
  __shared__ __align__(16) double tp[kTypeRows][TP_N];
  // block b serves beads g0 + [b, b + 1) * per, or with a block table (the
  // blocks of [0, g1) that hold a bead to step) beads [blk[b], blk[b] + 1) * per
  const int per = (int)blockDim.x / L, b = xcd_block();
  const int g = (blk ? blk[b] * per : g0 + b * per) + (int)threadIdx.x / L;
  const int l = (int)threadIdx.x & (L - 1);
  // the bead's own loads go out with the type tables', and its chain's
  // fields once the chain id is in, all before the barrier
  BeadHead h;
  ChainV cv;
  if (g < g1) bead_head_load<L>(D, cur ? D.ux[1] : D.ux[0], cur ? D.uy[1] : D.uy[0], cur ? D.uz[1] : D.uz[0], g, l, h);
  load_type_tables(T, tp);
  if (g < g1) {
    bead_head_share<L>(h);
    if (h.chain >= 0) cv = wave_chain_fields<ChainV>(D.cp, h.chain);
  }
  __syncthreads();
  if (g >= g1) return;
  force_bead<L, NEXT, true, EN>(D, T, tp, cur, pr, rel, flags, noise, g, l, &h, &cv);
}

// Observables (sop.cpp:200-280) and list counts, per bead.
enum { OBS_PHYS = 1, OBS_COUNTS = 2 };
// Rg^2 pair sum (sop.cpp:245-258) of bead b over its partners c > b, in
// kObsLanes strided parts (part l: c = b + 1 + l, b + 1 + l + kObsLanes, ...)
// combined by a fixed tree (part l adds part l + 8, then l + 4, l + 2, l + 1):
// one lane per part in k_observe, all parts in one thread in the resident
// kernel, the same bits either way.
constexpr int kObsLanes = 16;
__device__ __forceinline__ double rg_part(const double* ux, const double* uy, const double* uz, int off, int n,
                                          int b, int l, double boxl, double qbox) {
  const int g = off + b;
  const double x = ux[g], y = uy[g], z = uz[g];
  double rg = 0.0;
  for (int c = b + 1 + l; c < n; c += kObsLanes) {
    const int h = off + c;
    const double dx = mimg(x - ux[h], boxl, qbox);
    const double dy = mimg(y - uy[h], boxl, qbox);
    const double dz = mimg(z - uz[h], boxl, qbox);
    rg += (dx * dx + dy * dy + dz * dz);
  }
  return rg;
}

// Observables of bead g.  LO = kObsLanes: the bead's group of lanes shares the
// Rg^2 pairs (lane l its part l) and lane 0 does the rest; LO = 1: one thread.
template <int LO>
__device__ __forceinline__ void observe_bead(const Dev& D, int cur, int par, int rel, int all, int what, int g,
                                             int l) {
  const ChainP& P = D.cp[D.chain[g]];
  if (!all && (P.iup0 + rel) % P.nup != 0) return;
  const double* ux = cur ? D.ux[1] : D.ux[0];
  const double* uy = cur ? D.uy[1] : D.uy[0];
  const double* uz = cur ? D.uz[1] : D.uz[0];
  const double boxl = P.boxl, qbox = P.qbox;
  double rg = 0.0;
  if (what & OBS_PHYS) {
    const int b = g - P.off;
    if constexpr (LO == 1) {
      double s[kObsLanes];
      for (int k = 0; k < kObsLanes; k++) s[k] = rg_part(ux, uy, uz, P.off, P.n, b, k, boxl, qbox);
      for (int w = kObsLanes / 2; w > 0; w >>= 1)
        for (int k = 0; k < w; k++) s[k] += s[k + w];
      rg = s[0];
    } else {
      static_assert(LO == kObsLanes, "one lane per part");
      double s = rg_part(ux, uy, uz, P.off, P.n, b, l, boxl, qbox);
#pragma unroll
      for (int w = kObsLanes / 2; w > 0; w >>= 1) {
        const double o = __shfl_down(s, w, kObsLanes);
        if (l < w) s += o;
      }
      rg = s;
    }
  }
  if (l != 0) return;
  double* r = D.red + (size_t)g * kRed;
  int cont = 0, nla = 0, pla = 0, nlr = 0, plr = 0;
  const int a0 = D.att_beg[g], a1 = D.att_end[g];
  for (int q = a0; q < a1; q++) {
    const int code = D.att_code[q];
    if (code & 1) continue;  // each native pair once, from its i bead
    const int k = code >> 1, j = D.att_partner[q];
    const uint8_t f = D.att_flag[q];
    nla += f & 1;
    pla += (f >> (1 + par)) & 1;
    if (!(what & OBS_PHYS)) continue;
    // Q: d / (float) r_ij < 1.25 with d from unc_pos[i] - unc_pos[j] (sop.cpp:223-238)
    const double dx = mimg(ux[g] - ux[j], boxl, qbox);
    const double dy = mimg(uy[g] - uy[j], boxl, qbox);
    const double dz = mimg(uz[g] - uz[j], boxl, qbox);
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d / (double)D.am_rf[k] < 1.25) cont++;
  }
  const int cnt = D.cnt[g];
  const ListRow S{D.slot, g, D.cap, true};
  const uint32_t pbit = 1u << (30 + par);
  for (int q = 0; q < cnt; q++) {
    const uint32_t w = S[q];
    if ((w >> 25) & 1) continue;
    nlr++;
    if (w & pbit) plr++;
  }
  if (what & OBS_COUNTS) { r[R_NLA] = nla; r[R_PLA] = pla; r[R_NLR] = nlr; r[R_PLR] = plr; }
  if (!(what & OBS_PHYS)) return;
  // kinT source: velocities (underdamped, sop.cpp:263-269) or position
  // increments (overdamped, sop.cpp:271-278; scaled by zeta/(2h) on the host)
  const bool od = P.sim_type == 2;
  // the increment of the step that produced positions cur (the reference's
  // calculate_observables(incr) right after the iteration)
  const double* ix = cur ? D.ix[1] : D.ix[0];
  const double* iy = cur ? D.iy[1] : D.iy[0];
  const double* iz = cur ? D.iz[1] : D.iz[0];
  const double qx = od ? ix[g] : D.vx[g], qy = od ? iy[g] : D.vy[g], qz = od ? iz[g] : D.vz[g];
  const double v2 = qx * qx + qy * qy + qz * qz;
  r[R_CONT] = cont; r[R_RG] = rg; r[R_V2] = v2;
}

// kObsLanes lanes per bead (the Rg^2 pairs split between them).
__global__ void k_observe(Dev D, int cur, int par, int rel, int all, int what, int g0, int g1) {
  rel += *D.relb;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = g0 + t / kObsLanes, l = t & (kObsLanes - 1);
  if (g >= g1 || D.chain[g] < 0) return;  // past the end, or a free bead: the whole group
  if (!all && D.cp[D.chain[g]].resident) return;
  observe_bead<kObsLanes>(D, cur, par, rel, all, what, g, l);
}

// Sums of chain c's per-bead fields into out[kRed], in a fixed order: 256
// strided partial sums, then a tree -- the same order for any block size (a
// thread takes entries t, t + blockDim.x, ... of each level).  sh: 256
// doubles of LDS.
__device__ __forceinline__ void chain_reduce_block(const Dev& D, int c, double* out, double* sh) {
  const ChainP& P = D.cp[c];
  const int t = threadIdx.x, nt = blockDim.x;
  for (int f = 0; f < kRed; f++) {
    for (int u = t; u < 256; u += nt) {
      double s = 0.0;
      for (int b = u; b < P.n; b += 256) s += D.red[(size_t)(P.off + b) * kRed + f];
      sh[u] = s;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      for (int u = t; u < w; u += nt) sh[u] += sh[u + w];
      __syncthreads();
    }
    if (t == 0) out[f] = sh[0];
    __syncthreads();
  }
}

// Per-chain sums of the per-bead fields: one block of 256 per chain.
// Chains the resident kernel advances write their own sums during sopg_run.
__global__ void __launch_bounds__(256) k_chain_reduce(Dev D, int c0, int skip_resident, double* out) {
  const int c = c0 + blockIdx.x;
  if (skip_resident && D.cp[c].resident) return;
  __shared__ double sh[256];
  chain_reduce_block(D, c, out + (size_t)blockIdx.x * kRed, sh);
}

// ------------------------------------------------------------ resident kernel
//
// The persistent form of sopg_run for small chains: one block per chain (the
// chains[] list) advances that chain through every step of the run -- the
// run's first list rebuild, pair-list filter and position half, then per
// step the fused force / velocity / next-position pass, the record when due
// and the list rebuild when due -- with block barriers between the phases
// and no other launch.  It calls the per-bead code of the per-step kernels
// (force_bead, rebuild_scan_bead, pairfilter_bead, posupdate_bead,
// observe_bead, chain_reduce_block), so its results are the same bits; the
// rebuild is the full candidate scan, whose lists equal the binned ones.
// rec_slot[s] is the record buffer row of run step s (-1: no record).
template <int L>
__global__ void __launch_bounds__(512) k_resident(Dev D, const Tables* __restrict__ T, const int* __restrict__ chains,
                                                   int cur, int par, int nsteps, const int* __restrict__ rec_slot,
                                                   double* rec, int nc) {
  __shared__ __align__(16) double tp[kTypeRows][TP_N];
  __shared__ double sh[256];
  load_type_tables(T, tp);
  const int c = chains[blockIdx.x];
  const ChainP& P = D.cp[c];
  const int n = P.n, off = P.off;
  const int groups = (int)blockDim.x / L, grp = (int)threadIdx.x / L, l = (int)threadIdx.x & (L - 1);
  const int flags = F_STORE | F_VEL | F_ENERGY | F_NOISE | F_BOND | F_ANGLE | F_VDW | F_ADVANCE;
  auto rebuild = [&](int step) {
    if (P.cutoff == SOPG_CUTOFF_NONE || (P.phase + step) % P.nnlup != 0) return;
    if (P.cutoff == SOPG_CUTOFF_CELL) {
      for (int b = threadIdx.x; b < n; b += blockDim.x) cellcoords_bead(D, cur, off + b);
      __syncthreads();
    }
    for (int b = grp; b < n; b += groups) rebuild_scan_bead<L, true>(D, T, cur, par, step, -1, off + b, l);
    __syncthreads();
  };
  __syncthreads();
  // the run's first step: rebuild if due, pair list, position half
  rebuild(0);
  for (int b = grp; b < n; b += groups) pairfilter_bead<L>(D, T, cur, par, off + b, l);
  __syncthreads();
  for (int b = threadIdx.x; b < n; b += blockDim.x) posupdate_bead(D, cur, cur, off + b);
  __syncthreads();
  for (int s = 0; s < nsteps; s++) {
    const bool last = s == nsteps - 1;
    for (int b = grp; b < n; b += groups) {
      if (last) force_bead<L, false>(D, T, tp, cur, par, s, flags, nullptr, off + b, l);
      else force_bead<L, true>(D, T, tp, cur, par, s, flags, nullptr, off + b, l);
    }
    __syncthreads();
    if (rec_slot[s] >= 0 && (P.iup0 + s) % P.nup == 0) {
      for (int b = threadIdx.x; b < n; b += blockDim.x) observe_bead<1>(D, cur, par, s, 0, OBS_PHYS, off + b, 0);
      __syncthreads();
      chain_reduce_block(D, c, rec + ((size_t)rec_slot[s] * nc + c) * kRed, sh);
    }
    if (!last) {
      par ^= 1;
      rebuild(s + 1);
      cur ^= 1;
    }
  }
}

// ------------------------------------------------------------ fixed scratch
//
// Record staging of the allocation-free run (sopg_run_noalloc): a file-scope
// device array instead of context memory.  In host code `scratch` names a
// host-side stand-in, not a device address, so it is never passed to a kernel
// or a copy: scratch_address() resolves the device address of this GPU's copy
// once per device (with that device current) and every use goes through it.
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];
static bool g_scratch_done[kMaxDevices];
static std::mutex g_scratch_mu;

// The calling thread's current device must be `device`.
int scratch_address(int device, double** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_done[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) return -1;
    g_scratch_ptr[device] = (double*)p;
    g_scratch_done[device] = true;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

// Step base of a replayed graph segment (see Dev::relb).
__global__ void k_set_base(int* base, int value) { *base = value; }

// ------------------------------------------------------------------ launchers

static inline unsigned blocks_for(long threads) { return (unsigned)((threads + 255) / 256); }

template <int L>
static void launch_rebuild_L(const Dev& D, const Tables* T, int cur, int par, int rel, int mode, int g0, int g1,
                             hipStream_t s) {
  const unsigned blocks = blocks_for((long)(g1 - g0) * L);
  hipLaunchKernelGGL(k_rebuild<L>, dim3(blocks < 4096u ? blocks : 4096u), dim3(256), 0, s, D, T, cur, par, rel,
                     mode, g0, g1);
}
template <int L>
static void launch_pairfilter_L(const Dev& D, const Tables* T, int cur, int par, int g0, int g1, hipStream_t s) {
  hipLaunchKernelGGL(k_pairfilter<L>, dim3(blocks_for((long)(g1 - g0) * L)), dim3(256), 0, s, D, T, cur, par, g0,
                     g1);
}
template <int L>
static void launch_force_L(const Dev& D, const Tables* T, int cur, int pr, int rel, int flags, bool next,
                           const double* noise, int g0, int g1, const int* blk, int nblk, hipStream_t s) {
  const dim3 grid(blk ? (unsigned)nblk : blocks_for((long)(g1 - g0) * L)), block(256);
  const bool en = (flags & (F_ENERGY | F_ENERGY_ALL)) != 0;
  if (next && en)
    hipLaunchKernelGGL((k_force<L, true, true>), grid, block, 0, s, D, T, cur, pr, rel, flags, noise, g0, g1, blk);
  else if (next)
    hipLaunchKernelGGL((k_force<L, true, false>), grid, block, 0, s, D, T, cur, pr, rel, flags, noise, g0, g1, blk);
  else if (en)
    hipLaunchKernelGGL((k_force<L, false, true>), grid, block, 0, s, D, T, cur, pr, rel, flags, noise, g0, g1, blk);
  else
    hipLaunchKernelGGL((k_force<L, false, false>), grid, block, 0, s, D, T, cur, pr, rel, flags, noise, g0, g1, blk);
}

#define SOPG_DISPATCH_L(L, CALL)          \
  switch (L) {                            \
    case 1: CALL(1); break;               \
    case 2: CALL(2); break;               \
    case 4: CALL(4); break;               \
    case 8: CALL(8); break;               \
    case 16: CALL(16); break;             \
    case 32: CALL(32); break;             \
    default: CALL(64); break;             \
  }

void launch_cellcoords(const Dev& D, int cur, int g0, int g1, hipStream_t s) {
  hipLaunchKernelGGL(k_cellcoords, dim3(blocks_for(g1 - g0)), dim3(256), 0, s, D, cur, g0, g1);
}
// List rebuild of beads [g0, g1) = chains [c0, c0 + nc): the binned kernels for
// chains whose grid applies (binned != 0), the full scan for the others.
constexpr int kGridLanes = 16;                       // lanes per bead in the binned rebuild
constexpr int kGridBeads = 256 / kGridLanes;         // beads per block
constexpr int kFilterLanes = 8;                     // lanes per bead in the candidate filter
constexpr int kMaxBitmapWords = 512;                 // 16 beads x 2 bitmaps x 2 KiB = 64 KiB of LDS
void launch_rebuild(int L, const Dev& D, const Tables* T, int cur, int par, int rel, int mode, int g0, int g1,
                    int c0, int nc, int binned, int cand, hipStream_t s, const int* tiles, int ntiles) {
  const int words = D.bitmap_words < kMaxBitmapWords ? D.bitmap_words : kMaxBitmapWords;
  hipLaunchKernelGGL(k_grid, dim3(nc), dim3(kGridThreads), 0, s, D, T, cur, rel, mode, c0, binned ? 32 * words : 0);
  if (binned) {
    Dev Dg = D;
    Dg.bitmap_words = words;
    const unsigned blocks = tiles ? (unsigned)ntiles : (unsigned)((g1 - g0 + kGridBeads - 1) / kGridBeads);
    hipLaunchKernelGGL(k_rebuild_grid<kGridLanes>, dim3(blocks), dim3(256),
                       (size_t)kGridBeads * 2 * words * sizeof(uint32_t), s, Dg, T, cur, par, rel, mode, g0, g1,
                       tiles);
  }
  // the full scan (lists, or the candidate lists of chains too wide for the
  // grid), then the filter of every chain planned to filter its candidates
#define C_(X) launch_rebuild_L<X>(D, T, cur, par, rel, mode, g0, g1, s)
  SOPG_DISPATCH_L(L, C_)
#undef C_
  if (binned && cand) {
    const long threads = tiles ? (long)ntiles * kBeadTile * kFilterLanes : (long)(g1 - g0) * kFilterLanes;
    hipLaunchKernelGGL(k_cand_filter<kFilterLanes>, dim3(blocks_for(threads)), dim3(256), 0, s, D, T, cur, par, rel,
                       g0, g1, tiles, ntiles);
  }
}
void launch_pairfilter(int L, const Dev& D, const Tables* T, int cur, int par, int g0, int g1, hipStream_t s) {
#define C_(X) launch_pairfilter_L<X>(D, T, cur, par, g0, g1, s)
  SOPG_DISPATCH_L(L, C_)
#undef C_
}
void launch_force(int L, const Dev& D, const Tables* T, int cur, int pr, int rel, int flags, bool next,
                  const double* noise, int g0, int g1, hipStream_t s, const int* blk, int nblk) {
#define C_(X) launch_force_L<X>(D, T, cur, pr, rel, flags, next, noise, g0, g1, blk, nblk, s)
  SOPG_DISPATCH_L(L, C_)
#undef C_
}
void launch_posupdate(const Dev& D, int src, int dst, int g0, int g1, hipStream_t s) {
  hipLaunchKernelGGL(k_posupdate, dim3(blocks_for(g1 - g0)), dim3(256), 0, s, D, src, dst, g0, g1);
}
void launch_observe(const Dev& D, int cur, int par, int rel, int all, int what, int g0, int g1, hipStream_t s) {
  hipLaunchKernelGGL(k_observe, dim3(blocks_for((long)(g1 - g0) * kObsLanes)), dim3(256), 0, s, D, cur, par, rel, all,
                     what, g0, g1);
}
void launch_resident(int L, const Dev& D, const Tables* T, const int* chains, int nchains, int threads, int cur,
                     int par, int nsteps, const int* rec_slot, double* rec, int nc, hipStream_t s) {
#define C_(X) hipLaunchKernelGGL(k_resident<X>, dim3(nchains), dim3(threads), 0, s, D, T, chains, cur, par, nsteps, \
                                 rec_slot, rec, nc)
  SOPG_DISPATCH_L(L, C_)
#undef C_
}
void launch_set_base(int* base, int value, hipStream_t s) {
  hipLaunchKernelGGL(k_set_base, dim3(1), dim3(1), 0, s, base, value);
}
void launch_chain_reduce(const Dev& D, int c0, int nc, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_chain_reduce, dim3(nc), dim3(256), 0, s, D, c0, 1, out);
}

}  // namespace sopg
