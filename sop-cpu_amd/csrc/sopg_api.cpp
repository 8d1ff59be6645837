// sopg_api.cpp -- C ABI of the gfx950 SOP hot path (declared in include/sopg.h).
//
// Host side of a context: validates and copies topologies, builds the per-bead
// incidence lists (in the reference's list order), owns device memory, and
// sequences the kernels of sopg_kernels.hip on one HIP stream.
//
// Device memory is a set of pools (beads; bond, angle, native, non-native
// master entries and their per-bead incidences; chain slots).  Each pool is a
// first-fit range allocator whose device arrays grow on demand, so chains can
// join and leave a running context (continuous batching) without touching the
// other chains' memory; a free bead carries chain id -1 and is skipped.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sopg.h"
#include "sopg_device.h"
#include "sopg_internal.h"

namespace sopg {
void launch_cellcoords(const Dev& D, int cur, int g0, int g1, hipStream_t s);
// tiles: ntiles bead tiles (kBeadTile beads each) that cover the beads of
// [g0, g1) to rebuild, or null for every bead of [g0, g1)
void launch_rebuild(int L, const Dev& D, const Tables* T, int cur, int par, int rel, int mode, int g0, int g1,
                    int c0, int nc, int binned, int cand, hipStream_t s, const int* tiles = nullptr, int ntiles = 0);
void launch_pairfilter(int L, const Dev& D, const Tables* T, int cur, int par, int g0, int g1, hipStream_t s);
// blk: nblk block numbers of 256 / L beads each (blocks of [0, g1)), or null
// for every block of [g0, g1)
void launch_force(int L, const Dev& D, const Tables* T, int cur, int pr, int rel, int flags, bool next,
                  const double* noise, int g0, int g1, hipStream_t s, const int* blk = nullptr, int nblk = 0);
void launch_posupdate(const Dev& D, int src, int dst, int g0, int g1, hipStream_t s);
void launch_observe(const Dev& D, int cur, int par, int rel, int all, int what, int g0, int g1, hipStream_t s);
void launch_chain_reduce(const Dev& D, int c0, int nc, double* out, hipStream_t s);
void launch_set_base(int* base, int value, hipStream_t s);
int scratch_address(int device, double** out);
void launch_resident(int L, const Dev& D, const Tables* T, const int* chains, int nchains, int threads, int cur,
                     int par, int nsteps, const int* rec_slot, double* rec, int nc, hipStream_t s);
enum { F_STORE = 1, F_VEL = 2, F_ENERGY = 4, F_ENERGY_ALL = 8, F_NOISE = 16, F_BOND = 32, F_ANGLE = 64,
       F_VDW = 128, F_ACCUM = 256, F_ADVANCE = 512, F_ALLTERMS = 16 | 32 | 64 | 128 };
enum { OBS_PHYS = 1, OBS_COUNTS = 2 };
}  // namespace sopg

using namespace sopg;

namespace {

// First-fit allocator of index ranges in [0, cap).
struct Pool {
  size_t cap = 0, hw = 0;
  std::vector<std::pair<size_t, size_t>> fr;  // free [begin, end), sorted, merged
  bool take(size_t n, size_t& at) {
    if (n == 0) { at = 0; return true; }
    for (size_t k = 0; k < fr.size(); k++)
      if (fr[k].second - fr[k].first >= n) {
        at = fr[k].first;
        fr[k].first += n;
        if (fr[k].first == fr[k].second) fr.erase(fr.begin() + (long)k);
        hw = std::max(hw, at + n);
        return true;
      }
    return false;
  }
  void give(size_t at, size_t n) {
    if (n == 0) return;
    fr.push_back({at, at + n});
    merge();
    // a free range that reaches the end lowers the high-water mark
    if (!fr.empty() && fr.back().second == cap) hw = std::min(hw, fr.back().first);
  }
  void extend(size_t ncap) {
    if (ncap > cap) fr.push_back({cap, ncap});
    cap = ncap;
    merge();
  }
  void merge() {
    std::sort(fr.begin(), fr.end());
    std::vector<std::pair<size_t, size_t>> m;
    for (auto& r : fr) {
      if (!m.empty() && m.back().second >= r.first) m.back().second = std::max(m.back().second, r.second);
      else m.push_back(r);
    }
    fr.swap(m);
  }
};

// A device array living in a pool: per pool unit it holds `per` elements.
struct Arr {
  void** p;
  size_t esz, per;
};

enum { P_BEAD = 0, P_BINC, P_AINC, P_ATT, P_TINC, P_REP, P_RINC, P_CHAIN, NPOOL };

struct HostChain {
  bool active = false;
  sopg_chain_params p{};
  int n = 0;
  std::vector<int> bi, bj;  // 0-based, chain-local
  std::vector<double> br0;
  std::vector<int> gi, gk;
  std::vector<int> ai, aj, ait, ajt;
  std::vector<double> ar;
  bool rep_explicit = false;
  std::vector<int> ri, rj, rit, rjt;
  std::vector<int8_t> type;
  // derived once from the topology (fill_chainp runs at every sopg_run)
  double max_r2 = 0.0;  // largest lj_nat_pdb_dist^2
  int type_mask = 0;    // bit t: some bead has type t
  std::vector<int> q0, q1;  // device incidence index of each native entry's first / second bead
  double istep = 1.0;
  int inlup = 1, iup = 1;
  uint64_t noise_base = 0;
  bool on_device = false;
  // the pair list of parity `par` was built at the current positions (an
  // explicit list update since the last position change): sopg_run may skip
  // its first pair-list filter for this chain
  bool pl_fresh = false;
  bool resident = false;  // advanced by the resident kernel in the current sopg_run
  size_t at[NPOOL] = {0};  // range starts in each pool
};

}  // namespace

struct sopg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  sopg_tables tab{};
  Tables T{};
  int lanes = 8;        // lanes per bead, step / force kernels
  int list_lanes = 64;  // lanes per bead, full-scan list rebuild
  int binned = 1;       // binned list rebuild where a chain's extent allows it
  int cap_user = 0;
  double skin = 4.0;    // candidate lists of neighbor-list rebuilds (Angstrom; <= 0: off)
  size_t reserve_beads = 0, reserve_chains = 0;
  Tables* d_T = nullptr;
  std::vector<HostChain> chains;
  bool device_ready = false;
  Dev D{};
  std::vector<ChainP> cp;
  // Pinned staging of the chain records and of the error flag.  A full upload
  // of the chain records goes through cp_pin without waiting for the copy
  // (ev_cp marks it), and is skipped when the records equal what the device
  // holds (cp_dev, uploaded at generation cp_dev_gen) and gen has not moved.
  ChainP* cp_pin = nullptr;
  size_t cp_pin_cap = 0;
  bool cp_busy = false;
  hipEvent_t ev_cp = nullptr;
  std::vector<ChainP> cp_dev;
  long cp_dev_gen = -1;
  int* err_pin = nullptr;
  Pool pool[NPOOL];
  std::vector<Arr> arrs[NPOOL];
  std::vector<void*> allocs;  // fixed allocations (not in pools)
  int cur = 0, par = 0;
  double* d_out = nullptr;      // reduction output (chain slots * kRed)
  double* d_noise = nullptr;    // replay noise
  size_t noise_cap = 0;         // doubles
  double* d_rec = nullptr;      // per-record-step reductions
  double* d_scratch = nullptr;  // this GPU's file-scope record staging (sopg_run_noalloc)
  size_t rec_cap_d = 0;
  // graph replay of sopg_run's step segments
  int* d_relb = nullptr;        // [0] = 0 (direct launches), [1] = step base of the replayed segment
  bool graphs = true;
  struct GraphEntry {
    std::vector<char> key;
    hipGraphExec_t exec;
    long used;
  };
  std::vector<GraphEntry> gcache;
  long gclock = 0;
  long graph_launches = 0, graph_captures = 0;
  // Every call that can change positions, lists, parameters or the chain set
  // advances gen (the lane counts and the rebuild method do not: they change
  // how lists are built or walked, not what they hold).  A run leaves the next
  // step's pair list (at the positions it stops at) in the other parity; the
  // next run starts from it without a pair-list filter when gen has not moved
  // since (next_gen == gen).
  long gen = 0, next_gen = -1;
  // The last run's last step also wrote the position half that follows it
  // into buffer cur^1 (no chain was resident); with next_gen == gen the next
  // run starts by flipping cur instead of the in-place position update.
  bool adv_ready = false;
  // resident (persistent) kernel for small chains
  int resident_max = 0;         // chains of at most this many beads (0: off)
  int resident_lanes = 0;       // lanes per bead in it (0: automatic)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_res = nullptr;
  int* d_rlist = nullptr;
  size_t rlist_cap = 0;
  int* d_rslot = nullptr;
  size_t rslot_cap = 0;
  // block tables (block_table) of the step kernel and of the binned rebuild:
  // device array, its capacity, and what it holds
  struct BlockTable {
    int* d = nullptr;
    size_t cap = 0;
    std::vector<int> dev;
    // the last call's block size, block count, chain block ranges and result
    bool valid = false;
    int per = 0, total = 0;
    std::vector<std::pair<int, int>> ranges;
    const int* out = nullptr;
    int nout = 0;
  };
  BlockTable step_tab, tile_tab;
  // kernel timing (sopg_profile)
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<int, int>>> ev_pending;  // kernel, (start, stop) pool indices
  size_t ev_next = 0;
  double prof_ms[4] = {0, 0, 0, 0};
  long prof_n[4] = {0, 0, 0, 0};
};

namespace {

int fail(sopg_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

#define HIP_OK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, SOPG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------ timing

int prof_begin(sopg_ctx* c) {
  if (!c->prof) return -1;
  while (c->ev_pool.size() < c->ev_next + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    c->ev_pool.push_back(e);
  }
  const int i = (int)c->ev_next;
  c->ev_next += 2;
  (void)hipEventRecord(c->ev_pool[i], c->stream);
  return i;
}
void prof_end(sopg_ctx* c, int kernel, int i) {
  if (i < 0) return;
  (void)hipEventRecord(c->ev_pool[i + 1], c->stream);
  c->ev_pending.push_back({kernel, {i, i + 1}});
}
void prof_collect(sopg_ctx* c) {
  for (auto& p : c->ev_pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->ev_pool[p.second.first], c->ev_pool[p.second.second]) == hipSuccess) {
      c->prof_ms[p.first] += ms;
      c->prof_n[p.first]++;
    }
  }
  c->ev_pending.clear();
  c->ev_next = 0;
}

// ------------------------------------------------------------ memory

template <class T>
int dalloc(sopg_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return fail(c, SOPG_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->allocs.push_back(q);
  *p = (T*)q;
  return SOPG_OK;
}

void dfree(sopg_ctx* c, void* p) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(p);
}

// Host-to-device copy that has finished when it returns (the sources are
// often temporaries).
template <class T>
int upload_at(sopg_ctx* c, T* d, size_t at, const T* h, size_t n) {
  if (n == 0) return SOPG_OK;
  HIP_OK(c, hipMemcpyAsync(d + at, h, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return SOPG_OK;
}
template <class T>
int upload_at(sopg_ctx* c, T* d, size_t at, const std::vector<T>& h) {
  return upload_at(c, d, at, h.data(), h.size());
}

// Grow a pool to at least `need` units: new arrays, old contents copied.
int grow_pool(sopg_ctx* c, int pi, size_t need) {
  Pool& P = c->pool[pi];
  if (need <= P.cap) return SOPG_OK;
  size_t ncap = std::max(need, P.cap * 2 + 16);
  if (pi == P_BEAD) {
    if (need > kMaxPoolBeads) return fail(c, SOPG_ERR_ARG, "bead pool limit (2^28 beads per context) exceeded");
    // whole slot tiles, so that a grown array keeps the old tiles in place
    ncap = std::min((ncap + kSlotTile - 1) / kSlotTile * kSlotTile, kMaxPoolBeads);
  }
  HIP_OK(c, hipStreamSynchronize(c->stream));
  for (Arr& a : c->arrs[pi]) {
    void* q = nullptr;
    const size_t nb = ncap * a.per * a.esz, ob = P.cap * a.per * a.esz;
    hipError_t e = hipMalloc(&q, std::max<size_t>(nb, 1));
    if (e != hipSuccess) return fail(c, SOPG_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    HIP_OK(c, hipMemsetAsync(q, 0, std::max<size_t>(nb, 1), c->stream));
    if (*a.p && ob) HIP_OK(c, hipMemcpyAsync(q, *a.p, ob, hipMemcpyDeviceToDevice, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    if (*a.p) (void)hipFree(*a.p);
    *a.p = q;
  }
  P.extend(ncap);
  return SOPG_OK;
}

// Slot words of beads [off, off + n) as n rows of cap words.  The device keeps
// them in tiles of kSlotTile beads (slot_index), so a transfer covers whole
// tiles: [off, off + n) rounded out to tile bounds, inside the pool.
int slot_rows_get(sopg_ctx* c, size_t off, size_t n, std::vector<uint32_t>& rows) {
  const int cap = c->D.cap;
  const size_t t0 = off / kSlotTile * kSlotTile, t1 = (off + n + kSlotTile - 1) / kSlotTile * kSlotTile;
  std::vector<uint32_t> tiles((t1 - t0) * (size_t)cap);
  HIP_OK(c, hipMemcpyAsync(tiles.data(), c->D.slot + t0 * (size_t)cap, sizeof(uint32_t) * tiles.size(),
                           hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  rows.assign(n * (size_t)cap, 0u);
  const size_t base = t0 * (size_t)cap;
  for (size_t b = 0; b < n; b++)
    for (int q = 0; q < cap; q++) rows[b * cap + q] = tiles[slot_index((int)(off + b), q, cap) - base];
  return SOPG_OK;
}

// The inverse: write n rows of cap words as beads [off, off + n)'s slot words,
// leaving the other beads of the boundary tiles as they are.
int slot_rows_put(sopg_ctx* c, size_t off, size_t n, const std::vector<uint32_t>& rows) {
  const int cap = c->D.cap;
  const size_t t0 = off / kSlotTile * kSlotTile, t1 = (off + n + kSlotTile - 1) / kSlotTile * kSlotTile;
  std::vector<uint32_t> tiles((t1 - t0) * (size_t)cap);
  HIP_OK(c, hipMemcpyAsync(tiles.data(), c->D.slot + t0 * (size_t)cap, sizeof(uint32_t) * tiles.size(),
                           hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  const size_t base = t0 * (size_t)cap;
  for (size_t b = 0; b < n; b++)
    for (int q = 0; q < cap; q++) tiles[slot_index((int)(off + b), q, cap) - base] = rows[b * cap + q];
  return upload_at(c, c->D.slot, base, tiles);
}

int take_range(sopg_ctx* c, int pi, size_t n, size_t& at) {
  Pool& P = c->pool[pi];
  if (P.take(n, at)) return SOPG_OK;
  // grow past the current end (the last free range may be extended in place)
  size_t tail = 0;
  if (!P.fr.empty() && P.fr.back().second == P.cap) tail = P.fr.back().second - P.fr.back().first;
  int rc = grow_pool(c, pi, P.cap + n - tail);
  if (rc) return rc;
  if (!P.take(n, at)) return fail(c, SOPG_ERR_ARG, "pool allocation failed");
  return SOPG_OK;
}

void register_arrays(sopg_ctx* c) {
  Dev& D = c->D;
  auto add = [&](int pi, void* pp, size_t esz, size_t per) { c->arrs[pi].push_back({(void**)pp, esz, per}); };
  const size_t cap = (size_t)D.cap;
  for (int s = 0; s < 2; s++) {
    add(P_BEAD, &D.ux[s], 8, 1); add(P_BEAD, &D.uy[s], 8, 1); add(P_BEAD, &D.uz[s], 8, 1);
    add(P_BEAD, &D.px[s], 8, 1); add(P_BEAD, &D.py[s], 8, 1); add(P_BEAD, &D.pz[s], 8, 1);
    add(P_BEAD, &D.ix[s], 8, 1); add(P_BEAD, &D.iy[s], 8, 1); add(P_BEAD, &D.iz[s], 8, 1);
  }
  for (double** p : {&D.vx, &D.vy, &D.vz, &D.fx, &D.fy, &D.fz, &D.cx, &D.cy, &D.cz, &D.u0x,
                     &D.u0y, &D.u0z})
    add(P_BEAD, p, 8, 1);
  for (int** p : {&D.chain, &D.bond_beg, &D.bond_end, &D.ang_beg, &D.ang_end, &D.att_beg, &D.att_end, &D.rep_beg,
                  &D.rep_end, &D.cnt, &D.ccnt, &D.bead_cell, &D.cell_beads})
    add(P_BEAD, p, 4, 1);
  add(P_BEAD, &D.type, 1, 1);
  add(P_BEAD, &D.slot, 4, cap);
  add(P_BEAD, &D.cand, 4, cap);
  add(P_BEAD, &D.fpos, sizeof(float4), 1);
  add(P_BEAD, &D.slot_k, 4, cap);
  add(P_BEAD, &D.red, 8, kRed);
  add(P_BINC, &D.bond_pq, 4, 1); add(P_BINC, &D.bond_r0, 8, 1);
  add(P_AINC, &D.ang_pq, 4, 1);
  add(P_ATT, &D.am_i, 4, 1); add(P_ATT, &D.am_j, 4, 1); add(P_ATT, &D.am_q0, 4, 1); add(P_ATT, &D.am_t, 1, 1);
  for (double** p : {&D.am_r2, &D.am_nlrc2}) add(P_ATT, p, 8, 1);
  add(P_ATT, &D.am_rf, 4, 1);
  add(P_TINC, &D.att_code, 4, 1); add(P_TINC, &D.att_partner, 4, 1); add(P_TINC, &D.att_flag, 1, 1);
  add(P_TINC, &D.att_w, 4, 1); add(P_TINC, &D.att_r, 8, 1);
  add(P_REP, &D.rm_i, 4, 1); add(P_REP, &D.rm_j, 4, 1); add(P_REP, &D.rm_t, 1, 1);
  add(P_RINC, &D.rep_code, 4, 1);
  add(P_CHAIN, &D.cp, sizeof(ChainP), 1);
  add(P_CHAIN, &D.grid, sizeof(GridP), 1);
  add(P_CHAIN, &D.cell_start, 4, kGridCells);
  add(P_CHAIN, &D.cell_fill, 4, kGridCells);
  add(P_CHAIN, &c->d_out, 8, kRed);
}

// ------------------------------------------------------------ parameters

// skin <= 0: no candidate lists
void derive_tables(const sopg_tables& t, double skin, Tables& T) {
  double thr = 0.0, reach = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      T.fatt[i][j] = t.force_coeff_att[i][j];
      T.frep[i][j] = t.force_coeff_rep[i][j];
      T.catt[i][j] = t.coeff_att[i][j];
      T.crep[i][j] = t.coeff_rep[i][j];
      const double s = t.sigma_rep[i][j];
      // params.cpp:306-312
      T.s2[i][j] = s * s;
      T.s6[i][j] = T.s2[i][j] * T.s2[i][j] * T.s2[i][j];
      T.s12[i][j] = T.s6[i][j] * T.s6[i][j];
      // neighbor_list.cpp:97-100, sop.cpp:541-542
      const double rn = 3.2 * s, rp = 2.5 * s;
      T.nlrc2[i][j] = rn * rn;
      T.plrc2[i][j] = rp * rp;
      const double rc = rn + skin;
      T.candc2[i][j] = rc * rc;
      reach = std::max(reach, rn + 2.0 * skin + 1.0);
      thr = std::max(thr, T.s2[i][j] * kTol);
    }
  // A pair of the neighbor list at positions P (d < 3.2 sigma) was within
  // 3.2 sigma + skin of u0 if neither bead has moved more than skin/2 since;
  // the relative margin is far above the rounding of any distance here.
  T.cand_on = skin > 0.0 ? 1 : 0;
  const double half = 0.5 * skin * (1.0 - 1e-6);
  T.cand_disp2 = skin > 0.0 ? half * half : -1.0;
  T.cand_reach = reach;
  T.k_bnd = t.k_bnd;
  T.R0sq = t.R0sq;
  T.e_bnd_coeff = t.e_bnd_coeff;
  T.e_ang_ss_coeff = t.e_ang_ss_coeff;
  T.f_ang_ss_coeff = t.f_ang_ss_coeff;
  T.thr_rep = thr;
}

int check_params(sopg_ctx* c, const sopg_chain_params* p) {
  if (!p) return fail(c, SOPG_ERR_ARG, "null chain params");
  if (!(p->boxl > 0.0) || !std::isfinite(p->boxl)) return fail(c, SOPG_ERR_ARG, "boxl must be finite and > 0");
  if (p->cutoff < SOPG_CUTOFF_NONE || p->cutoff > SOPG_CUTOFF_CELL) return fail(c, SOPG_ERR_ARG, "bad cutoff");
  if (p->cutoff == SOPG_CUTOFF_CELL && (!(p->ncell > 0.0) || !(p->lcell > 0.0) || !std::isfinite(p->ncell)))
    return fail(c, SOPG_ERR_ARG, "cell list needs ncell > 0 and lcell > 0");
  if (p->nnlup < 1) return fail(c, SOPG_ERR_ARG, "nnlup must be >= 1 (the reference divides by it)");
  if (p->nup < 1) return fail(c, SOPG_ERR_ARG, "nup must be >= 1 (the reference divides by it)");
  if (p->noise_mode != SOPG_NOISE_REPLAY && p->noise_mode != SOPG_NOISE_PHILOX)
    return fail(c, SOPG_ERR_ARG, "bad noise mode");
  return SOPG_OK;
}

void fill_chainp(const HostChain& h, const Tables& T, ChainP& P) {
  if (!h.active || !h.on_device) {
    std::memset(&P, 0, sizeof P);
    P.cutoff = SOPG_CUTOFF_NONE;
    P.nnlup = P.nup = 1;
    P.phase = P.iup0 = 1;
    return;
  }
  const sopg_chain_params& p = h.p;
  P.off = (int)h.at[P_BEAD];
  P.n = h.n;
  P.att_off = (int)h.at[P_ATT];
  P.natt = (int)h.ai.size();
  P.rep_explicit = h.rep_explicit ? 1 : 0;
  P.thr_att = kTol * h.max_r2;
  P.cutoff = p.cutoff;
  P.nnlup = p.nnlup;
  P.nup = p.nup;
  P.phase = h.inlup;
  P.iup0 = h.iup;
  P.vel_on = (p.T < 1.0e-5) ? 0 : 1;
  P.noise_mode = p.noise_mode;
  P.boxl = p.boxl;
  P.qbox = 0.25 * p.boxl;
  P.lcell = p.lcell;
  P.ncell = p.ncell;
  P.qncell = 0.25 * p.ncell;
  P.a1 = p.a1; P.a2 = p.a2; P.a3 = p.a3; P.a4 = p.a4; P.a5 = p.a5;
  P.var = std::sqrt(2.0 * p.T * p.zeta / p.h);  // energy.cpp:456
  P.T = p.T;
  P.seed = p.seed;
  P.noise_base = h.noise_base;
  P.sim_type = p.sim_type == 2 ? 2 : 1;
  // grid edge of the binned rebuild: the largest neighbor-list cutoff among
  // the chain's bead types, with margin for the rounding of the cell index
  double rc2 = 0.0;
  bool present[3];
  for (int a = 0; a < 3; a++) present[a] = (h.type_mask >> a) & 1;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      if (present[a] && present[b]) rc2 = std::max(rc2, T.nlrc2[a][b]);
  P.gcell = std::sqrt(rc2) * 1.001 + 1e-9;
  double cc2 = 0.0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      if (present[a] && present[b]) cc2 = std::max(cc2, T.candc2[a][b]);
  P.gcand = std::sqrt(cc2) * 1.001 + 1e-9;
  P.resident = h.resident ? 1 : 0;
}

// Device copy of the per-chain parameters: chain `only`, or all (only < 0).
// may_alloc false: no allocation, not even of pinned host memory (a staging
// buffer too small for the records falls back to a waited copy).
int upload_chainp(sopg_ctx* c, int only = -1, bool may_alloc = true) {
  if (!c->device_ready) return SOPG_OK;
  const size_t cap = c->pool[P_CHAIN].cap;
  if (c->cp.size() != cap) {
    c->cp.resize(cap);
    only = -1;
  }
  const HostChain none;
  for (size_t i = 0; i < cap; i++) {
    if (only >= 0 && (int)i != only) continue;
    fill_chainp(i < c->chains.size() ? c->chains[i] : none, c->T, c->cp[i]);
  }
  if (only >= 0) {
    c->cp_dev_gen = -1;
    return upload_at(c, c->D.cp, (size_t)only, &c->cp[only], 1);
  }
  if (c->cp_dev_gen == c->gen && c->cp_dev.size() == cap &&
      !std::memcmp(c->cp_dev.data(), c->cp.data(), cap * sizeof(ChainP)))
    return SOPG_OK;
  if (cap > c->cp_pin_cap && may_alloc) {
    if (c->cp_busy) HIP_OK(c, hipEventSynchronize(c->ev_cp));
    c->cp_busy = false;
    if (c->cp_pin) HIP_OK(c, hipHostFree(c->cp_pin));
    c->cp_pin = nullptr;
    c->cp_pin_cap = 0;
    // without pinned memory the records go through the waited pageable copy below
    if (hipHostMalloc((void**)&c->cp_pin, 2 * cap * sizeof(ChainP), hipHostMallocDefault) == hipSuccess) {
      c->cp_pin_cap = 2 * cap;
    } else {
      c->cp_pin = nullptr;
      (void)hipGetLastError();  // not an error of a later launch
    }
  }
  c->cp_dev_gen = -1;
  if (cap > c->cp_pin_cap) return upload_at(c, c->D.cp, 0, c->cp);
  if (c->cp_busy) HIP_OK(c, hipEventSynchronize(c->ev_cp));
  std::memcpy(c->cp_pin, c->cp.data(), cap * sizeof(ChainP));
  HIP_OK(c, hipMemcpyAsync(c->D.cp, c->cp_pin, cap * sizeof(ChainP), hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, hipEventRecord(c->ev_cp, c->stream));
  c->cp_busy = true;
  c->cp_dev = c->cp;
  c->cp_dev_gen = c->gen;
  return SOPG_OK;
}

// Per-bead incidence lists of entries (a[e], b[e]), stable in entry order:
// entry e adds (base+e)<<1 to bead a[e], then (base+e)<<1|1 to bead b[e].
void build_csr(int n, const std::vector<int>& a, const std::vector<int>& b, size_t base, size_t inc_at, int boff,
               std::vector<int>& beg, std::vector<int>& end, std::vector<int>& code) {
  std::vector<int> cnt((size_t)n + 1, 0);
  for (size_t e = 0; e < a.size(); e++) { cnt[a[e] + 1]++; cnt[b[e] + 1]++; }
  for (int g = 0; g < n; g++) cnt[g + 1] += cnt[g];
  beg.resize(n);
  end.resize(n);
  for (int g = 0; g < n; g++) { beg[g] = (int)inc_at + cnt[g]; end[g] = (int)inc_at + cnt[g + 1]; }
  code.assign(cnt[n], 0);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (size_t e = 0; e < a.size(); e++) {
    code[fill[a[e]]++] = (int)((base + e) << 1);
    code[fill[b[e]]++] = (int)((base + e) << 1) | 1;
  }
  (void)boff;
}

int ensure_device(sopg_ctx* c);
int check_err_flag(sopg_ctx* c);

// Put one chain's topology and zeroed state into the pools.
int materialize(sopg_ctx* c, int id) {
  HostChain& h = c->chains[id];
  Dev& D = c->D;
  int rc;
  // a no-cutoff chain keeps every non-native partner of a bead (up to n - 4)
  if (h.p.cutoff == SOPG_CUTOFF_NONE && h.n > D.cap)
    return fail(c, SOPG_ERR_ARG, "a no-cutoff chain of " + std::to_string(h.n) +
                                     " beads exceeds the per-bead list capacity " + std::to_string(D.cap) +
                                     " (sopg_set_capacity before the first device operation)");
  const size_t nb = h.bi.size(), na = h.gi.size(), nt = h.ai.size(), nr = h.rep_explicit ? h.ri.size() : 0;
  if ((rc = take_range(c, P_BEAD, h.n, h.at[P_BEAD])) || (rc = take_range(c, P_BINC, 2 * nb, h.at[P_BINC])) ||
      (rc = take_range(c, P_AINC, 2 * na, h.at[P_AINC])) || (rc = take_range(c, P_ATT, nt, h.at[P_ATT])) ||
      (rc = take_range(c, P_TINC, 2 * nt, h.at[P_TINC])) || (rc = take_range(c, P_REP, nr, h.at[P_REP])) ||
      (rc = take_range(c, P_RINC, 2 * nr, h.at[P_RINC])))
    return rc;
  if ((size_t)id >= c->pool[P_CHAIN].cap && (rc = grow_pool(c, P_CHAIN, (size_t)id + 1))) return rc;
  const int off = (int)h.at[P_BEAD];
  const int n = h.n;
  D.bitmap_words = std::max(D.bitmap_words, (n + 31) / 32);
  // per-bead arrays
  std::vector<int> chain(n, id);
  std::vector<int8_t> type(h.type.begin(), h.type.end());
  std::vector<int> beg, end, code;
  auto glob = [&](const std::vector<int>& v) {
    std::vector<int> g(v.size());
    for (size_t k = 0; k < v.size(); k++) g[k] = off + v[k];
    return g;
  };
  if ((rc = upload_at(c, D.chain, off, chain)) || (rc = upload_at(c, D.type, off, type))) return rc;
  // bonds and angles: per incidence the partner (global) << 1 | side
  auto pq_of = [&](const std::vector<int>& code, const std::vector<int>& a, const std::vector<int>& b) {
    std::vector<int> pq(code.size());
    for (size_t q = 0; q < code.size(); q++) {
      const size_t k = (size_t)(code[q] >> 1), side = (size_t)(code[q] & 1);
      pq[q] = ((off + (side ? a[k] : b[k])) << 1) | (int)side;
    }
    return pq;
  };
  build_csr(n, h.bi, h.bj, 0, h.at[P_BINC], off, beg, end, code);
  std::vector<double> r0q(code.size());
  for (size_t q = 0; q < code.size(); q++) r0q[q] = h.br0[(size_t)(code[q] >> 1)];
  if ((rc = upload_at(c, D.bond_beg, off, beg)) || (rc = upload_at(c, D.bond_end, off, end)) ||
      (rc = upload_at(c, D.bond_pq, h.at[P_BINC], pq_of(code, h.bi, h.bj))) ||
      (rc = upload_at(c, D.bond_r0, h.at[P_BINC], r0q)))
    return rc;
  build_csr(n, h.gi, h.gk, 0, h.at[P_AINC], off, beg, end, code);
  if ((rc = upload_at(c, D.ang_beg, off, beg)) || (rc = upload_at(c, D.ang_end, off, end)) ||
      (rc = upload_at(c, D.ang_pq, h.at[P_AINC], pq_of(code, h.gi, h.gk))))
    return rc;
  // natives
  build_csr(n, h.ai, h.aj, h.at[P_ATT], h.at[P_TINC], off, beg, end, code);
  std::vector<int8_t> at(nt);
  std::vector<double> r2(nt), nlrc2(nt);
  std::vector<float> rf(nt);
  for (size_t k = 0; k < nt; k++) {
    at[k] = (int8_t)(h.ait[k] | (h.ajt[k] << 2));
    const double r = h.ar[k];
    // params.cpp:203-204 and neighbor_list.cpp:44-47 (sigma^6, sigma^12 and
    // the pair-list cutoff are computed from r on the device, nat_powers /
    // nat_plrc2)
    r2[k] = r * r;
    const double rn = 3.2 * r;
    nlrc2[k] = rn * rn;
    rf[k] = (float)r;
  }
  // per incidence: partner, and the pair's word and distance in the bead's own order
  const size_t nq = code.size();
  std::vector<int> partner(nq), q0(nt), q1(nt);
  std::vector<uint32_t> aw(nq);
  std::vector<double> ar(nq);
  for (size_t q = 0; q < nq; q++) {
    const size_t k = (size_t)(code[q] >> 1) - h.at[P_ATT];
    const int local = (code[q] & 1) ? h.ai[k] : h.aj[k];
    partner[q] = off + local;
    ((code[q] & 1) ? q1 : q0)[k] = (int)(h.at[P_TINC] + q);
    aw[q] = (uint32_t)local | ((uint32_t)(at[k] | ((code[q] & 1) << 4)) << kAttTShift);
    ar[q] = h.ar[k];
  }
  h.q0.assign(q0.begin(), q0.end());
  h.q1.assign(q1.begin(), q1.end());
  const size_t ta = h.at[P_ATT], tq = h.at[P_TINC];
  if ((rc = upload_at(c, D.att_beg, off, beg)) || (rc = upload_at(c, D.att_end, off, end)) ||
      (rc = upload_at(c, D.att_code, tq, code)) || (rc = upload_at(c, D.att_partner, tq, partner)) ||
      (rc = upload_at(c, D.att_w, tq, aw)) || (rc = upload_at(c, D.att_r, tq, ar)) ||
      (rc = upload_at(c, D.am_i, ta, glob(h.ai))) || (rc = upload_at(c, D.am_j, ta, glob(h.aj))) ||
      (rc = upload_at(c, D.am_q0, ta, q0)) || (rc = upload_at(c, D.am_t, ta, at)) ||
      (rc = upload_at(c, D.am_r2, ta, r2)) || (rc = upload_at(c, D.am_nlrc2, ta, nlrc2)) ||
      (rc = upload_at(c, D.am_rf, ta, rf)))
    return rc;
  if (nt) HIP_OK(c, hipMemsetAsync(D.att_flag + tq, 0, 2 * nt, c->stream));
  // explicit non-natives
  if (h.rep_explicit) {
    build_csr(n, h.ri, h.rj, h.at[P_REP], h.at[P_RINC], off, beg, end, code);
    std::vector<int8_t> rt(nr);
    for (size_t k = 0; k < nr; k++) rt[k] = (int8_t)(h.rit[k] | (h.rjt[k] << 2));
    if ((rc = upload_at(c, D.rep_beg, off, beg)) || (rc = upload_at(c, D.rep_end, off, end)) ||
        (rc = upload_at(c, D.rep_code, h.at[P_RINC], code)) || (rc = upload_at(c, D.rm_i, h.at[P_REP], glob(h.ri))) ||
        (rc = upload_at(c, D.rm_j, h.at[P_REP], glob(h.rj))) || (rc = upload_at(c, D.rm_t, h.at[P_REP], rt)))
      return rc;
  } else {
    std::vector<int> z(n, 0);
    if ((rc = upload_at(c, D.rep_beg, off, z)) || (rc = upload_at(c, D.rep_end, off, z))) return rc;
  }
  // zero state
  for (int s = 0; s < 2; s++)
    for (double* p : {D.ux[s], D.uy[s], D.uz[s], D.px[s], D.py[s], D.pz[s], D.ix[s], D.iy[s], D.iz[s]})
      HIP_OK(c, hipMemsetAsync(p + off, 0, sizeof(double) * n, c->stream));
  for (double* p : {D.vx, D.vy, D.vz, D.fx, D.fy, D.fz})
    HIP_OK(c, hipMemsetAsync(p + off, 0, sizeof(double) * n, c->stream));
  HIP_OK(c, hipMemsetAsync(D.cnt + off, 0, sizeof(int) * n, c->stream));
  HIP_OK(c, hipMemsetAsync(D.red + (size_t)off * kRed, 0, sizeof(double) * n * kRed, c->stream));
  HIP_OK(c, hipMemsetAsync(D.grid + id, 0, sizeof(GridP), c->stream));
  h.on_device = true;
  D.nb = (int)c->pool[P_BEAD].hw;
  h.pl_fresh = false;  // the new chain has no pair list yet
  if ((rc = upload_chainp(c, id))) return rc;
  HIP_OK(c, hipStreamSynchronize(c->stream));
  // No-cutoff chains interact through their full master lists, as the loader
  // leaves them (params.cpp:210-230): build that list once.
  if (h.p.cutoff == SOPG_CUTOFF_NONE) {
    launch_rebuild(c->list_lanes, D, c->d_T, c->cur, c->par, 0, SOPG_CUTOFF_NONE, off, off + n, id, 1, 0, 0, c->stream);
    HIP_OK(c, hipGetLastError());
    return check_err_flag(c);
  }
  return SOPG_OK;
}

// First device operation: allocate the pools (sized for the chains added so
// far, or the reservation) and materialise every pending chain.
int ensure_device(sopg_ctx* c) {
  if (c->device_ready) return SOPG_OK;
  HIP_OK(c, hipSetDevice(c->device));
  Dev& D = c->D;
  size_t beads = c->reserve_beads, nch = std::max(c->reserve_chains, c->chains.size());
  size_t nbnd = 0, nang = 0, natt = 0, nrep = 0;
  int maxn = 0;
  bool any_none = false;
  size_t pend = 0;
  for (auto& h : c->chains) {
    if (!h.active) continue;
    pend += h.n;
    nbnd += h.bi.size();
    nang += h.gi.size();
    natt += h.ai.size();
    if (h.rep_explicit) nrep += h.ri.size();
    maxn = std::max(maxn, h.n);
    any_none |= h.p.cutoff == SOPG_CUTOFF_NONE;
  }
  beads = std::max(beads, pend);
  int cap = c->cap_user;
  if (cap <= 0) {
    cap = 1024;
    if (any_none) cap = std::max(cap, maxn);
  }
  D.cap = (std::max(cap, 1) + 3) / 4 * 4;  // whole 4-entry chunks of the slot tiles
  D.bitmap_words = 1;
  register_arrays(c);
  int rc;
  const size_t init[NPOOL] = {beads, 2 * nbnd, 2 * nang, natt, 2 * natt, nrep, 2 * nrep, nch};
  for (int pi = 0; pi < NPOOL; pi++)
    if ((rc = grow_pool(c, pi, std::max<size_t>(init[pi], 1)))) return rc;
  if ((rc = dalloc(c, &D.err, 1)) || (rc = dalloc(c, &c->d_T, 1)) || (rc = dalloc(c, &c->d_relb, 2)) ||
      (rc = dalloc(c, &D.plan, 3)))
    return rc;
  HIP_OK(c, hipMemsetAsync(D.err, 0, sizeof(int), c->stream));
  HIP_OK(c, hipEventCreateWithFlags(&c->ev_cp, hipEventDisableTiming));
  HIP_OK(c, hipHostMalloc((void**)&c->err_pin, sizeof(int), hipHostMallocDefault));
  const int no_plan[3] = {INT_MIN, INT_MIN, INT_MIN};
  HIP_OK(c, hipMemcpyAsync(D.plan, no_plan, sizeof no_plan, hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  HIP_OK(c, hipMemsetAsync(c->d_relb, 0, 2 * sizeof(int), c->stream));
  D.relb = c->d_relb;
  HIP_OK(c, hipMemcpyAsync(c->d_T, &c->T, sizeof(Tables), hipMemcpyHostToDevice, c->stream));
  c->device_ready = true;
  for (size_t id = 0; id < c->chains.size(); id++)
    if (c->chains[id].active && !c->chains[id].on_device && (rc = materialize(c, (int)id))) return rc;
  return upload_chainp(c);
}

int get_chain(sopg_ctx* c, int chain, HostChain** h, ChainP** P) {
  if (!c) return SOPG_ERR_ARG;
  if (chain < 0 || chain >= (int)c->chains.size() || !c->chains[chain].active)
    return fail(c, SOPG_ERR_ARG, "bad chain id");
  int rc = ensure_device(c);
  if (rc) return rc;
  *h = &c->chains[chain];
  *P = &c->cp[chain];
  return SOPG_OK;
}

bool any_replay(const sopg_ctx* c) {
  for (auto& h : c->chains)
    if (h.active && h.p.noise_mode == SOPG_NOISE_REPLAY) return true;
  return false;
}

int ensure_noise(sopg_ctx* c, size_t n) {
  if (n <= c->noise_cap) return SOPG_OK;
  HIP_OK(c, hipStreamSynchronize(c->stream));
  dfree(c, c->d_noise);
  c->d_noise = nullptr;
  c->noise_cap = 0;
  int rc = dalloc(c, &c->d_noise, n);
  if (rc) return rc;
  c->noise_cap = n;
  return SOPG_OK;
}

// The error flag's verdict once its value e has been read.
int err_verdict(sopg_ctx* c, int e) {
  if (e & 1) {
    HIP_OK(c, hipMemsetAsync(c->D.err, 0, sizeof(int), c->stream));
    return fail(c, SOPG_ERR_OVERFLOW,
                "a per-bead non-native neighbor list exceeded its capacity; raise it with sopg_set_capacity");
  }
  return SOPG_OK;
}

int check_err_flag(sopg_ctx* c) {
  int e = 0;
  HIP_OK(c, hipMemcpyAsync(&e, c->D.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return err_verdict(c, e);
}

// Chain reduction of one chain into host array out[kRed].
int reduce_chain(sopg_ctx* c, int chain, double* out) {
  launch_chain_reduce(c->D, chain, 1, c->d_out, c->stream);
  HIP_OK(c, hipGetLastError());
  HIP_OK(c, hipMemcpyAsync(out, c->d_out, sizeof(double) * kRed, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return SOPG_OK;
}

int list_counts(sopg_ctx* c, int chain, double* out) {
  const ChainP& P = c->cp[chain];
  launch_observe(c->D, c->cur, c->par, 0, 1, OBS_COUNTS, P.off, P.off + P.n, c->stream);
  HIP_OK(c, hipGetLastError());
  return reduce_chain(c, chain, out);
}

int rebuild_one(sopg_ctx* c, int chain, int mode, int* nnl_att, int* nnl_rep) {
  if (c) c->gen++;
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if (mode == SOPG_CUTOFF_CELL && (!(h->p.ncell > 0.0) || !(h->p.lcell > 0.0)))
    return fail(c, SOPG_ERR_ARG, "cell list needs ncell > 0 and lcell > 0");
  if (mode == SOPG_CUTOFF_CELL) launch_cellcoords(c->D, c->cur, P->off, P->off + P->n, c->stream);
  // a neighbor-list rebuild also (re)builds the chain's candidate list, which
  // the rebuilds of the following runs filter
  launch_rebuild(c->list_lanes, c->D, c->d_T, c->cur, c->par, 0, mode, P->off, P->off + P->n, chain, 1, c->binned,
                 mode == SOPG_CUTOFF_NEIGHBOR && c->skin > 0.0, c->stream);
  HIP_OK(c, hipGetLastError());
  if ((rc = check_err_flag(c))) return rc;
  h->pl_fresh = true;
  double out[kRed];
  if ((rc = list_counts(c, chain, out))) return rc;
  if (nnl_att) *nnl_att = (int)out[R_NLA];
  if (nnl_rep) *nnl_rep = (int)out[R_NLR];
  return SOPG_OK;
}

// kinT (sop.cpp:262-279): sum v^2 / 3N underdamped; overdamped uses the
// position increments scaled by zeta / (2h).
double kin_temp(const sopg_chain_params& p, double sumsq, int n) {
  if (p.sim_type == 2) sumsq *= p.zeta / (2.0 * p.h);
  return sumsq / (3.0 * (double)n);
}

const double* chain_noise(sopg_ctx* c, const ChainP& P, const HostChain& h, const double* noise, int* rc) {
  *rc = SOPG_OK;
  if (h.p.noise_mode == SOPG_NOISE_REPLAY) {
    if (!noise) { *rc = fail(c, SOPG_ERR_ARG, "chain uses replayed noise: pass its 3N gaussians"); return nullptr; }
    if ((*rc = ensure_noise(c, 3 * (size_t)c->D.nb))) return nullptr;
    hipError_t e = hipMemcpyAsync(c->d_noise + 3 * (size_t)P.off, noise, sizeof(double) * 3 * (size_t)P.n,
                                  hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) { *rc = fail(c, SOPG_ERR_HIP, hipGetErrorString(e)); return nullptr; }
    return c->d_noise;
  }
  if (noise) { *rc = fail(c, SOPG_ERR_ARG, "chain uses the device generator: pass noise = NULL"); return nullptr; }
  return nullptr;
}

int parse_topology(sopg_ctx* c, const sopg_topology* t, HostChain& h) {
  const int n = t->nbead;
  if (n < 1 || n >= kMaxChainBeads) return fail(c, SOPG_ERR_ARG, "nbead out of range");
  h.n = n;
  auto bead = [&](int x) { return x >= 1 && x <= n; };
  auto typ = [&](int x) { return x >= 0 && x <= 2; };
  for (int k = 0; k < t->nbnd; k++) {
    if (!bead(t->ibnd[k]) || !bead(t->jbnd[k])) return fail(c, SOPG_ERR_ARG, "bond bead index out of range");
    h.bi.push_back(t->ibnd[k] - 1); h.bj.push_back(t->jbnd[k] - 1); h.br0.push_back(t->r0bnd[k]);
  }
  for (int k = 0; k < t->nang; k++) {
    if (!bead(t->iang[k]) || !bead(t->kang[k])) return fail(c, SOPG_ERR_ARG, "angle bead index out of range");
    h.gi.push_back(t->iang[k] - 1); h.gk.push_back(t->kang[k] - 1);
  }
  bool lex = true;
  long prev = -1;
  for (int k = 0; k < t->ncon_att; k++) {
    const int i = t->att_i[k], j = t->att_j[k];
    if (!bead(i) || !bead(j) || !typ(t->att_it[k]) || !typ(t->att_jt[k]))
      return fail(c, SOPG_ERR_ARG, "native contact index/type out of range");
    h.ai.push_back(i - 1); h.aj.push_back(j - 1); h.ait.push_back(t->att_it[k]); h.ajt.push_back(t->att_jt[k]);
    h.ar.push_back(t->att_r[k]);
    const long key = (long)(i - 1) * n + (j - 1);
    if (!(i < j && j - i > 2 && key > prev)) lex = false;
    prev = key;
  }
  h.type.assign(n, 1);
  if (t->bead_type)
    for (int b = 0; b < n; b++) {
      if (!typ(t->bead_type[b])) return fail(c, SOPG_ERR_ARG, "bead type out of range");
      h.type[b] = (int8_t)t->bead_type[b];
    }
  if (t->ncon_rep < 0) {
    // implicit complement list
    if (!lex) return fail(c, SOPG_ERR_ARG, "implicit non-native list needs natives with j-i>2 in (i,j) order");
    h.rep_explicit = false;
    return SOPG_OK;
  }
  for (int k = 0; k < t->ncon_rep; k++)
    if (!bead(t->rep_i[k]) || !bead(t->rep_j[k]) || !typ(t->rep_it[k]) || !typ(t->rep_jt[k]))
      return fail(c, SOPG_ERR_ARG, "non-native contact index/type out of range");
  // Is the given list exactly the implicit complement with per-bead types?
  bool same = lex;
  if (same) {
    std::vector<int8_t> ty(n, -1);
    auto settype = [&](int b, int tv) {
      if (ty[b] < 0) ty[b] = (int8_t)tv;
      return ty[b] == tv;
    };
    long k = 0, a = 0;
    const long na = (long)h.ai.size();
    for (int i = 0; i < n && same; i++)
      for (int j = i + 3; j < n && same; j++) {
        if (a < na && h.ai[a] == i && h.aj[a] == j) { a++; continue; }
        if (k >= t->ncon_rep || t->rep_i[k] != i + 1 || t->rep_j[k] != j + 1 || !settype(i, t->rep_it[k]) ||
            !settype(j, t->rep_jt[k]))
          same = false;
        k++;
      }
    if (same && k != t->ncon_rep) same = false;
    if (same && t->bead_type)
      for (int b = 0; b < n; b++)
        if (ty[b] >= 0 && ty[b] != h.type[b]) same = false;
    if (same && !t->bead_type)
      for (int b = 0; b < n; b++) h.type[b] = ty[b] >= 0 ? ty[b] : 1;
  }
  h.rep_explicit = !same;
  if (h.rep_explicit)
    for (int k = 0; k < t->ncon_rep; k++) {
      h.ri.push_back(t->rep_i[k] - 1); h.rj.push_back(t->rep_j[k] - 1);
      h.rit.push_back(t->rep_it[k]); h.rjt.push_back(t->rep_jt[k]);
    }
  return SOPG_OK;
}

}  // namespace

// =================================================================== C ABI

namespace {

// Captured step segments bake in which rebuild kernels run (the candidate
// filter only with a skin): a change of skin or tables drops them.
int drop_graphs(sopg_ctx* c) {
  if (c->gcache.empty()) return SOPG_OK;
  HIP_OK(c, hipStreamSynchronize(c->stream));
  for (auto& ge : c->gcache) (void)hipGraphExecDestroy(ge.exec);
  c->gcache.clear();
  return SOPG_OK;
}

// Tables of bead blocks (block k: beads [k, k + 1) * per) that hold a bead of
// an active chain (not a resident one, unless with_resident).  Chains that
// left free beads behind (a retired job) leave whole blocks of [0, nb) empty;
// a launch given the table covers its blocks only: the step kernel (256 / L
// beads per block) and the binned rebuild (kBeadTile beads per tile).
// *blk = null (every block) when no block is empty.
int block_table(sopg_ctx* c, sopg_ctx::BlockTable& bt, int per, bool with_resident, const int** blk, int* nblk) {
  const int total = (c->D.nb + per - 1) / per;
  // the block range [k0, k1] of each chain the launch serves, in chain order:
  // the table of the last call stands while these are unchanged
  std::vector<std::pair<int, int>> r;
  for (const HostChain& h : c->chains)
    if (h.active && h.on_device && h.n > 0 && (with_resident || !h.resident)) {
      const int b0 = (int)h.at[P_BEAD];
      r.push_back({b0 / per, (b0 + h.n - 1) / per});
    }
  if (bt.valid && bt.per == per && bt.total == total && bt.ranges == r) {
    *blk = bt.out;
    *nblk = bt.out ? bt.nout : total;
    return SOPG_OK;
  }
  bt.valid = false;  // set again once the table below is in place
  bt.ranges = r;
  *blk = nullptr;
  *nblk = total;
  std::sort(r.begin(), r.end());
  std::vector<std::pair<int, int>> m;  // merged ranges
  for (auto& x : r) {
    if (!m.empty() && x.first <= m.back().second + 1) m.back().second = std::max(m.back().second, x.second);
    else m.push_back(x);
  }
  int rc;
  if (!m.empty() && !(m.size() == 1 && m[0].first == 0 && m[0].second == total - 1)) {
    std::vector<int> t;
    for (auto& x : m)
      for (int k = x.first; k <= x.second; k++) t.push_back(k);
    if (t != bt.dev) {
      if (t.size() > bt.cap) {
        // cached graphs hold the old table's address
        if ((rc = drop_graphs(c))) return rc;
        HIP_OK(c, hipStreamSynchronize(c->stream));
        dfree(c, bt.d);
        bt.d = nullptr;
        bt.cap = 0;
        bt.dev.clear();
        const size_t cap = std::max<size_t>(t.size(), (size_t)total);
        if ((rc = dalloc(c, &bt.d, cap))) return rc;
        bt.cap = cap;
      }
      if ((rc = upload_at(c, bt.d, 0, t))) return rc;
      bt.dev.swap(t);
    }
    *blk = bt.d;
    *nblk = (int)bt.dev.size();
  }
  bt.valid = true;
  bt.per = per;
  bt.total = total;
  bt.out = *blk;
  bt.nout = *nblk;
  return SOPG_OK;
}
int step_blocks(sopg_ctx* c, int L, bool with_resident, const int** blk, int* nblk) {
  return block_table(c, c->step_tab, 256 / L, with_resident, blk, nblk);
}

// Candidate lists built with other cutoffs or another skin are rebuilt at
// each chain's next neighbor-list rebuild.
int invalidate_candidates(sopg_ctx* c, int only = -1) {
  if (!c->device_ready) return SOPG_OK;
  const int zero = 0;
  for (int id = 0; id < (int)c->chains.size(); id++)
    if (c->chains[id].active && c->chains[id].on_device && (only < 0 || id == only))
      HIP_OK(c, hipMemcpyAsync(&c->D.grid[id].cand_valid, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return SOPG_OK;
}

}  // namespace

extern "C" {

int sopg_create(int device, sopg_ctx** out) {
  if (!out) return SOPG_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SOPG_ERR_HIP;
  if (device < 0 || device >= n) return SOPG_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return SOPG_ERR_HIP;
  sopg_ctx* c = new sopg_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SOPG_ERR_HIP;
  }
  sopg_default_tables(&c->tab);
  derive_tables(c->tab, c->skin, c->T);
  // device address of this GPU's fixed record staging (the device is current);
  // on failure it stays null and sopg_run_noalloc reports it
  if (scratch_address(device, &c->d_scratch) != 0) c->d_scratch = nullptr;

  *out = c;
  return SOPG_OK;
}

int sopg_destroy(sopg_ctx* c) {
  if (!c) return SOPG_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int pi = 0; pi < NPOOL; pi++)
    for (Arr& a : c->arrs[pi])
      if (*a.p) (void)hipFree(*a.p);
  for (void* p : c->allocs) (void)hipFree(p);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->ev_res) (void)hipEventDestroy(c->ev_res);
  if (c->ev_cp) (void)hipEventDestroy(c->ev_cp);
  if (c->cp_pin) (void)hipHostFree(c->cp_pin);
  if (c->err_pin) (void)hipHostFree(c->err_pin);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  for (auto& ge : c->gcache) (void)hipGraphExecDestroy(ge.exec);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SOPG_OK;
}

const char* sopg_last_error(const sopg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int sopg_default_tables(sopg_tables* t) {
  if (!t) return SOPG_ERR_ARG;
  std::memset(t, 0, sizeof *t);
  // global.cpp:60-78
  const double catt[3][3] = {{0, 0, 0}, {0, 0.7, 0.8}, {0, 0.8, 1.0}};
  const double crep[3][3] = {{0, 0, 0}, {0, 1.0, 1.0}, {0, 1.0, 1.0}};
  const double fatt[3][3] = {{0, 0, 0}, {0, -12.0 * 1.0, -12.0 * 0.8}, {0, -12.0 * 0.8, -12.0 * 0.7}};
  const double frep[3][3] = {{0, 0, 0}, {0, -6.0 * 1.0, -6.0 * 1.0}, {0, -6.0 * 1.0, -6.0 * 1.0}};
  const double srep[3][3] = {{0, 0, 0}, {0, 3.8, 5.4}, {0, 5.4, 7.0}};
  std::memcpy(t->coeff_att, catt, sizeof catt);
  std::memcpy(t->coeff_rep, crep, sizeof crep);
  std::memcpy(t->force_coeff_att, fatt, sizeof fatt);
  std::memcpy(t->force_coeff_rep, frep, sizeof frep);
  std::memcpy(t->sigma_rep, srep, sizeof srep);
  // params.cpp:268-293
  t->k_bnd = 20.0;
  const double R0 = 2.0;
  t->R0sq = R0 * R0;
  t->e_bnd_coeff = t->k_bnd * t->R0sq / 2.0;
  const double sigma_ss6 = std::pow(3.5, 6.0);
  const double epsilon_ss = 1.0;
  t->e_ang_ss_coeff = epsilon_ss * sigma_ss6;
  t->f_ang_ss_coeff = 6.0 * t->e_ang_ss_coeff;
  return SOPG_OK;
}

int sopg_set_tables(sopg_ctx* c, const sopg_tables* t) {
  if (c) c->gen++;
  if (!c || !t) return SOPG_ERR_ARG;
  c->tab = *t;
  derive_tables(c->tab, c->skin, c->T);
  if (c->device_ready) {
    HIP_OK(c, hipMemcpyAsync(c->d_T, &c->T, sizeof(Tables), hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    int rc;
    if ((rc = upload_chainp(c)) || (rc = drop_graphs(c))) return rc;
    return invalidate_candidates(c);
  }
  return SOPG_OK;
}

int sopg_set_lanes(sopg_ctx* c, int step_lanes, int list_lanes) {
  if (!c) return SOPG_ERR_ARG;
  auto ok = [](int l) { return l >= 1 && l <= 64 && !(l & (l - 1)); };
  if (step_lanes == 0) step_lanes = 8;
  if (list_lanes == 0) list_lanes = 64;
  if (!ok(step_lanes) || !ok(list_lanes)) return fail(c, SOPG_ERR_ARG, "lanes: power of two in 1..64");
  c->lanes = step_lanes;
  c->list_lanes = list_lanes;
  return SOPG_OK;
}

int sopg_get_lanes(const sopg_ctx* c, int* step_lanes, int* list_lanes) {
  if (!c) return SOPG_ERR_ARG;
  if (step_lanes) *step_lanes = c->lanes;
  if (list_lanes) *list_lanes = c->list_lanes;
  return SOPG_OK;
}

int sopg_set_binned(sopg_ctx* c, int enable) {
  if (!c) return SOPG_ERR_ARG;
  c->binned = enable ? 1 : 0;
  return SOPG_OK;
}

int sopg_set_capacity(sopg_ctx* c, int cap) {
  if (!c) return SOPG_ERR_ARG;
  c->gen++;
  if (c->device_ready) return fail(c, SOPG_ERR_ARG, "capacity must be set before the first device operation");
  c->cap_user = cap;
  return SOPG_OK;
}

int sopg_reserve(sopg_ctx* c, long beads, int chains) {
  if (c) c->gen++;
  if (!c || beads < 0 || chains < 0) return SOPG_ERR_ARG;
  if (c->device_ready) {
    int rc;
    if ((rc = grow_pool(c, P_BEAD, (size_t)beads)) || (rc = grow_pool(c, P_CHAIN, (size_t)chains))) return rc;
    return upload_chainp(c);
  }
  c->reserve_beads = (size_t)beads;
  c->reserve_chains = (size_t)chains;
  return SOPG_OK;
}

int sopg_add_chain(sopg_ctx* c, const sopg_topology* t, const sopg_chain_params* p, int* id) {
  if (c) c->gen++;
  if (!c || !t) return SOPG_ERR_ARG;
  int rc = check_params(c, p);
  if (rc) return rc;
  HostChain h;
  h.p = *p;
  if ((rc = parse_topology(c, t, h))) return rc;
  for (double r : h.ar) h.max_r2 = std::max(h.max_r2, r * r);
  for (int8_t ty : h.type) h.type_mask |= 1 << ty;
  h.active = true;
  // reuse the lowest free chain slot
  size_t slot = c->chains.size();
  for (size_t k = 0; k < c->chains.size(); k++)
    if (!c->chains[k].active) { slot = k; break; }
  if (slot == c->chains.size()) c->chains.emplace_back();
  c->chains[slot] = std::move(h);
  if (c->device_ready && (rc = materialize(c, (int)slot))) {
    c->chains[slot].active = false;
    return rc;
  }
  if (id) *id = (int)slot;
  return SOPG_OK;
}

int sopg_remove_chain(sopg_ctx* c, int chain) {
  if (!c) return SOPG_ERR_ARG;
  c->gen++;
  if (chain < 0 || chain >= (int)c->chains.size() || !c->chains[chain].active)
    return fail(c, SOPG_ERR_ARG, "bad chain id");
  HostChain& h = c->chains[chain];
  if (h.on_device) {
    HIP_OK(c, hipStreamSynchronize(c->stream));
    const std::vector<int> none(h.n, -1);
    int rc = upload_at(c, c->D.chain, h.at[P_BEAD], none);
    if (rc) return rc;
    const size_t nb = h.bi.size(), na = h.gi.size(), nt = h.ai.size(), nr = h.rep_explicit ? h.ri.size() : 0;
    const size_t len[NPOOL] = {(size_t)h.n, 2 * nb, 2 * na, nt, 2 * nt, nr, 2 * nr, 0};
    for (int pi = 0; pi < P_CHAIN; pi++) c->pool[pi].give(h.at[pi], len[pi]);
    c->D.nb = (int)c->pool[P_BEAD].hw;  // lower when the chain was the last in the bead range
  }
  h = HostChain();
  return upload_chainp(c, chain);
}

int sopg_set_chain_params(sopg_ctx* c, int chain, const sopg_chain_params* p) {
  if (!c) return SOPG_ERR_ARG;
  c->gen++;
  if (chain < 0 || chain >= (int)c->chains.size() || !c->chains[chain].active)
    return fail(c, SOPG_ERR_ARG, "bad chain id");
  int rc = check_params(c, p);
  if (rc) return rc;
  c->chains[chain].p = *p;
  // a new box or cutoff changes what the chain's candidate list must hold
  if ((rc = upload_chainp(c, chain))) return rc;
  return invalidate_candidates(c, chain);
}

int sopg_nchains(const sopg_ctx* c) {
  if (!c) return 0;
  int n = 0;
  for (auto& h : c->chains) n += h.active ? 1 : 0;
  return n;
}

int sopg_chain_nbead(const sopg_ctx* c, int chain) {
  return (c && chain >= 0 && chain < (int)c->chains.size() && c->chains[chain].active) ? c->chains[chain].n : -1;
}

int sopg_put_state(sopg_ctx* c, int chain, const double* pos, const double* unc, const double* vel,
                   const double* force) {
  if (c) c->gen++;
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  const int n = P->n, off = P->off;
  std::vector<double> sx(n), sy(n), sz(n);
  auto put = [&](const double* a, double* dx, double* dy, double* dz) -> int {
    if (!a) return SOPG_OK;
    for (int b = 0; b < n; b++) { sx[b] = a[3 * b]; sy[b] = a[3 * b + 1]; sz[b] = a[3 * b + 2]; }
    HIP_OK(c, hipMemcpyAsync(dx + off, sx.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipMemcpyAsync(dy + off, sy.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipMemcpyAsync(dz + off, sz.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SOPG_OK;
  };
  Dev& D = c->D;
  if ((rc = put(pos, D.px[c->cur], D.py[c->cur], D.pz[c->cur])) ||
      (rc = put(unc, D.ux[c->cur], D.uy[c->cur], D.uz[c->cur])) || (rc = put(vel, D.vx, D.vy, D.vz)) ||
      (rc = put(force, D.fx, D.fy, D.fz)))
    return rc;
  if (unc) h->pl_fresh = false;
  return SOPG_OK;
}

int sopg_get_state(sopg_ctx* c, int chain, double* pos, double* unc, double* vel, double* force) {
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  const int n = P->n, off = P->off;
  std::vector<double> sx(n), sy(n), sz(n);
  auto get = [&](double* a, const double* dx, const double* dy, const double* dz) -> int {
    if (!a) return SOPG_OK;
    HIP_OK(c, hipMemcpyAsync(sx.data(), dx + off, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipMemcpyAsync(sy.data(), dy + off, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipMemcpyAsync(sz.data(), dz + off, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    for (int b = 0; b < n; b++) { a[3 * b] = sx[b]; a[3 * b + 1] = sy[b]; a[3 * b + 2] = sz[b]; }
    return SOPG_OK;
  };
  Dev& D = c->D;
  if ((rc = get(pos, D.px[c->cur], D.py[c->cur], D.pz[c->cur])) ||
      (rc = get(unc, D.ux[c->cur], D.uy[c->cur], D.uz[c->cur])) || (rc = get(vel, D.vx, D.vy, D.vz)) ||
      (rc = get(force, D.fx, D.fy, D.fz)))
    return rc;
  return SOPG_OK;
}

int sopg_update_neighbor_list(sopg_ctx* c, int chain, int* nnl_att, int* nnl_rep) {
  return rebuild_one(c, chain, SOPG_CUTOFF_NEIGHBOR, nnl_att, nnl_rep);
}

int sopg_update_cell_list(sopg_ctx* c, int chain, int* nnl_att, int* nnl_rep) {
  return rebuild_one(c, chain, SOPG_CUTOFF_CELL, nnl_att, nnl_rep);
}

int sopg_update_pair_list(sopg_ctx* c, int chain, int* nil_att, int* nil_rep) {
  if (c) c->gen++;
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  launch_pairfilter(c->list_lanes, c->D, c->d_T, c->cur, c->par, P->off, P->off + P->n, c->stream);
  HIP_OK(c, hipGetLastError());
  h->pl_fresh = true;
  double out[kRed];
  if ((rc = list_counts(c, chain, out))) return rc;
  if (nil_att) *nil_att = (int)out[R_PLA];
  if (nil_rep) *nil_rep = (int)out[R_PLR];
  return SOPG_OK;
}

int sopg_force_eval(sopg_ctx* c, int chain, const double* noise) {
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if ((rc = upload_chainp(c, chain))) return rc;
  const double* dn = chain_noise(c, *P, *h, noise, &rc);
  if (rc) return rc;
  launch_force(c->lanes, c->D, c->d_T, c->cur, c->par, 0, F_STORE | F_ALLTERMS, false, dn, P->off, P->off + P->n,
               c->stream);
  HIP_OK(c, hipGetLastError());
  HIP_OK(c, hipStreamSynchronize(c->stream));
  h->noise_base++;
  return SOPG_OK;
}

int sopg_force_terms(sopg_ctx* c, int chain, int terms, const double* noise) {
  if (c) c->gen++;
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if (terms & ~(SOPG_TERM_RANDOM | SOPG_TERM_FENE | SOPG_TERM_ANGULAR | SOPG_TERM_VDW))
    return fail(c, SOPG_ERR_ARG, "unknown force term");
  if ((rc = upload_chainp(c, chain))) return rc;
  const double* dn = nullptr;
  if (terms & SOPG_TERM_RANDOM) {
    dn = chain_noise(c, *P, *h, noise, &rc);
    if (rc) return rc;
  }
  int flags = F_STORE | F_ACCUM;
  if (terms & SOPG_TERM_RANDOM) flags |= F_NOISE;
  if (terms & SOPG_TERM_FENE) flags |= F_BOND;
  if (terms & SOPG_TERM_ANGULAR) flags |= F_ANGLE;
  if (terms & SOPG_TERM_VDW) flags |= F_VDW;
  launch_force(c->lanes, c->D, c->d_T, c->cur, c->par, 0, flags, false, dn, P->off, P->off + P->n, c->stream);
  HIP_OK(c, hipGetLastError());
  HIP_OK(c, hipStreamSynchronize(c->stream));
  if (terms & SOPG_TERM_RANDOM) h->noise_base++;
  return SOPG_OK;
}

int sopg_energy_eval(sopg_ctx* c, int chain, sopg_energies* out) {
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if ((rc = upload_chainp(c, chain))) return rc;
  launch_force(c->lanes, c->D, c->d_T, c->cur, c->par, 0, F_ENERGY_ALL | F_ALLTERMS, false, nullptr, P->off,
               P->off + P->n, c->stream);
  HIP_OK(c, hipGetLastError());
  double r[kRed];
  if ((rc = reduce_chain(c, chain, r))) return rc;
  if (out) {
    out->e_bnd = r[R_EBND] * -c->T.e_bnd_coeff;  // energy.cpp:148
    out->e_ang_ss = r[R_EANG];
    out->e_vdw_rr_att = r[R_EATT];
    out->e_vdw_rr_rep = r[R_EREP];
    out->e_vdw_rr = out->e_vdw_rr_att + out->e_vdw_rr_rep;
    out->rna_etot = out->e_bnd + out->e_ang_ss + out->e_vdw_rr;  // energy.cpp:16
  }
  return SOPG_OK;
}

int sopg_underdamped_iteration(sopg_ctx* c, int chain, const double* noise) {
  if (c) c->gen++;
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if ((rc = upload_chainp(c, chain))) return rc;
  const double* dn = chain_noise(c, *P, *h, noise, &rc);
  if (rc) return rc;
  launch_posupdate(c->D, c->cur, c->cur, P->off, P->off + P->n, c->stream);
  launch_force(c->lanes, c->D, c->d_T, c->cur, c->par, 0, F_STORE | F_VEL | F_ALLTERMS, false, dn, P->off,
               P->off + P->n, c->stream);
  HIP_OK(c, hipGetLastError());
  HIP_OK(c, hipStreamSynchronize(c->stream));
  h->noise_base++;
  h->pl_fresh = false;
  return SOPG_OK;
}

int sopg_overdamped_iteration(sopg_ctx* c, int chain, const double* noise) {
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if (h->p.sim_type != 2) return fail(c, SOPG_ERR_ARG, "chain is not overdamped (sim_type 2)");
  return sopg_underdamped_iteration(c, chain, noise);
}

int sopg_calculate_observables(sopg_ctx* c, int chain, sopg_observables* out) {
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  launch_observe(c->D, c->cur, c->par, 0, 1, OBS_PHYS, P->off, P->off + P->n, c->stream);
  HIP_OK(c, hipGetLastError());
  double r[kRed];
  if ((rc = reduce_chain(c, chain, r))) return rc;
  if (out) {
    out->contct_nat = (int)r[R_CONT];
    out->Q = (double)out->contct_nat / (double)P->natt;                // sop.cpp:240
    out->rgsq = r[R_RG] / ((double)P->n * (double)P->n);              // sop.cpp:258
    out->kinT = kin_temp(h->p, r[R_V2], P->n);
  }
  return SOPG_OK;
}

int sopg_get_list(sopg_ctx* c, int chain, int which, int* oi, int* oj, long cap, long* nout) {
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if (which < SOPG_LIST_NL_ATT || which > SOPG_LIST_PL_REP) return fail(c, SOPG_ERR_ARG, "bad list id");
  std::vector<std::pair<int, int>> L;
  if (which == SOPG_LIST_NL_ATT || which == SOPG_LIST_PL_ATT) {
    const size_t nq = 2 * (size_t)P->natt, tq = h->at[P_TINC];
    std::vector<uint8_t> f(nq);
    if (nq) HIP_OK(c, hipMemcpyAsync(f.data(), c->D.att_flag + tq, nq, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    const int bit = (which == SOPG_LIST_NL_ATT) ? 1 : (2 << c->par);
    for (int k = 0; k < P->natt; k++)
      if (f[(size_t)h->q0[k] - tq] & bit) L.push_back({h->ai[k] + 1, h->aj[k] + 1});
  } else {
    const int n = P->n, cap_s = c->D.cap;
    std::vector<int> cnt(n);
    std::vector<uint32_t> sl((size_t)n * cap_s);
    std::vector<int> sk;
    HIP_OK(c, hipMemcpyAsync(cnt.data(), c->D.cnt + P->off, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    rc = slot_rows_get(c, (size_t)P->off, (size_t)n, sl);
    if (rc) return rc;
    if (h->rep_explicit) {
      sk.resize((size_t)n * cap_s);
      HIP_OK(c, hipMemcpyAsync(sk.data(), c->D.slot_k + (size_t)P->off * cap_s, sizeof(int) * n * cap_s,
                               hipMemcpyDeviceToHost, c->stream));
    }
    HIP_OK(c, hipStreamSynchronize(c->stream));
    const uint32_t pbit = 1u << (30 + c->par);
    std::vector<std::pair<long, std::pair<int, int>>> ex;
    for (int b = 0; b < n; b++)
      for (int q = 0; q < cnt[b]; q++) {
        const uint32_t w = sl[(size_t)b * cap_s + q];
        if ((w >> 25) & 1) continue;  // each pair once, from its i bead
        if (which == SOPG_LIST_PL_REP && !(w & pbit)) continue;
        const int o = (int)(w & kPartnerMask);
        if (h->rep_explicit) ex.push_back({(long)sk[(size_t)b * cap_s + q], {b + 1, o + 1}});
        else L.push_back({b + 1, o + 1});
      }
    if (h->rep_explicit) {
      std::sort(ex.begin(), ex.end());
      for (auto& e : ex) L.push_back(e.second);
    }
  }
  if (nout) *nout = (long)L.size();
  for (long k = 0; k < (long)L.size() && k < cap; k++) {
    if (oi) oi[k] = L[k].first;
    if (oj) oj[k] = L[k].second;
  }
  return SOPG_OK;
}

int sopg_set_step(sopg_ctx* c, int chain, double istep, int inlup, int iup) {
  if (!c) return SOPG_ERR_ARG;
  c->gen++;
  if (chain < 0 || chain >= (int)c->chains.size() || !c->chains[chain].active)
    return fail(c, SOPG_ERR_ARG, "bad chain id");
  HostChain& h = c->chains[chain];
  // inlup = 0 (overdamped_ctrl starts from the global inlup = 0, sop.cpp:428)
  // rebuilds on the first step, as inlup = nnlup does
  if (inlup == 0) inlup = h.p.nnlup;
  if (inlup < 1 || inlup > h.p.nnlup || iup < 1 || iup > h.p.nup)
    return fail(c, SOPG_ERR_ARG, "inlup must be in [0, nnlup] and iup in [1, nup]");
  h.istep = istep;
  h.inlup = inlup;
  h.iup = iup;
  return upload_chainp(c, chain);
}

// ------------------------------------------------------------ checkpoints
//
// Blob: header, then pos, unc_pos, vel, force (AoS, 3N doubles each), the
// native NL flag of each master entry (bytes), the non-native slot count of
// each bead (ints), the slot words (pair-list bits cleared) and, for explicit
// master lists, their master indices.

namespace {
struct CkptHeader {
  uint32_t magic, version;
  int32_t nbead, natt, rep_explicit, pad;
  int64_t nslots;
  double istep;
  int32_t inlup, iup;
  uint64_t noise_base;
};
constexpr uint32_t kCkptMagic = 0x43504f53u;  // "SOPC"

int ckpt_lists(sopg_ctx* c, int chain, std::vector<uint8_t>& nl, std::vector<int>& cnt, std::vector<uint32_t>& w,
               std::vector<int>& k) {
  HostChain& h = c->chains[chain];
  const ChainP& P = c->cp[chain];
  const int n = P.n, cap = c->D.cap;
  const size_t nq = 2 * (size_t)P.natt, tq = h.at[P_TINC];
  std::vector<uint8_t> f(nq);
  std::vector<uint32_t> sl((size_t)n * cap);
  std::vector<int> sk;
  cnt.resize(n);
  if (nq) HIP_OK(c, hipMemcpyAsync(f.data(), c->D.att_flag + tq, nq, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipMemcpyAsync(cnt.data(), c->D.cnt + P.off, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
  int rc = slot_rows_get(c, (size_t)P.off, (size_t)n, sl);
  if (rc) return rc;
  if (h.rep_explicit) {
    sk.resize((size_t)n * cap);
    HIP_OK(c, hipMemcpyAsync(sk.data(), c->D.slot_k + (size_t)P.off * cap, sizeof(int) * n * cap,
                             hipMemcpyDeviceToHost, c->stream));
  }
  HIP_OK(c, hipStreamSynchronize(c->stream));
  nl.resize(P.natt);
  for (int a = 0; a < P.natt; a++) nl[a] = f[(size_t)h.q0[a] - tq] & 1;
  w.clear();
  k.clear();
  for (int b = 0; b < n; b++)
    for (int q = 0; q < cnt[b]; q++) {
      w.push_back(sl[(size_t)b * cap + q] & ~(3u << 30));
      if (h.rep_explicit) k.push_back(sk[(size_t)b * cap + q]);
    }
  return SOPG_OK;
}
}  // namespace

int sopg_checkpoint(sopg_ctx* c, int chain, void* buf, long cap, long* size) {
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  std::vector<uint8_t> nl;
  std::vector<int> cnt, k;
  std::vector<uint32_t> w;
  if ((rc = ckpt_lists(c, chain, nl, cnt, w, k))) return rc;
  const size_t n3 = 3 * (size_t)P->n;
  const size_t need = sizeof(CkptHeader) + 4 * n3 * sizeof(double) + nl.size() + cnt.size() * sizeof(int) +
                      w.size() * sizeof(uint32_t) + k.size() * sizeof(int);
  if (size) *size = (long)need;
  if (!buf) return SOPG_OK;
  if (cap < (long)need) return fail(c, SOPG_ERR_ARG, "checkpoint buffer too small (query the size with buf = NULL)");
  CkptHeader hd{kCkptMagic, 1, P->n, P->natt, h->rep_explicit ? 1 : 0, 0, (int64_t)w.size(),
                h->istep, h->inlup, h->iup, h->noise_base};
  char* o = (char*)buf;
  std::memcpy(o, &hd, sizeof hd);
  o += sizeof hd;
  double* st = (double*)o;
  if ((rc = sopg_get_state(c, chain, st, st + n3, st + 2 * n3, st + 3 * n3))) return rc;
  o += 4 * n3 * sizeof(double);
  auto put = [&](const void* p, size_t bytes) {
    if (bytes) std::memcpy(o, p, bytes);
    o += bytes;
  };
  put(nl.data(), nl.size());
  put(cnt.data(), cnt.size() * sizeof(int));
  put(w.data(), w.size() * sizeof(uint32_t));
  put(k.data(), k.size() * sizeof(int));
  return SOPG_OK;
}

int sopg_restore(sopg_ctx* c, int chain, const void* buf, long size) {
  if (c) c->gen++;
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if (!buf || size < (long)sizeof(CkptHeader)) return fail(c, SOPG_ERR_ARG, "checkpoint too short");
  CkptHeader hd;
  std::memcpy(&hd, buf, sizeof hd);
  if (hd.magic != kCkptMagic || hd.version != 1) return fail(c, SOPG_ERR_ARG, "not a chain checkpoint");
  if (hd.nbead != P->n || hd.natt != P->natt || hd.rep_explicit != (h->rep_explicit ? 1 : 0))
    return fail(c, SOPG_ERR_ARG, "checkpoint was taken from a chain of another topology");
  if (hd.inlup < 1 || hd.inlup > h->p.nnlup || hd.iup < 1 || hd.iup > h->p.nup)
    return fail(c, SOPG_ERR_ARG, "checkpoint step counters do not fit this chain's nnlup / nup");
  const size_t n = P->n, n3 = 3 * n, cap = c->D.cap;
  const size_t need = sizeof(CkptHeader) + 4 * n3 * sizeof(double) + (size_t)hd.natt + n * sizeof(int) +
                      (size_t)hd.nslots * sizeof(uint32_t) + (hd.rep_explicit ? (size_t)hd.nslots * sizeof(int) : 0);
  if ((size_t)size < need) return fail(c, SOPG_ERR_ARG, "checkpoint truncated");
  const char* in = (const char*)buf + sizeof hd;
  const double* st = (const double*)in;
  in += 4 * n3 * sizeof(double);
  const uint8_t* nl = (const uint8_t*)in;
  in += hd.natt;
  std::vector<int> cnt(n);
  std::memcpy(cnt.data(), in, n * sizeof(int));
  in += n * sizeof(int);
  const uint32_t* w = (const uint32_t*)in;
  in += (size_t)hd.nslots * sizeof(uint32_t);
  const int* kk = (const int*)in;
  // per-bead slots, validated against this chain and the context's capacity
  std::vector<uint32_t> sl(n * cap, 0u);
  std::vector<int> sk(h->rep_explicit ? n * cap : 0, 0);
  size_t q = 0;
  for (size_t b = 0; b < n; b++) {
    if (cnt[b] < 0 || (size_t)cnt[b] > cap || q + cnt[b] > (size_t)hd.nslots)
      return fail(c, SOPG_ERR_ARG, "checkpoint lists do not fit this context's per-bead capacity");
    for (int e = 0; e < cnt[b]; e++, q++) {
      if ((w[q] & kPartnerMask) >= n) return fail(c, SOPG_ERR_ARG, "checkpoint list entry out of range");
      sl[b * cap + e] = w[q] & ~(3u << 30);
      if (h->rep_explicit) sk[b * cap + e] = kk[q];
    }
  }
  if (q != (size_t)hd.nslots) return fail(c, SOPG_ERR_ARG, "checkpoint list counts inconsistent");
  if ((rc = sopg_put_state(c, chain, st, st + n3, st + 2 * n3, st + 3 * n3))) return rc;
  const size_t nq = 2 * (size_t)P->natt, tq = h->at[P_TINC];
  std::vector<uint8_t> f(nq, 0);
  for (int a = 0; a < P->natt; a++) {
    f[(size_t)h->q0[a] - tq] = nl[a] ? 1 : 0;
    f[(size_t)h->q1[a] - tq] = nl[a] ? 1 : 0;
  }
  if ((rc = upload_at(c, c->D.att_flag, tq, f)) || (rc = upload_at(c, c->D.cnt, (size_t)P->off, cnt)) ||
      (rc = slot_rows_put(c, (size_t)P->off, n, sl)))
    return rc;
  if (h->rep_explicit && (rc = upload_at(c, c->D.slot_k, (size_t)P->off * cap, sk))) return rc;
  h->istep = hd.istep;
  h->inlup = hd.inlup;
  h->iup = hd.iup;
  h->noise_base = hd.noise_base;
  h->pl_fresh = false;  // the pair list is refiltered from the restored lists before the next step
  return upload_chainp(c, chain);
}

int sopg_get_step(const sopg_ctx* c, int chain, double* istep, int* inlup, int* iup) {
  if (!c || chain < 0 || chain >= (int)c->chains.size() || !c->chains[chain].active) return SOPG_ERR_ARG;
  const HostChain& h = c->chains[chain];
  if (istep) *istep = h.istep;
  if (inlup) *inlup = h.inlup;
  if (iup) *iup = h.iup;
  return SOPG_OK;
}

int sopg_sync(sopg_ctx* c) {
  if (!c) return SOPG_ERR_ARG;
  if (c->stream) HIP_OK(c, hipStreamSynchronize(c->stream));
  return SOPG_OK;
}

int sopg_rebuild_plan(sopg_ctx* c, int chain, int* plan, int* cells) {
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  GridP G;
  HIP_OK(c, hipMemcpyAsync(&G, c->D.grid + chain, sizeof G, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  int p = SOPG_PLAN_NONE, n = 0;
  if (G.work & W_SCAN) p = SOPG_PLAN_SCAN;
  else if (G.work & W_BIN) { p = G.periodic ? SOPG_PLAN_PERIODIC : SOPG_PLAN_LOCAL; n = G.nx; }
  else if (G.work & W_FILTER) p = SOPG_PLAN_FILTER;
  if (plan) *plan = p;
  if (cells) *cells = n;
  return SOPG_OK;
}

int sopg_list_counts(sopg_ctx* c, int chain, int* nnl_att, int* nnl_rep, int* nil_att, int* nil_rep) {
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  double out[kRed];
  if ((rc = list_counts(c, chain, out))) return rc;
  if (nnl_att) *nnl_att = (int)out[R_NLA];
  if (nnl_rep) *nnl_rep = (int)out[R_NLR];
  if (nil_att) *nil_att = (int)out[R_PLA];
  if (nil_rep) *nil_rep = (int)out[R_PLR];
  return SOPG_OK;
}

int sopg_set_resident(sopg_ctx* c, int max_beads, int lanes) {
  if (c) c->gen++;
  if (!c || max_beads < 0 || lanes < 0 || lanes > 8 || (lanes & (lanes - 1)))
    return c ? fail(c, SOPG_ERR_ARG, "resident: max_beads >= 0, lanes 0 or a power of two up to 8") : SOPG_ERR_ARG;
  if (max_beads * (lanes ? lanes : 1) > 512)
    return fail(c, SOPG_ERR_ARG, "resident chains must fit one 512-thread block (max_beads * lanes <= 512)");
  c->resident_max = max_beads;
  c->resident_lanes = lanes;
  return SOPG_OK;
}

int sopg_set_skin(sopg_ctx* c, double skin) {
  if (!c) return SOPG_ERR_ARG;
  c->gen++;
  if (!(skin >= 0.0) || skin > 100.0) return fail(c, SOPG_ERR_ARG, "skin must be in [0, 100] Angstrom");
  c->skin = skin;
  derive_tables(c->tab, c->skin, c->T);
  if (c->device_ready) {
    HIP_OK(c, hipMemcpyAsync(c->d_T, &c->T, sizeof(Tables), hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    int rc;
    if ((rc = upload_chainp(c)) || (rc = drop_graphs(c)) || (rc = invalidate_candidates(c))) return rc;
  }
  return SOPG_OK;
}

int sopg_set_graphs(sopg_ctx* c, int enable) {
  if (!c) return SOPG_ERR_ARG;
  c->graphs = enable != 0;
  return SOPG_OK;
}

int sopg_graph_stats(const sopg_ctx* c, long* captures, long* launches) {
  if (!c) return SOPG_ERR_ARG;
  if (captures) *captures = c->graph_captures;
  if (launches) *launches = c->graph_launches;
  return SOPG_OK;
}

int sopg_profile(sopg_ctx* c, int enable) {
  if (!c) return SOPG_ERR_ARG;
  c->prof = enable != 0;
  c->ev_pending.clear();
  c->ev_next = 0;
  for (int k = 0; k < 4; k++) { c->prof_ms[k] = 0.0; c->prof_n[k] = 0; }
  return SOPG_OK;
}

int sopg_profile_read(sopg_ctx* c, int kernel, double* total_ms, long* launches) {
  if (!c || kernel < 0 || kernel > 3) return SOPG_ERR_ARG;
  if (total_ms) *total_ms = c->prof_ms[kernel];
  if (launches) *launches = c->prof_n[kernel];
  return SOPG_OK;
}

// Autotuner: times the step kernel (all force terms, no stores) at every lane
// count on the context's current chains and keeps the fastest.  Results do not
// depend on the lane count, so this only changes speed.
int sopg_autotune(sopg_ctx* c, int* step_lanes) {
  if (!c) return SOPG_ERR_ARG;
  c->gen++;
  int rc = ensure_device(c);
  if (rc) return rc;
  if ((rc = upload_chainp(c))) return rc;
  const int nb = c->D.nb;
  if (nb == 0) return fail(c, SOPG_ERR_ARG, "no chains");
  hipEvent_t e0, e1;
  HIP_OK(c, hipEventCreate(&e0));
  HIP_OK(c, hipEventCreate(&e1));
  double best = 1e300;
  int bestL = c->lanes;
  // The run's step kernel with its next-step pair-list test.  Without
  // F_STORE nothing of the state is written but the pair-list bits of the
  // other parity, which the next step of a run rewrites before any reads.
  for (int L : {1, 2, 4, 8, 16, 32, 64}) {
    const int* blk = nullptr;
    int nblk = 0;
    if ((rc = step_blocks(c, L, true, &blk, &nblk))) break;
    launch_force(L, c->D, c->d_T, c->cur, c->par, 0, F_ALLTERMS, true, nullptr, 0, nb, c->stream, blk, nblk);
    (void)hipEventRecord(e0, c->stream);
    for (int k = 0; k < 5; k++)
      launch_force(L, c->D, c->d_T, c->cur, c->par, 0, F_ALLTERMS, true, nullptr, 0, nb, c->stream, blk, nblk);
    (void)hipEventRecord(e1, c->stream);
    HIP_OK(c, hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_OK(c, hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) { best = ms; bestL = L; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  HIP_OK(c, hipGetLastError());
  c->lanes = bestL;
  if (step_lanes) *step_lanes = bestL;
  return SOPG_OK;
}

// The body of sopg_run and sopg_run_noalloc.  fixed: stage the records in
// this GPU's file-scope scratch array and allocate no device memory -- no
// replayed noise (it is staged in context memory), no graph capture (an
// instantiated graph holds device memory), no resident kernel (its chain and
// record-slot lists are context memory), no kernel timing.
static int run_impl(sopg_ctx* c, long nsteps, const double* noise, sopg_record* rec, long rec_cap, long* nrec,
                    bool fixed) {
  if (!c) return SOPG_ERR_ARG;
  if (nrec) *nrec = 0;
  if (nsteps < 0) return fail(c, SOPG_ERR_ARG, "nsteps < 0");
  int rc;
  if (fixed) {
    if (!c->device_ready)
      return fail(c, SOPG_ERR_ARG, "sopg_run_noalloc: the context allocates its device memory at its first device "
                                   "operation; add the chains and set their state first");
    if (!c->d_scratch) return fail(c, SOPG_ERR_HIP, "sopg_run_noalloc: no device address for the scratch array");
    if (c->prof) return fail(c, SOPG_ERR_ARG, "sopg_run_noalloc: kernel timing is on (sopg_profile)");
  } else if ((rc = ensure_device(c))) {
    return rc;
  }
  if (nsteps == 0 || sopg_nchains(c) == 0) return SOPG_OK;
  const bool replay = any_replay(c);
  if (fixed && replay)
    return fail(c, SOPG_ERR_ARG, "sopg_run_noalloc: replayed-noise chains need sopg_run (the noise is staged in "
                                 "context memory)");
  if (replay && !noise) return fail(c, SOPG_ERR_ARG, "replayed-noise chains need the run's noise");
  if (!replay && noise) return fail(c, SOPG_ERR_ARG, "all chains use the device generator: noise must be NULL");
  // small chains run in the resident kernel (not with replayed noise or kernel timing)
  std::vector<int> rlist;
  int rmaxn = 0;
  if (c->resident_max > 0 && !noise && !c->prof && !fixed)
    for (int id = 0; id < (int)c->chains.size(); id++) {
      HostChain& h = c->chains[id];
      if (h.active && h.n <= c->resident_max) {
        h.resident = true;
        rlist.push_back(id);
        rmaxn = std::max(rmaxn, h.n);
      }
    }
  if ((rc = upload_chainp(c, -1, !fixed))) return rc;
  const int nb = c->D.nb, nc = (int)c->pool[P_CHAIN].cap;
  hipStream_t s = c->stream;
  if (noise) {
    // caller order: per step, active chains by id, 3N each -> device layout 3 * bead index
    if ((rc = ensure_noise(c, (size_t)nsteps * 3 * nb))) return rc;
    std::vector<double> lay((size_t)nsteps * 3 * nb, 0.0);
    size_t q = 0;
    for (long r = 0; r < nsteps; r++)
      for (auto& h : c->chains) {
        if (!h.active) continue;
        std::memcpy(&lay[(size_t)r * 3 * nb + 3 * h.at[P_BEAD]], noise + q, sizeof(double) * 3 * h.n);
        q += 3 * (size_t)h.n;
      }
    HIP_OK(c, hipMemcpyAsync(c->d_noise, lay.data(), sizeof(double) * lay.size(), hipMemcpyHostToDevice, s));
    HIP_OK(c, hipStreamSynchronize(s));
  }
  // the per-step launches below serve the chains the resident kernel does not
  bool stepped = false;
  for (auto& h : c->chains) stepped |= h.active && !h.resident;
  // the previous run's last step wrote the pair list at these positions into
  // the other parity, and nothing has changed since: start from it
  const bool next_pl = stepped && c->next_gen == c->gen;
  if (next_pl) c->par ^= 1;
  // The chains whose list rebuild (rec: record) falls on run step r, as the
  // bead and chain-slot ranges that cover them: a step's rebuild and record
  // launches cover only these, so chains on other rebuild phases (jobs that
  // joined at different steps) cost no blocks there.
  struct Span {
    bool any = false, cell = false;
    int g0 = 0, g1 = 0, c0 = 0, c1 = 0;
  };
  auto span_of = [&](long r, bool rec) {
    Span sp;
    for (int id = 0; id < (int)c->chains.size(); id++) {
      const HostChain& h = c->chains[id];
      if (!h.active || h.resident) continue;
      const bool hit = rec ? (h.iup + r) % h.p.nup == 0
                           : h.p.cutoff != SOPG_CUTOFF_NONE && (h.inlup + r) % h.p.nnlup == 0;
      if (!hit) continue;
      const int b0 = (int)h.at[P_BEAD], b1 = b0 + h.n;
      sp.g0 = sp.any ? std::min(sp.g0, b0) : b0;
      sp.g1 = sp.any ? std::max(sp.g1, b1) : b1;
      sp.c0 = sp.any ? std::min(sp.c0, id) : id;
      sp.c1 = sp.any ? std::max(sp.c1, id + 1) : id + 1;
      sp.any = true;
      sp.cell |= h.p.cutoff == SOPG_CUTOFF_CELL;
    }
    // The binned kernels follow each bead's chain plan, which k_grid writes
    // for the chain slots of its launch (the plan of a chain not due: no
    // work).  Slot and bead orders differ once chains have left and joined,
    // so the slot range grows to every chain with beads in the bead range.
    if (sp.any && !rec)
      for (int id = 0; id < (int)c->chains.size(); id++) {
        const HostChain& h = c->chains[id];
        if (!h.active || !h.on_device) continue;
        const int b0 = (int)h.at[P_BEAD];
        if (b0 < sp.g1 && b0 + h.n > sp.g0) {
          sp.c0 = std::min(sp.c0, id);
          sp.c1 = std::max(sp.c1, id + 1);
        }
      }
    return sp;
  };
  auto due = [&](long r) { return span_of(r, false); };
  // the run steps at which some chain records (bit 0), some chain on the
  // per-step path (bit 1): chain h records at the steps r with
  // (iup + r) % nup == 0, every nup steps from the first
  std::vector<uint8_t> rec_at((size_t)nsteps, 0);
  for (auto& h : c->chains)
    if (h.active)
      for (long r = (h.p.nup - h.iup % h.p.nup) % h.p.nup; r < nsteps; r += h.p.nup)
        rec_at[(size_t)r] |= h.resident ? 1 : 3;
  auto records = [&](long r) { return r < nsteps && (rec_at[(size_t)r] & 2) != 0; };
  // record buffer rows: one per run step at which any chain records
  std::vector<long> rec_steps;
  std::vector<int> row_of(nsteps, -1);
  for (long r = 0; r < nsteps; r++)
    if (rec_at[(size_t)r]) {
      row_of[r] = (int)rec_steps.size();
      rec_steps.push_back(r);
    }
  const long nrs = (long)rec_steps.size();
  if (fixed && (size_t)nrs * nc * kRed > kScratchDoubles) {
    char m[256];
    std::snprintf(m, sizeof m,
                  "sopg_run_noalloc: %ld record steps x %d chain slots x %d fields exceed the %zu-double scratch "
                  "array; run fewer steps per call",
                  nrs, nc, (int)kRed, kScratchDoubles);
    return fail(c, SOPG_ERR_ARG, m);
  }
  double* const rec_buf = fixed ? c->d_scratch : nullptr;
  if (!fixed && (size_t)nrs * nc * kRed > c->rec_cap_d) {
    HIP_OK(c, hipStreamSynchronize(s));
    dfree(c, c->d_rec);
    c->d_rec = nullptr;
    c->rec_cap_d = 0;
    if ((rc = dalloc(c, &c->d_rec, (size_t)nrs * nc * kRed))) return rc;
    c->rec_cap_d = (size_t)nrs * nc * kRed;
  }
  const int L = c->lanes;
  Dev& D = c->D;
  if (!rlist.empty()) {
    if (!c->stream2) {
      HIP_OK(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
      HIP_OK(c, hipEventCreateWithFlags(&c->ev_res, hipEventDisableTiming));
    }
    if (rlist.size() > c->rlist_cap) {
      HIP_OK(c, hipStreamSynchronize(c->stream2));
      dfree(c, c->d_rlist);
      c->d_rlist = nullptr;
      if ((rc = dalloc(c, &c->d_rlist, rlist.size()))) return rc;
      c->rlist_cap = rlist.size();
    }
    if ((size_t)nsteps > c->rslot_cap) {
      HIP_OK(c, hipStreamSynchronize(c->stream2));
      dfree(c, c->d_rslot);
      c->d_rslot = nullptr;
      if ((rc = dalloc(c, &c->d_rslot, (size_t)nsteps))) return rc;
      c->rslot_cap = (size_t)nsteps;
    }
    if ((rc = upload_at(c, c->d_rlist, 0, rlist)) || (rc = upload_at(c, c->d_rslot, 0, row_of))) return rc;
    int rl = c->resident_lanes;
    if (rl <= 0)
      for (rl = 8; rl > 1 && rmaxn * rl > 512; rl >>= 1) {}
    const int threads = (rmaxn * rl + 63) / 64 * 64;
    if (c->cp_busy) HIP_OK(c, hipStreamWaitEvent(c->stream2, c->ev_cp, 0));  // the chain records it reads
    launch_resident(rl, D, c->d_T, c->d_rlist, (int)rlist.size(), threads, c->cur, c->par, (int)nsteps, c->d_rslot,
                    c->d_rec, nc, c->stream2);
    HIP_OK(c, hipGetLastError());
    HIP_OK(c, hipEventRecord(c->ev_res, c->stream2));
  }
  // the step kernel's blocks: those holding a bead that it steps (no block
  // table in sopg_run_noalloc, which allocates nothing)
  const int* blk = nullptr;
  int nblk = 0;
  if (stepped && !fixed && (rc = step_blocks(c, L, false, &blk, &nblk))) return rc;
  // and the binned rebuild's tiles: a rebuild of bead range [g0, g1) takes the
  // table's run of tiles that cover it
  const int* tiles = nullptr;
  int ntiles = 0;
  if (stepped && !fixed && (rc = block_table(c, c->tile_tab, kBeadTile, false, &tiles, &ntiles))) return rc;
  struct TileRun {
    const int* p = nullptr;
    int n = 0;
  };
  auto tiles_of = [&](const Span& d) {
    TileRun tr;
    if (!tiles || !d.any) return tr;
    const std::vector<int>& t = c->tile_tab.dev;
    const long i0 = std::lower_bound(t.begin(), t.end(), d.g0 / kBeadTile) - t.begin();
    const long i1 = std::lower_bound(t.begin(), t.end(), (d.g1 + kBeadTile - 1) / kBeadTile) - t.begin();
    if (i1 > i0) {
      tr.p = tiles + i0;
      tr.n = (int)(i1 - i0);
    }
    return tr;
  };
  // step 1 of the run: rebuild if due, pair list, position half
  auto dr = due(0);
  int pe;
  if (!stepped) {
    // every chain is resident: no per-step launches; the parities advance as they would
    if ((nsteps - 1) & 1) {
      c->cur ^= 1;
      c->par ^= 1;
    }
  } else if (dr.any) {
    pe = prof_begin(c);
    if (dr.cell) launch_cellcoords(D, c->cur, dr.g0, dr.g1, s);
    const TileRun tr = tiles_of(dr);
    launch_rebuild(c->list_lanes, D, c->d_T, c->cur, c->par, 0, -1, dr.g0, dr.g1, dr.c0, dr.c1 - dr.c0, c->binned,
                   c->skin > 0.0, s, tr.p, tr.n);
    prof_end(c, SOPG_KERNEL_REBUILD, pe);
  }
  // pair lists at the current positions; rerunning the filter on a chain whose
  // list is fresh reproduces it, so one launch covers every chain
  bool fresh = true;
  if (!next_pl)
    for (auto& h : c->chains)
      if (h.active && !h.resident && !h.pl_fresh) fresh = false;
  if (!fresh) {
    pe = prof_begin(c);
    launch_pairfilter(L, D, c->d_T, c->cur, c->par, 0, nb, s);  // a few dozen entries per bead: step lanes
    prof_end(c, SOPG_KERNEL_PAIRLIST, pe);
  }
  // the position half of the run's first step: computed by the previous
  // run's last step into buffer cur^1 when nothing has changed since
  if (stepped) {
    if (next_pl && c->adv_ready && rlist.empty()) c->cur ^= 1;
    else launch_posupdate(D, c->cur, c->cur, 0, nb, s);
  }
  // The launches of run step r.  Kernels see step k + *Dx.relb: direct
  // launches pass k = r with base 0; a captured segment passes k relative to
  // its first step, whose run step the replay writes to the base.
  auto one_step = [&](const Dev& Dx, long r, int k) {
    const bool last = (r == nsteps - 1);
    const double* nz = noise ? c->d_noise + (size_t)r * 3 * nb : nullptr;
    int pe = prof_begin(c);
    // energy terms only in a step at which some chain records
    const int fe = records(r) ? F_ENERGY : 0;
    // every step computes the next step's pair list and position half (into
    // buffer cur^1); the last one leaves cur at rest, and its pair list and
    // position half are where the next run starts
    launch_force(L, Dx, c->d_T, c->cur, c->par, k, F_STORE | F_VEL | fe | F_ALLTERMS | F_ADVANCE, true, nz, 0, nb, s,
                 blk, nblk);
    prof_end(c, SOPG_KERNEL_STEP, pe);
    const Span rs = fe ? span_of(r, true) : Span();
    if (rs.any) {
      // chain slot c's sums land in row row_of[r], column c
      pe = prof_begin(c);
      launch_observe(Dx, c->cur, c->par, k, 0, OBS_PHYS, rs.g0, rs.g1, s);
      launch_chain_reduce(Dx, rs.c0, rs.c1 - rs.c0,
                          (fixed ? rec_buf : c->d_rec) + ((size_t)row_of[r] * nc + (size_t)rs.c0) * kRed, s);
      prof_end(c, SOPG_KERNEL_OBSERVE, pe);
    }
    if (!last) {
      c->par ^= 1;
      const Span d = due(r + 1);
      if (d.any) {
        pe = prof_begin(c);
        if (d.cell) launch_cellcoords(Dx, c->cur, d.g0, d.g1, s);
        const TileRun tr = tiles_of(d);
        launch_rebuild(c->list_lanes, Dx, c->d_T, c->cur, c->par, k + 1, -1, d.g0, d.g1, d.c0, d.c1 - d.c0, c->binned,
                       c->skin > 0.0, s, tr.p, tr.n);
        prof_end(c, SOPG_KERNEL_REBUILD, pe);
      }
      c->cur ^= 1;
    }
  };
  // Segments of S steps -- the rebuild period (the lcm of the chains' nnlup
  // when small), so that the launch pattern repeats -- are captured once into
  // a graph and replayed.  A segment holding a record step or the run's last
  // step, replayed noise or kernel timing runs as direct launches.
  long S = 1;
  for (auto& h : c->chains)
    if (h.active && !h.resident && h.p.cutoff != SOPG_CUTOFF_NONE) {
      long a = S, b = h.p.nnlup;
      while (b) { const long t = a % b; a = b; b = t; }
      S = S / a * h.p.nnlup;
      if (S > 64) break;
    }
  if (S > 64 || S < 8) S = S > 64 ? 32 : S * ((8 + S - 1) / S);
  const bool use_graphs = c->graphs && !noise && !c->prof && !fixed;
  Dev Dg;
  std::memcpy(&Dg, &D, sizeof Dg);  // bytewise: the struct is part of the graph key
  Dg.relb = c->d_relb + 1;
  long r = stepped ? 0 : nsteps;
  while (r < nsteps) {
    const long seg = std::min(S, nsteps - r);
    bool graphable = use_graphs && seg == S && r + S < nsteps;
    for (long k = 0; graphable && k < S; k++) graphable = !records(r + k);
    if (!graphable) {
      for (long k = 0; k < seg; k++) one_step(D, r + k, (int)(r + k));
      r += seg;
      continue;
    }
    // key: launch shape, parities, the segment's rebuild pattern, device layout
    std::vector<char> key;
    auto add = [&](const void* p, size_t n) { key.insert(key.end(), (const char*)p, (const char*)p + n); };
    const int hdr[8] = {(int)S, c->cur, c->par, L, c->list_lanes, c->binned, c->skin > 0.0 ? 1 : 0, nblk};
    add(hdr, sizeof hdr);
    add(&blk, sizeof blk);
    for (long k = 0; k < S; k++) {
      const Span d = due(r + k + 1);
      const TileRun tr = tiles_of(d);
      const int f[6] = {(d.any ? 1 : 0) | (d.cell ? 2 : 0), d.g0, d.g1, d.c0, d.c1, tr.n};
      add(f, sizeof f);
      add(&tr.p, sizeof tr.p);
    }
    add(&Dg, sizeof Dg);
    add(&c->d_T, sizeof c->d_T);
    hipGraphExec_t exec = nullptr;
    for (auto& ge : c->gcache)
      if (ge.key == key) { exec = ge.exec; ge.used = ++c->gclock; break; }
    if (exec) {
      // replaying advances the parities as the captured steps did
      if (S & 1) { c->cur ^= 1; c->par ^= 1; }
    } else {
      hipGraph_t graph = nullptr;
      HIP_OK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      for (long k = 0; k < S; k++) one_step(Dg, r + k, (int)k);
      HIP_OK(c, hipStreamEndCapture(s, &graph));
      hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(c, SOPG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      if (c->gcache.size() >= 16) {
        auto lru = std::min_element(c->gcache.begin(), c->gcache.end(),
                                    [](const sopg_ctx::GraphEntry& a, const sopg_ctx::GraphEntry& b) { return a.used < b.used; });
        HIP_OK(c, hipStreamSynchronize(s));
        (void)hipGraphExecDestroy(lru->exec);
        c->gcache.erase(lru);
      }
      c->gcache.push_back({key, exec, ++c->gclock});
      c->graph_captures++;
    }
    launch_set_base(c->d_relb + 1, (int)r, s);
    HIP_OK(c, hipGraphLaunch(exec, s));
    c->graph_launches++;
    r += S;
  }
  HIP_OK(c, hipGetLastError());
  if (!rlist.empty()) HIP_OK(c, hipStreamWaitEvent(s, c->ev_res, 0));
  for (auto& h : c->chains) {
    h.pl_fresh = false;
    h.resident = false;
  }
  c->next_gen = stepped ? c->gen : -1;
  c->adv_ready = stepped && rlist.empty();
  std::vector<double> hr(rec_steps.size() * nc * kRed);
  if (!hr.empty())
    HIP_OK(c, hipMemcpyAsync(hr.data(), fixed ? rec_buf : c->d_rec, sizeof(double) * hr.size(),
                             hipMemcpyDeviceToHost, s));
  HIP_OK(c, hipMemcpyAsync(c->err_pin, c->D.err, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(c, hipStreamSynchronize(s));
  prof_collect(c);
  if ((rc = err_verdict(c, *c->err_pin))) return rc;
  long nr = 0;
  for (size_t q = 0; q < rec_steps.size(); q++) {
    const long r = rec_steps[q];
    for (int ic = 0; ic < (int)c->chains.size(); ic++) {
      HostChain& h = c->chains[ic];
      if (!h.active || (h.iup + r) % h.p.nup != 0) continue;
      if (rec && nr < rec_cap) {
        const double* v = &hr[(q * nc + ic) * kRed];
        sopg_record& o = rec[nr];
        o.chain = ic;
        o.istep = h.istep + (double)r;
        o.T = h.p.T;
        o.kinT = kin_temp(h.p, v[R_V2], h.n);
        o.e_bnd = v[R_EBND] * -c->T.e_bnd_coeff;
        o.e_ang_ss = v[R_EANG];
        o.e_vdw_rr = v[R_EATT] + v[R_EREP];
        o.rna_etot = o.e_bnd + o.e_ang_ss + o.e_vdw_rr;
        o.contct_nat = (int)v[R_CONT];
        o.Q = (double)o.contct_nat / (double)h.ai.size();
        o.rgsq = v[R_RG] / ((double)h.n * (double)h.n);
      }
      nr++;
    }
  }
  if (nrec) *nrec = nr;
  for (auto& h : c->chains) {
    if (!h.active) continue;
    h.istep += (double)nsteps;
    h.inlup = (int)(((long)h.inlup - 1 + nsteps) % h.p.nnlup) + 1;
    h.iup = (int)(((long)h.iup - 1 + nsteps) % h.p.nup) + 1;
    h.noise_base += (uint64_t)nsteps;
  }
  return upload_chainp(c, -1, !fixed);
}

int sopg_run(sopg_ctx* c, long nsteps, const double* noise, sopg_record* rec, long rec_cap, long* nrec) {
  return run_impl(c, nsteps, noise, rec, rec_cap, nrec, false);
}

int sopg_run_noalloc(sopg_ctx* c, long nsteps, sopg_record* rec, long rec_cap, long* nrec) {
  return run_impl(c, nsteps, nullptr, rec, rec_cap, nrec, true);
}

}  // extern "C"
