// sopg_api.cpp -- C ABI of the gfx950 SOP hot path (declared in include/sopg.h).
//
// Host side of a context: validates and copies topologies, builds the per-bead
// incidence lists (CSR, in the reference's list order), owns device memory,
// and sequences the kernels of sopg_kernels.hip on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
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
void launch_rebuild(int L, const Dev& D, const Tables* T, int cur, int par, int rel, int mode, int g0, int g1,
                    hipStream_t s);
void launch_pairfilter(int L, const Dev& D, const Tables* T, int cur, int par, int g0, int g1, hipStream_t s);
void launch_force(int L, const Dev& D, const Tables* T, int cur, int pr, int rel, int flags, bool next,
                  const double* noise, int g0, int g1, hipStream_t s);
void launch_posupdate(const Dev& D, int src, int dst, int g0, int g1, hipStream_t s);
void launch_velupdate(const Dev& D, int g0, int g1, hipStream_t s);
void launch_observe(const Dev& D, int cur, int par, int rel, int all, int what, int g0, int g1, hipStream_t s);
void launch_chain_reduce(const Dev& D, int c0, int nc, double* out, hipStream_t s);
enum { F_STORE = 1, F_VEL = 2, F_ENERGY = 4, F_ENERGY_ALL = 8, F_NOISE = 16, F_BOND = 32, F_ANGLE = 64,
       F_VDW = 128, F_ACCUM = 256, F_ALLTERMS = 16 | 32 | 64 | 128 };
enum { OBS_PHYS = 1, OBS_COUNTS = 2 };
}  // namespace sopg

using namespace sopg;

namespace {

struct HostChain {
  sopg_chain_params p{};
  int n = 0;
  std::vector<int> bi, bj;  // 0-based
  std::vector<double> br0;
  std::vector<int> gi, gk;
  std::vector<int> ai, aj, ait, ajt;
  std::vector<double> ar;
  bool rep_explicit = false;
  std::vector<int> ri, rj, rit, rjt;
  std::vector<int8_t> type;
  double istep = 1.0;
  int inlup = 1, iup = 1;
  uint64_t noise_base = 0;
};

}  // namespace

struct sopg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  sopg_tables tab{};
  Tables T{};
  int lanes = 16;       // lanes per bead, step / force kernels
  int list_lanes = 64;  // lanes per bead, list rebuild
  int cap_user = 0;
  Tables* d_T = nullptr;
  std::vector<HostChain> chains;
  bool committed = false;
  Dev D{};
  std::vector<ChainP> cp;
  std::vector<void*> allocs;
  int nb = 0;
  int cur = 0, par = 0;
  bool pl_fresh = false;
  double* d_out = nullptr;      // chain reduction output (nchains * kRed)
  double* d_noise = nullptr;    // replay noise
  size_t noise_cap = 0;         // doubles
  double* d_rec = nullptr;      // per-record-step reductions
  size_t rec_cap_d = 0;
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

// Event bracket around one launch when profiling; returns the pool index pair.
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
// After a stream synchronisation: fold the recorded brackets into the totals.
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

template <class T>
int dalloc(sopg_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return fail(c, SOPG_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->allocs.push_back(q);
  *p = (T*)q;
  return SOPG_OK;
}

template <class T>
int upload(sopg_ctx* c, T* d, const std::vector<T>& h) {
  if (h.empty()) return SOPG_OK;
  HIP_OK(c, hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return SOPG_OK;
}

void derive_tables(const sopg_tables& t, Tables& T) {
  double thr = 0.0;
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
      thr = std::max(thr, T.s2[i][j] * kTol);
    }
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

void fill_chainp(const HostChain& h, ChainP& P) {
  const sopg_chain_params& p = h.p;
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
  P.sim_type = p.sim_type == 2 ? 2 : 1;
  P.noise_base = h.noise_base;
}

// Counting sort of incidences by bead, stable in entry order: entry e adds
// (e<<1|0) to bead a[e] and then (e<<1|1) to bead b[e].
void build_csr(int nbtot, const std::vector<int>& a, const std::vector<int>& b, std::vector<int>& ptr,
               std::vector<int>& code) {
  ptr.assign((size_t)nbtot + 1, 0);
  for (size_t e = 0; e < a.size(); e++) { ptr[a[e] + 1]++; ptr[b[e] + 1]++; }
  for (int g = 0; g < nbtot; g++) ptr[g + 1] += ptr[g];
  code.assign(ptr[nbtot], 0);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (size_t e = 0; e < a.size(); e++) {
    code[fill[a[e]]++] = (int)(e << 1);
    code[fill[b[e]]++] = (int)(e << 1) | 1;
  }
}

bool chain_has_replay(const sopg_ctx* c) {
  for (auto& h : c->chains)
    if (h.p.noise_mode == SOPG_NOISE_REPLAY) return true;
  return false;
}

int upload_chainp(sopg_ctx* c) {
  for (size_t i = 0; i < c->chains.size(); i++) fill_chainp(c->chains[i], c->cp[i]);
  return upload(c, c->D.cp, c->cp);
}

int ensure_noise(sopg_ctx* c, size_t n) {
  if (n <= c->noise_cap) return SOPG_OK;
  if (c->d_noise) {
    (void)hipFree(c->d_noise);
    c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void*)c->d_noise));
  }
  c->d_noise = nullptr;
  c->noise_cap = 0;
  int rc = dalloc(c, &c->d_noise, n);
  if (rc) return rc;
  c->noise_cap = n;
  return SOPG_OK;
}

int commit(sopg_ctx* c) {
  if (c->committed) return SOPG_OK;
  if (c->chains.empty()) return fail(c, SOPG_ERR_ARG, "no chains");
  HIP_OK(c, hipSetDevice(c->device));
  Dev& D = c->D;
  // chain offsets
  int nb = 0, na = 0;

  int maxn = 0;
  bool any_cell = false, any_none = false, any_explicit = false;
  c->cp.assign(c->chains.size(), ChainP{});
  for (size_t ic = 0; ic < c->chains.size(); ic++) {
    HostChain& h = c->chains[ic];
    ChainP& P = c->cp[ic];
    P.off = nb;
    P.n = h.n;
    P.att_off = na;
    P.natt = (int)h.ai.size();
    P.rep_explicit = h.rep_explicit ? 1 : 0;
    double mx = 0.0;
    for (double r : h.ar) mx = std::max(mx, r * r);
    P.thr_att = kTol * mx;
    nb += h.n;
    na += (int)h.ai.size();



    maxn = std::max(maxn, h.n);
    any_cell |= h.p.cutoff == SOPG_CUTOFF_CELL;
    any_none |= h.p.cutoff == SOPG_CUTOFF_NONE;
    any_explicit |= h.rep_explicit;
  }
  c->nb = nb;
  int cap = c->cap_user;
  if (cap <= 0) {
    cap = std::min(maxn, any_cell ? 1024 : 512);
    if (any_none) cap = maxn;
  }
  cap = std::max(cap, 1);
  D.nb = nb;
  D.cap = cap;

  // concatenated host arrays
  std::vector<int> chain(nb);
  std::vector<int8_t> type(nb, 1);
  std::vector<int> bi, bj, gi, gk, ai, aj, ri, rj;
  std::vector<double> br0, r2, r6, r12, nlrc2, plrc2;
  std::vector<float> rf;
  std::vector<int8_t> at, rt;
  for (size_t ic = 0; ic < c->chains.size(); ic++) {
    HostChain& h = c->chains[ic];
    const int off = c->cp[ic].off;
    for (int b = 0; b < h.n; b++) { chain[off + b] = (int)ic; type[off + b] = h.type.empty() ? 1 : h.type[b]; }
    for (size_t k = 0; k < h.bi.size(); k++) { bi.push_back(off + h.bi[k]); bj.push_back(off + h.bj[k]); br0.push_back(h.br0[k]); }
    for (size_t k = 0; k < h.gi.size(); k++) { gi.push_back(off + h.gi[k]); gk.push_back(off + h.gk[k]); }
    for (size_t k = 0; k < h.ai.size(); k++) {
      ai.push_back(off + h.ai[k]);
      aj.push_back(off + h.aj[k]);
      at.push_back((int8_t)(h.ait[k] | (h.ajt[k] << 2)));
      const double r = h.ar[k];
      // params.cpp:203-208 and neighbor_list.cpp:44-47, sop.cpp:506-507
      const double rr2 = r * r;
      const double rr6 = rr2 * rr2 * rr2;
      r2.push_back(rr2);
      r6.push_back(rr6);
      r12.push_back(rr6 * rr6);
      const double rn = 3.2 * r, rp = 2.5 * r;
      nlrc2.push_back(rn * rn);
      plrc2.push_back(rp * rp);
      rf.push_back((float)r);
    }
    if (h.rep_explicit)
      for (size_t k = 0; k < h.ri.size(); k++) {
        ri.push_back(off + h.ri[k]);
        rj.push_back(off + h.rj[k]);
        rt.push_back((int8_t)(h.rit[k] | (h.rjt[k] << 2)));
      }
  }
  std::vector<int> bond_ptr, bond_code, ang_ptr, ang_code, att_ptr, att_code, rep_ptr, rep_code;
  build_csr(nb, bi, bj, bond_ptr, bond_code);
  build_csr(nb, gi, gk, ang_ptr, ang_code);
  build_csr(nb, ai, aj, att_ptr, att_code);
  std::vector<int> att_partner(att_code.size());
  for (size_t q = 0; q < att_code.size(); q++) {
    const int k = att_code[q] >> 1;
    att_partner[q] = (att_code[q] & 1) ? ai[k] : aj[k];
  }
  if (any_explicit) build_csr(nb, ri, rj, rep_ptr, rep_code);
  else rep_ptr.assign((size_t)nb + 1, 0);

  int rc = 0;
#define A(p, n) if ((rc = dalloc(c, &(p), (n)))) return rc
  for (int s = 0; s < 2; s++) {
    A(D.ux[s], nb); A(D.uy[s], nb); A(D.uz[s], nb);
    A(D.px[s], nb); A(D.py[s], nb); A(D.pz[s], nb);
  }
  A(D.vx, nb); A(D.vy, nb); A(D.vz, nb);
  A(D.fx, nb); A(D.fy, nb); A(D.fz, nb);
  A(D.ix, nb); A(D.iy, nb); A(D.iz, nb);
  A(D.chain, nb); A(D.type, nb);
  A(D.bond_ptr, nb + 1); A(D.bond_code, bond_code.size()); A(D.bi, bi.size()); A(D.bj, bj.size()); A(D.br0, br0.size());
  A(D.ang_ptr, nb + 1); A(D.ang_code, ang_code.size()); A(D.gi, gi.size()); A(D.gk, gk.size());
  A(D.att_ptr, nb + 1); A(D.att_code, att_code.size()); A(D.att_partner, att_partner.size());
  A(D.att_flag, att_code.size());
  A(D.am_i, ai.size()); A(D.am_j, aj.size()); A(D.am_t, at.size());
  A(D.am_r2, r2.size()); A(D.am_r6, r6.size()); A(D.am_r12, r12.size());
  A(D.am_nlrc2, nlrc2.size()); A(D.am_plrc2, plrc2.size()); A(D.am_rf, rf.size());
  A(D.am_nl, ai.size()); A(D.am_pl0, ai.size()); A(D.am_pl1, ai.size());
  A(D.rep_ptr, nb + 1);
  if (any_explicit) {
    A(D.rep_code, rep_code.size()); A(D.rm_i, ri.size()); A(D.rm_j, rj.size()); A(D.rm_t, rt.size());
    A(D.slot_k, (size_t)nb * cap);
  } else {
    D.rep_code = nullptr; D.rm_i = D.rm_j = nullptr; D.rm_t = nullptr; D.slot_k = nullptr;
  }
  A(D.slot, (size_t)nb * cap); A(D.cnt, nb);
  A(D.cx, nb); A(D.cy, nb); A(D.cz, nb);
  A(D.red, (size_t)nb * kRed);
  A(D.err, 1);
  A(D.cp, c->chains.size());
  A(c->d_T, 1);
  A(c->d_out, c->chains.size() * kRed);
#undef A
  hipStream_t s = c->stream;
  for (int k = 0; k < 2; k++) {
    HIP_OK(c, hipMemsetAsync(D.ux[k], 0, sizeof(double) * nb, s));
    HIP_OK(c, hipMemsetAsync(D.uy[k], 0, sizeof(double) * nb, s));
    HIP_OK(c, hipMemsetAsync(D.uz[k], 0, sizeof(double) * nb, s));
    HIP_OK(c, hipMemsetAsync(D.px[k], 0, sizeof(double) * nb, s));
    HIP_OK(c, hipMemsetAsync(D.py[k], 0, sizeof(double) * nb, s));
    HIP_OK(c, hipMemsetAsync(D.pz[k], 0, sizeof(double) * nb, s));
  }
  for (double* p : {D.vx, D.vy, D.vz, D.fx, D.fy, D.fz, D.ix, D.iy, D.iz})
    HIP_OK(c, hipMemsetAsync(p, 0, sizeof(double) * nb, s));
  HIP_OK(c, hipMemsetAsync(D.red, 0, sizeof(double) * nb * kRed, s));
  HIP_OK(c, hipMemsetAsync(D.cnt, 0, sizeof(int) * nb, s));
  HIP_OK(c, hipMemsetAsync(D.att_flag, 0, att_code.size() ? att_code.size() : 1, s));
  HIP_OK(c, hipMemsetAsync(D.am_nl, 0, ai.size() ? ai.size() : 1, s));
  HIP_OK(c, hipMemsetAsync(D.am_pl0, 0, ai.size() ? ai.size() : 1, s));
  HIP_OK(c, hipMemsetAsync(D.am_pl1, 0, ai.size() ? ai.size() : 1, s));
  HIP_OK(c, hipMemsetAsync(D.err, 0, sizeof(int), s));
  if ((rc = upload(c, D.chain, chain)) || (rc = upload(c, D.type, type)) || (rc = upload(c, D.bond_ptr, bond_ptr)) ||
      (rc = upload(c, D.bond_code, bond_code)) || (rc = upload(c, D.bi, bi)) || (rc = upload(c, D.bj, bj)) ||
      (rc = upload(c, D.br0, br0)) || (rc = upload(c, D.ang_ptr, ang_ptr)) || (rc = upload(c, D.ang_code, ang_code)) ||
      (rc = upload(c, D.gi, gi)) || (rc = upload(c, D.gk, gk)) || (rc = upload(c, D.att_ptr, att_ptr)) ||
      (rc = upload(c, D.att_code, att_code)) || (rc = upload(c, D.att_partner, att_partner)) ||
      (rc = upload(c, D.am_i, ai)) || (rc = upload(c, D.am_j, aj)) || (rc = upload(c, D.am_t, at)) ||
      (rc = upload(c, D.am_r2, r2)) || (rc = upload(c, D.am_r6, r6)) || (rc = upload(c, D.am_r12, r12)) ||
      (rc = upload(c, D.am_nlrc2, nlrc2)) || (rc = upload(c, D.am_plrc2, plrc2)) || (rc = upload(c, D.am_rf, rf)) ||
      (rc = upload(c, D.rep_ptr, rep_ptr)))
    return rc;
  if (any_explicit &&
      ((rc = upload(c, D.rep_code, rep_code)) || (rc = upload(c, D.rm_i, ri)) || (rc = upload(c, D.rm_j, rj)) ||
       (rc = upload(c, D.rm_t, rt))))
    return rc;
  if ((rc = upload_chainp(c))) return rc;
  HIP_OK(c, hipMemcpyAsync(c->d_T, &c->T, sizeof(Tables), hipMemcpyHostToDevice, s));
  HIP_OK(c, hipStreamSynchronize(s));
  c->committed = true;
  // No-cutoff chains interact through their full master lists, as the loader
  // leaves them (params.cpp:210-230): build that list once.
  if (any_none) {
    for (size_t ic = 0; ic < c->chains.size(); ic++)
      if (c->chains[ic].p.cutoff == SOPG_CUTOFF_NONE)
        launch_rebuild(c->list_lanes, D, c->d_T, c->cur, c->par, 0, SOPG_CUTOFF_NONE, c->cp[ic].off,
                       c->cp[ic].off + c->cp[ic].n, s);
    HIP_OK(c, hipGetLastError());
  }
  return SOPG_OK;
}

int get_chain(sopg_ctx* c, int chain, HostChain** h, ChainP** P) {
  if (!c) return SOPG_ERR_ARG;
  if (chain < 0 || chain >= (int)c->chains.size()) return fail(c, SOPG_ERR_ARG, "bad chain id");
  int rc = commit(c);
  if (rc) return rc;
  *h = &c->chains[chain];
  *P = &c->cp[chain];
  return SOPG_OK;
}

int check_err_flag(sopg_ctx* c) {
  int e = 0;
  HIP_OK(c, hipMemcpyAsync(&e, c->D.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  if (e & 1) {
    HIP_OK(c, hipMemsetAsync(c->D.err, 0, sizeof(int), c->stream));
    return fail(c, SOPG_ERR_OVERFLOW,
                "a per-bead non-native neighbor list exceeded its capacity; raise it with sopg_set_capacity");
  }
  return SOPG_OK;
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
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if (mode == SOPG_CUTOFF_CELL && (!(h->p.ncell > 0.0) || !(h->p.lcell > 0.0)))
    return fail(c, SOPG_ERR_ARG, "cell list needs ncell > 0 and lcell > 0");
  if (mode == SOPG_CUTOFF_CELL) launch_cellcoords(c->D, c->cur, P->off, P->off + P->n, c->stream);
  launch_rebuild(c->list_lanes, c->D, c->d_T, c->cur, c->par, 0, mode, P->off, P->off + P->n, c->stream);
  HIP_OK(c, hipGetLastError());
  if ((rc = check_err_flag(c))) return rc;
  c->pl_fresh = true;
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
    if ((*rc = ensure_noise(c, 3 * (size_t)c->nb))) return nullptr;
    hipError_t e = hipMemcpyAsync(c->d_noise + 3 * (size_t)P.off, noise, sizeof(double) * 3 * (size_t)P.n,
                                  hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) { *rc = fail(c, SOPG_ERR_HIP, hipGetErrorString(e)); return nullptr; }
    return c->d_noise;
  }
  if (noise) { *rc = fail(c, SOPG_ERR_ARG, "chain uses the device generator: pass noise = NULL"); return nullptr; }
  return nullptr;
}

}  // namespace

// =================================================================== C ABI

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
  derive_tables(c->tab, c->T);
  *out = c;
  return SOPG_OK;
}

int sopg_destroy(sopg_ctx* c) {
  if (!c) return SOPG_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
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
  if (!c || !t) return SOPG_ERR_ARG;
  c->tab = *t;
  derive_tables(c->tab, c->T);
  if (c->committed) {
    HIP_OK(c, hipMemcpyAsync(c->d_T, &c->T, sizeof(Tables), hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
  }
  return SOPG_OK;
}

int sopg_set_lanes(sopg_ctx* c, int step_lanes, int list_lanes) {
  if (!c) return SOPG_ERR_ARG;
  auto ok = [](int l) { return l >= 1 && l <= 64 && !(l & (l - 1)); };
  if (step_lanes == 0) step_lanes = 16;
  if (list_lanes == 0) list_lanes = 64;
  if (!ok(step_lanes) || !ok(list_lanes)) return fail(c, SOPG_ERR_ARG, "lanes: power of two in 1..64");
  c->lanes = step_lanes;
  c->list_lanes = list_lanes;
  return SOPG_OK;
}

int sopg_set_capacity(sopg_ctx* c, int cap) {
  if (!c) return SOPG_ERR_ARG;
  if (c->committed) return fail(c, SOPG_ERR_ARG, "capacity must be set before the first device operation");
  c->cap_user = cap;
  return SOPG_OK;
}

int sopg_add_chain(sopg_ctx* c, const sopg_topology* t, const sopg_chain_params* p, int* id) {
  if (!c || !t) return SOPG_ERR_ARG;
  if (c->committed) return fail(c, SOPG_ERR_ARG, "chains must be added before the first device operation");
  int rc = check_params(c, p);
  if (rc) return rc;
  const int n = t->nbead;
  if (n < 1 || n >= kMaxChainBeads) return fail(c, SOPG_ERR_ARG, "nbead out of range");
  HostChain h;
  h.p = *p;
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
  } else {
    for (int k = 0; k < t->ncon_rep; k++) {
      if (!bead(t->rep_i[k]) || !bead(t->rep_j[k]) || !typ(t->rep_it[k]) || !typ(t->rep_jt[k]))
        return fail(c, SOPG_ERR_ARG, "non-native contact index/type out of range");
    }
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
  }
  c->chains.push_back(std::move(h));
  if (id) *id = (int)c->chains.size() - 1;
  return SOPG_OK;
}

int sopg_set_chain_params(sopg_ctx* c, int chain, const sopg_chain_params* p) {
  if (!c) return SOPG_ERR_ARG;
  if (chain < 0 || chain >= (int)c->chains.size()) return fail(c, SOPG_ERR_ARG, "bad chain id");
  int rc = check_params(c, p);
  if (rc) return rc;
  c->chains[chain].p = *p;
  if (c->committed) return upload_chainp(c);
  return SOPG_OK;
}

int sopg_nchains(const sopg_ctx* c) { return c ? (int)c->chains.size() : 0; }
int sopg_chain_nbead(const sopg_ctx* c, int chain) {
  return (c && chain >= 0 && chain < (int)c->chains.size()) ? c->chains[chain].n : -1;
}

int sopg_put_state(sopg_ctx* c, int chain, const double* pos, const double* unc, const double* vel,
                   const double* force) {
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
  if ((rc = put(pos, D.px[c->cur], D.py[c->cur], D.pz[c->cur])) || (rc = put(unc, D.ux[c->cur], D.uy[c->cur], D.uz[c->cur])) ||
      (rc = put(vel, D.vx, D.vy, D.vz)) || (rc = put(force, D.fx, D.fy, D.fz)))
    return rc;
  if (unc) c->pl_fresh = false;
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
  if ((rc = get(pos, D.px[c->cur], D.py[c->cur], D.pz[c->cur])) || (rc = get(unc, D.ux[c->cur], D.uy[c->cur], D.uz[c->cur])) ||
      (rc = get(vel, D.vx, D.vy, D.vz)) || (rc = get(force, D.fx, D.fy, D.fz)))
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
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  launch_pairfilter(c->list_lanes, c->D, c->d_T, c->cur, c->par, P->off, P->off + P->n, c->stream);
  HIP_OK(c, hipGetLastError());
  c->pl_fresh = true;
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
  if ((rc = upload_chainp(c))) return rc;
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
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if (terms & ~(SOPG_TERM_RANDOM | SOPG_TERM_FENE | SOPG_TERM_ANGULAR | SOPG_TERM_VDW))
    return fail(c, SOPG_ERR_ARG, "unknown force term");
  if ((rc = upload_chainp(c))) return rc;
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
  if ((rc = upload_chainp(c))) return rc;
  launch_force(c->lanes, c->D, c->d_T, c->cur, c->par, 0, F_ENERGY_ALL | F_ALLTERMS, false, nullptr, P->off,
               P->off + P->n,
               c->stream);
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
  HostChain* h;
  ChainP* P;
  int rc = get_chain(c, chain, &h, &P);
  if (rc) return rc;
  if ((rc = upload_chainp(c))) return rc;
  const double* dn = chain_noise(c, *P, *h, noise, &rc);
  if (rc) return rc;
  launch_posupdate(c->D, c->cur, c->cur, P->off, P->off + P->n, c->stream);
  launch_force(c->lanes, c->D, c->d_T, c->cur, c->par, 0, F_STORE | F_VEL | F_ALLTERMS, false, dn, P->off,
               P->off + P->n,
               c->stream);
  HIP_OK(c, hipGetLastError());
  HIP_OK(c, hipStreamSynchronize(c->stream));
  h->noise_base++;
  c->pl_fresh = false;
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
    std::vector<uint8_t> f(P->natt);
    const uint8_t* src = (which == SOPG_LIST_NL_ATT) ? c->D.am_nl : (c->par ? c->D.am_pl1 : c->D.am_pl0);
    if (P->natt)
      HIP_OK(c, hipMemcpyAsync(f.data(), src + P->att_off, P->natt, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < P->natt; k++)
      if (f[k]) L.push_back({h->ai[k] + 1, h->aj[k] + 1});
  } else {
    const int n = P->n, cap_s = c->D.cap;
    std::vector<int> cnt(n);
    std::vector<uint32_t> sl((size_t)n * cap_s);
    std::vector<int> sk;
    HIP_OK(c, hipMemcpyAsync(cnt.data(), c->D.cnt + P->off, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipMemcpyAsync(sl.data(), c->D.slot + (size_t)P->off * cap_s, sizeof(uint32_t) * n * cap_s,
                             hipMemcpyDeviceToHost, c->stream));
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
  if (chain < 0 || chain >= (int)c->chains.size()) return fail(c, SOPG_ERR_ARG, "bad chain id");
  HostChain& h = c->chains[chain];
  // inlup = 0 (overdamped_ctrl starts from the global inlup = 0, sop.cpp:428)
  // rebuilds on the first step, as inlup = nnlup does
  if (inlup == 0) inlup = h.p.nnlup;
  if (inlup < 1 || inlup > h.p.nnlup || iup < 1 || iup > h.p.nup)
    return fail(c, SOPG_ERR_ARG, "inlup must be in [0, nnlup] and iup in [1, nup]");
  h.istep = istep;
  h.inlup = inlup;
  h.iup = iup;
  return SOPG_OK;
}

int sopg_sync(sopg_ctx* c) {
  if (!c) return SOPG_ERR_ARG;
  if (c->stream) HIP_OK(c, hipStreamSynchronize(c->stream));
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

int sopg_run(sopg_ctx* c, long nsteps, const double* noise, sopg_record* rec, long rec_cap, long* nrec) {
  if (!c) return SOPG_ERR_ARG;
  if (nrec) *nrec = 0;
  if (nsteps < 0) return fail(c, SOPG_ERR_ARG, "nsteps < 0");
  int rc = commit(c);
  if (rc) return rc;
  if (nsteps == 0) return SOPG_OK;
  const bool replay = chain_has_replay(c);
  if (replay && !noise) return fail(c, SOPG_ERR_ARG, "replayed-noise chains need the run's noise");
  if (!replay && noise) return fail(c, SOPG_ERR_ARG, "all chains use the device generator: noise must be NULL");
  if ((rc = upload_chainp(c))) return rc;
  const int nb = c->nb, nc = (int)c->chains.size();
  hipStream_t s = c->stream;
  if (noise) {
    if ((rc = ensure_noise(c, (size_t)nsteps * 3 * nb))) return rc;
    HIP_OK(c, hipMemcpyAsync(c->d_noise, noise, sizeof(double) * (size_t)nsteps * 3 * nb, hipMemcpyHostToDevice, s));
  }
  auto due = [&](long r) {
    bool any = false, cell = false;
    for (auto& h : c->chains)
      if (h.p.cutoff != SOPG_CUTOFF_NONE && (h.inlup + r) % h.p.nnlup == 0) { any = true; cell |= h.p.cutoff == SOPG_CUTOFF_CELL; }
    return std::make_pair(any, cell);
  };
  auto records = [&](long r) {
    for (auto& h : c->chains)
      if ((h.iup + r) % h.p.nup == 0) return true;
    return false;
  };
  long nrs = 0;
  for (long r = 0; r < nsteps; r++) nrs += records(r);
  if ((size_t)nrs * nc * kRed > c->rec_cap_d) {
    if (c->d_rec) {
      (void)hipFree(c->d_rec);
      c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void*)c->d_rec));
      c->d_rec = nullptr;
    }
    c->rec_cap_d = 0;
    if ((rc = dalloc(c, &c->d_rec, (size_t)nrs * nc * kRed))) return rc;
    c->rec_cap_d = (size_t)nrs * nc * kRed;
  }
  const int L = c->lanes;
  Dev& D = c->D;
  // step 1 of the run: rebuild if due, pair list, position half
  auto dr = due(0);
  int pe;
  if (dr.first) {
    pe = prof_begin(c);
    if (dr.second) launch_cellcoords(D, c->cur, 0, nb, s);
    launch_rebuild(c->list_lanes, D, c->d_T, c->cur, c->par, 0, -1, 0, nb, s);
    prof_end(c, SOPG_KERNEL_REBUILD, pe);
  }
  if (!c->pl_fresh) {
    pe = prof_begin(c);
    launch_pairfilter(c->list_lanes, D, c->d_T, c->cur, c->par, 0, nb, s);
    prof_end(c, SOPG_KERNEL_PAIRLIST, pe);
  }
  launch_posupdate(D, c->cur, c->cur, 0, nb, s);
  std::vector<long> rec_steps;
  for (long r = 0; r < nsteps; r++) {
    const bool last = (r == nsteps - 1);
    const double* nz = noise ? c->d_noise + (size_t)r * 3 * nb : nullptr;
    pe = prof_begin(c);
    launch_force(L, D, c->d_T, c->cur, c->par, (int)r, F_STORE | F_VEL | F_ENERGY | F_ALLTERMS, !last, nz, 0, nb, s);
    prof_end(c, SOPG_KERNEL_STEP, pe);
    if (records(r)) {
      pe = prof_begin(c);
      launch_observe(D, c->cur, c->par, (int)r, 0, OBS_PHYS, 0, nb, s);
      launch_chain_reduce(D, 0, nc, c->d_rec + rec_steps.size() * nc * kRed, s);
      prof_end(c, SOPG_KERNEL_OBSERVE, pe);
      rec_steps.push_back(r);
    }
    if (!last) {
      c->par ^= 1;
      dr = due(r + 1);
      if (dr.first) {
        pe = prof_begin(c);
        if (dr.second) launch_cellcoords(D, c->cur, 0, nb, s);
        launch_rebuild(c->list_lanes, D, c->d_T, c->cur, c->par, (int)(r + 1), -1, 0, nb, s);
        prof_end(c, SOPG_KERNEL_REBUILD, pe);
      }
      c->cur ^= 1;
    }
  }
  HIP_OK(c, hipGetLastError());
  c->pl_fresh = false;
  std::vector<double> hr(rec_steps.size() * nc * kRed);
  if (!hr.empty())
    HIP_OK(c, hipMemcpyAsync(hr.data(), c->d_rec, sizeof(double) * hr.size(), hipMemcpyDeviceToHost, s));
  HIP_OK(c, hipStreamSynchronize(s));
  prof_collect(c);
  if ((rc = check_err_flag(c))) return rc;
  long nr = 0;
  for (size_t q = 0; q < rec_steps.size(); q++) {
    const long r = rec_steps[q];
    for (int ic = 0; ic < nc; ic++) {
      HostChain& h = c->chains[ic];
      if ((h.iup + r) % h.p.nup != 0) continue;
      if (rec && nr < rec_cap) {
        const double* v = &hr[(q * nc + ic) * kRed];
        sopg_record& o = rec[nr];
        const ChainP& P = c->cp[ic];
        o.chain = ic;
        o.istep = h.istep + (double)r;
        o.T = h.p.T;
        o.kinT = kin_temp(h.p, v[R_V2], P.n);
        o.e_bnd = v[R_EBND] * -c->T.e_bnd_coeff;
        o.e_ang_ss = v[R_EANG];
        o.e_vdw_rr = v[R_EATT] + v[R_EREP];
        o.rna_etot = o.e_bnd + o.e_ang_ss + o.e_vdw_rr;
        o.contct_nat = (int)v[R_CONT];
        o.Q = (double)o.contct_nat / (double)P.natt;
        o.rgsq = v[R_RG] / ((double)P.n * (double)P.n);
      }
      nr++;
    }
  }
  if (nrec) *nrec = nr;
  for (auto& h : c->chains) {
    h.istep += (double)nsteps;
    h.inlup = (int)(((long)h.inlup - 1 + nsteps) % h.p.nnlup) + 1;
    h.iup = (int)(((long)h.iup - 1 + nsteps) % h.p.nup) + 1;
    h.noise_base += (uint64_t)nsteps;
  }
  return upload_chainp(c);
}

}  // extern "C"
