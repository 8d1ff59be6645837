// sop_shim.cpp -- the reference's hot-path entry points, on the GPU.
//
// Each function keeps the signature and the effect on the global state of the
// SOP-CPU function it replaces (energy.cpp, neighbor_list.cpp, cell_list.cpp,
// sop.cpp:200-363, 473-553) and runs through the C ABI of include/sopg.h:
//   * the device context holds one chain built from the host topology;
//     it is rebuilt when the topology arrays change;
//   * entry points called one at a time synchronise eagerly: positions go up
//     before, results (forces, lists, counts, energies, observables) come back
//     after, so host code sees exactly the globals the reference would leave;
//   * underdamped_ctrl / overdamped_ctrl (sop_ctrl.cpp) keep the state on the
//     device between output steps instead.
// The random force consumes `generator` exactly as random_force does (3N
// gasdev() per force evaluation, bead-major x,y,z), so trajectories reproduce
// the reference's; `set noise device;` switches to device Philox noise.
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/sopg.h"
#include "sop_shim.h"
#include "sop_state.h"

int sopg_noise_device = 0;
int sopg_lanes = 0;
int sopg_restart_lossless = 0;

namespace {

sopg_ctx* g_ctx = nullptr;
struct Signature {
  int nbead = -1, nbnd, nang, ncon_att, ncon_rep, cutoff, sim_type, lanes;
  const void* arrays[6];
  bool operator==(const Signature& o) const {
    if (nbead != o.nbead || nbnd != o.nbnd || nang != o.nang || ncon_att != o.ncon_att || ncon_rep != o.ncon_rep ||
        cutoff != o.cutoff || sim_type != o.sim_type || lanes != o.lanes)
      return false;
    for (int k = 0; k < 6; k++)
      if (arrays[k] != o.arrays[k]) return false;
    return true;
  }
} g_sig;
std::vector<double> g_buf[4];
std::vector<double> g_noise;

int cutoff_mode() {
  if (neighborlist == 1) return SOPG_CUTOFF_NEIGHBOR;
  if (celllist == 1) return SOPG_CUTOFF_CELL;
  return SOPG_CUTOFF_NONE;
}

Signature current_signature() {
  Signature s;
  s.nbead = nbead; s.nbnd = nbnd; s.nang = nang; s.ncon_att = ncon_att; s.ncon_rep = ncon_rep;
  s.cutoff = cutoff_mode(); s.sim_type = sim_type; s.lanes = sopg_lanes;
  const void* a[6] = {ibead_bnd, ibead_ang, ibead_lj_nat, ibead_lj_non_nat, jbead_lj_nat, jbead_lj_non_nat};
  for (int k = 0; k < 6; k++) s.arrays[k] = a[k];
  return s;
}

sopg_chain_params current_params() {
  sopg_chain_params p{};
  p.T = T; p.zeta = zeta; p.h = h;
  p.a1 = a1; p.a2 = a2; p.a3 = a3; p.a4 = a4; p.a5 = a5;
  p.boxl = boxl; p.ncell = ncell; p.lcell = lcell;
  p.cutoff = cutoff_mode();
  p.nnlup = nnlup;
  p.nup = nup;
  p.noise_mode = sopg_noise_device ? SOPG_NOISE_PHILOX : SOPG_NOISE_REPLAY;
  p.seed = 0x50504750ull * 1000003ull + (unsigned long long)run;
  p.sim_type = sim_type;
  return p;
}

double* xyz(const coord* a) { return const_cast<double*>(&a[1].x); }

}  // namespace

void sopg_shim_die(const char* what) {
  std::fprintf(stderr, "sopg: %s failed: %s\n", what, g_ctx ? sopg_last_error(g_ctx) : "no device context");
  std::exit(-1);
}

void sopg_shim_check(int rc, const char* what) {
  if (rc != SOPG_OK) sopg_shim_die(what);
}

sopg_ctx* sopg_shim_context() {
  const Signature s = current_signature();
  if (g_ctx && s == g_sig) {
    const sopg_chain_params p = current_params();
    sopg_shim_check(sopg_set_chain_params(g_ctx, 0, &p), "sopg_set_chain_params");
    return g_ctx;
  }
  if (g_ctx) sopg_destroy(g_ctx);
  g_ctx = nullptr;
  if (sopg_create(0, &g_ctx) != SOPG_OK) {
    std::fprintf(stderr, "sopg: no usable GPU\n");
    std::exit(-1);
  }
  if (sopg_lanes > 0) sopg_shim_check(sopg_set_lanes(g_ctx, sopg_lanes, 0), "sopg_set_lanes");
  sopg_tables t;
  sopg_default_tables(&t);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      t.coeff_att[i][j] = coeff_att[i][j];
      t.coeff_rep[i][j] = coeff_rep[i][j];
      t.force_coeff_att[i][j] = force_coeff_att[i][j];
      t.force_coeff_rep[i][j] = force_coeff_rep[i][j];
      t.sigma_rep[i][j] = sigma_rep[i][j];
    }
  t.k_bnd = k_bnd;
  t.R0sq = R0sq;
  t.e_bnd_coeff = e_bnd_coeff;
  t.e_ang_ss_coeff = e_ang_ss_coeff;
  t.f_ang_ss_coeff = f_ang_ss_coeff;
  sopg_shim_check(sopg_set_tables(g_ctx, &t), "sopg_set_tables");
  sopg_topology topo{};
  topo.nbead = nbead;
  topo.nbnd = nbnd; topo.ibnd = ibead_bnd + 1; topo.jbnd = jbead_bnd + 1; topo.r0bnd = pdb_dist + 1;
  topo.nang = nang; topo.iang = ibead_ang + 1; topo.jang = jbead_ang + 1; topo.kang = kbead_ang + 1;
  topo.ncon_att = ncon_att;
  topo.att_i = ibead_lj_nat + 1; topo.att_j = jbead_lj_nat + 1;
  topo.att_it = itype_lj_nat + 1; topo.att_jt = jtype_lj_nat + 1; topo.att_r = lj_nat_pdb_dist + 1;
  topo.ncon_rep = ncon_rep;
  topo.rep_i = ibead_lj_non_nat + 1; topo.rep_j = jbead_lj_non_nat + 1;
  topo.rep_it = itype_lj_non_nat + 1; topo.rep_jt = jtype_lj_non_nat + 1;
  const sopg_chain_params p = current_params();
  int id = 0;
  sopg_shim_check(sopg_add_chain(g_ctx, &topo, &p, &id), "sopg_add_chain");
  g_sig = s;
  return g_ctx;
}

void sopg_shim_put_state() {
  sopg_shim_check(sopg_put_state(sopg_shim_context(), 0, xyz(pos), xyz(unc_pos), xyz(vel), xyz(force)),
                  "sopg_put_state");
}

void sopg_shim_get_state() {
  sopg_shim_check(sopg_get_state(g_ctx, 0, xyz(pos), xyz(unc_pos), xyz(vel), xyz(force)), "sopg_get_state");
}

const double* sopg_shim_noise(long nforce_evals) {
  if (sopg_noise_device) return nullptr;
  g_noise.resize((size_t)nforce_evals * 3 * nbead);
  for (double& g : g_noise) g = generator.gasdev();  // random_force's order (energy.cpp:456-462)
  return g_noise.data();
}

// Lists back into the reference's arrays.  Neighbor and pair lists are
// order-preserving filters of the master lists, so each entry's attributes
// (types, native distance powers) are found by walking the master list.
static void download_lists(bool nl) {
  sopg_ctx* c = g_ctx;
  const int which_att = nl ? SOPG_LIST_NL_ATT : SOPG_LIST_PL_ATT;
  const int which_rep = nl ? SOPG_LIST_NL_REP : SOPG_LIST_PL_REP;
  long n = 0;
  std::vector<int> li, lj;
  sopg_shim_check(sopg_get_list(c, 0, which_att, nullptr, nullptr, 0, &n), "sopg_get_list");
  li.resize(n + 1); lj.resize(n + 1);
  sopg_shim_check(sopg_get_list(c, 0, which_att, li.data(), lj.data(), n, &n), "sopg_get_list");
  int* ia = nl ? ibead_neighbor_list_att : ibead_pair_list_att;
  int* ja = nl ? jbead_neighbor_list_att : jbead_pair_list_att;
  int* ta = nl ? itype_neighbor_list_att : itype_pair_list_att;
  int* ua = nl ? jtype_neighbor_list_att : jtype_pair_list_att;
  double* d1 = nl ? nl_lj_nat_pdb_dist : pl_lj_nat_pdb_dist;
  double* d2 = nl ? nl_lj_nat_pdb_dist2 : pl_lj_nat_pdb_dist2;
  double* d6 = nl ? nl_lj_nat_pdb_dist6 : pl_lj_nat_pdb_dist6;
  double* d12 = nl ? nl_lj_nat_pdb_dist12 : pl_lj_nat_pdb_dist12;
  int m = 1;
  for (long k = 0; k < n; k++) {
    while (m <= ncon_att && !(ibead_lj_nat[m] == li[k] && jbead_lj_nat[m] == lj[k])) m++;
    if (m > ncon_att) sopg_shim_die("native list download (entry not in master order)");
    ia[k + 1] = li[k]; ja[k + 1] = lj[k]; ta[k + 1] = itype_lj_nat[m]; ua[k + 1] = jtype_lj_nat[m];
    d1[k + 1] = lj_nat_pdb_dist[m]; d2[k + 1] = lj_nat_pdb_dist2[m];
    d6[k + 1] = lj_nat_pdb_dist6[m]; d12[k + 1] = lj_nat_pdb_dist12[m];
    m++;
  }
  if (nl) nnl_att = (int)n; else nil_att = (int)n;
  sopg_shim_check(sopg_get_list(c, 0, which_rep, nullptr, nullptr, 0, &n), "sopg_get_list");
  li.resize(n + 1); lj.resize(n + 1);
  sopg_shim_check(sopg_get_list(c, 0, which_rep, li.data(), lj.data(), n, &n), "sopg_get_list");
  int* ir = nl ? ibead_neighbor_list_rep : ibead_pair_list_rep;
  int* jr = nl ? jbead_neighbor_list_rep : jbead_pair_list_rep;
  int* tr = nl ? itype_neighbor_list_rep : itype_pair_list_rep;
  int* ur = nl ? jtype_neighbor_list_rep : jtype_pair_list_rep;
  m = 1;
  for (long k = 0; k < n; k++) {
    while (m <= ncon_rep && !(ibead_lj_non_nat[m] == li[k] && jbead_lj_non_nat[m] == lj[k])) m++;
    if (m > ncon_rep) sopg_shim_die("non-native list download (entry not in master order)");
    ir[k + 1] = li[k]; jr[k + 1] = lj[k]; tr[k + 1] = itype_lj_non_nat[m]; ur[k + 1] = jtype_lj_non_nat[m];
    m++;
  }
  if (nl) nnl_rep = (int)n; else nil_rep = (int)n;
}

// ------------------------------------------------------------ list updates

void update_neighbor_list() {   // neighbor_list.cpp:6-117
  sopg_shim_put_state();
  sopg_shim_check(sopg_update_neighbor_list(g_ctx, 0, &nnl_att, &nnl_rep), "sopg_update_neighbor_list");
  download_lists(true);
}

void update_cell_list() {       // cell_list.cpp:6-98
  sopg_shim_put_state();
  sopg_shim_check(sopg_update_cell_list(g_ctx, 0, &nnl_att, &nnl_rep), "sopg_update_cell_list");
  download_lists(true);
}

void update_pair_list() {       // sop.cpp:473-553
  sopg_shim_put_state();
  sopg_shim_check(sopg_update_pair_list(g_ctx, 0, &nil_att, &nil_rep), "sopg_update_pair_list");
  download_lists(false);
}

// ------------------------------------------------------------ forces

// The term table (energy.cpp:78-113): entries point at the GPU term functions.
void set_forces() {
  int iterm = 0;
  for (int i = 1; i <= mforce_term; i++) {
    if (!force_term_on[i]) continue;
    if (i == 1) force_term[++iterm] = &random_force;
    else if (i == 2) force_term[++iterm] = &fene_forces;
    else if (i == 3) force_term[++iterm] = &soft_sphere_angular_forces;
    else if (i == 6) force_term[++iterm] = &vdw_forces;
  }
  nforce_term = iterm;
}

void set_potential() {
  int iterm = 0;
  for (int i = 1; i <= mpot_term; i++) {
    if (!pot_term_on[i]) continue;
    if (i == 1) pot_term[++iterm] = &fene_energy;
    else if (i == 2) pot_term[++iterm] = &soft_sphere_angular_energy;
    else if (i == 5) pot_term[++iterm] = &vdw_energy;
  }
  npot_term = iterm;
}

void clear_forces() {           // energy.cpp:35-45
  for (int i = 1; i <= nbead; i++) force[i].x = force[i].y = force[i].z = 0.0;
}

static void add_terms(int terms) {
  sopg_shim_put_state();
  const double* nz = (terms & SOPG_TERM_RANDOM) ? sopg_shim_noise(1) : nullptr;
  sopg_shim_check(sopg_force_terms(g_ctx, 0, terms, nz), "sopg_force_terms");
  sopg_shim_get_state();
}

void random_force() { add_terms(SOPG_TERM_RANDOM); }
void fene_forces() { add_terms(SOPG_TERM_FENE); }
void soft_sphere_angular_forces() { add_terms(SOPG_TERM_ANGULAR); }
void vdw_forces() { add_terms(SOPG_TERM_VDW); }

// force_eval (energy.cpp:21-33).  With the default table every term runs in
// one fused device pass; a modified table runs term by term.
static bool default_force_table() {
  return nforce_term == 4 && force_term[1] == &random_force && force_term[2] == &fene_forces &&
         force_term[3] == &soft_sphere_angular_forces && force_term[4] == &vdw_forces;
}

void force_eval() {
  if (!default_force_table()) {
    clear_forces();
    for (int i = 1; i <= nforce_term; i++) force_term[i]();
    return;
  }
  sopg_shim_put_state();
  sopg_shim_check(sopg_force_eval(g_ctx, 0, sopg_shim_noise(1)), "sopg_force_eval");
  sopg_shim_get_state();
}

// ------------------------------------------------------------ energies

static sopg_energies device_energies() {
  sopg_shim_put_state();
  sopg_energies e;
  sopg_shim_check(sopg_energy_eval(g_ctx, 0, &e), "sopg_energy_eval");
  return e;
}

void fene_energy() { e_bnd = device_energies().e_bnd; }
void soft_sphere_angular_energy() { e_ang_ss = device_energies().e_ang_ss; }
void vdw_energy() {
  const sopg_energies e = device_energies();
  e_vdw_rr_att = e.e_vdw_rr_att;
  e_vdw_rr_rep = e.e_vdw_rr_rep;
  e_vdw_rr = e.e_vdw_rr;
}

void energy_eval() {            // energy.cpp:6-19
  const sopg_energies e = device_energies();
  e_bnd = e.e_bnd;
  e_ang_ss = e.e_ang_ss;
  e_vdw_rr_att = e.e_vdw_rr_att;
  e_vdw_rr_rep = e.e_vdw_rr_rep;
  e_vdw_rr = e.e_vdw_rr;
  rna_etot = e_bnd + e_ang_ss + e_vdw_rr;
  system_etot = rna_etot;
}

// ------------------------------------------------------------ dynamics

static void iterate(coord* incr) {
  // incr as the device computes it (sop.cpp:292-294 / 339-341), for callers
  // that read the increment array
  for (int i = 1; i <= nbead; i++) {
    if (sim_type == 2) {
      incr[i].x = a5 * force[i].x; incr[i].y = a5 * force[i].y; incr[i].z = a5 * force[i].z;
    } else {
      incr[i].x = a1 * vel[i].x + a2 * force[i].x;
      incr[i].y = a1 * vel[i].y + a2 * force[i].y;
      incr[i].z = a1 * vel[i].z + a2 * force[i].z;
    }
  }
  sopg_shim_put_state();
  sopg_shim_check(sopg_underdamped_iteration(g_ctx, 0, sopg_shim_noise(1)), "sopg_underdamped_iteration");
  sopg_shim_get_state();
}

void underdamped_iteration(coord* incr) { iterate(incr); }   // sop.cpp:282-329
void overdamped_iteration(coord* incr) { iterate(incr); }    // sop.cpp:331-363

void calculate_observables(coord*) {   // sop.cpp:200-280
  sopg_shim_put_state();
  sopg_observables o;
  sopg_shim_check(sopg_calculate_observables(g_ctx, 0, &o), "sopg_calculate_observables");
  Q = o.Q;
  contct_nat = o.contct_nat;
  rgsq = o.rgsq;
  kinT = o.kinT;
}
