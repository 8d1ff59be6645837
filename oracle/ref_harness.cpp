// Test-infrastructure harness around the compiled reference (SOP-CPU).
//
// This file is ORACLE code: it is linked with the reference's own sources,
// compiled where they lie (see oracle/Makefile), into oracle/_ref/libsopref.so.
// Only tests/, the golden-vector script and bench.py's cpu_baseline leg load it,
// and only as a checker or as the timed CPU baseline.  Nothing in the product
// links it.
//
// It exposes the reference's hot-path entry points and global state through a
// flat C ABI so that Python (ctypes) can drive them:
//   - commands are fed to the reference's own `load(icmd)` / `set_params(icmd)`
//     (params.cpp:12-101, 110-260) by filling its `cmd`/`opt`/`opt_ptr` tables;
//   - lists, coordinates, forces, energies and observables are read from the
//     globals declared in global.h.
#include <cstring>
#include <cstdlib>
#include <cstdio>
#include "global.h"
#include "energy.h"
#include "neighbor_list.h"
#include "cell_list.h"
#include "sop.h"
#include "params.h"
#include "io.h"

static coord* g_incr = 0;
static int g_incr_n = -1;

// The step scratch array the reference's ctrl loop owns (sop.cpp:104).
static coord* incr_array() {
  if (g_incr_n != nbead) {
    delete[] g_incr;
    g_incr = new coord[nbead + 1];
    g_incr_n = nbead;
  }
  return g_incr;
}

static void fill_command(const char* verb, const char* key, const char* val) {
  // One command at slot 1: cmd[1] = verb, options opt[1], opt[2].
  std::strncpy(cmd[1], verb, mwdsize - 1);
  opt_ptr[1] = 1;
  std::strncpy(opt[1], key, mwdsize - 1);
  std::strncpy(opt[2], val ? val : "", mwdsize - 1);
  opt_ptr[2] = 3;
  ncmd = 1;
}

extern "C" {

void ref_init() { alloc_arrays(); }

// "set <key> <val>" through the reference's set_params (params.cpp:12).
void ref_set(const char* key, const char* val) {
  fill_command("set", key, val);
  set_params(1);
}

// "load <kind> <path>" through the reference's load (params.cpp:110).
void ref_load(const char* kind, const char* path) {
  fill_command("load", kind, path);
  load(1);
}

// Full program path: read_input (io.cpp:9) + ex_cmds (sop.cpp:55).
void ref_read_input(const char* path) { read_input(path); }
void ref_ex_cmds() { ex_cmds(); }

// ex_cmds (sop.cpp:55-74) without `run`: the load / set commands of a script.
void ref_ex_cmds_norun() {
  for (int i = 1; i <= ncmd; i++) {
    if (!std::strcmp(cmd[i], "load")) load(i);
    else if (!std::strcmp(cmd[i], "set")) set_params(i);
  }
}
int ref_ncmd() { return ncmd; }

void ref_set_potential_forces() { set_potential(); set_forces(); }
void ref_update_neighbor_list() { update_neighbor_list(); }
void ref_update_cell_list() { update_cell_list(); }
void ref_update_pair_list() { update_pair_list(); }
void ref_force_eval() { force_eval(); }
void ref_clear_forces() { clear_forces(); }
void ref_random_force() { random_force(); }
void ref_fene_forces() { fene_forces(); }
void ref_angular_forces() { soft_sphere_angular_forces(); }
void ref_vdw_forces() { vdw_forces(); }
void ref_energy_eval() { energy_eval(); }

void ref_underdamped_iteration() { underdamped_iteration(incr_array()); }
void ref_overdamped_iteration() { overdamped_iteration(incr_array()); }
void ref_calculate_observables() { calculate_observables(incr_array()); }

// Reference Ran_Gen (random_generator.cpp:88-146) on the global generator.
void ref_set_seed(int s) { generator.set_seed(s); }
// Full generator state from a state file (random_generator.cpp:51-70).
void ref_rng_restart(const char* path) {
  char buf[2048];
  std::strncpy(buf, path, sizeof buf - 1);
  buf[sizeof buf - 1] = 0;
  generator.set_fname(buf);
  generator.restart();
}
double ref_ran2() { return generator.ran2(); }
double ref_gasdev() { return generator.gasdev(); }
void ref_gasdev_n(double* out, long n) {
  for (long i = 0; i < n; i++) out[i] = generator.gasdev();
}

int ref_get_int(const char* name) {
  if (!std::strcmp(name, "nbead")) return nbead;
  if (!std::strcmp(name, "nbnd")) return nbnd;
  if (!std::strcmp(name, "nang")) return nang;
  if (!std::strcmp(name, "ncon_att")) return ncon_att;
  if (!std::strcmp(name, "ncon_rep")) return ncon_rep;
  if (!std::strcmp(name, "nnl_att")) return nnl_att;
  if (!std::strcmp(name, "nnl_rep")) return nnl_rep;
  if (!std::strcmp(name, "nil_att")) return nil_att;
  if (!std::strcmp(name, "nil_rep")) return nil_rep;
  if (!std::strcmp(name, "contct_nat")) return contct_nat;
  if (!std::strcmp(name, "nup")) return nup;
  if (!std::strcmp(name, "nnlup")) return nnlup;
  if (!std::strcmp(name, "sim_type")) return sim_type;
  if (!std::strcmp(name, "neighborlist")) return neighborlist;
  if (!std::strcmp(name, "celllist")) return celllist;
  return -999999;
}

double ref_get_double(const char* name) {
  if (!std::strcmp(name, "e_bnd")) return e_bnd;
  if (!std::strcmp(name, "e_ang_ss")) return e_ang_ss;
  if (!std::strcmp(name, "e_vdw_rr")) return e_vdw_rr;
  if (!std::strcmp(name, "e_vdw_rr_att")) return e_vdw_rr_att;
  if (!std::strcmp(name, "e_vdw_rr_rep")) return e_vdw_rr_rep;
  if (!std::strcmp(name, "rna_etot")) return rna_etot;
  if (!std::strcmp(name, "Q")) return Q;
  if (!std::strcmp(name, "rgsq")) return rgsq;
  if (!std::strcmp(name, "kinT")) return kinT;
  if (!std::strcmp(name, "T")) return T;
  if (!std::strcmp(name, "h")) return h;
  if (!std::strcmp(name, "halfh")) return halfh;
  if (!std::strcmp(name, "zeta")) return zeta;
  if (!std::strcmp(name, "a1")) return a1;
  if (!std::strcmp(name, "a2")) return a2;
  if (!std::strcmp(name, "a3")) return a3;
  if (!std::strcmp(name, "a4")) return a4;
  if (!std::strcmp(name, "a5")) return a5;
  if (!std::strcmp(name, "boxl")) return boxl;
  if (!std::strcmp(name, "ncell")) return ncell;
  if (!std::strcmp(name, "lcell")) return lcell;
  if (!std::strcmp(name, "nstep")) return nstep;
  if (!std::strcmp(name, "e_ang_ss_coeff")) return e_ang_ss_coeff;
  if (!std::strcmp(name, "f_ang_ss_coeff")) return f_ang_ss_coeff;
  return -999999.0;
}

// which: 0 pos, 1 unc_pos, 2 vel, 3 force, 4 the ctrl loop's incr (sop.cpp:104).
// Layout out[3*(i-1)+c], beads 1..nbead.
static coord* coord_array(int which) {
  switch (which) {
    case 0: return pos;
    case 1: return unc_pos;
    case 2: return vel;
    case 3: return force;
    case 4: return incr_array();
    default: return 0;
  }
}
void ref_get_coords(int which, double* out) {
  coord* a = coord_array(which);
  for (int i = 1; i <= nbead; i++) {
    out[3 * (i - 1) + 0] = a[i].x;
    out[3 * (i - 1) + 1] = a[i].y;
    out[3 * (i - 1) + 2] = a[i].z;
  }
}
void ref_set_coords(int which, const double* in) {
  coord* a = coord_array(which);
  for (int i = 1; i <= nbead; i++) {
    a[i].x = in[3 * (i - 1) + 0];
    a[i].y = in[3 * (i - 1) + 1];
    a[i].z = in[3 * (i - 1) + 2];
  }
}

// which: 0 NL att, 1 NL rep, 2 PL att, 3 PL rep, 4 master att, 5 master rep.
// Writes entries 1..n of the ibead/jbead arrays into out_i/out_j; returns n.
int ref_get_list(int which, int* out_i, int* out_j) {
  int n = 0;
  int *ia = 0, *ja = 0;
  switch (which) {
    case 0: n = nnl_att; ia = ibead_neighbor_list_att; ja = jbead_neighbor_list_att; break;
    case 1: n = nnl_rep; ia = ibead_neighbor_list_rep; ja = jbead_neighbor_list_rep; break;
    case 2: n = nil_att; ia = ibead_pair_list_att; ja = jbead_pair_list_att; break;
    case 3: n = nil_rep; ia = ibead_pair_list_rep; ja = jbead_pair_list_rep; break;
    case 4: n = ncon_att; ia = ibead_lj_nat; ja = jbead_lj_nat; break;
    case 5: n = ncon_rep; ia = ibead_lj_non_nat; ja = jbead_lj_non_nat; break;
    default: return -1;
  }
  if (out_i && out_j) {
    for (int k = 1; k <= n; k++) { out_i[k - 1] = ia[k]; out_j[k - 1] = ja[k]; }
  }
  return n;
}

}  // extern "C"
