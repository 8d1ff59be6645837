// sop_state.cpp -- definitions, defaults, `set` and `load` commands of the
// host driver.  Semantics follow the reference: defaults of alloc_arrays
// (params.cpp:262-400), the order-sensitive `set` keys of set_params
// (params.cpp:12-101), and the four input formats of load (params.cpp:110-260),
// read with the same space-delimited tokenisation.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>

#include "sop_state.h"

coord::coord() : x(0.0), y(0.0), z(0.0) {}
coord::~coord() {}

char cmd[mcmd + 1][mwdsize];
char opt[mopt_tot + 1][mwdsize];
int opt_ptr[mcmd + 2];
int ncmd;
char pathname[MAXPATHLEN];

double k_bnd, R0, R0sq, e_bnd_coeff, e_bnd;
int nbnd;
int *ibead_bnd, *jbead_bnd;
double* pdb_dist;
int nang;
int *ibead_ang, *jbead_ang, *kbead_ang;
double* pdb_ang;
double sigma_ss, sigma_ss6, epsilon_ss, e_ang_ss_coeff, f_ang_ss_coeff, e_ang_ss;

int ncon_att, ncon_rep, nnl_att, nnl_rep, nil_att, nil_rep;
// type tables (global.cpp:60-92)
double coeff_att[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.7, 0.8}, {0.0, 0.8, 1.0}};
double coeff_rep[3][3] = {{0.0, 0.0, 0.0}, {0.0, 1.0, 1.0}, {0.0, 1.0, 1.0}};
double force_coeff_att[3][3] = {{0.0, 0.0, 0.0}, {0.0, -12.0 * 1.0, -12.0 * 0.8}, {0.0, -12.0 * 0.8, -12.0 * 0.7}};
double force_coeff_rep[3][3] = {{0.0, 0.0, 0.0}, {0.0, -6.0 * 1.0, -6.0 * 1.0}, {0.0, -6.0 * 1.0, -6.0 * 1.0}};
double sigma_rep[3][3] = {{0.0, 0.0, 0.0}, {0.0, 3.8, 5.4}, {0.0, 5.4, 7.0}};
double sigma_rep2[3][3], sigma_rep6[3][3], sigma_rep12[3][3];
double rcut_nat[3][3] = {{0.0, 0.0, 0.0}, {0.0, 8.0, 11.0}, {0.0, 11.0, 14.0}};
int *ibead_lj_nat, *jbead_lj_nat, *itype_lj_nat, *jtype_lj_nat;
double *lj_nat_pdb_dist, *lj_nat_pdb_dist2, *lj_nat_pdb_dist6, *lj_nat_pdb_dist12;
int *ibead_lj_non_nat, *jbead_lj_non_nat, *itype_lj_non_nat, *jtype_lj_non_nat;
int *ibead_neighbor_list_att, *jbead_neighbor_list_att, *itype_neighbor_list_att, *jtype_neighbor_list_att;
double *nl_lj_nat_pdb_dist, *nl_lj_nat_pdb_dist2, *nl_lj_nat_pdb_dist6, *nl_lj_nat_pdb_dist12;
int *ibead_neighbor_list_rep, *jbead_neighbor_list_rep, *itype_neighbor_list_rep, *jtype_neighbor_list_rep;
int *ibead_pair_list_att, *jbead_pair_list_att, *itype_pair_list_att, *jtype_pair_list_att;
double *pl_lj_nat_pdb_dist, *pl_lj_nat_pdb_dist2, *pl_lj_nat_pdb_dist6, *pl_lj_nat_pdb_dist12;
int *ibead_pair_list_rep, *jbead_pair_list_rep, *itype_pair_list_rep, *jtype_pair_list_rep;
double e_vdw_rr, e_vdw_rr_att, e_vdw_rr_rep;

int nbead;
coord *pos, *unc_pos, *vel, *force;

int run;
Ran_Gen generator;
int restart = 0, rgen_restart = 0, sim_type = 1;
double T;
int neighborlist = 0, celllist = 0;
double boxl, ncell, lcell, zeta, nstep, istep_restart = 0.0;
int nup, inlup, nnlup;
double h, halfh, a1, a2, a3, a4, a5;
char ufname[mwdsize + 1], rcfname[mwdsize + 1], cfname[mwdsize + 1], unccfname[mwdsize + 1];
char vfname[mwdsize + 1], binfname[mwdsize + 1], uncbinfname[mwdsize + 1];
int binsave = 1;

// default term tables (global.cpp:200-208): random, FENE, angular, vdw
int nforce_term = 4;
int force_term_on[mforce_term + 1] = {0, 1, 1, 1, 0, 0, 1, 0, 0, 0, 0};
force_term_Ptr force_term[mforce_term + 1];
int npot_term = 3;
int pot_term_on[mpot_term + 1] = {0, 1, 1, 0, 0, 1, 0, 0, 0, 0, 0};
pot_term_Ptr pot_term[mpot_term + 1];
double rna_etot, system_etot;

double Q, rgsq, kinT;
int contct_nat;


double rnd(double x) { return (x > 0) ? std::floor(x + 0.5) : std::ceil(x - 0.5); }

// ---------------------------------------------------------------- allocation

template <class T>
static T* fresh(int n) {
  return new T[(size_t)n + 1]();
}

void init_bonds(int n) {
  nbnd = n;
  ibead_bnd = fresh<int>(n);
  jbead_bnd = fresh<int>(n);
  pdb_dist = fresh<double>(n);
}

void init_angles(int n) {
  nang = n;
  ibead_ang = fresh<int>(n);
  jbead_ang = fresh<int>(n);
  kbead_ang = fresh<int>(n);
  pdb_ang = fresh<double>(n);
}

void init_lj(int na, int nr) {
  ncon_att = na;
  ncon_rep = nr;
  ibead_lj_nat = fresh<int>(na); jbead_lj_nat = fresh<int>(na);
  itype_lj_nat = fresh<int>(na); jtype_lj_nat = fresh<int>(na);
  lj_nat_pdb_dist = fresh<double>(na); lj_nat_pdb_dist2 = fresh<double>(na);
  lj_nat_pdb_dist6 = fresh<double>(na); lj_nat_pdb_dist12 = fresh<double>(na);
  ibead_lj_non_nat = fresh<int>(nr); jbead_lj_non_nat = fresh<int>(nr);
  itype_lj_non_nat = fresh<int>(nr); jtype_lj_non_nat = fresh<int>(nr);
  ibead_neighbor_list_att = fresh<int>(na); jbead_neighbor_list_att = fresh<int>(na);
  itype_neighbor_list_att = fresh<int>(na); jtype_neighbor_list_att = fresh<int>(na);
  nl_lj_nat_pdb_dist = fresh<double>(na); nl_lj_nat_pdb_dist2 = fresh<double>(na);
  nl_lj_nat_pdb_dist6 = fresh<double>(na); nl_lj_nat_pdb_dist12 = fresh<double>(na);
  ibead_neighbor_list_rep = fresh<int>(nr); jbead_neighbor_list_rep = fresh<int>(nr);
  itype_neighbor_list_rep = fresh<int>(nr); jtype_neighbor_list_rep = fresh<int>(nr);
  ibead_pair_list_att = fresh<int>(na); jbead_pair_list_att = fresh<int>(na);
  itype_pair_list_att = fresh<int>(na); jtype_pair_list_att = fresh<int>(na);
  pl_lj_nat_pdb_dist = fresh<double>(na); pl_lj_nat_pdb_dist2 = fresh<double>(na);
  pl_lj_nat_pdb_dist6 = fresh<double>(na); pl_lj_nat_pdb_dist12 = fresh<double>(na);
  ibead_pair_list_rep = fresh<int>(nr); jbead_pair_list_rep = fresh<int>(nr);
  itype_pair_list_rep = fresh<int>(nr); jtype_pair_list_rep = fresh<int>(nr);
}

void init_pos(int n) {
  nbead = n;
  pos = fresh<coord>(n);
  unc_pos = fresh<coord>(n);
  vel = fresh<coord>(n);
  force = fresh<coord>(n);
}

static void set_underdamped_coeffs() {
  halfh = h / 2.0;
  a1 = h * (1.0 - zeta * halfh);
  a2 = h * halfh;
  a3 = (1.0 - h * zeta / 2.0 + (h * zeta) * (h * zeta) / 4.0) / h;
  a4 = halfh * (1.0 - h * zeta / 2.0);
}

// Defaults (params.cpp:262-400).  Arrays are sized by the loaders.
void alloc_arrays() {
  k_bnd = 20.0;
  R0 = 2.0;
  R0sq = R0 * R0;
  e_bnd_coeff = k_bnd * R0sq / 2.0;
  sigma_ss = 3.5;
  sigma_ss6 = std::pow(sigma_ss, 6.0);
  epsilon_ss = 1.0;
  e_ang_ss_coeff = epsilon_ss * sigma_ss6;
  f_ang_ss_coeff = 6.0 * e_ang_ss_coeff;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      sigma_rep2[i][j] = sigma_rep[i][j] * sigma_rep[i][j];
      sigma_rep6[i][j] = sigma_rep2[i][j] * sigma_rep2[i][j] * sigma_rep2[i][j];
      sigma_rep12[i][j] = sigma_rep6[i][j] * sigma_rep6[i][j];
    }
  init_bonds(0);
  init_angles(0);
  init_lj(0, 0);
  init_pos(0);
  nnl_att = nnl_rep = nil_att = nil_rep = 0;
  run = 1;
  generator.set_seed(-100 - run);
  T = 0.6;
  neighborlist = 0;
  celllist = 0;
  boxl = 500.0;
  ncell = 55.0;
  lcell = boxl / ncell;
  zeta = 5.0e-2;
  nstep = 5e7;
  nup = 1000;
  nnlup = 50;
  h = 2.5e-3;
  set_underdamped_coeffs();
  a5 = h / zeta;
  std::strcpy(ufname, "update.out");
  std::strcpy(rcfname, "restart_c.dat");
  std::strcpy(cfname, "coord.out");
  std::strcpy(unccfname, "unccoord.out");
  std::strcpy(vfname, "veloc.out");
  std::strcpy(binfname, "traj.bin");
  std::strcpy(uncbinfname, "traj_uncorrected.bin");
}

void set_temp(double t) { T = t; }

// `set <key> <value>` (params.cpp:12-101).  Order matters exactly as in the
// reference: `dynamics` and `zeta` reset h; `zeta` re-derives a1/a3/a4/a5 but
// not halfh/a2; `ncell` derives lcell from the current boxl.
void set_params(int icmd) {
  const char* key = opt[opt_ptr[icmd]];
  const char* val = opt[opt_ptr[icmd] + 1];
  if (!std::strcmp(key, "dynamics")) {
    if (!std::strcmp(val, "underdamped")) {
      sim_type = 1;
      h = 2.5e-3;
      set_underdamped_coeffs();
    } else if (!std::strcmp(val, "overdamped")) {
      sim_type = 2;
      h = 0.02;
      a5 = h / zeta;
    }
  } else if (!std::strcmp(key, "temp")) {
    set_temp(std::atof(val));
  } else if (!std::strcmp(key, "nstep")) {
    nstep = std::atof(val);
  } else if (!std::strcmp(key, "istep_restart")) {
    istep_restart = std::atof(val);
  } else if (!std::strcmp(key, "nup")) {
    nup = std::atoi(val);
  } else if (!std::strcmp(key, "run")) {
    run = std::atoi(val);
    generator.set_seed(-100 - run);
  } else if (!std::strcmp(key, "ufname")) {
    std::strcpy(ufname, val);
  } else if (!std::strcmp(key, "rcfname")) {
    std::strcpy(rcfname, val);
  } else if (!std::strcmp(key, "cfname")) {
    std::strcpy(cfname, val);
  } else if (!std::strcmp(key, "rgenfname")) {
    generator.set_fname(opt[opt_ptr[icmd] + 1]);
  } else if (!std::strcmp(key, "unccfname")) {
    std::strcpy(unccfname, val);
  } else if (!std::strcmp(key, "vfname")) {
    std::strcpy(vfname, val);
  } else if (!std::strcmp(key, "binfname")) {
    std::strcpy(binfname, val);
  } else if (!std::strcmp(key, "uncbinfname")) {
    std::strcpy(uncbinfname, val);
  } else if (!std::strcmp(key, "cutofftype")) {
    if (!std::strcmp(val, "neighborlist")) neighborlist = 1;
    else if (!std::strcmp(val, "celllist")) celllist = 1;
  } else if (!std::strcmp(key, "nnlup")) {
    nnlup = std::atoi(val);
  } else if (!std::strcmp(key, "boxl")) {
    boxl = std::atof(val);
  } else if (!std::strcmp(key, "ncell")) {
    ncell = std::atof(val);
    lcell = boxl / ncell;
  } else if (!std::strcmp(key, "restart")) {
    restart = !std::strcmp(val, "on") ? 1 : 0;
  } else if (!std::strcmp(key, "rgen_restart")) {
    rgen_restart = !std::strcmp(val, "on") ? 1 : 0;
  } else if (!std::strcmp(key, "t_step")) {
    h = std::atof(val);
    set_underdamped_coeffs();
    a5 = h / zeta;
  } else if (!std::strcmp(key, "zeta")) {
    if (sim_type == 1) h = 2.5e-3;
    else if (sim_type == 2) h = 0.02;
    zeta = std::atof(val);
    a1 = h * (1.0 - zeta * halfh);
    a3 = (1.0 - h * zeta / 2.0 + (h * zeta) * (h * zeta) / 4.0) / h;
    a4 = halfh * (1.0 - h * zeta / 2.0);
    a5 = h / zeta;
  } else if (!std::strcmp(key, "noise")) {  // driver extension
    sopg_noise_device = !std::strcmp(val, "device") ? 1 : 0;
  } else if (!std::strcmp(key, "lanes")) {  // driver extension
    sopg_lanes = std::atoi(val);
  } else if (!std::strcmp(key, "restart_format")) {  // driver extension
    sopg_restart_lossless = !std::strcmp(val, "lossless") ? 1 : 0;
  }
}

// ---------------------------------------------------------------- loaders

namespace {
// One line split on single spaces (strtok semantics); missing fields read as "0".
struct Line {
  char buf[2048];
  char* tok[8];
  int ntok = 0;
  bool get(std::ifstream& in) {
    buf[0] = 0;
    in.getline(buf, sizeof buf);
    ntok = 0;
    for (char* t = std::strtok(buf, " "); t && ntok < 8; t = std::strtok(nullptr, " ")) tok[ntok++] = t;
    return (bool)in || buf[0];
  }
  const char* s(int i) const { return i < ntok ? tok[i] : "0"; }
  int i(int k) const { return std::atoi(s(k)); }
  double d(int k) const { return std::atof(s(k)); }
};
}  // namespace

void load(int icmd) {
  const char* kind = opt[opt_ptr[icmd]];
  const char* path = opt[opt_ptr[icmd] + 1];
  std::ifstream in(path);
  if (!in) {
    std::cerr << "CANNOT OPEN " << path << std::endl;
    std::exit(-1);
  }
  Line L;
  if (!std::strcmp(kind, "bonds")) {
    std::cout << "[Reading in bonds...]" << std::endl;
    L.get(in);
    init_bonds(L.i(1));
    for (int k = 1; k <= nbnd; k++) {
      L.get(in);
      ibead_bnd[k] = L.i(0);
      jbead_bnd[k] = L.i(1);
      pdb_dist[k] = L.d(2);
    }
    std::cout << "[Finished reading bonds (" << nbnd << ")]" << std::endl;
  } else if (!std::strcmp(kind, "angles")) {
    std::cout << "[Reading in angles...]" << std::endl;
    L.get(in);
    init_angles(L.i(1));
    for (int k = 1; k <= nang; k++) {
      L.get(in);
      ibead_ang[k] = L.i(0);
      jbead_ang[k] = L.i(1);
      kbead_ang[k] = L.i(2);
      pdb_ang[k] = L.d(3);
    }
    std::cout << "[Finished reading angles (" << nang << ")]" << std::endl;
  } else if (!std::strcmp(kind, "vdw")) {
    std::cout << "[Reading in VDW interactions...]" << std::endl;
    L.get(in);
    const int na = L.i(1), nr = L.i(3);
    init_lj(na, nr);
    int ia = 0, ir = 0;
    nil_att = nil_rep = 0;
    for (int k = 1; k <= na + nr; k++) {
      L.get(in);
      const int ib = L.i(0), jb = L.i(1), it = L.i(3), jt = L.i(4);
      const double r = L.d(2);
      if (it < 0 || it > 2 || jt < 0 || jt > 2) {
        std::cerr << "vdw: bead type out of range at line " << k + 1 << std::endl;
        std::exit(-1);
      }
      // native if closer than rcut_nat (params.cpp:197); the pair list starts
      // as the full master list (params.cpp:210-230)
      if (r < rcut_nat[it][jt]) {
        if (ia == na) { std::cerr << "vdw: more native pairs than the header declares" << std::endl; std::exit(-1); }
        ia++;
        ibead_lj_nat[ia] = ib; jbead_lj_nat[ia] = jb; itype_lj_nat[ia] = it; jtype_lj_nat[ia] = jt;
        lj_nat_pdb_dist[ia] = r;
        lj_nat_pdb_dist2[ia] = r * r;
        lj_nat_pdb_dist6[ia] = lj_nat_pdb_dist2[ia] * lj_nat_pdb_dist2[ia] * lj_nat_pdb_dist2[ia];
        lj_nat_pdb_dist12[ia] = lj_nat_pdb_dist6[ia] * lj_nat_pdb_dist6[ia];
        nil_att++;
        ibead_pair_list_att[nil_att] = ib; jbead_pair_list_att[nil_att] = jb;
        itype_pair_list_att[nil_att] = it; jtype_pair_list_att[nil_att] = jt;
        pl_lj_nat_pdb_dist[nil_att] = r;
        pl_lj_nat_pdb_dist2[nil_att] = lj_nat_pdb_dist2[ia];
        pl_lj_nat_pdb_dist6[nil_att] = lj_nat_pdb_dist6[ia];
        pl_lj_nat_pdb_dist12[nil_att] = lj_nat_pdb_dist12[ia];
      } else {
        if (ir == nr) { std::cerr << "vdw: more non-native pairs than the header declares" << std::endl; std::exit(-1); }
        ir++;
        ibead_lj_non_nat[ir] = ib; jbead_lj_non_nat[ir] = jb; itype_lj_non_nat[ir] = it; jtype_lj_non_nat[ir] = jt;
        nil_rep++;
        ibead_pair_list_rep[nil_rep] = ib; jbead_pair_list_rep[nil_rep] = jb;
        itype_pair_list_rep[nil_rep] = it; jtype_pair_list_rep[nil_rep] = jt;
      }
    }
    std::cout << "[Finished reading VDW interactions (" << ia << "/" << ir << ")]" << std::endl;
    // a header that overstates the counts leaves the tail unread (the
    // reference reads past EOF into uninitialised entries): refuse it
    if (ia != na || ir != nr) {
      std::cerr << "vdw: header counts " << na << "/" << nr << " disagree with the data " << ia << "/" << ir
                << std::endl;
      std::exit(-1);
    }
  } else if (!std::strcmp(kind, "init")) {
    std::cout << "[Reading in initial coordinates...]" << std::endl;
    L.get(in);
    init_pos(L.i(1));
    for (int k = 1; k <= nbead; k++) {
      L.get(in);
      pos[k].x = unc_pos[k].x = L.d(1);
      pos[k].y = unc_pos[k].y = L.d(2);
      pos[k].z = unc_pos[k].z = L.d(3);
    }
    std::cout << "[Finished reading initial coordinates (" << nbead << ")]" << std::endl;
  }
}
