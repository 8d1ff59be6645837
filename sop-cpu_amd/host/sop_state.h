// sop_state.h -- host-side program state and entry points of the SOP driver.
//
// The names, types and 1-based array conventions are those of SOP-CPU's
// global state (global.h:6-227) and entry-point headers (energy.h:4-22,
// neighbor_list.h:4, cell_list.h:4, sop.h:12-27, params.h:4-14, io.h:5-12), so
// that host code written against the reference links against this driver's
// shim unchanged.  The hot-path entry points (sop_shim.cpp) run on the GPU
// through the C ABI of include/sopg.h.
#ifndef SOPG_SOP_STATE_H
#define SOPG_SOP_STATE_H

#include <cstddef>

#ifdef SOPG_REFERENCE_GLOBALS
// Built into SOP-CPU's own program (the shim and the step loops linked with
// the reference's objects): the reference's headers declare the state and the
// entry points, and its global.cpp / params.cpp / io.cpp define them.
#include "global.h"
#include "energy.h"
#include "neighbor_list.h"
#include "cell_list.h"
#include "sop.h"
#include "params.h"
#include "io.h"
#else

#include "ran_gen.h"

class coord {
 public:
  coord();
  ~coord();
  double x;
  double y;
  double z;
};

const int mcmd = 100;
const int mopt = 10;
const int mopt_tot = mcmd * mopt;
const int mwdsize = 1024;
const size_t MAXPATHLEN = 2048;

extern char cmd[][mwdsize];
extern char opt[][mwdsize];
extern int opt_ptr[];
extern int ncmd;
extern char pathname[];

// bonds / angles
extern double k_bnd, R0, R0sq, e_bnd_coeff, e_bnd;
extern int nbnd;
extern int *ibead_bnd, *jbead_bnd;
extern double* pdb_dist;
extern int nang;
extern int *ibead_ang, *jbead_ang, *kbead_ang;
extern double* pdb_ang;
extern double sigma_ss, sigma_ss6, epsilon_ss, e_ang_ss_coeff, f_ang_ss_coeff, e_ang_ss;

// native / non-native master lists, neighbor list, pair list
extern int ncon_att, ncon_rep, nnl_att, nnl_rep, nil_att, nil_rep;
extern double coeff_att[][3], coeff_rep[][3], force_coeff_att[][3], force_coeff_rep[][3];
extern double sigma_rep[][3], sigma_rep2[][3], sigma_rep6[][3], sigma_rep12[][3], rcut_nat[][3];
extern int *ibead_lj_nat, *jbead_lj_nat, *itype_lj_nat, *jtype_lj_nat;
extern double *lj_nat_pdb_dist, *lj_nat_pdb_dist2, *lj_nat_pdb_dist6, *lj_nat_pdb_dist12;
extern int *ibead_lj_non_nat, *jbead_lj_non_nat, *itype_lj_non_nat, *jtype_lj_non_nat;
extern int *ibead_neighbor_list_att, *jbead_neighbor_list_att, *itype_neighbor_list_att, *jtype_neighbor_list_att;
extern double *nl_lj_nat_pdb_dist, *nl_lj_nat_pdb_dist2, *nl_lj_nat_pdb_dist6, *nl_lj_nat_pdb_dist12;
extern int *ibead_neighbor_list_rep, *jbead_neighbor_list_rep, *itype_neighbor_list_rep, *jtype_neighbor_list_rep;
extern int *ibead_pair_list_att, *jbead_pair_list_att, *itype_pair_list_att, *jtype_pair_list_att;
extern double *pl_lj_nat_pdb_dist, *pl_lj_nat_pdb_dist2, *pl_lj_nat_pdb_dist6, *pl_lj_nat_pdb_dist12;
extern int *ibead_pair_list_rep, *jbead_pair_list_rep, *itype_pair_list_rep, *jtype_pair_list_rep;
extern double e_vdw_rr, e_vdw_rr_att, e_vdw_rr_rep;

// beads
extern int nbead;
extern coord *pos, *unc_pos, *vel, *force;

// run parameters
extern int run;
extern Ran_Gen generator;
extern int restart, rgen_restart, sim_type;
extern double T;
extern int neighborlist, celllist;
extern double boxl, ncell, lcell, zeta, nstep, istep_restart;
extern int nup, inlup, nnlup;
extern double h, halfh, a1, a2, a3, a4, a5;
extern char ufname[], rcfname[], cfname[], unccfname[], vfname[], binfname[], uncbinfname[];
extern int binsave;

// potential / force term tables (the reference's plug-in hook, global.h:185-198)
extern int npot_term;
const int mpot_term = 10;
extern int pot_term_on[];
typedef void (*pot_term_Ptr)();
extern pot_term_Ptr pot_term[];
extern double rna_etot, system_etot;
extern int nforce_term;
const int mforce_term = 10;
extern int force_term_on[];
typedef void (*force_term_Ptr)();
extern force_term_Ptr force_term[];

// observables
extern double Q, rgsq, kinT;
extern int contct_nat;

double rnd(double);

// energy.h
void set_potential();
void energy_eval();
void fene_energy();
void soft_sphere_angular_energy();
void vdw_energy();
void set_forces();
void clear_forces();
void force_eval();
void random_force();
void fene_forces();
void soft_sphere_angular_forces();
void vdw_forces();
// neighbor_list.h / cell_list.h
void update_neighbor_list();
void update_cell_list();
// sop.h
void ex_cmds();
void simulation_ctrl();
void underdamped_ctrl();
void overdamped_ctrl();
void underdamped_iteration(coord*);
void overdamped_iteration(coord*);
void calculate_observables(coord*);
void update_pair_list();
// params.h
void load(int);
void set_temp(double);
void alloc_arrays();
void init_bonds(int);
void init_angles(int);
void init_pos(int);
void init_lj(int, int);
void set_params(int);
// io.h
void read_input(const char* const);
void record_traj(char*, char*);
void save_coords(char*, char*);
void load_coords(char*, char*);
void save_unccoords(char*);
void save_vels(char*);
void load_vels(char*);
void print_sim_params();

#endif  // SOPG_REFERENCE_GLOBALS

// Driver extensions (defined in sop_shim.cpp; set only by this driver's
// set_params -- the reference skips unknown `set` keys, so in its program
// they keep these defaults):
//   set noise device;             Philox noise on the GPU instead of replaying `generator`
//   set lanes L;                  lanes per bead for the step kernels
//   set restart_format lossless;  coordinate, velocity and generator-state files at %.17g
//                                 (the reference's %f loses digits; reading takes either)
extern int sopg_noise_device;
extern int sopg_lanes;
extern int sopg_restart_lossless;

#endif
