/*
 * sop_oracle.h -- CPU restatement of the SOP-CPU hot path (TEST INFRASTRUCTURE).
 *
 * This is the oracle: a from-scratch C restatement of the reference's
 * arithmetic, used only by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg, and only as the checker (or the timed CPU baseline).  The
 * product never links or calls it.
 *
 * Every function follows one reference function operation-for-operation, so
 * that, compiled with gcc -O3 and no -march (no FMA contraction, SSE2 double),
 * it reproduces the reference bit for bit.  Arrays are 1-based like the
 * reference's (index 0 unused).  Citations are file:line in SOP-CPU.
 *
 * Pinning (see tests/test_oracle.py): every function here is checked against
 * the compiled reference (oracle/_ref) on generated chains, against the
 * committed golden fixtures, and against closed forms / brute force.
 */
#ifndef SOP_ORACLE_H
#define SOP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y, z; } orc_vec;

/* Ran_Gen state (random_generator.h:4-33). */
typedef struct {
  int mseed, iset, idum2, iy;
  int iv[32];
  double gset;
} orc_rng;

typedef struct orc_sys {
  /* beads (global.h:133-136) */
  int nbead;
  orc_vec *pos, *unc_pos, *vel, *force, *incr;
  /* bonds, angles (global.h:29-57) */
  int nbnd; int *ibnd, *jbnd; double *pdb_dist;
  int nang; int *iang, *jang, *kang; double *pdb_ang;
  /* master lists (global.h:77-89) */
  int ncon_att, ncon_rep;
  int *ai, *aj, *ait, *ajt; double *ar, *ar2, *ar6, *ar12;
  int *ri, *rj, *rit, *rjt;
  /* neighbor list (global.h:91-105) */
  int nnl_att, nnl_rep;
  int *nai, *naj, *nait, *najt; double *nar, *nar2, *nar6, *nar12;
  int *nri, *nrj, *nrit, *nrjt;
  /* pair list (global.h:107-121) */
  int nil_att, nil_rep;
  int *pai, *paj, *pait, *pajt; double *par, *par2, *par6, *par12;
  int *pri, *prj, *prit, *prjt;
  /* parameters (params.cpp:262-400, global.cpp:60-92) */
  double k_bnd, R0, R0sq, e_bnd_coeff;
  double sigma_ss, sigma_ss6, epsilon_ss, e_ang_ss_coeff, f_ang_ss_coeff;
  double coeff_att[3][3], coeff_rep[3][3], force_coeff_att[3][3], force_coeff_rep[3][3];
  double sigma_rep[3][3], sigma_rep2[3][3], sigma_rep6[3][3], sigma_rep12[3][3];
  double rcut_nat[3][3];
  int run, restart, rgen_restart, sim_type, neighborlist, celllist;
  double T, boxl, ncell, lcell, zeta, nstep, istep_restart, h, halfh;
  double a1, a2, a3, a4, a5;
  int nup, nnlup;
  /* energies / observables */
  double e_bnd, e_ang_ss, e_vdw_rr, e_vdw_rr_att, e_vdw_rr_rep, rna_etot;
  double Q, rgsq, kinT; int contct_nat;
  /* noise: generator, or a replay buffer of 3*nbead gaussians per force_eval */
  orc_rng rng;
  const double* noise; long noise_pos;
} orc_sys;

/* life cycle / parameters */
orc_sys* orc_create(void);                 /* alloc_arrays defaults (params.cpp:262-400) */
void orc_destroy(orc_sys* s);
int orc_set(orc_sys* s, const char* key, const char* val);   /* set_params (params.cpp:12-101) */
int orc_load(orc_sys* s, const char* kind, const char* path); /* load (params.cpp:110-260) */

/* helpers */
extern int orc_pow_mode;                 /* 0: glibc pow (reference), 1: correctly rounded */
double orc_pow_cr(double d, int k);        /* correctly rounded d^6 / d^8 */
double orc_rnd(double x);                  /* global.cpp:223-227 */
void orc_rng_seed(orc_rng* g, int seed);   /* Ran_Gen ctor + set_seed */
double orc_ran2(orc_rng* g);               /* random_generator.cpp:114-146 */
double orc_gasdev(orc_rng* g);             /* random_generator.cpp:88-112 */

/* hot path */
void orc_update_neighbor_list(orc_sys* s); /* neighbor_list.cpp:6-117 */
void orc_update_cell_list(orc_sys* s);     /* cell_list.cpp:6-98 */
void orc_update_pair_list(orc_sys* s);     /* sop.cpp:473-553 */
void orc_clear_forces(orc_sys* s);         /* energy.cpp:35-45 */
void orc_random_force(orc_sys* s);         /* energy.cpp:449-465 */
void orc_fene_forces(orc_sys* s);          /* energy.cpp:402-447 */
void orc_angular_forces(orc_sys* s);       /* energy.cpp:354-400 */
void orc_vdw_forces(orc_sys* s);           /* energy.cpp:259-352 */
void orc_force_eval(orc_sys* s);           /* energy.cpp:21-33 with set_forces order */
void orc_fene_energy(orc_sys* s);          /* energy.cpp:115-152 */
void orc_angular_energy(orc_sys* s);       /* energy.cpp:154-187 */
void orc_vdw_energy(orc_sys* s);           /* energy.cpp:189-257 */
void orc_energy_eval(orc_sys* s);          /* energy.cpp:6-19 */
void orc_underdamped_iteration(orc_sys* s);/* sop.cpp:282-329 */
void orc_overdamped_iteration(orc_sys* s); /* sop.cpp:331-363 */
void orc_calculate_observables(orc_sys* s);/* sop.cpp:200-280 */

/* flat accessors for ctypes */
int orc_get_int(orc_sys* s, const char* name);
double orc_get_double(orc_sys* s, const char* name);
void orc_get_coords(orc_sys* s, int which, double* out);
void orc_set_coords(orc_sys* s, int which, const double* in);
int orc_get_list(orc_sys* s, int which, int* out_i, int* out_j);
void orc_set_noise(orc_sys* s, const double* noise);
void orc_gasdev_n(orc_sys* s, double* out, long n);
double orc_sys_ran2(orc_sys* s);
void orc_set_seed(orc_sys* s, int seed);

#ifdef __cplusplus
}
#endif
#endif
