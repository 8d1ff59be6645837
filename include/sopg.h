/*
 * sopg.h -- C ABI of the MI355X (gfx950) SOP Langevin-dynamics hot path.
 *
 * Drop-in boundary for SOP-CPU's hot path.  Each entry point below replaces
 * one reference function (cited file:line in SOP-CPU); the C++ shim in
 * sop-cpu_amd/host/ defines the reference's own signatures (energy.h,
 * neighbor_list.h, cell_list.h, sop.h) on top of these.
 *
 * Conventions
 *   - Every function returns an int status: SOPG_OK (0) or a negative code;
 *     sopg_last_error() gives the message.  No exceptions cross the ABI.
 *   - Plain pointers and sizes only.  Bead indices in topology arrays are
 *     1-based, as in the reference's files; arrays themselves are 0-based
 *     (pass &ibead_bnd[1] for a reference 1-based array).
 *   - Coordinates are AoS x,y,z triples per bead (the reference's `coord`
 *     layout, global.h:6-13): pass &pos[1].x.
 *   - A context owns all device memory of one GPU stream and holds one or more
 *     chains (a ragged batch).  One host thread per context.
 *   - FP64 throughout.  Device code is built without FMA contraction and with
 *     IEEE division/sqrt so that forces, lists and trajectories reproduce the
 *     reference's arithmetic.
 */
#ifndef SOPG_H
#define SOPG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SOPG_OK 0
#define SOPG_ERR_ARG -1       /* invalid argument / state */
#define SOPG_ERR_HIP -2       /* HIP runtime error */
#define SOPG_ERR_OVERFLOW -3  /* a per-bead neighbor list exceeded its capacity */
#define SOPG_ERR_UNSUPPORTED -4

#define SOPG_CUTOFF_NONE 0    /* no cutoff: the loader's full master lists (params.cpp:210-230) */
#define SOPG_CUTOFF_NEIGHBOR 1 /* cutofftype neighborlist (neighbor_list.cpp) */
#define SOPG_CUTOFF_CELL 2    /* cutofftype celllist (cell_list.cpp) */

#define SOPG_NOISE_REPLAY 0   /* caller supplies the 3N gaussians of each force_eval */
#define SOPG_NOISE_PHILOX 1   /* device Philox4x32-10 + Box-Muller, keyed by chain seed */

/* Which list sopg_get_list materialises (reference order). */
#define SOPG_LIST_NL_ATT 0    /* ibead/jbead_neighbor_list_att, nnl_att */
#define SOPG_LIST_NL_REP 1    /* ibead/jbead_neighbor_list_rep, nnl_rep */
#define SOPG_LIST_PL_ATT 2    /* ibead/jbead_pair_list_att, nil_att */
#define SOPG_LIST_PL_REP 3    /* ibead/jbead_pair_list_rep, nil_rep */

typedef struct sopg_ctx sopg_ctx;

/* Type tables and force-field constants (global.cpp:60-92, params.cpp:268-312).
 * sopg_default_tables() fills the reference's values. */
typedef struct {
  double coeff_att[3][3], coeff_rep[3][3];
  double force_coeff_att[3][3], force_coeff_rep[3][3];
  double sigma_rep[3][3];
  double k_bnd, R0sq, e_bnd_coeff;     /* FENE: k=20, R0^2=4, kR0^2/2 */
  double e_ang_ss_coeff, f_ang_ss_coeff; /* eps*sigma_ss^6, 6*eps*sigma_ss^6 */
} sopg_tables;

/* Per-chain run parameters (params.cpp:12-101 semantics are applied by the
 * caller: pass the already-derived a1..a5 exactly as set_params left them). */
typedef struct {
  double T, zeta, h;                   /* temperature, friction, time step */
  double a1, a2, a3, a4, a5;           /* integrator coefficients */
  double boxl, ncell, lcell;           /* periodic box, cells per edge, cell edge */
  int cutoff;                          /* SOPG_CUTOFF_* */
  int nnlup;                           /* list rebuild period (steps) */
  int nup;                             /* observable/record period (steps) */
  int noise_mode;                      /* SOPG_NOISE_* */
  uint64_t seed;                       /* Philox key (SOPG_NOISE_PHILOX) */
  int sim_type;                        /* 1 underdamped (sop.cpp:282-329), 2 overdamped (sop.cpp:331-363) */
} sopg_chain_params;

/* One chain's topology.  Arrays are 0-based with 1-based bead indices.
 * rep_* may be NULL with ncon_rep = -1: the non-native master list is then the
 * complement of the native list among all pairs j-i>2 in (i,j) lexicographic
 * order -- exactly the list a complete lexicographic vdw file yields -- and
 * bead_type (may be NULL: all type 1) gives itype/jtype. */
typedef struct {
  int nbead;
  int nbnd; const int* ibnd; const int* jbnd; const double* r0bnd;
  int nang; const int* iang; const int* jang; const int* kang;
  int ncon_att; const int* att_i; const int* att_j; const int* att_it; const int* att_jt;
  const double* att_r;                 /* lj_nat_pdb_dist */
  int ncon_rep; const int* rep_i; const int* rep_j; const int* rep_it; const int* rep_jt;
  const int* bead_type;
} sopg_topology;

typedef struct { double e_bnd, e_ang_ss, e_vdw_rr_att, e_vdw_rr_rep, e_vdw_rr, rna_etot; } sopg_energies;
typedef struct { double Q; int contct_nat; double rgsq, kinT; } sopg_observables;

/* One `update.out` record (sop.cpp:176-179) produced by sopg_run. */
typedef struct {
  int chain; double istep, T, kinT, e_bnd, e_ang_ss, e_vdw_rr, rna_etot, Q; int contct_nat; double rgsq;
} sopg_record;

/* ---------------------------------------------------------------- context */
int sopg_create(int device, sopg_ctx** out);
int sopg_destroy(sopg_ctx* ctx);
const char* sopg_last_error(const sopg_ctx* ctx);
int sopg_default_tables(sopg_tables* t);
int sopg_set_tables(sopg_ctx* ctx, const sopg_tables* t);
/* Lanes (1..64, power of two; 0 = default) that cooperate on one bead in the
 * step (force) kernels and in the list-rebuild kernels.  Results (states,
 * lists, records) do not depend on either. */
int sopg_set_lanes(sopg_ctx* ctx, int step_lanes, int list_lanes);
int sopg_get_lanes(const sopg_ctx* ctx, int* step_lanes, int* list_lanes);
/* Time the step kernel at every lane count on the context's current chains
 * and keep the fastest (no counterpart in the reference: the CPU loop has no
 * launch shape).  *step_lanes = the choice. */
int sopg_autotune(sopg_ctx* ctx, int* step_lanes);
/* enable = 0 disables the binned (cell-grid) list rebuild and always scans
 * every candidate; the lists are identical either way (default: enabled). */
int sopg_set_binned(sopg_ctx* ctx, int enable);
/* Per-bead capacity of the non-native neighbor list (0 = automatic: 1024, or
 * the largest no-cutoff chain).  Fixed at the first device operation. */
int sopg_set_capacity(sopg_ctx* ctx, int cap);
/* Pre-size the device pools for `beads` beads in `chains` chains (optional:
 * pools grow on demand, but growing copies every array once).  A context
 * holds at most 2^28 beads (SOPG_ERR_ARG beyond) and a chain 2^25. */
int sopg_reserve(sopg_ctx* ctx, long beads, int chains);

/* ---------------------------------------------------------------- chains */
/* Topology copy of the loaders (params.cpp:110-260); returns chain id in *id
 * (the lowest free id).  A chain may join a context at any time, also between
 * sopg_run calls: its beads start at zero (set them with sopg_put_state). */
int sopg_add_chain(sopg_ctx* ctx, const sopg_topology* topo, const sopg_chain_params* p, int* id);
/* Release a chain's device memory and id (continuous batching: a finished
 * chain leaves, a queued one joins; the other chains are not touched). */
int sopg_remove_chain(sopg_ctx* ctx, int chain);
int sopg_set_chain_params(sopg_ctx* ctx, int chain, const sopg_chain_params* p);
/* State of one chain: pos, unc_pos, vel, force as AoS 3N doubles (any may be NULL). */
int sopg_put_state(sopg_ctx* ctx, int chain, const double* pos, const double* unc_pos,
                   const double* vel, const double* force);
int sopg_get_state(sopg_ctx* ctx, int chain, double* pos, double* unc_pos, double* vel, double* force);
int sopg_nchains(const sopg_ctx* ctx);
int sopg_chain_nbead(const sopg_ctx* ctx, int chain);

/* ------------------------------------------- reference entry points (one chain) */
/* update_neighbor_list()  neighbor_list.cpp:6-117 */
int sopg_update_neighbor_list(sopg_ctx* ctx, int chain, int* nnl_att, int* nnl_rep);
/* update_cell_list()      cell_list.cpp:6-98 */
int sopg_update_cell_list(sopg_ctx* ctx, int chain, int* nnl_att, int* nnl_rep);
/* update_pair_list()      sop.cpp:473-553 */
int sopg_update_pair_list(sopg_ctx* ctx, int chain, int* nil_att, int* nil_rep);
/* force_eval()            energy.cpp:21-33 (clear, random, FENE, angular, vdw).
 * noise: the 3N gaussians this call consumes (bead-major x,y,z), or NULL for
 * the chain's device generator. */
int sopg_force_eval(sopg_ctx* ctx, int chain, const double* noise);
/* The individual force terms of the reference's force_term[] table, ADDED to
 * the chain's current force (each reference term accumulates into force[]):
 * random_force (energy.cpp:449-465), fene_forces (:402-447),
 * soft_sphere_angular_forces (:354-400), vdw_forces (:259-352).  Terms run in
 * the table's order (random, FENE, angular, vdw) within one call. */
#define SOPG_TERM_RANDOM 1
#define SOPG_TERM_FENE 2
#define SOPG_TERM_ANGULAR 4
#define SOPG_TERM_VDW 8
int sopg_force_terms(sopg_ctx* ctx, int chain, int terms, const double* noise);
/* energy_eval()           energy.cpp:6-19 */
int sopg_energy_eval(sopg_ctx* ctx, int chain, sopg_energies* out);
/* underdamped_iteration() sop.cpp:282-329 */
int sopg_underdamped_iteration(sopg_ctx* ctx, int chain, const double* noise);
/* overdamped_iteration()  sop.cpp:331-363 (chain with sim_type 2) */
int sopg_overdamped_iteration(sopg_ctx* ctx, int chain, const double* noise);
/* calculate_observables() sop.cpp:200-280 */
int sopg_calculate_observables(sopg_ctx* ctx, int chain, sopg_observables* out);
/* The current NL / PL of a chain in the reference's order (1-based bead ids).
 * Writes at most cap entries; *n = list length. */
int sopg_get_list(sopg_ctx* ctx, int chain, int which, int* ibead, int* jbead, long cap, long* n);

/* ---------------------------------------------------------------- hot loop */
/* Advance every chain of the context by nsteps steps of the underdamped_ctrl
 * loop body (sop.cpp:153-189): list rebuild every nnlup steps, pair list every
 * step, underdamped_iteration, and an energy + observables record every nup
 * steps.  Step counters continue across calls (sopg_set_step).  noise: NULL
 * (device generator) or nsteps*3*Ntot gaussians, step-major; within a step the
 * chains in ascending id order, 3N each, bead-major.  Records go to rec (up to
 * rec_cap), in step order and ascending chain id; *nrec = count produced. */
int sopg_run(sopg_ctx* ctx, long nsteps, const double* noise, sopg_record* rec, long rec_cap, long* nrec);
/* sopg_run without any device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * steps, states and records as sopg_run, with these limits: every chain uses
 * the device generator (no replayed noise), the context has already made its
 * first device operation (it allocates there), kernel timing is off, and the
 * records of one call fit a fixed 4 MiB per-GPU scratch array: (record steps
 * in the call) x (chain slots) x 12 doubles <= 2^19, else SOPG_ERR_ARG.  It
 * launches every kernel directly (no graph capture, no resident kernel).
 * Every context on one GPU shares that GPU's scratch array: calls that
 * overlap on one GPU (two host threads) race on it. */
int sopg_run_noalloc(sopg_ctx* ctx, long nsteps, sopg_record* rec, long rec_cap, long* nrec);
/* Set the step counters of a chain: istep of the next step, and the
 * reference's inlup / iup counters as they stand before that step. */
int sopg_set_step(sopg_ctx* ctx, int chain, double istep, int inlup, int iup);
int sopg_get_step(const sopg_ctx* ctx, int chain, double* istep, int* inlup, int* iup);
/* Lossless checkpoint of one chain (no counterpart in the reference, whose
 * restart files -- coordinates, velocities and the generator state, io.cpp
 * and random_generator.cpp:24-70 -- rebuild the lists at restart): the state,
 * the step and noise counters and the current neighbor lists, so that
 * restoring it into a chain of the same topology, in any context, continues
 * the run exactly as if it had not stopped.  buf = NULL: *size = bytes needed. */
int sopg_checkpoint(sopg_ctx* ctx, int chain, void* buf, long cap, long* size);
int sopg_restore(sopg_ctx* ctx, int chain, const void* buf, long size);
/* sopg_run advances every chain of at most max_beads beads (0: none, the
 * default) in one persistent kernel: a block per chain runs all of the run's
 * steps -- force, integration, records, list rebuilds -- with block barriers
 * only, concurrently with the per-step kernels of the other chains.  lanes:
 * lanes per bead in it (0: automatic); max_beads * lanes <= 512.  Results do
 * not depend on it.  No counterpart in the reference. */
int sopg_set_resident(sopg_ctx* ctx, int max_beads, int lanes);
/* Skin of the candidate lists of neighbor-list rebuilds during sopg_run, in
 * Angstrom (default 4; 0: off).  A candidate list holds a bead's non-native
 * pairs within 3.2 sigma + skin of the positions it was built at; while no
 * bead of the chain has moved more than skin/2 from them, a rebuild filters
 * the candidates instead of scanning the grid, otherwise it rebuilds them.
 * Lists and results do not depend on it.  No counterpart in the reference. */
int sopg_set_skin(sopg_ctx* ctx, double skin);
/* sopg_run replays repeating step segments (one list-rebuild period, no
 * record, not the run's last step) from captured HIP graphs instead of
 * launching their kernels one by one.  Results do not depend on it (default:
 * enabled; segments with replayed noise or kernel timing always run direct).
 * No counterpart in the reference. */
int sopg_set_graphs(sopg_ctx* ctx, int enable);
/* Segments captured and graph replays so far. */
int sopg_graph_stats(const sopg_ctx* ctx, long* captures, long* launches);
/* Block until all work queued on the context's stream has finished. */
int sopg_sync(sopg_ctx* ctx);

/* Current list sizes of a chain (nnl_att, nnl_rep, nil_att, nil_rep: the
 * reference's counters, global.h:96-101) without changing any state. */
int sopg_list_counts(sopg_ctx* ctx, int chain, int* nnl_att, int* nnl_rep, int* nil_att, int* nil_rep);

/* How the latest list rebuild that planned this chain built its lists (no
 * counterpart in the reference, whose rebuild always scans every master pair):
 *   SOPG_PLAN_NONE      no rebuild was due for the chain in it
 *   SOPG_PLAN_LOCAL     binned on a grid over the chain's extent
 *   SOPG_PLAN_PERIODIC  binned on a grid over the box (a chain wider than the
 *                       box less one grid edge; for the cell list, than ncell
 *                       less two cells)
 *   SOPG_PLAN_SCAN      the full scan (a box under three grid edges, a chain
 *                       beyond the binned rebuild's bead limit, non-finite
 *                       positions, an explicit master list)
 *   SOPG_PLAN_FILTER    the neighbor list filtered from the chain's candidate
 *                       list, which was still valid
 * cells: cells per axis of the grid (0 without one). */
#define SOPG_PLAN_NONE 0
#define SOPG_PLAN_LOCAL 1
#define SOPG_PLAN_PERIODIC 2
#define SOPG_PLAN_SCAN 3
#define SOPG_PLAN_FILTER 4
int sopg_rebuild_plan(sopg_ctx* ctx, int chain, int* plan, int* cells);

/* ------------------------------------------------------------- measurement */
#define SOPG_KERNEL_STEP 0     /* fused force / velocity / next-position kernel */
#define SOPG_KERNEL_REBUILD 1  /* neighbor / cell list rebuild */
#define SOPG_KERNEL_PAIRLIST 2 /* standalone pair-list filter */
#define SOPG_KERNEL_OBSERVE 3  /* observables + reductions */
/* enable != 0: bracket every kernel launched by sopg_run with HIP events on
 * the context's stream and accumulate their durations (reset on enable). */
int sopg_profile(sopg_ctx* ctx, int enable);
int sopg_profile_read(sopg_ctx* ctx, int kernel, double* total_ms, long* launches);

#ifdef __cplusplus
}
#endif
#endif /* SOPG_H */
