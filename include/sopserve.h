/*
 * sopserve.h -- continuous-batching job server over one sopg context.
 *
 * No counterpart in SOP-CPU, which runs one chain per process
 * (sop.cpp:20-60: main -> ex_cmds -> simulation_ctrl).  A job is one such
 * simulation: a chain (topology, parameters, initial coordinates and
 * velocities) advanced by nsteps steps of underdamped_ctrl (sop.cpp:119-189),
 * with its energy/Q/Rg record every nup steps streamed back as it is made.
 *
 * Jobs queue in arrival order.  A scheduling round (sopg_server_poll) admits
 * queued jobs while their device footprint fits the HBM budget and a chain
 * slot is free, advances every active job together by one chunk of steps
 * (sopg_run over the whole batch), collects records, and retires finished
 * jobs, whose memory the next round reuses.  Every job's results are those
 * of the same chain run alone (sopg_run is per-chain deterministic and
 * independent of the batch it runs in).
 *
 * Times are seconds on the server's steady clock (sopg_server_now).
 */
#ifndef SOPSERVE_H
#define SOPSERVE_H

#include "sopg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sopg_server sopg_server;

typedef struct {
  double hbm_budget;  /* bytes of device memory jobs may hold together (0: 64 GiB) */
  int max_chains;     /* jobs resident at once (0: 4096) */
  long chunk;         /* steps per scheduling round, at most (0: 1000) */
  int lanes;          /* step-kernel lanes per bead (0: autotuned as the resident set changes) */
  int capacity;       /* per-bead non-native list capacity (0: library default) */
} sopg_server_config;

#define SOPG_JOB_QUEUED 0
#define SOPG_JOB_ACTIVE 1
#define SOPG_JOB_DONE 2
#define SOPG_JOB_REJECTED 3 /* larger than the whole budget */

typedef struct {
  int state;                  /* SOPG_JOB_* */
  int chain;                  /* chain id while active, else -1 */
  long nbead, nsteps, steps_done;
  double t_arrival, t_admit, t_finish;
  double footprint;           /* bytes charged against the budget */
  long nrec;                  /* records collected so far */
  long rounds_waited;         /* scheduling rounds that ended with the job still queued */
} sopg_job_info;

int sopg_server_create(int device, const sopg_server_config* cfg, sopg_server** out);
int sopg_server_destroy(sopg_server* srv);
const char* sopg_server_last_error(const sopg_server* srv);
double sopg_server_now(const sopg_server* srv);
/* Device bytes a job of this topology is charged (and holds while active). */
double sopg_server_footprint(const sopg_server* srv, const sopg_topology* topo);
/* Queue a job.  The topology and initial state are copied.  pos: 3N
 * coordinates (used as pos and unc_pos), vel: 3N velocities.  The chain must
 * use the device noise generator (SOPG_NOISE_PHILOX).  arrival: the job's
 * arrival time on the server clock (< 0: now); latency counts from it. */
int sopg_server_submit(sopg_server* srv, const sopg_topology* topo, const sopg_chain_params* p,
                       const double* pos, const double* vel, long nsteps, double arrival, long* job);
/* One scheduling round (admit, advance one chunk, collect, retire).
 * *active / *queued: jobs resident / waiting afterwards. */
int sopg_server_poll(sopg_server* srv, int* active, int* queued);
int sopg_server_job(const sopg_server* srv, long job, sopg_job_info* info);
/* The job's records so far (up to cap); *n = count. */
int sopg_server_job_records(const sopg_server* srv, long job, sopg_record* rec, long cap, long* n);
/* Final pos, unc_pos, vel, force of a finished job (3N each; any may be NULL). */
int sopg_server_job_state(const sopg_server* srv, long job, double* pos, double* unc, double* vel, double* force);
/* Preempt a resident job: checkpoint it losslessly (sopg_checkpoint), free its
 * device memory and put it at the head of the queue.  When it is admitted
 * again it continues exactly where it stopped. */
int sopg_server_evict(sopg_server* srv, long job);
/* Drop the host copy of a finished job's results. */
int sopg_server_release(sopg_server* srv, long job);
/* Totals: bead-steps advanced, rounds run, seconds spent inside rounds. */
int sopg_server_stats(const sopg_server* srv, double* bead_steps, long* rounds, double* busy_s);

#ifdef __cplusplus
}
#endif
#endif /* SOPSERVE_H */
