// sop_serve.cpp -- continuous-batching job server (include/sopserve.h).
//
// One sopg context holds every resident job as one chain of a ragged batch.
// A round: admit queued jobs in arrival order while the HBM budget and the
// chain slots allow (stopping at the first that does not fit, so no job is
// overtaken forever), run the underdamped_ctrl preamble for each newcomer
// (sop.cpp:119-147: list build, pair list, first force), advance the whole
// batch by one chunk with sopg_run, file the records by job, and retire
// finished jobs.  The chunk never runs a job past its last step: it is the
// smallest remaining step count among resident jobs, capped by the config.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/sopserve.h"

namespace {

struct Job {
  int state = SOPG_JOB_QUEUED;
  int chain = -1;
  long nbead = 0, nsteps = 0, done = 0;
  long waited = 0;  // rounds that ended with this job still queued
  double t_arrival = 0, t_admit = -1, t_finish = -1, footprint = 0;
  sopg_chain_params p{};
  // topology copy
  std::vector<int> ibnd, jbnd, iang, jang, kang, ai, aj, ait, ajt, ri, rj, rit, rjt, btype;
  std::vector<double> r0, ar;
  int ncon_rep = -1;
  bool has_btype = false;
  std::vector<double> pos, vel;  // initial state, then final pos
  std::vector<double> unc, fvel, force;
  std::vector<sopg_record> rec;
  std::vector<char> ckpt;  // set while the job is evicted: resumes from here
  long evictions = 0;
  sopg_topology topo() const {
    sopg_topology t{};
    t.nbead = (int)nbead;
    t.nbnd = (int)ibnd.size(); t.ibnd = ibnd.data(); t.jbnd = jbnd.data(); t.r0bnd = r0.data();
    t.nang = (int)iang.size(); t.iang = iang.data(); t.jang = jang.data(); t.kang = kang.data();
    t.ncon_att = (int)ai.size(); t.att_i = ai.data(); t.att_j = aj.data(); t.att_it = ait.data();
    t.att_jt = ajt.data(); t.att_r = ar.data();
    t.ncon_rep = ncon_rep;
    if (ncon_rep >= 0) { t.rep_i = ri.data(); t.rep_j = rj.data(); t.rep_it = rit.data(); t.rep_jt = rjt.data(); }
    t.bead_type = has_btype ? btype.data() : nullptr;
    return t;
  }
};

template <class T>
std::vector<T> copy(const T* p, long n) {
  return (p && n > 0) ? std::vector<T>(p, p + n) : std::vector<T>();
}

}  // namespace

struct sopg_server {
  sopg_ctx* ctx = nullptr;
  sopg_server_config cfg{};
  int cap = 1024;
  std::string err;
  std::chrono::steady_clock::time_point t0;
  std::vector<Job> jobs;
  std::deque<long> queue;
  std::vector<long> by_chain;  // chain id -> job
  double used = 0;
  int resident = 0;
  double bead_steps = 0, busy = 0;
  long rounds = 0;
  // step-kernel lanes per bead by resident-bead bucket (floor(log2(beads))):
  // each bucket is autotuned once, the first time the batch falls into it
  std::map<int, int> lanes_by_bucket;
  std::vector<sopg_record> recbuf;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

int sfail(sopg_server* s, int code, const std::string& m) {
  s->err = m;
  return code;
}

int ctx_fail(sopg_server* s, int rc, const char* what) {
  return sfail(s, rc, std::string(what) + ": " + sopg_last_error(s->ctx));
}

// Bytes the context's pools hold for a chain: per bead, per bond, angle,
// native and explicit non-native entry (both incidences included), plus the
// chain slot and its binning grid.
double footprint(int cap, const sopg_topology* t) {
  const double per_bead = 12 * 8 + 15 * 8 + 12 * 4 + 1 + 8.0 * cap + 12 * 8;
  const double per_chain = 512 + 3.0 * 32 * 32 * 32 * 4 + 12 * 8;
  const double nrep = t->ncon_rep > 0 ? t->ncon_rep : 0;
  const double per_native = 4 + 4 + 4 + 1 + 8 + 8 + 4 + 2 * (4 + 4 + 1 + 1 + 8 + 8 + 8);
  return per_bead * t->nbead + 24.0 * t->nbnd + 8.0 * t->nang + per_native * t->ncon_att + 17.0 * nrep + per_chain;
}

int admit(sopg_server* s, Job& j, long id) {
  const sopg_topology t = j.topo();
  int chain = -1;
  int rc = sopg_add_chain(s->ctx, &t, &j.p, &chain);
  if (rc) return ctx_fail(s, rc, "sopg_add_chain");
  j.chain = chain;
  if (!j.ckpt.empty()) {
    // an evicted job continues from its checkpoint
    if ((rc = sopg_restore(s->ctx, chain, j.ckpt.data(), (long)j.ckpt.size()))) return ctx_fail(s, rc, "sopg_restore");
    j.ckpt.clear();
    j.ckpt.shrink_to_fit();
    if ((size_t)chain >= s->by_chain.size()) s->by_chain.resize(chain + 1, -1);
    s->by_chain[chain] = id;
    j.state = SOPG_JOB_ACTIVE;
    s->used += j.footprint;
    s->resident++;
    return SOPG_OK;
  }
  const std::vector<double> zero(3 * j.nbead, 0.0);
  if ((rc = sopg_put_state(s->ctx, chain, j.pos.data(), j.pos.data(), j.vel.data(), zero.data())))
    return ctx_fail(s, rc, "sopg_put_state");
  // underdamped_ctrl preamble (sop.cpp:119-147)
  if (j.p.cutoff == SOPG_CUTOFF_CELL) rc = sopg_update_cell_list(s->ctx, chain, nullptr, nullptr);
  else if (j.p.cutoff == SOPG_CUTOFF_NEIGHBOR) rc = sopg_update_neighbor_list(s->ctx, chain, nullptr, nullptr);
  if (rc) return ctx_fail(s, rc, "list build");
  if ((rc = sopg_update_pair_list(s->ctx, chain, nullptr, nullptr))) return ctx_fail(s, rc, "sopg_update_pair_list");
  if ((rc = sopg_force_eval(s->ctx, chain, nullptr))) return ctx_fail(s, rc, "sopg_force_eval");
  if ((size_t)chain >= s->by_chain.size()) s->by_chain.resize(chain + 1, -1);
  s->by_chain[chain] = id;
  j.chain = chain;
  j.state = SOPG_JOB_ACTIVE;
  j.t_admit = s->now();
  s->used += j.footprint;
  s->resident++;
  return SOPG_OK;
}

int retire(sopg_server* s, Job& j) {
  const size_t n3 = 3 * (size_t)j.nbead;
  j.unc.resize(n3);
  j.fvel.resize(n3);
  j.force.resize(n3);
  int rc = sopg_get_state(s->ctx, j.chain, j.pos.data(), j.unc.data(), j.fvel.data(), j.force.data());
  if (rc) return ctx_fail(s, rc, "sopg_get_state");
  if ((rc = sopg_remove_chain(s->ctx, j.chain))) return ctx_fail(s, rc, "sopg_remove_chain");
  s->by_chain[j.chain] = -1;
  j.chain = -1;
  j.state = SOPG_JOB_DONE;
  j.t_finish = s->now();
  s->used -= j.footprint;
  s->resident--;
  j.vel.clear();
  j.vel.shrink_to_fit();
  return SOPG_OK;
}

const Job* find(const sopg_server* s, long id) {
  return (s && id >= 0 && id < (long)s->jobs.size()) ? &s->jobs[id] : nullptr;
}

}  // namespace

extern "C" {

int sopg_server_create(int device, const sopg_server_config* cfg, sopg_server** out) {
  if (!out) return SOPG_ERR_ARG;
  *out = nullptr;
  sopg_server* s = new sopg_server();
  if (cfg) s->cfg = *cfg;
  if (s->cfg.hbm_budget <= 0) s->cfg.hbm_budget = 64.0 * (1ull << 30);
  if (s->cfg.max_chains <= 0) s->cfg.max_chains = 4096;
  if (s->cfg.chunk <= 0) s->cfg.chunk = 1000;
  int rc = sopg_create(device, &s->ctx);
  if (rc) {
    delete s;
    return rc;
  }
  if (s->cfg.lanes > 0 && (rc = sopg_set_lanes(s->ctx, s->cfg.lanes, 0))) {
    sopg_destroy(s->ctx);
    delete s;
    return rc;
  }
  if (s->cfg.capacity > 0) {
    sopg_set_capacity(s->ctx, s->cfg.capacity);
    s->cap = s->cfg.capacity;
  }
  // size the bead pool for the budget once, so admissions do not regrow it
  const double per_bead = 12 * 8 + 15 * 8 + 12 * 4 + 1 + 8.0 * s->cap + 12 * 8;
  const long beads = (long)std::min(s->cfg.hbm_budget / per_bead, 4.0e7);
  sopg_reserve(s->ctx, beads, std::min(s->cfg.max_chains, 65536));
  s->t0 = std::chrono::steady_clock::now();
  *out = s;
  return SOPG_OK;
}

int sopg_server_destroy(sopg_server* s) {
  if (!s) return SOPG_OK;
  sopg_destroy(s->ctx);
  delete s;
  return SOPG_OK;
}

const char* sopg_server_last_error(const sopg_server* s) { return s ? s->err.c_str() : "null server"; }

double sopg_server_now(const sopg_server* s) { return s ? s->now() : 0.0; }

double sopg_server_footprint(const sopg_server* s, const sopg_topology* t) {
  return (s && t) ? footprint(s->cap, t) : 0.0;
}

int sopg_server_submit(sopg_server* s, const sopg_topology* t, const sopg_chain_params* p, const double* pos,
                       const double* vel, long nsteps, double arrival, long* job) {
  if (!s || !t || !p || !pos || !vel) return SOPG_ERR_ARG;
  if (nsteps < 1) return sfail(s, SOPG_ERR_ARG, "nsteps must be >= 1");
  if (p->noise_mode != SOPG_NOISE_PHILOX)
    return sfail(s, SOPG_ERR_ARG, "served jobs use the device noise generator (SOPG_NOISE_PHILOX)");
  if (t->nbead < 1) return sfail(s, SOPG_ERR_ARG, "nbead < 1");
  Job j;
  j.nbead = t->nbead;
  j.nsteps = nsteps;
  j.p = *p;
  j.t_arrival = arrival < 0 ? s->now() : arrival;
  j.ibnd = copy(t->ibnd, t->nbnd); j.jbnd = copy(t->jbnd, t->nbnd); j.r0 = copy(t->r0bnd, t->nbnd);
  j.iang = copy(t->iang, t->nang); j.jang = copy(t->jang, t->nang); j.kang = copy(t->kang, t->nang);
  j.ai = copy(t->att_i, t->ncon_att); j.aj = copy(t->att_j, t->ncon_att);
  j.ait = copy(t->att_it, t->ncon_att); j.ajt = copy(t->att_jt, t->ncon_att); j.ar = copy(t->att_r, t->ncon_att);
  j.ncon_rep = t->ncon_rep;
  if (t->ncon_rep >= 0) {
    j.ri = copy(t->rep_i, t->ncon_rep); j.rj = copy(t->rep_j, t->ncon_rep);
    j.rit = copy(t->rep_it, t->ncon_rep); j.rjt = copy(t->rep_jt, t->ncon_rep);
  }
  if (t->bead_type) { j.btype = copy(t->bead_type, t->nbead); j.has_btype = true; }
  j.pos = copy(pos, 3L * t->nbead);
  j.vel = copy(vel, 3L * t->nbead);
  j.footprint = footprint(s->cap, t);
  const long id = (long)s->jobs.size();
  if (j.footprint > s->cfg.hbm_budget) {
    j.state = SOPG_JOB_REJECTED;
    s->jobs.push_back(std::move(j));
    if (job) *job = id;
    return sfail(s, SOPG_ERR_ARG, "job footprint exceeds the server's HBM budget");
  }
  s->jobs.push_back(std::move(j));
  s->queue.push_back(id);
  if (job) *job = id;
  return SOPG_OK;
}

int sopg_server_poll(sopg_server* s, int* active, int* queued) {
  if (!s) return SOPG_ERR_ARG;
  const double t_in = s->now();
  int rc;
  // admit in arrival order while the head fits
  while (!s->queue.empty()) {
    Job& j = s->jobs[s->queue.front()];
    if (s->resident >= s->cfg.max_chains || s->used + j.footprint > s->cfg.hbm_budget) break;
    const long id = s->queue.front();
    s->queue.pop_front();
    if ((rc = admit(s, j, id))) {
      // a chain the context cannot hold (e.g. a no-cutoff chain above the
      // list capacity) is rejected; a device error ends the round
      if (j.chain >= 0) sopg_remove_chain(s->ctx, j.chain);
      j.chain = -1;
      j.state = SOPG_JOB_REJECTED;
      if (rc == SOPG_ERR_HIP) return rc;
    }
  }
  for (long id : s->queue) s->jobs[id].waited++;
  if (s->resident > 0) {
    long chunk = s->cfg.chunk;
    double beads = 0;
    for (long id : s->by_chain)
      if (id >= 0) {
        const Job& j = s->jobs[id];
        chunk = std::min(chunk, j.nsteps - j.done);
        beads += (double)j.nbead;
      }
    // the fastest launch shape depends on the batch size: a bucket of resident
    // bead counts (a factor of two wide) is autotuned once and then reused
    if (s->cfg.lanes <= 0) {
      const int bucket = beads >= 1.0 ? (int)std::floor(std::log2(beads)) : 0;
      auto it = s->lanes_by_bucket.find(bucket);
      if (it == s->lanes_by_bucket.end()) {
        int lanes = 0;
        if ((rc = sopg_autotune(s->ctx, &lanes))) return ctx_fail(s, rc, "sopg_autotune");
        s->lanes_by_bucket[bucket] = lanes;
      } else if ((rc = sopg_set_lanes(s->ctx, it->second, 0))) {
        return ctx_fail(s, rc, "sopg_set_lanes");
      }
    }
    long nrec_max = 0;
    for (long id : s->by_chain)
      if (id >= 0) nrec_max += chunk / s->jobs[id].p.nup + 1;
    if ((long)s->recbuf.size() < nrec_max) s->recbuf.resize(nrec_max);
    long nrec = 0;
    if ((rc = sopg_run(s->ctx, chunk, nullptr, s->recbuf.data(), (long)s->recbuf.size(), &nrec)))
      return ctx_fail(s, rc, "sopg_run");
    for (long k = 0; k < std::min(nrec, (long)s->recbuf.size()); k++) {
      const sopg_record& r = s->recbuf[k];
      s->jobs[s->by_chain[r.chain]].rec.push_back(r);
    }
    s->bead_steps += beads * (double)chunk;
    for (size_t c = 0; c < s->by_chain.size(); c++) {
      const long id = s->by_chain[c];
      if (id < 0) continue;
      Job& j = s->jobs[id];
      j.done += chunk;
      if (j.done >= j.nsteps && (rc = retire(s, j))) return rc;
    }
    s->rounds++;
  }
  s->busy += s->now() - t_in;
  if (active) *active = s->resident;
  if (queued) *queued = (int)s->queue.size();
  return SOPG_OK;
}

int sopg_server_job(const sopg_server* s, long id, sopg_job_info* info) {
  const Job* j = find(s, id);
  if (!j || !info) return SOPG_ERR_ARG;
  info->state = j->state;
  info->chain = j->chain;
  info->nbead = j->nbead;
  info->nsteps = j->nsteps;
  info->steps_done = j->done;
  info->t_arrival = j->t_arrival;
  info->t_admit = j->t_admit;
  info->t_finish = j->t_finish;
  info->footprint = j->footprint;
  info->nrec = (long)j->rec.size();
  info->rounds_waited = j->waited;
  return SOPG_OK;
}

int sopg_server_job_records(const sopg_server* s, long id, sopg_record* rec, long cap, long* n) {
  const Job* j = find(s, id);
  if (!j) return SOPG_ERR_ARG;
  if (n) *n = (long)j->rec.size();
  if (rec)
    for (long k = 0; k < std::min(cap, (long)j->rec.size()); k++) rec[k] = j->rec[k];
  return SOPG_OK;
}

int sopg_server_job_state(const sopg_server* s, long id, double* pos, double* unc, double* vel, double* force) {
  const Job* j = find(s, id);
  if (!j) return SOPG_ERR_ARG;
  if (j->state != SOPG_JOB_DONE || j->unc.empty()) return SOPG_ERR_ARG;
  const size_t b = sizeof(double) * 3 * (size_t)j->nbead;
  if (pos) std::memcpy(pos, j->pos.data(), b);
  if (unc) std::memcpy(unc, j->unc.data(), b);
  if (vel) std::memcpy(vel, j->fvel.data(), b);
  if (force) std::memcpy(force, j->force.data(), b);
  return SOPG_OK;
}

int sopg_server_release(sopg_server* s, long id) {
  if (!find(s, id)) return SOPG_ERR_ARG;
  Job& j = s->jobs[id];
  if (j.state != SOPG_JOB_DONE && j.state != SOPG_JOB_REJECTED) return sfail(s, SOPG_ERR_ARG, "job not finished");
  Job light;
  light.state = j.state;
  light.nbead = j.nbead;
  light.nsteps = j.nsteps;
  light.done = j.done;
  light.t_arrival = j.t_arrival;
  light.t_admit = j.t_admit;
  light.t_finish = j.t_finish;
  light.footprint = j.footprint;
  light.waited = j.waited;
  j = std::move(light);
  return SOPG_OK;
}

int sopg_server_evict(sopg_server* s, long id) {
  if (!find(s, id)) return SOPG_ERR_ARG;
  Job& j = s->jobs[id];
  if (j.state != SOPG_JOB_ACTIVE) return sfail(s, SOPG_ERR_ARG, "only a resident job can be evicted");
  long n = 0;
  int rc = sopg_checkpoint(s->ctx, j.chain, nullptr, 0, &n);
  if (rc) return ctx_fail(s, rc, "sopg_checkpoint");
  j.ckpt.resize((size_t)n);
  if ((rc = sopg_checkpoint(s->ctx, j.chain, j.ckpt.data(), n, &n))) return ctx_fail(s, rc, "sopg_checkpoint");
  if ((rc = sopg_remove_chain(s->ctx, j.chain))) return ctx_fail(s, rc, "sopg_remove_chain");
  s->by_chain[j.chain] = -1;
  j.chain = -1;
  j.state = SOPG_JOB_QUEUED;
  j.evictions++;
  s->used -= j.footprint;
  s->resident--;
  s->queue.push_front(id);
  return SOPG_OK;
}

int sopg_server_stats(const sopg_server* s, double* bead_steps, long* rounds, double* busy_s) {
  if (!s) return SOPG_ERR_ARG;
  if (bead_steps) *bead_steps = s->bead_steps;
  if (rounds) *rounds = s->rounds;
  if (busy_s) *busy_s = s->busy;
  return SOPG_OK;
}

}  // extern "C"
