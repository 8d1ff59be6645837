"""Continuous-batching server on the GPU: every job's trajectory and records
are bit-identical to the same chain run alone, and admission respects the
HBM budget (jobs wait while the resident ones hold it)."""
import numpy as np
import pytest

import sopg
from conftest import chain_inputs

pytestmark = pytest.mark.gpu

JOBS = [  # (N, seed, nsteps, kT)
    (60, 71, 250, 0.6),
    (210, 72, 1200, 1.2),
    (130, 73, 700, 0.9),
    (95, 74, 1000, 0.5),
    (170, 75, 333, 1.3),
]


def _params(kT, seed):
    return sopg.chain_params(T=kT, cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, nup=100, noise_mode=sopg.NOISE_PHILOX,
                             seed=seed)


def _alone(x, v, kT, seed, nsteps):
    ctx = sopg.Context()
    ctx.add_chain(sopg.chain_topology(x), _params(kT, seed))
    ctx.put_state(0, pos=x, unc=x, vel=v, force=np.zeros_like(x))
    ctx.update_neighbor_list(0)
    ctx.update_pair_list(0)
    ctx.force_eval(0)
    recs = ctx.run(nsteps)
    st = ctx.get_state(0)
    ctx.close()
    return st, recs


def test_served_jobs_equal_alone(inputs_root):
    data = [chain_inputs(inputs_root, n, seed)[:2] for (n, seed, _, _) in JOBS]
    topos = [sopg.chain_topology(x) for x, _ in data]
    probe = sopg.Server(hbm_budget=1 << 30)
    fp = [probe.footprint(t) for t in topos]
    probe.close()
    budget = fp[0] + fp[1] + 0.5 * min(fp)  # two resident jobs at a time, at most
    srv = sopg.Server(hbm_budget=budget, chunk=400)
    ids = []
    for (n, seed, nsteps, kT), (x, v), t in zip(JOBS, data, topos):
        ids.append(srv.submit(t, _params(kT, 900 + seed), x, v, nsteps))
    print("[gpu] serving 5 jobs", flush=True)
    for _ in range(1000):
        active, queued = srv.poll()
        if active == 0 and queued == 0:
            break
    info = [srv.job(j) for j in ids]
    assert all(i["state"] == sopg.JOB_DONE for i in info)
    assert all(i["steps_done"] == nsteps for i, (_, _, nsteps, _) in zip(info, JOBS))
    # the budget held at every admission: resident footprints never exceed it
    for i in info:
        t = i["t_admit"]
        held = sum(k["footprint"] for k in info if k["t_admit"] <= t < k["t_finish"])
        assert held <= budget
    assert max(i["t_admit"] - i["t_arrival"] for i in info) > 0.0  # someone waited
    assert any(i["rounds_waited"] > 0 for i in info)  # ... for memory, across rounds
    for (n, seed, nsteps, kT), (x, v), j in zip(JOBS, data, ids):
        print(f"[gpu] job {j} alone N={n} steps={nsteps}", flush=True)
        st, recs = _alone(x, v, kT, 900 + seed, nsteps)
        got = srv.state(j)
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(got[k], st[k]), (j, k)
        chain = srv.records(j)[0]["chain"] if recs else 0
        assert srv.records(j) == [dict(r, chain=chain) for r in recs]
        assert len(recs) == nsteps // 100
    srv.close()


def test_evicted_job_resumes_exactly(inputs_root):
    """A job preempted mid-run (checkpointed, its memory freed, requeued) and
    later resumed ends bit-identically to the same job run alone."""
    specs = [(80, 91, 900, 0.7), (150, 92, 1500, 1.0), (60, 93, 700, 0.8)]
    data = [chain_inputs(inputs_root, n, seed)[:2] for (n, seed, _, _) in specs]
    srv = sopg.Server(chunk=335)  # two rounds end between list rebuilds (670 % 20 != 0)
    ids = [srv.submit(sopg.chain_topology(x), _params(kT, 700 + seed), x, v, nsteps)
           for (n, seed, nsteps, kT), (x, v) in zip(specs, data)]
    print("[gpu] evict and resume", flush=True)
    srv.poll()
    srv.poll()  # 670 steps done
    srv.evict(ids[1])
    assert srv.job(ids[1])["state"] == sopg.JOB_QUEUED and srv.job(ids[1])["steps_done"] == 670
    for _ in range(100):
        active, queued = srv.poll()
        if active == 0 and queued == 0:
            break
    for (n, seed, nsteps, kT), (x, v), j in zip(specs, data, ids):
        st, recs = _alone(x, v, kT, 700 + seed, nsteps)
        got = srv.state(j)
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(got[k], st[k]), (j, k)
        assert [dict(r, chain=0) for r in srv.records(j)] == recs
    srv.close()
