"""Parity at the benchmark configurations' own sizes (the workloads of
sop-cpu_amd/workloads.py), against the oracle with the reference's noise
replayed where the configuration allows it:

  plain3000    N=3000, 12 A cell list (boxl 504, ncell 42) at kT 0.6 and 1.2, and
               the neighbor-list variant: 200 steps, lists entry for entry after
               every 40, forces bit-identical, positions within 1e-9 A
  replicas1024 1024 replicas in one batch, replica r drawing from Ran_Gen seeded
               -100-r (`set run r`, params.cpp:43-45); three compared with the oracle
  mixed256     the full 256-chain batch (N up to 4964), 120 steps (6 rebuilds):
               every chain's lists and state equal a full-scan context without
               candidate lists; the smallest, median and largest chains run with
               replayed noise and equal the oracle
  serving      a stream with an N=10^4 folding job (from an unfolded start) and
               N=50 jobs at kT 0.5 and 1.3, arriving while others run: each job's
               result equals the same job run alone
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import sopg
import workloads  # noqa: E402
from conftest import chain_inputs
from refs import Oracle, setup_system, std_cmds
from test_gpu_parity import TRAJ_ABS, check_lists, compare_records, per_bead_rel

pytestmark = pytest.mark.gpu


@pytest.fixture
def pow_cr():
    Oracle.set_pow_mode(1)
    yield
    Oracle.set_pow_mode(0)


class OracleCtrl:
    """underdamped_ctrl's loop body (sop.cpp:153-189) on the oracle, with its counters kept
    across calls, so that the oracle can be compared with the device after every chunk."""

    def __init__(self, orc, nnlup, nup, cutoff="neighborlist"):
        self.orc, self.nnlup, self.nup = orc, nnlup, nup
        self.upd = "update_cell_list" if cutoff == "celllist" else "update_neighbor_list"
        self.inlup, self.iup, self.istep = 1, 1, 1

    def run(self, nsteps):
        o, recs = self.orc, []
        for _ in range(nsteps):
            if self.inlup % self.nnlup == 0:
                o.call(self.upd)
                self.inlup = 0
            self.inlup += 1
            o.call("update_pair_list")
            o.call("underdamped_iteration")
            if self.iup % self.nup == 0:
                o.call("energy_eval")
                o.call("calculate_observables")
                recs.append(dict(istep=self.istep, e_bnd=o.getd("e_bnd"), e_ang_ss=o.getd("e_ang_ss"),
                                 e_vdw_rr=o.getd("e_vdw_rr"), rna_etot=o.getd("rna_etot"), Q=o.getd("Q"),
                                 contct_nat=o.geti("contct_nat"), rgsq=o.getd("rgsq"), kinT=o.getd("kinT")))
                self.iup = 0
            self.iup += 1
            self.istep += 1
        return recs


def _all_lists_equal(ctx, orc, chain=0):
    for w in range(4):
        check_lists(ctx, orc, w, w, chain)


@pytest.mark.parametrize("cutoff,kt", [("celllist", 0.6), ("celllist", 1.2), ("neighborlist", 0.6)])
def test_plain3000(inputs_root, pow_cr, cutoff, kt):
    n, nnlup, nup, steps, chunk = 3000, 20, 50, 200, 40
    seed = 3001 if kt == 0.6 else 3002
    x, v, d = chain_inputs(inputs_root, n, seed, kT=kt)
    w = workloads.WORKLOADS["plain3000" if cutoff == "celllist" else "plain3000nl"]
    orc = Oracle()
    setup_system(orc, d, std_cmds(temp=kt, cutoff=cutoff, nnlup=nnlup, nup=nup,
                                  extra=[("boxl", "504"), ("ncell", "42")]))
    orc.set_coords(2, v)
    assert orc.getd("lcell") == 12.0
    noise = Oracle().gasdev(3 * n * (steps + 1))
    ctx = sopg.Context()
    print(f"[gpu] N=3000 {cutoff} kT={kt}: preamble", flush=True)
    workloads.setup_ctx(ctx, [(x, v, seed)], w, kt, noise_mode=sopg.NOISE_REPLAY, nup=nup,
                        preamble_noise=[noise[:3 * n]])
    orc.call("update_cell_list" if cutoff == "celllist" else "update_neighbor_list")
    orc.call("update_pair_list")
    _all_lists_equal(ctx, orc)
    orc.set_noise(noise)
    orc.call("energy_eval")
    orc.call("force_eval")
    assert np.array_equal(ctx.get_state(0)["force"], orc.coords(3))
    oc = OracleCtrl(orc, nnlup, nup, cutoff)
    drecs, orecs = [], []
    for k in range(steps // chunk):
        q = 3 * n * (1 + k * chunk)
        drecs += ctx.run(chunk, noise[q:q + 3 * n * chunk])
        orecs += oc.run(chunk)
        print(f"[gpu] N=3000 {cutoff} kT={kt}: {(k + 1) * chunk} steps", flush=True)
        _all_lists_equal(ctx, orc)  # rebuilt at steps 20, 40, ... (the last one in this chunk)
        s = ctx.get_state(0)
        assert np.abs(s["unc"] - orc.coords(1)).max() <= TRAJ_ABS
    s = ctx.get_state(0)
    assert np.array_equal(s["unc"], orc.coords(1)), "expected bit-identical trajectory"
    assert np.array_equal(s["force"], orc.coords(3)), per_bead_rel(s["force"], orc.coords(3))
    compare_records(drecs, orecs)
    ctx.close()


def test_replicas1024_reference_streams(tmp_path):
    """1024 replicas in one replayed-noise batch, replica r with the reference's own
    generator stream for `set run r`; replicas 0, 511 and 1023 equal the oracle running
    that generator itself."""
    R, n, steps, nnlup, nup = 1024, 76, 100, 50, 50
    w = workloads.WORKLOADS["replicas1024"]
    chains = workloads.make_chains("replicas1024", 16)
    streams = np.empty((R, steps + 1, 3 * n))
    for r in range(R):
        o = Oracle()
        o.set("run", r)
        streams[r] = o.gasdev(3 * n * (steps + 1)).reshape(steps + 1, 3 * n)
    ctx = sopg.Context()
    print("[gpu] replicas1024: preamble", flush=True)
    workloads.setup_ctx(ctx, chains, w, 0.6, noise_mode=sopg.NOISE_REPLAY, nup=nup,
                        preamble_noise=[streams[r, 0] for r in range(R)])
    print("[gpu] replicas1024: run", flush=True)
    drecs = ctx.run(steps, np.ascontiguousarray(streams[:, 1:].transpose(1, 0, 2)).ravel())
    assert len(drecs) == R * (steps // nup)
    x = chains[0][0]
    d = str(tmp_path / "rep")
    sopg.write_inputs(x, d)
    for r in (0, 511, 1023):
        orc = Oracle()
        setup_system(orc, d, std_cmds(nnlup=nnlup, nup=nup, extra=[("run", str(r))]))
        orc.set_coords(2, chains[r][1])
        orc.call("update_neighbor_list")
        orc.call("update_pair_list")
        orc.call("energy_eval")
        orc.call("force_eval")   # the generator's first 3N deviates, as the device's preamble
        orecs = OracleCtrl(orc, nnlup, nup).run(steps)
        s = ctx.get_state(r)
        assert np.abs(s["unc"] - orc.coords(1)).max() <= TRAJ_ABS, r
        assert per_bead_rel(s["force"], orc.coords(3)) <= 1e-12, r
        compare_records([q for q in drecs if q["chain"] == r], orecs)
    ctx.close()


def test_mixed256_full_batch(tmp_path, pow_cr):
    """The full mixed256 batch for 120 steps (rebuilds at 20, 40, ..., 120): with binned
    rebuilds and candidate lists (the defaults) every chain's state and lists equal a context
    that scans every candidate and keeps no candidate lists; the smallest, median and largest
    chains, run with replayed noise inside the batch, equal the oracle."""
    steps, nup = 120, 40
    w = workloads.WORKLOADS["mixed256"]
    chains = workloads.make_chains("mixed256", 16)
    sizes = np.array([len(x) // 3 for x, _, _ in chains])
    order = np.argsort(sizes, kind="stable")
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    assert sizes[picks[-1]] == 4964
    off = np.concatenate([[0], np.cumsum(sizes)])
    nb = int(off[-1])
    streams = {}
    for c in picks:  # each chain its own reference stream (`set run c`)
        o = Oracle()
        o.set("run", c)
        streams[c] = o.gasdev(3 * int(sizes[c]) * (steps + 1))
    noise = np.zeros((steps, 3 * nb))
    for c in picks:
        noise[:, 3 * off[c]:3 * off[c + 1]] = streams[c][3 * sizes[c]:].reshape(steps, -1)
    noise = noise.ravel()
    topos = [sopg.chain_topology(x) for x, _, _ in chains]
    results = []
    for binned, skin in ((True, None), (False, 0.0)):
        ctx = sopg.Context(binned=binned)
        if skin is not None:
            ctx.set_skin(skin)
        for c, ((x, v, seed), t) in enumerate(zip(chains, topos)):
            mode = sopg.NOISE_REPLAY if c in picks else sopg.NOISE_PHILOX
            ctx.add_chain(t, workloads.params(w, 0.6, seed, mode, nup))
        for c, (x, v, seed) in enumerate(chains):
            ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
            ctx.update_neighbor_list(c)
            ctx.update_pair_list(c)
            ctx.force_eval(c, streams[c][:3 * sizes[c]] if c in picks else None)
        print(f"[gpu] mixed256 binned={binned} skin={skin}: run", flush=True)
        recs = ctx.run(steps, noise)
        results.append(([ctx.get_state(c) for c in range(len(chains))],
                         [[ctx.get_list(k, c) for k in range(4)] for c in range(len(chains))], recs))
        if binned:
            keep = ctx
        else:
            ctx.close()
    (sa, la, ra), (sb, lb, rb) = results
    for c in range(len(chains)):
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(sa[c][k], sb[c][k]), (c, k)
        for (ai, aj), (bi, bj) in zip(la[c], lb[c]):
            assert np.array_equal(ai, bi) and np.array_equal(aj, bj), c
    assert ra == rb
    for c in picks:
        n = int(sizes[c])
        d = str(tmp_path / f"c{c}")
        x, v, _ = chains[c]
        sopg.write_inputs(x, d)
        print(f"[gpu] mixed256 chain {c} (N={n}) on the oracle", flush=True)
        orc = Oracle()
        setup_system(orc, d, std_cmds(nnlup=20, nup=nup))
        orc.set_coords(2, v)
        orc.call("update_neighbor_list")
        orc.call("update_pair_list")
        orc.set_noise(streams[c])
        orc.call("energy_eval")
        orc.call("force_eval")
        orecs = OracleCtrl(orc, 20, nup).run(steps)
        s = sa[c]
        assert np.array_equal(s["unc"], orc.coords(1)), c
        assert np.array_equal(s["force"], orc.coords(3)), c
        compare_records([q for q in ra if q["chain"] == c], orecs)
        _all_lists_equal(keep, orc, c)
        os.remove(os.path.join(d, "vdw.dat"))
    keep.close()


SERVE_JOBS = [  # (N, kind, kT, nsteps, submitted after this many rounds)
    (10000, "fold", 0.9, 2000, 0),
    (50, "unfold", 0.5, 3000, 0),
    (50, "fold", 1.3, 2500, 1),
    (1200, "unfold", 1.3, 1500, 1),
    (3000, "fold", 0.5, 1000, 2),
]


def _job_inputs(j, n, kind, kt):
    nat = sopg.gen_native(n, 40_000 + j)
    x0 = nat if kind == "unfold" else sopg.gen_unfolded(nat, 50_000 + j)
    return x0, sopg.gen_maxwell(n, kt, 40_000 + j), sopg.chain_topology(nat)


def test_served_stream_with_large_and_small_jobs():
    """Jobs arriving while others run, up to N=10^4 folding from an unfolded start and
    N=50 at kT 0.5 and 1.3: each served job's final state and records equal the job alone."""
    w = workloads.SERVE
    with ThreadPoolExecutor(max_workers=8) as ex:
        built = list(ex.map(lambda a: _job_inputs(a[0], *a[1][:3]), enumerate(SERVE_JOBS)))
    params = [workloads.params(w, kt, 60_000 + j, nup=500) for j, (_, _, kt, _, _) in enumerate(SERVE_JOBS)]
    srv = sopg.Server(chunk=700)
    ids = {}
    print("[gpu] serving stream with N up to 10^4", flush=True)
    for rnd in range(1000):
        for j, (n, kind, kt, nsteps, at) in enumerate(SERVE_JOBS):
            if at == rnd:
                x0, v, t = built[j]
                ids[j] = srv.submit(t, params[j], x0, v, nsteps)
        active, queued = srv.poll()
        if len(ids) == len(SERVE_JOBS) and active == 0 and queued == 0:
            break
    info = {j: srv.job(i) for j, i in ids.items()}
    assert all(k["state"] == sopg.JOB_DONE for k in info.values())
    late = [j for j, (*_, at) in enumerate(SERVE_JOBS) if at > 0]
    assert all(info[j]["t_admit"] < info[0]["t_finish"] for j in late)  # admitted while the big job ran
    for j, (n, kind, kt, nsteps, at) in enumerate(SERVE_JOBS):
        x0, v, t = built[j]
        print(f"[gpu] job {j} alone N={n} {kind} kT={kt}", flush=True)
        ctx = sopg.Context()
        ctx.add_chain(t, params[j])
        ctx.put_state(0, pos=x0, unc=x0, vel=v, force=np.zeros_like(x0))
        ctx.update_neighbor_list(0)
        ctx.update_pair_list(0)
        ctx.force_eval(0)
        recs = ctx.run(nsteps)
        st = ctx.get_state(0)
        ctx.close()
        got = srv.state(ids[j])
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(got[k], st[k]), (j, k)
        # bit-equal records, NaN included: a bond stretched past R0 leaves FENE's log1p
        # (energy.cpp:144) without a real value, in the reference as here
        got_r = [dict(r, chain=0) for r in srv.records(ids[j])]
        assert len(got_r) == len(recs) == nsteps // 500
        for a, b in zip(got_r, recs):
            assert a.keys() == b.keys()
            for k in a:
                assert a[k] == b[k] or (a[k] != a[k] and b[k] != b[k]), (j, k, a[k], b[k])
    srv.close()
