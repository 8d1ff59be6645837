#!/usr/bin/env python3
"""Throughput benchmark of the SOP Langevin hot path on one MI355X.

Default workload (mixed256): 256 independent chains, N drawn log-uniform in
[50, 5000] from seed 4, each its own synthetic native structure (seeded SAW +
MC collapse into 3.0*N^(1/3) A), all advanced together in one ragged batch:
FP64, distance neighbor list rebuilt every 20 steps, pair list every step,
underdamped Langevin h=0.05, zeta=50, kT=0.6, records (energy + observables)
every 1000 steps, device Philox noise.  Other workloads: --workload plain3000
(one N=3000 chain, 12 A cell list, rebuild every 20 steps) and replicas1024
(1024 trajectories of one N=76 chain, rebuild every 50 steps).

serve: a Poisson stream of jobs (20 jobs/s; N log-uniform in [50, 10000];
steps log-uniform in [1e4, 1e6] scaled by --serve-step-scale (default 0.1)
so the stream finishes in seconds, rounded to a multiple of 1000; kT uniform
in [0.5, 1.3]; half folding from a SAW, half unfolding from native) replayed
in real time against
the continuous-batching server (include/sopserve.h); p50/p99 job latency =
finish - arrival, throughput = job bead-steps / (last finish - first arrival).

One step = the reference's underdamped_ctrl loop body (sop.cpp:153-189) for
every chain.  value = bead-steps/s = sum(N) * K / wall seconds of exactly K
timed steps of the batched form, inputs resident in HBM.  plain = the same
chains advanced one chain per context (one chain per launch), K_plain steps
each; speedup_vs_plain = value / plain.

The timed run replays repeating step segments from captured HIP graphs (the
library default; --no-graphs launches every kernel directly).

roofline: the fused step kernel (SOPG_KERNEL_STEP), timed with HIP events on
the library's stream over a second run of the same batch (min(steps, 400)
steps, direct launches); algorithmic bytes per launch =
sum over chains of 192 N + 16 nbnd + 8 nang + 40 nnl_att + 8 nnl_rep, against
8 TB/s HBM.

cpu_baseline: the reference itself (oracle/_ref, compiled from SOP-CPU's
sources with its own flags) as --cpu-workers single-thread processes at once
(one chain each, at evenly spaced size quantiles), on a bounded sample of the
same workload (named in `sample`).
"""
import argparse
import json
import os
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sop-cpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import sopg  # noqa: E402

HBM_PEAK_GBS = 8000.0
KT = 0.6
NUP = 1000

WORKLOADS = {
    "mixed256": dict(desc="256 chains, N log-uniform in [50, 5000] (seed 4), each its own native structure; "
                          "neighbor list every 20 steps; kT {kT}, h 0.05, zeta 50; Philox noise",
                     cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, boxl=500.0, ncell=55.0),
    "plain3000": dict(desc="one chain N=3000; cell list 12 A (boxl 504, ncell 42) rebuilt every 20 steps; "
                           "kT {kT}, h 0.05, zeta 50; Philox noise",
                      cutoff=sopg.CUTOFF_CELL, nnlup=20, boxl=504.0, ncell=42.0),
    "replicas1024": dict(desc="1024 replicas of one N=76 chain, own seeds and Maxwell velocities; "
                              "neighbor list every 50 steps; kT {kT}, h 0.05, zeta 50; Philox noise",
                         cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=50, boxl=500.0, ncell=55.0),
}


def chain_sizes(workload):
    if workload == "mixed256":
        rng = np.random.default_rng(4)
        return [int(v) for v in np.round(np.exp(rng.uniform(np.log(50), np.log(5000), 256)))]
    if workload == "plain3000":
        return [3000]
    return [76] * 1024


def make_chains(workload, threads, kt=KT):
    sizes = chain_sizes(workload)
    if workload == "replicas1024":
        x = sopg.gen_native(76, 76)
        return [(x, sopg.gen_maxwell(76, kt, 10_000 + r), 10_000 + r) for r in range(len(sizes))]

    def one(c):
        n = sizes[c]
        return (sopg.gen_native(n, 1000 + c), sopg.gen_maxwell(n, kt, 1000 + c), 1000 + c)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, range(len(sizes))))


def build_ctx(chains, w, lanes=0, list_lanes=0, topo_cache=None):
    ctx = sopg.Context(device=0, lanes=lanes, list_lanes=list_lanes)
    topos = []
    for (x, v, seed) in chains:
        key = id(x)
        t = topo_cache.get(key) if topo_cache is not None else None
        if t is None:
            t = sopg.chain_topology(x)
            if topo_cache is not None:
                topo_cache[key] = t
        topos.append(t)
        ctx.add_chain(t, sopg.chain_params(T=w.get("kT", KT), boxl=w["boxl"], ncell=w["ncell"], cutoff=w["cutoff"],
                                           nnlup=w["nnlup"], nup=NUP, noise_mode=sopg.NOISE_PHILOX, seed=seed))
    for c, (x, v, seed) in enumerate(chains):
        ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
    # underdamped_ctrl preamble (sop.cpp:119-147): lists, energy, first force (untimed)
    for c in range(len(chains)):
        if w["cutoff"] == sopg.CUTOFF_CELL:
            ctx.update_cell_list(c)
        else:
            ctx.update_neighbor_list(c)
        ctx.update_pair_list(c)
        ctx.force_eval(c)
    return ctx, topos


def timed_run(ctx, steps, warmup, profile, autotune=False):
    if warmup:
        ctx.run(warmup)
    if autotune:
        ctx.autotune()
    ctx.sync()
    if profile:
        ctx.profile(True)
    t0 = time.perf_counter()
    recs = ctx.run(steps)
    ctx.sync()
    t1 = time.perf_counter()
    prof = {k: ctx.profile_read(k) for k in range(4)} if profile else None
    if profile:
        ctx.profile(False)
    return t1 - t0, prof, recs


def gpu_batched(chains, w, steps, warmup, lanes, list_lanes, topo_cache, graphs=True, resident=(0, 0), skin=None):
    ctx, topos = build_ctx(chains, w, lanes, list_lanes, topo_cache)
    ctx.set_graphs(graphs)
    ctx.set_resident(*resident)
    if skin is not None:
        ctx.set_skin(skin)
    # value: the production path (graph replay on, no per-kernel timing)
    seconds, _, recs = timed_run(ctx, steps, warmup, False, autotune=not lanes)
    step_lanes = ctx.lanes()[0]
    graph_stats = ctx.graph_stats()
    # roofline: the same batch, each kernel bracketed by HIP events (direct launches)
    _, prof, _ = timed_run(ctx, min(steps, 400), 0, True)
    bytes_step = 0
    counts = []
    for c, t in enumerate(topos):
        k = ctx.list_counts(c)
        counts.append(k)
        bytes_step += (192 * t["n"] + 16 * len(t["ib"]) + 8 * len(t["ia"]) + 40 * k["nnl_att"] + 8 * k["nnl_rep"])
    ctx.close()
    return dict(seconds=seconds, prof=prof, bytes_step=bytes_step, counts=counts, records=len(recs),
                lanes=step_lanes, graphs=graph_stats)


def gpu_plain(chains, w, steps, warmup, topo_cache, lanes=0, list_lanes=0):
    """One chain per context: every launch advances a single chain."""
    total_bs, total_t = 0.0, 0.0
    for ch in chains:
        ctx, topos = build_ctx([ch], w, lanes, list_lanes, topo_cache)
        seconds, _, _ = timed_run(ctx, steps, warmup, False, autotune=not lanes)
        total_bs += topos[0]["n"] * steps
        total_t += seconds
        ctx.close()
    return total_bs / total_t, total_t


def serve_jobs(njobs, step_scale, threads, seed=5, rate=20.0):
    """The serving stream: Poisson arrivals (rate jobs/s), N log-uniform in
    [50, 10000], steps log-uniform in [1e4, 1e6] times step_scale and rounded
    to a multiple of 1000 (the record period), kT uniform in [0.5, 1.3].
    Each job has its own native structure (the topology); unfolding jobs
    (even ids) start from it, folding jobs (odd ids) from a self-avoiding walk."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, njobs))
    sizes = np.round(np.exp(rng.uniform(np.log(50), np.log(10000), njobs))).astype(int)
    raw = np.exp(rng.uniform(np.log(1e4), np.log(1e6), njobs)) * step_scale
    steps = np.maximum(1000, np.round(raw / 1000.0) * 1000).astype(int)
    kts = rng.uniform(0.5, 1.3, njobs)

    def one(j):
        n = int(sizes[j])
        nat = sopg.gen_native(n, 20_000 + j)
        x0 = nat if j % 2 == 0 else sopg.gen_saw(n, 30_000 + j)
        return x0, sopg.gen_maxwell(n, float(kts[j]), 20_000 + j), sopg.chain_topology(nat)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        built = list(ex.map(one, range(njobs)))
    return [dict(arrival=float(arrivals[j]), n=int(sizes[j]), steps=int(steps[j]), kT=float(kts[j]), x=built[j][0],
                 v=built[j][1], topo=built[j][2], seed=20_000 + j, kind="unfold" if j % 2 == 0 else "fold")
            for j in range(njobs)]


def gpu_serve(jobs, w, hbm_budget, max_chains, chunk, lanes):
    """Replay the arrival stream against the continuous-batching server in real
    time; latency = finish - arrival of each job."""
    srv = sopg.Server(hbm_budget=hbm_budget, max_chains=max_chains, chunk=chunk, lanes=lanes)
    t0 = srv.now()
    i, active, queued = 0, 0, 0
    ids = []
    peak_active = 0
    while i < len(jobs) or active or queued:
        now = srv.now() - t0
        while i < len(jobs) and jobs[i]["arrival"] <= now:
            jb = jobs[i]
            p = sopg.chain_params(T=jb["kT"], boxl=w["boxl"], ncell=w["ncell"], cutoff=w["cutoff"],
                                  nnlup=w["nnlup"], nup=NUP, noise_mode=sopg.NOISE_PHILOX, seed=jb["seed"])
            ids.append(srv.submit(jb["topo"], p, jb["x"], jb["v"], jb["steps"], arrival=t0 + jb["arrival"]))
            queued += 1
            i += 1
        if active == 0 and queued == 0:
            time.sleep(max(0.0, min(0.01, jobs[i]["arrival"] - (srv.now() - t0))))
            continue
        active, queued = srv.poll()
        peak_active = max(peak_active, active)
    info = [srv.job(j) for j in ids]
    st = srv.stats()
    nrec = sum(len(srv.records(j)) for j in ids)
    srv.close()
    done = [k for k in info if k["state"] == sopg.JOB_DONE]
    offered = sum(j["n"] * j["steps"] for j in jobs) / (jobs[-1]["arrival"] - jobs[0]["arrival"])
    lat = np.array([k["t_finish"] - k["t_arrival"] for k in done])
    wait = np.array([k["t_admit"] - k["t_arrival"] for k in done])
    span = max(k["t_finish"] for k in done) - (t0 + jobs[0]["arrival"])
    work = sum(k["nbead"] * k["nsteps"] for k in done)
    return dict(value=work / span, unit="bead-steps/s", jobs=len(jobs), done=len(done),
                rejected=sum(k["state"] == sopg.JOB_REJECTED for k in info),
                p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                p50_wait_s=float(np.percentile(wait, 50)), p99_wait_s=float(np.percentile(wait, 99)),
                makespan_s=span, rounds=st["rounds"], busy_s=st["busy_s"], peak_resident=peak_active,
                records=nrec, offered_load=offered,
                jobs_waited_for_capacity=sum(k["rounds_waited"] > 0 for k in done),
                max_rounds_waited=max(k["rounds_waited"] for k in done))


def _cpu_worker(job):
    """One reference process: set up one chain from its files and run the
    underdamped_ctrl loop body (sop.cpp:153-189) until the time budget is
    spent (at least one rebuild period).  Returns (N, steps, seconds)."""
    d, n, v, cut, nnlup, extra, budget, kt = job
    # the reference reports its loading on stdout; keep this process's stdout
    # (shared with the benchmark's single JSON line) clean
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refs import Ref, setup_system, std_cmds
    r = Ref()
    setup_system(r, d, std_cmds(temp=kt, cutoff=cut, nnlup=nnlup, nup=NUP, extra=extra))
    r.set_coords(2, v)
    rebuild = "update_cell_list" if cut == "celllist" else "update_neighbor_list"
    r.call(rebuild)
    r.call("update_pair_list")
    r.call("set_potential_forces")
    r.call("energy_eval")
    r.call("force_eval")
    inlup, steps = 1, 0
    t0 = time.perf_counter()
    while steps < nnlup or time.perf_counter() - t0 < budget:
        if inlup % nnlup == 0:
            r.call(rebuild)
            inlup = 0
        inlup += 1
        r.call("update_pair_list")
        r.call("underdamped_iteration")
        steps += 1
    return n, steps, time.perf_counter() - t0


def cpu_baseline(chains, w, workers, budget_s=15.0, max_n=3200):
    """The reference (oracle/_ref: SOP-CPU compiled from its own sources with
    its own flags) on `workers` host cores at once: one single-thread process
    per chain, chains at evenly spaced size quantiles of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refs import Ref
    if not Ref.available():
        return None
    import multiprocessing as mp
    order = sorted((c for c in range(len(chains)) if len(chains[c][0]) // 3 <= max_n),
                   key=lambda c: len(chains[c][0]))
    k = min(workers, len(order))
    picks = sorted({order[min(int((i + 0.5) * len(order) / k), len(order) - 1)] for i in range(k)})
    cut = "celllist" if w["cutoff"] == sopg.CUTOFF_CELL else "neighborlist"
    extra = [("boxl", str(int(w["boxl"]))), ("ncell", str(int(w["ncell"])))] if cut == "celllist" else []
    with tempfile.TemporaryDirectory() as tmp:
        jobs = []
        for c in picks:
            x, v, _ = chains[c]
            jobs.append((os.path.join(tmp, f"c{c}"), len(x) // 3, v, cut, w["nnlup"], extra, budget_s,
                         w.get("kT", KT)))
        with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
            list(ex.map(lambda jc: sopg.write_inputs(chains[jc[1]][0], jc[0][0], vel=chains[jc[1]][1]),
                        zip(jobs, picks)))
        t0 = time.perf_counter()
        with mp.get_context("spawn").Pool(len(jobs)) as pool:
            res = pool.map(_cpu_worker, jobs)
        wall = time.perf_counter() - t0
    rate = sum(n * st / sec for n, st, sec in res)
    desc = ", ".join(f"N={n}:{st} steps/{sec:.1f} s" for n, st, sec in res)
    return dict(value=rate, unit="bead-steps/s", cores=len(jobs), kind="reference",
                sample=f"SOP-CPU compiled from its sources (g++ -O3, its own flags), {len(jobs)} single-thread "
                       f"processes at once, one chain each at evenly spaced size quantiles of the workload "
                       f"(N <= {max_n}; {desc}); value = sum over processes of N*steps/seconds "
                       f"({wall:.1f} s wall with process start and setup)")


def profiled_traffic(workload):
    """HBM bytes per k_force launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
    (scaled by a copy kernel of known size) for this workload, as committed in
    profiles/, with the launch time that profile saw."""
    path = os.path.join(ROOT, "profiles", f"k_force_{workload}.md")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and line.split(":")[0] in ("traffic_bytes_per_launch", "avg_launch_us", "valu_issue_frac",
                                                  "fp64_tflops"):
            k, v = line.split(":", 1)
            vals[k] = float(v)
    if "traffic_bytes_per_launch" not in vals:
        return None
    return dict(bytes=vals["traffic_bytes_per_launch"], avg_launch_us=vals.get("avg_launch_us"),
                valu_issue_frac=vals.get("valu_issue_frac"), fp64_tflops=vals.get("fp64_tflops"),
                source=os.path.relpath(path, ROOT))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="mixed256", choices=sorted(WORKLOADS))
    ap.add_argument("--kt", type=float, default=KT, help="temperature (kcal/mol) of the workload's chains")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--list-lanes", type=int, default=0)
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="reference processes run at once for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch every kernel directly (no graph replay)")
    ap.add_argument("--resident", type=int, default=0,
                    help="chains of at most this many beads run in the persistent resident kernel (0: none)")
    ap.add_argument("--resident-lanes", type=int, default=0)
    ap.add_argument("--skin", type=float, default=None,
                    help="candidate-list skin of neighbor-list rebuilds in Angstrom (default: the library's; 0: off)")
    ap.add_argument("--serve-jobs", type=int, default=100)
    ap.add_argument("--serve-step-scale", type=float, default=0.1,
                    help="job step counts are log-uniform in [1e4, 1e6] times this factor")
    ap.add_argument("--serve-budget-gib", type=float, default=8.0)
    ap.add_argument("--serve-chunk", type=int, default=1000)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, copy this many MiB of FP64 device memory three times with a plain "
                         "elementwise kernel (the known-size reference for HBM counter scaling)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (independent chains shard across GPUs with no "
              f"per-step exchange); --gpus {a.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    w = dict(WORKLOADS[a.workload], kT=a.kt)
    w["desc"] = w["desc"].format(kT=f"{a.kt:g}")
    t0 = time.perf_counter()
    chains = make_chains(a.workload, a.threads, a.kt)
    print(f"[bench] generated {len(chains)} chains in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    sizes = [len(x) // 3 for x, _, _ in chains]
    nb = sum(sizes)
    topo_cache = {}
    g = gpu_batched(chains, w, a.steps, a.warmup, a.lanes, a.list_lanes, topo_cache, graphs=not a.no_graphs,
                    resident=(a.resident, a.resident_lanes), skin=a.skin)
    value = nb * a.steps / g["seconds"]
    print(f"[bench] batched {value:.4g} bead-steps/s", file=sys.stderr, flush=True)
    plain, plain_t = (None, None) if a.no_plain else gpu_plain(chains, w, a.plain_steps, 10, topo_cache, a.lanes,
                                                              a.list_lanes)
    if plain:
        print(f"[bench] plain {plain:.4g} bead-steps/s ({plain_t:.1f} s)", file=sys.stderr, flush=True)
    step_ms, step_n = g["prof"][sopg.KERNEL_STEP]
    reb_ms, reb_n = g["prof"][sopg.KERNEL_REBUILD]
    obs_ms, obs_n = g["prof"][sopg.KERNEL_OBSERVE]
    avg_s = step_ms / max(step_n, 1) / 1e3
    achieved = g["bytes_step"] / avg_s / 1e9
    serve = None
    if not a.no_serve:
        t1 = time.perf_counter()
        jobs = serve_jobs(a.serve_jobs, a.serve_step_scale, a.threads)
        print(f"[bench] generated {len(jobs)} jobs in {time.perf_counter() - t1:.1f} s", file=sys.stderr, flush=True)
        serve = gpu_serve(jobs, WORKLOADS["mixed256"], a.serve_budget_gib * 2**30, 1024, a.serve_chunk, a.lanes)
        serve.update(stream=f"Poisson arrivals 20 jobs/s, {a.serve_jobs} jobs (seed 5); N log-uniform [50, 10000]; "
                            f"steps log-uniform [1e4, 1e6] x {a.serve_step_scale:g}, rounded to a multiple of 1000; "
                            f"kT uniform [0.5, 1.3]; half unfolding (from native), half folding (from a SAW); "
                            f"neighbor list every 20 steps; Q/Rg/energy record every {NUP} steps",
                     hbm_budget_gib=a.serve_budget_gib, chunk=a.serve_chunk,
                     # saturated: work offered faster than the batched throughput of this
                     # mixture (the server's capacity) can complete it
                     capacity_estimate=value,
                     saturated=bool(serve["offered_load"] > value),
                     offered_load_at_full_step_counts=serve["offered_load"] / a.serve_step_scale,
                     saturated_at_full_step_counts=bool(serve["offered_load"] / a.serve_step_scale > value))
        print(f"[bench] serve {serve['value']:.4g} bead-steps/s p50 {serve['p50_latency_s']:.3f} s "
              f"p99 {serve['p99_latency_s']:.3f} s", file=sys.stderr, flush=True)
    cpu = None if a.no_cpu else cpu_baseline(chains, w, a.cpu_workers)
    if cpu:
        print(f"[bench] cpu baseline {cpu['value']:.4g} bead-steps/s on {cpu['cores']} cores", file=sys.stderr,
              flush=True)
    tot = lambda k: sum(c[k] for c in g["counts"])
    prof_t = profiled_traffic(a.workload)
    line = {
        "metric": "bead-steps/s (batched vs plain) at FP64 force parity; p50/p99 job latency",
        "value": value,
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": g["seconds"] * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: {w['desc']}", "chains": len(chains), "beads": nb,
                   "n_min": min(sizes), "n_max": max(sizes), "form": "batched (all chains per launch)",
                   "nnl_att": tot("nnl_att"), "nnl_rep": tot("nnl_rep"), "nil_att": tot("nil_att"),
                   "nil_rep": tot("nil_rep"), "records": g["records"],
                   "step_lanes": g["lanes"], "lanes_from": "flag" if a.lanes else "autotuner",
                   "graph_segments": {"captured": g["graphs"][0], "replayed": g["graphs"][1]}},
        "plain": None if plain is None else {"value": plain, "unit": "bead-steps/s",
                                             "form": "plain (one chain per launch)", "steps_per_chain": a.plain_steps},
        "speedup_vs_plain": None if plain is None else value / plain,
        "serve": serve,
        "p50_latency_s": None if serve is None else serve["p50_latency_s"],
        "p99_latency_s": None if serve is None else serve["p99_latency_s"],
        "roofline": {"bound": "hbm", "kernel": "k_force (fused step)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if prof_t is None else prof_t["bytes"],
                     "traffic_source": None if prof_t is None else
                     f"{prof_t['source']} (rocprofv3 FETCH_SIZE + WRITE_SIZE per k_force launch, scaled by a "
                     f"1 GiB copy; that profile's k_force average {prof_t['avg_launch_us']} us)",
                     "valu_issue_frac": None if prof_t is None else prof_t["valu_issue_frac"],
                     "fp64_tflops_profiled": None if prof_t is None else prof_t["fp64_tflops"],
                     "fp64_peak_tflops": 78.6,
                     "bytes_per_launch": g["bytes_step"], "avg_launch_us": avg_s * 1e6,
                     "launches": step_n, "rebuild_avg_us": reb_ms / max(reb_n, 1) * 1e3,
                     "rebuilds": reb_n, "observe_avg_us": obs_ms / obs_n * 1e3 if obs_n else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if a.calib_copy_mib:
        calib_copy(a.calib_copy_mib)


def calib_copy(mib):
    """Three device-to-device copies of `mib` MiB (HIP's blit kernel): each
    reads and writes exactly mib MiB, the reference for HBM counter scaling."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = ctypes.c_size_t(mib * 2**20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), nbytes):
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, nbytes)
    for _ in range(3):
        if hip.hipMemcpy(dst, src, nbytes, 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"[bench] calibration copy: 3 x {mib} MiB read + written", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
