#!/usr/bin/env python3
"""Throughput benchmark of the SOP Langevin hot path on one MI355X.

One step = the reference's underdamped_ctrl loop body (sop.cpp:153-189) for
every chain: list rebuild when due, pair list, force (noise, FENE, angular,
native LJ, non-native repulsion), integration; energies and Q/Rg every 1000
steps.  FP64 throughout, inputs resident in HBM before every timed region,
device Philox noise.  The workloads are those of sop-cpu_amd/workloads.py.

Headline (`value`): mixed256 -- 256 chains, N log-uniform in [50, 5000] (seed
4), each its own native structure, neighbor list every 20 steps, kT 0.6 --
advanced together in one ragged batch: sum(N) * K / seconds of exactly K timed
steps after W untimed warmup steps (the lane autotuner and the preamble run
before the warmup).  The same line carries the other configurations:

  plain76        one N=76 chain, neighbor list every 50 steps
  plain3000      one N=3000 chain, 12 A cell list, kT 0.6 and 1.2, and the
                 neighbor-list variant (rebuilds every 20 steps)
  replicas1024   1024 replicas of one N=76 chain: batched (per-step kernels,
                 graph replay), the resident kernel, and plain (one replica
                 per context, a sample of them)
  mixed256       batched (the headline) and plain (one chain per context)
  serve          the Poisson job stream (20 jobs/s, N log-uniform in
                 [50, 10000], steps log-uniform in [1e4, 1e6], kT in [0.5, 1.3],
                 half folding from a self-avoiding walk that keeps native pairs
                 at least their native distance apart) replayed in real time
                 against the continuous-batching server: throughput, p50/p99
                 latency and wait, saturation against the server's own
                 measured capacity

Every configuration has the reference CPU path beside it: SOP-CPU compiled
from its own sources with its own flags (oracle/_ref), single-thread
processes on the host's cores in two waves of at most --cpu-workers at once.

roofline: the fused step kernel's algorithmic bytes per launch (per chain
192 N + 16 nbnd + 8 nang + 40 nnl_att + 8 nnl_rep) over its average launch
time, measured with HIP events on the library's stream over a second run of
the same batch (200 steps, ten list rebuilds); the list rebuild's (per chain
and rebuild 40 ncon_att + 24 N + 40 nnl_att + 8 nnl_rep) likewise.  Both
against 8 TB/s.
"""
import argparse
import json
import math
import multiprocessing as mp
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
import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "bead-steps/s (batched vs plain) at FP64 force parity; p50/p99 job latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------------ GPU legs

def build_ctx(chains, w, kt, lanes=0, topos=None):
    ctx = sopg.Context(device=0, lanes=lanes)
    topos = W.setup_ctx(ctx, chains, w, kt, topos=topos)
    return ctx, topos


def step_bytes(ctx, topos):
    """Algorithmic bytes of one step-kernel launch and of one list rebuild over the batch."""
    b_step = b_reb = 0
    counts = []
    for c, t in enumerate(topos):
        k = ctx.list_counts(c)
        counts.append(k)
        n = t["n"]
        b_step += 192 * n + 16 * len(t["ib"]) + 8 * len(t["ia"]) + 40 * k["nnl_att"] + 8 * k["nnl_rep"]
        b_reb += 40 * len(t["att_i"]) + 24 * n + 40 * k["nnl_att"] + 8 * k["nnl_rep"]
    return b_step, b_reb, counts


def timed(ctx, steps, noalloc=False):
    ctx.sync()
    t0 = time.perf_counter()
    recs = ctx.run_noalloc(steps) if noalloc else ctx.run(steps)
    ctx.sync()
    return time.perf_counter() - t0, recs


def gpu_batched(chains, w, kt, steps, warmup, lanes=0, topos=None, graphs=True, resident=(0, 0), skin=None,
                profile_steps=0):
    """One context holding every chain; lanes autotuned (setup), W warmup steps, K timed."""
    ctx, topos = build_ctx(chains, w, kt, lanes, topos)
    ctx.set_graphs(graphs)
    ctx.set_resident(*resident)
    if skin is not None:
        ctx.set_skin(skin)
    if not lanes:
        ctx.autotune()
    if warmup:
        ctx.run(warmup)
    seconds, recs = timed(ctx, steps)
    out = dict(seconds=seconds, records=len(recs), lanes=ctx.lanes()[0], graphs=ctx.graph_stats(),
               beads=sum(t["n"] for t in topos))
    if profile_steps:
        # the same batch, every kernel bracketed by HIP events on the library's stream
        ctx.profile(True)
        ctx.run(profile_steps)
        ctx.sync()
        out["prof"] = {k: ctx.profile_read(k) for k in range(4)}
        ctx.profile(False)
        out["bytes_step"], out["bytes_rebuild"], out["counts"] = step_bytes(ctx, topos)
    ctx.close()
    return out, topos


def gpu_plain(chains, w, kt, steps, warmup, topos=None):
    """One chain per context (every launch advances a single chain), each timed alone."""
    total_bs, total_t = 0.0, 0.0
    for i, ch in enumerate(chains):
        ctx, tp = build_ctx([ch], w, kt, 0, None if topos is None else [topos[i]])
        ctx.autotune()
        if warmup:
            ctx.run(warmup)
        seconds, _ = timed(ctx, steps)
        total_bs += tp[0]["n"] * steps
        total_t += seconds
        ctx.close()
    return total_bs / total_t, total_t


def gpu_serve(jobs, hbm_budget, chunk):
    """Replay the arrival stream in real time against the server; latency = finish - arrival."""
    w = W.SERVE
    srv = sopg.Server(hbm_budget=hbm_budget, max_chains=1024, chunk=chunk)
    t0 = srv.now()
    i, active, queued, peak = 0, 0, 0, 0
    ids = []
    last_log = time.perf_counter()
    while i < len(jobs) or active or queued:
        now = srv.now() - t0
        while i < len(jobs) and jobs[i]["arrival"] <= now:
            jb = jobs[i]
            p = sopg.chain_params(T=jb["kT"], boxl=w["boxl"], ncell=w["ncell"], cutoff=w["cutoff"],
                                  nnlup=w["nnlup"], nup=W.NUP, noise_mode=sopg.NOISE_PHILOX, seed=jb["seed"])
            ids.append(srv.submit(jb["topo"], p, jb["x"], jb["v"], jb["steps"], arrival=t0 + jb["arrival"]))
            queued += 1
            i += 1
        if active == 0 and queued == 0:
            time.sleep(max(0.0, min(0.01, jobs[i]["arrival"] - (srv.now() - t0))))
            continue
        active, queued = srv.poll()
        peak = max(peak, active)
        if time.perf_counter() - last_log > 15.0:
            done = sum(srv.job(j)["state"] == sopg.JOB_DONE for j in ids)
            log(f"serve: {done}/{len(jobs)} jobs done, {active} resident, {queued} queued")
            last_log = time.perf_counter()
    info = [srv.job(j) for j in ids]
    st = srv.stats()
    nrec = sum(len(srv.records(j)) for j in ids)
    srv.close()
    done = [k for k in info if k["state"] == sopg.JOB_DONE]
    span_arr = jobs[-1]["arrival"] - jobs[0]["arrival"]
    offered = sum(j["n"] * j["steps"] for j in jobs) / span_arr
    lat = np.array([k["t_finish"] - k["t_arrival"] for k in done])
    wait = np.array([k["t_admit"] - k["t_arrival"] for k in done])
    span = max(k["t_finish"] for k in done) - (t0 + jobs[0]["arrival"])
    work = sum(k["nbead"] * k["nsteps"] for k in done)
    capacity = st["bead_steps"] / st["busy_s"]
    return dict(value=work / span, unit="bead-steps/s", jobs=len(jobs), done=len(done),
                rejected=sum(k["state"] == sopg.JOB_REJECTED for k in info),
                p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                p50_wait_s=float(np.percentile(wait, 50)), p99_wait_s=float(np.percentile(wait, 99)),
                makespan_s=span, rounds=st["rounds"], busy_s=st["busy_s"], peak_resident=peak, records=nrec,
                offered_load=offered,
                capacity=capacity,
                capacity_source="the server's own bead-steps over the seconds spent inside its rounds",
                saturated=bool(offered > capacity),
                jobs_waited_for_capacity=sum(k["rounds_waited"] > 0 for k in done),
                max_rounds_waited=max(k["rounds_waited"] for k in done))


# ------------------------------------------------------------------ CPU legs

def _cpu_worker(job):
    """One reference process: set up one chain from its files, run underdamped_ctrl's loop
    body (sop.cpp:153-189) for whole rebuild periods until the budget is spent.
    Returns (tag, N, steps, seconds)."""
    tag, d, n, x, v, cut, nnlup, extra, budget, kt = job
    # the reference reports its loading on stdout; keep this process's stdout clean
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "sop-cpu_amd"))
    import sopg as sg
    from refs import Ref, setup_system, std_cmds
    sg.write_inputs(x, d, vel=v)
    r = Ref()
    setup_system(r, d, std_cmds(temp=kt, cutoff=cut, nnlup=nnlup, nup=W.NUP, extra=extra))
    r.set_coords(2, v)
    for f in os.listdir(d):
        os.remove(os.path.join(d, f))
    rebuild = "update_cell_list" if cut == "celllist" else "update_neighbor_list"
    r.call(rebuild)
    r.call("update_pair_list")
    r.call("set_potential_forces")
    r.call("energy_eval")
    r.call("force_eval")
    inlup, steps = 1, 0
    t0 = time.perf_counter()
    while steps < nnlup or steps % nnlup or time.perf_counter() - t0 < budget:
        if inlup % nnlup == 0:
            r.call(rebuild)
            inlup = 0
        inlup += 1
        r.call("update_pair_list")
        r.call("underdamped_iteration")
        steps += 1
    return tag, n, steps, time.perf_counter() - t0


def cpu_wave(jobs):
    with mp.get_context("spawn").Pool(len(jobs)) as pool:
        return pool.map(_cpu_worker, jobs)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def quantile_picks(sizes, k):
    order = sorted(range(len(sizes)), key=lambda c: sizes[c])
    return sorted({order[min(int((i + 0.5) * len(order) / k), len(order) - 1)] for i in range(k)})


def cpu_baselines(data, workers, budget):
    """The reference on the host: wave A holds plain76 (1 process), plain3000's three variants
    (1 process each) and replicas1024 (the rest of the workers, one replica each); wave B holds
    mixed256 (one process per chain at evenly spaced size quantiles).  Throughput per
    configuration = sum over its processes of N * steps / seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refs import Ref
    if not Ref.available():
        return None
    cl = [("boxl", "504"), ("ncell", "42")]
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        def job(tag, k, x, v, cut, nnlup, extra, kt):
            return (tag, os.path.join(tmp, f"{tag}_{k}"), len(x) // 3, x, v, cut, nnlup, extra, budget, kt)
        a = []
        x, v, _ = data["plain76"][0]
        a.append(job("plain76", 0, x, v, "neighborlist", 50, [], 0.6))
        for tag, key, cut, kt in (("plain3000_cl06", "plain3000", "celllist", 0.6),
                                  ("plain3000_cl12", "plain3000_kt12", "celllist", 1.2),
                                  ("plain3000_nl06", "plain3000", "neighborlist", 0.6)):
            x, v, _ = data[key][0]
            a.append(job(tag, 0, x, v, cut, 20, cl, kt))
        reps = data["replicas1024"]
        for k in range(max(1, workers - len(a))):
            x, v, _ = reps[k]
            a.append(job("replicas1024", k, x, v, "neighborlist", 50, [], 0.6))
        t0 = time.perf_counter()
        out_a = cpu_wave(a)
        wall_a = time.perf_counter() - t0
        log(f"cpu wave A: {len(a)} processes, {wall_a:.1f} s")
        mixed = data["mixed256"]
        picks = quantile_picks([len(x) // 3 for x, _, _ in mixed], workers)
        b = [job("mixed256", c, mixed[c][0], mixed[c][1], "neighborlist", 20, [], 0.6) for c in picks]
        t0 = time.perf_counter()
        out_b = cpu_wave(b)
        wall_b = time.perf_counter() - t0
        log(f"cpu wave B: {len(b)} processes, {wall_b:.1f} s")
    for tag, n, st, sec in out_a + out_b:
        res.setdefault(tag, []).append((n, st, sec))
    model = cpu_model()

    def entry(tag, kind_note):
        rows = res[tag]
        rate = sum(n * st / sec for n, st, sec in rows)
        desc = ", ".join(f"N={n}: {st} steps in {sec:.1f} s" for n, st, sec in rows)
        return dict(value=rate, unit="bead-steps/s", cores=len(rows), kind="reference",
                    sample=f"SOP-CPU compiled from its sources (g++ -O3, its own flags), {len(rows)} single-thread "
                           f"process(es) at once {kind_note}; whole rebuild periods for at least {budget:g} s each "
                           f"({desc}); host: {model}, {os.cpu_count()} logical cores")
    return dict(
        plain76=entry("plain76", "(the reference is single-threaded)"),
        plain3000_cl06=entry("plain3000_cl06", "(cell list, kT 0.6)"),
        plain3000_cl12=entry("plain3000_cl12", "(cell list, kT 1.2)"),
        plain3000_nl06=entry("plain3000_nl06", "(neighbor list, kT 0.6)"),
        replicas1024=entry("replicas1024", "(one replica each, alongside wave A's other processes)"),
        mixed256=entry("mixed256", "(one chain each, at evenly spaced size quantiles of the 256, N up to 4964)"),
        per_n=sorted((n, n * st / sec) for n, st, sec in res["mixed256"]),
        cores_used=max(len(out_a), len(out_b)))


def cpu_serve_model(per_n, jobs, cores):
    """The reference's throughput on the serving stream's job mix with `cores` single-thread
    processes: each job alone at the per-N rate measured in wave B (log-log interpolation,
    extrapolated beyond the largest size measured), perfectly packed onto the cores."""
    ln = np.log([n for n, _ in per_n])
    lr = np.log([r for _, r in per_n])
    fit = np.polyfit(ln, lr, 1)

    def rate(n):
        if n <= per_n[-1][0]:
            return float(np.exp(np.interp(np.log(n), ln, lr)))
        return float(np.exp(np.polyval(fit, np.log(n))))
    work = sum(j["n"] * j["steps"] for j in jobs)
    secs = sum(j["n"] * j["steps"] / rate(j["n"]) for j in jobs)
    return dict(value=work / (secs / cores), unit="bead-steps/s", cores=cores, kind="reference",
                sample=f"modelled from the reference's measured per-N rates of the mixed256 wave (log-log "
                       f"interpolation, a power-law fit beyond N={per_n[-1][0]}): the stream's {len(jobs)} jobs, "
                       f"each alone on one core, packed onto {cores} cores; {secs:.0f} core-seconds of work")


# ------------------------------------------------------------------ profile hook

def profiled_traffic(workload):
    """HBM bytes per k_force launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a
    copy kernel of known size) for this workload, as committed in profiles/, with the launch
    time that profile saw."""
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


def calib_copy(mib):
    """Three device-to-device copies of `mib` MiB (HIP's blit kernel): each reads and writes
    exactly mib MiB, the reference for HBM counter scaling."""
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
    log(f"calibration copy: 3 x {mib} MiB read + written")


# ------------------------------------------------------------------ main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed steps of the headline (mixed256 batched)")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--leg-steps", type=int, default=2000, help="timed steps of the other batched / plain legs")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="reference processes run at once for the CPU baselines")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--only-headline", action="store_true", help="the mixed256 batched leg and its roofline only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch every kernel directly (no graph replay)")
    ap.add_argument("--skin", type=float, default=None, help="candidate-list skin in Angstrom (default: library's)")
    ap.add_argument("--serve-jobs", type=int, default=100)
    ap.add_argument("--serve-step-scale", type=float, default=1.0,
                    help="job step counts are log-uniform in [1e4, 1e6] times this factor")
    ap.add_argument("--serve-budget-gib", type=float, default=8.0)
    ap.add_argument("--serve-chunk", type=int, default=1000)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, copy this many MiB with HIP's blit kernel (HBM counter scaling)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (independent chains shard across GPUs with no "
              f"per-step exchange); --gpus {a.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    extra_legs = not a.only_headline
    t0 = time.perf_counter()
    data = {"mixed256": W.make_chains("mixed256", a.threads)}
    if extra_legs:
        data["replicas1024"] = W.make_chains("replicas1024", a.threads)
        data["plain76"] = data["replicas1024"][:1]
        data["plain3000"] = W.make_chains("plain3000", a.threads)
        x3 = data["plain3000"][0][0]
        data["plain3000_kt12"] = [(x3, sopg.gen_maxwell(3000, 1.2, 1000), 1000)]
    log(f"generated the chains in {time.perf_counter() - t0:.1f} s")

    # -- headline: mixed256 batched
    mixed = data["mixed256"]
    topo_mixed = [sopg.chain_topology(x) for x, _, _ in mixed]
    w = W.WORKLOADS["mixed256"]
    g, _ = gpu_batched(mixed, w, W.KT, a.steps, a.warmup, a.lanes, topo_mixed, graphs=not a.no_graphs, skin=a.skin,
                       profile_steps=200)
    nb = g["beads"]
    value = nb * a.steps / g["seconds"]
    log(f"mixed256 batched {value:.4g} bead-steps/s ({g['seconds'] * 1e3 / a.steps:.4f} ms/step)")
    step_ms, step_n = g["prof"][sopg.KERNEL_STEP]
    reb_ms, reb_n = g["prof"][sopg.KERNEL_REBUILD]
    obs_ms, obs_n = g["prof"][sopg.KERNEL_OBSERVE]
    avg_s = step_ms / max(step_n, 1) / 1e3
    achieved = g["bytes_step"] / avg_s / 1e9
    reb_s = reb_ms / max(reb_n, 1) / 1e3
    configs = {}
    plain = None
    if not a.no_plain:
        plain, plain_t = gpu_plain(mixed, w, W.KT, 100, 10, topo_mixed)
        log(f"mixed256 plain {plain:.4g} bead-steps/s")
    configs["mixed256"] = dict(batched=value, plain=plain, speedup_vs_plain=None if plain is None else value / plain,
                               plain_form="one chain per context, 100 timed steps each after 10 warmup")

    if extra_legs:
        ls = a.leg_steps
        # cfg1: one N=76 chain
        r1, _ = gpu_batched(data["plain76"], W.WORKLOADS["plain76"], W.KT, ls, 100)
        configs["plain76"] = dict(value=76 * ls / r1["seconds"], ms_per_step=r1["seconds"] * 1e3 / ls, steps=ls,
                                  step_lanes=r1["lanes"])
        log(f"plain76 {configs['plain76']['value']:.4g} bead-steps/s")
        # cfg2: one N=3000 chain, cell list at kT 0.6 and 1.2, neighbor list at 0.6
        c2 = {}
        for tag, key, wl, kt in (("cl_kT0.6", "plain3000", "plain3000", 0.6),
                                 ("cl_kT1.2", "plain3000_kt12", "plain3000", 1.2),
                                 ("nl_kT0.6", "plain3000", "plain3000nl", 0.6)):
            r2, _ = gpu_batched(data[key], W.WORKLOADS[wl], kt, ls, 100)
            c2[tag] = dict(value=3000 * ls / r2["seconds"], ms_per_step=r2["seconds"] * 1e3 / ls,
                           step_lanes=r2["lanes"])
            log(f"plain3000 {tag} {c2[tag]['value']:.4g} bead-steps/s")
        configs["plain3000"] = dict(c2, steps=ls)
        # cfg3: 1024 replicas of the N=76 chain
        reps = data["replicas1024"]
        topo_r = [sopg.chain_topology(reps[0][0])] * len(reps)
        wr = W.WORKLOADS["replicas1024"]
        r3, _ = gpu_batched(reps, wr, W.KT, ls, 100, topos=topo_r)
        r3r, _ = gpu_batched(reps, wr, W.KT, ls, 100, topos=topo_r, resident=(128, 0))
        nrep_plain = 64
        p3, _ = (None, None) if a.no_plain else gpu_plain(reps[:nrep_plain], wr, W.KT, 200, 10, topo_r[:nrep_plain])
        b3 = 76 * len(reps) * ls / r3["seconds"]
        configs["replicas1024"] = dict(
            batched=b3, ms_per_step=r3["seconds"] * 1e3 / ls, step_lanes=r3["lanes"],
            graph_segments={"captured": r3["graphs"][0], "replayed": r3["graphs"][1]},
            resident=76 * len(reps) * ls / r3r["seconds"], resident_ms_per_step=r3r["seconds"] * 1e3 / ls,
            plain=p3, plain_form=f"{nrep_plain} of the replicas, one per context, 200 timed steps each",
            speedup_vs_plain=None if p3 is None else b3 / p3, steps=ls)
        log(f"replicas1024 batched {b3:.4g}, resident {configs['replicas1024']['resident']:.4g}, plain {p3}")

    serve = None
    if not a.no_serve:
        t1 = time.perf_counter()
        jobs = W.serve_jobs(a.serve_jobs, a.serve_step_scale, a.threads)
        log(f"generated {len(jobs)} jobs in {time.perf_counter() - t1:.1f} s")
        serve = gpu_serve(jobs, a.serve_budget_gib * 2**30, a.serve_chunk)
        serve.update(stream=f"Poisson arrivals 20 jobs/s, the first {a.serve_jobs} jobs of the stream (seed 5); "
                            f"N log-uniform [50, 10000]; steps log-uniform [1e4, 1e6]"
                            + ("" if a.serve_step_scale == 1.0 else f" x {a.serve_step_scale:g}")
                            + ", rounded to a multiple of 1000; kT uniform [0.5, 1.3]; half unfolding (from "
                              "native), half folding (from a self-avoiding walk within 200 A that keeps every native pair at "
                              "least its native distance apart); neighbor list every 20 steps; "
                              f"Q/Rg/energy record every {W.NUP} steps",
                     hbm_budget_gib=a.serve_budget_gib, chunk=a.serve_chunk)
        log(f"serve {serve['value']:.4g} bead-steps/s p50 {serve['p50_latency_s']:.3f} s "
            f"p99 {serve['p99_latency_s']:.3f} s, capacity {serve['capacity']:.4g}")
        configs["serve"] = serve

    cpu = None
    if not a.no_cpu and extra_legs:
        cpu_all = cpu_baselines(data, a.cpu_workers, a.cpu_seconds)
        if cpu_all:
            cpu = cpu_all["mixed256"]
            configs["mixed256"]["cpu_baseline"] = cpu
            configs["plain76"]["cpu_baseline"] = cpu_all["plain76"]
            configs["plain3000"]["cl_kT0.6"]["cpu_baseline"] = cpu_all["plain3000_cl06"]
            configs["plain3000"]["cl_kT1.2"]["cpu_baseline"] = cpu_all["plain3000_cl12"]
            configs["plain3000"]["nl_kT0.6"]["cpu_baseline"] = cpu_all["plain3000_nl06"]
            configs["replicas1024"]["cpu_baseline"] = cpu_all["replicas1024"]
            if serve is not None:
                configs["serve"]["cpu_baseline"] = cpu_serve_model(cpu_all["per_n"], jobs, a.cpu_workers)
            log(f"cpu baseline mixed256 {cpu['value']:.4g} bead-steps/s on {cpu['cores']} cores")

    tot = lambda k: sum(c[k] for c in g["counts"])
    prof_t = profiled_traffic("mixed256")
    line = {
        "metric": METRIC,
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
        "config": {"workload": f"mixed256: {w['desc'].format(kT='0.6')}", "chains": len(mixed), "beads": nb,
                   "n_min": min(t["n"] for t in topo_mixed), "n_max": max(t["n"] for t in topo_mixed),
                   "form": "batched (all chains per launch)", "nnl_att": tot("nnl_att"), "nnl_rep": tot("nnl_rep"),
                   "nil_att": tot("nil_att"), "nil_rep": tot("nil_rep"), "records": g["records"],
                   "step_lanes": g["lanes"], "lanes_from": "flag" if a.lanes else "autotuner",
                   "graph_segments": {"captured": g["graphs"][0], "replayed": g["graphs"][1]}},
        "plain": None if plain is None else {"value": plain, "unit": "bead-steps/s",
                                             "form": "plain (one chain per launch)", "steps_per_chain": 100},
        "speedup_vs_plain": None if plain is None else value / plain,
        "serve": serve,
        "p50_latency_s": None if serve is None else serve["p50_latency_s"],
        "p99_latency_s": None if serve is None else serve["p99_latency_s"],
        "roofline": {"bound": "hbm", "kernel": "k_force (fused step)", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if prof_t is None else prof_t["bytes"],
                     "traffic_source": None if prof_t is None else
                     f"{prof_t['source']} (rocprofv3 FETCH_SIZE + WRITE_SIZE per k_force launch, scaled by a "
                     f"1 GiB copy; that profile's k_force average {prof_t['avg_launch_us']} us)",
                     "valu_issue_frac": None if prof_t is None else prof_t["valu_issue_frac"],
                     "fp64_tflops_profiled": None if prof_t is None else prof_t["fp64_tflops"],
                     "fp64_peak_tflops": FP64_PEAK_TFLOPS, "bytes_per_launch": g["bytes_step"],
                     "avg_launch_us": avg_s * 1e6, "launches": step_n,
                     "observe_avg_us": obs_ms / obs_n * 1e3 if obs_n else None},
        "rebuild_roofline": {"bound": "hbm", "kernels": "k_grid + k_rebuild_grid + k_cand_filter + k_rebuild "
                                                        "(one list rebuild of the batch)",
                             "achieved": g["bytes_rebuild"] / reb_s / 1e9 if reb_n else None,
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": g["bytes_rebuild"] / reb_s / 1e9 / HBM_PEAK_GBS if reb_n else None,
                             "bytes_per_rebuild": g["bytes_rebuild"], "avg_rebuild_us": reb_s * 1e6,
                             "rebuilds": reb_n},
        "cpu_baseline": cpu,
        "configs": configs,
    }
    print(json.dumps(line), flush=True)
    if a.calib_copy_mib:
        calib_copy(a.calib_copy_mib)


if __name__ == "__main__":
    main()
