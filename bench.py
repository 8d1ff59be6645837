#!/usr/bin/env python3
"""Throughput benchmark of the SOP Langevin hot path on one MI355X.

Workload: plain SOP single chain, N=3000 beads (seeded SAW + MC collapse into a
sphere of 3.0*N^(1/3) A), FP64, cell list with 12 A cells (boxl 504, ncell 42)
rebuilt every 20 steps, pair list every step, underdamped Langevin h=0.05,
zeta=50, kT=0.6, records (energy + observables) every 1000 steps, device
(Philox) noise.  Inputs are resident in HBM before the timed region.

One step = the reference's underdamped_ctrl loop body (sop.cpp:153-189).
value = bead-steps/s = N * K / wall seconds of exactly K timed steps.

roofline: the fused step kernel (SOPG_KERNEL_STEP), timed with HIP events on
the library's stream over the timed region; algorithmic bytes per launch =
192 N + 16 nbnd + 8 nang + 40 nnl_att + 8 nnl_rep (bead state read+write, bond
and angle indices, neighbor list streamed once), against 8 TB/s HBM.

cpu_baseline: the reference itself (oracle/_ref, compiled from its sources with
its own flags) on a bounded sample of the same workload, one host core.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sop-cpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import sopg  # noqa: E402

HBM_PEAK_GBS = 8000.0
N_BEADS = 3000
SEED = 2024
KT = 0.6
NNLUP = 20
NUP = 1000
BOXL, NCELL = 504.0, 42.0


def gpu_run(x, v, steps, warmup, lanes, list_lanes=0):
    topo = sopg.chain_topology(x)
    p = sopg.chain_params(T=KT, boxl=BOXL, ncell=NCELL, cutoff=sopg.CUTOFF_CELL, nnlup=NNLUP, nup=NUP,
                          noise_mode=sopg.NOISE_PHILOX, seed=SEED)
    ctx = sopg.Context(device=0, lanes=lanes, list_lanes=list_lanes)
    ctx.add_chain(topo, p)
    ctx.put_state(0, pos=x, unc=x, vel=v, force=np.zeros_like(x))
    # underdamped_ctrl preamble: lists, energy, first force (untimed)
    ctx.update_cell_list()
    ctx.update_pair_list()
    ctx.energy_eval()
    ctx.force_eval()
    if warmup:
        ctx.run(warmup)
    counts = ctx.list_counts()
    ctx.sync()
    ctx.profile(True)
    t0 = time.perf_counter()
    recs = ctx.run(steps)
    ctx.sync()
    t1 = time.perf_counter()
    step_ms, step_n = ctx.profile_read(0)
    reb_ms, reb_n = ctx.profile_read(1)
    counts_end = ctx.list_counts()
    ctx.profile(False)
    ctx.close()
    return dict(seconds=t1 - t0, step_ms=step_ms, step_n=step_n, reb_ms=reb_ms, reb_n=reb_n,
                counts=counts, counts_end=counts_end, records=len(recs), topo=topo)


def cpu_baseline(x, v, steps):
    """The reference (oracle/_ref) on `steps` steps of the same workload, one core."""
    from refs import Ref, setup_system, std_cmds
    if not Ref.available():
        return None
    n = len(x) // 3
    with tempfile.TemporaryDirectory() as d:
        sopg.write_inputs(x, d, vel=v)
        r = Ref()
        setup_system(r, d, std_cmds(temp=KT, cutoff="celllist", nnlup=NNLUP, nup=NUP,
                                    extra=[("boxl", str(int(BOXL))), ("ncell", str(int(NCELL)))]))
    r.set_coords(2, v)
    r.call("update_cell_list")
    r.call("update_pair_list")
    r.call("set_potential_forces")
    r.call("energy_eval")
    r.call("force_eval")
    inlup = 1
    t0 = time.perf_counter()
    for _ in range(steps):
        if inlup % NNLUP == 0:
            r.call("update_cell_list")
            inlup = 0
        inlup += 1
        r.call("update_pair_list")
        r.call("underdamped_iteration")
    dt = time.perf_counter() - t0
    return dict(value=n * steps / dt, unit="bead-steps/s", cores=1, kind="reference",
                sample=f"{steps} steps ({steps // NNLUP} cell-list periods) of the N={n} workload, "
                       f"SOP-CPU compiled from its sources with g++ -O3, {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--list-lanes", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=800)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (independent chains shard across GPUs with no "
              f"per-step exchange); --gpus {a.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    x = sopg.gen_native(N_BEADS, SEED)
    v = sopg.gen_maxwell(N_BEADS, KT, SEED)
    g = gpu_run(x, v, a.steps, a.warmup, a.lanes, a.list_lanes)
    value = N_BEADS * a.steps / g["seconds"]
    c = g["counts"]
    topo = g["topo"]
    nbnd, nang = len(topo["ib"]), len(topo["ia"])
    bytes_per_launch = 192 * N_BEADS + 16 * nbnd + 8 * nang + 40 * c["nnl_att"] + 8 * c["nnl_rep"]
    avg_s = g["step_ms"] / max(g["step_n"], 1) / 1e3
    achieved = bytes_per_launch / avg_s / 1e9
    cpu = None if a.no_cpu else cpu_baseline(x, v, a.cpu_steps)
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
        "config": {"workload": "plain SOP single chain N=3000, cell list 12 A (boxl 504, ncell 42), "
                               "list rebuild every 20 steps, kT 0.6, h 0.05, zeta 50, Philox noise",
                   "n_beads": N_BEADS, "form": "plain (one chain per launch)",
                   "nnl_att": c["nnl_att"], "nnl_rep": c["nnl_rep"], "nil_att": c["nil_att"],
                   "nil_rep": c["nil_rep"], "records": g["records"]},
        "roofline": {"bound": "hbm", "kernel": "k_force (fused step)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": avg_s * 1e6, "launches": g["step_n"],
                     "rebuild_avg_us": g["reb_ms"] / max(g["reb_n"], 1) * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
