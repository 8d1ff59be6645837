"""The synthetic workloads the benchmark and the configuration tests run.

Every workload is built from the generator of libsopg.so (seeded self-avoiding
walk + Monte Carlo collapse into a sphere of 3.0*N^(1/3) A, natives = pairs
|i-j| > 2 closer than 8 A) with the dynamics the reference's scripts set in
this order (params.cpp:12-101): `set dynamics underdamped; set zeta 50;
set t_step 0.05` (h = 0.05, zeta = 50), then the temperature.

  plain76       one chain N=76, neighbor list every 50 steps
  plain3000     one chain N=3000, 12 A cell list (boxl 504, ncell 42) every 20
                steps; plain3000nl: the same chain with the neighbor list
  replicas1024  1024 trajectories of one N=76 chain, each with its own seed and
                Maxwell velocities, neighbor list every 50 steps
  mixed256      256 chains, N log-uniform in [50, 5000] (seed 4), each its own
                native structure, neighbor list every 20 steps
  serve stream  Poisson arrivals (20 jobs/s, seed 5), N log-uniform in
                [50, 10000], steps log-uniform in [1e4, 1e6] (rounded to a
                multiple of 1000), kT uniform in [0.5, 1.3]; even jobs unfold
                (start from their native structure), odd jobs fold (start from a
                self-avoiding walk within 200 A of the origin that keeps every
                native pair at least its native distance apart); neighbor list
                every 20 steps
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import sopg

KT = 0.6
NUP = 1000

WORKLOADS = {
    "plain76": dict(desc="one chain N=76; neighbor list every 50 steps; kT {kT}, h 0.05, zeta 50",
                    cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=50, boxl=500.0, ncell=55.0),
    "plain3000": dict(desc="one chain N=3000; cell list 12 A (boxl 504, ncell 42) rebuilt every 20 steps; "
                           "kT {kT}, h 0.05, zeta 50",
                      cutoff=sopg.CUTOFF_CELL, nnlup=20, boxl=504.0, ncell=42.0),
    "plain3000nl": dict(desc="one chain N=3000; neighbor list every 20 steps; kT {kT}, h 0.05, zeta 50",
                        cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, boxl=504.0, ncell=42.0),
    "replicas1024": dict(desc="1024 replicas of one N=76 chain, own seeds and Maxwell velocities; "
                              "neighbor list every 50 steps; kT {kT}, h 0.05, zeta 50",
                         cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=50, boxl=500.0, ncell=55.0),
    "mixed256": dict(desc="256 chains, N log-uniform in [50, 5000] (seed 4), each its own native structure; "
                          "neighbor list every 20 steps; kT {kT}, h 0.05, zeta 50",
                     cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, boxl=500.0, ncell=55.0),
}
SERVE = dict(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, boxl=500.0, ncell=55.0)


def chain_sizes(workload):
    if workload == "mixed256":
        rng = np.random.default_rng(4)
        return [int(v) for v in np.round(np.exp(rng.uniform(np.log(50), np.log(5000), 256)))]
    if workload.startswith("plain3000"):
        return [3000]
    if workload == "plain76":
        return [76]
    return [76] * 1024


def make_chains(workload, threads, kt=KT):
    """[(native xyz, Maxwell velocities, seed)] of the workload's chains."""
    sizes = chain_sizes(workload)
    if workload in ("replicas1024", "plain76"):
        x = sopg.gen_native(76, 76)
        return [(x, sopg.gen_maxwell(76, kt, 10_000 + r), 10_000 + r) for r in range(len(sizes))]

    def one(c):
        n = sizes[c]
        return (sopg.gen_native(n, 1000 + c), sopg.gen_maxwell(n, kt, 1000 + c), 1000 + c)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, range(len(sizes))))


def params(w, kt, seed, noise_mode=sopg.NOISE_PHILOX, nup=NUP):
    return sopg.chain_params(T=kt, boxl=w["boxl"], ncell=w["ncell"], cutoff=w["cutoff"], nnlup=w["nnlup"],
                             nup=nup, noise_mode=noise_mode, seed=seed)


def setup_ctx(ctx, chains, w, kt, topos=None, noise_mode=sopg.NOISE_PHILOX, nup=NUP, preamble_noise=None):
    """Add the chains and run underdamped_ctrl's preamble (sop.cpp:119-147) for each: lists,
    pair list, first force.  preamble_noise[c]: the 3N gaussians of chain c's first force
    (replayed-noise chains).  Returns the topologies."""
    topos = topos or [sopg.chain_topology(x) for (x, _, _) in chains]
    for (x, v, seed), t in zip(chains, topos):
        ctx.add_chain(t, params(w, kt, seed, noise_mode, nup))
    for c, (x, v, seed) in enumerate(chains):
        ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
        if w["cutoff"] == sopg.CUTOFF_CELL:
            ctx.update_cell_list(c)
        else:
            ctx.update_neighbor_list(c)
        ctx.update_pair_list(c)
        ctx.force_eval(c, None if preamble_noise is None else preamble_noise[c])
    return topos


def serve_jobs(njobs, step_scale, threads, seed=5, rate=20.0):
    """The serving stream's first njobs jobs (step counts times step_scale)."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, njobs))
    sizes = np.round(np.exp(rng.uniform(np.log(50), np.log(10000), njobs))).astype(int)
    raw = np.exp(rng.uniform(np.log(1e4), np.log(1e6), njobs)) * step_scale
    steps = np.maximum(1000, np.round(raw / 1000.0) * 1000).astype(int)
    kts = rng.uniform(0.5, 1.3, njobs)

    def one(j):
        n = int(sizes[j])
        nat = sopg.gen_native(n, 20_000 + j)
        x0 = nat if j % 2 == 0 else sopg.gen_unfolded(nat, 30_000 + j)
        return x0, sopg.gen_maxwell(n, float(kts[j]), 20_000 + j), sopg.chain_topology(nat)
    # largest first, so that the slowest structures start early
    order = sorted(range(njobs), key=lambda j: -sizes[j])
    with ThreadPoolExecutor(max_workers=threads) as ex:
        built = dict(zip(order, ex.map(one, order)))
    return [dict(arrival=float(arrivals[j]), n=int(sizes[j]), steps=int(steps[j]), kT=float(kts[j]), x=built[j][0],
                 v=built[j][1], topo=built[j][2], seed=20_000 + j, kind="unfold" if j % 2 == 0 else "fold")
            for j in range(njobs)]
