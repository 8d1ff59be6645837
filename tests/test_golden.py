"""Golden fixtures written by the compiled reference (tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture bit for bit (this pins the oracle
without needing the reference at test time).  GPU: the device path reproduces
them through the C ABI -- lists exactly, forces and the final state bit for bit
(glibc pow vs correctly rounded pow never changed a bit on these inputs;
bars: forces 1e-12 relative, positions 1e-9 A), records within tolerance.
"""
import os

import numpy as np
import pytest

import sopg
from refs import Oracle, setup_system, std_cmds

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {
    "n76": (76, "neighborlist", 50, 100, 1000, []),
    "n120cl": (120, "celllist", 20, 50, 300, [("boxl", "504"), ("ncell", "42")]),
}


def load(name):
    return dict(np.load(os.path.join(HERE, f"golden_{name}.npz")))


def read_xyz(path):
    return np.loadtxt(path)[:, 1:4].ravel()


def test_ran_gen_known_answers():
    k = np.load(os.path.join(HERE, "ran_gen_kat.npz"))
    o = Oracle()
    o.L.orc_set_seed(o.s, -101)
    assert np.array_equal([o.L.orc_sys_ran2(o.s) for _ in range(64)], k["ran2_seed_m101"])
    o = Oracle()
    assert np.array_equal(o.gasdev(1024), k["gasdev_seed_m101"])
    # first values as published by the generator for this seed: a loose sanity anchor
    assert 0.0 < k["ran2_seed_m101"].min() and k["ran2_seed_m101"].max() < 1.0


def _oracle_case(name):
    n, cutoff, nnlup, nup, steps, extra = CASES[name]
    d = os.path.join(HERE, name)
    o = Oracle()
    setup_system(o, d, std_cmds(cutoff=cutoff, nnlup=nnlup, nup=nup, extra=extra))
    o.set_coords(2, read_xyz(os.path.join(d, "vel.xyz")))
    return o, d


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_golden(name):
    g = load(name)
    n, cutoff, nnlup, nup, steps, extra = CASES[name]
    o, d = _oracle_case(name)
    upd = "update_neighbor_list" if cutoff == "neighborlist" else "update_cell_list"
    o.call(upd)
    o.call("update_pair_list")
    for w, k in enumerate(("nl_att", "nl_rep", "pl_att", "pl_rep")):
        assert np.array_equal(np.stack(o.lst(w)), g[k]), k
    o.call("energy_eval")
    assert np.array_equal([o.getd(k) for k in ("e_bnd", "e_ang_ss", "e_vdw_rr", "rna_etot")], g["energy0"])
    o.set_noise(Oracle().gasdev(3 * n * (steps + 1)))
    o.call("force_eval")
    assert np.array_equal(o.coords(3), g["force0"])
    recs = []
    inlup = iup = 1
    for istep in range(1, steps + 1):
        if inlup % nnlup == 0:
            o.call(upd)
            inlup = 0
        inlup += 1
        o.call("update_pair_list")
        o.call("underdamped_iteration")
        if iup % nup == 0:
            o.call("energy_eval")
            o.call("calculate_observables")
            recs.append([istep, o.getd("kinT"), o.getd("e_bnd"), o.getd("e_ang_ss"), o.getd("e_vdw_rr"),
                         o.getd("rna_etot"), o.getd("Q"), o.geti("contct_nat"), o.getd("rgsq")])
            iup = 0
        iup += 1
    assert np.array_equal(np.array(recs), g["records"])
    for w, k in enumerate(("pos", "unc", "vel", "force")):
        assert np.array_equal(o.coords(w), g[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_device_reproduces_golden(name):
    g = load(name)
    n, cutoff, nnlup, nup, steps, extra = CASES[name]
    d = os.path.join(HERE, name)
    o, _ = _oracle_case(name)
    x = o.coords(1)
    v = read_xyz(os.path.join(d, "vel.xyz"))
    topo = sopg.chain_topology(x)
    ai, aj = o.lst(4)
    assert np.array_equal(ai, topo["att_i"]) and np.array_equal(aj, topo["att_j"])
    ex = dict(extra)
    boxl = float(ex.get("boxl", 500.0))
    ncell = float(ex.get("ncell", 55.0))
    cut = sopg.CUTOFF_NEIGHBOR if cutoff == "neighborlist" else sopg.CUTOFF_CELL
    ctx = sopg.Context()
    ctx.add_chain(topo, sopg.chain_params(boxl=boxl, ncell=ncell, cutoff=cut, nnlup=nnlup, nup=nup,
                                          noise_mode=sopg.NOISE_REPLAY))
    ctx.put_state(0, pos=x, unc=x, vel=v, force=np.zeros_like(x))
    print(f"[gpu] golden {name}", flush=True)
    if cut == sopg.CUTOFF_NEIGHBOR:
        ctx.update_neighbor_list()
    else:
        ctx.update_cell_list()
    ctx.update_pair_list()
    for w, k in enumerate(("nl_att", "nl_rep", "pl_att", "pl_rep")):
        assert np.array_equal(np.stack(ctx.get_list(w)), g[k]), k
    e = ctx.energy_eval()
    for i, k in enumerate(("e_bnd", "e_ang_ss", "e_vdw_rr", "rna_etot")):
        assert abs(e[k] - g["energy0"][i]) <= 1e-10 * abs(g["energy0"][i]), k
    noise = Oracle().gasdev(3 * n * (steps + 1))
    ctx.force_eval(0, noise[:3 * n])
    f = ctx.get_state(0)["force"]
    rel = np.linalg.norm((f - g["force0"]).reshape(-1, 3), axis=1) / np.linalg.norm(g["force0"].reshape(-1, 3), axis=1)
    assert rel.max() <= 1e-12
    recs = ctx.run(steps, noise[3 * n:])
    s = ctx.get_state(0)
    assert np.abs(s["unc"] - g["unc"]).max() <= 1e-9
    assert np.abs(s["pos"] - g["pos"]).max() <= 1e-9
    dr = np.array([[r["istep"], r["kinT"], r["e_bnd"], r["e_ang_ss"], r["e_vdw_rr"], r["rna_etot"], r["Q"],
                    r["contct_nat"], r["rgsq"]] for r in recs])
    assert dr.shape == g["records"].shape
    assert np.array_equal(dr[:, [0, 6, 7]], g["records"][:, [0, 6, 7]])
    np.testing.assert_allclose(dr[:, [2, 3, 4, 5]], g["records"][:, [2, 3, 4, 5]], rtol=1e-10)
    np.testing.assert_allclose(dr[:, [1, 8]], g["records"][:, [1, 8]], rtol=1e-12)
    for w, k in enumerate(("nl_att_end", "nl_rep_end", "pl_att_end", "pl_rep_end")):
        assert np.array_equal(np.stack(ctx.get_list(w)), g[k]), k
