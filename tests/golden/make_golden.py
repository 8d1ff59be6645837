#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Runs SOP-CPU itself (oracle/_ref/libsopref.so, compiled from the reference's
sources by oracle/Makefile) on generated inputs and stores what it computes:
Ran_Gen known-answer streams, neighbor / cell / pair lists, energies, forces,
`update.out` records and the state after a run.  Inputs are committed next to
the fixtures (n76/, n120cl/), so the fixtures do not depend on the generator.

Usage (in the development container, where the reference sources exist):
    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "sop-cpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import sopg  # noqa: E402
from refs import Ref, setup_system, std_cmds  # noqa: E402

CASES = {
    # name: (n, seed, cutoff, nnlup, nup, steps, extra set commands)
    "n76": (76, 7, "neighborlist", 50, 100, 1000, []),
    "n120cl": (120, 13, "celllist", 20, 50, 300, [("boxl", "504"), ("ncell", "42")]),
}


def make_inputs(name, n, seed):
    d = os.path.join(HERE, name)
    if not os.path.exists(os.path.join(d, "vdw.dat")):
        x = sopg.gen_native(n, seed)
        v = sopg.gen_maxwell(n, 0.6, seed)
        sopg.write_inputs(x, d, vel=v)
        os.remove(os.path.join(d, "native.xyz"))
    return d


def read_xyz(path):
    return np.loadtxt(path)[:, 1:4].ravel()


def run_case(name):
    n, seed, cutoff, nnlup, nup, steps, extra = CASES[name]
    d = make_inputs(name, n, seed)
    r = Ref()
    tmp = tempfile.mkdtemp()
    setup_system(r, d, std_cmds(cutoff=cutoff, nnlup=nnlup, nup=nup, extra=extra))
    v = read_xyz(os.path.join(d, "vel.xyz"))
    r.set_coords(2, v)
    out = {}
    upd = "update_neighbor_list" if cutoff == "neighborlist" else "update_cell_list"
    r.call(upd)
    r.call("update_pair_list")
    for w, k in enumerate(("nl_att", "nl_rep", "pl_att", "pl_rep")):
        i, j = r.lst(w)
        out[k] = np.stack([i, j]).astype(np.int32)
    r.call("set_potential_forces")
    r.call("energy_eval")
    out["energy0"] = np.array([r.getd(k) for k in ("e_bnd", "e_ang_ss", "e_vdw_rr", "rna_etot")])
    r.reset_rng(-101, tmp)
    r.call("force_eval")
    out["force0"] = r.coords(3)
    recs = []
    inlup, iup = 1, 1
    for istep in range(1, steps + 1):
        if inlup % nnlup == 0:
            r.call(upd)
            inlup = 0
        inlup += 1
        r.call("update_pair_list")
        r.call("underdamped_iteration")
        if iup % nup == 0:
            r.call("energy_eval")
            r.call("calculate_observables")
            recs.append([istep, r.getd("kinT"), r.getd("e_bnd"), r.getd("e_ang_ss"), r.getd("e_vdw_rr"),
                         r.getd("rna_etot"), r.getd("Q"), r.geti("contct_nat"), r.getd("rgsq")])
            iup = 0
        iup += 1
    out["records"] = np.array(recs)
    for w, k in enumerate(("pos", "unc", "vel", "force")):
        out[k] = r.coords(w)
    for w, k in enumerate(("nl_att_end", "nl_rep_end", "pl_att_end", "pl_rep_end")):
        i, j = r.lst(w)
        out[k] = np.stack([i, j]).astype(np.int32)
    return out


def kat():
    r = Ref()
    tmp = tempfile.mkdtemp()
    r.reset_rng(-101, tmp)
    ran2 = np.array([r.L.ref_ran2() for _ in range(64)])
    r.reset_rng(-101, tmp)
    g = np.zeros(1024)
    r.L.ref_gasdev_n(g.ctypes.data_as(r.L.ref_gasdev_n.argtypes[0]), len(g))
    return dict(ran2_seed_m101=ran2, gasdev_seed_m101=g)


if __name__ == "__main__":
    assert Ref.available(), "build oracle/_ref first: make -C oracle ref"
    np.savez_compressed(os.path.join(HERE, "ran_gen_kat.npz"), **kat())
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, f"golden_{name}.npz"), **run_case(name))
        print("wrote", name)
