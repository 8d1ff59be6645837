"""The C-ABI library loads and exports every entry point its headers declare
(no device calls: these run without a GPU)."""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import sopg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b((?:sopg|sopgen)_\w+)\s*\(", src, flags=re.M):
            syms.add(m.group(1))
    return syms


def test_headers_declare_entry_points():
    syms = declared_symbols()
    assert "sopg_force_eval" in syms and "sopg_update_neighbor_list" in syms and "sopgen_native" in syms
    assert len(syms) >= 30


def test_library_exports_every_declared_symbol():
    assert os.path.exists(sopg.LIB_PATH), "libsopg.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", sopg.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(declared_symbols() - exported)
    assert not missing, missing
    lib = ctypes.CDLL(sopg.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert set(sopg.EXPORTS) <= declared_symbols()


def test_no_gpu_errors_are_clean():
    """Without a usable device, context creation fails with an error code (no fallback)."""
    try:
        ctx = sopg.Context()
    except sopg.SopgError:
        return
    ctx.close()  # a GPU is present: nothing to check here


def test_generator_constraints():
    for n, seed in ((76, 1), (500, 2)):
        x = sopg.gen_native(n, seed)
        a, b, c = (ctypes.c_double() for _ in range(3))
        sopg.lib().sopgen_check(n, x.ctypes.data_as(ctypes.c_void_p), ctypes.byref(a), ctypes.byref(b),
                                ctypes.byref(c))
        assert a.value >= 4.0 - 1e-9          # non-bonded pairs >= 4 A
        assert b.value < 1e-9                 # bonds 3.8 A
        assert c.value <= 3.0 * n ** (1 / 3) + 1e-9  # inside the collapse sphere
    s = sopg.gen_saw(300, 4)
    X = s.reshape(-1, 3)
    bond = np.linalg.norm(X[1:] - X[:-1], axis=1)
    assert np.allclose(bond, 3.8, atol=1e-9)
    assert np.array_equal(sopg.gen_native(76, 1), sopg.gen_native(76, 1))  # seeded, deterministic


def test_unfolded_start_keeps_native_pairs_apart():
    """The folding runs' start: a self-avoiding walk (3.8 A bonds, non-bonded pairs >= 4 A)
    within rmax that keeps every native pair at least its native distance apart, where a
    plain walk puts some at about half of it."""
    n = 1500
    nat = sopg.gen_native(n, 7)
    t = sopg.chain_topology(nat)
    i, j, r0 = t["att_i"] - 1, t["att_j"] - 1, t["att_r"]
    x = sopg.gen_unfolded(nat, 8, rmax=120.0)
    a, b, c = (ctypes.c_double() for _ in range(3))
    sopg.lib().sopgen_check(n, x.ctypes.data_as(ctypes.c_void_p), ctypes.byref(a), ctypes.byref(b),
                            ctypes.byref(c))
    assert a.value >= 4.0 - 1e-9 and b.value < 1e-9 and c.value <= 120.0 + 1e-9
    X = x.reshape(-1, 3)
    assert (np.linalg.norm(X[i] - X[j], axis=1) >= r0 - 1e-9).all()
    s = sopg.gen_saw(n, 8).reshape(-1, 3)
    assert (np.linalg.norm(s[i] - s[j], axis=1) / r0).min() < 0.8  # what the floor prevents
    assert np.array_equal(x, sopg.gen_unfolded(nat, 8, rmax=120.0))


def test_written_files_roundtrip(tmp_path):
    """Files written by the generator load (through the oracle's restatement of the
    reference loaders) into exactly the arrays chain_topology() hands the device."""
    from refs import Oracle
    x = sopg.gen_native(60, 3)
    natt = sopg.write_inputs(x, str(tmp_path))
    o = Oracle()
    for k in ("bonds", "angles", "vdw", "init"):
        o.load(k, str(tmp_path / (k + ".dat")))
    t = sopg.chain_topology(x)
    ai, aj = o.lst(4)
    assert natt == len(ai) == len(t["att_i"])
    assert np.array_equal(ai, t["att_i"]) and np.array_equal(aj, t["att_j"])
    assert np.array_equal(o.coords(1), x)
    assert o.geti("ncon_rep") == (60 - 2) * (60 - 3) // 2 - natt


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asking for more exits with a message before
    any device work."""
    r = subprocess.run(["python3", os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr and r.stdout == ""
