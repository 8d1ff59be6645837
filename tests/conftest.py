import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "sop-cpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsopg.so on the device)")


def gpu_available():
    try:
        import sopg
        n = sopg.C.c_int()
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


_cache = {}


def chain_inputs(tmp_root, n, seed, kT=0.6):
    """Generated native chain (+ Maxwell velocities) written in the reference's formats."""
    import sopg
    key = (n, seed)
    if key not in _cache:
        x = sopg.gen_native(n, seed)
        v = sopg.gen_maxwell(n, kT, seed)
        d = os.path.join(str(tmp_root), f"chain_{n}_{seed}")
        sopg.write_inputs(x, d, vel=v)
        _cache[key] = (x, v, d)
    return _cache[key]


@pytest.fixture(scope="session")
def inputs_root(tmp_path_factory):
    return tmp_path_factory.mktemp("inputs")


def quirk_system(d):
    """Ten beads placed so that both vdw_forces quirks decide (energy.cpp:292-293, 331), written
    in the reference's formats into directory d (no bonds, no angles; explicit master lists):

      natives (file order)  k1 (1,5) r 7, far apart; k2 (2,6) r 4, far apart;
                            k3 (3,7) r 4, d^2 = 3e-6: pair-list position 1, whose skip test reads
                            lj_nat_pdb_dist2[1] = 49 (tol * 49 = 4.9e-6: skipped; with its own
                            r^2 = 16 it would not be);
                            k4 (4,8) r 5, types 2 2, d 5.5: the last native pair-list entry, so
                            rep_tol = sigma_rep2[2][2] * tol = 4.9e-6
      non-natives           (1,9) types 1 1, d^2 = 2e-6: skipped by that rep_tol (its own types
                            would give 1.444e-6: not skipped); (2,10) d 5: an ordinary term.
    Returns (xyz, topology dict for the device)."""
    import os
    X = np.zeros((10, 3))
    X[0] = (0.0, 0.0, 0.0)
    X[8] = (np.sqrt(2e-6), 0.0, 0.0)
    X[4] = (60.0, 0.0, 0.0)
    X[1] = (0.0, 30.0, 0.0)
    X[5] = (0.0, 90.0, 0.0)
    X[9] = (5.0, 30.0, 0.0)
    X[2] = (0.0, 0.0, 40.0)
    X[6] = (np.sqrt(3e-6), 0.0, 40.0)
    X[3] = (40.0, 40.0, 0.0)
    X[7] = (45.5, 40.0, 0.0)
    att = [(1, 5, 7.0, 1, 1), (2, 6, 4.0, 1, 1), (3, 7, 4.0, 1, 1), (4, 8, 5.0, 2, 2)]
    rep = [(1, 9, 1, 1), (2, 10, 1, 1)]
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "bonds.dat"), "w") as f:
        f.write("bonds 0\n")
    with open(os.path.join(d, "angles.dat"), "w") as f:
        f.write("angles 0\n")
    with open(os.path.join(d, "vdw.dat"), "w") as f:
        f.write(f"natt {len(att)} nrep {len(rep)}\n")
        for (i, j, r, it, jt) in att:
            f.write(f"{i} {j} {r!r} {it} {jt}\n")
        for (i, j, it, jt) in rep:
            f.write(f"{i} {j} 20.0 {it} {jt}\n")
    with open(os.path.join(d, "init.dat"), "w") as f:
        f.write("beads 10\n" + "".join(f"{k + 1} {float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n" for k, p in enumerate(X)))
    e = np.zeros(0, np.int32)
    topo = dict(n=10, ib=e, jb=e, r0=np.zeros(0), ia=e, ja=e, ka=e,
                att_i=np.array([a[0] for a in att], np.int32), att_j=np.array([a[1] for a in att], np.int32),
                att_r=np.array([a[2] for a in att]), att_it=np.array([a[3] for a in att], np.int32),
                att_jt=np.array([a[4] for a in att], np.int32),
                rep_i=np.array([r[0] for r in rep], np.int32), rep_j=np.array([r[1] for r in rep], np.int32),
                rep_it=np.array([r[2] for r in rep], np.int32), rep_jt=np.array([r[3] for r in rep], np.int32))
    return X.ravel().copy(), topo
