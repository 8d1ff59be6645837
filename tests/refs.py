"""ctypes access to the checkers: the C oracle (oracle/liboracle.so) and the
compiled reference (oracle/_ref/libsopref.so).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsopref.so")

_vp = C.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp)


class Oracle:
    """One oracle system (orc_sys) -- the CPU restatement of the reference."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = C.CDLL(ORACLE_SO)
            L.orc_create.restype = _vp
            L.orc_destroy.argtypes = [_vp]
            for f in ("orc_set", "orc_load"):
                getattr(L, f).argtypes = [_vp, C.c_char_p, C.c_char_p]
            L.orc_get_int.argtypes = [_vp, C.c_char_p]
            L.orc_get_double.argtypes = [_vp, C.c_char_p]
            L.orc_get_double.restype = C.c_double
            L.orc_get_coords.argtypes = [_vp, C.c_int, _vp]
            L.orc_set_coords.argtypes = [_vp, C.c_int, _vp]
            L.orc_get_list.argtypes = [_vp, C.c_int, _vp, _vp]
            L.orc_set_noise.argtypes = [_vp, _vp]
            L.orc_gasdev_n.argtypes = [_vp, _vp, C.c_long]
            L.orc_set_seed.argtypes = [_vp, C.c_int]
            for f in ("orc_update_neighbor_list", "orc_update_cell_list", "orc_update_pair_list",
                      "orc_force_eval", "orc_energy_eval", "orc_underdamped_iteration",
                      "orc_overdamped_iteration", "orc_calculate_observables", "orc_clear_forces",
                      "orc_random_force", "orc_fene_forces", "orc_angular_forces", "orc_vdw_forces"):
                getattr(L, f).argtypes = [_vp]
            L.orc_rnd.argtypes = [C.c_double]
            L.orc_rnd.restype = C.c_double
            L.orc_pow_cr.argtypes = [C.c_double, C.c_int]
            L.orc_pow_cr.restype = C.c_double
            L.orc_sys_ran2.argtypes = [_vp]
            L.orc_sys_ran2.restype = C.c_double
            cls._lib = L
        return cls._lib

    @classmethod
    def set_pow_mode(cls, mode):
        C.c_int.in_dll(cls.lib(), "orc_pow_mode").value = mode

    def __init__(self):
        self.L = self.lib()
        self.s = _vp(self.L.orc_create())
        self._noise = None

    def __del__(self):
        try:
            self.L.orc_destroy(self.s)
        except Exception:
            pass

    def set(self, key, val):
        self.L.orc_set(self.s, key.encode(), str(val).encode())

    def load(self, kind, path):
        rc = self.L.orc_load(self.s, kind.encode(), str(path).encode())
        assert rc == 0, (kind, path, rc)

    def geti(self, name):
        return self.L.orc_get_int(self.s, name.encode())

    def getd(self, name):
        return self.L.orc_get_double(self.s, name.encode())

    def coords(self, which):  # 0 pos 1 unc 2 vel 3 force 4 incr
        n = self.geti("nbead")
        a = np.zeros(3 * n)
        self.L.orc_get_coords(self.s, which, _p(a))
        return a

    def set_coords(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.L.orc_set_coords(self.s, which, _p(a))

    def lst(self, which):  # 0 NL att 1 NL rep 2 PL att 3 PL rep 4 master att 5 master rep
        n = self.L.orc_get_list(self.s, which, None, None)
        i = np.zeros(max(n, 1), np.int32)
        j = np.zeros(max(n, 1), np.int32)
        self.L.orc_get_list(self.s, which, _p(i), _p(j))
        return i[:n], j[:n]

    def gasdev(self, n):
        out = np.zeros(n)
        self.L.orc_gasdev_n(self.s, _p(out), n)
        return out

    def set_noise(self, noise):
        self._noise = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
        self.L.orc_set_noise(self.s, None if noise is None else _p(self._noise))

    def call(self, fn):
        getattr(self.L, "orc_" + fn)(self.s)


class Ref:
    """The compiled reference (global state: one system per process)."""

    _lib = None

    @classmethod
    def available(cls):
        return os.path.exists(REF_SO)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = C.CDLL(REF_SO)
            L.ref_set.argtypes = [C.c_char_p, C.c_char_p]
            L.ref_load.argtypes = [C.c_char_p, C.c_char_p]
            L.ref_get_int.argtypes = [C.c_char_p]
            L.ref_get_double.argtypes = [C.c_char_p]
            L.ref_get_double.restype = C.c_double
            L.ref_get_coords.argtypes = [C.c_int, _vp]
            L.ref_set_coords.argtypes = [C.c_int, _vp]
            L.ref_get_list.argtypes = [C.c_int, _vp, _vp]
            L.ref_gasdev_n.argtypes = [_vp, C.c_long]
            L.ref_set_seed.argtypes = [C.c_int]
            L.ref_rng_restart.argtypes = [C.c_char_p]
            L.ref_ran2.restype = C.c_double
            L.ref_gasdev.restype = C.c_double
            cls._lib = L
        return cls._lib

    def __init__(self):
        self.L = self.lib()
        self.L.ref_init()

    def reset_rng(self, seed, tmpdir):
        """Fresh generator state (mseed = seed, no cached deviate) via Ran_Gen::restart."""
        path = os.path.join(str(tmpdir), "ran_gen_state.dat")
        with open(path, "w") as f:
            f.write(f"{seed}\n0\n0.0\n123456789\n0\n" + "0\n" * 32)
        self.L.ref_rng_restart(path.encode())

    def set(self, key, val):
        self.L.ref_set(key.encode(), str(val).encode())

    def load(self, kind, path):
        self.L.ref_load(kind.encode(), str(path).encode())

    def geti(self, name):
        return self.L.ref_get_int(name.encode())

    def getd(self, name):
        return self.L.ref_get_double(name.encode())

    def coords(self, which):
        n = self.geti("nbead")
        a = np.zeros(3 * n)
        self.L.ref_get_coords(which, _p(a))
        return a

    def set_coords(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.L.ref_set_coords(which, _p(a))

    def lst(self, which):
        n = self.L.ref_get_list(which, None, None)
        i = np.zeros(max(n, 1), np.int32)
        j = np.zeros(max(n, 1), np.int32)
        self.L.ref_get_list(which, _p(i), _p(j))
        return i[:n], j[:n]

    def call(self, fn):
        getattr(self.L, "ref_" + fn)()


def setup_system(sysobj, d, cmds, files=("bonds", "angles", "vdw", "init")):
    """Apply `set` commands (ordered) and load the four input files from directory d."""
    for k, v in cmds:
        sysobj.set(k, v)
    for kind in files:
        sysobj.load(kind, os.path.join(d, kind + ".dat"))


# Standard dynamics settings of the workloads, in the order the scripts must use (set_params is
# order-sensitive: `set zeta` resets h, then `set t_step` re-derives a1..a5).
def std_cmds(temp=0.6, cutoff="neighborlist", nnlup=50, nup=1000, extra=()):
    c = [("dynamics", "underdamped"), ("zeta", "50"), ("t_step", "0.05"), ("temp", str(temp)),
         ("cutofftype", cutoff), ("nnlup", str(nnlup)), ("nup", str(nup))]
    return c + list(extra)
