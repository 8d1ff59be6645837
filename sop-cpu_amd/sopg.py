"""ctypes binding of libsopg.so (the C ABI declared in include/sopg.h).

This is a thin host-side convenience layer used by the tests and bench.py; the
product interface is the C ABI itself.  Loading fails loudly when the built
library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsopg.so")

OK = 0
CUTOFF_NONE, CUTOFF_NEIGHBOR, CUTOFF_CELL = 0, 1, 2
NOISE_REPLAY, NOISE_PHILOX = 0, 1
LIST_NL_ATT, LIST_NL_REP, LIST_PL_ATT, LIST_PL_REP = 0, 1, 2, 3
PLAN_NONE, PLAN_LOCAL, PLAN_PERIODIC, PLAN_SCAN, PLAN_FILTER = 0, 1, 2, 3, 4
KERNEL_STEP, KERNEL_REBUILD, KERNEL_PAIRLIST, KERNEL_OBSERVE = 0, 1, 2, 3

_D33 = (C.c_double * 3) * 3


class Tables(C.Structure):
    _fields_ = [("coeff_att", _D33), ("coeff_rep", _D33), ("force_coeff_att", _D33),
                ("force_coeff_rep", _D33), ("sigma_rep", _D33), ("k_bnd", C.c_double),
                ("R0sq", C.c_double), ("e_bnd_coeff", C.c_double), ("e_ang_ss_coeff", C.c_double),
                ("f_ang_ss_coeff", C.c_double)]


class ChainParams(C.Structure):
    _fields_ = [("T", C.c_double), ("zeta", C.c_double), ("h", C.c_double),
                ("a1", C.c_double), ("a2", C.c_double), ("a3", C.c_double), ("a4", C.c_double),
                ("a5", C.c_double), ("boxl", C.c_double), ("ncell", C.c_double), ("lcell", C.c_double),
                ("cutoff", C.c_int), ("nnlup", C.c_int), ("nup", C.c_int), ("noise_mode", C.c_int),
                ("seed", C.c_uint64), ("sim_type", C.c_int)]


_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)


class Topology(C.Structure):
    _fields_ = [("nbead", C.c_int),
                ("nbnd", C.c_int), ("ibnd", _IP), ("jbnd", _IP), ("r0bnd", _DP),
                ("nang", C.c_int), ("iang", _IP), ("jang", _IP), ("kang", _IP),
                ("ncon_att", C.c_int), ("att_i", _IP), ("att_j", _IP), ("att_it", _IP), ("att_jt", _IP),
                ("att_r", _DP),
                ("ncon_rep", C.c_int), ("rep_i", _IP), ("rep_j", _IP), ("rep_it", _IP), ("rep_jt", _IP),
                ("bead_type", _IP)]


class Energies(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("e_bnd", "e_ang_ss", "e_vdw_rr_att", "e_vdw_rr_rep",
                                           "e_vdw_rr", "rna_etot")]


class Observables(C.Structure):
    _fields_ = [("Q", C.c_double), ("contct_nat", C.c_int), ("rgsq", C.c_double), ("kinT", C.c_double)]


class Record(C.Structure):
    _fields_ = [("chain", C.c_int), ("istep", C.c_double), ("T", C.c_double), ("kinT", C.c_double),
                ("e_bnd", C.c_double), ("e_ang_ss", C.c_double), ("e_vdw_rr", C.c_double),
                ("rna_etot", C.c_double), ("Q", C.c_double), ("contct_nat", C.c_int), ("rgsq", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_double), ("max_chains", C.c_int), ("chunk", C.c_long), ("lanes", C.c_int),
                ("capacity", C.c_int)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int), ("chain", C.c_int), ("nbead", C.c_long), ("nsteps", C.c_long),
                ("steps_done", C.c_long), ("t_arrival", C.c_double), ("t_admit", C.c_double),
                ("t_finish", C.c_double), ("footprint", C.c_double), ("nrec", C.c_long),
                ("rounds_waited", C.c_long)]


JOB_QUEUED, JOB_ACTIVE, JOB_DONE, JOB_REJECTED = 0, 1, 2, 3

# every entry point of include/sopg.h, include/sopgen.h and include/sopserve.h
EXPORTS = [
    "sopg_create", "sopg_destroy", "sopg_last_error", "sopg_default_tables", "sopg_set_tables",
    "sopg_set_lanes", "sopg_set_capacity", "sopg_add_chain", "sopg_set_chain_params", "sopg_put_state",
    "sopg_get_state", "sopg_nchains", "sopg_chain_nbead", "sopg_update_neighbor_list",
    "sopg_update_cell_list", "sopg_update_pair_list", "sopg_force_eval", "sopg_energy_eval",
    "sopg_underdamped_iteration", "sopg_overdamped_iteration", "sopg_calculate_observables", "sopg_get_list", "sopg_run",
    "sopg_set_step", "sopg_sync", "sopg_list_counts", "sopg_profile", "sopg_profile_read",
    "sopgen_saw", "sopgen_unfolded", "sopgen_native", "sopgen_maxwell", "sopgen_check", "sopgen_write_topology",
    "sopgen_write_xyz", "sopgen_natives", "sopg_force_terms", "sopg_set_binned",
    "sopg_get_lanes", "sopg_autotune", "sopg_reserve", "sopg_remove_chain", "sopg_get_step",
    "sopg_checkpoint", "sopg_restore", "sopg_set_graphs", "sopg_graph_stats", "sopg_set_resident", "sopg_set_skin",
    "sopg_server_create", "sopg_server_destroy", "sopg_server_last_error", "sopg_server_now",
    "sopg_server_footprint", "sopg_server_submit", "sopg_server_poll", "sopg_server_job",
    "sopg_server_job_records", "sopg_server_job_state", "sopg_server_release", "sopg_server_stats",
    "sopg_server_evict", "sopg_run_noalloc", "sopg_rebuild_plan",
]

_lib = None


def lib():
    """Load libsopg.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.sopg_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.sopg_destroy.argtypes = [vp]
        L.sopg_last_error.argtypes = [vp]
        L.sopg_last_error.restype = C.c_char_p
        L.sopg_default_tables.argtypes = [C.POINTER(Tables)]
        L.sopg_set_tables.argtypes = [vp, C.POINTER(Tables)]
        L.sopg_set_lanes.argtypes = [vp, C.c_int, C.c_int]
        L.sopg_set_capacity.argtypes = [vp, C.c_int]
        L.sopg_set_binned.argtypes = [vp, C.c_int]
        L.sopg_get_lanes.argtypes = [vp, _IP, _IP]
        L.sopg_autotune.argtypes = [vp, _IP]
        L.sopg_reserve.argtypes = [vp, C.c_long, C.c_int]
        L.sopg_remove_chain.argtypes = [vp, C.c_int]
        L.sopg_get_step.argtypes = [vp, C.c_int, _DP, _IP, _IP]
        L.sopg_checkpoint.argtypes = [vp, C.c_int, vp, C.c_long, C.POINTER(C.c_long)]
        L.sopg_restore.argtypes = [vp, C.c_int, vp, C.c_long]
        L.sopg_set_graphs.argtypes = [vp, C.c_int]
        L.sopg_set_resident.argtypes = [vp, C.c_int, C.c_int]
        L.sopg_set_skin.argtypes = [vp, C.c_double]
        L.sopg_graph_stats.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.sopg_force_terms.argtypes = [vp, C.c_int, C.c_int, vp]
        L.sopg_overdamped_iteration.argtypes = [vp, C.c_int, vp]
        L.sopg_add_chain.argtypes = [vp, C.POINTER(Topology), C.POINTER(ChainParams), _IP]
        L.sopg_set_chain_params.argtypes = [vp, C.c_int, C.POINTER(ChainParams)]
        L.sopg_put_state.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.sopg_get_state.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.sopg_nchains.argtypes = [vp]
        L.sopg_chain_nbead.argtypes = [vp, C.c_int]
        for f in ("sopg_update_neighbor_list", "sopg_update_cell_list", "sopg_update_pair_list"):
            getattr(L, f).argtypes = [vp, C.c_int, _IP, _IP]
        L.sopg_force_eval.argtypes = [vp, C.c_int, vp]
        L.sopg_energy_eval.argtypes = [vp, C.c_int, C.POINTER(Energies)]
        L.sopg_underdamped_iteration.argtypes = [vp, C.c_int, vp]
        L.sopg_calculate_observables.argtypes = [vp, C.c_int, C.POINTER(Observables)]
        L.sopg_get_list.argtypes = [vp, C.c_int, C.c_int, _IP, _IP, C.c_long, C.POINTER(C.c_long)]
        L.sopg_run.argtypes = [vp, C.c_long, vp, C.POINTER(Record), C.c_long, C.POINTER(C.c_long)]
        L.sopg_run_noalloc.argtypes = [vp, C.c_long, C.POINTER(Record), C.c_long, C.POINTER(C.c_long)]
        L.sopg_set_step.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_int]
        L.sopg_sync.argtypes = [vp]
        L.sopg_list_counts.argtypes = [vp, C.c_int, _IP, _IP, _IP, _IP]
        L.sopg_rebuild_plan.argtypes = [vp, C.c_int, _IP, _IP]
        L.sopg_profile.argtypes = [vp, C.c_int]
        L.sopg_profile_read.argtypes = [vp, C.c_int, _DP, C.POINTER(C.c_long)]
        L.sopg_server_create.argtypes = [C.c_int, C.POINTER(ServerConfig), C.POINTER(vp)]
        L.sopg_server_destroy.argtypes = [vp]
        L.sopg_server_last_error.argtypes = [vp]
        L.sopg_server_last_error.restype = C.c_char_p
        L.sopg_server_now.argtypes = [vp]
        L.sopg_server_now.restype = C.c_double
        L.sopg_server_footprint.argtypes = [vp, C.POINTER(Topology)]
        L.sopg_server_footprint.restype = C.c_double
        L.sopg_server_submit.argtypes = [vp, C.POINTER(Topology), C.POINTER(ChainParams), vp, vp, C.c_long,
                                         C.c_double, C.POINTER(C.c_long)]
        L.sopg_server_poll.argtypes = [vp, _IP, _IP]
        L.sopg_server_job.argtypes = [vp, C.c_long, C.POINTER(JobInfo)]
        L.sopg_server_job_records.argtypes = [vp, C.c_long, C.POINTER(Record), C.c_long, C.POINTER(C.c_long)]
        L.sopg_server_job_state.argtypes = [vp, C.c_long, vp, vp, vp, vp]
        L.sopg_server_release.argtypes = [vp, C.c_long]
        L.sopg_server_evict.argtypes = [vp, C.c_long]
        L.sopg_server_stats.argtypes = [vp, _DP, C.POINTER(C.c_long), _DP]
        L.sopgen_saw.argtypes = [C.c_int, C.c_uint64, vp]
        L.sopgen_unfolded.argtypes = [C.c_int, C.c_uint64, vp, C.c_double, vp]
        L.sopgen_native.argtypes = [C.c_int, C.c_uint64, vp]
        L.sopgen_maxwell.argtypes = [C.c_int, C.c_double, C.c_uint64, vp]
        L.sopgen_check.argtypes = [C.c_int, vp, _DP, _DP, _DP]
        L.sopgen_write_topology.argtypes = [C.c_int, vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                            C.c_double]
        L.sopgen_write_xyz.argtypes = [C.c_int, vp, C.c_char_p]
        L.sopgen_natives.argtypes = [C.c_int, vp, C.c_double, _IP, _IP, _DP, C.c_long]
        L.sopgen_natives.restype = C.c_long
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _ip(a):
    return a.ctypes.data_as(_IP)


def _dp(a):
    return a.ctypes.data_as(_DP)


class SopgError(RuntimeError):
    pass


# ------------------------------------------------------------------ generator

def gen_native(n, seed):
    x = np.zeros(3 * n)
    rc = lib().sopgen_native(n, seed, _ptr(x))
    if rc:
        raise SopgError(f"sopgen_native({n}, {seed}) failed: {rc}")
    return x


def gen_saw(n, seed):
    x = np.zeros(3 * n)
    rc = lib().sopgen_saw(n, seed, _ptr(x))
    if rc:
        raise SopgError(f"sopgen_saw failed: {rc}")
    return x


def gen_unfolded(native, seed, rmax=200.0):
    """Unfolded start for a folding run of the chain whose native structure is `native`: a
    self-avoiding walk within rmax of the origin that keeps every native pair at least its
    native distance apart."""
    native = np.ascontiguousarray(native, dtype=np.float64)
    n = native.size // 3
    x = np.zeros(3 * n)
    rc = lib().sopgen_unfolded(n, seed, _ptr(native), rmax, _ptr(x))
    if rc:
        raise SopgError(f"sopgen_unfolded failed: {rc}")
    return x


def gen_maxwell(n, kT, seed):
    v = np.zeros(3 * n)
    rc = lib().sopgen_maxwell(n, kT, seed, _ptr(v))
    if rc:
        raise SopgError("sopgen_maxwell failed")
    return v


def write_inputs(xyz, directory, vel=None, rcut_nat=8.0):
    """Write bonds/angles/vdw/init (+ native.xyz, vel.xyz) in the reference's formats."""
    n = len(xyz) // 3
    os.makedirs(directory, exist_ok=True)
    p = lambda f: os.path.join(directory, f).encode()
    natt = lib().sopgen_write_topology(n, _ptr(xyz), p("bonds.dat"), p("angles.dat"), p("vdw.dat"),
                                       p("init.dat"), rcut_nat)
    if natt < 0:
        raise SopgError("write_topology failed")
    lib().sopgen_write_xyz(n, _ptr(xyz), p("native.xyz"))
    if vel is not None:
        lib().sopgen_write_xyz(n, _ptr(vel), p("vel.xyz"))
    return natt


def chain_topology(xyz, rcut_nat=8.0):
    """Topology arrays of a generated chain (as the loaders would read its files):
    bonds (i,i+1), angles (i,i+1,i+2), natives = pairs j-i>2 with r < rcut (lexicographic).
    The r values are the %.17g round trip of the native distances, exactly like the files."""
    X = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
    n = len(X)
    ib = np.arange(1, n, dtype=np.int32)
    jb = ib + 1
    # same arithmetic as the generator: sqrt(dx*dx + dy*dy + dz*dz) in order
    D = X[1:] - X[:-1]
    r0 = np.sqrt(D[:, 0] * D[:, 0] + D[:, 1] * D[:, 1] + D[:, 2] * D[:, 2])
    ia = np.arange(1, n - 1, dtype=np.int32)
    ja = ia + 1
    ka = ia + 2
    L = lib()
    cnt = L.sopgen_natives(n, _ptr(X), rcut_nat, None, None, None, 0)
    att_i = np.zeros(max(cnt, 1), np.int32)
    att_j = np.zeros(max(cnt, 1), np.int32)
    att_r = np.zeros(max(cnt, 1))
    L.sopgen_natives(n, _ptr(X), rcut_nat, _ip(att_i), _ip(att_j), _dp(att_r), cnt)
    return dict(n=n, ib=ib, jb=jb, r0=r0, ia=ia, ja=ja, ka=ka, att_i=att_i[:cnt], att_j=att_j[:cnt],
                att_r=att_r[:cnt])


def _dist(X, i, j):
    dx, dy, dz = X[j] - X[i]
    return float(np.sqrt(dx * dx + dy * dy + dz * dz))


def underdamped_coeffs(h=0.05, zeta=50.0):
    """a1..a5 as set_params leaves them for `set dynamics underdamped; set zeta Z; set t_step H`
    (params.cpp:19-32, 84-98)."""
    halfh = h / 2.0
    a1 = h * (1.0 - zeta * halfh)
    a2 = h * halfh
    a3 = (1.0 - h * zeta / 2.0 + (h * zeta) * (h * zeta) / 4.0) / h
    a4 = halfh * (1.0 - h * zeta / 2.0)
    a5 = h / zeta
    return a1, a2, a3, a4, a5


def chain_params(T=0.6, zeta=50.0, h=0.05, boxl=500.0, ncell=55.0, lcell=None, cutoff=CUTOFF_NEIGHBOR,
                 nnlup=50, nup=1000, noise_mode=NOISE_PHILOX, seed=1, coeffs=None, sim_type=1):
    a = coeffs if coeffs is not None else underdamped_coeffs(h, zeta)
    p = ChainParams()
    p.T, p.zeta, p.h = T, zeta, h
    p.a1, p.a2, p.a3, p.a4, p.a5 = a
    p.boxl, p.ncell = boxl, ncell
    p.lcell = boxl / ncell if lcell is None else lcell
    p.cutoff, p.nnlup, p.nup, p.noise_mode, p.seed = cutoff, nnlup, nup, noise_mode, seed
    p.sim_type = sim_type
    return p


def make_topology(topo, implicit=True):
    """Topology struct of a chain_topology() dict; returns (struct, arrays it points into)."""
    t = Topology()
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a
    n = int(topo["n"])
    t.nbead = n
    ib, jb, r0 = arr(topo["ib"], np.int32), arr(topo["jb"], np.int32), arr(topo["r0"], np.float64)
    t.nbnd, t.ibnd, t.jbnd, t.r0bnd = len(ib), _ip(ib), _ip(jb), _dp(r0)
    ia, ja, ka = arr(topo["ia"], np.int32), arr(topo["ja"], np.int32), arr(topo["ka"], np.int32)
    t.nang, t.iang, t.jang, t.kang = len(ia), _ip(ia), _ip(ja), _ip(ka)
    ai, aj, ar = arr(topo["att_i"], np.int32), arr(topo["att_j"], np.int32), arr(topo["att_r"], np.float64)
    at = arr(topo.get("att_it", np.ones(len(ai))), np.int32)
    ajt = arr(topo.get("att_jt", np.ones(len(ai))), np.int32)
    t.ncon_att, t.att_i, t.att_j, t.att_it, t.att_jt, t.att_r = len(ai), _ip(ai), _ip(aj), _ip(at), _ip(ajt), _dp(ar)
    if "rep_i" in topo and not implicit:
        ri, rj = arr(topo["rep_i"], np.int32), arr(topo["rep_j"], np.int32)
        rit = arr(topo.get("rep_it", np.ones(len(ri))), np.int32)
        rjt = arr(topo.get("rep_jt", np.ones(len(ri))), np.int32)
        t.ncon_rep, t.rep_i, t.rep_j, t.rep_it, t.rep_jt = len(ri), _ip(ri), _ip(rj), _ip(rit), _ip(rjt)
    else:
        t.ncon_rep = -1
    if "bead_type" in topo:
        bt = arr(topo["bead_type"], np.int32)
        t.bead_type = _ip(bt)
    return t, keep


class Context:
    """One device context (a ragged batch of chains) on one GPU."""

    def __init__(self, device=0, lanes=0, capacity=0, list_lanes=0, binned=True):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.sopg_create(device, C.byref(h))
        if rc:
            raise SopgError(f"sopg_create failed ({rc}): no usable GPU")
        self.h = h
        self._keep = []
        if lanes or list_lanes:
            self._chk(self.L.sopg_set_lanes(self.h, lanes, list_lanes))
        if capacity:
            self._chk(self.L.sopg_set_capacity(self.h, capacity))
        if not binned:
            self._chk(self.L.sopg_set_binned(self.h, 0))

    def _chk(self, rc):
        if rc != OK:
            raise SopgError(f"sopg error {rc}: {self.L.sopg_last_error(self.h).decode()}")
        return rc

    def close(self):
        if self.h:
            self.L.sopg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_chain(self, topo, params, implicit=True):
        """topo: dict from chain_topology() (+ optional 'rep_i','rep_j' explicit list)."""
        t, keep = make_topology(topo, implicit)
        cid = C.c_int()
        self._chk(self.L.sopg_add_chain(self.h, C.byref(t), C.byref(params), C.byref(cid)))
        del keep
        return cid.value

    def remove_chain(self, chain):
        self._chk(self.L.sopg_remove_chain(self.h, chain))

    def reserve(self, beads, chains):
        self._chk(self.L.sopg_reserve(self.h, beads, chains))

    def autotune(self):
        """Pick the fastest step-kernel lane count for the current chains; returns it."""
        lanes = C.c_int()
        self._chk(self.L.sopg_autotune(self.h, C.byref(lanes)))
        return lanes.value

    def set_skin(self, skin):
        self._chk(self.L.sopg_set_skin(self.h, float(skin)))

    def set_resident(self, max_beads, lanes=0):
        self._chk(self.L.sopg_set_resident(self.h, max_beads, lanes))

    def set_graphs(self, enable):
        self._chk(self.L.sopg_set_graphs(self.h, 1 if enable else 0))

    def graph_stats(self):
        a, b = C.c_long(), C.c_long()
        self._chk(self.L.sopg_graph_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def lanes(self):
        a, b = C.c_int(), C.c_int()
        self._chk(self.L.sopg_get_lanes(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def nchains(self):
        return self.L.sopg_nchains(self.h)

    def put_state(self, chain, pos=None, unc=None, vel=None, force=None):
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (pos, unc, vel, force)]
        self._chk(self.L.sopg_put_state(self.h, chain, *[_ptr(x) for x in a]))

    def get_state(self, chain):
        n = self.L.sopg_chain_nbead(self.h, chain)
        out = [np.zeros(3 * n) for _ in range(4)]
        self._chk(self.L.sopg_get_state(self.h, chain, *[_ptr(x) for x in out]))
        return dict(pos=out[0], unc=out[1], vel=out[2], force=out[3])

    def update_neighbor_list(self, chain=0):
        a, r = C.c_int(), C.c_int()
        self._chk(self.L.sopg_update_neighbor_list(self.h, chain, C.byref(a), C.byref(r)))
        return a.value, r.value

    def update_cell_list(self, chain=0):
        a, r = C.c_int(), C.c_int()
        self._chk(self.L.sopg_update_cell_list(self.h, chain, C.byref(a), C.byref(r)))
        return a.value, r.value

    def update_pair_list(self, chain=0):
        a, r = C.c_int(), C.c_int()
        self._chk(self.L.sopg_update_pair_list(self.h, chain, C.byref(a), C.byref(r)))
        return a.value, r.value

    def force_eval(self, chain=0, noise=None):
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
        self._chk(self.L.sopg_force_eval(self.h, chain, _ptr(nz)))

    def underdamped_iteration(self, chain=0, noise=None):
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
        self._chk(self.L.sopg_underdamped_iteration(self.h, chain, _ptr(nz)))

    def overdamped_iteration(self, chain=0, noise=None):
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
        self._chk(self.L.sopg_overdamped_iteration(self.h, chain, _ptr(nz)))

    def energy_eval(self, chain=0):
        e = Energies()
        self._chk(self.L.sopg_energy_eval(self.h, chain, C.byref(e)))
        return {f: getattr(e, f) for f, _ in Energies._fields_}

    def observables(self, chain=0):
        o = Observables()
        self._chk(self.L.sopg_calculate_observables(self.h, chain, C.byref(o)))
        return dict(Q=o.Q, contct_nat=o.contct_nat, rgsq=o.rgsq, kinT=o.kinT)

    def get_list(self, which, chain=0):
        n = C.c_long()
        self._chk(self.L.sopg_get_list(self.h, chain, which, None, None, 0, C.byref(n)))
        i = np.zeros(max(n.value, 1), np.int32)
        j = np.zeros(max(n.value, 1), np.int32)
        self._chk(self.L.sopg_get_list(self.h, chain, which, _ip(i), _ip(j), n.value, C.byref(n)))
        return i[:n.value], j[:n.value]

    def _records(self, rec_cap):
        # one record buffer per context, reused across runs (allocating and
        # zeroing megabytes per call would land inside every timed run)
        rec_cap = max(rec_cap, 1)
        if getattr(self, "_recbuf", None) is None or len(self._recbuf) < rec_cap:
            self._recbuf = (Record * rec_cap)()
        return self._recbuf

    def run(self, nsteps, noise=None, rec_cap=100000):
        recs = self._records(rec_cap)
        nr = C.c_long()
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
        self._chk(self.L.sopg_run(self.h, nsteps, _ptr(nz), recs, rec_cap, C.byref(nr)))
        out = []
        for k in range(min(nr.value, rec_cap)):
            r = recs[k]
            out.append({f: getattr(r, f) for f, _ in Record._fields_})
        return out

    def run_noalloc(self, nsteps, rec_cap=100000):
        """sopg_run_noalloc: the same steps with no device allocation (device noise only)."""
        recs = self._records(rec_cap)
        nr = C.c_long()
        self._chk(self.L.sopg_run_noalloc(self.h, nsteps, recs, rec_cap, C.byref(nr)))
        return [{f: getattr(recs[k], f) for f, _ in Record._fields_} for k in range(min(nr.value, rec_cap))]

    def set_step(self, chain, istep, inlup, iup):
        self._chk(self.L.sopg_set_step(self.h, chain, istep, inlup, iup))

    def checkpoint(self, chain):
        """Lossless checkpoint of one chain as bytes."""
        n = C.c_long()
        self._chk(self.L.sopg_checkpoint(self.h, chain, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._chk(self.L.sopg_checkpoint(self.h, chain, buf, n.value, C.byref(n)))
        return buf.raw[:n.value]

    def restore(self, chain, blob):
        buf = C.create_string_buffer(blob, len(blob))
        self._chk(self.L.sopg_restore(self.h, chain, buf, len(blob)))

    def get_step(self, chain):
        istep, inlup, iup = C.c_double(), C.c_int(), C.c_int()
        self._chk(self.L.sopg_get_step(self.h, chain, C.byref(istep), C.byref(inlup), C.byref(iup)))
        return istep.value, inlup.value, iup.value

    def sync(self):
        self._chk(self.L.sopg_sync(self.h))

    def list_counts(self, chain=0):
        v = [C.c_int() for _ in range(4)]
        self._chk(self.L.sopg_list_counts(self.h, chain, *[C.byref(x) for x in v]))
        return dict(zip(("nnl_att", "nnl_rep", "nil_att", "nil_rep"), (x.value for x in v)))

    def rebuild_plan(self, chain=0):
        """(plan, cells per axis) of the chain's latest list rebuild: one of PLAN_*."""
        p, n = C.c_int(), C.c_int()
        self._chk(self.L.sopg_rebuild_plan(self.h, chain, C.byref(p), C.byref(n)))
        return p.value, n.value

    def profile(self, enable=True):
        self._chk(self.L.sopg_profile(self.h, 1 if enable else 0))

    def profile_read(self, kernel):
        ms, n = C.c_double(), C.c_long()
        self._chk(self.L.sopg_profile_read(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Server:
    """Continuous-batching job server (include/sopserve.h) on one GPU."""

    def __init__(self, device=0, hbm_budget=0.0, max_chains=0, chunk=0, lanes=0, capacity=0):
        self.L = lib()
        cfg = ServerConfig(float(hbm_budget), max_chains, chunk, lanes, capacity)
        h = C.c_void_p()
        rc = self.L.sopg_server_create(device, C.byref(cfg), C.byref(h))
        if rc:
            raise SopgError(f"sopg_server_create failed ({rc})")
        self.h = h

    def _chk(self, rc):
        if rc != OK:
            raise SopgError(f"server error {rc}: {self.L.sopg_server_last_error(self.h).decode()}")
        return rc

    def close(self):
        if self.h:
            self.L.sopg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self):
        return self.L.sopg_server_now(self.h)

    def footprint(self, topo):
        t, keep = make_topology(topo)
        return self.L.sopg_server_footprint(self.h, C.byref(t))

    def submit(self, topo, params, pos, vel, nsteps, arrival=-1.0):
        t, keep = make_topology(topo)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        job = C.c_long()
        self._chk(self.L.sopg_server_submit(self.h, C.byref(t), C.byref(params), _ptr(pos), _ptr(vel), nsteps,
                                            float(arrival), C.byref(job)))
        return job.value

    def poll(self):
        a, q = C.c_int(), C.c_int()
        self._chk(self.L.sopg_server_poll(self.h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def job(self, job):
        i = JobInfo()
        self._chk(self.L.sopg_server_job(self.h, job, C.byref(i)))
        return {f: getattr(i, f) for f, _ in JobInfo._fields_}

    def records(self, job):
        n = C.c_long()
        self._chk(self.L.sopg_server_job_records(self.h, job, None, 0, C.byref(n)))
        recs = (Record * max(n.value, 1))()
        self._chk(self.L.sopg_server_job_records(self.h, job, recs, n.value, C.byref(n)))
        return [{f: getattr(recs[k], f) for f, _ in Record._fields_} for k in range(n.value)]

    def state(self, job):
        n = self.job(job)["nbead"]
        out = [np.zeros(3 * n) for _ in range(4)]
        self._chk(self.L.sopg_server_job_state(self.h, job, *[_ptr(x) for x in out]))
        return dict(pos=out[0], unc=out[1], vel=out[2], force=out[3])

    def release(self, job):
        self._chk(self.L.sopg_server_release(self.h, job))

    def evict(self, job):
        self._chk(self.L.sopg_server_evict(self.h, job))

    def stats(self):
        bs, busy = C.c_double(), C.c_double()
        rounds = C.c_long()
        self._chk(self.L.sopg_server_stats(self.h, C.byref(bs), C.byref(rounds), C.byref(busy)))
        return dict(bead_steps=bs.value, rounds=rounds.value, busy_s=busy.value)
