"""Pins the oracle (CPU restatement) before it is trusted as the GPU checker.

* against the compiled reference (oracle/_ref, built from SOP-CPU's own sources):
  bit-identical lists, forces, energies, observables and 1000-step trajectories;
* against closed forms and brute force, chosen so that a plausible mistake
  (a dropped term, a wrong sign or index, a transposed operand) fails one;
* against the committed golden fixtures (tests/golden/), which also pin the
  reference's Ran_Gen stream as a known-answer test.
"""
import math
import os

import numpy as np
import pytest

import sopg
from conftest import chain_inputs
from refs import Oracle, Ref, setup_system, std_cmds

needs_ref = pytest.mark.skipif(not Ref.available(), reason="oracle/_ref not built")


def both(root, n, seed, cmds):
    x, v, d = chain_inputs(root, n, seed)
    o = Oracle()
    r = Ref()
    setup_system(o, d, cmds)
    setup_system(r, d, cmds)
    o.set_coords(2, v)
    r.set_coords(2, v)
    return x, v, d, o, r


def same_lists(o, r, which):
    a, b = o.lst(which), r.lst(which)
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@needs_ref
@pytest.mark.parametrize("cutoff", ["neighborlist", "celllist"])
def test_oracle_equals_reference_trajectory(inputs_root, cutoff):
    """Oracle vs compiled reference: bit-identical for 400 steps (lists, state, energies, observables)."""
    n = 120
    extra = [("boxl", "504"), ("ncell", "42")] if cutoff == "celllist" else []
    x, v, d, o, r = both(inputs_root, n, 13, std_cmds(cutoff=cutoff, nnlup=20, nup=50, extra=extra))
    upd = "update_neighbor_list" if cutoff == "neighborlist" else "update_cell_list"
    noise = Oracle().gasdev(3 * n * 401)
    o.set_noise(noise)
    r.reset_rng(-101, inputs_root)  # the reference draws its own stream: the same one
    for f in (upd, "update_pair_list"):
        o.call(f)
        r.call(f)
    for w in range(6):
        assert same_lists(o, r, w), w
    r.call("set_potential_forces")
    for f in ("energy_eval", "force_eval"):
        o.call(f)
        r.call(f)
    assert np.array_equal(o.coords(3), r.coords(3))
    inlup = 1
    for step in range(1, 401):
        if inlup % 20 == 0:
            o.call(upd)
            r.call(upd)
            inlup = 0
        inlup += 1
        o.call("update_pair_list")
        r.call("update_pair_list")
        o.call("underdamped_iteration")
        r.call("underdamped_iteration")
        if step % 50 == 0:
            for f in ("energy_eval", "calculate_observables"):
                o.call(f)
                r.call(f)
            for k in ("e_bnd", "e_ang_ss", "e_vdw_rr", "rna_etot", "Q", "rgsq", "kinT"):
                assert o.getd(k) == r.getd(k), (step, k)
    for w in range(4):
        assert np.array_equal(o.coords(w), r.coords(w)), w
    for w in range(4):
        assert same_lists(o, r, w), w


def od_cmds(temp=0.6, cutoff="neighborlist", nnlup=20, nup=1000, extra=()):
    """Overdamped dynamics settings (params.cpp:30-33: h = 0.02, a5 = h / zeta)."""
    c = [("dynamics", "overdamped"), ("zeta", "50"), ("temp", str(temp)), ("cutofftype", cutoff),
         ("nnlup", str(nnlup)), ("nup", str(nup))]
    return c + list(extra)


@needs_ref
@pytest.mark.parametrize("cutoff", ["neighborlist", "celllist"])
def test_oracle_equals_reference_overdamped(inputs_root, cutoff):
    """Overdamped BD (sop.cpp:331-363, 365-471): oracle vs compiled reference bit-identical for
    400 steps -- increments, state, lists, energies and observables (kinT from the increments,
    sop.cpp:271-278).  The schedule is overdamped_ctrl's: its first rebuild falls on step 1
    (the global inlup starts at 0, sop.cpp:428)."""
    n = 120
    extra = [("boxl", "504"), ("ncell", "42")] if cutoff == "celllist" else []
    x, v, d, o, r = both(inputs_root, n, 31, od_cmds(cutoff=cutoff, extra=extra))
    assert o.geti("sim_type") == 2 and r.geti("sim_type") == 2
    assert o.getd("a5") == r.getd("a5") == 0.02 / 50.0
    upd = "update_neighbor_list" if cutoff == "neighborlist" else "update_cell_list"
    noise = Oracle().gasdev(3 * n * 401)
    o.set_noise(noise)
    r.reset_rng(-101, inputs_root)
    for f in (upd, "update_pair_list"):
        o.call(f)
        r.call(f)
    r.call("set_potential_forces")
    for f in ("energy_eval", "force_eval"):
        o.call(f)
        r.call(f)
    assert np.array_equal(o.coords(3), r.coords(3))
    inlup = 0
    for step in range(1, 401):
        if inlup % 20 == 0:
            o.call(upd)
            r.call(upd)
            inlup = 0
        inlup += 1
        o.call("update_pair_list")
        r.call("update_pair_list")
        o.call("overdamped_iteration")
        r.call("overdamped_iteration")
        if step % 50 == 0 or step in (1, 7, 333):
            assert np.array_equal(o.coords(4), r.coords(4)), step  # incr
            for f in ("energy_eval", "calculate_observables"):
                o.call(f)
                r.call(f)
            for k in ("e_bnd", "e_ang_ss", "e_vdw_rr", "rna_etot", "Q", "rgsq", "kinT"):
                assert o.getd(k) == r.getd(k), (step, k)
            assert o.getd("kinT") > 0.0
    for w in range(5):
        assert np.array_equal(o.coords(w), r.coords(w)), w
    for w in range(4):
        assert same_lists(o, r, w), w
    # the dynamics moved the chain: the lists are not the native ones any more
    assert not np.array_equal(o.coords(1), x)


def test_overdamped_free_particle_closed_form(tmp_path):
    """F = const: one overdamped step moves a bead by a5 F (sop.cpp:339-351); kinT is
    zeta / (2 h) * |incr|^2 / 3 (sop.cpp:271-278)."""
    with open(tmp_path / "bonds.dat", "w") as fh:
        fh.write("bonds 0\n")
    with open(tmp_path / "angles.dat", "w") as fh:
        fh.write("angles 0\n")
    with open(tmp_path / "vdw.dat", "w") as fh:
        fh.write("natt 0 nrep 0\n")
    with open(tmp_path / "init.dat", "w") as fh:
        fh.write("beads 1\n1 1.0 2.0 3.0\n")
    o = Oracle()
    setup_system(o, str(tmp_path), [("dynamics", "overdamped"), ("zeta", "50"), ("temp", "0.5")])
    g = np.array([0.7, -1.1, 0.3])
    o.set_noise(np.concatenate([g, g]))
    F0 = np.array([2.0, -4.0, 6.0])
    o.set_coords(3, F0)
    a5 = 0.02 / 50.0
    o.call("overdamped_iteration")
    assert np.array_equal(o.coords(4), a5 * F0)
    assert np.array_equal(o.coords(1), np.array([1.0, 2.0, 3.0]) + a5 * F0)
    var = math.sqrt(2.0 * 0.5 * 50.0 / 0.02)
    assert np.allclose(o.coords(3), var * g, rtol=1e-15)  # the new force: the noise term alone
    o.call("calculate_observables")
    want = (a5 * F0 * (a5 * F0)).sum() * 50.0 / (2.0 * 0.02) / 3.0
    assert o.getd("kinT") == pytest.approx(want, rel=1e-14)


def _quirk_forces(sysobj):
    for f in ("update_neighbor_list", "update_pair_list", "clear_forces", "vdw_forces"):
        sysobj.call(f)
    return sysobj.coords(3).reshape(-1, 3)


def test_vdw_quirks_closed_form(tmp_path):
    """Both skip tests of vdw_forces as coded (energy.cpp:292-293, 331) on the constructed case:
    the native pair at pair-list position 1 is skipped through lj_nat_pdb_dist2[1] of ANOTHER
    master entry, and the non-native pair through the rep_tol of the last native pair-list
    entry's types; only (4,8) and (2,10) act, with their analytic pair forces."""
    from conftest import quirk_system
    x, topo = quirk_system(str(tmp_path / "q"))
    o = Oracle()
    setup_system(o, str(tmp_path / "q"), std_cmds())
    F = _quirk_forces(o)
    ai, aj = o.lst(2)
    assert list(zip(ai, aj)) == [(3, 7), (4, 8)]   # pair-list order: (3,7) sits at position 1
    ri, rj = o.lst(3)
    assert list(zip(ri, rj)) == [(1, 9), (2, 10)]
    for b in (1, 3, 7, 9):                           # the skipped pairs' beads feel nothing
        assert np.all(F[b - 1] == 0.0), b
    # native (4,8), types 2 2: force_coeff_att = -12 * 0.7, r = 5, d = 5.5 along +x
    d = 5.5
    co = -12.0 * 0.7 / d ** 2 * ((5.0 / d) ** 12 - (5.0 / d) ** 6)
    assert F[3, 0] == pytest.approx(co * d, rel=1e-13) and F[7, 0] == pytest.approx(-co * d, rel=1e-13)
    # non-native (2,10), types 1 1: sigma 3.8, d = 5 along +x
    d, s = 5.0, 3.8
    co = -6.0 / d ** 2 * (2 * (s / d) ** 12 + (s / d) ** 6)
    assert F[1, 0] == pytest.approx(co * d, rel=1e-13) and F[9, 0] == pytest.approx(-co * d, rel=1e-13)


@needs_ref
def test_vdw_quirks_equal_reference(tmp_path):
    """The constructed quirk case: oracle forces and energies bit-identical to the reference's."""
    from conftest import quirk_system
    quirk_system(str(tmp_path / "q"))
    o, r = Oracle(), Ref()
    setup_system(o, str(tmp_path / "q"), std_cmds())
    setup_system(r, str(tmp_path / "q"), std_cmds())
    r.call("set_potential_forces")
    assert np.array_equal(_quirk_forces(o), _quirk_forces(r))
    for w in range(4):
        assert same_lists(o, r, w), w
    o.call("energy_eval")
    r.call("energy_eval")
    for k in ("e_bnd", "e_ang_ss", "e_vdw_rr", "rna_etot"):
        assert o.getd(k) == r.getd(k), k


@needs_ref
def test_ran2_gasdev_match_reference(tmp_path):
    o = Oracle()
    r = Ref()
    r.reset_rng(-101, tmp_path)
    for seed in (-101, -1, 0, 7, -2147483647):
        o.L.orc_set_seed(o.s, seed)
        r.L.ref_set_seed(seed)
        for _ in range(2000):
            assert o.L.orc_sys_ran2(o.s) == r.L.ref_ran2()
    r.reset_rng(-103, tmp_path)
    o = Oracle()
    o.L.orc_set_seed(o.s, -103)
    a = o.gasdev(5001)
    b = np.zeros(5001)
    r.L.ref_gasdev_n(b.ctypes.data_as(r.L.ref_gasdev_n.argtypes[0]), 5001)
    assert np.array_equal(a, b)


@needs_ref
def test_set_params_order_semantics():
    """`set zeta` resets h to the dynamics default but not halfh/a2 (params.cpp:92-98)."""
    o, r = Oracle(), Ref()
    for k, val in [("dynamics", "underdamped"), ("t_step", "0.01"), ("zeta", "3"), ("temp", "1.1"),
                   ("boxl", "300"), ("ncell", "25"), ("run", "4")]:
        o.set(k, val)
        r.set(k, val)
        for nm in ("a1", "a2", "a3", "a4", "a5", "h", "halfh", "zeta", "T", "boxl", "ncell", "lcell"):
            assert o.getd(nm) == r.getd(nm), (k, nm)
    assert o.getd("h") == 2.5e-3 and o.getd("halfh") == 0.005  # the reference's inconsistency, kept


def test_rnd_closed_form():
    L = Oracle.lib()
    cases = {0.5: 1.0, -0.5: -1.0, 1.49: 1.0, -1.5: -2.0, 0.0: -0.0, 2.5: 3.0, -2.5: -3.0, 1e-300: 0.0}
    for x, want in cases.items():
        assert L.orc_rnd(x) == want, x
    assert math.copysign(1.0, L.orc_rnd(0.0)) == -1.0  # ceil(-0.5) = -0
    assert L.orc_rnd(0.49999999999999994) == 1.0        # x + 0.5 rounds up to 1


def test_pow_cr_is_correctly_rounded():
    """orc_pow_cr (the device's power) against exact rational arithmetic."""
    from fractions import Fraction
    L = Oracle.lib()
    rng = np.random.default_rng(5)
    for d in rng.uniform(3.0, 12.0, 3000):
        d = float(d)
        for k in (6, 8):
            exact = Fraction(d) ** k
            got = L.orc_pow_cr(d, k)
            # correctly rounded: no double strictly closer to the exact value
            err = abs(Fraction(got) - exact)
            ulp = Fraction(math.ulp(got))
            assert err <= ulp / 2, (d, k)


def _two_bead_system(tmp, rij, d, native):
    """Beads 1 and 4 (|i-j| = 3) at distance d; beads 2, 3 far away with no bonds/angles."""
    os.makedirs(tmp, exist_ok=True)
    pos = np.array([[0, 0, 0], [100, 0, 0], [200, 0, 0], [d, 0, 0]], float)
    with open(os.path.join(tmp, "bonds.dat"), "w") as f:
        f.write("bonds 0\n")
    with open(os.path.join(tmp, "angles.dat"), "w") as f:
        f.write("angles 0\n")
    with open(os.path.join(tmp, "vdw.dat"), "w") as f:
        f.write(f"natt {1 if native else 0} nrep {0 if native else 1}\n1 4 {rij!r} 1 1\n")
    with open(os.path.join(tmp, "init.dat"), "w") as f:
        f.write("beads 4\n" + "".join(f"{i+1} {float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n"
                                      for i, p in enumerate(pos)))


@pytest.mark.parametrize("native", [True, False])
def test_pair_force_closed_form(tmp_path, native):
    """Two beads: the vdw force equals the analytic derivative form (energy.cpp:297, 335)."""
    d, rij = 5.3, (5.0 if native else 9.0)
    _two_bead_system(str(tmp_path), rij, d, native)
    o = Oracle()
    setup_system(o, str(tmp_path), [("temp", "0")])
    o.call("clear_forces")
    o.call("vdw_forces")
    f = o.coords(3).reshape(-1, 3)
    if native:
        # F_i = -12 / d^2 * ((r/d)^12 - (r/d)^6) * (r_j - r_i)
        co = -12.0 / d ** 2 * ((rij / d) ** 12 - (rij / d) ** 6)
    else:
        s = 3.8
        co = -6.0 / d ** 2 * (2 * (s / d) ** 12 + (s / d) ** 6)
    assert f[0, 0] == pytest.approx(co * d, rel=1e-13)
    assert f[3, 0] == pytest.approx(-co * d, rel=1e-13)
    assert np.all(f[1:3] == 0)


def test_fene_and_angular_closed_form(tmp_path):
    d, r0 = 4.1, 3.8
    with open(tmp_path / "bonds.dat", "w") as fh:
        fh.write(f"bonds 1\n1 2 {r0!r}\n")
    with open(tmp_path / "angles.dat", "w") as fh:
        fh.write("angles 1\n1 2 3 0.0\n")
    with open(tmp_path / "vdw.dat", "w") as fh:
        fh.write("natt 0 nrep 0\n")
    with open(tmp_path / "init.dat", "w") as fh:
        fh.write(f"beads 3\n1 0 0 0\n2 {d!r} 0 0\n3 0 {2 * d!r} 0\n")
    o = Oracle()
    setup_system(o, str(tmp_path), [("temp", "0")])
    o.call("clear_forces")
    o.call("fene_forces")
    f = o.coords(3).reshape(-1, 3)
    dev = d - r0
    temp = -20.0 * dev / d / (1 - dev * dev / 4.0)       # FENE magnitude along +x from bead 1
    assert f[0, 0] == pytest.approx(-temp * d, rel=1e-14)
    assert f[1, 0] == pytest.approx(temp * d, rel=1e-14)
    o.call("clear_forces")
    o.call("angular_forces")
    f = o.coords(3).reshape(-1, 3)
    dik = 2 * d
    co = 6.0 * 3.5 ** 6 / dik ** 8                       # soft sphere 6 eps sigma^6 / d^8 (i, i+2)
    assert f[0, 1] == pytest.approx(-co * dik, rel=1e-13)
    assert f[2, 1] == pytest.approx(co * dik, rel=1e-13)


def test_forces_sum_to_zero_without_noise(inputs_root):
    """Newton's third law: every deterministic term is pairwise +f/-f."""
    x, v, d = chain_inputs(inputs_root, 150, 9)
    o = Oracle()
    setup_system(o, d, std_cmds())
    o.call("update_neighbor_list")
    o.call("update_pair_list")
    o.call("clear_forces")
    for f in ("fene_forces", "angular_forces", "vdw_forces"):
        o.call(f)
    F = o.coords(3).reshape(-1, 3)
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).max()


def test_neighbor_list_brute_force(inputs_root):
    """NL / PL membership and order equal an O(N^2) numpy scan of the master lists."""
    n = 200
    x, v, d = chain_inputs(inputs_root, n, 17)
    o = Oracle()
    setup_system(o, d, std_cmds())
    X = x.reshape(-1, 3) * 1.07  # perturb: lists are not the native ones
    o.set_coords(1, X.ravel())
    o.call("update_neighbor_list")
    ai, aj = o.lst(4)
    ri, rj = o.lst(5)

    def d2(i, j):
        dd = X[j - 1] - X[i - 1]
        return dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1] + dd[:, 2] * dd[:, 2]
    topo = sopg.chain_topology(x)
    r = dict(zip(zip(topo["att_i"], topo["att_j"]), topo["att_r"]))
    ar = np.array([r[(i, j)] for i, j in zip(ai, aj)])
    keep_a = d2(ai, aj) < (3.2 * ar) ** 2
    keep_r = d2(ri, rj) < (3.2 * 3.8) ** 2
    ni, nj = o.lst(0)
    assert np.array_equal(ni, ai[keep_a]) and np.array_equal(nj, aj[keep_a])
    ni, nj = o.lst(1)
    assert np.array_equal(ni, ri[keep_r]) and np.array_equal(nj, rj[keep_r])
    # complement structure of the generated master lists
    assert len(ai) + len(ri) == (n - 2) * (n - 3) // 2


def test_free_particle_integrator(tmp_path):
    """F = 0 (T = 0, no terms): x1 = x0 + a1 v0, v1 = a3 a1 v0 (sop.cpp:292-326)."""
    with open(tmp_path / "bonds.dat", "w") as fh:
        fh.write("bonds 0\n")
    with open(tmp_path / "angles.dat", "w") as fh:
        fh.write("angles 0\n")
    with open(tmp_path / "vdw.dat", "w") as fh:
        fh.write("natt 0 nrep 0\n")
    with open(tmp_path / "init.dat", "w") as fh:
        fh.write("beads 1\n1 1.0 2.0 3.0\n")
    o = Oracle()
    setup_system(o, str(tmp_path), [("dynamics", "underdamped"), ("zeta", "50"), ("t_step", "0.05"),
                                    ("temp", "0.5")])
    o.set_noise(np.zeros(30))
    v0 = np.array([0.3, -0.2, 0.1])
    o.set_coords(2, v0)
    o.set_coords(3, np.zeros(3))
    a1, a2, a3, a4, _ = sopg.underdamped_coeffs(0.05, 50.0)
    o.call("underdamped_iteration")
    assert np.allclose(o.coords(1), np.array([1, 2, 3]) + a1 * v0, rtol=0, atol=1e-15)
    assert np.allclose(o.coords(2), a3 * a1 * v0, rtol=1e-15)


def test_observables_closed_form(tmp_path):
    """Q and Rg^2 on hand-built coordinates (sop.cpp:219-258)."""
    _two_bead_system(str(tmp_path), 5.0, 6.0, True)  # native r=5, d=6: 6/5 = 1.2 < 1.25
    o = Oracle()
    setup_system(o, str(tmp_path), [("temp", "0.6")])
    o.call("calculate_observables")
    assert o.geti("contct_nat") == 1 and o.getd("Q") == 1.0
    X = o.coords(1).reshape(-1, 3)
    want = sum(((X[i] - X[j]) ** 2).sum() for i in range(4) for j in range(i + 1, 4)) / 16.0
    assert o.getd("rgsq") == pytest.approx(want, rel=1e-15)
