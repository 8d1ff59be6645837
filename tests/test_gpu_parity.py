"""GPU parity of the gfx950 hot path against the oracle, through the C ABI.

Bars (from the project's parity contract):
  * neighbor / cell / pair lists: identical, entry for entry, in the reference's order;
  * per-bead forces: bit-identical to the oracle running the correctly rounded
    power (the device's one arithmetic difference from glibc pow), and within
    1e-12 relative of the oracle with glibc pow (= the reference itself);
  * energies within 1e-10 relative; observables Q exact, Rg^2 / kinT 1e-12;
  * trajectories (replayed reference noise) within 1e-9 A after 1000 steps.
"""
import os

import numpy as np
import pytest

import sopg
from conftest import chain_inputs
from refs import Oracle, setup_system, std_cmds

pytestmark = pytest.mark.gpu

FORCE_REL = 1e-12
ENERGY_REL = 1e-10
TRAJ_ABS = 1e-9


def per_bead_rel(a, b):
    a = a.reshape(-1, 3)
    b = b.reshape(-1, 3)
    num = np.linalg.norm(a - b, axis=1)
    den = np.maximum(np.linalg.norm(b, axis=1), 1e-300)
    return float((num / den).max())


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.fixture
def pow_cr():
    Oracle.set_pow_mode(1)
    yield
    Oracle.set_pow_mode(0)


def make_pair(root, n, seed, cmds, params, coords_vel=True, lanes=0, topo=None):
    x, v, d = chain_inputs(root, n, seed)
    orc = Oracle()
    setup_system(orc, d, cmds)
    ctx = sopg.Context(lanes=lanes, list_lanes=lanes)
    ctx.add_chain(topo or sopg.chain_topology(x), params)
    if coords_vel:
        orc.set_coords(2, v)
        ctx.put_state(0, pos=x, unc=x, vel=v, force=np.zeros_like(x))
    else:
        ctx.put_state(0, pos=x, unc=x, vel=np.zeros_like(x), force=np.zeros_like(x))
    return x, v, d, orc, ctx


def check_lists(ctx, orc, which_dev, which_orc, chain=0):
    di, dj = ctx.get_list(which_dev, chain)
    oi, oj = orc.lst(which_orc)
    assert len(di) == len(oi), (which_dev, len(di), len(oi))
    assert np.array_equal(di, oi) and np.array_equal(dj, oj), which_dev


def run_oracle_ctrl(orc, nsteps, nnlup, nup, cutoff="nl"):
    """Hot loop of underdamped_ctrl (sop.cpp:153-189) on the oracle; returns records."""
    recs = []
    inlup, iup = 1, 1
    for istep in range(1, nsteps + 1):
        if inlup % nnlup == 0:
            orc.call("update_neighbor_list" if cutoff == "nl" else "update_cell_list")
            inlup = 0
        inlup += 1
        orc.call("update_pair_list")
        orc.call("underdamped_iteration")
        if iup % nup == 0:
            orc.call("energy_eval")
            orc.call("calculate_observables")
            recs.append(dict(istep=istep, e_bnd=orc.getd("e_bnd"), e_ang_ss=orc.getd("e_ang_ss"),
                             e_vdw_rr=orc.getd("e_vdw_rr"), rna_etot=orc.getd("rna_etot"),
                             Q=orc.getd("Q"), contct_nat=orc.geti("contct_nat"),
                             rgsq=orc.getd("rgsq"), kinT=orc.getd("kinT")))
            iup = 0
        iup += 1
    return recs


def compare_records(drecs, orecs):
    assert len(drecs) == len(orecs)
    for a, b in zip(drecs, orecs):
        assert a["istep"] == b["istep"]
        assert a["contct_nat"] == b["contct_nat"]
        for k in ("e_bnd", "e_ang_ss", "e_vdw_rr", "rna_etot"):
            assert rel(a[k], b[k]) <= ENERGY_REL, (k, a[k], b[k])
        for k in ("rgsq", "kinT", "Q"):
            assert rel(a[k], b[k]) <= 1e-12, (k, a[k], b[k])


@pytest.mark.parametrize("lanes", [1, 4, 8])
def test_lists_forces_energy_nl76(inputs_root, pow_cr, lanes):
    n = 76
    params = sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=50, nup=100, noise_mode=sopg.NOISE_REPLAY)
    x, v, d, orc, ctx = make_pair(inputs_root, n, 7, std_cmds(nup=100), params, lanes=lanes)
    print(f"[gpu] N={n} lanes={lanes}: lists", flush=True)
    na, nr = ctx.update_neighbor_list()
    orc.call("update_neighbor_list")
    assert (na, nr) == (orc.geti("nnl_att"), orc.geti("nnl_rep"))
    check_lists(ctx, orc, sopg.LIST_NL_ATT, 0)
    check_lists(ctx, orc, sopg.LIST_NL_REP, 1)
    pa, pr = ctx.update_pair_list()
    orc.call("update_pair_list")
    assert (pa, pr) == (orc.geti("nil_att"), orc.geti("nil_rep"))
    check_lists(ctx, orc, sopg.LIST_PL_ATT, 2)
    check_lists(ctx, orc, sopg.LIST_PL_REP, 3)
    e = ctx.energy_eval()
    orc.call("energy_eval")
    for k in ("e_bnd", "e_ang_ss", "e_vdw_rr", "rna_etot"):
        assert rel(e[k], orc.getd(k)) <= ENERGY_REL, k
    noise = Oracle().gasdev(3 * n)
    orc.set_noise(noise)
    orc.call("force_eval")
    ctx.force_eval(0, noise)
    fd = ctx.get_state(0)["force"]
    fo = orc.coords(3)
    assert np.array_equal(fd, fo), per_bead_rel(fd, fo)
    ob = ctx.observables()
    orc.call("calculate_observables")
    assert ob["contct_nat"] == orc.geti("contct_nat")
    assert rel(ob["rgsq"], orc.getd("rgsq")) <= 1e-12
    assert rel(ob["kinT"], orc.getd("kinT")) <= 1e-12


def test_force_vs_glibc_pow_reference(inputs_root):
    """Against the oracle with glibc pow (bit-identical to the reference): 1e-12 relative."""
    n = 300
    params = sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, noise_mode=sopg.NOISE_REPLAY)
    x, v, d, orc, ctx = make_pair(inputs_root, n, 11, std_cmds(), params)
    ctx.update_neighbor_list(); ctx.update_pair_list()
    orc.call("update_neighbor_list"); orc.call("update_pair_list")
    noise = Oracle().gasdev(3 * n)
    orc.set_noise(noise)
    orc.call("force_eval")
    ctx.force_eval(0, noise)
    assert per_bead_rel(ctx.get_state(0)["force"], orc.coords(3)) <= FORCE_REL


def test_trajectory_1000_steps_nl76(inputs_root, pow_cr):
    n, steps, nnlup, nup = 76, 1000, 50, 100
    params = sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=nnlup, nup=nup, noise_mode=sopg.NOISE_REPLAY)
    x, v, d, orc, ctx = make_pair(inputs_root, n, 7, std_cmds(nnlup=nnlup, nup=nup), params)
    print(f"[gpu] trajectory N={n} steps={steps}", flush=True)
    noise = Oracle().gasdev(3 * n * (steps + 1))
    # underdamped_ctrl preamble (sop.cpp:119-147): lists, energy, force
    orc.call("update_neighbor_list"); orc.call("update_pair_list")
    ctx.update_neighbor_list(); ctx.update_pair_list()
    orc.set_noise(noise)
    orc.call("energy_eval"); orc.call("force_eval")
    ctx.force_eval(0, noise[:3 * n])
    orecs = run_oracle_ctrl(orc, steps, nnlup, nup)
    drecs = ctx.run(steps, noise[3 * n:])
    s = ctx.get_state(0)
    for w, k in ((1, "unc"), (0, "pos"), (2, "vel")):
        diff = np.abs(s[k] - orc.coords(w)).max()
        assert diff <= TRAJ_ABS, (k, diff)
    assert np.array_equal(s["unc"], orc.coords(1)), "expected bit-identical trajectory"
    assert per_bead_rel(s["force"], orc.coords(3)) <= FORCE_REL
    compare_records(drecs, orecs)
    check_lists(ctx, orc, sopg.LIST_PL_REP, 3)


def test_cell_list_n300(inputs_root, pow_cr):
    n, steps, nnlup, nup = 300, 200, 20, 50
    extra = [("boxl", "504"), ("ncell", "42")]
    cmds = std_cmds(cutoff="celllist", nnlup=nnlup, nup=nup, extra=extra)
    params = sopg.chain_params(boxl=504.0, ncell=42.0, cutoff=sopg.CUTOFF_CELL, nnlup=nnlup, nup=nup,
                               noise_mode=sopg.NOISE_REPLAY)
    x, v, d, orc, ctx = make_pair(inputs_root, n, 5, cmds, params)
    print(f"[gpu] cell list N={n}", flush=True)
    assert orc.getd("lcell") == params.lcell
    na, nr = ctx.update_cell_list()
    orc.call("update_cell_list")
    assert (na, nr) == (orc.geti("nnl_att"), orc.geti("nnl_rep"))
    check_lists(ctx, orc, sopg.LIST_NL_ATT, 0)
    check_lists(ctx, orc, sopg.LIST_NL_REP, 1)
    ctx.update_pair_list(); orc.call("update_pair_list")
    check_lists(ctx, orc, sopg.LIST_PL_REP, 3)
    noise = Oracle().gasdev(3 * n * (steps + 1))
    orc.set_noise(noise)
    orc.call("energy_eval"); orc.call("force_eval")
    ctx.force_eval(0, noise[:3 * n])
    orecs = run_oracle_ctrl(orc, steps, nnlup, nup, cutoff="cl")
    drecs = ctx.run(steps, noise[3 * n:])
    s = ctx.get_state(0)
    assert np.abs(s["unc"] - orc.coords(1)).max() <= TRAJ_ABS
    compare_records(drecs, orecs)
    check_lists(ctx, orc, sopg.LIST_NL_REP, 1)


def _write_vdw(path, entries, natt, nrep):
    with open(path, "w") as f:
        f.write(f"natt {natt} nrep {nrep}\n")
        for (i, j, r) in entries:
            f.write(f"{i} {j} {r!r} 1 1\n")


def test_explicit_master_lists(inputs_root, pow_cr, tmp_path):
    """A vdw file in arbitrary order with some reversed pairs: explicit lists."""
    n, steps, nnlup, nup = 90, 150, 10, 50
    x, v, d = chain_inputs(inputs_root, n, 3)
    X = x.reshape(-1, 3)
    rng = np.random.default_rng(0)
    ent = []
    for i in range(n):
        for j in range(i + 3, n):
            dx, dy, dz = X[j] - X[i]
            r = float(np.sqrt(dx * dx + dy * dy + dz * dz))
            ent.append((i + 1, j + 1, r) if rng.random() < 0.8 else (j + 1, i + 1, r))
    rng.shuffle(ent)
    natt = sum(1 for e in ent if e[2] < 8.0)
    dd = tmp_path / "explicit"
    dd.mkdir()
    for f in ("bonds.dat", "angles.dat", "init.dat"):
        (dd / f).write_bytes(open(os.path.join(d, f), "rb").read())
    _write_vdw(dd / "vdw.dat", ent, natt, len(ent) - natt)
    orc = Oracle()
    setup_system(orc, str(dd), std_cmds(nnlup=nnlup, nup=nup))
    orc.set_coords(2, v)
    ai, aj = orc.lst(4)
    ri, rj = orc.lst(5)
    att = [e for e in ent if e[2] < 8.0]
    topo = sopg.chain_topology(x)
    topo.update(att_i=ai, att_j=aj, att_r=np.array([e[2] for e in att]), rep_i=ri, rep_j=rj)
    params = sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=nnlup, nup=nup, noise_mode=sopg.NOISE_REPLAY)
    ctx = sopg.Context()
    ctx.add_chain(topo, params, implicit=False)
    ctx.put_state(0, pos=x, unc=x, vel=v, force=np.zeros_like(x))
    print("[gpu] explicit master lists", flush=True)
    ctx.update_neighbor_list(); orc.call("update_neighbor_list")
    for w in (0, 1):
        check_lists(ctx, orc, w, w)
    ctx.update_pair_list(); orc.call("update_pair_list")
    for w in (2, 3):
        check_lists(ctx, orc, w, w)
    noise = Oracle().gasdev(3 * n * (steps + 1))
    orc.set_noise(noise)
    orc.call("energy_eval"); orc.call("force_eval")
    ctx.force_eval(0, noise[:3 * n])
    assert np.array_equal(ctx.get_state(0)["force"], orc.coords(3))
    orecs = run_oracle_ctrl(orc, steps, nnlup, nup)
    drecs = ctx.run(steps, noise[3 * n:])
    assert np.abs(ctx.get_state(0)["unc"] - orc.coords(1)).max() <= TRAJ_ABS
    compare_records(drecs, orecs)


def _philox_ctx(chains, lanes=0):
    ctx = sopg.Context(lanes=lanes, list_lanes=lanes)
    for (x, v, seed) in chains:
        ctx.add_chain(sopg.chain_topology(x), sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, nup=100,
                                                                 noise_mode=sopg.NOISE_PHILOX, seed=seed))
    for c, (x, v, seed) in enumerate(chains):
        ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
        ctx.update_neighbor_list(c)
        ctx.update_pair_list(c)
        ctx.force_eval(c)
    return ctx


def test_batch_equals_alone(inputs_root):
    """Each chain of a ragged batch evolves exactly as it does alone."""
    chains = []
    for n, seed in ((60, 21), (131, 22), (205, 23)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        chains.append((x, v, 1000 + seed))
    print("[gpu] batch vs alone", flush=True)
    mix = _philox_ctx(chains)
    rm = mix.run(300)
    for c, ch in enumerate(chains):
        alone = _philox_ctx([ch])
        ra = alone.run(300)
        a, b = mix.get_state(c), alone.get_state(0)
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(a[k], b[k]), (c, k)
        assert [r for r in rm if r["chain"] == c] == [dict(r, chain=c) for r in ra]


@pytest.mark.parametrize("lanes", [1, 2, 4, 16, 64])
def test_lanes_invariance(inputs_root, lanes):
    x, v, _ = chain_inputs(inputs_root, 150, 31)
    base = _philox_ctx([(x, v, 5)], lanes=8)
    base.run(120)
    other = _philox_ctx([(x, v, 5)], lanes=lanes)
    other.run(120)
    a, b = base.get_state(0), other.get_state(0)
    for k in ("unc", "vel", "force"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("cutoff", [sopg.CUTOFF_NEIGHBOR, sopg.CUTOFF_CELL])
def test_binned_rebuild_equals_full_scan(inputs_root, cutoff):
    """The cell-grid list rebuild and the full candidate scan give identical lists
    and therefore bit-identical trajectories (batch of sizes 80..1500)."""
    chains = []
    for n, seed in ((80, 41), (400, 42), (1500, 43)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        chains.append((x, v, 77 + seed))
    # self-avoiding walks wider than half the box (278 and 450 A in a 504 A box): the
    # grid applies up to the box length less the grid edge
    for n, seed in ((3000, 8), (4000, 9)):
        chains.append((sopg.gen_saw(n, seed), sopg.gen_maxwell(n, 0.6, seed), 77 + seed))
    res = []
    for binned in (True, False):
        ctx = sopg.Context(binned=binned)
        for (x, v, seed) in chains:
            ctx.add_chain(sopg.chain_topology(x),
                          sopg.chain_params(boxl=504.0, ncell=42.0, cutoff=cutoff, nnlup=10, nup=50,
                                            noise_mode=sopg.NOISE_PHILOX, seed=seed))
        for c, (x, v, seed) in enumerate(chains):
            ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
            (ctx.update_cell_list if cutoff == sopg.CUTOFF_CELL else ctx.update_neighbor_list)(c)
            ctx.update_pair_list(c)
            ctx.force_eval(c)
        print(f"[gpu] binned={binned} cutoff={cutoff}", flush=True)
        ctx.run(200)
        res.append([(ctx.get_state(c), ctx.get_list(sopg.LIST_NL_REP, c), ctx.get_list(sopg.LIST_PL_REP, c))
                    for c in range(len(chains))])
        ctx.close()
    for (sa, na, pa), (sb, nb, pb) in zip(*res):
        for k in ("unc", "vel", "force"):
            assert np.array_equal(sa[k], sb[k]), k
        assert np.array_equal(na[0], nb[0]) and np.array_equal(na[1], nb[1])
        assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])


def zigzag_chain(n, direction=(1.0, 0.05, 0.0), step=3.6, wiggle=1.0, origin=(0.0, 0.0, 0.0)):
    """n beads along a tilted axis, step A apart along it and +-wiggle across: in a box
    shorter than the chain, consecutive wraps along x lie 6 A apart (direction's y / x x the
    box), so wrapped segments meet within the list cutoffs only through the min image."""
    u = np.asarray(direction, float)
    u /= np.linalg.norm(u)
    w = np.cross(u, (0.0, 0.0, 1.0))
    w /= np.linalg.norm(w)
    k = np.arange(n)[:, None]
    x = np.asarray(origin, float) + k * step * u + wiggle * np.where(k % 2 == 0, 1.0, -1.0) * w
    return x.reshape(-1).copy()


def _lists(ctx, c):
    return [ctx.get_list(w, c) for w in (sopg.LIST_NL_ATT, sopg.LIST_NL_REP, sopg.LIST_PL_ATT, sopg.LIST_PL_REP)]


@pytest.mark.parametrize("crit,skin", [("nl", 4.0), ("nl", 0.0), ("cl", 0.0)])
def test_periodic_grid_lists(tmp_path, crit, skin):
    """Chains wider than the period less a grid edge are binned on a grid over the period
    (wrapped coordinates, neighbouring cells taken modulo the cell count): the same lists as the
    full scan, pairs that only the min image brings within the cutoffs (or the periodic cell
    adjacency) included, and the oracle's lists for one of them.  Box 120 A with 12 A cells: a
    900 A zig-zag, a 1500-bead self-avoiding walk (whose periodic images overlap it), and a
    native chain 10^4 A from the origin (local grid)."""
    box = 120.0
    cutoff = sopg.CUTOFF_NEIGHBOR if crit == "nl" else sopg.CUTOFF_CELL
    zz = zigzag_chain(250)
    saw = sopg.gen_saw(1500, 21)
    nat = sopg.gen_native(400, 22) + np.tile([1.0e4, -3.0e3, 50.0], 400)
    chains = [zz, saw, nat]
    res = []
    for binned in (True, False):
        print(f"[gpu] periodic grid lists {crit} skin={skin} binned={binned}", flush=True)
        ctx = sopg.Context(binned=binned)
        ctx.set_skin(skin)
        for c, x in enumerate(chains):
            ctx.add_chain(sopg.chain_topology(x),
                          sopg.chain_params(boxl=box, ncell=10.0, cutoff=cutoff, nnlup=10, nup=50,
                                            noise_mode=sopg.NOISE_PHILOX, seed=5 + c))
        plans = []
        for c, x in enumerate(chains):
            ctx.put_state(c, pos=x, unc=x, vel=np.zeros_like(x), force=np.zeros_like(x))
            (ctx.update_neighbor_list if crit == "nl" else ctx.update_cell_list)(c)
            plans.append(ctx.rebuild_plan(c))
            ctx.update_pair_list(c)
        # cells of at least the grid edge (3.2 sigma_rep, plus the skin for candidates; one
        # reference cell for the cell list)
        n = int(box // (3.2 * 3.8 * 1.001 + skin)) if crit == "nl" else 10
        want = [(sopg.PLAN_PERIODIC, n), (sopg.PLAN_PERIODIC, n), (sopg.PLAN_LOCAL, None)] if binned else \
            [(sopg.PLAN_SCAN, 0)] * 3
        for (p, k), (wp, wk) in zip(plans, want):
            assert p == wp and (wk is None or k == wk), (plans, want)
        res.append([_lists(ctx, c) for c in range(len(chains))])
        ctx.close()
    for la, lb in zip(*res):
        for (ai, aj), (bi, bj) in zip(la, lb):
            assert np.array_equal(ai, bi) and np.array_equal(aj, bj)
    # pairs that only the min image brings together: plain separation beyond half the box
    zi, zj = res[0][0][1]
    assert len(zi) > 0
    X = zz.reshape(-1, 3)
    plain = np.abs(X[zj - 1] - X[zi - 1]).max(axis=1)
    assert (plain > box / 2).sum() > 100
    # the oracle (the reference's algorithm) on the zig-zag alone
    d = str(tmp_path / "zz")
    sopg.write_inputs(zz, d)
    orc = Oracle()
    setup_system(orc, d, std_cmds(cutoff="neighborlist" if crit == "nl" else "celllist", nnlup=10, nup=50,
                                  extra=[("boxl", "120"), ("ncell", "10")]))
    orc.call("update_neighbor_list" if crit == "nl" else "update_cell_list")
    orc.call("update_pair_list")
    for k in range(4):
        oi, oj = orc.lst(k)
        di, dj = res[0][0][k]
        assert np.array_equal(di, oi) and np.array_equal(dj, oj), k


def test_wide_local_grid_lists():
    """A chain that needs more than 32 grid edges per axis but fits the box (a 900 A zig-zag in
    a 2000 A box) is binned on a local grid of wider cubic cells: the full scan's lists."""
    zz = zigzag_chain(250)
    res = []
    for binned in (True, False):
        print(f"[gpu] wide local grid binned={binned}", flush=True)
        ctx = sopg.Context(binned=binned)
        ctx.add_chain(sopg.chain_topology(zz),
                      sopg.chain_params(boxl=2000.0, ncell=100.0, cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=10, nup=50,
                                        noise_mode=sopg.NOISE_PHILOX, seed=3))
        ctx.put_state(0, pos=zz, unc=zz, vel=np.zeros_like(zz), force=np.zeros_like(zz))
        ctx.update_neighbor_list(0)
        plan = ctx.rebuild_plan(0)
        # 895 A along x over at most 32 cells (the 16.2 A edge alone would need 56)
        assert (plan[0] == sopg.PLAN_LOCAL and 30 <= plan[1] <= 32) if binned else plan == (sopg.PLAN_SCAN, 0), plan
        ctx.update_pair_list(0)
        res.append(_lists(ctx, 0))
        ctx.close()
    for (ai, aj), (bi, bj) in zip(*res):
        assert np.array_equal(ai, bi) and np.array_equal(aj, bj)
    assert len(res[0][1][0]) > 0


def test_periodic_grid_run_equals_full_scan():
    """A run on the periodic grid (a zig-zag wrapping the 120 A box about seven times, beside a
    chain on the local grid): lists and trajectories bit-identical to the full scan's."""
    box = 120.0
    chains = [zigzag_chain(250), sopg.gen_native(400, 23) + np.tile([-2.0e3, 40.0, 7.0e3], 400)]
    res = []
    for binned in (True, False):
        print(f"[gpu] periodic grid run binned={binned}", flush=True)
        ctx = sopg.Context(binned=binned)
        for c, x in enumerate(chains):
            ctx.add_chain(sopg.chain_topology(x),
                          sopg.chain_params(boxl=box, ncell=10.0, cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=10, nup=50,
                                            noise_mode=sopg.NOISE_PHILOX, seed=31 + c))
        for c, x in enumerate(chains):
            ctx.put_state(c, pos=x, unc=x, vel=sopg.gen_maxwell(len(x) // 3, 0.6, 40 + c), force=np.zeros_like(x))
            ctx.update_neighbor_list(c)
            ctx.update_pair_list(c)
            ctx.force_eval(c)
        ctx.run(60)
        res.append([(ctx.get_state(c), _lists(ctx, c)) for c in range(len(chains))])
        ctx.close()
    for (sa, la), (sb, lb) in zip(*res):
        for k in ("unc", "vel", "force"):
            assert np.array_equal(sa[k], sb[k]), k
        for (ai, aj), (bi, bj) in zip(la, lb):
            assert np.array_equal(ai, bi) and np.array_equal(aj, bj)


def test_philox_noise_statistics(inputs_root):
    """Device noise: zero force field replaced by pure noise statistics check."""
    n = 2000
    x = sopg.gen_saw(n, 3)
    topo = sopg.chain_topology(x)
    ctx = sopg.Context()
    p = sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, noise_mode=sopg.NOISE_PHILOX, seed=99)
    ctx.add_chain(topo, p)
    ctx.put_state(0, pos=x, unc=x, vel=np.zeros_like(x), force=np.zeros_like(x))
    ctx.update_neighbor_list(); ctx.update_pair_list()
    # energy-free part: the random force alone = force_eval minus the deterministic terms
    ctx.force_eval(0)
    f1 = ctx.get_state(0)["force"]
    ctx.force_eval(0)
    f2 = ctx.get_state(0)["force"]
    var = np.sqrt(2.0 * 0.6 * 50.0 / 0.05)
    g = (f2 - f1) / var  # difference of two independent deviates: N(0, 2)
    assert abs(g.mean()) < 0.05
    assert abs(g.var() - 2.0) < 0.1


def _setup_chain(ctx, c, x, v):
    ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
    ctx.update_neighbor_list(c)
    ctx.update_pair_list(c)
    ctx.force_eval(c)


def test_continuous_batching_join_leave(inputs_root):
    """Chains join and leave a running context (continuous batching): a chain
    that stays, and one that joins mid-run into the memory a finished chain
    freed, each evolve bit-identically to the same chain alone."""
    spec = {"A": (90, 51), "B": (240, 52), "C": (130, 53), "D": (170, 54)}
    data = {}
    for k, (n, seed) in spec.items():
        x, v, _ = chain_inputs(inputs_root, n, seed)
        data[k] = (x, v, 500 + seed)
    print("[gpu] continuous batching", flush=True)
    mix = _philox_ctx([data["A"], data["B"], data["C"]])
    r1 = mix.run(100)
    mix.remove_chain(1)  # B leaves
    assert mix.nchains() == 2
    x, v, seed = data["D"]
    d = mix.add_chain(sopg.chain_topology(x), sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, nup=100,
                                                                noise_mode=sopg.NOISE_PHILOX, seed=seed))
    assert d == 1  # the freed id is reused
    _setup_chain(mix, d, x, v)
    r2 = mix.run(150)
    alone_a = _philox_ctx([data["A"]])
    ra = alone_a.run(100) + alone_a.run(150)
    alone_d = _philox_ctx([data["D"]])
    rd = alone_d.run(150)
    for got, ref in ((mix.get_state(0), alone_a.get_state(0)), (mix.get_state(d), alone_d.get_state(0))):
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(got[k], ref[k]), k
    assert [r for r in r1 + r2 if r["chain"] == 0] == ra
    assert [r for r in r2 if r["chain"] == d] == [dict(r, chain=d) for r in rd]


@pytest.mark.parametrize("lanes", [1, 4, 64])
def test_freed_beads_skipped(inputs_root, lanes):
    """Chains that leave free beads behind: a hole of whole step-kernel blocks
    (which the step launch then skips) and a freed last chain (the high-water
    mark falls), with a resident chain beside them.  The chains that stay
    evolve bit-identically to the same chains alone."""
    data = []
    for n, seed in ((90, 71), (1500, 72), (150, 73), (700, 74)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        data.append((x, v, 700 + seed))
    print(f"[gpu] freed beads L={lanes}", flush=True)
    mix = _philox_ctx(data, lanes=lanes)
    mix.set_resident(100)  # chain 0 runs in the resident kernel
    r1 = mix.run(40)
    mix.remove_chain(1)
    mix.remove_chain(3)
    if lanes == 4:
        mix.autotune()  # times every lane count on the block tables of the holed batch
    r2 = mix.run(70)
    for c, d in ((0, data[0]), (2, data[2])):
        alone = _philox_ctx([d], lanes=lanes)
        alone.set_resident(100)
        ra = alone.run(40) + alone.run(70)
        got, ref = mix.get_state(c), alone.get_state(0)
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(got[k], ref[k]), (c, k)
        assert [r for r in r1 + r2 if r["chain"] == c] == [dict(r, chain=c) for r in ra]


def test_rebuild_span_leaves_chains_not_due(inputs_root):
    """A step's list rebuild covers the bead range of the chains due at that step.  A
    chain inside that range that is not due keeps its lists until its own rebuild step,
    also when its slot id lies outside the due chains' slots (slot and bead orders differ
    after a chain left and a larger one took its slot at the end of the bead range)."""
    data = []
    for n, seed in ((100, 81), (400, 82), (900, 83), (600, 84)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        data.append((x, v, 800 + seed))
    print("[gpu] rebuild span", flush=True)
    mix = _philox_ctx(data[:3])
    mix.remove_chain(0)
    x, v, seed = data[3]
    d = mix.add_chain(sopg.chain_topology(x), sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, nup=100,
                                                                noise_mode=sopg.NOISE_PHILOX, seed=seed))
    assert d == 0  # slot 0, beads after chain 2's
    _setup_chain(mix, d, x, v)
    # chains 0 and 1 rebuild at run steps 19 and 39, chain 2 at steps 9 and 29
    phase = {0: 1, 1: 1, 2: 11}
    for c, ph in phase.items():
        mix.set_step(c, 1.0, ph, 1)
    mix.run(25)
    for c, k in ((0, 3), (1, 1), (2, 2)):
        alone = _philox_ctx([data[k]])
        alone.set_step(0, 1.0, phase[c], 1)
        alone.run(25)
        got, ref = mix.get_state(c), alone.get_state(0)
        for key in ("unc", "vel", "force", "pos"):
            assert np.array_equal(got[key], ref[key]), (c, key)
        for k, ((mi, mj), (ai, aj)) in enumerate(zip(_lists(mix, c), _lists(alone, 0))):
            assert np.array_equal(mi, ai) and np.array_equal(mj, aj), (c, k)


def test_autotune_keeps_results(inputs_root):
    """The autotuner changes the launch shape only: results stay bit-identical."""
    chains = []
    for n, seed in ((70, 61), (900, 62)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        chains.append((x, v, 9 + seed))
    ref = _philox_ctx(chains, lanes=8)
    ref.run(60)
    tuned = _philox_ctx(chains, lanes=8)
    lanes = tuned.autotune()
    assert lanes in (1, 2, 4, 8, 16, 32, 64) and tuned.lanes()[0] == lanes
    tuned.run(60)
    for c in range(len(chains)):
        a, b = ref.get_state(c), tuned.get_state(c)
        for k in ("unc", "vel", "force"):
            assert np.array_equal(a[k], b[k]), k


def test_checkpoint_restore_lossless(inputs_root):
    """A chain checkpointed mid-run (between list rebuilds) and restored into
    another context -- at another offset, beside another chain -- continues
    bit-identically to the uninterrupted run: state, records, noise stream."""
    a_in = chain_inputs(inputs_root, 110, 81)[:2]
    b_in = chain_inputs(inputs_root, 190, 82)[:2]
    other = chain_inputs(inputs_root, 70, 83)[:2]
    chains = [(a_in[0], a_in[1], 11), (b_in[0], b_in[1], 12)]
    print("[gpu] checkpoint / restore", flush=True)
    ctx = _philox_ctx(chains)
    ctx.run(130)  # 130 is not a multiple of the rebuild period 20
    blob = ctx.checkpoint(1)
    tail = ctx.run(170)
    fresh = _philox_ctx([(other[0], other[1], 13)])
    cid = fresh.add_chain(sopg.chain_topology(b_in[0]),
                          sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, nup=100,
                                            noise_mode=sopg.NOISE_PHILOX, seed=12))
    fresh.restore(cid, blob)
    resumed = fresh.run(170)
    assert fresh.get_step(cid) == ctx.get_step(1)
    got, ref = fresh.get_state(cid), ctx.get_state(1)
    for k in ("unc", "vel", "force", "pos"):
        assert np.array_equal(got[k], ref[k]), k
    assert [r for r in resumed if r["chain"] == cid] == [dict(r, chain=cid) for r in tail if r["chain"] == 1]
    with pytest.raises(sopg.SopgError):
        fresh.restore(0, blob)  # a chain of another topology


@pytest.mark.parametrize("cutoff", [sopg.CUTOFF_NEIGHBOR, sopg.CUTOFF_CELL])
def test_graph_replay_equals_direct(inputs_root, cutoff):
    """Step segments replayed from captured HIP graphs give the same bits as
    direct launches, run after run, with eager single-chain calls between."""
    chains = []
    for n, seed in ((70, 101), (160, 102), (240, 103)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        chains.append((x, v, 40 + seed))
    results = []
    for graphs in (False, True):
        ctx = sopg.Context()
        for (x, v, seed) in chains:
            ctx.add_chain(sopg.chain_topology(x),
                          sopg.chain_params(boxl=504.0, ncell=42.0, cutoff=cutoff, nnlup=20, nup=100,
                                            noise_mode=sopg.NOISE_PHILOX, seed=seed))
        ctx.set_graphs(graphs)
        for c, (x, v, seed) in enumerate(chains):
            ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
            (ctx.update_cell_list if cutoff == sopg.CUTOFF_CELL else ctx.update_neighbor_list)(c)
            ctx.update_pair_list(c)
            ctx.force_eval(c)
        print(f"[gpu] graphs={graphs} cutoff={cutoff}", flush=True)
        per_run = []
        recs = []
        for nsteps in (250, 250, 170):
            recs += ctx.run(nsteps)
            ctx.energy_eval(1)  # eager work between replays
            ctx.observables(2)
            per_run.append([ctx.get_state(c) for c in range(len(chains))])
        if graphs:
            caps, launches = ctx.graph_stats()
            assert caps >= 1 and launches > caps  # segments were replayed, not only captured
        else:
            assert ctx.graph_stats() == (0, 0)
        results.append((per_run, recs))
        ctx.close()
    (da, ra), (ga, rg) = results
    for run_a, run_g in zip(da, ga):
        for sa, sg in zip(run_a, run_g):
            for k in ("unc", "vel", "force", "pos"):
                assert np.array_equal(sa[k], sg[k]), k
    assert ra == rg


@pytest.mark.parametrize("cutoff", [sopg.CUTOFF_NEIGHBOR, sopg.CUTOFF_CELL])
def test_split_runs_equal_one_run(inputs_root, cutoff):
    """A run leaves the pair list at the positions it stops at in the other
    parity, and the next run starts from it.  Runs split at every phase of the
    rebuild period (ends on and next to rebuild steps, single steps) give the
    bits of one unbroken run -- also when calls between runs change the state
    (a pair-list update, a state upload), after which the run filters again."""
    chains = []
    for n, seed in ((70, 111), (160, 112), (240, 113)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        chains.append((x, v, 60 + seed))

    def make():
        ctx = sopg.Context()
        for (x, v, seed) in chains:
            ctx.add_chain(sopg.chain_topology(x),
                          sopg.chain_params(boxl=504.0, ncell=42.0, cutoff=cutoff, nnlup=20, nup=50,
                                            noise_mode=sopg.NOISE_PHILOX, seed=seed))
        for c, (x, v, seed) in enumerate(chains):
            ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
            (ctx.update_cell_list if cutoff == sopg.CUTOFF_CELL else ctx.update_neighbor_list)(c)
            ctx.update_pair_list(c)
            ctx.force_eval(c)
        return ctx

    def snapshot(ctx):
        out = []
        for c in range(len(chains)):
            st = ctx.get_state(c)
            lists = [ctx.get_list(w, c) for w in (sopg.LIST_NL_ATT, sopg.LIST_NL_REP, sopg.LIST_PL_ATT,
                                                  sopg.LIST_PL_REP)]
            out.append((st, lists))
        return out

    print(f"[gpu] split runs cutoff={cutoff}", flush=True)
    one = make()
    rec_one = one.run(120)
    ref = snapshot(one)
    one.close()
    splits = (7, 13, 20, 1, 19, 40, 20)
    for disturb in (False, True):
        ctx = make()
        recs = []
        for i, k in enumerate(splits):
            recs += ctx.run(k)
            if disturb and i == 1:
                ctx.update_pair_list(0)
            if disturb and i == 3:
                st = ctx.get_state(2)
                ctx.put_state(2, pos=st["pos"], unc=st["unc"], vel=st["vel"], force=st["force"])
            if i == 4:
                ctx.energy_eval(1)  # eager work that changes nothing
        got = snapshot(ctx)
        ctx.close()
        assert recs == rec_one
        for (sa, la), (sb, lb) in zip(ref, got):
            for k in ("unc", "vel", "force", "pos"):
                assert np.array_equal(sa[k], sb[k]), (disturb, k)
            for (ai, aj), (bi, bj) in zip(la, lb):
                assert np.array_equal(ai, bi) and np.array_equal(aj, bj), disturb


@pytest.mark.parametrize("cutoff,lanes", [(sopg.CUTOFF_NEIGHBOR, 0), (sopg.CUTOFF_CELL, 0),
                                          (sopg.CUTOFF_NEIGHBOR, 1), (sopg.CUTOFF_NEIGHBOR, 4)])
def test_resident_kernel_equals_per_step(inputs_root, cutoff, lanes):
    """Small chains advanced by the persistent resident kernel (all steps of a
    run in one launch, next to the per-step kernels of the large chain) end
    with the same bits, and give the same records, as the per-step path --
    over several runs, with their own rebuild and record periods."""
    chains = []
    for n, seed, nnlup, nup in ((50, 111, 20, 100), (76, 112, 15, 50), (120, 113, 20, 100), (300, 114, 20, 75)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        chains.append((x, v, 60 + seed, nnlup, nup))
    results = []
    for resident in (0, 128):
        ctx = sopg.Context()
        for (x, v, seed, nnlup, nup) in chains:
            ctx.add_chain(sopg.chain_topology(x),
                          sopg.chain_params(boxl=504.0, ncell=42.0, cutoff=cutoff, nnlup=nnlup, nup=nup,
                                            noise_mode=sopg.NOISE_PHILOX, seed=seed))
        ctx.set_resident(resident, lanes)
        for c, (x, v, *_rest) in enumerate(chains):
            ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
            (ctx.update_cell_list if cutoff == sopg.CUTOFF_CELL else ctx.update_neighbor_list)(c)
            ctx.update_pair_list(c)
            ctx.force_eval(c)
        print(f"[gpu] resident={resident} lanes={lanes} cutoff={cutoff}", flush=True)
        per_run, recs = [], []
        for nsteps in (250, 251, 170):
            recs += ctx.run(nsteps)
            ctx.energy_eval(1)  # eager work between runs
            per_run.append([ctx.get_state(c) for c in range(len(chains))])
        results.append((per_run, recs))
        ctx.close()
    (da, ra), (ga, rg) = results
    for run_a, run_g in zip(da, ga):
        for c, (sa, sg) in enumerate(zip(run_a, run_g)):
            for k in ("unc", "vel", "force", "pos"):
                assert np.array_equal(sa[k], sg[k]), (c, k)
    assert {r["chain"] for r in ra} == {0, 1, 2, 3}
    assert sorted(ra, key=lambda r: (r["istep"], r["chain"])) == sorted(rg, key=lambda r: (r["istep"], r["chain"]))


def test_resident_only_run_all_chains_resident(inputs_root):
    """A run whose chains are all resident launches no per-step kernels; the
    result equals the per-step path's, and a following per-step run (resident
    off) continues from it exactly."""
    chains = []
    for n, seed in ((60, 121), (76, 122), (90, 123)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        chains.append((x, v, 80 + seed))
    results = []
    for resident in (0, 100):
        ctx = _philox_ctx(chains)
        ctx.set_resident(resident)
        print(f"[gpu] all-resident={resident}", flush=True)
        recs = ctx.run(333)
        ctx.set_resident(0)
        recs += ctx.run(120)
        results.append(([ctx.get_state(c) for c in range(len(chains))], recs))
        ctx.close()
    (sa, ra), (sg, rg) = results
    for a, g in zip(sa, sg):
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(a[k], g[k]), k
    assert ra == rg


def test_long_run_statistics_match_reference(inputs_root):
    """The device generator (Philox) is not the reference's stream, so its
    trajectories differ from the reference's; long-run averages of Q and Rg^2
    must agree within their statistical error bars.  16 independent replicas
    on each side: the device with 16 Philox keys, the oracle (bit-identical to
    the reference) with its own generator at 16 `run` seeds."""
    n, kT, nnlup, nup, equil, steps, reps = 76, 0.9, 50, 100, 2000, 10000, 16
    x, v, d = chain_inputs(inputs_root, n, 5)
    topo = sopg.chain_topology(x)
    ctx = sopg.Context()
    for r in range(reps):
        ctx.add_chain(topo, sopg.chain_params(T=kT, cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=nnlup, nup=nup,
                                              noise_mode=sopg.NOISE_PHILOX, seed=5000 + r))
    for r in range(reps):
        _setup_chain(ctx, r, x, v)
    print("[gpu] long-run statistics", flush=True)
    recs = ctx.run(equil + steps)
    dev = {k: np.array([np.mean([rec[k] for rec in recs if rec["chain"] == r and rec["istep"] > equil])
                        for r in range(reps)]) for k in ("Q", "rgsq")}
    ref = {"Q": [], "rgsq": []}
    for r in range(reps):
        orc = Oracle()
        setup_system(orc, d, std_cmds(temp=kT, nnlup=nnlup, nup=nup, extra=[("run", str(r + 1))]))
        orc.set_coords(2, v)
        orc.call("update_neighbor_list"); orc.call("update_pair_list")
        orc.call("energy_eval"); orc.call("force_eval")
        orecs = run_oracle_ctrl(orc, equil + steps, nnlup, nup)
        for k in ("Q", "rgsq"):
            ref[k].append(np.mean([rec[k] for rec in orecs if rec["istep"] > equil]))
    for k in ("Q", "rgsq"):
        a, b = dev[k], np.array(ref[k])
        se = np.sqrt(a.var(ddof=1) / reps + b.var(ddof=1) / reps)
        z = abs(a.mean() - b.mean()) / max(se, 1e-12)
        print(f"{k}: device {a.mean():.5f} +- {a.std(ddof=1) / np.sqrt(reps):.5f}, "
              f"reference {b.mean():.5f} +- {b.std(ddof=1) / np.sqrt(reps):.5f}, z = {z:.2f}")
        assert z < 4.0, (k, a.mean(), b.mean(), se)


def test_list_capacity_overflow_is_an_error(inputs_root):
    """A per-bead non-native list longer than the context's capacity is a
    clean error (the kernels never write past the capacity), not a fault."""
    x, v, _ = chain_inputs(inputs_root, 300, 111)
    ctx = sopg.Context(capacity=8)
    ctx.add_chain(sopg.chain_topology(x), sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, nup=100))
    ctx.put_state(0, pos=x, unc=x, vel=v, force=np.zeros_like(x))
    with pytest.raises(sopg.SopgError, match="capacity"):
        ctx.update_neighbor_list(0)
    ctx.close()


def test_empty_runs(inputs_root):
    """Zero steps and a context without chains are no-ops."""
    ctx = sopg.Context()
    assert ctx.run(10) == []
    x, v, _ = chain_inputs(inputs_root, 50, 112)
    ctx.add_chain(sopg.chain_topology(x), sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, nup=100))
    _setup_chain(ctx, 0, x, v)
    before = ctx.get_state(0)
    assert ctx.run(0) == []
    after = ctx.get_state(0)
    for k in ("unc", "vel", "force"):
        assert np.array_equal(before[k], after[k])
    ctx.remove_chain(0)
    assert ctx.nchains() == 0 and ctx.run(5) == []
    ctx.close()


def test_resident_kernel_overdamped(inputs_root):
    """Overdamped chains (sim_type 2, first rebuild at step 1) in the resident
    kernel give the per-step path's bits and records."""
    chains = []
    for n, seed in ((64, 141), (100, 142)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        chains.append((x, v, 20 + seed))
    results = []
    for resident in (0, 128):
        ctx = sopg.Context()
        for (x, v, seed) in chains:
            ctx.add_chain(sopg.chain_topology(x),
                          sopg.chain_params(T=0.9, cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=10, nup=40,
                                            noise_mode=sopg.NOISE_PHILOX, seed=seed, sim_type=2))
        ctx.set_resident(resident)
        for c, (x, v, seed) in enumerate(chains):
            ctx.put_state(c, pos=x, unc=x, vel=np.zeros_like(x), force=np.zeros_like(x))
            ctx.set_step(c, 1.0, 0, 1)  # overdamped_ctrl: inlup = 0, the list rebuilt on the first step
            ctx.force_eval(c)
        print(f"[gpu] overdamped resident={resident}", flush=True)
        recs = ctx.run(157) + ctx.run(43)
        results.append(([ctx.get_state(c) for c in range(len(chains))], recs))
        ctx.close()
    (sa, ra), (sb, rb) = results
    for a, b in zip(sa, sb):
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(a[k], b[k]), k
    assert ra == rb and len(ra) == 2 * (200 // 40)


@pytest.mark.parametrize("offset", [0.0, 2.0e4])
def test_candidate_lists_keep_lists_and_results(inputs_root, offset):
    """Neighbor lists filtered from candidate lists (reused while no bead has
    moved more than half the skin since their build) are the lists of a full
    rebuild: states, records and lists agree bit for bit for skin 0 (off), a
    skin small enough to rebuild the candidates nearly every time, and the
    default -- with positions moved from the host between runs, a cell-list
    chain beside, and runs starting between rebuilds.  The chains far from the
    origin (offset) widen the single-precision error band of the candidate
    build and filter, so that more pairs are decided in double precision."""
    specs = [(90, 151, sopg.CUTOFF_NEIGHBOR), (260, 152, sopg.CUTOFF_NEIGHBOR), (700, 153, sopg.CUTOFF_NEIGHBOR),
             (150, 154, sopg.CUTOFF_CELL)]
    data = [chain_inputs(inputs_root, n, seed)[:2] for (n, seed, _) in specs]
    data = [(x + offset, v) for (x, v) in data]
    results = []
    for skin in (0.0, 0.2, 3.0):
        ctx = sopg.Context()
        for (n, seed, cut), (x, v) in zip(specs, data):
            ctx.add_chain(sopg.chain_topology(x),
                          sopg.chain_params(T=1.2, boxl=504.0, ncell=42.0, cutoff=cut, nnlup=20, nup=50,
                                            noise_mode=sopg.NOISE_PHILOX, seed=10 + seed))
        ctx.set_skin(skin)
        for c, ((n, seed, cut), (x, v)) in enumerate(zip(specs, data)):
            ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
            (ctx.update_cell_list if cut == sopg.CUTOFF_CELL else ctx.update_neighbor_list)(c)
            ctx.update_pair_list(c)
            ctx.force_eval(c)
        print(f"[gpu] candidate lists skin={skin} offset={offset}", flush=True)
        recs = ctx.run(130)
        # move chain 1 from the host (to its native positions again) mid-run
        st = ctx.get_state(1)
        ctx.put_state(1, pos=data[1][0], unc=data[1][0], vel=st["vel"], force=st["force"])
        recs += ctx.run(207)
        recs += ctx.run(95)
        lists = [[ctx.get_list(w, c) for w in (sopg.LIST_NL_REP, sopg.LIST_PL_REP, sopg.LIST_NL_ATT)]
                 for c in range(len(specs))]
        results.append(([ctx.get_state(c) for c in range(len(specs))], recs, lists))
        ctx.close()
    (s0, r0, l0) = results[0]
    for (s1, r1, l1) in results[1:]:
        for a, b in zip(s0, s1):
            for k in ("unc", "vel", "force", "pos"):
                assert np.array_equal(a[k], b[k]), k
        assert r0 == r1
        for la, lb in zip(l0, l1):
            for (ai, aj), (bi, bj) in zip(la, lb):
                assert np.array_equal(ai, bi) and np.array_equal(aj, bj)


def test_bead_pool_limit_refused_and_context_usable():
    """Reserving beyond the bead pool's 2^28-bead limit fails with a clear error
    before anything is allocated; the context keeps working."""
    x = sopg.gen_native(60, 61)
    v = sopg.gen_maxwell(60, 0.6, 61)
    ctx = sopg.Context()
    c = ctx.add_chain(sopg.chain_topology(x), sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR,
                                                                 noise_mode=sopg.NOISE_PHILOX, seed=61))
    _setup_chain(ctx, c, x, v)
    with pytest.raises(sopg.SopgError, match="bead pool limit"):
        ctx.reserve((1 << 28) + 1, 4)
    ctx.run(5)
    st = ctx.get_state(c)
    assert np.all(np.isfinite(st["unc"]))
    ctx.close()


def run_oracle_od_ctrl(orc, nsteps, nnlup, nup, inlup=0, iup=1, istep0=1):
    """Hot loop of overdamped_ctrl (sop.cpp:428-466) on the oracle: the global inlup starts at 0,
    so the list is rebuilt on the first step; records use the step's increments (kinT)."""
    recs = []
    for istep in range(istep0, istep0 + nsteps):
        if inlup % nnlup == 0:
            orc.call("update_neighbor_list")
            inlup = 0
        inlup += 1
        orc.call("update_pair_list")
        orc.call("overdamped_iteration")
        if iup % nup == 0:
            orc.call("energy_eval")
            orc.call("calculate_observables")
            recs.append(dict(istep=istep, e_bnd=orc.getd("e_bnd"), e_ang_ss=orc.getd("e_ang_ss"),
                             e_vdw_rr=orc.getd("e_vdw_rr"), rna_etot=orc.getd("rna_etot"),
                             Q=orc.getd("Q"), contct_nat=orc.geti("contct_nat"),
                             rgsq=orc.getd("rgsq"), kinT=orc.getd("kinT")))
            iup = 0
        iup += 1
    return recs


def od_cmds(nnlup, nup, temp=0.6):
    return [("dynamics", "overdamped"), ("zeta", "50"), ("temp", str(temp)), ("cutofftype", "neighborlist"),
            ("nnlup", str(nnlup)), ("nup", str(nup))]


@pytest.mark.parametrize("lanes", [1, 4])
def test_overdamped_trajectory_vs_oracle(inputs_root, pow_cr, lanes):
    """Overdamped BD (sop.cpp:331-471) with the reference's noise replayed: 1000 steps in runs
    of 157 and 843 steps, so that records (every 40 steps, kinT from the step's increments,
    sop.cpp:271-278) fall in the middle of a run, where the next step's increment is already
    computed.  Forces bit-identical, positions within 1e-9 A, records as the oracle's."""
    n, nnlup, nup = 120, 20, 40
    params = sopg.chain_params(h=0.02, coeffs=(0.0, 0.0, 0.0, 0.0, 0.02 / 50.0), cutoff=sopg.CUTOFF_NEIGHBOR,
                               nnlup=nnlup, nup=nup, noise_mode=sopg.NOISE_REPLAY, sim_type=2)
    x, v, d, orc, ctx = make_pair(inputs_root, n, 19, od_cmds(nnlup, nup), params, coords_vel=False, lanes=lanes)
    assert orc.getd("a5") == params.a5 and orc.getd("h") == params.h
    print(f"[gpu] overdamped N={n} lanes={lanes}", flush=True)
    noise = Oracle().gasdev(3 * n * 1002)
    # overdamped_ctrl preamble (sop.cpp:396-424): lists, energy, first force
    orc.call("update_neighbor_list"); orc.call("update_pair_list")
    ctx.update_neighbor_list(); ctx.update_pair_list()
    orc.set_noise(noise)
    orc.call("energy_eval"); orc.call("force_eval")
    ctx.force_eval(0, noise[:3 * n])
    assert np.array_equal(ctx.get_state(0)["force"], orc.coords(3))
    ctx.set_step(0, 1.0, 0, 1)  # the global inlup = 0: rebuild on the first step
    orecs = run_oracle_od_ctrl(orc, 1000, nnlup, nup)
    q = 3 * n
    drecs = ctx.run(157, noise[q:q + 157 * 3 * n])
    q += 157 * 3 * n
    drecs += ctx.run(843, noise[q:q + 843 * 3 * n])
    s = ctx.get_state(0)
    for w, k in ((1, "unc"), (0, "pos")):
        assert np.abs(s[k] - orc.coords(w)).max() <= TRAJ_ABS, k
    assert np.array_equal(s["unc"], orc.coords(1)), "expected bit-identical trajectory"
    assert np.array_equal(s["force"], orc.coords(3))
    compare_records(drecs, orecs)
    assert len(drecs) == 25 and all(r["kinT"] > 0.0 for r in drecs)
    # the standalone entry point (sop.cpp:331-363) continues the same trajectory
    q = 3 * n * 1001
    orc.call("update_pair_list")
    orc.call("overdamped_iteration")
    ctx.update_pair_list()
    ctx.overdamped_iteration(0, noise[q:q + 3 * n])
    s = ctx.get_state(0)
    assert np.array_equal(s["unc"], orc.coords(1)) and np.array_equal(s["force"], orc.coords(3))
    orc.call("calculate_observables")
    ob = ctx.observables()
    assert rel(ob["kinT"], orc.getd("kinT")) <= 1e-12


@pytest.mark.parametrize("lanes", [1, 4, 16])
def test_vdw_quirks_vs_oracle(tmp_path, lanes):
    """The reference's two skip tests as coded (energy.cpp:292-293, 331) on a constructed case in
    which both decide: a native pair-list entry skipped through another master entry's
    lj_nat_pdb_dist2, and a non-native pair skipped through the last native pair-list entry's
    rep_tol.  At T = 0 the skipped beads stay at rest, so the quirks fire at every step of a
    50-step run too: forces, energies and positions equal the oracle's bit for bit."""
    from conftest import quirk_system
    d = str(tmp_path / "q")
    x, topo = quirk_system(d)
    nnlup = 10
    orc = Oracle()
    setup_system(orc, d, std_cmds(temp=0.0, nnlup=nnlup, nup=10))
    params = sopg.chain_params(T=0.0, cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=nnlup, nup=10,
                               noise_mode=sopg.NOISE_REPLAY)
    ctx = sopg.Context(lanes=lanes, list_lanes=lanes)
    ctx.add_chain(topo, params, implicit=False)
    z = np.zeros_like(x)
    ctx.put_state(0, pos=x, unc=x, vel=z, force=z)
    orc.set_coords(2, z)
    print(f"[gpu] vdw quirks lanes={lanes}", flush=True)
    assert ctx.update_neighbor_list() == (2, 2)  # (1,5) and (2,6) are too far for the list
    ctx.update_pair_list()
    orc.call("update_neighbor_list"); orc.call("update_pair_list")
    for w in range(4):
        check_lists(ctx, orc, w, w)
    noise = Oracle().gasdev(3 * 10 * 52)
    orc.set_noise(noise)
    orc.call("energy_eval"); orc.call("force_eval")
    ctx.force_eval(0, noise[:30])
    f = ctx.get_state(0)["force"]
    assert np.array_equal(f, orc.coords(3))
    F = f.reshape(-1, 3)
    assert all(np.all(F[b - 1] == 0.0) for b in (1, 3, 7, 9))  # skipped: no force at all
    e = ctx.energy_eval()
    for k in ("e_vdw_rr", "rna_etot"):
        assert rel(e[k], orc.getd(k)) <= ENERGY_REL, k
    orecs = run_oracle_ctrl(orc, 50, nnlup, 10)
    drecs = ctx.run(50, noise[30:30 + 50 * 30])
    s = ctx.get_state(0)
    assert np.array_equal(s["unc"], orc.coords(1)) and np.array_equal(s["force"], orc.coords(3))
    assert np.array_equal(s["unc"].reshape(-1, 3)[[0, 2, 6, 8]], x.reshape(-1, 3)[[0, 2, 6, 8]])
    compare_records(drecs, orecs)


def test_skin_change_between_runs(inputs_root):
    """Runs with captured graphs at skin 0, then at skin 3 on the same context (the second needs
    the candidate filter the first's graphs never launched), equal a context that ran at the
    default skin throughout, bit for bit."""
    chains = []
    for n, seed in ((80, 161), (220, 162), (400, 163)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        chains.append((x, v, 30 + seed))
    print("[gpu] skin change between runs", flush=True)
    a = _philox_ctx(chains)
    a.set_skin(0.0)
    ra = a.run(140)
    a.set_skin(3.0)
    ra += a.run(200)
    b = _philox_ctx(chains)
    rb = b.run(140) + b.run(200)
    assert a.graph_stats()[1] > 0
    for c in range(len(chains)):
        sa, sb = a.get_state(c), b.get_state(c)
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(sa[k], sb[k]), (c, k)
        for w in (sopg.LIST_NL_REP, sopg.LIST_PL_REP):
            (ai, aj), (bi, bj) = a.get_list(w, c), b.get_list(w, c)
            assert np.array_equal(ai, bi) and np.array_equal(aj, bj)
    assert ra == rb


def test_candidate_overflow_is_an_error_every_time(inputs_root):
    """A candidate list longer than the capacity (its radius is wider than the neighbor list's)
    is reported by the run that builds it and by every later run: a truncated candidate list
    never serves a rebuild."""
    x, v, _ = chain_inputs(inputs_root, 300, 111)
    ctx = sopg.Context(capacity=96)
    ctx.set_skin(0.0)
    ctx.add_chain(sopg.chain_topology(x), sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, nup=100,
                                                             noise_mode=sopg.NOISE_PHILOX, seed=5))
    _setup_chain(ctx, 0, x, v)
    print("[gpu] candidate overflow", flush=True)
    ctx.run(40)  # neighbor lists alone fit
    ctx.set_skin(4.0)
    for _ in range(3):
        with pytest.raises(sopg.SopgError, match="capacity"):
            ctx.run(40)
    ctx.close()
