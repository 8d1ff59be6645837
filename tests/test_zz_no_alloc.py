"""sopg_run_noalloc (no device allocation; records staged in a fixed per-GPU
scratch array) against sopg_run: the same states and records, bit for bit, on
two calls in a row, the second on a smaller batch and fewer steps."""
import numpy as np
import pytest

import sopg
from conftest import chain_inputs

pytestmark = pytest.mark.gpu


def _ctx(chains, nup=50):
    ctx = sopg.Context()
    for (x, v, seed) in chains:
        ctx.add_chain(sopg.chain_topology(x), sopg.chain_params(cutoff=sopg.CUTOFF_NEIGHBOR, nnlup=20, nup=nup,
                                                                 noise_mode=sopg.NOISE_PHILOX, seed=seed))
    for c, (x, v, seed) in enumerate(chains):
        ctx.put_state(c, pos=x, unc=x, vel=v, force=np.zeros_like(x))
        ctx.update_neighbor_list(c)
        ctx.update_pair_list(c)
        ctx.force_eval(c)
    return ctx


def test_noalloc_run_equals_run(inputs_root):
    chains = []
    for n, seed in ((70, 171), (260, 172), (150, 173)):
        x, v, _ = chain_inputs(inputs_root, n, seed)
        chains.append((x, v, 90 + seed))
    print("[gpu] sopg_run_noalloc vs sopg_run", flush=True)
    a = _ctx(chains)
    b = _ctx(chains)
    ra = a.run(130)
    rb = b.run_noalloc(130)
    assert ra == rb and len(ra) == 3 * 2
    # second call: a chain has left, fewer steps
    a.remove_chain(1)
    b.remove_chain(1)
    ra = a.run(57)
    rb = b.run_noalloc(57)
    assert ra == rb and {r["chain"] for r in ra} == {0, 2}
    for c in (0, 2):
        sa, sb = a.get_state(c), b.get_state(c)
        for k in ("unc", "vel", "force", "pos"):
            assert np.array_equal(sa[k], sb[k]), (c, k)
        assert a.get_step(c) == b.get_step(c)
    a.close()
    b.close()


def test_noalloc_run_limits(inputs_root):
    x, v, _ = chain_inputs(inputs_root, 60, 174)
    fresh = sopg.Context()
    fresh.add_chain(sopg.chain_topology(x), sopg.chain_params(noise_mode=sopg.NOISE_PHILOX, seed=1))
    with pytest.raises(sopg.SopgError, match="first device operation"):
        fresh.run_noalloc(10)
    fresh.close()
    ctx = _ctx([(x, v, 5)], nup=1)
    with pytest.raises(sopg.SopgError, match="scratch array"):
        ctx.run_noalloc(200000)  # one record row per step: far beyond 2^19 doubles
    before = ctx.get_state(0)
    ctx.run_noalloc(30)          # still usable, and the refused call changed nothing first
    ref = _ctx([(x, v, 5)], nup=1)
    ref.run(30)
    after, want = ctx.get_state(0), ref.get_state(0)
    assert not np.array_equal(before["unc"], after["unc"])
    for k in ("unc", "vel", "force"):
        assert np.array_equal(after[k], want[k]), k
    rp = sopg.Context()
    rp.add_chain(sopg.chain_topology(x), sopg.chain_params(noise_mode=sopg.NOISE_REPLAY))
    rp.put_state(0, pos=x, unc=x, vel=v, force=np.zeros_like(x))
    rp.update_neighbor_list(0)
    with pytest.raises(sopg.SopgError, match="replayed-noise"):
        rp.run_noalloc(5)
    for o in (ctx, ref, rp):
        o.close()
