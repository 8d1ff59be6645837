"""Floating-point identities the step kernel relies on for bit parity, checked
on the host in IEEE double (numpy): the kernel's rewritten expressions equal
the reference's on every input, signed zeros included."""
import numpy as np


def _bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def _rnd(x):
    # global.cpp:223-227: round half away from zero through floor / ceil
    return np.where(x > 0, np.floor(x + 0.5), np.ceil(x - 0.5))


def _mimg(d, boxl):
    # min image as the kernels apply it: d + 0.0 when |d| <= boxl/4, else d - boxl*rnd(d/boxl)
    return np.where(np.abs(d) <= 0.25 * boxl, d + 0.0, d - boxl * _rnd(d / boxl))


def _samples(rng, n):
    c = rng.uniform(-600.0, 600.0, n)
    o = c + rng.normal(0.0, 30.0, n)
    o[: n // 8] = c[: n // 8]                              # equal coordinates (c - o = 0)
    o[n // 8: n // 4] = c[n // 8: n // 4] + rng.uniform(-300, 300, n // 8)  # beyond boxl/4
    c[n // 4: n // 4 + 16] = 0.0                           # signed zeros
    o[n // 4: n // 4 + 8] = -0.0
    o[n // 4 + 8: n // 4 + 16] = 0.0
    return c, o


def test_side_as_sign_factor_equals_reverse_subtraction():
    """mimg((o - c) * -1.0) == mimg(c - o) bit for bit (the side-1 displacement)."""
    rng = np.random.default_rng(11)
    c, o = _samples(rng, 1 << 18)
    for boxl in (500.0, 504.0, 123.25):
        ref = _mimg(c - o, boxl)
        got = _mimg((o - c) * -1.0, boxl)
        assert np.array_equal(_bits(ref), _bits(got))
        # the side-0 form is unchanged
        assert np.array_equal(_bits(_mimg(o - c, boxl)), _bits(_mimg((o - c) * 1.0, boxl)))


def test_negative_zero_is_additive_identity():
    """x + (-0.0) == x bit for bit for every x (the fold adds -0.0 for lanes
    without a term), while x + (+0.0) turns -0.0 into +0.0."""
    rng = np.random.default_rng(12)
    special = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 5e-324, -5e-324, 2.2250738585072014e-308,
                        1.7976931348623157e308, -1.7976931348623157e308])
    x = np.concatenate([special, rng.normal(0, 1e3, 1 << 16), rng.standard_cauchy(1 << 12)])
    assert np.array_equal(_bits(x + (-0.0)), _bits(x))
    assert _bits(np.float64(-0.0) + 0.0) != _bits(np.float64(-0.0))


def test_ordered_sum_with_neutral_terms_unchanged():
    """Adding a sequence in order with -0.0 in the gaps gives the same bits as
    adding only the contributing terms (the fold at 2 and 4 lanes)."""
    rng = np.random.default_rng(13)
    for _ in range(2000):
        n = rng.integers(1, 40)
        f = rng.normal(0, 1, n) * 10.0 ** rng.integers(-8, 8, n)
        keep = rng.random(n) < 0.3
        start = rng.normal(0, 5)
        a = np.float64(start)
        for v, k in zip(f, keep):
            if k:
                a = a + v
        b = np.float64(start)
        for v, k in zip(f, keep):
            b = b + (v if k else -0.0)
        assert _bits(a) == _bits(b)


def test_single_precision_candidate_distance_decides_as_double():
    """The candidate filter decides from single-precision positions when the
    squared distance lies farther than eps2 from every cutoff, and in double
    precision otherwise: every fast decision equals the double one."""
    rng = np.random.default_rng(14)
    n = 1 << 20
    R = 22.4 + 2 * 4.0 + 1.0                       # largest 3.2 sigma + 2 skin + 1
    c = rng.uniform(-3000.0, 3000.0, (n, 3))
    c[: n // 4] = rng.uniform(-60.0, 60.0, (n // 4, 3))
    cuts = np.array([3.2 * 3.8, 2.5 * 3.8, 3.2 * 7.0, 2.5 * 5.4])
    r = rng.choice(cuts, n) + rng.normal(0.0, 1e-4, n)   # distances at the cutoffs
    r[: n // 8] = rng.uniform(0.5, R - 1.0, n // 8)
    u = rng.normal(size=(n, 3))
    p = c + u / np.linalg.norm(u, axis=1)[:, None] * r[:, None]
    d = p - c
    d2 = (d * d).sum(1)                              # the double computation
    m = np.abs(c).max(1)
    ec = 2.0 ** -24 * (2.0 * m + R)
    eps2 = 2.0 * R * np.sqrt(3.0) * ec + 3.0 * ec * ec + 1e-9
    s = p.astype(np.float32).astype(np.float64) - c.astype(np.float32).astype(np.float64)
    s2 = (s * s).sum(1)
    for cut in cuts:
        cn = cut * cut
        fast = np.abs(s2 - cn) > eps2
        assert fast.mean() > 0.5
        assert np.array_equal((s2 < cn)[fast], (d2 < cn)[fast])
