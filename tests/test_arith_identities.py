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
