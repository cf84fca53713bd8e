"""The single-precision screen of the sphere-union double fold (cc_device.h screen_record and
screen_skips), restated in numpy, against the double skip test it must imply (cc_device.h
fold_sphere): wherever the screen skips a primitive, fold_sphere's own test skips it too, so the
fold is bitwise the oracle's.  The samples sit on both sides of the skip boundary, closer to it than
the screen's margin, over a range of coordinate magnitudes; a margin dropped from the record or the
test makes them fail.  The screen must also skip what lies clearly beyond the boundary, or it saves
nothing."""
import numpy as np

EPS = np.finfo(np.float64).eps
f32 = np.float32


def fp64_skip(acc, c, r, x, k):
    """fold_sphere's test, operation for operation."""
    d = x - c
    s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    t = (acc + k) + r
    tm = t + 16.0 * EPS * (((np.abs(acc) + k) + r) + np.abs(t))
    return (tm <= 0.0) | (s >= tm * tm)


def fma32(a, b, c):
    # a*b is exact in double for float operands; the sum then rounds once to double and once to
    # float, which can differ from a true fused multiply-add by one float ulp in rare ties -- far
    # inside the margins checked here
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(f32)


def screen_skips(acc, c, r, x, k, o, R, split=False):
    """split: R left out of the record and added as the separate margin m = fl32up(1e-6 R), as the
    plain fold of the whole grid does (records about the origin, R the grid's largest corner norm)."""
    cx = c - o
    kr = k + r
    Rw = 0.0 if split else R
    w = kr + 1e-6 * ((np.abs(kr) + Rw) + np.sqrt((cx[:, 0] * cx[:, 0] + cx[:, 1] * cx[:, 1]) + cx[:, 2] * cx[:, 2]))
    with np.errstate(invalid="ignore"):
        bad = ~(r >= 0.0) | ~(w < 1e30) | ~(w >= 1e-18)
    w = np.where(bad, np.inf, w)
    q = cx.astype(f32)
    qw = w.astype(f32)
    xf = (x - o).astype(f32)
    accf = acc.astype(f32)
    with np.errstate(over="ignore", invalid="ignore"):
        d = xf - q
        s = fma32(d[:, 2], d[:, 2], fma32(d[:, 1], d[:, 1], d[:, 0] * d[:, 0]))
        tp = fma32(np.abs(accf), np.full_like(accf, 2.0 ** -18), accf + qw)
        if split:
            m = np.nextafter((1e-6 * np.asarray(R) * (1.0 + 1e-9)).astype(f32), f32(np.inf))
            tp = tp + m
        return (np.abs(tp) < f32(1e18)) & ((tp <= f32(0)) | ((tp >= f32(2.0 ** -60)) & (s >= tp * tp)))


def samples(rng, n, scale):
    """Points x within R of an origin o (a brick's first node, coordinates up to `scale`), spheres
    near them, and an acc placing acc + k + r at |x - c| times (1 + e), e log-uniform in magnitude
    from 1e-12 to 1e-2 and of either sign (e < 0: the skip test holds)."""
    h = rng.uniform(1e-3, 1e-2, n) * max(scale, 1.0) ** 0.5
    R = 16.0 * np.sqrt(3.0) * h
    o = rng.uniform(-scale, scale, (n, 3))
    u = rng.normal(size=(n, 3))
    x = o + u / np.linalg.norm(u, axis=1)[:, None] * (rng.uniform(0, 1, n) * R / np.sqrt(3.0))[:, None]
    v = rng.normal(size=(n, 3))
    c = x + v / np.linalg.norm(v, axis=1)[:, None] * rng.uniform(0.0, 0.5, n)[:, None]
    r = rng.uniform(0.0, 0.3, n)
    k = rng.uniform(1e-3, 0.1, n)
    D = np.linalg.norm(x - c, axis=1)
    e = np.sign(rng.uniform(-1, 1, n)) * 10.0 ** rng.uniform(-12, -2, n)
    acc = D * (1.0 + e) - k - r
    return acc, c, r, x, k, o, R


def test_the_screen_implies_the_double_skip_test():
    rng = np.random.default_rng(7)
    for scale, split in ((0.5, False), (2.0, False), (1e3, False), (2.0, True), (1e3, True)):
        acc, c, r, x, k, o, R = samples(rng, 400_000, scale)
        if split:  # records about the origin: R bounds |x|
            R = np.linalg.norm(x, axis=1) * (1.0 + 1e-12)
            o = np.zeros_like(o)
        sk = screen_skips(acc, c, r, x, k, o, R, split)
        ok = fp64_skip(acc, c, r, x, k)
        assert not np.any(sk & ~ok), (scale, int(np.count_nonzero(sk & ~ok)))
        # nothing inside the screen's margin of the boundary is skipped, and it is not vacuous:
        # beyond a gap of 1e-4 of the magnitudes (and 1e-5 of the record's extent) it skips
        # everything the double test skips
        D = np.linalg.norm(x - c, axis=1)
        gap = (D - (acc + k + r)) / (np.abs(acc) + k + r + D)
        assert not np.any(sk & (gap < 1e-7))
        far = (D - (acc + k + r)) > 1e-4 * (np.abs(acc) + k + r + D) + 1e-5 * (R + np.linalg.norm(c - o, axis=1))
        assert np.all(sk[far]), (scale, int(np.count_nonzero(~sk[far])))


def test_deep_inside_and_degenerate_records():
    rng = np.random.default_rng(11)
    n = 100_000
    acc, c, r, x, k, o, R = samples(rng, n, 1.0)
    # acc + k + r below zero by a relative margin: both tests skip (the centre is k above acc)
    acc2 = -(k + r) * (1.0 + 10.0 ** rng.uniform(-4, 0, n))
    assert np.all(screen_skips(acc2, c, r, x, k, o, R) <= fp64_skip(acc2, c, r, x, k))
    assert np.mean(screen_skips(acc2, c, r, x, k, o, R)) > 0.99
    # acc + k + r just below or above zero, within the screen's margin: whatever the screen skips,
    # the double test skips (it holds wherever |x - c| >= acc + k + r up to its margin)
    for rel in (-1e-9, -1e-14, 0.0, 1e-14, 1e-9):
        acc3 = -(k + r) * (1.0 + rel)
        assert np.all(screen_skips(acc3, c, r, x, k, o, R) <= fp64_skip(acc3, c, r, x, k)), rel
    # the centre itself (|x - c| = 0) with acc + k + r barely positive: neither skips
    acc4 = -(k + r) + 1e-12
    assert not np.any(screen_skips(acc4, x.copy(), r, x, k, o, R))
    # negative radii, infinite or NaN accumulators: no screen skip
    assert not np.any(screen_skips(acc, c, -r - 1e-3, x, k, o, R))
    assert not np.any(screen_skips(np.full(n, np.inf), c, r, x, k, o, R))
    assert not np.any(screen_skips(np.full(n, -np.inf), c, r, x, k, o, R))
    assert not np.any(screen_skips(np.full(n, np.nan), c, r, x, k, o, R))


def test_a_margin_dropped_from_the_screen_is_caught():
    """The check above has teeth: with the record's 1e-6 term and the 2^-18 |acc| term removed, the
    boundary samples make the screen skip where the double test does not."""
    rng = np.random.default_rng(7)
    acc, c, r, x, k, o, R = samples(rng, 400_000, 2.0)
    cx = c - o
    w = (k + r).astype(f32)
    xf = (x - o).astype(f32)
    d = xf - cx.astype(f32)
    s = fma32(d[:, 2], d[:, 2], fma32(d[:, 1], d[:, 1], d[:, 0] * d[:, 0]))
    tp = acc.astype(f32) + w
    slip = (tp <= 0) | (s >= tp * tp)
    assert np.any(slip & ~fp64_skip(acc, c, r, x, k))
