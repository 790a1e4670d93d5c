"""Identities the update kernel relies on to skip work bit for bit.

The FP32 update divides many numerators by one divisor (the body-frame
gradient's ga0·a / (2·a0), src/gradients.jl:347-362) as RN32(x · RN64(1/d)):
one binary64 reciprocal shared by all numerators. The kernels rely on that
being bit for bit the IEEE binary32 quotient x / d. This checks the identity on
the CPU over random binary32 operands spanning the whole exponent range,
subnormals, zeros, infinities and the values the update actually sees."""
import numpy as np


def shared(x, d):
    rd = 1.0 / d.astype(np.float64)
    with np.errstate(over="ignore", invalid="ignore"):
        return (x.astype(np.float64) * rd).astype(np.float32)


def ieee(x, d):
    with np.errstate(over="ignore", invalid="ignore", divide="ignore", under="ignore"):
        return x / d


def same_bits(a, b):
    nan = np.isnan(a) & np.isnan(b)
    return np.all(nan | (a.view(np.uint32) == b.view(np.uint32)))


def test_random_bit_patterns():
    rng = np.random.default_rng(20261204)
    n = 2_000_000
    x = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32).view(np.float32)
    d = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32).view(np.float32)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore", under="ignore"):
        assert same_bits(shared(x, d), ieee(x, d))


def test_operands_of_the_update():
    # numerators ga0·a with ga0 ~ 1e-4..1e1 and a ~ 1e-6..0.8; divisors 2·a0, a0 in (0.5, 1]
    rng = np.random.default_rng(7)
    n = 2_000_000
    x = (rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-10, 1, n)).astype(np.float32)
    d = (2.0 * rng.uniform(0.5, 1.0, n)).astype(np.float32)
    assert same_bits(shared(x, d), ieee(x, d))


def test_hard_cases():
    f32 = np.float32
    tiny, big = np.finfo(f32).tiny, np.finfo(f32).max
    sub = f32(1e-45)
    vals = np.array([0.0, -0.0, 1.0, -1.0, 3.0, tiny, -tiny, sub, -sub, big, -big, np.inf, -np.inf, np.nan,
                     f32(1) + np.finfo(f32).eps, f32(1) - np.finfo(f32).epsneg, 1.5, 0.1, 1 / 3], dtype=f32)
    x, d = np.meshgrid(vals, vals)
    x, d = x.ravel(), d.ravel()
    # quotients at the edges: near overflow, in the subnormal range, exact
    rng = np.random.default_rng(3)
    m = 500_000
    sig = rng.uniform(1.0, 2.0, m).astype(np.float32)
    xe = np.concatenate([x, sig * f32(2.0 ** 127), sig * f32(2.0 ** -126), sig * f32(2.0 ** -140)])
    de = np.concatenate([d, rng.uniform(0.5, 2.0, m).astype(np.float32),
                         rng.uniform(1.0, 2.0 ** 20, m).astype(np.float32),
                         rng.uniform(0.25, 4.0, m).astype(np.float32)])
    with np.errstate(divide="ignore", over="ignore", invalid="ignore", under="ignore"):
        assert same_bits(shared(xe, de), ieee(xe, de))


# ---- the FP64 position wrap (src/box.jl:21-30) returns x unchanged for 0 < x < L
# without evaluating floor(x / L); that relies on x / L rounding below 1 there.

def ref_wrap(x, L):
    s = np.floor(x / L)
    xp = x - s * L
    over = xp >= L
    return np.where(over, 0.0, xp), np.where(over, s + 1.0, s)


def test_wrap_inside_the_box_is_identity():
    rng = np.random.default_rng(11)
    Ls = np.concatenate([rng.uniform(1.0, 2000.0, 200), [975.168, 844.521, 62.531, 1.0, 2.0, 1024.0, 3.0]])
    for L in Ls:
        top = np.nextafter(L, 0.0)
        x = np.concatenate([rng.uniform(0.0, L, 2000), [top, np.nextafter(top, 0.0), 5e-324, 1e-300,
                                                        np.nextafter(L / 2, 0.0), L / 2]])
        x = x[(x > 0.0) & (x < L)]
        xp, s = ref_wrap(x, L)
        assert np.all(s == 0.0)
        assert np.array_equal(xp.view(np.uint64), x.view(np.uint64))


# ---- the skin check's minimum image (src/box.jl:14) is d itself for |d| < L/4
# (only d² is used, so the sign of a zero d does not matter)

def test_min_image_small_offsets():
    rng = np.random.default_rng(12)
    for L in np.concatenate([rng.uniform(1.0, 2000.0, 100), [975.168, 4.0]]):
        q = 0.25 * L
        d = np.concatenate([rng.uniform(-q, q, 5000), [np.nextafter(q, 0.0), -np.nextafter(q, 0.0), 0.0, 1e-300]])
        d = d[np.abs(d) < q]
        full = d - L * np.rint(d / L)
        assert np.array_equal(full * full, d * d)
