"""Elementary functions of the path (log, sin, cos, atan2, acos) against the C library.

The device and the oracle share one implementation built from correctly rounded IEEE
operations (so they agree bit for bit, which the GPU tests check); these host tests pin
that implementation against the platform libm, called through Python's math module:
at most 2 ulp apart on random arguments over the domains the path uses and beyond, and
IEEE special cases (zeros, infinities, NaN, quadrants) as C99 Annex F gives them.
"""
import math
import struct

import numpy as np
import pytest

import _oracle as O


def ulp_diff(a, b):
    """Distance in units in the last place between two finite doubles of equal sign."""
    ia = struct.unpack("<q", struct.pack("<d", a))[0]
    ib = struct.unpack("<q", struct.pack("<d", b))[0]
    if ia < 0:
        ia = -(ia & 0x7FFFFFFFFFFFFFFF)
    if ib < 0:
        ib = -(ib & 0x7FFFFFFFFFFFFFFF)
    return abs(ia - ib)


def _max_ulp(fn, xs, ys=None):
    got = O.elemfn(fn, xs, ys)
    ref_f = {"log": math.log, "sin": math.sin, "cos": math.cos, "acos": math.acos}.get(fn)
    worst = 0
    for i, x in enumerate(xs):
        ref = math.atan2(x, ys[i]) if fn == "atan2" else ref_f(x)
        if ref == 0.0:
            assert got[i] == 0.0, (fn, x)
            continue
        worst = max(worst, ulp_diff(float(got[i]), ref))
    return worst


def test_log_random_domains():
    rng = np.random.default_rng(11)
    xs = np.concatenate([
        rng.uniform(1e-9, 1.0, 20000),                 # d / dhat of the barrier
        np.exp(rng.uniform(-700, 700, 20000)),         # whole normal range
        1.0 + rng.uniform(-1e-6, 1e-6, 5000),          # near 1 (cancellation)
        rng.uniform(0.70, 1.45, 5000),                 # around the sqrt(2) split
        np.array([5e-324, 1e-310, 2.2250738585072014e-308, 0.5, 2.0, 1.0, 1e308]),
    ])
    assert _max_ulp("log", xs) <= 1


def test_log_special_values():
    out = O.elemfn("log", np.array([0.0, -0.0, -1.0, np.inf, np.nan, 1.0]))
    assert out[0] == -np.inf and out[1] == -np.inf
    assert np.isnan(out[2]) and out[3] == np.inf and np.isnan(out[4]) and out[5] == 0.0


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_sin_cos_random(fn):
    rng = np.random.default_rng(12)
    xs = np.concatenate([
        rng.uniform(-np.pi, np.pi, 20000),             # half angles of a rotation update
        rng.uniform(-1e-3, 1e-3, 5000),                # small rotations
        rng.uniform(-200.0, 200.0, 5000),
        np.array([np.pi / 4, np.pi / 2, 3.0, 1e-300, -1e-8]),
    ])
    # results that cancel to near zero (x close to a multiple of pi/2, other than 0)
    # carry the reduction's absolute error; compare those in absolute terms
    ref = np.sin(xs) if fn == "sin" else np.cos(xs)
    big = np.abs(ref) > 1e-3
    assert _max_ulp(fn, xs[big]) <= 2
    got = O.elemfn(fn, xs[~big])
    assert np.max(np.abs(got - ref[~big]), initial=0.0) <= 1e-18 + 1e-16 * np.max(np.abs(xs[~big]), initial=0.0)


def test_sin_cos_special_values():
    s = O.elemfn("sin", np.array([0.0, -0.0, np.inf, np.nan]))
    c = O.elemfn("cos", np.array([0.0, -0.0, np.inf, np.nan]))
    assert s[0] == 0.0 and math.copysign(1.0, s[1]) == -1.0 and np.isnan(s[2]) and np.isnan(s[3])
    assert c[0] == 1.0 and c[1] == 1.0 and np.isnan(c[2]) and np.isnan(c[3])
    # sin^2 + cos^2 = 1 and the quadrant signs over four turns
    xs = np.linspace(-4 * np.pi, 4 * np.pi, 4001)
    s, c = O.elemfn("sin", xs), O.elemfn("cos", xs)
    assert np.max(np.abs(s * s + c * c - 1.0)) <= 4e-16
    assert np.all(np.sign(np.round(s, 12)) == np.sign(np.round(np.sin(xs), 12)))


def test_atan2_random_quadrants():
    rng = np.random.default_rng(13)
    ys = np.concatenate([rng.normal(size=20000), np.abs(rng.normal(size=5000)) * 1e-6])
    xs = np.concatenate([rng.normal(size=20000), np.abs(rng.normal(size=5000))])
    assert _max_ulp("atan2", ys, xs) <= 2
    # |v|, |w| of unit quaternions (the rotation-vector use)
    ang = rng.uniform(0, np.pi, 10000)
    assert _max_ulp("atan2", np.sin(ang / 2), np.abs(np.cos(ang / 2))) <= 2


def test_atan2_special_values():
    inf = np.inf
    cases = [(0.0, 1.0), (-0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (0.0, 0.0), (0.0, -0.0), (-0.0, -0.0),
             (1.0, 0.0), (-1.0, 0.0), (inf, 1.0), (-inf, 1.0), (1.0, inf), (1.0, -inf), (-1.0, -inf),
             (inf, inf), (inf, -inf), (-inf, -inf), (3.0, 3.0), (-2.0, -2.0)]
    ys = np.array([c[0] for c in cases])
    xs = np.array([c[1] for c in cases])
    got = O.elemfn("atan2", ys, xs)
    for (y, x), g in zip(cases, got):
        ref = math.atan2(y, x)
        assert ulp_diff(float(g), ref) <= 1 and math.copysign(1.0, g) == math.copysign(1.0, ref), (y, x, g, ref)
    assert np.isnan(O.elemfn("atan2", np.array([np.nan]), np.array([1.0]))[0])


def test_acos_random():
    rng = np.random.default_rng(14)
    xs = np.concatenate([
        rng.uniform(-1.0, 1.0, 20000),
        1.0 - np.abs(rng.normal(0, 1e-6, 5000)),       # slerp between close rotations
        np.array([-1.0, -0.5, 0.0, 0.5, 1.0, 0.9999999999999999]),
    ])
    xs = xs[xs <= 1.0]
    assert _max_ulp("acos", xs) <= 2
    assert np.isnan(O.elemfn("acos", np.array([1.0000000000000002]))[0])
