"""Pins the rounding of the oracle's squared-distance restatements with an
exact-rational emulator (CPU only).

fuel's distance (src/distance/squared_euclidean.rs) has three arithmetic forms:
  a1  f32, AVX+FMA (:13-39): 8 lanes interleaved over the coordinates, each
      lane acc_j = fma(t, t, acc_j) from +0 with t = fl(a_i - b_i), a left-fold
      lane sum (Rust's f32 Sum), then a fused tail s = fma(t, t, s);
  a2  f64, AVX+FMA (:41-67): the same with 4 lanes;
  a3  the fallback (:99-133) of every host without AVX+FMA, f32 and f64:
      8 lanes of vsum_j = fl(t * t) + vsum_j (unfused), a left fold from +0,
      an unfused tail.
The emulator below performs every operation on exact rationals and rounds once
per IEEE operation (round to nearest, ties to even), so it fixes each rounding
of those forms independently of any floating-point library.  The inputs mix
magnitudes (1e-3 .. 3e3 per coordinate) so that the plausible mistakes --
an unfused lane update, 4 instead of 8 lanes (or 8 instead of 4), a single
sequential accumulator, a reversed or pairwise lane fold, an unfused or fused
tail in the wrong form -- each change the result bits of some pair; the tests
assert that every such variant is observable on these inputs and that the
oracle equals the reference form bit for bit.
"""
from fractions import Fraction

import numpy as np
import pytest


def _rn32(fr):
    """Exact round-to-nearest-even of a Fraction to float32."""
    f = np.float32(float(fr))
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        if not np.isfinite(c):
            continue
        err = abs(Fraction(float(c)) - fr)
        if best is None or err < best[0] or (err == best[0] and (int(c.view(np.uint32)) & 1) == 0):
            best = (err, c)
    return Fraction(float(best[1]))


def _rn64(fr):
    # int / int true division in CPython is correctly rounded (ties to even)
    return Fraction(fr.numerator / fr.denominator) if fr.denominator != 1 else Fraction(float(fr))


def emulate(a, b, lanes, fused, tail_fused, fold="left", rn=_rn32):
    """Squared distance of a and b with the given arithmetic form, exactly."""
    a = [Fraction(float(x)) for x in a]
    b = [Fraction(float(x)) for x in b]
    d = len(a)
    sd = (d // lanes) * lanes if lanes else 0
    acc = [Fraction(0)] * max(lanes, 1)
    for i in range(0, sd, lanes):
        for j in range(lanes):
            t = rn(a[i + j] - b[i + j])
            acc[j] = rn(t * t + acc[j]) if fused else rn(rn(t * t) + acc[j])
    if fold == "left":
        s = Fraction(0)
        for j in range(lanes):
            s = rn(s + acc[j])
    elif fold == "right":
        s = Fraction(0)
        for j in reversed(range(lanes)):
            s = rn(s + acc[j])
    else:  # pairwise tree (a horizontal-add reduction)
        v = list(acc[:lanes])
        while len(v) > 1:
            v = [rn(v[2 * i] + v[2 * i + 1]) for i in range(len(v) // 2)]
        s = v[0]
    for i in range(sd, d):
        t = rn(a[i] - b[i])
        s = rn(t * t + s) if tail_fused else rn(rn(t * t) + s)
    return s


def _pairs(rng, d, count, dtype):
    out = []
    for _ in range(count):
        sa = rng.choice([1e-3, 1.0, 3e3], d)
        sb = rng.choice([1e-3, 1.0, 3e3], d)
        a = (rng.standard_normal(d) * sa).astype(dtype)
        b = (rng.standard_normal(d) * sb).astype(dtype)
        out.append((a, b))
    return out


def _bits(x, dtype):
    return np.array(float(x), dtype).view(np.uint32 if dtype == np.float32 else np.uint64).item()


DIMS = [7, 8, 9, 31, 128, 784]

# plausible mistakes of the a1/a2 restatement (lanes, fused lanes, fused tail, fold)
VARIANTS_AVX = {
    "unfused lanes": dict(fused=False),
    "half the lanes": dict(lanes_scale=0.5),
    "twice the lanes": dict(lanes_scale=2),
    "one sequential accumulator": dict(lanes=1),
    "unfused tail": dict(tail_fused=False),
    "reversed lane fold": dict(fold="right"),
    "pairwise lane fold": dict(fold="tree"),
}


def _variant_value(a, b, base_lanes, var, rn, fused=True, tail_fused=True):
    lanes = var.get("lanes", int(base_lanes * var.get("lanes_scale", 1)))
    return emulate(a, b, lanes, var.get("fused", fused), var.get("tail_fused", tail_fused), var.get("fold", "left"),
                   rn)


def _which_observable(d, base_lanes):
    """Variants that can differ at dimension d (a lane variant needs >= 2 lane steps
    of the form it changes, a tail variant needs a tail)."""
    names = []
    for name, var in VARIANTS_AVX.items():
        lanes = var.get("lanes", int(base_lanes * var.get("lanes_scale", 1)))
        tail = d - (d // base_lanes) * base_lanes
        if name == "unfused tail" and tail == 0:
            continue
        if name == "unfused lanes" and d < 2 * base_lanes:  # one step from +0 rounds once either way
            continue
        if name in ("reversed lane fold", "pairwise lane fold") and d < base_lanes:
            continue
        if name in ("half the lanes", "twice the lanes") and d < min(lanes, base_lanes):
            continue
        if name == "one sequential accumulator" and d < base_lanes:  # below, both are one fused chain
            continue
        names.append(name)
    return names


@pytest.mark.parametrize("d", DIMS)
def test_a1_f32_avx_bits_equal_exact_emulation(oracle, d):
    rng = np.random.default_rng(1000 + d)
    pairs = _pairs(rng, d, 64 if d <= 31 else (24 if d <= 128 else 8), np.float32)
    seen = set()
    for a, b in pairs:
        exp = emulate(a, b, 8, True, True, "left", _rn32)
        got = oracle.sqdist(a, b)
        assert got.view(np.uint32) == _bits(exp, np.float32), d
        for name in _which_observable(d, 8):
            if _variant_value(a, b, 8, VARIANTS_AVX[name], _rn32) != exp:
                seen.add(name)
    assert seen == set(_which_observable(d, 8)), (d, sorted(set(_which_observable(d, 8)) - seen))


@pytest.mark.parametrize("d", DIMS)
def test_a2_f64_avx_bits_equal_exact_emulation(oracle, d):
    rng = np.random.default_rng(2000 + d)
    pairs = _pairs(rng, d, 64 if d <= 31 else (24 if d <= 128 else 8), np.float64)
    seen = set()
    for a, b in pairs:
        exp = emulate(a, b, 4, True, True, "left", _rn64)
        got = oracle.sqdist(a, b)
        assert np.float64(got).view(np.uint64) == _bits(exp, np.float64), d
        for name in _which_observable(d, 4):
            if _variant_value(a, b, 4, VARIANTS_AVX[name], _rn64) != exp:
                seen.add(name)
    assert seen == set(_which_observable(d, 4)), (d, sorted(set(_which_observable(d, 4)) - seen))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", DIMS)
def test_a3_fallback_bits_equal_exact_emulation(oracle, d, dtype):
    rn = _rn32 if dtype == np.float32 else _rn64
    rng = np.random.default_rng(3000 + d + (7 if dtype == np.float64 else 0))
    pairs = _pairs(rng, d, 64 if d <= 31 else (24 if d <= 128 else 8), dtype)
    fused_seen = False
    for a, b in pairs:
        exp = emulate(a, b, 8, False, False, "left", rn)
        got = oracle.sqdist(a, b, "fallback")
        assert _bits(got, dtype) == _bits(exp, dtype), d
        # the fallback differs from the fused form on these inputs
        if emulate(a, b, 8, True, True, "left", rn) != exp:
            fused_seen = True
    # with d = 8 every lane takes one step from +0 and there is no tail: then
    # both forms round each square once and coincide
    assert fused_seen == (d % 8 != 0 or d >= 16), d


def test_emulator_agrees_with_ieee_on_single_operations():
    # the emulator's rounding is IEEE's: one f32 fma and one f32 multiply-add
    rng = np.random.default_rng(5)
    for _ in range(200):
        t, c = (rng.standard_normal(2) * rng.choice([1e-3, 1.0, 3e3], 2)).astype(np.float32)
        prod = np.float32(t) * np.float32(t)
        assert _rn32(Fraction(float(t)) ** 2) == Fraction(float(prod))
        assert _rn32(Fraction(float(prod)) + Fraction(float(c))) == Fraction(float(np.float32(prod + c)))
        fma = np.float32(np.float64(t) * np.float64(t) + np.float64(c))  # exact product in f64, one
        exact = Fraction(float(t)) ** 2 + Fraction(float(c))            # f64 rounding of the sum
        if Fraction(float(np.float64(t) * np.float64(t) + np.float64(c))) == exact:
            assert _rn32(exact) == Fraction(float(fma))
        x, y = rng.standard_normal(2)
        assert _rn64(Fraction(float(x)) * Fraction(float(y))) == Fraction(float(x * y))
        assert _rn64(Fraction(float(x)) + Fraction(float(y))) == Fraction(float(x + y))
