"""CPU checks of the arithmetic arguments the IEEE build's shortcuts rest on (device code in
linon_amd/csrc/device/wgsl_lib.hpp), in exact rational arithmetic or with float64 standing in
for an FMA where that is exact.  The GPU self-tests check the device code itself
(tests/test_gpu_parity.py); these pin the mathematics independently of the device."""
from fractions import Fraction as F

import numpy as np

C = np.float32(0.001)


def _exact(x: np.float32) -> F:
    b = int(np.float32(x).view(np.uint32))
    e = (b >> 23) & 0xFF
    m = b & 0x7FFFFF
    v = F(m | 0x800000) * F(2) ** (e - 150) if e else F(m) * F(2) ** -149
    return -v if b >> 31 else v


def _round_bits(x: F, bits: int) -> F:
    """x > 0 rounded to `bits` significant bits, ties to even."""
    e = 0
    while x >= 2 ** bits:
        x /= 2
        e += 1
    while x < 2 ** (bits - 1):
        x *= 2
        e -= 1
    n, r = int(x), x - int(x)
    if r > F(1, 2) or (r == F(1, 2) and n % 2):
        n += 1
    return F(n) * F(2) ** e


def _round_int(x: F) -> int:
    n, r = int(x), x - int(x)
    if r > F(1, 2) or (r == F(1, 2) and n % 2):
        n += 1
    return n


def test_lens_floor_reciprocal_constant():
    """Y_C of quot_lens_floor is RN(1 / RN(0.001)): 0x4479ffff."""
    c = _exact(C)
    y = np.float32(999.99993896484375)
    assert int(y.view(np.uint32)) == 0x4479FFFF
    below = _exact(np.nextafter(y, np.float32(0)))
    above = _exact(np.nextafter(y, np.float32(2000)))
    target = 1 / c
    assert abs(_exact(y) - target) < abs(below - target) and abs(_exact(y) - target) < abs(above - target)


def test_lens_floor_subnormal_quotients():
    """For every subnormal numerator a = j 2^-149 whose quotient a / c is subnormal (j <= 8388),
    rounding the 24-bit quotient Q = RN(a 2^48 / c) once more to the subnormal grid gives the
    correctly rounded a / c, and that is 1000 j; j = 8389 is the first normal quotient."""
    c = _exact(C)
    for j in range(1, 8389):
        x = F(j) / c  # a / c in units of 2^-149
        assert x < 2 ** 23
        assert _round_int(x) == 1000 * j
        assert _round_int(_round_bits(x, 24)) == 1000 * j, j
    assert F(8389) / c >= 2 ** 23


def test_refraction_k_gap():
    """k = RN(1 - e t) (one FMA) for floats e and t in [0, 1] is never in (0, 2^-48): near
    e t = 1 the product is a multiple of 2^-48.  float64 forms e t (48-bit product) and 1 - e t
    (within a factor 2 of 1) exactly here, so float32 of it is the FMA.  Drawn near the critical
    product: e = 1 + a 2^-23, t = 1 - b 2^-24 with b within 3 of 2a (e t = 1 + (2a - b) 2^-24 -
    a b 2^-47), and random e in [0.5, 2) with t within 16 ulps of 1 / e."""
    a = np.arange(0, 4096, dtype=np.int64)
    d = np.arange(-3, 4, dtype=np.int64)
    A = np.repeat(a, d.size)
    B = np.clip(2 * A + np.tile(d, a.size), 0, 2 ** 23)
    e1 = (np.float64(1.0) + A * 2.0 ** -23).astype(np.float32)
    t1 = (np.float64(1.0) - B * 2.0 ** -24).astype(np.float32)
    rng = np.random.default_rng(7)
    e2 = (0.5 + 1.5 * rng.random(400_000)).astype(np.float32)
    t0 = (np.float32(1.0) / e2).astype(np.float32)
    off = rng.integers(-16, 17, size=e2.size).astype(np.int64)
    t2 = (t0.view(np.uint32).astype(np.int64) + off).astype(np.uint32).view(np.float32)
    e = np.concatenate([e1, e2])
    t = np.minimum(np.concatenate([t1, t2]), np.float32(1.0))
    k = (1.0 - e.astype(np.float64) * t.astype(np.float64)).astype(np.float32)
    assert not ((k != 0) & (np.abs(k) < np.float32(2.0 ** -48))).any()
    nz = np.abs(k[k != 0])
    assert (k == 0).any() and nz.min() <= 2.0 ** -46  # the critical cases occur in the draw


def _div6_device_form(s: np.ndarray) -> np.ndarray:
    """div6 of the ray-cast kernels (q0 = RN(s y), r' = RN(6 q0 - s), q = RN(q0 - r' y), y =
    RN(1/6)) for |s| < 1.5 2^-124, each FMA formed in float64 and rounded once to float32.  At
    these magnitudes every operand is a multiple of 2^-149 (or y's 2^-26) with few significant
    bits, so float64 forms the FMAs exactly."""
    y = np.float32(1.0) / np.float32(6.0)
    q0 = (s.astype(np.float64) * np.float64(y)).astype(np.float32)
    rn = (6.0 * q0.astype(np.float64) - s.astype(np.float64)).astype(np.float32)
    return (q0.astype(np.float64) - rn.astype(np.float64) * np.float64(y)).astype(np.float32)


def test_rk4_division_by_six():
    """div6 equals IEEE s / 6 (zeros' signs included) at every s whose quotient is subnormal
    (|s| < 1.5 2^-124, both signs) except the documented ties: s = (6 m + 3) 2^-149 with m odd,
    where it returns m 2^-149 instead of the even (m + 1) 2^-149.  (The device self-test compares
    the device's own div6 with the division at every float.)"""
    top = int(np.float32(1.5 * 2.0 ** -124).view(np.uint32))
    ties = 0
    for lo in range(0, top, 1 << 23):
        bits = np.arange(lo, min(lo + (1 << 23), top), dtype=np.uint32)
        for sign in (0, 0x80000000):
            sv = (bits | np.uint32(sign)).view(np.float32)
            got = _div6_device_form(sv)
            want = sv / np.float32(6.0)
            bad = got.view(np.uint32) != want.view(np.uint32)
            jj = np.rint(np.abs(sv.astype(np.float64)) * 2.0 ** 149).astype(np.int64)  # s = j 2^-149
            tie_odd = (jj % 6 == 3) & (((jj - 3) // 6) % 2 == 1)
            assert np.array_equal(bad, tie_odd), lo
            assert (np.abs(want[bad].astype(np.float64) - got[bad].astype(np.float64)) == 2.0 ** -149).all()
            ties += int(bad.sum())
    assert ties > 0


def test_hot_air_temperature_accuracy_does_not_reach_the_index_ratio():
    """Hot air's default variant takes its sigmoid at WGSL's accuracy (the hardware 2^t of
    RN(y log2 e) and the hardware reciprocal).  The temperature T = 20 + 80 s enters the step only
    through the refraction indices n(T) = c1 P (1 + P (60.1 - 0.972 T) 1e-10) / (1 + 0.00366 T)
    (src/main_view.wgsl:111-146).  Checked here: RN(log2 e) and the product's rounding err by at
    most 7.3e-8 relative; that error moves T by at most 80 max(y e^-y) 7.3e-8 <= 2.2e-6 degrees
    and an ulp of s by 80 2^-24; n's relative slope over hot air's [20, 100] degrees is at most
    3.5e-3 per degree; so either moves n(T) by less than 2e-8 relative, below a half ulp of
    the index ratio (2^-24 relative at least 5.9e-8)."""
    L = F(int(np.float32(np.log2(np.e)).view(np.uint32)) & 0x7FFFFF | 0x800000, 2 ** 23)
    rel_L = abs(float(L) - np.log2(np.e)) / np.log2(np.e)
    rel = rel_L + 2.0 ** -24
    assert rel <= 7.3e-8, rel
    y = np.linspace(0.0, 100.0, 200001)
    dT_exp = 80.0 * np.max(y * np.exp(-y)) * 7.3e-8
    assert dT_exp <= 2.2e-6, dT_exp
    dT_ulp = 80.0 * 2.0 ** -24
    P, c1 = 101325.0, 0.0000104
    T = np.linspace(20.0, 100.0, 80001)
    n = c1 * P * (1.0 + P * (60.1 - 0.972 * T) * 1e-10) / (1.0 + 0.00366 * T)
    slope = np.max(np.abs(np.gradient(np.log(n), T)))
    assert slope <= 3.5e-3, slope
    moved = slope * max(dT_exp, dT_ulp)
    assert moved < 2e-8 < 2.0 ** -24, moved
