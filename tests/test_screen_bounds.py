"""Rounding margins of the distance screens, checked on the CPU with exact
arithmetic.

The FP64 screen tests the expanded form s = a.(-2b) + |b|^2 (three fused
multiply-adds) against T - |a|^2 + delta, delta = 32 u (T + 4 m^2), u = 2^-53,
m = the largest |coordinate|.  Every pair whose reference d2
(dx*dx + dy*dy + dz*dz, distance_analysis.cu:202-338) is <= T must pass.  The
worst case is T equal to the pair's own d2 with heavy cancellation (centres far
from the origin, close to each other); the fused multiply-adds are emulated
exactly (rational arithmetic, one correctly rounded result).

The matrix-core form of the same test (v_mfma_f64_16x16x4f64: the 4-term dot
product [ax ay az 1].[-2bx -2by -2bz |b|^2]) has an undocumented internal
order; the margin is checked against the largest value over every order of
rounded products and sums, pairwise trees and fused chains.

A planted weakening of the margin must fail on the same samples, so the check
has teeth.
"""
import itertools
from fractions import Fraction

import numpy as np

U53 = 2.0 ** -53


def fma(x, y, z):
    """fl(x*y + z) with one rounding (Fraction -> float is correctly rounded)."""
    return float(Fraction(x) * Fraction(y) + Fraction(z))


def ref_d2(a, b):
    dx, dy, dz = a[0] - b[0], a[1] - b[1], a[2] - b[2]
    return (dx * dx + dy * dy) + dz * dz


def expanded_passes(a, b, T, m, coef_t=32.0, coef_m=4.0):
    # side-2 record as staged: (-2b, |b|^2)
    bx, by, bz = -2.0 * b[0], -2.0 * b[1], -2.0 * b[2]
    nb = (b[0] * b[0] + b[1] * b[1]) + b[2] * b[2]
    s = fma(a[0], bx, fma(a[1], by, fma(a[2], bz, nb)))
    na = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]
    delta = coef_t * U53 * (T + coef_m * m * m)
    thr = (T - na) + delta
    return s <= thr


def rnd(x):
    return float(x)  # Fraction -> nearest double


def dot4_worst(a, b):
    """Largest computed value of [ax ay az 1].[-2bx -2by -2bz |b|^2] over the
    evaluation orders a matrix core may use (each operation correctly rounded)."""
    x = [a[0], a[1], a[2], 1.0]
    y = [-2.0 * b[0], -2.0 * b[1], -2.0 * b[2], (b[0] * b[0] + b[1] * b[1]) + b[2] * b[2]]
    prods = [Fraction(x[k]) * Fraction(y[k]) for k in range(4)]
    p = [rnd(v) for v in prods]
    vals = []
    for perm in itertools.permutations(range(4)):
        acc = p[perm[0]]  # rounded products, summed left to right
        for k in perm[1:]:
            acc = rnd(Fraction(acc) + Fraction(p[k]))
        vals.append(acc)
        acc = p[perm[3]]  # fused chain
        for k in (perm[2], perm[1], perm[0]):
            acc = rnd(prods[k] + Fraction(acc))
        vals.append(acc)
    for (i, j), (k, l) in (((0, 1), (2, 3)), ((0, 2), (1, 3)), ((0, 3), (1, 2))):
        vals.append(rnd(Fraction(rnd(Fraction(p[i]) + Fraction(p[j]))) + Fraction(rnd(Fraction(p[k]) + Fraction(p[l])))))
    vals.append(rnd(sum(prods)))  # one rounding
    return max(vals)


def mfma_passes(a, b, T, m, coef_t=32.0, coef_m=4.0):
    na = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]
    thr = (T - na) + coef_t * U53 * (T + coef_m * m * m)
    return dot4_worst(a, b) <= thr


def samples(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        scale = 10.0 ** rng.uniform(-3, 3)
        a = rng.uniform(-scale, scale, 3)
        rel = 10.0 ** rng.uniform(-9, -1)  # distance relative to the coordinates
        b = a + rng.normal(0, scale * rel, 3)
        if rng.random() < 0.2:  # exactly representable grid values, ties in the roundings
            a = np.round(a * 1024) / 1024
            b = np.round(b * 1024) / 1024
        out.append((a.astype(np.float64), b.astype(np.float64)))
    return out


def margin_violations(coef_t, coef_m, n=6000, seed=7, test=None):
    test = test or expanded_passes
    bad = 0
    for a, b in samples(seed, n):
        T = ref_d2(a, b)
        m = float(max(np.max(np.abs(a)), np.max(np.abs(b))))
        a_t, b_t = tuple(float(x) for x in a), tuple(float(x) for x in b)
        if not test(a_t, b_t, T, m, coef_t, coef_m):
            bad += 1
    return bad


def test_expanded_screen_margin_keeps_every_pair_at_threshold():
    assert margin_violations(32.0, 4.0) == 0


def test_planted_margin_too_small_is_caught():
    # no margin at all, and a margin without the |coordinate| term: the
    # cancellation in |a|^2 + |b|^2 - 2 a.b is then not covered
    assert margin_violations(0.0, 0.0) > 0
    assert margin_violations(32.0, 0.0) > 0


def test_fma_emulation_is_single_rounding():
    # 1 + 2^-53 is not representable: fl(x*y + z) keeps the product exact
    x = 1.0 + 2.0 ** -30
    exact = Fraction(x) * Fraction(x) - 1
    assert fma(x, x, -1.0) == float(exact)
    assert fma(x, x, -1.0) != (x * x) - 1.0


def test_matrix_core_form_keeps_every_pair_at_threshold():
    assert margin_violations(32.0, 4.0, n=1500, seed=11, test=mfma_passes) == 0


def test_matrix_core_planted_margin_too_small_is_caught():
    assert margin_violations(0.0, 0.0, n=1500, seed=11, test=mfma_passes) > 0
    assert margin_violations(32.0, 0.0, n=1500, seed=11, test=mfma_passes) > 0
