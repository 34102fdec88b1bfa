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


# ---------------------------------------------------------------------------
# FP32SCREEN: the float screen of round 1 reads centres rounded to float and
# a float threshold t widened by k_inflate (jsi_kernels.hip):
#   inflate(t) = ru32((sqrt(t) g + 4 sqrt(3) E g)^2 g + 1e-36), g = 1 + 16u,
#   u = 2^-24, E = 2^-24 max |centre coordinate|.
# t is at least the float d2 of k real pairs (seed pairs, or the k-th record
# of a capped screen), so the FP64 k-th d2 D is at most the largest FP64 d2 of
# those pairs; every pair of the FP64 top-k has FP64 d2 <= D and must pass the
# float test fl32 d2 <= inflate(t).  The float d2 may be evaluated with fused
# multiply-adds in any order (the screen is compiled with contraction on).
# ---------------------------------------------------------------------------
U24 = 2.0 ** -24


def f32(x):
    """A double rounded to the nearest float32 (the centre copies of the screen)."""
    return float(np.float32(x))


def r32(q):
    """Correct rounding of an exact rational to float32 (via the exact double
    nearest, then a check of the two float neighbours)."""
    d = float(q)
    c = np.float32(d)
    best = None
    for cand in (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))):
        err = abs(Fraction(float(cand)) - q)
        if best is None or err < best[0] or (err == best[0] and int(np.float32(cand).view(np.uint32)) % 2 == 0):
            best = (err, float(cand))
    return best[1]


def float_d2_orders(a, b):
    """Every value of (dx^2 + dy^2) + dz^2 over float evaluation orders with
    products rounded or fused into the next add; a, b are float32 centres."""
    d = [r32(Fraction(a[c]) - Fraction(b[c])) for c in range(3)]
    sq = [Fraction(x) * Fraction(x) for x in d]
    out = []
    for perm in itertools.permutations(range(3)):
        p0, p1, p2 = perm
        # products rounded, sum left to right
        out.append(r32(Fraction(r32(Fraction(r32(sq[p0])) + Fraction(r32(sq[p1])))) + Fraction(r32(sq[p2]))))
        # fused chain fma(d2, d2, fma(d1, d1, d0*d0))
        out.append(r32(sq[p2] + Fraction(r32(sq[p1] + Fraction(r32(sq[p0]))))))
        # one fused add
        out.append(r32(sq[p2] + Fraction(r32(Fraction(r32(sq[p0])) + Fraction(r32(sq[p1]))))))
        out.append(r32(Fraction(r32(sq[p2])) + Fraction(r32(sq[p1] + Fraction(r32(sq[p0]))))))
    return out


def inflate(t, E, e_coef=4.0, g_coef=16.0):
    """k_inflate's arithmetic in double, rounded up to float."""
    g = 1.0 + g_coef * U24
    r = np.sqrt(np.float64(t)) * g + e_coef * 1.7320508075688772 * E * g
    v = r * r * g + 1e-36
    f = np.float32(v)
    if float(f) < v:
        f = np.nextafter(f, np.float32(np.inf))
    return float(f)


def fp32screen_samples(seed, n):
    """(A, B, C, D): C, D give the threshold pair, A, B a pair whose FP64 d2
    is just below C, D's (the hardest pair the FP64 top-k can hold)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        scale = 10.0 ** rng.uniform(-2, 3)
        rel = 10.0 ** rng.uniform(-7, -1)
        c = rng.uniform(-scale, scale, 3)
        dvec = rng.normal(0, scale * rel, 3)
        dd = c + dvec
        a = c + rng.uniform(-scale * rel, scale * rel, 3) * 5
        dir_ = rng.normal(0, 1, 3)
        dir_ /= np.linalg.norm(dir_)
        b = a + dir_ * np.linalg.norm(dvec) * (1 - 10.0 ** rng.uniform(-12, -6))
        if rng.random() < 0.25:  # coordinates on a coarse grid: ties in the float roundings
            q = 2.0 ** int(rng.integers(8, 22))
            a, b, c, dd = (np.round(x * q) / q for x in (a, b, c, dd))
        out.append(tuple(x.astype(np.float64) for x in (a, b, c, dd)))
    return out


def fp32screen_violations(n=1200, seed=3, **kw):
    bad = 0
    for A, B, Cc, D in fp32screen_samples(seed, n):
        if ref_d2(A, B) > ref_d2(Cc, D):
            continue
        E = float(max(np.max(np.abs(x)) for x in (A, B, Cc, D))) * U24
        a, b, c, d = ([f32(v) for v in x] for x in (A, B, Cc, D))
        t = min(float_d2_orders(c, d))  # the threshold pair at its smallest float d2
        if max(float_d2_orders(a, b)) > inflate(t, E, **kw):
            bad += 1
    return bad


def test_fp32screen_inflate_keeps_every_fp64_topk_pair():
    assert fp32screen_violations() == 0


def test_fp32screen_planted_weaker_inflate_is_caught():
    # a widening that covers the centre rounding only once, and none at all
    assert fp32screen_violations(n=400, e_coef=1.0) > 0
    assert fp32screen_violations(n=400, e_coef=0.0, g_coef=0.0) > 0


# ---------------------------------------------------------------------------
# FP32SCREEN refinement rounds (k_dist_filter_mfma32, jsi_kernels.hip): each
# side relative to its own origin, a' = fl32(a - o1), b' = fl32(b - o2),
# c = fl64(o1 - o2); the matrix core forms
#   D = [-2b' 1] . [a' fl32(P - kd)] + fl32(Q),
#   P = 2 c.a' + |a'|^2, Q = |b'|^2 - 2 c.b' (float fused chains, c in float),
# and a pair passes when D <= 0.  kd = ru32(K + delta) with
#   R = sqrt(T (1 + 1e-12)) (1 + 1e-14) + 1.01 u (A + B) + 1.01 u64 |c|_1,
#   K = R^2 - |c|^2, delta = (16 u (2AB + B^2 + 2|c|B + 2|c|A + A^2 + |K|)
#                             + 8 u64 (R^2 + |c|^2 + 2|c|A + A^2)) (1 + 1e-6) + 1e-30.
# Every pair whose FP64 d2 is at most T (1 + 1e-13) must pass for every
# evaluation order of the 5-term sum.  A, B are the smallest bounds the kernel
# could use (|a'|, |b'| themselves; the kernel's box diagonals are larger).
# ---------------------------------------------------------------------------


def fmaf(x, y, z):
    return r32(Fraction(x) * Fraction(y) + Fraction(z))


def ru32(v):
    f = np.float32(v)
    if float(f) < v:
        f = np.nextafter(f, np.float32(np.inf))
    return float(f)


def refine32_kd(T, A, B, c, delta_coef=16.0, eps=True):
    cx, cy, cz = c
    sqt = np.sqrt(T * (1.0 + 1e-12)) * (1.0 + 1e-14)
    cn = (abs(cx) + abs(cy) + abs(cz)) * (1.0 + 1e-15)
    R = sqt + ((1.01 * U24 * (A + B) + 1.01 * U53 * cn) if eps else 0.0)
    R2 = R * R
    c2 = cx * cx + cy * cy + cz * cz
    K = R2 - c2
    delta = (delta_coef * U24 * (2.0 * A * B + B * B + 2.0 * cn * B + 2.0 * cn * A + A * A + abs(K)) +
             8.0 * U53 * (R2 + c2 + 2.0 * cn * A + A * A)) * (1.0 + 1e-6) + 1e-30
    return ru32(K + delta)


def refine32_worst(a, b, o1, o2, T, **kw):
    """Largest computed D of the pair over the orders a matrix core may use."""
    c = tuple(float(Fraction(o1[k]) - Fraction(o2[k])) for k in range(3))
    ap = [f32(float(Fraction(a[k]) - Fraction(o1[k]))) for k in range(3)]
    bp = [f32(float(Fraction(b[k]) - Fraction(o2[k]))) for k in range(3)]
    cf = [f32(x) for x in c]
    A = float(np.sqrt(float(sum(Fraction(x) * Fraction(x) for x in ap)))) * (1.0 + 1e-15)
    B = float(np.sqrt(float(sum(Fraction(x) * Fraction(x) for x in bp)))) * (1.0 + 1e-15)
    kd = refine32_kd(T, A, B, c, **kw)
    a2 = fmaf(ap[0], ap[0], fmaf(ap[1], ap[1], r32(Fraction(ap[2]) * Fraction(ap[2]))))
    pa = fmaf(2.0 * cf[0], ap[0], fmaf(2.0 * cf[1], ap[1], fmaf(2.0 * cf[2], ap[2], a2)))
    b2 = fmaf(bp[0], bp[0], fmaf(bp[1], bp[1], r32(Fraction(bp[2]) * Fraction(bp[2]))))
    q = fmaf(-2.0 * cf[0], bp[0], fmaf(-2.0 * cf[1], bp[1], fmaf(-2.0 * cf[2], bp[2], b2)))
    x = [ap[0], ap[1], ap[2], r32(Fraction(pa) - Fraction(kd))]
    y = [-2.0 * bp[0], -2.0 * bp[1], -2.0 * bp[2], 1.0]
    items = [Fraction(x[k]) * Fraction(y[k]) for k in range(4)] + [Fraction(q)]
    rounded = [r32(v) for v in items]
    worst = r32(sum(items))  # one rounding
    for perm in itertools.permutations(range(5)):
        acc = rounded[perm[0]]  # rounded terms, summed in this order
        for k in perm[1:]:
            acc = r32(Fraction(acc) + Fraction(rounded[k]))
        worst = max(worst, acc)
        acc = rounded[perm[0]]  # each product fused into the running sum
        for k in perm[1:]:
            acc = r32(items[k] + Fraction(acc))
        worst = max(worst, acc)
    return worst


def refine32_samples(seed, n):
    """Two refined units at distance |c| with cells of extent `ext` around
    their origins, world coordinates up to 1e3; a pair and its FP64 d2."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        world = 10.0 ** rng.uniform(-2, 3)
        gap = world * 10.0 ** rng.uniform(-6, -1)
        ext = gap * 10.0 ** rng.uniform(-6, 0.3)
        o1 = rng.uniform(-world, world, 3)
        d = rng.normal(0, 1, 3)
        o2 = o1 + d / np.linalg.norm(d) * gap
        a = o1 + rng.uniform(-ext, ext, 3)
        b = o2 + rng.uniform(-ext, ext, 3)
        if rng.random() < 0.25:  # coordinates on a grid: ties in the roundings
            s = 2.0 ** int(rng.integers(10, 30))
            o1, o2, a, b = (np.round(v * s) / s for v in (o1, o2, a, b))
        out.append(tuple(tuple(float(x) for x in v) for v in (a, b, o1, o2)))
    return out


def refine32_violations(n=300, seed=5, rel=(1.0, 1.0 - 1e-13), **kw):
    bad = 0
    for a, b, o1, o2 in refine32_samples(seed, n):
        d2 = ref_d2(a, b)
        for r in rel:  # threshold at the pair's own d2 and just below (pairs up to T (1 + 1e-13))
            if refine32_worst(a, b, o1, o2, d2 * r, **kw) > 0.0:
                bad += 1
    return bad


def test_refine32_screen_keeps_every_pair_at_threshold():
    assert refine32_violations() == 0


def test_refine32_planted_weaker_margins_are_caught():
    # no rounding margin at all, and the margin without the widening of R
    assert refine32_violations(n=120, delta_coef=0.0, eps=False) > 0
    assert refine32_violations(n=120, delta_coef=0.0) > 0


# ---------------------------------------------------------------------------
# FP32SCREEN intersection (k_isect_boxes32, jsi_kernels.hip): per axis, the
# float record (c, r) of a cell covers the hull of the FP64 acceptance
# interval [c64 -+ (r64 (1 + 4 u64) + 4 u64 (|c64| + r64))] and, in round 1,
# VERSION5's float one [c32 -+ (r32 (1 + 3u) + 4 u64 (|c32| + r32))]:
#   c = fl32(mid), r = ru32((half + |c - mid|)(1 + 3u) + 4 u64 (|mid| + half) + 1e-30).
# Every pair that the FP64 test fl64(|c1 - c2|) < fl64(r1 + r2) accepts, or
# that the float test accepts on the VERSION5 values, must pass the float test
# on the records.
# ---------------------------------------------------------------------------


def isect_record(c64, r64, c32=None, r32=None, k=3.0, a=4.0):
    w = r64 * (1.0 + 4.0 * U53) + a * U53 * (abs(c64) + r64)
    lo, hi = c64 - w, c64 + w
    if c32 is not None:
        w32 = r32 * (1.0 + 3.0 * U24) + a * U53 * (abs(c32) + r32)
        lo, hi = min(lo, c32 - w32), max(hi, c32 + w32)
    mid, half = 0.5 * (lo + hi), 0.5 * (hi - lo)
    cf = f32(mid)
    return cf, ru32((half + abs(cf - mid)) * (1.0 + k * U24) + a * U53 * (abs(mid) + half) + 1e-30)


def float_overlap(c1, r1, c2, r2):
    """k_isect_pass<float>: fl32(|c1 - c2|) < fl32(r1 + r2) on float values."""
    return r32(abs(Fraction(c1) - Fraction(c2))) < r32(Fraction(r1) + Fraction(r2))


def isect_samples(seed, n):
    """Cells at the edge of the overlap test: centres |c1 - c2| just inside
    r1 + r2 (relative gap 1e-1 .. 1e-16), world coordinates 1e-3 .. 1e3,
    radii 1e-9 .. 1e-1 of them; each with a nearby float (VERSION5-like) copy."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        world = 10.0 ** rng.uniform(-3, 3)
        r1 = world * 10.0 ** rng.uniform(-9, -1)
        r2 = r1 * 10.0 ** rng.uniform(-1, 1)
        c1 = rng.uniform(-world, world)
        c2 = c1 + rng.choice([-1.0, 1.0]) * (r1 + r2) * (1.0 - 10.0 ** rng.uniform(-16, -1))
        if rng.random() < 0.25:  # grid values: ties in the roundings
            q = 2.0 ** int(rng.integers(10, 40))
            c1, c2, r1, r2 = (np.round(x * q) / q for x in (c1, c2, r1, r2))
        noise = lambda x: f32(x * (1.0 + rng.normal(0, 1e-6)))
        out.append(tuple(float(x) for x in (c1, r1, c2, r2)) + (noise(c1), noise(r1), noise(c2), noise(r2)))
    return out


def isect_violations(n=4000, seed=17, **kw):
    bad = 0
    for c1, r1, c2, r2, v1, w1, v2, w2 in isect_samples(seed, n):
        fp64 = abs(c1 - c2) < (r1 + r2)
        v5 = float_overlap(v1, w1, v2, w2)
        if not (fp64 or v5):
            continue
        x1, y1 = isect_record(c1, r1, v1, w1, **kw)
        x2, y2 = isect_record(c2, r2, v2, w2, **kw)
        if not float_overlap(x1, y1, x2, y2):
            bad += 1
        if fp64:  # later rounds: records from the FP64 values alone
            x1, y1 = isect_record(c1, r1, **kw)
            x2, y2 = isect_record(c2, r2, **kw)
            if not float_overlap(x1, y1, x2, y2):
                bad += 1
    return bad


def test_isect32_records_keep_every_accepted_pair():
    assert isect_violations() == 0


def test_isect32_planted_weaker_records_are_caught():
    # records without the relative widening, and without any widening
    assert isect_violations(n=1500, k=0.0) > 0
    assert isect_violations(n=1500, k=0.0, a=0.0) > 0
