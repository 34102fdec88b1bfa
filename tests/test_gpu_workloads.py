"""The benchmark's workloads at their own sizes.

* Hard-wave pair scaled 4x at nuv 1024, k 4096 (1.1e12 round-1 cell pairs,
  beyond what the oracle finishes in seconds): FP32SCREEN reproduces FP64
  exactly (every round's ordered list, the k-th distance, the polish), the
  final list is totally ordered under (d2, slot, i, j), and the unscaled pair
  gives the same lists with d2 / 16 (scaling by a power of two is exact);
  distances at nuv <= 128 are checked against the oracle elsewhere.
* The uneven batch (256 pairs of random_pair's distribution at their drawn
  nuv up to 512): every batched result equals the same query run alone, and
  the pairs at nuv <= 128 equal the oracle.
* Poisson arrivals of the uneven mix plus 10% 4x hard-wave pairs at nuv 1024
  through the server: each result equals the same query run alone.
"""
import time

import numpy as np
import pytest

from jsi_cad_next_versions_amd import surfaces as S
from jsi_cad_next_versions_amd._abi import QUERY_DISTANCE, QUERY_INTERSECT, default_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from jsi_cad_next_versions_amd.runtime import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc():
    from oracle import Oracle
    return Oracle()


def same(a, b):
    assert a.rounds == b.rounds and a.npairs == b.npairs and a.distance == b.distance
    assert a.rounds_list() == b.rounds_list()
    assert a.closest.distance == b.closest.distance and a.closest.candidate == b.closest.candidate


def totally_ordered(pairs):
    key = np.rec.fromarrays([pairs["d2"], pairs["slot"], pairs["i"], pairs["j"]])
    for x, y in zip(key[:-1], key[1:]):
        assert tuple(x) < tuple(y)


def test_hardwave_4x_nuv1024_fp32screen_equals_fp64(ctx):
    a, b = S.hard_wave_pair(scale=4.0)
    r1, r2 = S.full_region(a, 1024, 1), S.full_region(b, 1024, 2)
    print("hard wave 4x, nuv 1024, k 4096: fp64", flush=True)
    r64, p64 = ctx.minimum_distance(a, r1, b, r2, default_params(k=4096, precision="fp64", trace=1, polish=64))
    print("hard wave 4x, nuv 1024, k 4096: fp32screen", flush=True)
    r32, p32 = ctx.minimum_distance(a, r1, b, r2, default_params(k=4096, precision="fp32screen", trace=1, polish=64))
    same(r32, r64)
    np.testing.assert_array_equal(p32, p64)
    assert len(p64) == 4096 and r64.rounds == 4
    totally_ordered(p64)
    # Scaling the control points by 4 (a power of two, exact in float32)
    # scales every evaluated point, centre and rounding bound exactly, so the
    # unscaled pair gives the same ordered lists with d2 / 16, and the same
    # closest points at a quarter of the distance.
    a1, b1 = S.hard_wave_pair()
    print("hard wave 1x, nuv 1024, k 4096: fp64", flush=True)
    q64, s64 = ctx.minimum_distance(a1, S.full_region(a1, 1024, 1), b1, S.full_region(b1, 1024, 2),
                                    default_params(k=4096, precision="fp64", trace=1, polish=64))
    for x, y in zip(r64.rounds_list(), q64.rounds_list()):
        assert (x["npairs"], x["slots"], x["units1"], x["units2"]) == (y["npairs"], y["slots"], y["units1"],
                                                                     y["units2"])
        assert x["kth_d2"] == 16.0 * y["kth_d2"]
    np.testing.assert_array_equal(p64[["slot", "i", "j"]], s64[["slot", "i", "j"]])
    np.testing.assert_array_equal(p64["d2"], 16.0 * s64["d2"])
    assert abs(r64.closest.distance - 4.0 * q64.closest.distance) <= 1e-10 * r64.closest.distance
    # a polished distance is the distance of two surface points, so it cannot
    # fall below the surfaces' true minimum gap (1e-3 unscaled, 4e-3 here)
    assert r64.closest.distance >= 4.0e-3 * (1.0 - 1e-4)
    assert r64.closest.flags & 8  # converged


def uneven(ctx, n, seed0):
    qs, meta = [], []
    for kind, a, b, nuv in S.uneven_workload(n, seed0):
        r1, r2 = S.full_region(a, nuv, 1), S.full_region(b, nuv, 2)
        if kind == "dist":
            p = default_params(k=512, precision="fp32screen", polish=64, trace=1)
            qs.append(ctx.query(QUERY_DISTANCE, a, r1, b, r2, p))
        else:
            p = default_params(precision="fp64", trace=1)
            qs.append(ctx.query(QUERY_INTERSECT, a, r1, b, r2, p))
        meta.append((kind, a, b, r1, r2, p))
    return qs, meta


def solo(ctx, m):
    kind, a, b, r1, r2, p = m
    if kind == "dist":
        return ctx.minimum_distance(a, r1, b, r2, p)[0]
    return ctx.intersection(a, r1, b, r2, p)[0]


def test_uneven_batch_256_equals_solo_and_oracle(ctx, orc):
    qs, meta = uneven(ctx, 256, 0)
    print("uneven batch: 256 queries", flush=True)
    res, st = ctx.run_batch(qs, concurrency=8)
    assert st == [0] * len(qs)
    print("uneven batch: solo runs", flush=True)
    checked = 0
    for r, m in zip(res, meta):
        s = solo(ctx, m)
        same(r, s)
        kind, a, b, r1, r2, p = m
        if r1.nuv <= 128 and checked < 6:
            print(f"uneven batch: oracle {kind} nuv {r1.nuv}", flush=True)
            o = orc.min_distance(a, r1, b, r2, p)[0] if kind == "dist" else orc.intersect(a, r1, b, r2, p)[0]
            assert r.rounds_list() == o.rounds_list()
            assert r.distance == o.distance
            if kind == "dist":
                assert abs(r.closest.distance - o.closest.distance) <= 1e-12
            checked += 1
    assert checked == 6


def test_poisson_mix_with_4x_hardwave_alone_and_in_the_mix(ctx):
    from jsi_cad_next_versions_amd.runtime import Server
    qs, meta = uneven(ctx, 36, 10000)
    ha, hb = S.hard_wave_pair(scale=4.0)
    h1, h2 = S.full_region(ha, 1024, 1), S.full_region(hb, 1024, 2)
    hp = default_params(k=4096, precision="fp32screen", polish=64, trace=1)
    hq = ctx.query(QUERY_DISTANCE, ha, h1, hb, h2, hp)
    mix, mmeta = [], []
    for i in range(40):
        if i % 10 == 9:
            mix.append(hq)
            mmeta.append(("dist", ha, hb, h1, h2, hp))
        else:
            mix.append(qs[i - i // 10])
            mmeta.append(meta[i - i // 10])
    rng = np.random.default_rng(11)
    arrivals = np.cumsum(rng.exponential(1.0 / 400.0, len(mix)))  # 400 requests/s
    print("poisson mix: 40 requests at 400/s", flush=True)
    results = {}
    with Server(ctx, slots=8) as srv:
        t0 = time.perf_counter()
        tickets = []
        for q, t in zip(mix, arrivals):
            while time.perf_counter() - t0 < t:
                time.sleep(1e-4)
            tk = srv.submit(q)
            while tk is None:  # admission holds it back
                time.sleep(1e-3)
                tk = srv.submit(q)
            tickets.append(tk)
        for i, tk in enumerate(tickets):
            results[i] = srv.poll(tk)
    print("poisson mix: solo runs", flush=True)
    for i, m in enumerate(mmeta):
        same(results[i], solo(ctx, m))
