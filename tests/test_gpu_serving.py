"""Batches and the serving loop: every result equals the same query run
alone, whatever runs beside it; admission refuses work beyond the memory cap;
the surface cache evicts least-recently-used surfaces."""
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


def mixed_queries(ctx, n=24):
    qs, solo = [], []
    for i in range(n):
        a, b, nuv = S.random_pair(i, "dist" if i % 3 else "isect")
        nuv = min(nuv, 96)
        r1, r2 = S.full_region(a, nuv, 1), S.full_region(b, nuv, 2)
        if i % 3:
            p = default_params(k=[8, 64, 512][i % 3], precision=["exact32", "fp64", "fp32screen"][i % 3], trace=1)
            kind = QUERY_DISTANCE
        else:
            p = default_params(precision=["exact32", "fp64"][i % 2], isect_max_rounds=3, isect_refine_nuv=8, trace=1)
            kind = QUERY_INTERSECT
        qs.append(ctx.query(kind, a, r1, b, r2, p))
        if kind == QUERY_DISTANCE:
            res, _ = ctx.minimum_distance(a, r1, b, r2, p)
        else:
            res, _ = ctx.intersection(a, r1, b, r2, p)
        solo.append(res)
    return qs, solo


def same(a, b):
    assert a.rounds == b.rounds and a.npairs == b.npairs and a.distance == b.distance
    assert a.rounds_list() == b.rounds_list()


def test_batch_equals_solo(ctx):
    qs, solo = mixed_queries(ctx)
    res, st = ctx.run_batch(qs, concurrency=6)
    assert st == [0] * len(qs)
    for r, s in zip(res, solo):
        same(r, s)


def test_server_alone_and_in_the_mix(ctx):
    from jsi_cad_next_versions_amd.runtime import Server
    qs, solo = mixed_queries(ctx, 18)
    with Server(ctx, slots=4) as srv:
        # staggered arrivals: later requests are admitted while earlier ones run
        tickets = []
        for q in qs:
            tickets.append(srv.submit(q))
            time.sleep(0.001)
        for t, s in zip(tickets, solo):
            same(srv.poll(t), s)
        c = srv.counters()
        assert c["admitted"] == len(qs) and c["completed"] == len(qs)


def test_server_admission_refuses_beyond_cap(ctx):
    from jsi_cad_next_versions_amd.runtime import Server
    a, b = S.hard_wave_pair()
    q = ctx.query(QUERY_DISTANCE, a, S.full_region(a, 256, 1), b, S.full_region(b, 256, 2), default_params(k=64))
    # a cap below one query's estimate: the first is admitted (nothing in
    # flight), the next ones are refused while it runs
    with Server(ctx, slots=1, hbm_cap_bytes=1 << 20) as srv:
        t1 = srv.submit(q)
        assert t1 is not None
        refused = srv.submit(q) is None or srv.submit(q) is None
        assert refused
        assert srv.poll(t1).rounds == 4
        assert srv.counters()["refused"] >= 1


def test_server_surface_cache_lru(ctx):
    from jsi_cad_next_versions_amd.runtime import Server
    surfs = [S.random_pair(i)[0] for i in range(4)]
    cap = sum(s.ndata * 4 for s in surfs[:2]) + 16
    with Server(ctx, slots=1, cache_cap_bytes=cap) as srv:
        h0 = srv.surface(100, surfs[0])
        assert srv.surface(100, surfs[0]) == h0
        srv.surface(101, surfs[1])
        srv.surface(102, surfs[2])  # evicts key 100 (least recently used)
        c = srv.counters()
        assert c["cache_hits"] == 1 and c["cache_evictions"] >= 1
