"""The allocation-free distance entry point (jsi_min_distance_noalloc)
against jsi_min_distance, through the C ABI: identical results (every
round's list, the distance, the closest points, the final pairs) on two calls
in a row, the second smaller, so the scratch array's leftovers from the first
call cannot leak into the second; a query beyond the scratch limit is
refused with JSI_ERR_NOMEM and a message."""
import numpy as np
import pytest

from jsi_cad_next_versions_amd import surfaces as S
from jsi_cad_next_versions_amd._abi import default_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from jsi_cad_next_versions_amd.runtime import Context
    c = Context(0)
    yield c
    c.close()


def same(a, b):
    assert a.rounds == b.rounds and a.npairs == b.npairs and a.distance == b.distance
    assert a.rounds_list() == b.rounds_list()
    assert a.closest.distance == b.closest.distance and a.closest.candidate == b.closest.candidate
    assert (a.closest.u1, a.closest.v1, a.closest.u2, a.closest.v2) == (b.closest.u1, b.closest.v1, b.closest.u2,
                                                                        b.closest.v2)


CASES = [  # (pair, nuv, params): the benchmark query, then smaller ones
    ("hardwave", 512, dict(k=4096, precision="fp32screen", polish=64, trace=1)),
    ("hardwave", 96, dict(k=256, precision="fp64", polish=16, trace=1)),
    ("wave", 64, dict(k=8, precision="exact32", trace=1)),
]


@pytest.mark.parametrize("order", ["large_then_small", "small_then_large"])
def test_noalloc_equals_min_distance(ctx, order):
    cases = CASES if order == "large_then_small" else CASES[::-1]
    for name, nuv, kw in cases:
        a, b = S.hard_wave_pair() if name == "hardwave" else S.wave_pair_small()
        r1, r2 = S.full_region(a, nuv, 1), S.full_region(b, nuv, 2)
        p = default_params(**kw)
        print(f"noalloc: {name} nuv {nuv} {kw}", flush=True)
        want, wpairs = ctx.minimum_distance(a, r1, b, r2, p)
        got, gpairs = ctx.minimum_distance_noalloc(a, r1, b, r2, p)
        same(got, want)
        np.testing.assert_array_equal(gpairs, wpairs)


def test_noalloc_candidate_overflow_path(ctx):
    """A small candidate capacity forces the repeated screen inside the scratch."""
    a, b = S.hard_wave_pair()
    r1, r2 = S.full_region(a, 256, 1), S.full_region(b, 256, 2)
    p = default_params(k=512, precision="fp64", cand_cap=1024, trace=1)
    want, wpairs = ctx.minimum_distance(a, r1, b, r2, p)
    got, gpairs = ctx.minimum_distance_noalloc(a, r1, b, r2, p)
    same(got, want)
    np.testing.assert_array_equal(gpairs, wpairs)


def test_noalloc_refuses_beyond_scratch(ctx):
    from jsi_cad_next_versions_amd.runtime import JsiError
    a, b = S.hard_wave_pair(scale=4.0)
    r1, r2 = S.full_region(a, 1024, 1), S.full_region(b, 1024, 2)
    with pytest.raises(JsiError) as e:
        ctx.minimum_distance_noalloc(a, r1, b, r2, default_params(k=4096, precision="fp32screen"))
    assert e.value.status == -2 and "scratch" in str(e.value)
    # the context still serves queries afterwards
    a2, b2 = S.hard_wave_pair()
    q1, q2 = S.full_region(a2, 64, 1), S.full_region(b2, 64, 2)
    p = default_params(k=64, precision="fp64", trace=1)
    same(ctx.minimum_distance_noalloc(a2, q1, b2, q2, p)[0], ctx.minimum_distance(a2, q1, b2, q2, p)[0])
