"""Evaluation kernel variants against the oracle: degree 3 in both directions
(the fixed-degree basis recurrence) and other degrees (the general one), knot
vectors that fit the on-chip copy (at most 128 knots per direction) and longer
ones read from device memory, single and multi-tile grids.  Grid values and
every round's pair lists must be bit-identical with the oracle."""
import numpy as np
import pytest

from jsi_cad_next_versions_amd import surfaces as S
from jsi_cad_next_versions_amd._abi import default_params

pytestmark = pytest.mark.gpu


def wavy(n, deg, seed, lift=0.0, rational=True, periods=3.0):
    """n x n net of degree `deg`, a seeded sinusoidal height field over [0,1]^2."""
    k = S.clamped_uniform_knots(n, deg)
    g = np.arange(n) / (n - 1)
    X, Y = np.meshgrid(g, g)
    rng = S.Rng(seed)
    Z = 0.05 * np.sin(2 * np.pi * periods * X) * np.sin(2 * np.pi * periods * Y)
    Z = Z + rng.normal(n * n, 0.002).reshape(n, n) + lift
    w = rng.uniform(n * n, 0.8, 1.2).reshape(n, n) if rational else None
    return S.make_surface(deg, deg, k, k, np.stack([X, Y, Z], -1), w, f"wavy{n}_d{deg}_s{seed}")


# (net size, degree): knots = n + deg + 1
CASES = [(160, 3), (20, 2), (150, 2), (12, 5)]


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


@pytest.mark.parametrize("n,deg", CASES)
@pytest.mark.parametrize("nuv", [13, 40, 77])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_eval_grid_matches_oracle(ctx, orc, n, deg, nuv, dtype):
    s = wavy(n, deg, seed=n + deg)
    us, vs = dtype(1.0 / nuv), dtype(1.0 / nuv)
    a = ctx.eval_grid(s, us, vs, dtype(0.0), dtype(0.0), nuv, dtype=dtype)
    b = orc.eval_grid(s, us, vs, dtype(0.0), dtype(0.0), nuv, dtype=dtype)
    assert a.dtype == b.dtype and a.shape == b.shape
    np.testing.assert_array_equal(a.view(np.uint32 if dtype == np.float32 else np.uint64),
                                  b.view(np.uint32 if dtype == np.float32 else np.uint64))


@pytest.mark.parametrize("n,deg", [(160, 3), (150, 2)])
@pytest.mark.parametrize("prec", ["exact32", "fp64"])
def test_min_distance_long_knot_vectors_match_oracle(ctx, orc, n, deg, prec):
    a, b = wavy(n, deg, seed=7), wavy(n, deg, seed=8, lift=0.02)
    r1, r2 = S.full_region(a, 48, 1), S.full_region(b, 48, 2)
    p = default_params(k=256, precision=prec, trace=1)
    res, pairs = ctx.minimum_distance(a, r1, b, r2, p)
    ores, opairs = orc.min_distance(a, r1, b, r2, p)
    assert res.rounds_list() == ores.rounds_list()
    assert res.distance == ores.distance
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("n,deg", [(160, 3), (150, 2)])
def test_intersection_long_knot_vectors_matches_oracle(ctx, orc, n, deg):
    a, b = wavy(n, deg, seed=9), wavy(40, 3, seed=10, lift=0.04, periods=1.0)
    r1, r2 = S.full_region(a, 24, 1), S.full_region(b, 24, 2)
    p = default_params(precision="fp64", trace=1)
    res, pairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
    ores, opairs = orc.intersect(a, r1, b, r2, p, cap=1 << 20)
    assert res.rounds_list() == ores.rounds_list()
    assert res.rounds_list()[0]["npairs"] > 0
    np.testing.assert_array_equal(pairs, opairs)
