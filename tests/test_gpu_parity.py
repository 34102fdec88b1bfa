"""Parity of the HIP library (through the C ABI) with the oracle and with the
reference's golden outputs.  Integer results (pair indices, counts, unit
maps) must be identical; floats bit-identical (EXACT32 against VERSION5,
FP64 against the FP64 oracle, both built without fused multiply-add)."""
import numpy as np
import pytest

import golden_data as G
from jsi_cad_next_versions_amd import surfaces as S
from jsi_cad_next_versions_amd._abi import default_params

pytestmark = pytest.mark.gpu

GOLD, ARR = G.load()
SURF = G.surfaces()


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


def regions(s1, s2, nuv1, nuv2=None):
    return S.full_region(s1, nuv1, 1), S.full_region(s2, nuv2 or nuv1, 2)


def same_rounds(a, b):
    ra, rb = a.rounds_list(), b.rounds_list()
    assert len(ra) == len(rb)
    for x, y in zip(ra, rb):
        assert x == y, (x, y)


@pytest.mark.parametrize("e", GOLD["eval"], ids=lambda e: e["key"])
def test_eval_grid_f32_matches_reference(ctx, e):
    if e["surface"] not in SURF:
        pytest.skip("surface needs the reference build")
    f = G.bits_to_f32
    out = ctx.eval_grid(SURF[e["surface"]], f(e["ustep"]), f(e["vstep"]), f(e["ustart"]), f(e["vstart"]), e["nuv"])
    np.testing.assert_array_equal(out.view(np.uint32), ARR[e["key"]].view(np.uint32))


@pytest.mark.parametrize("name", ["hardwave_A", "wave16_B", "saddle", "inter_case3_obj2"])
@pytest.mark.parametrize("nuv", [9, 64])
def test_eval_grid_f64_matches_oracle(ctx, orc, name, nuv):
    s = SURF[name]
    us = (s.uend() - s.ubegin()) / nuv
    vs = (s.vend() - s.vbegin()) / nuv
    a = ctx.eval_grid(s, us, vs, s.ubegin(), s.vbegin(), nuv, dtype=np.float64)
    b = orc.eval_grid(s, us, vs, s.ubegin(), s.vbegin(), nuv, dtype=np.float64)
    np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("e", GOLD["min_distance"], ids=lambda e: f"{e['a']}-{e['nuv']}")
def test_min_distance_exact32_matches_reference(ctx, orc, e):
    s1, s2 = SURF[e["a"]], SURF[e["b"]]
    r1, r2 = regions(s1, s2, e["nuv"])
    p = default_params(trace=1)
    res, pairs = ctx.minimum_distance(s1, r1, s2, r2, p)
    assert np.float32(res.distance).view(np.uint32) == e["distance"]
    ores, opairs = orc.min_distance(s1, r1, s2, r2, p)
    same_rounds(res, ores)
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("prec", ["exact32", "fp64", "fp32screen"])
@pytest.mark.parametrize("k", [1, 64, 4096])
def test_min_distance_hardwave_matches_oracle(ctx, orc, prec, k):
    """FP32SCREEN must reproduce the FP64 oracle exactly (its FP32 candidate
    set is a superset, re-ranked in FP64)."""
    a, b = S.hard_wave_pair()
    r1, r2 = regions(a, b, 96)
    p = default_params(k=k, precision=prec, trace=1)
    res, pairs = ctx.minimum_distance(a, r1, b, r2, p)
    ores, opairs = orc.min_distance(a, r1, b, r2, p)
    same_rounds(res, ores)
    assert res.distance == ores.distance
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("prec", ["exact32", "fp64", "fp32screen"])
@pytest.mark.parametrize("k", [4032, 4095, 4096])
def test_min_distance_k_near_all_pairs(ctx, orc, prec, k):
    """nuv=9: 64 x 64 = 4096 cell pairs in round 1, so rank k lies at or near
    the largest seed key: the seed histogram's last bin (keys above the
    sampled range) must give no threshold rather than a cut below rank k."""
    a, b = S.hard_wave_pair()
    r1, r2 = regions(a, b, 9)
    p = default_params(k=k, precision=prec, trace=1)
    res, pairs = ctx.minimum_distance(a, r1, b, r2, p)
    ores, opairs = orc.min_distance(a, r1, b, r2, p)
    same_rounds(res, ores)
    assert res.distance == ores.distance
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("prec", ["fp64", "fp32screen"])
@pytest.mark.parametrize("nref", [3, 4, 5, 9, 13, 16, 17])
def test_min_distance_refine_grids_match_oracle(ctx, orc, prec, nref):
    """Refinement units of (nref - 1)^2 cells: the refinement filter splits
    them into full 16-cell blocks (matrix-core products dealt to the four
    wavefronts) and partial ones (vector test); nref = 3 and 4 have no full
    block, 5 one, 9 four, 13 nine, 16 fourteen plus one cell, 17 sixteen."""
    a, b = S.hard_wave_pair()
    r1, r2 = regions(a, b, 64)
    p = default_params(k=512, precision=prec, trace=1)
    p.dist_refine_nuv = nref
    p.dist_max_rounds = 3
    res, pairs = ctx.minimum_distance(a, r1, b, r2, p)
    ores, opairs = orc.min_distance(a, r1, b, r2, p)
    same_rounds(res, ores)
    assert res.distance == ores.distance
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("prec", ["exact32", "fp32screen"])
@pytest.mark.parametrize("cull", [0, 1])
def test_culling_leaves_results_unchanged(ctx, orc, cull, prec):
    a, b = S.wave_pair_small()
    r1, r2 = regions(a, b, 80, 72)
    p = default_params(k=300, cull=cull, precision=prec, trace=1)
    res, pairs = ctx.minimum_distance(a, r1, b, r2, p)
    ores, opairs = orc.min_distance(a, r1, b, r2, p)
    same_rounds(res, ores)
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("prec", ["exact32", "fp64", "fp32screen"])
def test_candidate_overflow_path_matches_oracle(ctx, orc, prec):
    """A candidate capacity below the screen's candidate count forces the
    overflow path (k-th bound of the captured candidates, repeated screen)."""
    a, b = S.hard_wave_pair()
    r1, r2 = regions(a, b, 128)
    k = 512
    p = default_params(k=k, precision=prec, trace=1, cand_cap=2 * k)
    res, pairs = ctx.minimum_distance(a, r1, b, r2, p)
    ores, opairs = orc.min_distance(a, r1, b, r2, default_params(k=k, precision=prec, trace=1))
    same_rounds(res, ores)
    assert res.distance == ores.distance
    np.testing.assert_array_equal(pairs, opairs)


def dense_pair(nu, nv, deg=3):
    """Hard-wave shaped surfaces on a nu x nv control net: long knot vectors
    (read from HBM by the evaluation when over 128 knots) and tiles that cover
    many knot spans (control rows processed in several chunks)."""
    ku, kv = S.clamped_uniform_knots(nu, deg), S.clamped_uniform_knots(nv, deg)
    X, Y = np.meshgrid(np.arange(nu) / (nu - 1), np.arange(nv) / (nv - 1))
    za = 0.05 * np.sin(16 * np.pi * X) * np.sin(16 * np.pi * Y)
    zb = 0.05 * np.sin(16 * np.pi * X + np.pi / 7) * np.sin(16 * np.pi * Y) + 0.06
    wa = S.Rng(11).uniform(nu * nv, 0.8, 1.2).reshape(nv, nu)
    wb = S.Rng(12).uniform(nu * nv, 0.8, 1.2).reshape(nv, nu)
    a = S.make_surface(deg, deg, ku, kv, np.stack([X, Y, za], -1), wa, f"dense{nu}x{nv}_A")
    b = S.make_surface(deg, deg, ku, kv, np.stack([X, Y, zb], -1), wb, f"dense{nu}x{nv}_B")
    return a, b


@pytest.mark.parametrize("prec", ["exact32", "fp64"])
@pytest.mark.parametrize("net", [(200, 200), (150, 20), (20, 150)])
def test_min_distance_dense_nets_match_oracle(ctx, orc, prec, net):
    a, b = dense_pair(*net)
    r1, r2 = regions(a, b, 64, 48)
    p = default_params(k=64, precision=prec, trace=1)
    res, pairs = ctx.minimum_distance(a, r1, b, r2, p)
    ores, opairs = orc.min_distance(a, r1, b, r2, p)
    same_rounds(res, ores)
    assert res.distance == ores.distance
    np.testing.assert_array_equal(pairs, opairs)


def test_min_distance_more_k_than_pairs(ctx, orc):
    """k larger than the number of cell pairs: every pair is kept, sorted."""
    a, b = S.wave_pair_small()
    r1, r2 = regions(a, b, 4, 3)  # 9 x 4 = 36 pairs
    p = default_params(k=100, dist_max_rounds=1, trace=1)
    res, pairs = ctx.minimum_distance(a, r1, b, r2, p)
    ores, opairs = orc.min_distance(a, r1, b, r2, p)
    assert res.npairs == 36
    np.testing.assert_array_equal(pairs, opairs)


def test_min_distance_single_cell(ctx, orc):
    a, b = S.wave_pair_small()
    r1, r2 = regions(a, b, 2)
    p = default_params(trace=1)
    res, pairs = ctx.minimum_distance(a, r1, b, r2, p)
    ores, opairs = orc.min_distance(a, r1, b, r2, p)
    same_rounds(res, ores)
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("e", GOLD["intersection"], ids=lambda e: f"{e['a']}-{e['nuv']}-{e['use_k']}")
def test_intersection_exact32_matches_reference(ctx, orc, e):
    s1, s2 = SURF[e["a"]], SURF[e["b"]]
    r1, r2 = regions(s1, s2, e["nuv"])
    p = default_params(use_compute_k=int(e["use_k"]), isect_refine_nuv=e["refine_nuv_in"], trace=1)
    res, pairs = ctx.intersection(s1, r1, s2, r2, p, cap=1 << 20)
    assert [r["npairs"] for r in res.rounds_list()] == e["counts"]
    ores, opairs = orc.intersect(s1, r1, s2, r2, p, cap=1 << 20)
    same_rounds(res, ores)
    assert res.isect_refine_nuv == ores.isect_refine_nuv
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("cull", [0, 1])
def test_intersection_fp64_matches_oracle(ctx, orc, cull):
    a, b = S.saddle_wave_pair()
    r1, r2 = regions(a, b, 64)
    p = default_params(precision="fp64", cull=cull, trace=1)
    res, pairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
    ores, opairs = orc.intersect(a, r1, b, r2, p, cap=1 << 20)
    same_rounds(res, ores)
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("seed", [1, 13])
@pytest.mark.parametrize("prec", ["exact32", "fp64"])
def test_intersection_dense_overlap_matches_oracle(ctx, orc, seed, prec):
    """Wavy random pairs whose second-derivative padding makes ~15% of the
    round-1 cell pairs overlap (2.3M pairs, ~580 per row): the long rows take
    the segmented radix sort; every round's ordered list must equal the
    oracle's."""
    a, b, nuv = S.random_pair(seed, "isect")
    assert nuv == 64
    r1, r2 = regions(a, b, nuv)
    p = default_params(precision=prec, trace=1)
    res, pairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 16)
    ores, opairs = orc.intersect(a, r1, b, r2, p, cap=1 << 16)
    assert res.round[0].npairs > 16 * (nuv - 1) ** 2
    same_rounds(res, ores)
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("k", [64, 8192])
def test_min_distance_maximum_grid_fp32screen_equals_fp64(ctx, k):
    """nuv = 1025 (the largest grid: 1024 x 1024 cells per side, cell indices
    of 20 bits) and k up to the largest k: beyond what the oracle finishes in
    seconds, so the check is the property FP32SCREEN is built on, identical
    pair lists and distance to the FP64 path, plus the order of the list."""
    a, b = S.hard_wave_pair()
    r1, r2 = regions(a, b, 1025)
    g64, p64 = ctx.minimum_distance(a, r1, b, r2, default_params(k=k, precision="fp64", trace=1))
    g32, p32 = ctx.minimum_distance(a, r1, b, r2, default_params(k=k, precision="fp32screen", trace=1))
    same_rounds(g64, g32)
    assert g64.distance == g32.distance
    np.testing.assert_array_equal(p64, p32)
    assert len(p64) == k
    key = list(zip(p64["d2"], p64["slot"], p64["i"], p64["j"]))
    assert key == sorted(key)  # total order (d2, slot, i, j)
    assert p64["i"].max() < 1024 * 1024 and p64["j"].max() < 1024 * 1024


def test_invalid_arguments_raise(ctx):
    from jsi_cad_next_versions_amd.runtime import JsiError
    a, b = S.wave_pair_small()
    r1, r2 = regions(a, b, 1)
    with pytest.raises(JsiError):
        ctx.minimum_distance(a, r1, b, r2)
    r1, r2 = regions(a, b, 8)
    with pytest.raises(JsiError):
        ctx.minimum_distance(a, r1, b, r2, default_params(k=0))
    with pytest.raises(JsiError):
        ctx.minimum_distance(a, r1, b, r2, default_params(k=8193))
    r1, r2 = regions(a, b, 1026)  # one past the largest grid
    with pytest.raises(JsiError):
        ctx.minimum_distance(a, r1, b, r2)
