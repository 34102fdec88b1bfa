"""Intersection refinement units of nuv <= 8 (the fused evaluation, cell AABB
and tile-box pass that holds several units per wavefront) against the
oracle and against the grid-point path it replaces (JSI_ISECT_CELLS=0):
every round's ordered pair list, the final pairs and the final refinement
nuv must be identical.  The cases cover every unit grid from 2 x 2 to 8 x 8
points (lane groups of 4 to 64, 1 to 8 units per wavefront), the compute_K
padding on and off, the three precisions, degrees 2-5 and 11, rational and
polynomial surfaces, and knot vectors longer than the on-chip copy."""
import time

import numpy as np
import pytest

import golden_data as G
from jsi_cad_next_versions_amd import surfaces as S
from jsi_cad_next_versions_amd._abi import default_params

pytestmark = pytest.mark.gpu

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


def wavy(n, deg, seed, lift=0.0, rational=True, periods=3.0, tilt=0.0):
    k = S.clamped_uniform_knots(n, deg)
    g = np.arange(n) / (n - 1)
    X, Y = np.meshgrid(g, g)
    rng = S.Rng(seed)
    Z = 0.05 * np.sin(2 * np.pi * periods * X) * np.sin(2 * np.pi * periods * Y)
    Z = Z + rng.normal(n * n, 0.002).reshape(n, n) + lift + tilt * (X - 0.5)
    w = rng.uniform(n * n, 0.8, 1.2).reshape(n, n) if rational else None
    return S.make_surface(deg, deg, k, k, np.stack([X, Y, Z], -1), w, f"wavy{n}_d{deg}_s{seed}")


def run_three(ctx, orc, a, b, nuv, monkeypatch, **kw):
    """The fused path (slots screened directly, and listed after the unit-box
    test as rounds with many slots are; the emit pass reading the count
    pass's row masks, and repeating the tests), the grid-point path and the
    oracle on one query."""
    r1, r2 = S.full_region(a, nuv, 1), S.full_region(b, nuv, 2)
    p = default_params(trace=1, **kw)
    res, pairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
    monkeypatch.setenv("JSI_SLOT_LIST_MIN", "1")
    lres, lpairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
    monkeypatch.delenv("JSI_SLOT_LIST_MIN")
    monkeypatch.setenv("JSI_SLOT_MASKS", "0")
    mres, mpairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
    monkeypatch.delenv("JSI_SLOT_MASKS")
    monkeypatch.setenv("JSI_ISECT_CELLS", "0")
    pres, ppairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
    monkeypatch.delenv("JSI_ISECT_CELLS")
    ores, opairs = orc.intersect(a, r1, b, r2, p, cap=1 << 20)
    for other, opl in ((lres, lpairs), (mres, mpairs), (pres, ppairs), (ores, opairs)):
        assert res.rounds_list() == other.rounds_list()
        assert res.isect_refine_nuv == other.isect_refine_nuv
        np.testing.assert_array_equal(pairs, opl)
    return res


# (refine nuv, min nuv, halve above, rounds) -> the unit grids of the rounds
# after the first: 8 then 4; 8; 7; 6; 5; 3; 2; 16 (the grid-point path, so
# rounds of both paths meet in one query)
BIG = 1 << 30
GRIDS = [(16, 4, 64, 6), (16, 8, 64, 4), (7, 7, BIG, 3), (6, 6, BIG, 3), (5, 5, BIG, 3), (6, 3, 64, 6),
         (5, 2, 64, 6), (16, 16, BIG, 3)]


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: f"r{g[0]}m{g[1]}")
@pytest.mark.parametrize("prec", ["exact32", "fp64", "fp32screen"])
def test_saddle_unit_grids(ctx, orc, monkeypatch, grid, prec):
    refine, min_nuv, halve, rounds = grid
    a, b = S.saddle_wave_pair()
    res = run_three(ctx, orc, a, b, 48, monkeypatch, precision=prec, isect_refine_nuv=refine,
                    isect_min_nuv=min_nuv, isect_halve_above=halve, isect_max_rounds=rounds)
    assert res.rounds == rounds and res.npairs > 0


@pytest.mark.parametrize("use_k", [0, 1])
@pytest.mark.parametrize("rational", [True, False])
def test_padding_and_rational(ctx, orc, monkeypatch, use_k, rational):
    a = wavy(24, 3, seed=21, rational=rational)
    b = wavy(20, 3, seed=22, rational=rational, tilt=0.2, periods=1.0)
    res = run_three(ctx, orc, a, b, 40, monkeypatch, precision="fp64", use_compute_k=use_k)
    assert res.rounds >= 3 and res.npairs > 0


@pytest.mark.parametrize("seed", [3, 8, 11, 19])
def test_uneven_degrees(ctx, orc, monkeypatch, seed):
    """random_pair's intersection pairs: degrees 2-5 per direction, nets 8-96."""
    a, b, _ = S.random_pair(seed, "isect")
    res = run_three(ctx, orc, a, b, 32, monkeypatch, precision="fp64")
    assert res.npairs > 0


@pytest.mark.parametrize("n,deg", [(160, 3), (150, 2)])
def test_long_knot_vectors(ctx, orc, monkeypatch, n, deg):
    a, b = wavy(n, deg, seed=9), wavy(40, 3, seed=10, lift=0.04, periods=1.0)
    res = run_three(ctx, orc, a, b, 24, monkeypatch, precision="fp64")
    assert res.npairs > 0


@pytest.mark.parametrize("case", [1, 2])
@pytest.mark.parametrize("prec", ["exact32", "fp64"])
def test_degree11_cad_faces(ctx, orc, monkeypatch, case, prec):
    n1, n2 = f"inter_case{case}_obj1", f"inter_case{case}_obj2"
    if n1 not in SURF:
        pytest.skip("the degree-11 faces need the reference build")
    run_three(ctx, orc, SURF[n1], SURF[n2], 16, monkeypatch, precision=prec, isect_max_rounds=4)


def test_large_round_matches_grid_point_path(ctx, monkeypatch):
    """The benchmark intersection (saddle / wave plane, nuv 512, six rounds,
    up to 114k units per side in a round): beyond what the oracle finishes in
    seconds, so the check is the grid-point path's lists."""
    a, b = S.saddle_wave_pair()
    r1, r2 = S.full_region(a, 512, 1), S.full_region(b, 512, 2)
    for prec in ("fp64", "fp32screen"):
        p = default_params(precision=prec, trace=1)
        res, pairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
        monkeypatch.setenv("JSI_ISECT_CELLS", "0")
        pres, ppairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
        monkeypatch.delenv("JSI_ISECT_CELLS")
        assert res.rounds_list() == pres.rounds_list()
        assert [r["npairs"] for r in res.rounds_list()] == [13193, 51328, 64566, 81276, 103552, 131148]
        np.testing.assert_array_equal(pairs, ppairs)


@pytest.mark.parametrize("nuv", [2, 5, 9])
@pytest.mark.parametrize("prec", ["exact32", "fp64"])
def test_small_first_round(ctx, orc, monkeypatch, nuv, prec):
    """Initial grids of at most 9 x 9 points: the first round's screen takes
    the hit-mask form too (one tile per side, compute_K of the two regions
    from the grid-point evaluation)."""
    a, b = S.saddle_wave_pair()
    res = run_three(ctx, orc, a, b, nuv, monkeypatch, precision=prec)
    assert res.round[0].npairs > 0


@pytest.mark.parametrize("nuv1,nuv2", [(40, 24), (9, 33)])
@pytest.mark.parametrize("prec", ["exact32", "fp64"])
def test_unequal_first_grids(ctx, orc, monkeypatch, nuv1, nuv2, prec):
    """Different first grids on the two sides: each side's child regions
    split its own parent's grid (make_intersect_refine_task uses the parent
    AABB result's own nuv and width)."""
    a, b = S.saddle_wave_pair()
    r1, r2 = S.full_region(a, nuv1, 1), S.full_region(b, nuv2, 2)
    p = default_params(trace=1, precision=prec)
    res, pairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
    monkeypatch.setenv("JSI_SLOT_LIST_MIN", "1")
    lres, lpairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
    monkeypatch.delenv("JSI_SLOT_LIST_MIN")
    monkeypatch.setenv("JSI_ISECT_CELLS", "0")
    pres, ppairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 20)
    monkeypatch.delenv("JSI_ISECT_CELLS")
    ores, opairs = orc.intersect(a, r1, b, r2, p, cap=1 << 20)
    for other, opl in ((lres, lpairs), (pres, ppairs), (ores, opairs)):
        assert res.rounds_list() == other.rounds_list()
        np.testing.assert_array_equal(pairs, opl)
    assert res.rounds == 6 and res.round[1].units1 != res.round[1].units2


def test_dense_first_round_ten_million_slots(ctx, monkeypatch):
    """An uneven-workload pair whose 64 x 64 first grids overlap on 10.4M cell
    pairs: the second round has 10.4M slots of 49 x 49 cells, almost all
    without a hit (the per-slot screen keeps 12 bytes per slot)."""
    kind, a, b, nuv = S.uneven_workload(37, 0)[36]
    assert kind == "isect" and nuv == 64
    r1, r2 = S.full_region(a, nuv, 1), S.full_region(b, nuv, 2)
    p = default_params(trace=1, precision="fp64")
    res, pairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 16)
    monkeypatch.setenv("JSI_SLOT_LIST_MIN", str(1 << 40))  # every round screened directly
    dres, dpairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 16)
    monkeypatch.delenv("JSI_SLOT_LIST_MIN")
    monkeypatch.setenv("JSI_ISECT_CELLS", "0")
    pres, ppairs = ctx.intersection(a, r1, b, r2, p, cap=1 << 16)
    monkeypatch.delenv("JSI_ISECT_CELLS")
    # FP32SCREEN: float records and float unit boxes pick the candidates of
    # the 10.4M-slot round, the FP64 test decides them
    sres, spairs = ctx.intersection(a, r1, b, r2, default_params(trace=1, precision="fp32screen"), cap=1 << 16)
    for other, opl in ((dres, dpairs), (pres, ppairs), (sres, spairs)):
        assert res.rounds_list() == other.rounds_list()
        np.testing.assert_array_equal(pairs, opl)
    assert [r["npairs"] for r in res.rounds_list()] == [10389328, 15836, 14360, 16715, 20375, 25402]


def test_server_intersections_only(ctx):
    """Every intersection query of the first 400 uneven-workload pairs (nuv
    64-512, round-1 lists up to 10M pairs) queued at once on 8 server
    workers: all succeed and equal the same query run alone."""
    from jsi_cad_next_versions_amd._abi import QUERY_INTERSECT
    from jsi_cad_next_versions_amd.runtime import Server
    meta = [(a, b, nuv) for kind, a, b, nuv in S.uneven_workload(400, 0) if kind == "isect"]
    p = default_params(precision="fp64", trace=1)
    qs = [ctx.query(QUERY_INTERSECT, a, S.full_region(a, nuv, 1), b, S.full_region(b, nuv, 2), p)
          for a, b, nuv in meta]
    print(f"server: {len(qs)} intersection queries on 8 workers", flush=True)
    with Server(ctx, slots=8) as srv:
        tickets = []
        for q in qs:
            tk = srv.submit(q)
            while tk is None:  # admission holds it back
                time.sleep(1e-3)
                tk = srv.submit(q)
            tickets.append(tk)
        results = [srv.poll(tk) for tk in tickets]
    print("server: solo runs", flush=True)
    for r, (a, b, nuv) in zip(results, meta):
        s, _ = ctx.intersection(a, S.full_region(a, nuv, 1), b, S.full_region(b, nuv, 2), p)
        assert r.rounds_list() == s.rounds_list()
