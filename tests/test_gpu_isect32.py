"""FP32SCREEN intersection: each round screens float records that cover every
cell's FP64 acceptance interval (and, in round 1, VERSION5's own float boxes
of the same grid), then re-tests the float candidates with the FP64 overlap
test.  The round-1 candidates must contain both the FP64 list and the EXACT32
(bit-exact VERSION5) list; every round's list after the re-test, the final
pairs and the traced curves must equal FP64 mode's (and the oracle's)."""
import numpy as np
import pytest

from jsi_cad_next_versions_amd import surfaces as S
from jsi_cad_next_versions_amd._abi import default_params

pytestmark = pytest.mark.gpu
CAP = 1 << 22


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


def keys(pairs):
    return set(zip(pairs["slot"].tolist(), pairs["i"].tolist(), pairs["j"].tolist()))


def regions(nuv):
    a, b = S.saddle_wave_pair()
    return a, b, S.full_region(a, nuv, 1), S.full_region(b, nuv, 2)


def test_round1_candidates_contain_fp64_and_version5_lists(ctx):
    a, b, r1, r2 = regions(512)
    one = dict(isect_max_rounds=1, trace=1)
    print("saddle / wave-plane nuv 512, round 1: fp32screen candidates", flush=True)
    rc, cand = ctx.intersection(a, r1, b, r2, default_params(precision="fp32screen", screen_out=1, **one), cap=CAP)
    print("round 1: fp64, exact32, fp32screen", flush=True)
    r64, p64 = ctx.intersection(a, r1, b, r2, default_params(precision="fp64", **one), cap=CAP)
    r32, p32 = ctx.intersection(a, r1, b, r2, default_params(precision="exact32", **one), cap=CAP)
    rs, ps = ctx.intersection(a, r1, b, r2, default_params(precision="fp32screen", **one), cap=CAP)
    kc = keys(cand)
    assert len(kc) == len(cand) == rc.npairs  # no duplicates, all returned
    assert keys(p64) <= kc and keys(p32) <= kc
    assert len(p64) > 1000 and len(kc) >= len(p64)
    # after the FP64 re-test: the FP64 list, in order
    np.testing.assert_array_equal(ps, p64)
    assert rs.rounds_list() == r64.rounds_list()


def test_rounds_equal_fp64_and_oracle(ctx, orc):
    for nuv in (128, 512):
        a, b, r1, r2 = regions(nuv)
        print(f"saddle / wave-plane nuv {nuv}: 6 rounds fp64 / fp32screen", flush=True)
        r64, p64 = ctx.intersection(a, r1, b, r2, default_params(precision="fp64", trace=1), cap=CAP)
        rs, ps = ctx.intersection(a, r1, b, r2, default_params(precision="fp32screen", trace=1), cap=CAP)
        assert rs.rounds == r64.rounds == 6
        assert rs.rounds_list() == r64.rounds_list()
        np.testing.assert_array_equal(ps, p64)
        if nuv == 128:
            print("nuv 128: oracle", flush=True)
            ro, po = orc.intersect(a, r1, b, r2, default_params(precision="fp64", trace=1), cap=CAP)
            assert rs.rounds_list() == ro.rounds_list()
            np.testing.assert_array_equal(ps[["slot", "i", "j"]], po[["slot", "i", "j"]])


def test_curves_identical_to_fp64(ctx):
    a, b, r1, r2 = regions(512)
    print("saddle / wave-plane nuv 512: curves fp64 / fp32screen", flush=True)
    g64 = ctx.intersection_curves(a, r1, b, r2, default_params(precision="fp64"))
    g32 = ctx.intersection_curves(a, r1, b, r2, default_params(precision="fp32screen"))
    assert g32[2] == g64[2] == 2
    np.testing.assert_array_equal(g32[1], g64[1])


def test_small_grids_and_disjoint(ctx, orc):
    """Coarse grids (single tile, partial tiles) and surfaces that do not meet."""
    for nuv in (5, 17, 40):
        a, b, r1, r2 = regions(nuv)
        r64, p64 = ctx.intersection(a, r1, b, r2, default_params(precision="fp64", trace=1), cap=CAP)
        rs, ps = ctx.intersection(a, r1, b, r2, default_params(precision="fp32screen", trace=1), cap=CAP)
        assert rs.rounds_list() == r64.rounds_list()
        np.testing.assert_array_equal(ps, p64)
    import shapes
    a, b = shapes.plane(0, 0, 0), shapes.plane(0, 0, 4.0)
    r1, r2 = S.full_region(a, 32, 1), S.full_region(b, 32, 2)
    rs, ps = ctx.intersection(a, r1, b, r2, default_params(precision="fp32screen", trace=1), cap=CAP)
    assert rs.npairs == 0 and len(ps) == 0 and rs.rounds == 1
