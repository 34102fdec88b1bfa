"""The hotspot operator ABI (jsi_ascend_distance_topk_mask /
jsi_ascend_intersection_mask): the oracle against the reference's own
implementation (compiled from source, CPU), and the HIP library against the
oracle (GPU)."""
import os

import numpy as np
import pytest

from oracle import ASCEND_REF_SO, Oracle, RefHotspots


def cells(seed, n, spread=1.0, grid=False):
    rng = np.random.default_rng(seed)
    if grid:  # many exactly equal distances
        return rng.integers(0, 6, (n, 3)).astype(np.float32) * np.float32(0.25)
    return rng.uniform(0, spread, (n, 3)).astype(np.float32)


CASES = [(1, 300, 411, 8, False), (2, 1000, 777, 50, False), (3, 64, 81, 500, True), (4, 5, 3, 100, False),
         (5, 2500, 2000, 4096, False), (6, 1, 1, 1, False)]


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.mark.skipif(not os.path.exists(ASCEND_REF_SO), reason="reference operator library not built")
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"s{c[0]}")
def test_oracle_matches_reference_operator_library(orc, case):
    seed, n1, n2, k, grid = case
    c1, c2 = cells(seed, n1, grid=grid), cells(seed + 100, n2, grid=grid)
    ref = RefHotspots()
    a = orc.distance_topk_mask(c1, c2, k)
    b = ref.distance_topk_mask(c1, c2, k)
    assert np.float32(a[0]) == np.float32(b[0])
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])
    r1, r2 = cells(seed + 7, n1, 0.05), cells(seed + 8, n2, 0.05)
    for x, y in zip(orc.intersection_mask(c1, r1, c2, r2), ref.intersection_mask(c1, r1, c2, r2)):
        np.testing.assert_array_equal(x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"s{c[0]}")
def test_gpu_operator_abi_matches_oracle(orc, case):
    from jsi_cad_next_versions_amd import runtime
    seed, n1, n2, k, grid = case
    c1, c2 = cells(seed, n1, grid=grid), cells(seed + 100, n2, grid=grid)
    a = runtime.distance_topk_mask(c1, c2, k)
    b = orc.distance_topk_mask(c1, c2, k)
    assert np.float32(a[0]) == np.float32(b[0])
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])
    r1, r2 = cells(seed + 7, n1, 0.05), cells(seed + 8, n2, 0.05)
    for x, y in zip(runtime.intersection_mask(c1, r1, c2, r2), orc.intersection_mask(c1, r1, c2, r2)):
        np.testing.assert_array_equal(x, y)
