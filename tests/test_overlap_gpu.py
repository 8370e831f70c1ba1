"""Grid overlap estimator (updateOverlapArea: model.cpp:824-847, kernels.cuh:879-979) on the
GPU against the oracle's restatement of the reference kernel (orc_overlap_count, pinned by the
analytic lens area of two discs in tests/test_oracle.py): every point's count is equal, for
edge-sized, diameter-sized and 2x2 cell grids, a packing whose 2x2 grid puts more than 513
polygons around every point (the reference's limit on polygons visited per point), and a
context of several packings (counter laid out packing-major)."""
import numpy as np
import pytest

from oracle import oracle as O
from pycudapolygon_amd import capi
from pycudapolygon_amd.generate import make_packing, box_size_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("grid", ["edge", "diameter", "2x2"])
def test_counter_equals_oracle_c1(seed, grid):
    p = make_packing(64, 12, 0.80, seed=seed)
    B = {"edge": box_size_for(p["maxEdgeLength"]), "diameter": box_size_for(p["diameter"]), "2x2": 2}[grid]
    D = 96
    print(f"[overlap] seed={seed} grid={grid} B={B} D={D}", flush=True)
    s = capi.System([p], [B])
    s.cells()
    sums, cnt = s.overlap_count(D)
    so, co = O.overlap_count(p["positions"], p["startIndices"], B, D)
    assert np.array_equal(cnt[0], co) and int(sums[0]) == so
    assert so > 0
    s.close()


def test_counter_beyond_the_visiting_limit():
    """N=1024 octagons on the 2x2 grid: every point's 3x3 block holds all 1024 polygons, the
    reference tests only the first 513 it meets; the counts follow it."""
    p = make_packing(1024, 8, 0.85, seed=4)
    D = 40
    s = capi.System([p], [2])
    s.cells()
    sums, cnt = s.overlap_count(D)
    so, co = O.overlap_count(p["positions"], p["startIndices"], 2, D)
    assert np.array_equal(cnt[0], co) and int(sums[0]) == so
    # the limit matters here: with an edge-sized grid (all polygons near a point visited)
    # the same packing counts more
    B = box_size_for(p["maxEdgeLength"])
    s2 = capi.System([p], [B])
    s2.cells()
    full, _ = s2.overlap_count(D, counter=False)
    assert int(full[0]) == O.overlap_count(p["positions"], p["startIndices"], B, D)[0] > so
    s.close()
    s2.close()


def test_counter_several_packings():
    packs = [make_packing(N, n, phi, seed=k) for k, (N, n, phi) in enumerate(((64, 12, 0.8), (100, 9, 0.85),
                                                                             (40, 16, 0.75)))]
    Bs = [box_size_for(q["maxEdgeLength"]) for q in packs]
    Bs[1] = 3
    D = 64
    s = capi.System(packs, Bs)
    s.cells()
    sums, cnt = s.overlap_count(D)
    for k, q in enumerate(packs):
        so, co = O.overlap_count(q["positions"], q["startIndices"], Bs[k], D)
        assert np.array_equal(cnt[k], co) and int(sums[k]) == so, k
    s.close()
