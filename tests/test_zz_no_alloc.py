"""pcp_overlap_count_noalloc (the estimator with its partial sums in the library's fixed
device array, for callers that cannot allocate) against pcp_overlap_count, through the C ABI:
equal counters and sums on two calls in a row, the second on a smaller input (fewer cells
and points, so stale partials from the first call would show), and a clear error beyond its
limit on cells."""
import ctypes

import numpy as np
import pytest

from pycudapolygon_amd import capi
from pycudapolygon_amd.generate import make_packing, box_size_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


def _both(s, D):
    ref_sum, ref_cnt = s.overlap_count(D)
    cnt = capi.DeviceArray((s.P * D * D,), np.int32)
    cnt.upload(np.full(s.P * D * D, -7, np.int32))   # every entry must be written
    sums = np.zeros(s.P, np.int64)
    s.ctx.call("pcp_overlap_count_noalloc", s.pos.ptr, int(D), cnt.ptr, sums.ctypes.data_as(ctypes.c_void_p))
    return ref_sum, ref_cnt, sums, cnt.get().reshape(s.P, D * D)


def test_noalloc_equals_overlap_count_twice():
    p1 = make_packing(256, 12, 0.85, seed=3)
    p2 = make_packing(40, 9, 0.80, seed=5)
    big = capi.System([p1, p1], [box_size_for(p1["maxEdgeLength"]), 2])
    big.cells()
    print(f"[noalloc] call 1: P=2 cells={big.boxSizes.tolist()} D=120", flush=True)
    rs, rc, ns, nc = _both(big, 120)
    assert np.array_equal(rs, ns) and np.array_equal(rc, nc) and rs[0] > 0
    small = capi.System([p2], [3])
    small.cells()
    print("[noalloc] call 2: P=1 cells=3 D=50", flush=True)
    rs, rc, ns, nc = _both(small, 50)
    assert np.array_equal(rs, ns) and np.array_equal(rc, nc) and rs[0] > 0
    big.close()
    small.close()


def test_noalloc_rejects_more_cells_than_its_array():
    p = make_packing(32, 8, 0.7, seed=1)
    s = capi.System([p], [800])   # 640,000 cells > PCP_OVERLAP_NOALLOC_MAX_CELLS = 524,288
    sums = np.zeros(1, np.int64)
    with pytest.raises(capi.PcpError, match="at most 524288 cells"):
        s.ctx.call("pcp_overlap_count_noalloc", s.pos.ptr, 16, None, sums.ctypes.data_as(ctypes.c_void_p))
    s.close()
