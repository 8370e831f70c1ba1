"""The odd-crossing window, on the CPU oracle.

Packing 41 of the 256-packing uneven batch (seed 3; N=65 x 16) with squared lens areas:
FIRE stalls (dt collapses through energy rollbacks) with max|F| ~ 5e-8.  At the stalled
state vertex 2 of polygon 22 lies within rounding of edge 13 of polygon 58; moving every
vertex by 1e-16 along the force (scaled to max 1) makes the reference's segment test
(kernels.cuh:760-790) count the crossing of one of the vertex's two edges with that edge and
not of the other, so the pair has 3 crossings, a record is paired with a wrong exit and the
energy jumps by 3.7e-9; by 1e-15 both crossings are gone (one lens of the pair left) and the
energy is 2.2e-12 higher (two lenses merged into one under squared lens areas).  The test
documents that behaviour and the odd-count check that flags it (on the GPU:
test_odd_crossing_records_flag_the_window in tests/test_fire_oracle_gpu.py)."""
import numpy as np
import pytest

from oracle import oracle as O
from pycudapolygon_amd.generate import make_batch, box_size_for

SQUARED = 0x200


@pytest.fixture(scope="module")
def stalled():
    p = make_batch(256, 3, min_center_frac=0.5)[41]
    B = box_size_for(p["diameter"])
    sysO = O.System(p["startIndices"], B, O.NORMAL | SQUARED, p["targetEdgeV"], p["targetAreas"], solver="lanes")
    x, f, st, steps = sysO.fire_minimize(p["positions"], fTol=1e-10, dt=1e-3, maxSteps=5000, dtMax=100.0)
    return p, B, sysO, x, f, st, steps


def _records(p, B, x):
    return O.outersections(p["startIndices"], O.neighbors(x, p["startIndices"], B))


def test_fire_stalls_with_even_crossings(stalled):
    p, B, sysO, x, f, st, steps = stalled
    assert steps == 5000 and np.abs(f).max() > 1e-8 and st.dt < 1e-5
    rec = _records(p, B, x)
    assert rec["numIntersections"] == 294
    assert O.odd_crossing_records(rec["intersections"]) == (0, [])


@pytest.mark.parametrize("h,records,odd,jump", [(1e-16, 293, (3, [(22, 58)]), (3.0e-9, 4.5e-9)),
                                                (1e-15, 292, (0, []), (1.5e-12, 3.0e-12))])
def test_a_step_along_the_force_crosses_the_window(stalled, h, records, odd, jump):
    p, B, sysO, x, f, st, steps = stalled
    E0, F0, _ = sysO.force_eval(x)
    y = x + h * F0 / np.abs(F0).max()
    y -= np.floor(y)
    E1, _, _ = sysO.force_eval(y)
    rec = _records(p, B, y)
    assert rec["numIntersections"] == records
    assert O.odd_crossing_records(rec["intersections"]) == odd
    assert jump[0] < E1 - E0 < jump[1]


def test_odd_check_on_hand_built_keys():
    """keys [sj|si|i|j]: pair (1, 2) with three crossings (two from 1, one from 2), pair
    (0, 3) with two, polygon 4 with itself once"""
    def key(si, sj, i, j):
        return (sj << 48) | (si << 32) | (i << 16) | j
    keys = [key(1, 2, 0, 1), key(1, 2, 3, 1), key(2, 1, 5, 2), key(0, 3, 1, 1), key(3, 0, 2, 2), key(4, 4, 0, 5)]
    assert O.odd_crossing_records(keys) == (3, [(1, 2)])


# ---- the side-consistent decision (ORC_CONSISTENT / PCP_CONSISTENT) -----------------------

def _spike_configuration(rng):
    """A square B (random turn) and a kite A whose tip v lies within a few ulps of B's first
    edge, v's two edges leaving on the same side of that edge's line (both cross it or neither,
    geometrically) or across it"""
    th = rng.uniform(0, 2 * np.pi)
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    B = (np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]]) * 0.1) @ R.T + 0.5
    b0, b1 = B[0], B[1]
    v = b0 + rng.uniform(0.2, 0.8) * (b1 - b0)
    v = v + rng.integers(-3, 4, 2) * np.spacing(v)
    e = (b1 - b0) / np.linalg.norm(b1 - b0)
    n = np.array([-e[1], e[0]])
    side = rng.choice([-1.0, 1.0])
    vm, vp, w = v - 0.02 * e + side * 0.03 * n, v + 0.02 * e + side * 0.03 * n, v + side * 0.06 * n
    A = np.array([v, vp, w, vm]) if side < 0 else np.array([v, vm, w, vp])
    return np.concatenate([A.reshape(-1), B.reshape(-1)]), np.array([0, 4, 8], np.int32)


def test_side_consistent_crossings_are_even_near_a_vertex_on_an_edge():
    """Two closed polygons cross an even number of times.  With a vertex within a few ulps of
    the other polygon's edge the parameter test (the reference's) counts an odd number in a
    sizeable share of configurations; the side-consistent decision never does (a side computed
    from anything but the edge's own vertices and the point, e.g. from P + R, fails this)."""
    rng = np.random.default_rng(0)
    odd = {False: 0, True: 0}
    for _ in range(1000):
        pos, start = _spike_configuration(rng)
        for cons in (False, True):
            odd[cons] += int(O.neighbors(pos, start, 1, consistent=cons)["numNeighbors"].sum()) % 2
    assert odd[False] > 50 and odd[True] == 0


def test_side_consistent_closes_the_window_at_the_stall(stalled):
    """At the packing-41 stall, 1e-16 along the force: with the side-consistent decision both
    crossings of the tiny lens are gone together (292 records, every pair even), and the energy
    changes only by the lens merge of squared lens areas (+2.2e-12), not by +3.7e-9."""
    p, B, sysO, x, f, st, steps = stalled
    sysC = O.System(p["startIndices"], B, O.NORMAL | SQUARED | O.CONSISTENT, p["targetEdgeV"], p["targetAreas"],
                    solver="lanes")
    E0, F0, n0 = sysC.force_eval(x)
    assert n0 == 294
    for h in (1e-17, 1e-16):
        y = x + h * F0 / np.abs(F0).max()
        y -= np.floor(y)
        E1, _, n1 = sysC.force_eval(y)
        rec = O.outersections(p["startIndices"], O.neighbors(y, p["startIndices"], B, consistent=True))
        assert n1 == 292 and rec["numIntersections"] == 292
        assert O.odd_crossing_records(rec["intersections"]) == (0, [])
        assert 1.5e-12 < E1 - E0 < 3.0e-12


@pytest.mark.parametrize("N,n,phi,seed", [(64, 12, 0.80, 0), (1024, 16, 0.88, 5)])
def test_side_consistent_equals_the_reference_on_generic_packings(N, n, phi, seed):
    from pycudapolygon_amd.generate import make_packing
    p = make_packing(N, n, phi, seed=seed, min_center_frac=0.5)
    B = box_size_for(p["diameter"])
    a = O.neighbors(p["positions"], p["startIndices"], B)
    b = O.neighbors(p["positions"], p["startIndices"], B, consistent=True)
    for k in ("neighbors", "inside", "tuSlots", "numNeighbors"):
        assert np.array_equal(a[k], b[k]), k
