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
