"""Side-consistent crossings (PCP_CONSISTENT) on the GPU against the oracle's ORC_CONSISTENT,
through the C ABI, under both detection schedules.  The decision itself is pinned on the CPU
(tests/test_odd_crossings.py: even crossing counts near a vertex on an edge, equal to the
reference's parameter test on generic packings, the packing-41 window closed)."""
import numpy as np
import pytest

from oracle import oracle as O
from pycudapolygon_amd import capi
from pycudapolygon_amd.generate import make_packing, make_batch, box_size_for
from tests.test_odd_crossings import _spike_configuration

pytestmark = pytest.mark.gpu

SQUARED, CONS = 0x200, 0x800
KEYS = ("intersections", "outersections", "tu", "ut")


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


@pytest.mark.parametrize("sched", [0, 1])
@pytest.mark.parametrize("N,n,phi,seed", [(64, 12, 0.80, 0), (1024, 16, 0.85, 3)])
def test_records_and_energy_equal_the_oracle(sched, N, n, phi, seed):
    p = make_packing(N, n, phi, seed=seed, min_center_frac=0.5)
    B = box_size_for(p["diameter"])
    st, x = p["startIndices"], p["positions"]
    print(f"[cons] N={N} schedule {sched}", flush=True)
    s = capi.System([p], [B], model=capi.NORMAL | CONS)
    s.set_detect_schedule(sched)
    s.force_eval()
    r = s.get_records(0)
    rec = O.outersections(st, O.neighbors(x, st, B, consistent=True))
    for k in KEYS:
        np.testing.assert_array_equal(r[k], rec[k])
    s.close()
    s = capi.System([p], [B], model=capi.NORMAL | SQUARED | CONS)
    s.set_detect_schedule(sched)
    s.force_eval()
    Eo, Fo, _ = O.System(st, B, O.NORMAL | SQUARED | CONS, p["targetEdgeV"], p["targetAreas"],
                         solver="lanes").force_eval(x)
    assert s.get_energy()[0] == Eo and np.array_equal(s.get_forces(0), Fo)
    s.close()


@pytest.mark.parametrize("sched", [0, 1])
def test_even_crossings_near_a_vertex_on_an_edge(sched):
    """200 kite/square pairs with the kite's tip within a few ulps of the square's edge, one
    packing each in one context: every pair crosses an even number of times, with the
    oracle's records, while the parameter test leaves odd counts in some of them."""
    rng = np.random.default_rng(1)
    packs = []
    for _ in range(200):
        pos, start = _spike_configuration(rng)
        packs.append(dict(startIndices=start, positions=pos, targetEdgeV=np.zeros(8), targetAreas=np.zeros(2)))
    odd = {}
    for model in (capi.NORMAL, capi.NORMAL | CONS):
        print(f"[cons] spikes model {model:#x} schedule {sched}", flush=True)
        s = capi.System(packs, [1] * len(packs), model=model)
        s.set_detect_schedule(sched)
        s.force_eval()
        odd[model] = 0
        for k, q in enumerate(packs):
            r = s.get_records(k)
            odd[model] += r["numIntersections"] % 2
            o = O.outersections(q["startIndices"], O.neighbors(q["positions"], q["startIndices"], 1,
                                                               consistent=bool(model & CONS)))
            np.testing.assert_array_equal(r["intersections"], o["intersections"])
        s.close()
    assert odd[capi.NORMAL] > 0 and odd[capi.NORMAL | CONS] == 0


def test_window_closed_at_the_stall():
    """The packing-41 stall (squared lens areas; the GPU reaches the oracle's positions, see
    test_odd_crossing_records_flag_the_window), 1e-16 along the force: with PCP_CONSISTENT no
    record is in an odd pair and the energy equals the oracle's ORC_CONSISTENT energy."""
    p = make_batch(256, 3, min_center_frac=0.5)[41]
    B = box_size_for(p["diameter"])
    s = capi.System([p], [B], model=capi.NORMAL | SQUARED)
    prm = s.fire_params(fTol=1e-10, dtMax=100.0)
    s.fire_init(prm, 1e-3)
    s.fire_run(prm, 5000, 8, True)
    x = s.get_positions(0).copy()
    s.force_eval()
    F = s.get_forces(0)
    y = x + 1e-16 * F / np.abs(F).max()
    y -= np.floor(y)
    s.close()
    c = capi.System([dict(p, positions=y)], [B], model=capi.NORMAL | SQUARED | CONS)
    c.force_eval()
    Eo, Fo, nI = O.System(p["startIndices"], B, O.NORMAL | SQUARED | CONS, p["targetEdgeV"],
                          p["targetAreas"]).force_eval(y)
    assert nI == 292 and c.get_records(0)["numIntersections"] == 292
    assert c.odd_crossing_records() == 0 and c.get_energy()[0] == Eo
    c.close()
