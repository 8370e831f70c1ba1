"""GPU FIRE against the FIRE oracle run with the lane-order solves (oracle
System(solver="lanes"): the projection and SHAKE in the kernels' operation order,
pinned against the dense solvers by tests/test_oracle.py).

The projection and SHAKE are bit-identical to that oracle, so FIRE (squared lens areas,
rigid polygons) follows the oracle's trajectory step for step: same rollbacks, same
dt, same positions, and at max|F| <= 1e-10 the same step count and final energies
within 1e-12 relative (expected identical), with and without the learned warm start."""
import numpy as np
import pytest

from oracle import oracle as O
from pycudapolygon_amd import capi
from pycudapolygon_amd import warmstart as W
from pycudapolygon_amd.generate import make_packing, make_batch, box_size_for

pytestmark = pytest.mark.gpu

SQUARED = 0x200
PAIR = 0x400


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


def _wrapdiff(a, b):
    return np.abs((a - b + 0.5) % 1.0 - 0.5).max()


@pytest.mark.parametrize("seed", range(3))
def test_projection_and_shake_bit_identical_to_lane_order_oracle(seed):
    """12-gons (16-lane segments), 16-gons (DPP rows) and a context of uneven packings
    (n 8..32, 32-lane segments): projected forces, SHAKE positions and iteration counts
    equal the lane-order oracle bit for bit; the full force evaluation of the normal
    model (projection fused into the gather) equals the oracle's."""
    rng = np.random.default_rng(seed)
    for n in (12, 16):
        p = make_packing(64, n, 0.8, seed=seed)
        B = box_size_for(p["maxEdgeLength"])
        print(f"[lanes] seed={seed} n={n}", flush=True)
        s = capi.System([p], [B])
        F = rng.normal(size=p["positions"].shape)
        s.force.upload(F)
        s.project()
        assert np.array_equal(s.get_forces(0), O.project_force_lanes(p["positions"], p["startIndices"], F))
        noisy = (p["positions"] + rng.normal(scale=2e-4, size=p["positions"].shape)) % 1.0
        s.pos.upload(noisy)
        it = s.shake(50)
        xo, ito = O.shake_lanes(noisy, p["startIndices"], p["targetEdgeV"], p["targetAreas"], 50)
        assert np.array_equal(s.get_positions(0), xo) and int(it[0]) == ito
        s2 = capi.System([p], [B])
        s2.force_eval()
        sysO = O.System(p["startIndices"], B, "normal", p["targetEdgeV"], p["targetAreas"], solver="lanes")
        Eo, Fo, _ = sysO.force_eval(p["positions"])
        assert np.array_equal(s2.get_forces(0), Fo) and s2.get_energy()[0] == Eo
    packs = make_batch(4, 100 + seed)
    print(f"[lanes] seed={seed} batch n={[int(q['startIndices'][1]) for q in packs]}", flush=True)
    s = capi.System(packs, [box_size_for(q["maxEdgeLength"]) for q in packs])
    Fs = [rng.normal(size=q["positions"].shape) for q in packs]
    s.force.upload(np.concatenate(Fs))
    s.project()
    for k, q in enumerate(packs):
        assert np.array_equal(s.get_forces(k), O.project_force_lanes(q["positions"], q["startIndices"], Fs[k]))


def test_fire_follows_the_oracle_step_by_step_c1():
    """N=64 12-gons, phi=0.80, squared lens areas, dtMax 100: every FIRE step's
    decisions and positions equal the oracle's until both reach max|F| <= 1e-10."""
    p = make_packing(64, 12, 0.80, seed=0, min_center_frac=0.5)
    B = box_size_for(p["diameter"])
    s = capi.System([p], [B], model=capi.NORMAL | SQUARED)
    prm = s.fire_params(fTol=1e-10, dtMax=100.0)
    s.fire_init(prm, 1e-3)
    sysO = O.System(p["startIndices"], B, SQUARED, p["targetEdgeV"], p["targetAreas"], solver="lanes")
    pos = p["positions"].copy()
    st, vel, f = sysO.fire_init(pos, 1e-3, 0.1)
    for k in range(5000):
        gs = s.fire_step(prm)[0]
        sysO.fire_step(pos, vel, f, st, dtMax=100.0)
        xg = s.get_positions(0)
        same = (gs.rolledBack == st.rolledBack and gs.shakeFailed == st.shakeFailed and gs.nPos == st.nPos
                and gs.dt == st.dt and gs.energy == st.energy and np.array_equal(xg, pos))
        assert same, (f"step {k + 1}: rolledBack {gs.rolledBack}/{st.rolledBack} nPos {gs.nPos}/{st.nPos} "
                      f"dt {gs.dt!r}/{st.dt!r} E {gs.energy!r}/{st.energy!r} |dx| {_wrapdiff(xg, pos):.3e}")
        mo = np.abs(f).max()
        if gs.converged or mo <= 1e-10:
            assert gs.converged and mo <= 1e-10 and gs.maxForce == mo
            print(f"[fire-oracle] c1: {k + 1} steps, E {gs.energy:.6e} on both", flush=True)
            return
    pytest.fail("no convergence in 5000 steps")


def test_fire_final_energy_matches_oracle_n4096_plain_and_warm_start():
    """The N=4096 configuration (16-gons, phi=0.85, squared lens areas, dtMax 100): FIRE
    from the generated packing, and from the packing after the committed learned warm
    start, reach max|F| <= 1e-10 in the same number of steps as the oracle started from
    the same positions, with final energies within 1e-12 relative (and identical final
    positions)."""
    p = make_packing(4096, 16, 0.85, seed=0, min_center_frac=0.5)
    B = box_size_for(p["diameter"])
    ws = W.WarmStart.default()
    for warm in (None, ws):
        s, prm = W.fire_system([p], dtMax=W.DT_MAX)
        s.force_eval()
        s.sync()
        if warm is not None:
            warm.apply(s, W.packing_length_units([p]))
        x0 = s.get_positions(0).copy()
        s.fire_init(prm, 1e-3)
        gs = s.fire_minimize(prm, 3000)[0]
        xg = s.get_positions(0)
        sysO = O.System(p["startIndices"], B, SQUARED, p["targetEdgeV"], p["targetAreas"], solver="lanes")
        xo, fo, sto, steps = sysO.fire_minimize(x0, fTol=1e-10, dt=1e-3, maxSteps=3000, dtMax=W.DT_MAX)
        tag = "warm start" if warm is not None else "plain"
        print(f"[fire-oracle] n4096 {tag}: GPU {gs.steps} steps E {gs.energy!r}; oracle {steps} steps "
              f"E {sto.energy!r}; |dx| {_wrapdiff(xg, xo):.3e}", flush=True)
        assert gs.converged and gs.maxForce <= 1e-10 and np.abs(fo).max() <= 1e-10
        assert gs.steps == steps
        assert abs(gs.energy - sto.energy) <= 1e-12 * abs(sto.energy)
        assert np.array_equal(xg, xo)


def test_pair_areas_bit_identical_with_two_lenses_per_pair():
    """A U-shaped polygon under a bar: two lenses of one polygon pair, then one after they
    merge.  Squared pair areas (and squared lens areas) on the GPU equal the oracle bit
    for bit at every bar height."""
    from tests.test_oracle import _u_and_bar
    from tests import _geom as G
    for yc in np.linspace(3.1, 1.2, 12):
        U, bar, _ = _u_and_bar(yc)
        start, pos = G.to_model([U, bar])
        pk = dict(startIndices=start, positions=pos, targetEdgeV=np.zeros(12), targetAreas=np.zeros(2))
        for flag in (SQUARED, PAIR):
            print(f"[pair] yc={yc:.3f} flag={flag:#x}", flush=True)
            s = capi.System([pk], [2])
            s.geometry()
            s.cells()
            s.neighbors()
            s.records(exact_pairing=True)
            s.forces(capi.NORMAL | flag)
            F, E = s.get_forces(0), s.get_energy()[0]
            Eo, Fo, nIo = O.System(start, 2, flag | 5, np.zeros(12), np.zeros(2)).force_eval(pos)
            assert np.array_equal(F, Fo) and E == Eo, (yc, flag, np.abs(F - Fo).max(), E, Eo)


def test_pair_areas_fire_converges_where_lens_areas_stall():
    """Packing 41 of the 256-packing uneven batch (seed 3; N=65 x 16): with squared lens
    areas FIRE stalls (two lenses of one pair merge where a vertex meets an edge, and the
    energy jumps there); with squared pair areas it reaches max|F| <= 1e-10, on the GPU
    in the oracle's number of steps and at the oracle's final energy."""
    p = make_batch(256, 3, min_center_frac=0.5)[41]
    B = box_size_for(p["diameter"])
    s = capi.System([p], [B], model=capi.NORMAL | PAIR)
    prm = s.fire_params(fTol=1e-10, dtMax=100.0)
    s.fire_init(prm, 1e-3)
    gs = s.fire_run(prm, 5000, 8, True)[0]
    sysO = O.System(p["startIndices"], B, PAIR, p["targetEdgeV"], p["targetAreas"], solver="lanes")
    xo, fo, sto, steps = sysO.fire_minimize(p["positions"], fTol=1e-10, dt=1e-3, maxSteps=5000, dtMax=100.0)
    print(f"[pair] packing 41: GPU {gs.steps} steps E {gs.energy!r}; oracle {steps} steps E {sto.energy!r}",
          flush=True)
    assert gs.converged and gs.maxForce <= 1e-10 and np.abs(fo).max() <= 1e-10
    assert gs.steps == steps
    assert abs(gs.energy - sto.energy) <= 1e-12 * abs(sto.energy)


def test_odd_crossing_records_flag_the_window():
    """Packing 41 with squared lens areas: the GPU's FIRE stalls where the oracle's does (same
    positions after 5000 steps) with every polygon pair crossing an even number of times;
    moved by 1e-16 along the force, pair (22, 58) crosses 3 times and pcp_odd_crossing_records
    reports its 3 records, as the oracle's check does, at the oracle's energy (the oracle's side
    is pinned in tests/test_odd_crossings.py)."""
    p = make_batch(256, 3, min_center_frac=0.5)[41]
    B = box_size_for(p["diameter"])
    s = capi.System([p], [B], model=capi.NORMAL | SQUARED)
    prm = s.fire_params(fTol=1e-10, dtMax=100.0)
    s.fire_init(prm, 1e-3)
    gs = s.fire_run(prm, 5000, 8, True)[0]
    sysO = O.System(p["startIndices"], B, SQUARED, p["targetEdgeV"], p["targetAreas"], solver="lanes")
    xo, fo, sto, steps = sysO.fire_minimize(p["positions"], fTol=1e-10, dt=1e-3, maxSteps=5000, dtMax=100.0)
    print(f"[odd] packing 41: GPU {gs.steps} steps E {gs.energy!r}; oracle {steps} steps E {sto.energy!r}",
          flush=True)
    assert not gs.converged and steps == 5000
    x = s.get_positions(0)
    assert np.array_equal(x, xo)
    s.force_eval()
    assert s.odd_crossing_records() == 0
    F = s.get_forces(0)
    y = x + 1e-16 * F / np.abs(F).max()
    y -= np.floor(y)
    s.pos.upload(y)
    s.force_eval()
    Eo, _, _ = sysO.force_eval(y)
    rec = O.outersections(p["startIndices"], O.neighbors(y, p["startIndices"], B))
    assert O.odd_crossing_records(rec["intersections"]) == (3, [(22, 58)])
    assert s.odd_crossing_records() == 3 and s.get_energy()[0] == Eo
    s.close()
