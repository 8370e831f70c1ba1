"""GPU XPBD (pcp_xpbd, pcp_xpbd_ex) and FIRE with XPBD springs against the oracle's
lane-order restatement (orc_xpbd_lanes, orc_fire_step_ex), through the C ABI.

Reference: xpbdProjectCUDA src/model.cu:209-278, kernels.cuh:532-623 (hard edge and area
projection, never called by the reference's Model); compliance after Macklin et al. 2016.
The kernel keeps the oracle's operation order, so positions and iteration counts are equal
bit for bit; the oracle itself is pinned by tests/test_xpbd_oracle.py."""
import numpy as np
import pytest

from oracle import oracle as O
from pycudapolygon_amd import capi
from pycudapolygon_amd.generate import make_packing, make_batch, box_size_for

pytestmark = pytest.mark.gpu

PAIR = 0x400
INF = float("inf")


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


def _all(packs):
    """one oracle call over every packing of a context: polygons are independent"""
    starts, pos, tev, ta, off = [0], [], [], [], 0
    for q in packs:
        st = np.asarray(q["startIndices"])
        starts.extend((st[1:] + off).tolist())
        off += int(st[-1])
        pos.append(q["positions"]); tev.append(q["targetEdgeV"]); ta.append(q["targetAreas"])
    return np.asarray(starts, np.int32), np.concatenate(pos), np.concatenate(tev), np.concatenate(ta)


@pytest.mark.parametrize("n", [12, 13])
def test_reference_xpbd_bit_exact(n):
    """pcp_xpbd (hard edges with one target per polygon, hard area): positions equal the
    lane-order oracle bit for bit, for a fixed iteration count and under a tolerance (same
    per-packing iteration counts), for even and odd n."""
    p = make_packing(64, n, 0.8, seed=n)
    rng = np.random.default_rng(n)
    start = p["startIndices"]
    noisy = (p["positions"] + rng.normal(scale=3e-4, size=p["positions"].shape)) % 1.0
    # equilateral targets (the reference keeps one edge length per polygon) with an area the
    # equilateral polygon can have: 0.97 of the regular polygon's
    Lp = np.add.reduceat(p["targetEdgeV"], start[:-1]) / n
    A0 = 0.97 * n * Lp ** 2 / (4 * np.tan(np.pi / n))
    for nIter, tol in ((7, 0.0), (600, 1e-13)):
        print(f"[xpbd] n={n} nIter={nIter} tol={tol}", flush=True)
        q = dict(p, targetAreas=A0)
        s = capi.System([q], [box_size_for(p["maxEdgeLength"])])
        s.pos.upload(noisy)
        it = s.xpbd(nIter, tol, targetEdgeP=Lp)
        xo, ito = O.xpbd_lanes(noisy, start, INF, INF, 0.0, 1.0, nIter, tol, targetEdgeP=Lp, targetAreas=A0)
        assert np.array_equal(s.get_positions(0), xo)
        assert int(it[0]) == (nIter if tol <= 0 else ito)
        if tol > 0:
            assert ito < nIter
        s.close()


@pytest.mark.parametrize("kE,kA,kP", [(1e3, 1e3, 0.0), (0.0, 1e3, 1e3), (1e3, 1e3, 1e3), (INF, 0.0, 1e2),
                                      (0.0, INF, INF)])
def test_compliant_xpbd_bit_exact_uneven(kE, kA, kP):
    """pcp_xpbd_ex on a context of uneven packings (n 8..32): every constraint set, hard and
    compliant, equals the oracle bit for bit, with the per-packing iteration counts."""
    packs = make_batch(4, 7)
    start, pos, tev, ta = _all(packs)
    rng = np.random.default_rng(1)
    noisy = (pos + rng.normal(scale=2e-4, size=pos.shape)) % 1.0
    s = capi.System(packs, [box_size_for(q["maxEdgeLength"]) for q in packs])
    s.pos.upload(noisy)
    print(f"[xpbd-ex] k={kE},{kA},{kP}", flush=True)
    it = s.xpbd_ex(kE, kA, kP, 0.05, 60, 1e-14)
    xo, _ = O.xpbd_lanes(noisy, start, kE, kA, kP, 0.05, 60, 1e-14, targetEdgeV=tev, targetAreas=ta)
    assert np.array_equal(s.get_positions(), xo)
    for k, q in enumerate(packs):
        a, b = s.vertex_range(k)
        sk = np.asarray(q["startIndices"])
        _, itk = O.xpbd_lanes(noisy[2 * a:2 * b], sk, kE, kA, kP, 0.05, 60, 1e-14, targetEdgeV=q["targetEdgeV"],
                              targetAreas=q["targetAreas"])
        assert int(it[k]) == itk
    s.close()


def test_perimeter_spring_force_bit_exact():
    """hybrid model with squared pair areas and area + perimeter springs: forces and
    energy equal the oracle's bit for bit."""
    p = make_packing(64, 12, 0.80, seed=0, min_center_frac=0.5)
    B = box_size_for(p["diameter"])
    s = capi.System([p], [B], model=capi.HYBRID | PAIR, stiffness=3.0, compressibility=1e3, perimStiffness=1e3)
    s.pos.upload((p["positions"] * 1.0001) % 1.0)
    s.force_eval()
    sysO = O.System(p["startIndices"], B, PAIR | O.HYBRID, p["targetEdgeV"], p["targetAreas"], stiffness=3.0,
                    compressibility=1e3, perimStiffness=1e3)
    Eo, Fo, _ = sysO.force_eval((p["positions"] * 1.0001) % 1.0)
    assert s.get_energy()[0] == Eo and np.array_equal(s.get_forces(0), Fo)
    s.close()


def test_fire_with_xpbd_springs_follows_the_oracle_c1():
    """C1 (N=64 12-gons, phi=0.80), hybrid model: squared pair areas with area and perimeter
    constraints at stiffness 1e3 integrated by XPBD (20 iterations, tol 1e-15), dtMax 100.
    Every step's decisions, energy and positions equal the oracle's until both reach
    max|F| <= 1e-10; FIRE in graph-replayed blocks ends at the same state."""
    p = make_packing(64, 12, 0.80, seed=0, min_center_frac=0.5)
    B = box_size_for(p["diameter"])
    kw = dict(stiffness=0.0, compressibility=1e3, perimStiffness=1e3)
    s = capi.System([p], [B], model=capi.HYBRID | PAIR, **kw)
    prm = s.fire_params(fTol=1e-10, dtMax=100.0, shakeIter=0, xpbdIter=20, xpbdTol=1e-15)
    s.fire_init(prm, 1e-3)
    sysO = O.System(p["startIndices"], B, PAIR | O.HYBRID, p["targetEdgeV"], p["targetAreas"], **kw)
    pos = p["positions"].copy()
    st, vel, f = sysO.fire_init(pos, 1e-3, 0.1)
    steps = None
    for k in range(6000):
        gs = s.fire_step(prm)[0]
        sysO.fire_step(pos, vel, f, st, dtMax=100.0, shakeIter=0, xpbdIter=20, xpbdTol=1e-15)
        same = (gs.rolledBack == st.rolledBack and gs.nPos == st.nPos and gs.dt == st.dt
                and gs.energy == st.energy and gs.lastShakeIters == st.lastShakeIters)
        if not same or (k % 100 == 0 and not np.array_equal(s.get_positions(0), pos)):
            pytest.fail(f"step {k + 1}: rolledBack {gs.rolledBack}/{st.rolledBack} nPos {gs.nPos}/{st.nPos} "
                        f"dt {gs.dt!r}/{st.dt!r} E {gs.energy!r}/{st.energy!r} "
                        f"it {gs.lastShakeIters}/{st.lastShakeIters}")
        if gs.converged or np.abs(f).max() <= 1e-10:
            assert gs.converged and np.abs(f).max() <= 1e-10
            assert np.array_equal(s.get_positions(0), pos)
            steps = k + 1
            print(f"[xpbd-fire] c1: {steps} steps, E {gs.energy:.6e} on both", flush=True)
            break
    assert steps is not None, "no convergence in 6000 steps"
    b = capi.System([p], [B], model=capi.HYBRID | PAIR, **kw)
    b.fire_init(prm, 1e-3)
    gb = b.fire_run(prm, 6000, 8, True)[0]
    assert gb.converged and gb.steps == steps and gb.energy == st.energy
    assert np.array_equal(b.get_positions(0), pos)
    s.close()
    b.close()


def test_fire_with_xpbd_rejects_rigid_model():
    p = make_packing(16, 8, 0.7, seed=1)
    s = capi.System([p], [box_size_for(p["maxEdgeLength"])])
    prm = s.fire_params(xpbdIter=5)
    with pytest.raises(capi.PcpError, match="XPBD"):
        s.fire_init(prm, 1e-3)
    s.close()
