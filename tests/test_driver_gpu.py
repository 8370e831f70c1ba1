"""The reference's Python front end (pyCudaPolygon.model, mirrored by
pycudapolygon_amd.pyCudaPolygon) driving the GPU, against the oracle.

* updatePositions / saveTentativePositions / getMaxEffectiveForce (kernels.cuh:1433-1438,
  625-633; model.cpp:328-335) and a few minimizeGDStep / minimizeGD calls
  (pyCudaPolygon.py:573-647) against the oracle's force, update and SHAKE, bit for bit.
* minimizeFIRE with its GD kick and its flat-energy stop (pyCudaPolygon.py:674-731), with the
  reference's energy (where FIRE stalls and the flat-energy stop ends the run) and with squared
  pair areas (where it converges), against the oracle's FIRE driven by the same rules.
* minimizeSoftBody (pyCudaPolygon.py:764-773).
* saveModel / loadModel through the native Model, getConstrainedForcePython against the
  oracle's projection.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pycudapolygon_amd import capi
from pycudapolygon_amd.generate import make_packing

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


def _mirror(p, modelType="normal", stiffness=0, compressibility=0):
    from pycudapolygon_amd.pyCudaPolygon import model
    V = int(p["startIndices"][-1])
    m = model(size=V, modelType=modelType, stiffness=stiffness, compressibility=compressibility)
    m.setNumVertices(V)
    m.setnArray(np.diff(p["startIndices"]))
    m.setPositions(p["positions"])
    m.setTargetEdgeLengthsPerEdge(p["targetEdgeV"].tolist())
    m.setTargetAreas(p["targetAreas"])
    m.initializeNeighborCells()   # the cell side goes to the 0.5 default: a 2x2 grid (K5)
    return m


def _eval_chain(m):
    m.updateNeighborCells()
    m.updateNeighbors()
    m.updateOutersections()
    m.updatePolygonGeometry()
    m.updateForceEnergy()


def test_update_positions_and_effective_force():
    p = make_packing(64, 12, 0.80, seed=1)
    m = _mirror(p)
    _eval_chain(m)
    F = m.getForces()
    x0 = m.getPositions()
    m.saveTentativePositions()
    dt = 2e-3
    m.updatePositions(dt)
    x1 = m.getPositions()
    assert np.array_equal(x1, O.update_positions(x0, F, dt))
    for kind, scale in (("GD", 1.0 / dt), ("FIRE", 2.0 / (dt * dt))):
        assert m.getMaxEffectiveForce(dt, kind) == O.max_effective_force(x1, x0, F, scale), kind


@pytest.mark.parametrize("modelType", ["normal", "softBody"])
def test_gd_steps_follow_the_oracle(modelType):
    """minimizeGDStep (fixed dt) six times: the positions after each step equal the oracle's
    force (projected for the normal model), update and SHAKE (lane order) bit for bit."""
    p = make_packing(64, 12, 0.80, seed=2)
    k, c = (0.0, 0.0) if modelType == "normal" else (1.0, 1.0)
    m = _mirror(p, modelType, k, c)
    sysO = O.System(p["startIndices"], 2, modelType, p["targetEdgeV"], p["targetAreas"], stiffness=k,
                    compressibility=c, solver="lanes")
    x = p["positions"].copy()
    dt = 1e-3 if modelType == "normal" else 0.05
    for step in range(6):
        E, _ = m.minimizeGDStep(dt=dt, maxDisplacement=1.0)
        Eo, Fo, _ = sysO.force_eval(x)
        assert E == Eo, step
        x = O.update_positions(x, Fo, dt)
        if modelType == "normal":
            x, _ = O.shake_lanes(x, p["startIndices"], p["targetEdgeV"], p["targetAreas"], 100, 1e-15)
        assert np.array_equal(m.getPositions(), x), step


def test_minimize_gd_adaptive_dt_follows_the_oracle():
    """minimizeGD (softBody, 40 steps): dt halves by 2.1 after an energy rise and grows by 1.9
    after more than five successes, and the run returns with dt 0 once max|F| <= the threshold
    (pyCudaPolygon.py:620-640); the returned energy and dt and the final positions equal the
    same rules applied to the oracle."""
    p = make_packing(32, 10, 0.75, seed=3)
    m = _mirror(p, "softBody", 1.0, 1.0)
    E, dt, _ = m.minimizeGD(maxUnbalancedForceThreshold=1e-14, dt=0.02, maxSteps=40, maxDisplacement=1.0)
    sysO = O.System(p["startIndices"], 2, "softBody", p["targetEdgeV"], p["targetAreas"], stiffness=1.0,
                    compressibility=1.0)
    x, d, succ, prev = p["positions"].copy(), 0.02, 0, None
    for _ in range(40):
        e, f, _ = sysO.force_eval(x)
        x = O.update_positions(x, f, d)
        if np.abs(f).max() <= 1e-14:   # the force threshold returns (energy, 0, ...)
            d = 0
            break
        if prev is not None and e > prev:
            d, succ = max(1e-32, d / 2.1), 0
        else:
            succ, prev = succ + 1, e
        if succ > 5:
            succ, d = 0, min(1e4, d * 1.9)
    assert E == e and dt == d
    assert np.array_equal(m.getPositions(), x)


def _oracle_front_end(p, sysO, thr, dt, maxSteps, dtMax=0.1):
    """minimizeFIRE's rules (pyCudaPolygon.py:674-731) over the oracle's FIRE step"""
    E0, F0, _ = sysO.force_eval(p["positions"])
    x = O.update_positions(p["positions"], F0, dt)   # the GD kick
    st, vel, f = sysO.fire_init(x, dt, 0.1)
    flat, prev, energy, steps = 0, 1e9, st.energy, 0
    for k in range(maxSteps):
        ePre = st.energy
        sysO.fire_step(x, vel, f, st, dtMax=dtMax)
        # a rolled-back step reports the energy before the step (model.cpp:469)
        energy, steps = (ePre if st.rolledBack else st.energy), k + 1
        if np.abs(f).max() <= thr:
            break
        if abs(energy - prev) < 1e-15:
            flat += 1
        if flat > 5:
            break
        prev = energy
    return energy, st.dt, steps, x


@pytest.mark.parametrize("energy", ["reference", "squaredPair"])
def test_front_end_minimize_fire_c1(energy):
    p = make_packing(64, 12, 0.80, seed=0, min_center_frac=0.5)
    m = _mirror(p)
    m.setOverlapEnergy(energy)
    dtMax = 0.1 if energy == "reference" else 100.0
    E, dt, steps, _ = m.minimizeFIRE(maxForceThreshold=1e-10, dt=1e-3, maxSteps=3000, dtMax=dtMax)
    flag = {"reference": O.NORMAL, "squaredPair": 0x400}[energy]
    sysO = O.System(p["startIndices"], 2, flag, p["targetEdgeV"], p["targetAreas"], solver="lanes")
    Eo, dto, stepso, xo = _oracle_front_end(p, sysO, 1e-10, 1e-3, 3000, dtMax=dtMax)
    print(f"[front-end] {energy}: {steps} steps E {E!r} (oracle {stepso}, {Eo!r})", flush=True)
    assert steps == stepso and E == Eo and dt == dto
    assert np.array_equal(m.getPositions(), xo)
    if energy == "squaredPair":
        assert m.getMaxUnbalancedForce() <= 1e-10
    else:
        assert steps < 3000   # stopped by the flat-energy rule: FIRE stalls on the reference energy


def test_minimize_soft_body():
    """minimizeSoftBody: FIRE on the edge and area springs (stiffness and compressibility 1,
    dt 0.1), then every polygon rescaled to its target area; the model type comes back."""
    p = make_packing(16, 8, 0.7, seed=4)
    rng = np.random.default_rng(0)
    p = dict(p, targetEdgeV=p["targetEdgeV"] * rng.uniform(0.97, 1.03, p["targetEdgeV"].shape))
    a = _mirror(p)
    a.setModelEnum("softBody")
    a.setStiffness(1)
    a.setCompressibility(1)
    a.minimizeFIRE(dt=0.1, maxForceThreshold=1e-14)
    a.resetAreas()
    a.updatePolygonGeometry()
    b = _mirror(p)
    b.minimizeSoftBody()
    assert b.getModelEnum() == "normal" and b.getStiffness() == 1 and b.getCompressibility() == 1
    assert np.array_equal(b.getPositions(), a.getPositions())
    b.updatePolygonGeometry()
    np.testing.assert_allclose(b.getAreas(), p["targetAreas"], rtol=1e-12, atol=0)
    el0 = O.geometry(p["positions"], p["startIndices"])["edgeLengths"]
    r0 = np.abs(el0 / p["targetEdgeV"] - 1).max()
    r1 = np.abs(b.getEdgeLengths() / p["targetEdgeV"] - 1).max()
    assert r1 < r0


def test_save_load_through_the_native_model(tmp_path):
    p = make_packing(64, 12, 0.80, seed=5)
    m = _mirror(p, "softBody", 2.0, 3.0)
    # one target edge length per polygon: the reference's format keeps that much
    m.setTargetEdgeLengths(np.add.reduceat(p["targetEdgeV"], p["startIndices"][:-1]) / 12)
    m.setMaxEdgeLength(0.1)
    _eval_chain(m)
    d = str(tmp_path / "m")
    m.saveModel(d)
    from pycudapolygon_amd.pyCudaPolygon import model
    n = model()
    n.loadModel(d)
    assert n.getModelEnum() == "softBody" and n.getStiffness() == 2.0 and n.getCompressibility() == 3.0
    assert np.array_equal(n.getPositions(), m.getPositions())
    assert np.array_equal(n.getnArray(), m.getnArray())
    assert np.array_equal(n.getTargetAreas(), m.getTargetAreas())
    assert np.array_equal(n.getTargetEdgeLengths(), m.getTargetEdgeLengths())
    assert n.getMaxEdgeLength() == 0.5   # loadModel re-initializes the cells (reference)
    n.updatePolygonGeometry()
    n.updateForceEnergy()
    m.updateForceEnergy()
    assert n.getEnergy() == m.getEnergy()


def test_constrained_force_python_equals_the_projection():
    p = make_packing(64, 12, 0.80, seed=6)
    m = _mirror(p)
    f = np.random.default_rng(1).normal(size=p["positions"].shape)
    g = m.getConstrainedForcePython(f)
    go = O.project_force(p["positions"], p["startIndices"], f)
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-12 * np.abs(f).max())
