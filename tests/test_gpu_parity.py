"""GPU parity: the HIP path (through the C ABI, include/pcp.h) against the CPU oracle
on the same seeded inputs.  Integer work (cells, crossings, records, pairing) must be
bit-identical; floating-point results are compared with the tolerance written in each
test (the reference-convention forces are expected bit-identical as well, since both
sides accumulate in the same fixed order without FMA contraction)."""
import numpy as np
import pytest

from oracle import oracle as O
from pycudapolygon_amd import capi
from pycudapolygon_amd.generate import make_packing, make_batch, box_size_for

pytestmark = pytest.mark.gpu

AREA = 0x100
SQUARED = 0x200
PAIR = 0x400


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


def c1(seed, B=None):
    p = make_packing(64, 12, 0.80, seed=seed)
    return p, (B if B is not None else box_size_for(p["maxEdgeLength"]))


def run_gpu(p, B, model=capi.NORMAL, exact=False):
    s = capi.System([p], [B], model=model)
    s.geometry()
    s.cells()
    mx = s.neighbors()
    nI = s.records(exact_pairing=exact)
    return s, mx, nI


def oracle_records(p, B):
    start, pos = p["startIndices"], p["positions"]
    nb = O.neighbors(pos, start, B)
    return nb, O.outersections(start, nb)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("which", ["edge", "diam", "two"])
def test_cells_neighbors_records_bit_exact(seed, which):
    p = make_packing(64, 12, 0.80, seed=seed)
    B = {"edge": box_size_for(p["maxEdgeLength"]), "diam": box_size_for(p["diameter"]), "two": 2}[which]
    print(f"[parity] seed={seed} B={B}", flush=True)
    s, mx, nI = run_gpu(p, B)
    start, pos = p["startIndices"], p["positions"]
    cl, ni, cpb = s.get_cells(0)
    ocl, oni, ocpb = O.cells(pos, start, B)
    np.testing.assert_array_equal(cl, ocl)
    np.testing.assert_array_equal(ni, oni)
    np.testing.assert_array_equal(cpb, ocpb)
    nb, rec = oracle_records(p, B)
    g = s.get_neighbor_lists(0)
    M = g["maxNeighbors"]
    assert M == nb["maxNeighbors"]
    np.testing.assert_array_equal(g["numNeighbors"], nb["numNeighbors"])
    for i in np.nonzero(nb["numNeighbors"])[0]:
        k = nb["numNeighbors"][i]
        np.testing.assert_array_equal(g["neighbors"][i * M:i * M + k], nb["neighbors"][i * M:i * M + k])
        np.testing.assert_array_equal(g["inside"][i * M:i * M + k], nb["inside"][i * M:i * M + k])
    assert mx == int(nb["numNeighbors"].max())
    r = s.get_records(0)
    assert nI[0] == rec["numIntersections"] == r["numIntersections"]
    for key in ("intersections", "outersections", "tu", "ut", "shapeCounts"):
        np.testing.assert_array_equal(r[key], rec[key], err_msg=key)
    # compaction order: edge by edge, slots in the order of the stencil scan
    comp = s.get_compacted_records(0)
    want = np.concatenate([nb["tuSlots"][2 * i * M:2 * (i * M + nb["numNeighbors"][i])]
                           for i in range(len(nb["numNeighbors"]))] + [np.zeros(0)])
    np.testing.assert_array_equal(comp["tu"], want.astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(np.sort(comp["intersections"]), np.sort(rec["intersections"]))


@pytest.mark.parametrize("seed", range(6))
def test_geometry_bit_exact(seed):
    p, B = c1(seed)
    s = capi.System([p], [B])
    s.geometry()
    s.sync()
    g = O.geometry(p["positions"], p["startIndices"])
    np.testing.assert_array_equal(s.edgeLengths.get(), g["edgeLengths"])
    np.testing.assert_array_equal(s.areas.get(), g["areas"])
    np.testing.assert_array_equal(s.comX.get(), g["comX"])
    np.testing.assert_array_equal(s.comY.get(), g["comY"])
    np.testing.assert_array_equal(s.constraints.get(), g["constraints"])
    assert s.maxEdge.get()[0] == g["maxEdgeLength"]


@pytest.mark.parametrize("n_range", [(3, 4), (5, 8), (9, 16), (17, 32), (33, 64), (60, 80)],
                         ids=["seg4", "seg8", "seg16", "seg32", "seg64", "thread-per-polygon"])
def test_geometry_bit_exact_uneven(n_range):
    """pcp_update_geometry on uneven packings (lane segments of 4..64 lanes, the thread per
    polygon beyond 64 vertices) against the oracle's updatePolygonGeometry, packing by packing."""
    from pycudapolygon_amd.generate import make_batch
    pk = make_batch(3, 90 + n_range[1], N_range=(20, 60), n_range=n_range)   # centres uniform
    s = capi.System(pk, [box_size_for(p["diameter"]) for p in pk])
    s.geometry()
    s.sync()
    el, ar, cx, cy, cons, mx = (s.edgeLengths.get(), s.areas.get(), s.comX.get(), s.comY.get(),
                                s.constraints.get(), s.maxEdge.get())
    for k, p in enumerate(pk):
        g = O.geometry(p["positions"], p["startIndices"])
        a, b = s.vertex_range(k)
        u, w = int(s.packPoly[k]), int(s.packPoly[k + 1])
        np.testing.assert_array_equal(el[a:b], g["edgeLengths"])
        np.testing.assert_array_equal(ar[u:w], g["areas"])
        np.testing.assert_array_equal(cx[u:w], g["comX"])
        np.testing.assert_array_equal(cy[u:w], g["comY"])
        np.testing.assert_array_equal(cons[6 * a:6 * b], g["constraints"])
        assert mx[k] == g["maxEdgeLength"]


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("model", [capi.NORMAL, capi.HYBRID, capi.SOFT_BODY, capi.EDGE_ONLY, capi.AREA_ONLY])
def test_force_energy_parity(seed, model):
    """Tolerance 1e-12 relative (norm-wise) on forces and energy; expected exact."""
    p, B = c1(seed)
    s, mx, nI = run_gpu(p, B, model=model)
    s.stiffness, s.compressibility = 3.0, 7.0
    s.forces(model)
    F = s.get_forces(0)
    E = s.get_energy()[0]
    nb, rec = oracle_records(p, B)
    g = O.geometry(p["positions"], p["startIndices"])
    Eo, Fo = O.force_energy(model, p["positions"], p["startIndices"], rec, g, p["targetEdgeV"],
                            p["targetAreas"], 3.0, 7.0)
    scale = max(np.abs(Fo).max(), 1e-300)
    assert np.abs(F - Fo).max() <= 1e-12 * scale
    assert abs(E - Eo) <= 1e-12 * max(abs(Eo), 1e-300)
    assert np.array_equal(F, Fo) and E == Eo, "fixed-order accumulation should be bit-identical"


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("flag", [AREA, SQUARED, PAIR])
def test_area_energy_parity(seed, flag):
    """Area energy, squared lens areas and squared pair areas: GPU vs oracle, bit for bit
    (same exact crossings, cycle origins, lens representatives and accumulation order)."""
    p, B = c1(seed)
    s, mx, nI = run_gpu(p, B, exact=True)
    s.forces(capi.NORMAL | flag)
    F, E = s.get_forces(0), s.get_energy()[0]
    start, pos = p["startIndices"], p["positions"]
    sysO = O.System(start, B, flag | 5, p["targetEdgeV"], p["targetAreas"])   # hybrid: no projection
    Eo, Fo, nIo = sysO.force_eval(pos)
    assert nIo == nI[0]
    assert np.array_equal(F, Fo), np.abs(F - Fo).max()
    assert E == Eo


def test_fire_squared_lens_areas_converges_c1():
    """FIRE on squared lens areas (N=64 12-gons, phi=0.80, no polygon starting inside
    another) reaches max|F| <= 1e-10; the oracle's FIRE ends at the same energy scale."""
    p = make_packing(64, 12, 0.80, seed=0, min_center_frac=0.5)
    B = box_size_for(p["maxEdgeLength"])
    s = capi.System([p], [B], model=capi.NORMAL | SQUARED)
    prm = s.fire_params(fTol=1e-10)
    s.fire_init(prm, 1e-3)
    st = s.fire_minimize(prm, 20000)[0]
    print(f"[fire] steps={st.steps} E={st.energy:.3e} maxF={st.maxForce:.3e}", flush=True)
    assert st.converged and st.maxForce <= 1e-10
    sysO = O.System(p["startIndices"], B, SQUARED, p["targetEdgeV"], p["targetAreas"])
    xo, fo, sto, steps = sysO.fire_minimize(p["positions"], fTol=1e-10, maxSteps=20000)
    assert np.abs(fo).max() <= 1e-10
    assert st.energy < 1e-10 and sto.energy < 1e-10


@pytest.mark.parametrize("seed", range(4))
def test_projection_parity(seed):
    """Projection: GPU (bordered cyclic-tridiagonal normal equations) vs oracle
    (Householder QR); tolerance 1e-12 of the unprojected force magnitude."""
    p, B = c1(seed)
    s, mx, nI = run_gpu(p, B)
    s.forces(capi.NORMAL)
    Fu = s.get_forces(0).copy()
    s.project()
    Fp = s.get_forces(0)
    Fo = O.project_force(p["positions"], p["startIndices"], Fu)
    assert np.abs(Fp - Fo).max() <= 1e-12 * np.abs(Fu).max()


@pytest.mark.parametrize("seed", range(3))
def test_full_force_eval_matches_oracle(seed):
    p, B = c1(seed)
    s = capi.System([p], [B])
    s.force_eval()
    F, E = s.get_forces(0), s.get_energy()[0]
    sysO = O.System(p["startIndices"], B, "normal", p["targetEdgeV"], p["targetAreas"])
    Eo, Fo, nI = sysO.force_eval(p["positions"])
    assert E == Eo
    assert np.abs(F - Fo).max() <= 1e-12 * np.abs(Fo).max()


def test_deterministic_and_cell_size_independent_c2():
    """N=4096, n=16, phi=0.85: two evaluations are bit-identical; edge-sized, diameter
    and 2x2 cell grids give identical records and forces; integer parity vs oracle."""
    p = make_packing(4096, 16, 0.85, seed=1)
    res = []
    for B in (box_size_for(p["maxEdgeLength"]), box_size_for(p["diameter"])):
        s = capi.System([p], [B])
        s.force_eval()
        r = s.get_records(0)
        F = s.get_forces(0).copy()
        s.force_eval()
        assert np.array_equal(s.get_forces(0), F)
        res.append((r, F, s.get_energy()[0]))
    for key in ("intersections", "outersections", "tu", "ut"):
        np.testing.assert_array_equal(res[0][0][key], res[1][0][key])
    assert np.array_equal(res[0][1], res[1][1]) and res[0][2] == res[1][2]
    nb, rec = oracle_records(p, box_size_for(p["maxEdgeLength"]))
    for key in ("intersections", "outersections", "tu", "ut"):
        np.testing.assert_array_equal(res[0][0][key], rec[key])
    sysO = O.System(p["startIndices"], box_size_for(p["maxEdgeLength"]), "normal", p["targetEdgeV"],
                    p["targetAreas"])
    Eo, Fo, _ = sysO.force_eval(p["positions"])
    assert res[0][2] == Eo
    assert np.abs(res[0][1] - Fo).max() <= 1e-12 * np.abs(Fo).max()


def test_batched_packings_match_single_packing_oracle():
    """Uneven packings (N, n, phi differ) in one context: each equals its own oracle."""
    packs = make_batch(5, seed=3, N_range=(40, 200), n_range=(6, 20))
    Bs = [box_size_for(p["maxEdgeLength"]) for p in packs]
    s = capi.System(packs, Bs)
    s.force_eval()
    E = s.get_energy()
    for k, p in enumerate(packs):
        print(f"[batch] packing {k} N={len(p['startIndices'])-1} B={Bs[k]}", flush=True)
        nb, rec = oracle_records(p, Bs[k])
        r = s.get_records(k)
        for key in ("intersections", "outersections", "tu", "ut", "shapeCounts"):
            np.testing.assert_array_equal(r[key], rec[key], err_msg=key)
        sysO = O.System(p["startIndices"], Bs[k], "normal", p["targetEdgeV"], p["targetAreas"])
        Eo, Fo, _ = sysO.force_eval(p["positions"])
        assert E[k] == Eo
        assert np.abs(s.get_forces(k) - Fo).max() <= 1e-12 * np.abs(Fo).max()


def test_fixture_known_answer_through_c_abi():
    from tests.test_oracle import fixture
    d, start, pos = fixture()
    pk = dict(startIndices=start, positions=pos, targetEdgeV=np.zeros(10), targetAreas=np.zeros(2))
    s = capi.System([pk], [4])
    s.cells()
    s.neighbors()
    g = s.get_neighbor_lists(0)
    M = g["maxNeighbors"]
    assert list(np.nonzero(g["numNeighbors"])[0]) == [1]
    assert g["neighbors"][M:M + 2].tolist() == [7, 8]
    assert g["inside"][M:M + 2].tolist() == [1, 0]


@pytest.mark.parametrize("seed", range(3))
def test_shake_parity(seed):
    """SHAKE: structured solve (GPU) vs dense elimination (oracle); positions within
    1e-13 absolute (box units), constraints satisfied to 1e-15."""
    p = make_packing(64, 12, 0.8, seed=seed)
    rng = np.random.default_rng(seed)
    noisy = (p["positions"] + rng.normal(scale=2e-4, size=p["positions"].shape)) % 1.0
    pk = dict(p, positions=noisy)
    s = capi.System([pk], [2])
    it = s.shake(50)
    xg = s.get_positions(0)
    xo, ito = O.shake(noisy, p["startIndices"], p["targetEdgeV"], p["targetAreas"], 50)
    d = (xg - xo + 0.5) % 1.0 - 0.5
    assert np.abs(d).max() < 1e-13
    assert abs(int(it[0]) - ito) <= 1
    g = O.geometry(xg, p["startIndices"])
    assert np.abs(g["edgeLengths"] - p["targetEdgeV"]).max() < 1e-15
    assert np.abs(g["areas"] - p["targetAreas"]).max() < 1e-15


@pytest.mark.parametrize("seed", range(2))
def test_sixteen_vertex_solves_bit_identical_to_general_path(seed):
    """Packings whose polygons all have 16 vertices run the projection and SHAKE with
    DPP row moves; with one 15-vertex polygon in another packing of the same context they
    take the general shuffle path.  The 16-vertex packing's results must be bit-identical
    between the two, and match the oracle within the tolerances of the tests above."""
    p = make_packing(64, 16, 0.8, seed=seed)
    extra = make_packing(4, 15, 0.3, seed=seed + 10)
    B, Bx = box_size_for(p["maxEdgeLength"]), box_size_for(extra["maxEdgeLength"])
    rng = np.random.default_rng(seed)
    F = rng.normal(size=p["positions"].shape)
    Fx = rng.normal(size=extra["positions"].shape)
    noisy = (p["positions"] + rng.normal(scale=2e-4, size=p["positions"].shape)) % 1.0
    res = []
    for packs, Bs, forces in (([p], [B], F), ([p, extra], [B, Bx], np.concatenate([F, Fx]))):
        print(f"[row16] seed={seed} packings={len(packs)}", flush=True)
        s = capi.System(packs, Bs)
        s.force.upload(forces)
        s.project()
        Fp = s.get_forces(0).copy()
        s.pos.upload(np.concatenate([noisy] + [q["positions"] for q in packs[1:]]))
        it = s.shake(50)
        res.append((Fp, s.get_positions(0).copy(), int(it[0])))
        s2 = capi.System(packs, Bs)
        prm = s2.fire_params(fTol=1e-10)
        s2.fire_init(prm, 1e-3)
        for _ in range(10):
            s2.fire_step(prm)
        res[-1] += (s2.get_positions(0).copy(),)
    (Fa, xa, ia, fa), (Fb, xb, ib, fb) = res
    assert np.array_equal(Fa, Fb) and np.array_equal(xa, xb) and ia == ib
    assert np.array_equal(fa, fb)
    Fo = O.project_force(p["positions"], p["startIndices"], F)
    assert np.abs(Fa - Fo).max() <= 1e-12 * np.abs(F).max()
    xo, ito = O.shake(noisy, p["startIndices"], p["targetEdgeV"], p["targetAreas"], 50)
    d = (xa - xo + 0.5) % 1.0 - 0.5
    assert np.abs(d).max() < 1e-13 and abs(ia - ito) <= 1


def test_fire_steps_track_oracle():
    """First FIRE steps (reference convention) vs the oracle's FIRE: same step
    decisions and positions within 1e-12 (projection/SHAKE differ in rounding only)."""
    p, B = c1(0)
    s = capi.System([p], [B])
    prm = s.fire_params(fTol=1e-10)
    s.fire_init(prm, 1e-3)
    sysO = O.System(p["startIndices"], B, "normal", p["targetEdgeV"], p["targetAreas"])
    pos = p["positions"].copy()
    st, vel, f = sysO.fire_init(pos, 1e-3, 0.1)
    for k in range(20):
        gs = s.fire_step(prm)[0]
        sysO.fire_step(pos, vel, f, st)
        assert gs.rolledBack == st.rolledBack and gs.shakeFailed == st.shakeFailed
        assert abs(gs.dt - st.dt) <= 1e-15 * st.dt and gs.nPos == st.nPos
        d = (s.get_positions(0) - pos + 0.5) % 1.0 - 0.5
        assert np.abs(d).max() < 1e-12, (k, np.abs(d).max())
        np.testing.assert_allclose(gs.energy, st.energy, rtol=1e-10)


def test_fire_minimize_two_polygons_area_energy():
    from tests import _geom as G
    polys = [G.regular(8, 0.06, (0.4, 0.4), 0.1), G.regular(8, 0.06, (0.49, 0.42), 0.3)]
    start, pos = G.to_model(polys)
    tev = np.full(16, 0.06 * 2 * np.sin(np.pi / 8))
    pk = dict(startIndices=start, positions=pos, targetEdgeV=tev,
              targetAreas=np.array([G.shoelace(polys[0]), G.shoelace(polys[1])]))
    s = capi.System([pk], [2], model=capi.NORMAL | AREA)
    prm = s.fire_params(fTol=1e-10)
    s.fire_init(prm, 1e-3)
    st = s.fire_minimize(prm, 20000)[0]
    assert st.converged and st.maxForce <= 1e-10
    sysO = O.System(start, 2, AREA, tev, pk["targetAreas"])
    xo, fo, sto, steps = sysO.fire_minimize(pos, fTol=1e-10, maxSteps=20000)
    assert np.abs(fo).max() <= 1e-10
    assert abs(st.energy - sto.energy) <= 1e-12 * max(abs(sto.energy), 1e-300)


def test_model_binding_smoke():
    """The pybind11 module (reference Model API) on the GPU."""
    from pycudapolygon_amd.pyCudaPolygonLink import libpyCudaPolygon as lpcp
    p, B = c1(2)
    m = lpcp.Model(0)
    V = int(p["startIndices"][-1])
    m.setNumVertices(V)
    m.setStartIndices(p["startIndices"].tolist())
    m.setPositions(p["positions"].tolist())
    m.setMaxEdgeLength(1.0 / B)
    m.initializeNeighborCells()
    m.updateNeighborCells()
    m.updateNeighbors()
    m.updateOutersections()
    m.updatePolygonGeometry()
    m.updateForceEnergy()
    nb, rec = oracle_records(p, B)
    assert m.getNumIntersections() == rec["numIntersections"]
    np.testing.assert_array_equal(np.array(m.getIntersections(), dtype=np.uint64), rec["intersections"])
    g = O.geometry(p["positions"], p["startIndices"])
    Eo, Fo = O.force_energy("normal", p["positions"], p["startIndices"], rec, g)
    assert m.getEnergy() == Eo
    np.testing.assert_array_equal(np.array(m.getForces()), Fo)


def _two_noisy_polygons(n, noise, seed):
    """Two noisy n-gons of radius 0.1 almost on top of each other (centres 0.002 apart, the
    second turned by half a vertex step): their boundaries cross at many edges."""
    from tests import _geom as G
    rng = np.random.default_rng(seed)
    c = np.array([0.5, 0.5])

    def noisy(cen, th):
        P = G.regular(n, 0.1, cen, th)
        return cen + (P - cen) * (1 + rng.uniform(-noise, noise, n))[:, None]

    polys = [noisy(c, 0.0), noisy(c + 0.002, np.pi / n)]
    start, pos = G.to_model(polys)
    g = O.geometry(pos, start)
    return dict(startIndices=start, positions=pos, targetEdgeV=g["edgeLengths"].copy(),
                targetAreas=g["areas"].copy())


def test_bucket_overflow_regrows_and_polygons_beyond_64_vertices():
    """Two noisy 384-gons crossing 168 times (84 records per polygon, above the initial
    16-entry record buckets): the first evaluation reports PCP_ERR_CAPACITY at the next
    synchronisation with the buckets enlarged; the evaluation after it gives the oracle's
    records bit for bit and its energy exactly, and forces within 1e-12 (polygons beyond 64
    vertices are projected by the thread-serial solver).  SHAKE of those polygons matches
    the dense oracle to 1e-13.  The reference warns and resizes (model.cpp:264-305)."""
    pk = _two_noisy_polygons(384, 0.05, 5)
    start, pos = pk["startIndices"], pk["positions"]
    nb, rec = oracle_records(pk, 2)
    assert rec["numIntersections"] == 168 and list(rec["shapeCounts"]) == [84, 84]
    s = capi.System([pk], [2])
    s.force_eval()
    with pytest.raises(capi.PcpError, match="overflow"):
        s.sync()
    s.force_eval()
    s.sync()
    r = s.get_records(0)
    for key in ("intersections", "outersections", "tu", "ut", "shapeCounts"):
        np.testing.assert_array_equal(r[key], rec[key], err_msg=key)
    F, E = s.get_forces(0), s.get_energy()[0]
    Eo, Fo, nIo = O.System(start, 2, "normal", pk["targetEdgeV"], pk["targetAreas"]).force_eval(pos)
    assert nIo == 168 and E == Eo
    assert np.abs(F - Fo).max() <= 1e-12 * np.abs(Fo).max()
    rng = np.random.default_rng(6)
    noisy = (pos + rng.normal(scale=2e-5, size=pos.shape)) % 1.0
    s.pos.upload(noisy)
    it = s.shake(50)
    xo, ito = O.shake(noisy, start, pk["targetEdgeV"], pk["targetAreas"], 50)
    assert np.abs((s.get_positions(0) - xo + 0.5) % 1.0 - 0.5).max() < 1e-13 and abs(int(it[0]) - ito) <= 1


@pytest.mark.parametrize("flag", [0, 0x400])
def test_buckets_beyond_the_lds_stage(flag):
    """Two noisy 3000-gons crossing 1444 times (722 records per polygon, where the earlier
    design stopped at 512): the buckets grow past what k_exterior_poly can stage in LDS
    (Kp * 116 bytes > 64 KiB), so the exterior forces stage the records in global memory.
    The hybrid model (overlap forces with zero springs, no projection) shows the exterior
    forces themselves: records, energy and forces equal the oracle's bit for bit, with the
    reference energy and with squared pair areas."""
    pk = _two_noisy_polygons(3000, 0.05, 5)
    start, pos = pk["startIndices"], pk["positions"]
    nb, rec = oracle_records(pk, 2)
    assert rec["numIntersections"] == 1444 and list(rec["shapeCounts"]) == [722, 722]
    print(f"[buckets] 2 x 3000-gons, flag {flag:#x}", flush=True)
    s = capi.System([pk], [2], model=capi.HYBRID | flag)
    s.force_eval()
    with pytest.raises(capi.PcpError, match="overflow"):
        s.sync()
    s.force_eval()
    s.sync()
    r = s.get_records(0)
    for key in ("intersections", "shapeCounts"):
        np.testing.assert_array_equal(r[key], rec[key], err_msg=key)
    if flag == 0:
        for key in ("outersections", "tu", "ut"):
            np.testing.assert_array_equal(r[key], rec[key], err_msg=key)
    F, E = s.get_forces(0), s.get_energy()[0]
    Eo, Fo, nIo = O.System(start, 2, O.HYBRID | flag, pk["targetEdgeV"], pk["targetAreas"]).force_eval(pos)
    assert nIo == 1444 and E == Eo and np.array_equal(F, Fo)
    s.close()


def test_packing_without_overlaps():
    """A dilute packing (phi = 0.3, no two centres closer than the sum of the radii): no
    crossings, zero records, zero energy and force on the GPU as in the oracle, and FIRE
    stops after its first step with max|F| = 0."""
    p = make_packing(64, 12, 0.3, seed=1, min_center_frac=1.0)
    B = box_size_for(p["diameter"])
    nb, rec = oracle_records(p, B)
    assert rec["numIntersections"] == 0
    for model in (capi.NORMAL, capi.NORMAL | SQUARED, capi.NORMAL | PAIR):
        s = capi.System([p], [B], model=model)
        s.force_eval()
        s.sync()
        assert s.get_records(0)["numIntersections"] == 0
        F, E = s.get_forces(0), s.get_energy()[0]
        Eo, Fo, nIo = O.System(p["startIndices"], B, model, p["targetEdgeV"], p["targetAreas"]).force_eval(
            p["positions"])
        assert nIo == 0 and E == Eo == 0.0 and np.array_equal(F, Fo) and not F.any()
        prm = s.fire_params(fTol=1e-10, dtMax=100.0)
        s.fire_init(prm, 1e-3)
        st = s.fire_minimize(prm, 10)[0]
        assert st.converged and st.steps == 1 and st.maxForce == 0.0


def test_model_overlap_energy_selection():
    """Model.setOverlapEnergy (added to the reference's Model API; no bound name removed):
    with "squaredPair", minimizeFIRE through the reference-style calls reaches
    max|F| <= 1e-10 in the same number of steps and at the same energy, bit for bit, as the
    C-ABI FIRE of the same packing and cell grid; unknown names raise ValueError and
    "reference" restores the reference's energy."""
    from pycudapolygon_amd.pyCudaPolygonLink import libpyCudaPolygon as lpcp
    p = make_packing(64, 12, 0.80, seed=0, min_center_frac=0.5)
    B = box_size_for(p["diameter"])
    m = lpcp.Model(0)
    V = int(p["startIndices"][-1])
    m.setNumVertices(V)
    m.setStartIndices(p["startIndices"].tolist())
    m.setPositions(p["positions"].tolist())
    m.setMaxEdgeLength(1.0 / B)
    m.setTargetEdgeLengthsPerEdge(p["targetEdgeV"].tolist())
    m.setTargetAreas(p["targetAreas"].tolist())
    m.setModelEnum(lpcp.modelEnum.normal)
    m.initializeNeighborCells()
    assert m.getOverlapEnergy() == "reference"
    with pytest.raises(ValueError):
        m.setOverlapEnergy("lens")
    m.setOverlapEnergy("squaredPair")
    assert m.getOverlapEnergy() == "squaredPair"
    E, dt, steps = m.minimizeFIRE(1e-10, 1e-3, 5000, 100.0, 0.1, 0.99, 1.1, 0.5, 5, 5)
    assert np.abs(np.array(m.getForces())).max() <= 1e-10 and steps < 5000
    s = capi.System([p], [B], model=capi.NORMAL | PAIR)
    prm = s.fire_params(fTol=1e-10, dtMax=100.0)
    s.fire_init(prm, 1e-3)
    st = s.fire_minimize(prm, 5000)[0]
    print(f"[model] squaredPair: Model {steps} steps E {E!r}; C ABI {st.steps} steps E {st.energy!r}", flush=True)
    assert st.converged and steps == st.steps and E == st.energy
    np.testing.assert_array_equal(np.array(m.getPositions()), s.get_positions(0))
    m.setOverlapEnergy("reference")
    assert m.getOverlapEnergy() == "reference"
