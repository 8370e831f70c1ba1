"""CPU tests that pin the oracle (oracle/pcp_oracle.c) before it is trusted as the
checker: the reference's own fixture and recorded outputs, closed forms derived from
plane geometry (computed here with independent numpy code), brute force on small
inputs, and invariances.  No GPU needed."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from pycudapolygon_amd.generate import make_packing, box_size_for
from tests import _geom as G

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture():
    with open(os.path.join(HERE, "golden", "testsaver.json")) as f:
        d = json.load(f)
    start = np.concatenate([[0], np.cumsum(d["nArray"])]).astype(np.int32)
    return d, start, np.array(d["positions"])


def neighbor_dict(nb):
    M, num = nb["maxNeighbors"], nb["numNeighbors"]
    out = {}
    for i in range(len(num)):
        if num[i]:
            out[i] = (nb["neighbors"][i * M:i * M + num[i]].tolist(),
                      [bool(x) for x in nb["inside"][i * M:i * M + num[i]]])
    return out


# ---------------------------------------------------------------- reference fixture
def test_fixture_neighbors_match_recorded_notebook_output():
    # tests/neighborTesting.ipynb: setMaxEdgeLength(0.25) -> {1: [7, 8]}, inside [True, False]
    d, start, pos = fixture()
    B = int(np.floor(1.0 / d["maxEdgeLength_notebook"]))
    nd = neighbor_dict(O.neighbors(pos, start, B))
    assert nd == {1: ([7, 8], [True, False])}


def test_fixture_neighbors_idempotent_and_translation_invariant():
    # neighborTesting.ipynb cells 9 and 12-14: repeated updates and a periodic shift
    d, start, pos = fixture()
    B = 4
    ref = neighbor_dict(O.neighbors(pos, start, B))
    for _ in range(5):
        assert neighbor_dict(O.neighbors(pos, start, B)) == ref
    p = pos.reshape(-1, 2)
    shifted = ((p - p.mean(axis=0)) % 1.0).reshape(-1)
    got = neighbor_dict(O.neighbors(shifted, start, B))
    assert {k: v[0] for k, v in got.items()} == {k: v[0] for k, v in ref.items()}


def test_fixture_crossing_parameters_brute_force():
    # edge 1 crosses edges 7 and 8; (t, u) from an independent segment intersection
    d, start, pos = fixture()
    P = pos.reshape(-1, 2)
    nb = O.neighbors(pos, start, 4)
    M = nb["maxNeighbors"]
    for c, other in enumerate([7, 8]):
        a, b = P[1], P[2]
        q, r = P[other], P[other + 1 if other + 1 < 10 else 5]
        den = np.cross(b - a, r - q)
        t = np.cross(q - a, r - q) / den
        u = np.cross(q - a, b - a) / den
        slot = 1 * M + c
        got = nb["tuSlots"][2 * slot:2 * slot + 2]
        exp = (t, u) if den > 0 else (u, t)
        np.testing.assert_allclose(got, exp, rtol=0, atol=1e-14)


def test_record_parameters_are_float_rounded_slot_parameters():
    # compaction narrows the double slot parameters to float (kernels.cuh:857-858):
    # as multisets, the record (t, u) equal the float32 rounding of the slot values
    from pycudapolygon_amd.generate import make_packing, box_size_for
    p = make_packing(64, 12, 0.80, seed=1)
    B = box_size_for(p["maxEdgeLength"])
    nb = O.neighbors(p["positions"], p["startIndices"], B)
    rec = O.outersections(p["startIndices"], nb)
    M = nb["maxNeighbors"]
    slots = []
    for v, cnt in enumerate(nb["numNeighbors"]):
        for c in range(cnt):
            slots.append(nb["tuSlots"][2 * (v * M + c):2 * (v * M + c) + 2])
    slots = np.array(slots)
    tu = np.asarray(rec["tu"]).reshape(-1, 2)
    assert len(tu) == rec["numIntersections"] == len(slots) > 0
    assert np.all(tu.astype(np.float32).astype(np.float64) == tu)
    assert not np.all(slots.astype(np.float32).astype(np.float64) == slots)
    want = np.sort(slots.astype(np.float32).astype(np.float64).view(np.complex128).ravel())
    got = np.sort(np.concatenate([tu[:, 0] + 1j * tu[:, 1]]))
    np.testing.assert_array_equal(got, want)


# ---------------------------------------------------------------- closed forms
def _raw(polys, B=2):
    start, pos = G.to_model(polys)
    g = O.geometry(pos, start)
    rec = O.outersections(start, O.neighbors(pos, start, B))
    E, f = O.force_energy("normal", pos, start, rec, g)
    return E, f, rec


def test_two_axis_aligned_squares_have_zero_energy():
    # E = A + (Qx-Px)(Qy-Py): overlap [1,2]^2 with crossings (1,2),(2,1) gives 1 - 1 = 0
    s, off = 0.05, np.array([0.3, 0.3])
    A = np.array([[0, 0], [2, 0], [2, 2], [0, 2]]) * s + off
    Bq = np.array([[1, 1], [3, 1], [3, 3], [1, 3]]) * s + off
    for B in (1, 2, 7):
        E, f, rec = _raw([A, Bq], B)
        assert rec["numIntersections"] == 2
        assert abs(E) < 1e-17


def test_rotated_square_energy_equals_overlap_area():
    # diamond over a square so that the crossings share an x coordinate: E = overlap area
    s, c = 0.05, np.array([0.5, 0.5])
    sq = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]]) * s + c
    dia = np.array([[0, -1], [1, 0], [0, 1], [-1, 0]]) * s * 1.3 + c + np.array([1.4 * s, 0.0])
    E, f, rec = _raw([sq, dia])
    X = G.crossings(sq, dia)
    assert len(X) == 2 and abs(X[0][0] - X[1][0]) < 1e-15
    # crossing parameters are stored as float (kernels.cuh:857-858), so the energy
    # carries a relative error of order 2^-24 of the lens geometry
    np.testing.assert_allclose(E, G.overlap_area(sq, dia), rtol=1e-6)


@pytest.mark.parametrize("seed", range(6))
def test_single_lens_energy_and_force_closed_form(seed):
    """One lens between convex polygons with crossings a, b (independent numpy geometry):
    energy E = A_ov + (bx-ax)(by-ay); force (start point frozen, forces not halved,
    kernels.cuh:1143-1144 vs :1271) F = -2 grad A_ov - 2 (bx-ax) grad (by-ay)."""
    rng = np.random.default_rng(seed)
    n1, n2 = rng.integers(5, 12), rng.integers(5, 12)
    ang = rng.uniform(0, 2 * np.pi)
    d = rng.uniform(0.06, 0.09)
    p1 = G.regular(n1, 0.06, (0.4, 0.4), rng.uniform(0, 6))
    p2 = G.regular(n2, 0.05, (0.4 + d * np.cos(ang), 0.4 + d * np.sin(ang)), rng.uniform(0, 6))
    X = G.crossings(p1, p2)
    assert len(X) == 2
    a, b = X
    E, f, rec = _raw([p1, p2])
    # float crossing parameters (kernels.cuh:857-858): relative error of order 2^-24
    np.testing.assert_allclose(E, G.overlap_area(p1, p2) + (b[0] - a[0]) * (b[1] - a[1]),
                               rtol=1e-6, atol=1e-17)
    x0 = np.concatenate([p1.reshape(-1), p2.reshape(-1)])
    k1 = 2 * n1

    def split(x):
        return x[:k1].reshape(-1, 2), x[k1:].reshape(-1, 2)

    h = 1e-7
    fe = np.zeros_like(x0)
    for k in range(len(x0)):
        xp, xm = x0.copy(), x0.copy()
        xp[k] += h
        xm[k] -= h
        Pp, Pm = split(xp), split(xm)
        dA = (G.overlap_area(*Pp) - G.overlap_area(*Pm)) / (2 * h)
        ap, bp = sorted(G.crossings(*Pp), key=lambda q: np.linalg.norm(q - a))
        am, bm = sorted(G.crossings(*Pm), key=lambda q: np.linalg.norm(q - a))
        dDy = ((bp[1] - ap[1]) - (bm[1] - am[1])) / (2 * h)
        fe[k] = -2 * dA - 2 * (b[0] - a[0]) * dDy
    np.testing.assert_allclose(f, fe, rtol=0, atol=2e-7 * np.abs(fe).max())


def test_records_pair_mutually_for_lenses():
    rng = np.random.default_rng(3)
    polys = [G.regular(8, 0.05, (0.2 + 0.6 * i / 3, 0.5), rng.uniform(0, 6)) for i in range(4)]
    polys += [G.regular(6, 0.04, (0.2 + 0.6 * i / 3 + 0.06, 0.52), rng.uniform(0, 6)) for i in range(4)]
    start, pos = G.to_model(polys)
    rec = O.outersections(start, O.neighbors(pos, start, 3))
    inter, outer = rec["intersections"], rec["outersections"]
    idx = {int(k): r for r, k in enumerate(inter)}
    assert rec["numIntersections"] == 8
    for r in range(len(inter)):
        k, o = int(inter[r]), int(outer[r])
        assert (o >> 48, (o >> 32) & 0xFFFF) == ((k >> 32) & 0xFFFF, k >> 48)
        assert int(outer[idx[o]]) == k


# ---------------------------------------------------------------- record order / cells
def test_records_sorted_by_si_i_j_sj_and_cell_size_independent():
    p = make_packing(64, 12, 0.8, seed=5)
    start, pos = p["startIndices"], p["positions"]
    recs = []
    for B in (1, 2, 5, box_size_for(p["maxEdgeLength"]), box_size_for(p["diameter"])):
        recs.append(O.outersections(start, O.neighbors(pos, start, B)))
    k = recs[0]["intersections"].astype(np.uint64)
    order = [((int(x) >> 32) & 0xFFFF, (int(x) >> 16) & 0xFFFF, int(x) & 0xFFFF, int(x) >> 48) for x in k]
    assert order == sorted(order) and len(set(order)) == len(order)
    for r in recs[1:]:
        for key in ("intersections", "outersections", "tu", "ut"):
            np.testing.assert_array_equal(r[key], recs[0][key])


def test_cells_are_a_stable_counting_sort():
    p = make_packing(64, 12, 0.8, seed=2)
    start, pos = p["startIndices"], p["positions"]
    B = 9
    cl, ni, cpb = O.cells(pos, start, B)
    np.testing.assert_array_equal(ni, np.argsort(cl, kind="stable"))
    np.testing.assert_array_equal(cpb, np.searchsorted(np.sort(cl), np.arange(B * B), side="left"))
    # midpoint cell (brute force in numpy)
    P = pos.reshape(-1, 2)
    nxt = np.array([start[np.searchsorted(start, v, side="right") - 1] if v + 1 in start[1:] else v + 1
                    for v in range(len(P))])
    d = (P[nxt] - P + 0.5) % 1.0 - 0.5
    mid = (P + d / 2) % 1.0
    cell = np.floor(mid[:, 1] * B).astype(int) * B + np.floor(mid[:, 0] * B).astype(int)
    assert (cell == cl).mean() > 0.999


def test_neighbor_brute_force_all_pairs():
    p = make_packing(24, 8, 0.85, seed=11)
    start, pos = p["startIndices"], p["positions"]
    V = int(start[-1])
    P = pos.reshape(-1, 2)
    sid = np.repeat(np.arange(len(start) - 1), np.diff(start))
    nxt = np.array([start[sid[v]] if v + 1 == start[sid[v] + 1] else v + 1 for v in range(V)])
    wrap = lambda x: (x + 1.5) % 1.0 - 0.5
    brute = set()
    for i in range(V):
        for j in range(i + 1, V):
            if j == nxt[i] or nxt[j] == i or nxt[j] == nxt[i]:
                continue
            r, s, g = wrap(P[nxt[i]] - P[i]), wrap(P[nxt[j]] - P[j]), wrap(P[j] - P[i])
            den = r[0] * s[1] - r[1] * s[0]
            if abs(den) < 1e-12:
                continue
            t = (g[0] * s[1] - g[1] * s[0]) / den
            u = (g[0] * r[1] - g[1] * r[0]) / den
            if 0 < t < 1 and 0 < u < 1:
                brute.add((i, j))
    nb = O.neighbors(pos, start, 3)
    got = {(i, int(n)) for i, (ns, _) in neighbor_dict(nb).items() for n in ns}
    assert got == brute and len(brute) > 10


# ---------------------------------------------------------------- invariances
def test_energy_invariant_under_cyclic_relabeling():
    p = make_packing(32, 10, 0.85, seed=4)
    start, pos = p["startIndices"], p["positions"]
    B = box_size_for(p["maxEdgeLength"])
    E0 = O.System(start, B).force_eval(pos)[0]
    P = pos.reshape(-1, 2).copy()
    rng = np.random.default_rng(0)
    for s in range(len(start) - 1):
        a, b = start[s], start[s + 1]
        P[a:b] = np.roll(P[a:b], rng.integers(0, b - a), axis=0)
    E1 = O.System(start, B).force_eval(P.reshape(-1))[0]
    np.testing.assert_allclose(E1, E0, rtol=1e-11)


# ---------------------------------------------------------------- projection / SHAKE
def _edge_area_grads(pos, start, s):
    st, en = start[s], start[s + 1]
    n = en - st
    P = pos.reshape(-1, 2)[st:en]
    wrap = lambda x: (x + 1.5) % 1.0 - 0.5
    A = np.zeros((2 * n, n + 1))
    for e in range(n):
        d = wrap(P[(e + 1) % n] - P[e])
        t = d / np.linalg.norm(d)
        A[2 * e:2 * e + 2, e] -= t
        A[2 * ((e + 1) % n):2 * ((e + 1) % n) + 2, e] += t
    for k in range(n):
        A[2 * k, n] = wrap(P[(k + 1) % n][1] - P[k - 1][1])
        A[2 * k + 1, n] = wrap(P[k - 1][0] - P[(k + 1) % n][0])
    return A


def test_projection_removes_constraint_directions_and_is_idempotent():
    p = make_packing(16, 9, 0.8, seed=7)
    start, pos = p["startIndices"], p["positions"]
    rng = np.random.default_rng(1)
    f = rng.normal(size=len(pos))
    g = O.project_force(pos, start, f)
    np.testing.assert_allclose(O.project_force(pos, start, g), g, atol=1e-13)
    for s in range(len(start) - 1):
        A = _edge_area_grads(pos, start, s)
        seg = g[2 * start[s]:2 * start[s + 1]]
        assert np.abs(A.T @ seg).max() < 1e-12
        # rigid translation survives
        t = np.tile([1.0, 0.0], start[s + 1] - start[s])
        pt = O.project_force(pos, start, np.tile([1.0, 0.0], len(pos) // 2))[2 * start[s]:2 * start[s + 1]]
        np.testing.assert_allclose(pt, t, atol=1e-12)


def test_shake_satisfies_constraints():
    p = make_packing(16, 12, 0.8, seed=3)
    start, pos = p["startIndices"], p["positions"]
    rng = np.random.default_rng(2)
    noisy = (pos + rng.normal(scale=2e-4, size=pos.shape)) % 1.0
    out, it = O.shake(noisy, start, p["targetEdgeV"], p["targetAreas"], 50)
    assert 0 < it < 50
    g = O.geometry(out, start)
    np.testing.assert_allclose(g["edgeLengths"], p["targetEdgeV"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(g["areas"], p["targetAreas"], rtol=0, atol=1e-15)
    # already satisfied: zero iterations, no motion
    out2, it2 = O.shake(out, start, p["targetEdgeV"], p["targetAreas"], 50)
    assert it2 == 0 and np.array_equal(out2, out)


@pytest.mark.parametrize("n", [6, 7, 8, 12, 16, 23, 32, 64])
def test_lane_order_solves_equal_dense_solves(n):
    """The lane-order projection and SHAKE (parallel cyclic reduction, butterfly sums:
    the HIP kernels' operation order) against the independent dense solvers (Householder
    QR; Gauss-Newton with Gaussian elimination): the same projected force to 1e-12
    relative, the same SHAKE positions to 1e-13, constraints met, same iteration counts,
    and the lane-order projection is idempotent and keeps rigid translations.  A wrong
    index, sign or dropped Schur term in the reduction fails the dense comparison.
    (Near-regular 4- and 5-gons are left out: they are close to the cyclic polygon, where
    the area is stationary under the polygon's one or two flexes, so the area row is
    nearly dependent on the edge rows and two exact methods differ by up to ~1e-9.)"""
    p = make_packing(32, n, 0.8, seed=n)
    start, pos = p["startIndices"], p["positions"]
    rng = np.random.default_rng(n)
    f = rng.normal(size=len(pos))
    fd = O.project_force(pos, start, f)
    fl = O.project_force_lanes(pos, start, f)
    assert np.abs(fl - fd).max() <= 1e-12 * np.abs(f).max()
    np.testing.assert_allclose(O.project_force_lanes(pos, start, fl), fl, rtol=0, atol=1e-13 * np.abs(f).max())
    tr = np.tile([1.0, 0.0], len(pos) // 2)
    np.testing.assert_allclose(O.project_force_lanes(pos, start, tr), tr, rtol=0, atol=1e-12)
    noisy = (pos + rng.normal(scale=2e-4, size=pos.shape)) % 1.0
    xd, itd = O.shake(noisy, start, p["targetEdgeV"], p["targetAreas"], 50)
    xl, itl = O.shake_lanes(noisy, start, p["targetEdgeV"], p["targetAreas"], 50)
    d = (xl - xd + 0.5) % 1.0 - 0.5
    assert np.abs(d).max() < 1e-13 and 0 < itl < 50 and abs(itl - itd) <= 1
    g = O.geometry(xl, start)
    np.testing.assert_allclose(g["edgeLengths"], p["targetEdgeV"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(g["areas"], p["targetAreas"], rtol=0, atol=1e-15)
    x2, it2 = O.shake_lanes(xl, start, p["targetEdgeV"], p["targetAreas"], 50)
    assert it2 == 0 and np.array_equal(x2, xl)


def test_lane_order_solves_reject_polygons_beyond_64():
    p = make_packing(4, 65, 0.5, seed=0)
    with pytest.raises(ValueError):
        O.project_force_lanes(p["positions"], p["startIndices"], np.zeros(len(p["positions"])))
    with pytest.raises(ValueError):
        O.shake_lanes(p["positions"], p["startIndices"], p["targetEdgeV"], p["targetAreas"], 5)


def test_fire_with_lane_order_solves_tracks_dense_fire():
    """The FIRE oracle with either solver pair: same step decisions and positions within
    1e-12 over the first 30 steps (squared lens areas, rigid polygons)."""
    p = make_packing(64, 12, 0.80, seed=0, min_center_frac=0.5)
    B = box_size_for(p["maxEdgeLength"])
    runs = []
    for solver in ("dense", "lanes"):
        sysO = O.System(p["startIndices"], B, 0x200, p["targetEdgeV"], p["targetAreas"], solver=solver)
        pos = p["positions"].copy()
        st, vel, f = sysO.fire_init(pos, 1e-3, 0.1)
        hist = []
        for _ in range(30):
            sysO.fire_step(pos, vel, f, st, dtMax=100.0)
            hist.append((st.rolledBack, st.shakeFailed, st.nPos, st.dt, pos.copy(), st.energy))
        runs.append(hist)
    for a, b in zip(*runs):
        assert a[:4] == b[:4]
        assert np.abs((a[4] - b[4] + 0.5) % 1.0 - 0.5).max() < 1e-12
        assert abs(a[5] - b[5]) <= 1e-9 * abs(b[5])


@pytest.mark.parametrize("seed", range(4))
def test_squared_lens_energy_closed_form(seed):
    """Squared lens areas on one lens between convex polygons: E = A^2/2 with A the
    overlap area of an independent clipping geometry, F = -grad E by central
    differences of that geometry."""
    SQUARED = 0x200
    rng = np.random.default_rng(seed)
    n1, n2 = rng.integers(5, 12), rng.integers(5, 12)
    ang = rng.uniform(0, 2 * np.pi)
    d = rng.uniform(0.06, 0.09)
    p1 = G.regular(n1, 0.06, (0.4, 0.4), rng.uniform(0, 6))
    p2 = G.regular(n2, 0.05, (0.4 + d * np.cos(ang), 0.4 + d * np.sin(ang)), rng.uniform(0, 6))
    assert len(G.crossings(p1, p2)) == 2
    start, pos = G.to_model([p1, p2])
    sysO = O.System(start, 2, SQUARED | 5, np.zeros(n1 + n2), np.zeros(2))   # hybrid: no projection
    E, f, nI = sysO.force_eval(pos)
    A = G.overlap_area(p1, p2)
    np.testing.assert_allclose(E, A * A / 2, rtol=1e-10)
    k1 = 2 * n1
    h = 1e-7
    fe = np.zeros_like(pos)
    for q in range(len(pos)):
        xp, xm = pos.copy(), pos.copy()
        xp[q] += h
        xm[q] -= h
        Ap = G.overlap_area(xp[:k1].reshape(-1, 2), xp[k1:].reshape(-1, 2))
        Am = G.overlap_area(xm[:k1].reshape(-1, 2), xm[k1:].reshape(-1, 2))
        fe[q] = -(Ap * Ap / 2 - Am * Am / 2) / (2 * h)
    np.testing.assert_allclose(f, fe, rtol=0, atol=1e-6 * np.abs(fe).max())


def _u_and_bar(yc, th=0.05):
    """A U-shaped 8-gon (arms x in [0,1] and [2,3], y in [1,3], base y in [0,1], unit
    s = 0.02) and a slightly rotated bar centred at height yc (in units of s) across both
    arms: the bar overlaps U in two separate regions while its lower edge is above the
    notch bottom (y = 1), in one region after it passes below."""
    s, off = 0.02, np.array([0.4, 0.4])
    U = np.array([[0, 0], [3, 0], [3, 3], [2, 3], [2, 1], [1, 1], [1, 3], [0, 3]], float) * s + off
    hw, hh = 2.2 * s, 0.5 * s
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    bar = off + np.array([1.5 * s, yc * s]) + np.array([[-hw, -hh], [hw, -hh], [hw, hh], [-hw, hh]]) @ R.T
    arms = [np.array([[0, 1], [1, 1], [1, 3], [0, 3]], float) * s + off,
            np.array([[2, 1], [3, 1], [3, 3], [2, 3]], float) * s + off]
    return U, bar, arms


def test_pair_squared_energy_is_the_squared_pair_overlap_through_a_merge():
    """Squared pair areas: E = (A1 + A2)^2/2 with A the overlap of an independent clipping
    geometry, for two separate lenses of one polygon pair and after they merge into one
    as the bar slides down (the energy is continuous through the merge).  Squared lens
    areas give (A1^2 + A2^2)/2 before the merge and A^2/2 after: they jump by A1*A2."""
    for yc in np.linspace(3.1, 1.2, 39):
        U, bar, arms = _u_and_bar(yc)
        start, pos = G.to_model([U, bar])
        A = G.overlap_area(U, bar)
        E = {}
        for flag in (0x200, 0x400):
            sysO = O.System(start, 2, flag | 5, np.zeros(12), np.zeros(2))   # hybrid: no projection
            E[flag] = sysO.force_eval(pos)[0]
        np.testing.assert_allclose(E[0x400], A * A / 2, rtol=1e-10)
        if yc - 0.5 - 0.11 > 1.0:   # the bar's lower edge (rotation included) above the notch
            A1, A2 = (G.overlap_area(a, bar) for a in arms)
            np.testing.assert_allclose(A1 + A2, A, rtol=1e-12)
            np.testing.assert_allclose(E[0x200], (A1 * A1 + A2 * A2) / 2, rtol=1e-10)
            assert E[0x200] < 0.51 * E[0x400]
        elif yc - 0.5 + 0.11 < 1.0:   # the whole lower edge below the notch: merged
            np.testing.assert_allclose(E[0x200], A * A / 2, rtol=1e-10)


def test_pair_squared_force_is_minus_gradient_with_two_lenses():
    """Two lenses of one pair (the U and the bar at yc = 2.0, 8 crossings): F = -grad of
    (overlap area)^2/2 by central differences of the clipping geometry."""
    U, bar, _ = _u_and_bar(2.0)
    start, pos = G.to_model([U, bar])
    sysO = O.System(start, 2, 0x400 | 5, np.zeros(12), np.zeros(2))
    E, f, nI = sysO.force_eval(pos)
    assert nI == 8
    h = 1e-8
    fe = np.zeros_like(pos)
    for q in range(len(pos)):
        xp, xm = pos.copy(), pos.copy()
        xp[q] += h
        xm[q] -= h
        Ap = G.overlap_area(xp[:16].reshape(-1, 2), xp[16:].reshape(-1, 2))
        Am = G.overlap_area(xm[:16].reshape(-1, 2), xm[16:].reshape(-1, 2))
        fe[q] = -(Ap * Ap / 2 - Am * Am / 2) / (2 * h)
    np.testing.assert_allclose(f, fe, rtol=0, atol=1e-6 * np.abs(fe).max())


@pytest.mark.parametrize("seed", range(3))
def test_pair_squared_equals_lens_squared_for_single_lens_pairs(seed):
    """With one lens per polygon pair the two squared energies are the same numbers,
    bit for bit (the pair sum of one lens is that lens's area)."""
    rng = np.random.default_rng(seed)
    p1 = G.regular(int(rng.integers(5, 12)), 0.06, (0.4, 0.4), rng.uniform(0, 6))
    p2 = G.regular(int(rng.integers(5, 12)), 0.05, (0.47, 0.42), rng.uniform(0, 6))
    start, pos = G.to_model([p1, p2])
    V = len(pos) // 2
    a = O.System(start, 2, 0x200 | 5, np.zeros(V), np.zeros(2)).force_eval(pos)
    b = O.System(start, 2, 0x400 | 5, np.zeros(V), np.zeros(2)).force_eval(pos)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and a[2] == b[2] == 2


def test_squared_lens_fire_converges_where_area_energy_stalls():
    """C1 (N=64, phi=0.80, no polygon starting inside another): FIRE on the area energy
    stalls at a kink (flush contacts, rollbacks shrink dt); on squared lens areas it
    reaches max|F| <= 1e-10."""
    from pycudapolygon_amd.generate import make_packing, box_size_for
    p = make_packing(64, 12, 0.80, seed=0, min_center_frac=0.5)
    B = box_size_for(p["maxEdgeLength"])
    res = {}
    for flag in (0x100, 0x200):
        sysO = O.System(p["startIndices"], B, flag, p["targetEdgeV"], p["targetAreas"])
        x, f, st, steps = sysO.fire_minimize(p["positions"], fTol=1e-10, maxSteps=3000)
        res[flag] = (np.abs(f).max(), st.dt)
    assert res[0x200][0] <= 1e-10
    assert res[0x100][0] > 1e-6 and res[0x100][1] < 1e-4


def test_soft_body_penalties_closed_form():
    """softBody = edge springs + area springs only (kernels.cuh:1284-1352):
    E = sum_m k (l_m - l0_m)^2 / 2 + sum_s c (A_s - A0_s)^2 / 2 (A0 >= 1e-14), computed here
    from an independent numpy geometry, and F = -grad E by central differences."""
    rng = np.random.default_rng(5)
    polys = [G.regular(int(n), 0.05, (0.2 + 0.3 * k, 0.5), rng.uniform(0, 6)) for k, n in enumerate((5, 7, 9))]
    polys = [P + rng.normal(scale=3e-3, size=P.shape) for P in polys]
    start, pos = G.to_model(polys)
    V = int(start[-1])
    l0 = rng.uniform(0.02, 0.04, V)
    A0 = np.array([G.shoelace(P) * f for P, f in zip(polys, (0.9, 1.1, 1.0))])
    k, c = 3.0, 7.0

    def energy_np(x):
        e = 0.0
        for s, P in enumerate(np.split(x.reshape(-1, 2), start[1:-1])):
            d = np.roll(P, -1, axis=0) - P
            e += (k * (np.linalg.norm(d, axis=1) - l0[start[s]:start[s + 1]]) ** 2 / 2).sum()
            e += c * (G.shoelace(P) - A0[s]) ** 2 / 2
        return e

    def oracle(x):
        g = O.geometry(x, start)
        return O.force_energy("softBody", x, start, None, g, l0, A0, k, c)

    E, f = oracle(pos)
    np.testing.assert_allclose(E, energy_np(pos), rtol=1e-12)
    h = 1e-7
    fe = np.zeros_like(pos)
    for q in range(len(pos)):
        xp, xm = pos.copy(), pos.copy()
        xp[q] += h
        xm[q] -= h
        fe[q] = -(energy_np(xp) - energy_np(xm)) / (2 * h)
    np.testing.assert_allclose(f, fe, rtol=0, atol=1e-6 * np.abs(fe).max())


def test_det_sum_is_a_fixed_tree():
    rng = np.random.default_rng(0)
    for n in (1, 7, 256, 257, 5000, 70000):
        x = rng.normal(size=n)
        np.testing.assert_allclose(O.det_sum(x), x.sum(), rtol=1e-12, atol=1e-12)
        ints = rng.integers(-1000, 1000, n).astype(float)
        assert O.det_sum(ints) == ints.sum()


def test_geometry_closed_forms():
    polys = [G.regular(n, 0.05, (0.3 + 0.1 * k, 0.5), 0.3 * k) for k, n in enumerate([3, 5, 8, 16])]
    start, pos = G.to_model(polys)
    g = O.geometry(pos, start)
    for k, n in enumerate([3, 5, 8, 16]):
        np.testing.assert_allclose(g["areas"][k], 0.5 * n * 0.05 ** 2 * np.sin(2 * np.pi / n), rtol=1e-12)
        np.testing.assert_allclose(g["edgeLengths"][start[k]:start[k + 1]], 2 * 0.05 * np.sin(np.pi / n),
                                   rtol=1e-12)
        np.testing.assert_allclose([g["comX"][k], g["comY"][k]], [0.3 + 0.1 * k, 0.5], atol=1e-15)


def test_overlap_counter_estimates_area_of_two_discs():
    # updateOverlapAreaKernel grid estimator vs the exact overlap of two 64-gons
    p1 = G.regular(64, 0.1, (0.45, 0.5))
    p2 = G.regular(64, 0.1, (0.55, 0.5))
    start, pos = G.to_model([p1, p2])
    s, cnt = O.overlap_count(pos, start, 2, 400)
    np.testing.assert_allclose(s / 400 ** 2, G.overlap_area(p1, p2), rtol=2e-2)


AREA = 0x100


def _two_octagons():
    polys = [G.regular(8, 0.06, (0.4, 0.4), 0.1), G.regular(8, 0.06, (0.49, 0.42), 0.3)]
    start, pos = G.to_model(polys)
    tev = O.per_vertex_targets(start, [0.06 * 2 * np.sin(np.pi / 8)] * 2)
    return polys, start, pos, tev, [G.shoelace(polys[0]), G.shoelace(polys[1])]


def test_reference_energy_fire_stalls_on_a_single_lens():
    """Documented behaviour of the reference convention: its energy A + dx*dy is not the
    potential of its force, FIRE's energy test rejects steps, dt collapses."""
    polys, start, pos, tev, ta = _two_octagons()
    x, f, st, steps = O.System(start, 2, "normal", tev, ta).fire_minimize(pos, fTol=1e-10, maxSteps=3000)
    assert steps == 3000 and np.abs(f).max() > 1e-6 and st.dt < 1e-6


@pytest.mark.parametrize("seed", range(4))
def test_area_energy_is_overlap_area_and_force_its_gradient(seed):
    rng = np.random.default_rng(seed)
    n1, n2 = rng.integers(5, 12), rng.integers(5, 12)
    ang = rng.uniform(0, 2 * np.pi)
    p1 = G.regular(n1, 0.06, (0.4, 0.4), rng.uniform(0, 6))
    p2 = G.regular(n2, 0.05, (0.4 + 0.075 * np.cos(ang), 0.4 + 0.075 * np.sin(ang)), rng.uniform(0, 6))
    start, pos = G.to_model([p1, p2])
    E, f, nI = O.System(start, 2, AREA | 5).force_eval(pos)   # hybrid, zero springs: no projection
    np.testing.assert_allclose(E, G.overlap_area(p1, p2), rtol=1e-11)
    x0 = np.concatenate([p1.reshape(-1), p2.reshape(-1)])
    h, k1 = 1e-7, 2 * n1
    fd = np.zeros_like(x0)
    for k in range(len(x0)):
        xp, xm = x0.copy(), x0.copy()
        xp[k] += h
        xm[k] -= h
        fd[k] = -(G.overlap_area(xp[:k1].reshape(-1, 2), xp[k1:].reshape(-1, 2)) -
                  G.overlap_area(xm[:k1].reshape(-1, 2), xm[k1:].reshape(-1, 2))) / (2 * h)
    np.testing.assert_allclose(f, fd, rtol=0, atol=2e-7 * np.abs(fd).max())


def test_area_energy_fire_separates_two_polygons():
    polys, start, pos, tev, ta = _two_octagons()
    E0 = O.System(start, 2, AREA, tev, ta).force_eval(pos)[0]
    x, f, st, steps = O.System(start, 2, AREA, tev, ta).fire_minimize(pos, fTol=1e-10, maxSteps=20000)
    assert np.abs(f).max() <= 1e-10 and st.energy == 0.0 and E0 > 0 and steps < 20000


def _rest_and_moved(seed, n, c0, theta, scale, noise=0.0):
    """A noisy n-gon rest shape r (offsets from its vertex mean) and the polygon
    c0 + scale * R(theta) r (+ Gaussian vertex noise), wrapped into the unit box."""
    rng = np.random.default_rng(seed)
    P = G.regular(n, 0.05, (0.0, 0.0), rng.uniform(0, 6))
    P = P * (1 + rng.uniform(-0.05, 0.05, n))[:, None]
    r = P - P.mean(axis=0)
    R = np.array([[np.cos(theta), -np.sin(theta)], [np.sin(theta), np.cos(theta)]])
    x = np.asarray(c0) + scale * r @ R.T + rng.normal(scale=noise, size=r.shape)
    return r, x % 1.0


def test_shape_projection_keeps_a_moved_rest_shape():
    """A rest shape turned, scaled and moved (across the periodic boundary) projects onto
    itself: with the target area of the scaled shape the result equals the input to
    1e-15, its area is the target and a second projection changes nothing."""
    for seed, (c0, th, sc) in enumerate([((0.5, 0.5), 0.3, 1.0), ((0.99, 0.01), -2.0, 1.3), ((0.02, 0.6), 3.0, 0.7)]):
        r, x = _rest_and_moved(seed, 12, c0, th, sc)
        start = np.array([0, 12], np.int32)
        ta = np.array([sc * sc * G.shoelace(r)])
        y = O.shape_project(x.reshape(-1), start, r.reshape(-1), ta)
        assert np.abs((y - x.reshape(-1) + 0.5) % 1.0 - 0.5).max() < 1e-15
        d = (y.reshape(-1, 2) - y[:2] + 0.5) % 1.0 - 0.5
        np.testing.assert_allclose(G.shoelace(d), ta[0], rtol=1e-14)
        assert np.abs(O.shape_project(y, start, r.reshape(-1), ta) - y).max() <= 2.3e-16   # 2 ulp at 0.5


@pytest.mark.parametrize("seed", range(4))
def test_shape_projection_is_the_least_squares_rigid_motion(seed):
    """Noisy vertices: the result is the rest shape (unscaled) at the vertex mean of the
    input and at the rotation angle that minimizes sum |q_i - R(theta) r_i|^2, found here
    by a brute-force scan plus golden-section refinement of theta (independent of the
    closed form used by the oracle)."""
    r, x = _rest_and_moved(10 + seed, 16, (0.3, 0.7), 1.0 + seed, 1.0, noise=3e-3)
    start = np.array([0, 16], np.int32)
    y = O.shape_project(x.reshape(-1), start, r.reshape(-1)).reshape(-1, 2)
    q = (x - x[0] + 0.5) % 1.0 - 0.5
    q = q - q.mean(axis=0)

    def cost(t):
        R = np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
        return np.sum((q - r @ R.T) ** 2)
    ts = np.linspace(-np.pi, np.pi, 20001)
    t0 = ts[np.argmin([cost(t) for t in ts])]
    lo, hi = t0 - 1e-3, t0 + 1e-3
    g = (np.sqrt(5) - 1) / 2
    for _ in range(80):
        a, b = hi - g * (hi - lo), lo + g * (hi - lo)
        if cost(a) < cost(b):
            hi = b
        else:
            lo = a
    tbest = 0.5 * (lo + hi)
    yd = (y - x[0] + 0.5) % 1.0 - 0.5
    np.testing.assert_allclose(yd.mean(axis=0), ((x - x[0] + 0.5) % 1.0 - 0.5).mean(axis=0), atol=1e-15)
    yo = yd - yd.mean(axis=0)
    tgot = np.arctan2(np.sum(r[:, 0] * yo[:, 1] - r[:, 1] * yo[:, 0]), np.sum(r[:, 0] * yo[:, 0] + r[:, 1] * yo[:, 1]))
    assert abs((tgot - tbest + np.pi) % (2 * np.pi) - np.pi) < 1e-8
    np.testing.assert_allclose(np.linalg.norm(yo, axis=1), np.linalg.norm(r, axis=1), rtol=1e-13)


def test_shape_projection_rejects_polygons_beyond_64():
    p = make_packing(4, 65, 0.5, seed=0)
    with pytest.raises(ValueError):
        O.shape_project(p["positions"], p["startIndices"], np.zeros(len(p["positions"])))
