"""The oracle's compliant XPBD (orc_xpbd_lanes) and FIRE with XPBD springs, pinned on the CPU.

Reference: xpbdProjectCUDA src/model.cu:209-278 and kernels.cuh:532-623 (hard edge and area
projections); compliance as in Macklin, Mueller & Chentanez, "XPBD", MIG 2016:
dlambda = -(C + at lambda) / (sum |grad C|^2 + at), at = 1/(k dt^2).

Pins (each computed here from the geometry, independently of the oracle's code):
* one edge of an axis-aligned rectangle: its length after one iteration is
  (at l + 2 L)/(2 + at) -- the single-constraint step with w1 = w2 = 1;
* the perimeter of a regular polygon: one iteration scales it about its centre by
  s = (R at + 2 sin(pi/n) P0)/(R at + 2 sin(pi/n) P1) and a second changes nothing;
* the area step on a noisy polygon: x + grad A dlambda with the numpy shoelace gradient;
* area and perimeter together: one step of the 2x2 system (J J^T + diag(at)) dl = -r;
* k = INFINITY is the reference's hard projection (orc_xpbd, its kernels restated in their
  own order) for even and odd n; k -> 0 leaves the positions where they are;
* the perimeter spring of the force model is minus the gradient of its energy;
* FIRE with XPBD springs ends at a minimum of the total energy.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _geom as G

INF = float("inf")


def _grad_area(P):
    """d(shoelace area)/d(vertex), (n, 2)"""
    nxt, prv = np.roll(P, -1, axis=0), np.roll(P, 1, axis=0)
    return 0.5 * np.stack([nxt[:, 1] - prv[:, 1], prv[:, 0] - nxt[:, 0]], axis=1)


def _perimeter(P):
    d = np.roll(P, -1, axis=0) - P
    return float(np.hypot(d[:, 0], d[:, 1]).sum())


def _grad_perimeter(P):
    d = np.roll(P, -1, axis=0) - P
    t = d / np.hypot(d[:, 0], d[:, 1])[:, None]
    return np.roll(t, 1, axis=0) - t


def _noisy(n, R=0.05, seed=0, c=(0.5, 0.5)):
    rng = np.random.default_rng(seed)
    P = G.regular(n, R, c, rng.uniform(0, 6))
    return c + (P - c) * (1 + rng.uniform(-0.05, 0.05, size=(n, 1)))


@pytest.mark.parametrize("k,dt", [(1e3, 0.05), (INF, 0.05), (10.0, 0.3)])
def test_edge_step_of_a_rectangle(k, dt):
    """Axis-aligned rectangle, edges only: the even edges (horizontal) move their ends
    horizontally and the odd edges (vertical) vertically, so after one iteration every edge
    has taken exactly one step dlambda = -(l - L)/(2 + at) from lambda = 0 and its length is
    l + 2 dlambda = (at l + 2 L)/(2 + at)."""
    a, b = 0.1, 0.06
    P = np.array([[0.4, 0.45], [0.4 + a, 0.45], [0.4 + a, 0.45 + b], [0.4, 0.45 + b]])
    start, pos = G.to_model([P])
    L = np.array([0.08, 0.07, 0.08, 0.07])
    at = 1.0 / (k * dt * dt)
    out, it = O.xpbd_lanes(pos, start, k, 0.0, 0.0, dt, 1, targetEdgeV=L)
    Q = out.reshape(-1, 2)
    sides = np.hypot(*(np.roll(Q, -1, axis=0) - Q).T)
    expect = (at * np.array([a, b, a, b]) + 2 * L) / (2 + at)
    np.testing.assert_allclose(sides, expect, rtol=1e-14, atol=0)
    # still a rectangle with the same centre
    np.testing.assert_allclose(Q.mean(axis=0), P.mean(axis=0), rtol=0, atol=1e-15)
    assert abs(Q[0, 1] - Q[1, 1]) < 1e-15 and abs(Q[0, 0] - Q[3, 0]) < 1e-15


@pytest.mark.parametrize("n", [6, 9, 16])
@pytest.mark.parametrize("k", [1e3, INF])
def test_perimeter_of_a_regular_polygon(n, k):
    R, dt = 0.05, 0.1
    P = G.regular(n, R, (0.5, 0.5), 0.4)
    start, pos = G.to_model([P])
    P1 = 2 * n * R * np.sin(np.pi / n)
    P0 = 1.03 * P1
    L = np.full(n, P0 / n)
    at = 1.0 / (k * dt * dt)
    sn = 2 * np.sin(np.pi / n)
    s = (R * at + sn * P0) / (R * at + sn * P1)
    one, _ = O.xpbd_lanes(pos, start, 0.0, 0.0, k, dt, 1, targetEdgeV=L)
    np.testing.assert_allclose(one.reshape(-1, 2), 0.5 + s * (P - 0.5), rtol=0, atol=2e-16)
    two, _ = O.xpbd_lanes(pos, start, 0.0, 0.0, k, dt, 2, targetEdgeV=L)
    np.testing.assert_allclose(two, one, rtol=0, atol=2e-16)
    if k == INF:
        assert abs(_perimeter(one.reshape(-1, 2)) - P0) < 1e-15


@pytest.mark.parametrize("k", [1e3, 5.0, INF])
def test_area_step_on_a_noisy_polygon(k):
    dt = 0.2
    P = _noisy(11, seed=2)
    start, pos = G.to_model([P])
    A1 = G.shoelace(P)
    A0 = 0.97 * A1
    at = 1.0 / (k * dt * dt)
    g = _grad_area(P)
    dl = -(A1 - A0) / (np.sum(g * g) + at)
    out, _ = O.xpbd_lanes(pos, start, 0.0, k, 0.0, dt, 1, targetAreas=[A0])
    np.testing.assert_allclose(out.reshape(-1, 2), P + g * dl, rtol=0, atol=1e-16)


@pytest.mark.parametrize("k", [1e3, INF])
def test_area_and_perimeter_block_step(k):
    """Both polygon-wide constraints on: one step solves (J J^T + diag(at)) dl = -r with
    J = [grad A; grad P] at the current positions, and moves x by J^T dl."""
    dt = 0.1
    P = _noisy(12, seed=5)
    start, pos = G.to_model([P])
    # a longer perimeter around a smaller area (a rounder polygon may not exist)
    L = np.hypot(*(np.roll(P, -1, axis=0) - P).T) * 1.01
    A0 = G.shoelace(P) * 0.99
    at = 1.0 / (k * dt * dt)
    gA, gP = _grad_area(P).reshape(-1), _grad_perimeter(P).reshape(-1)
    J = np.stack([gA, gP])
    r = np.array([G.shoelace(P) - A0, _perimeter(P) - L.sum()])
    dl = np.linalg.solve(J @ J.T + at * np.eye(2), -r)
    out, _ = O.xpbd_lanes(pos, start, 0.0, k, k, dt, 1, targetEdgeV=L, targetAreas=[A0])
    step = J.T @ dl
    # the two gradients are nearly parallel, so rounding is amplified by the 2x2 solve
    np.testing.assert_allclose(out - pos, step, rtol=0, atol=1e-9 * np.abs(step).max())
    if k == INF:   # converged: both constraints hold
        conv, it = O.xpbd_lanes(pos, start, 0.0, k, k, dt, 50, tol=1e-15, targetEdgeV=L, targetAreas=[A0])
        Q = conv.reshape(-1, 2)
        assert it < 50
        assert abs(G.shoelace(Q) - A0) < 1e-15 and abs(_perimeter(Q) - L.sum()) < 1e-15


@pytest.mark.parametrize("n", [8, 9, 12, 13])
def test_hard_limit_is_the_reference_projection(n):
    """k = INFINITY, edges + area, per-polygon edge target: the reference's kernels
    (restated in their own order by orc_xpbd) to rounding, for even and odd n, and the same
    iteration count under a tolerance."""
    polys = [_noisy(n, R=0.04, seed=s, c=(0.25 + 0.5 * (s % 2), 0.25 + 0.5 * (s // 2))) for s in range(4)]
    start, pos = G.to_model(polys)
    Lp = np.array([np.hypot(*(np.roll(Q, -1, axis=0) - Q).T).mean() for Q in polys])
    A0 = np.array([G.shoelace(Q) for Q in polys]) * 0.98
    ref, _ = O.xpbd(pos, start, Lp, A0, 6, 0.0)
    lan, it = O.xpbd_lanes(pos, start, INF, INF, 0.0, 1.0, 6, targetEdgeP=Lp, targetAreas=A0)
    assert it == 6
    np.testing.assert_allclose(lan, ref, rtol=0, atol=1e-14)
    ref2, it_r = O.xpbd(pos, start, Lp, A0, 500, 1e-12)
    lan2, it_l = O.xpbd_lanes(pos, start, INF, INF, 0.0, 1.0, 500, 1e-12, targetEdgeP=Lp, targetAreas=A0)
    assert it_l == it_r < 500
    # the reference keeps iterating converged polygons until all have converged; here each
    # stops on its own, so the two differ by what the extra iterations move (below ~tol)
    np.testing.assert_allclose(lan2, ref2, rtol=0, atol=1e-11)


def test_soft_limit_leaves_positions():
    P = _noisy(10, seed=7)
    start, pos = G.to_model([P])
    L = np.hypot(*(np.roll(P, -1, axis=0) - P).T) * 0.9
    A0 = G.shoelace(P) * 0.9
    soft, _ = O.xpbd_lanes(pos, start, 1e-12, 1e-12, 1e-12, 0.1, 5, targetEdgeV=L, targetAreas=[A0])
    np.testing.assert_allclose(soft, pos, rtol=0, atol=1e-15)
    off, it = O.xpbd_lanes(pos, start, 0.0, 0.0, 0.0, 0.1, 5, tol=1e-3, targetEdgeV=L, targetAreas=[A0])
    assert np.array_equal(off, pos) and it == 0


def test_perimeter_spring_is_minus_gradient():
    """softBody with the perimeter spring: E = sum k (l - l0)^2/2 + c (A - A0)^2/2
    + kP (P - P0)^2/2, F = -grad E by central differences (P0 = sum of the edge targets)."""
    rng = np.random.default_rng(3)
    polys = [_noisy(n, R=0.05, seed=s, c=(0.25 + 0.45 * s, 0.5)) for s, n in enumerate((7, 10))]
    start, pos = G.to_model(polys)
    V = int(start[-1])
    l0 = rng.uniform(0.025, 0.035, V)
    A0 = np.array([G.shoelace(Q) for Q in polys]) * 1.05
    k, c, kP = 2.0, 5.0, 7.0

    def energy_np(x):
        E = 0.0
        for s in range(2):
            Q = x[2 * start[s]:2 * start[s + 1]].reshape(-1, 2)
            d = np.hypot(*(np.roll(Q, -1, axis=0) - Q).T)
            E += np.sum(k * (d - l0[start[s]:start[s + 1]]) ** 2 / 2)
            E += c * (G.shoelace(Q) - A0[s]) ** 2 / 2
            E += kP * (d.sum() - l0[start[s]:start[s + 1]].sum()) ** 2 / 2
        return E

    geom = O.geometry(pos, start)
    E, F, fe = O.force_energy("softBody", pos, start, None, geom, l0, A0, k, c, perimStiffness=kP,
                              want_fext=True)
    assert abs(E - energy_np(pos)) <= 1e-12 * E
    assert np.all(fe == 0.0)
    h = 1e-7
    for q in range(0, 2 * V, 3):
        xp, xm = pos.copy(), pos.copy()
        xp[q] += h; xm[q] -= h
        fd = -(energy_np(xp) - energy_np(xm)) / (2 * h)
        assert abs(F[q] - fd) <= 1e-6 * np.abs(F).max(), (q, F[q], fd)


def test_fire_with_xpbd_springs_ends_at_a_minimum():
    """Two overlapping polygons, hybrid model (squared pair area + area and perimeter
    springs k = 1e3, integrated by XPBD): FIRE reaches max|F| <= 1e-10 and the total energy
    does not go down along random directions there."""
    A = _noisy(10, R=0.06, seed=11, c=(0.45, 0.5))
    B = _noisy(12, R=0.05, seed=12, c=(0.55, 0.52))
    start, pos = G.to_model([A, B])
    V = int(start[-1])
    tev = np.concatenate([np.hypot(*(np.roll(Q, -1, axis=0) - Q).T) for Q in (A, B)])
    ta = np.array([G.shoelace(A), G.shoelace(B)])
    sysO = O.System(start, 2, 0x400 | O.HYBRID, tev, ta, stiffness=0.0, compressibility=1e3,
                    solver="lanes", perimStiffness=1e3)
    x, f, st, steps = sysO.fire_minimize(pos, fTol=1e-10, dt=1e-3, maxSteps=20000, dtMax=100.0,
                                         xpbdIter=20, xpbdTol=1e-15)
    assert np.abs(f).max() <= 1e-10 and steps < 20000
    E, F, fe, nI = sysO.force_eval_ext(x)
    E0 = sysO.force_eval(pos)[0]
    assert np.array_equal(F, f) and 0 <= E < 1e-5 * E0
    rng = np.random.default_rng(0)
    for _ in range(5):
        d = rng.normal(size=x.shape)
        d *= 1e-6 / np.linalg.norm(d)
        assert sysO.force_eval(x + d)[0] >= E - 1e-22 and sysO.force_eval(x - d)[0] >= E - 1e-22
