"""Known-answer tests of the oracle's FIRE state machine and of where its SHAKE lands.

Reference: Model::minimizeFIREStep src/model.cpp:405-486 (rollback tolerance :453,
SHAKE retry/rollback :418-438, dt/alpha/nPos update :471-485) and Bitzek et al.,
PRL 97, 170201 (2006).  Each test drives orc_fire_step on a small system whose
forces come from closed forms written here (edge springs: E = sum k (l - l0)^2 / 2),
and checks the step against values fixed by the reference's text: the
velocity-Verlet recursion, the mixing v <- (1-a) v + a |v| F/|F|, the dt / alpha
sequence, the two sides of the rollback threshold max(1e-10 |E|, 1e-14), the P <= 0
branch, and the 10x SHAKE retry.  A mistake such as a wrong rollback tolerance,
alpha decayed without dt growth, a dropped retry or a half-kick with the wrong force
fails one of them.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _geom as G

SOFT = O.SOFT_BODY


def _square(side, center=(0.5, 0.5)):
    c = np.asarray(center)
    h = side / 2.0
    return np.array([c + [-h, -h], c + [h, -h], c + [h, h], c + [-h, h]])


def _spring_system(poly, l0, k=1.0):
    start, pos = G.to_model([poly])
    V = int(start[-1])
    sysO = O.System(start, 2, SOFT, np.full(V, l0), np.zeros(1), stiffness=k, compressibility=0.0)
    return sysO, start, pos


def _spring_force(pos, l0, k):
    """-grad of sum_e k (|p_{e+1} - p_e| - l0)^2 / 2 (no wrap needed: polygons stay inside)."""
    P = pos.reshape(-1, 2)
    n = len(P)
    F = np.zeros_like(P)
    for e in range(n):
        d = P[(e + 1) % n] - P[e]
        l = np.hypot(d[0], d[1])
        g = k * (l - l0) * d / l   # = -dE/dp_e ; +dE/dp_{e+1}
        F[e] += g
        F[(e + 1) % n] -= g
    return F.reshape(-1)


def _spring_energy(pos, l0, k):
    P = pos.reshape(-1, 2)
    d = np.roll(P, -1, axis=0) - P
    return float(np.sum(k * (np.hypot(d[:, 0], d[:, 1]) - l0) ** 2 / 2))


def test_closed_form_springs_agree_with_the_oracle():
    sysO, start, pos = _spring_system(_square(0.1), 0.08, k=1.0)
    E, F, _ = sysO.force_eval(pos)
    assert abs(E - _spring_energy(pos, 0.08, 1.0)) <= 1e-13 * E
    np.testing.assert_allclose(F, _spring_force(pos, 0.08, 1.0), rtol=0, atol=1e-15)


def test_verlet_recursion_with_alpha_zero():
    """alphaStart = 0 and a contracting stretched square (E falls, P > 0 at every step):
    x_{n+1} = x_n + dt_n (v_n + dt_n/2 F_n), v_{n+1} = v_n + dt_n/2 (F_n + F_{n+1}),
    dt_{n+1} = min(fInc dt_n, dtMax) (nMin = 1), alpha stays 0."""
    k, l0 = 1.0, 0.08
    sysO, start, pos = _spring_system(_square(0.1), l0, k)
    x = pos.copy()
    st, vel, f = sysO.fire_init(x, 0.05, 0.0)
    xr, vr, dt, fInc = pos.copy(), np.zeros_like(pos), 0.05, 1.1
    Er = _spring_energy(xr, l0, k)
    for step in range(8):
        sysO.fire_step(x, vel, f, st, dtMax=1.0, alphaStart=0.0, fAlpha=0.99, fInc=fInc, fDec=0.5, nMin=1,
                       shakeIter=0)
        Fr = _spring_force(xr, l0, k)
        vh = vr + 0.5 * dt * Fr
        xr = xr + dt * vh
        Fn = _spring_force(xr, l0, k)
        vr = vh + 0.5 * dt * Fn
        En = _spring_energy(xr, l0, k)
        assert En < Er and np.dot(Fn, vr) > 0   # the regime the test is built for
        Er = En
        dt = min(dt * fInc, 1.0)
        assert st.rolledBack == 0 and st.nPos == step + 1
        assert st.dt == dt and st.alpha == 0.0
        np.testing.assert_allclose(x, xr, rtol=0, atol=1e-14)
        np.testing.assert_allclose(vel, vr, rtol=0, atol=1e-13)
        np.testing.assert_allclose(f, Fn, rtol=0, atol=1e-15)


def test_mixing_and_dt_alpha_sequence():
    """nMin = 3: dt and alpha stay for the first two positive steps; then dt grows by fInc
    up to dtMax and alpha decays by fAlpha only in the steps where dt grew
    (model.cpp:473-477); the velocity is mixed as v <- (1-a) v + (a |v| / |F|) F."""
    k, l0 = 1.0, 0.08
    sysO, start, pos = _spring_system(_square(0.1), l0, k)
    x = pos.copy()
    dt0, a0, fInc, fAlpha = 0.05, 0.1, 1.1, 0.99
    dtMax = (dt0 * fInc) * fInc
    st, vel, f = sysO.fire_init(x, dt0, a0)
    xr, vr = pos.copy(), np.zeros_like(pos)
    expect_dt = [dt0, dt0, dt0 * fInc, dtMax, dtMax, dtMax]
    expect_a = [a0, a0, a0 * fAlpha, a0 * fAlpha * fAlpha, a0 * fAlpha * fAlpha, a0 * fAlpha * fAlpha]
    dt, a = dt0, a0
    for step in range(6):
        sysO.fire_step(x, vel, f, st, dtMax=dtMax, alphaStart=a0, fAlpha=fAlpha, fInc=fInc, fDec=0.5, nMin=3,
                       shakeIter=0)
        vh = vr + 0.5 * dt * _spring_force(xr, l0, k)
        xr = xr + dt * vh
        Fn = _spring_force(xr, l0, k)
        vr = vh + 0.5 * dt * Fn
        assert np.dot(Fn, vr) > 0
        vr = (1 - a) * vr + (a * np.linalg.norm(vr) / np.linalg.norm(Fn)) * Fn
        assert st.rolledBack == 0 and st.nPos == step + 1
        assert st.dt == expect_dt[step] and st.alpha == expect_a[step], (step, st.dt, st.alpha)
        dt, a = st.dt, st.alpha
        np.testing.assert_allclose(x, xr, rtol=0, atol=1e-14)
        np.testing.assert_allclose(vel, vr, rtol=0, atol=1e-13)


def test_power_not_positive_zeroes_velocity_and_resets():
    """An accepted step that overshoots the minimum (E falls, P = F.v < 0) zeroes v,
    multiplies dt by fDec and resets alpha and nPos (model.cpp:479-484)."""
    k, l0 = 1.0, 0.08
    sysO, start, pos = _spring_system(_square(0.1), l0, k)
    x = pos.copy()
    dt = 0.1
    st, vel, f = sysO.fire_init(x, dt, 0.07)
    st.nPos = 4
    target = _square(0.07).reshape(-1)
    vh = (target - pos) / dt
    vel[:] = vh - 0.5 * dt * f
    sysO.fire_step(x, vel, f, st, dtMax=1.0, alphaStart=0.1, fAlpha=0.99, fInc=1.1, fDec=0.5, nMin=2, shakeIter=0)
    assert st.rolledBack == 0
    assert st.energy < _spring_energy(pos, l0, k)
    np.testing.assert_allclose(x, target, rtol=0, atol=1e-15)
    assert np.all(vel == 0.0)
    assert st.dt == dt * 0.5 and st.alpha == 0.1 and st.nPos == 0


def _rise_step(sysO, pos, f, dt, d, s):
    """positions after a step whose half-kicked velocity is s * d (the oracle's arithmetic)"""
    vel = s * d - 0.5 * dt * f
    vh = vel + (0.5 * dt) * f
    p = pos + dt * vh
    return vel, p - np.floor(p)


@pytest.mark.parametrize("stretch,branch", [(0.02, "relative"), (1e-6, "absolute")])
def test_rollback_threshold_both_sides(stretch, branch):
    """The step is rolled back iff E_post > E_pre + max(1e-10 |E_pre|, 1e-14)
    (model.cpp:453): energy rises 0.1 % above / below the threshold, for a threshold set
    by the relative term (E ~ 1e-3) and by the 1e-14 floor (E ~ 1e-12)."""
    k = 1.0
    l0 = 0.1 - stretch
    sysO, start, pos = _spring_system(_square(0.1), l0, k)
    E0, f0, _ = sysO.force_eval(pos)
    thr = max(1e-10 * abs(E0), 1e-14)
    assert (thr > 1e-14) == (branch == "relative")
    dt = 1e-3
    d = np.zeros_like(pos)
    d[0:2] = [-1.0, -1.0]   # vertex 0 outwards: both of its edges stretch, E rises
    d /= np.linalg.norm(d)

    def rise(s):
        _, p = _rise_step(sysO, pos, f0, dt, d, s)
        return sysO.force_eval(p)[0] - E0

    for side in (+1, -1):
        goal = thr * (1 + side * 1e-3)
        lo, hi = 0.0, 1.0
        while rise(hi) < goal:
            hi *= 2
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if rise(mid) < goal:
                lo = mid
            else:
                hi = mid
        s = hi if side > 0 else lo
        r = rise(s)
        assert (r > thr * (1 + 5e-4)) if side > 0 else (thr * (1 - 5e-3) < r < thr * (1 - 5e-4))
        x = pos.copy()
        st, vel, f = sysO.fire_init(x, dt, 0.1)
        st.nPos = 7
        vel[:], xp = _rise_step(sysO, pos, f, dt, d, s)
        sysO.fire_step(x, vel, f, st, dtMax=1.0, alphaStart=0.1, fAlpha=0.99, fInc=1.1, fDec=0.5, nMin=5,
                       shakeIter=0)
        if side > 0:
            assert st.rolledBack == 1
            assert np.array_equal(x, pos) and np.all(vel == 0.0)
            assert st.energy == E0 and np.array_equal(f, f0)
            assert st.dt == dt * 0.5 and st.alpha == 0.1 and st.nPos == 0
        else:
            assert st.rolledBack == 0
            assert np.array_equal(x, xp)
            assert st.energy == E0 + r


def _ngon_system(n=12, R=0.05, seed=0, area_scale=1.0):
    rng = np.random.default_rng(seed)
    P = G.regular(n, R, (0.5, 0.5), 0.3)
    P = P + rng.uniform(-0.05, 0.05, size=(n, 1)) * (P - 0.5)
    start, pos = G.to_model([P])
    g = O.geometry(pos, start)
    sysO = O.System(start, 2, O.NORMAL, g["edgeLengths"], g["areas"] * area_scale, solver="dense")
    return sysO, start, pos, g


def test_shake_retry_at_ten_times_the_iterations():
    """shakeIter = 1: one Gauss-Newton iteration does not converge, so SHAKE runs again with
    10 iterations from where it stopped (model.cpp:420-429); it converges there, the step
    is kept and v is re-derived from the SHAKE displacement."""
    sysO, start, pos, g = _ngon_system()
    rng = np.random.default_rng(1)
    dt = 0.01
    x = pos.copy()
    st, vel, f = sysO.fire_init(x, dt, 0.1)
    vel[:] = rng.normal(scale=0.05, size=pos.shape)   # |dt v| ~ 5e-4 against edges ~ 0.026
    v0 = vel.copy()
    pre = pos + dt * v0           # f = 0 for a lone polygon
    pre = pre - np.floor(pre)
    sysO.fire_step(x, vel, f, st, dtMax=1.0, alphaStart=0.1, fAlpha=0.99, fInc=1.1, fDec=0.5, nMin=5, shakeIter=1)
    assert st.shakeFailed == 0 and 1 < st.lastShakeIters < 10
    once, it1 = O.shake(pre, start, g["edgeLengths"], g["areas"], 1)
    twice, it2 = O.shake(once, start, g["edgeLengths"], g["areas"], 10)
    assert it1 == 1 and it2 == st.lastShakeIters
    assert np.array_equal(x, twice)
    gg = O.geometry(x, start)
    assert np.abs(gg["edgeLengths"] - g["edgeLengths"]).max() < 1e-15
    assert abs(gg["areas"][0] - g["areas"][0]) < 1e-15
    wrap = lambda a: (a + 1.5) % 1.0 - 0.5
    # f = 0 here, so P = 0 and the velocity is zeroed at the end of the step; the
    # re-derived velocity is checked through the oracle's SHAKE displacement instead
    assert np.abs(wrap(x - pre)).max() > 0
    assert np.all(vel == 0.0) and st.dt == dt * 0.5


def test_shake_failure_after_the_retry_rolls_back():
    """A target area no polygon with these edges can have: SHAKE fails at shakeIter and
    at 10 shakeIter, so the step is undone: x restored, v = 0, dt *= fDec, alpha and nPos
    reset (model.cpp:430-435, 445-450), lastShakeIters = 10 shakeIter."""
    sysO, start, pos, g = _ngon_system(area_scale=1.5)
    dt = 0.01
    x = pos.copy()
    st, vel, f = sysO.fire_init(x, dt, 0.05)
    st.nPos = 3
    vel[:] = 0.01
    sysO.fire_step(x, vel, f, st, dtMax=1.0, alphaStart=0.1, fAlpha=0.99, fInc=1.1, fDec=0.5, nMin=5, shakeIter=2)
    assert st.shakeFailed == 1 and st.lastShakeIters == 20
    assert np.array_equal(x, pos) and np.all(vel == 0.0)
    assert st.dt == dt * 0.5 and st.alpha == 0.1 and st.nPos == 0


def _tangent_projector(pos, start):
    from tests.test_oracle import _edge_area_grads
    J = _edge_area_grads(pos, start, 0)          # columns: gradients of the n+1 constraints
    Q, _ = np.linalg.qr(J)
    return np.eye(J.shape[0]) - Q @ Q.T


@pytest.mark.parametrize("solver", ["dense", "lanes"])
def test_shake_lands_on_the_nearest_manifold_point(solver):
    """For a small displacement delta off the constraint manifold, SHAKE's result is
    x0 + P_T delta + O(|delta|^2), P_T the orthogonal projector onto the manifold's tangent
    space at x0: it lands on the nearest point to first order (corrections along the
    constraint gradients).  A rhombus scaled about its centre comes back to itself."""
    sysO, start, pos, g = _ngon_system(n=12, seed=3)
    rng = np.random.default_rng(4)
    delta = rng.normal(scale=1e-8, size=pos.shape)
    fn = O.shake if solver == "dense" else O.shake_lanes
    out, it = fn(pos + delta, start, g["edgeLengths"], g["areas"], 50)
    assert 0 < it < 50
    PT = _tangent_projector(pos, start)
    lin = PT @ delta
    err = np.linalg.norm((out - pos) - lin)
    assert err <= 2e-5 * np.linalg.norm(lin), err / np.linalg.norm(lin)   # O(|delta|) relative
    # the normal component is removed: the projected displacement is not the raw one
    assert np.linalg.norm(delta - lin) > 0.1 * np.linalg.norm(delta)

    # a rhombus (not cyclic, so the area row is independent of the edge rows) scaled about
    # its centre: the corrections keep its two mirror symmetries, and the one rhombus with
    # these mirrors, edge length and area is the original
    L, th = 0.06, np.pi / 3
    h1, h2 = L * np.cos(th / 2), L * np.sin(th / 2)
    rh = np.array([[0.5 + h1, 0.5], [0.5, 0.5 + h2], [0.5 - h1, 0.5], [0.5, 0.5 - h2]])
    st4, pos4 = G.to_model([rh])
    A = L * L * np.sin(th)
    scaled = (0.5 + 1.02 * (rh - 0.5)).reshape(-1)
    out4, it4 = fn(scaled, st4, np.full(4, L), np.array([A]), 50)
    assert 0 < it4 < 50
    np.testing.assert_allclose(out4, pos4, rtol=0, atol=2e-15)
