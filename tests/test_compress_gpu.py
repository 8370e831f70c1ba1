"""Compression protocol pieces on the GPU: resetAreas (pcp_reset_areas, kernels.cuh:1440-1455)
pinned by what it must achieve -- every polygon's area equals its new target, its centroid
stays put and every edge is scaled by the same factor sqrt(target / area) -- and a short
packing-fraction ramp whose every step ends minimized (max|F| <= 1e-10)."""
import numpy as np
import pytest

from pycudapolygon_amd import capi, compress
from pycudapolygon_amd.generate import make_packing, make_batch, box_size_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


def unwrapped(x, start):
    """vertices of each polygon as minimum-image offsets from its first vertex (host)"""
    v = x.reshape(-1, 2)
    out = []
    for a, b in zip(start[:-1], start[1:]):
        d = v[a:b] - v[a]
        d -= np.round(d)
        out.append(d)
    return out


def shoelace(d):
    x, y = d[:, 0], d[:, 1]
    return 0.5 * float((x * np.roll(y, -1) - np.roll(x, -1) * y).sum())


def test_reset_areas_reaches_targets_about_fixed_centroids():
    pk = make_batch(3, 51, N_range=(40, 120), n_range=(8, 24), min_center_frac=0.5)
    s = capi.System(pk, [box_size_for(p["diameter"]) for p in pk], model=capi.NORMAL | 0x200)
    start = s.polyStart
    s.geometry(constraints=False)
    x0 = s.get_positions()
    com0 = np.stack([s.comX.get(), s.comY.get()], 1)
    rng = np.random.default_rng(0)
    r = rng.uniform(0.9, 1.1, s.N)
    ta = s.ta.get() * r
    s.ta.upload(ta)
    a0 = np.array([shoelace(d) for d in unwrapped(x0, start)])
    s.ctx.call("pcp_reset_areas", s.pos.ptr, s.areas.ptr, s.ta.ptr, s.comX.ptr, s.comY.ptr)
    s.sync()
    x1 = s.get_positions()
    assert ((x1 >= 0) & (x1 < 1)).all()
    u0, u1 = unwrapped(x0, start), unwrapped(x1, start)
    a1 = np.array([shoelace(d) for d in u1])
    np.testing.assert_allclose(a1, ta, rtol=1e-12, atol=0)
    for k, (d0, d1) in enumerate(zip(u0, u1)):
        e0 = np.linalg.norm(np.roll(d0, -1, 0) - d0, axis=1)
        e1 = np.linalg.norm(np.roll(d1, -1, 0) - d1, axis=1)
        np.testing.assert_allclose(e1, e0 * np.sqrt(ta[k] / a0[k]), rtol=1e-12, atol=0)
    s.geometry(constraints=False)
    com1 = np.stack([s.comX.get(), s.comY.get()], 1)
    dc = com1 - com0
    dc -= np.round(dc)
    assert np.abs(dc).max() <= 1e-14


def test_scale_targets_equals_host_scaling():
    """pcp_scale_targets multiplies each packing's target areas and edge targets by its two
    factors on the device: bit-identical to the host's numpy products (setPhi's scaling,
    pyCudaPolygon.py:223-234), uneven packings, a NULL array left untouched."""
    pk = make_batch(4, 52, N_range=(30, 90), n_range=(8, 24), min_center_frac=0.5)
    s = capi.System(pk, [box_size_for(p["diameter"]) for p in pk], model=capi.NORMAL | 0x200)
    ta0, tev0 = s.ta.get(), s.tev.get()
    phi = compress.packing_fractions(s)
    r = np.ascontiguousarray((phi + np.array([0.01, 0.002, -0.003, 0.0])) / phi)
    q = np.sqrt(r)
    s.ctx.call("pcp_scale_targets", capi._ptr(r), capi._ptr(q), s.ta.ptr, s.tev.ptr)
    s.sync()
    ta_h, tev_h = ta0.copy(), tev0.copy()
    for p in range(s.P):
        a, b = int(s.packPoly[p]), int(s.packPoly[p + 1])
        ta_h[a:b] *= r[p]
        va, vb = s.vertex_range(p)
        tev_h[va:vb] *= np.sqrt(r[p])
    assert np.array_equal(s.ta.get(), ta_h) and np.array_equal(s.tev.get(), tev_h)
    # edge targets only: the areas stay as they are
    s.ctx.call("pcp_scale_targets", None, capi._ptr(q), None, s.tev.ptr)
    s.sync()
    assert np.array_equal(s.ta.get(), ta_h)
    for p in range(s.P):
        va, vb = s.vertex_range(p)
        tev_h[va:vb] *= np.sqrt(r[p])
    assert np.array_equal(s.tev.get(), tev_h)


def test_short_compression_ramp_minimizes_every_step():
    pk = [make_packing(128, 16, 0.80, seed=60 + k, min_center_frac=0.5) for k in range(2)]
    steps, s = compress.ramp(pk, 0.804, 1e-3, dict(fTol=1e-10, dtMax=100.0), 4000,
                             log=lambda m: print(m, flush=True))
    assert len(steps) == 5
    for rec in steps:
        assert all(rec["converged"]) and max(rec["max_force"]) <= 1e-10
    np.testing.assert_allclose(compress.packing_fractions(s), 0.804, rtol=1e-12)
    s.geometry(constraints=False)
    np.testing.assert_allclose(s.areas.get(), s.ta.get(), rtol=1e-9)   # SHAKE keeps the areas


def test_shape_projection_bit_identical_to_oracle():
    """pcp_shape_project (2x2 polar decomposition per polygon) equals the oracle's shape
    matching bit for bit: 16-gons (16-lane segments) and uneven packings (n 8..32), with
    and without scaling to target areas, on positions moved away from the rest shapes."""
    from oracle import oracle as O
    rng = np.random.default_rng(7)
    cases = [[make_packing(64, 16, 0.8, seed=3)], make_batch(4, 61, N_range=(40, 200), n_range=(8, 32))]
    for packs in cases:
        ref = compress.rest_shapes(packs)
        s = capi.System(packs, [box_size_for(p["diameter"]) for p in packs])
        x = (np.concatenate([p["positions"] for p in packs]) + rng.normal(scale=1e-3, size=len(ref))) % 1.0
        ta = np.concatenate([p["targetAreas"] for p in packs]) * 1.01
        dref, dta = capi.DeviceArray.from_numpy(ref), capi.DeviceArray.from_numpy(ta)
        for targets in (None, dta):
            print(f"[shape] packings={len(packs)} scaled={targets is not None}", flush=True)
            s.pos.upload(x)
            s.ctx.call("pcp_shape_project", s.pos.ptr, dref.ptr, targets.ptr if targets is not None else None)
            s.sync()
            y = s.pos.get()
            for k, p in enumerate(packs):
                a, b = s.vertex_range(k)
                pa, pb = int(s.packPoly[k]), int(s.packPoly[k + 1])
                yo = O.shape_project(x[2 * a:2 * b], p["startIndices"], ref[2 * a:2 * b],
                                     ta[pa:pb] if targets is not None else None)
                assert np.array_equal(y[2 * a:2 * b], yo)


def test_compression_ramp_with_shape_matching():
    """Ramp with shape matching after every step: right after a step every polygon is its
    initial shape (edge lengths scaled by sqrt(phi'/phi0), area = target), and every step
    ends minimized."""
    pk = [make_packing(256, 12, 0.80, seed=71 + k, min_center_frac=0.5) for k in range(2)]
    ref = compress.rest_shapes(pk)
    s = capi.System(pk, [box_size_for(p["diameter"] * 1.05) for p in pk], model=capi.NORMAL | 0x200)
    dref = capi.DeviceArray.from_numpy(ref)
    phi0 = compress.packing_fractions(s)
    compress.set_phi(s, phi0 + 0.01, dref)
    x = s.pos.get()
    e0 = np.concatenate([np.linalg.norm(np.roll(d, -1, axis=0) - d, axis=1)
                         for d in unwrapped(np.concatenate([p["positions"] for p in pk]), s.polyStart)])
    e1 = np.concatenate([np.linalg.norm(np.roll(d, -1, axis=0) - d, axis=1) for d in unwrapped(x, s.polyStart)])
    f = np.repeat(np.sqrt((phi0 + 0.01) / phi0), [int(s.polyStart[s.packPoly[k + 1]] - s.polyStart[s.packPoly[k]])
                                                   for k in range(s.P)])
    np.testing.assert_allclose(e1, e0 * f, rtol=1e-12)
    areas = np.array([shoelace(d) for d in unwrapped(x, s.polyStart)])
    np.testing.assert_allclose(areas, s.ta.get(), rtol=1e-12)
    steps, s2 = compress.ramp(pk, 0.82, 1e-2, dict(fTol=1e-10, dtMax=100.0), 5000, dt0=100.0, shape_match=True)
    assert len(steps) == 3
    for r in steps:
        assert all(r["converged"]) and max(r["max_force"]) <= 1e-10


def test_shape_projection_rejects_polygons_beyond_64():
    """pcp_shape_project takes polygons of at most 64 vertices and says so (PCP_ERR_ARG)."""
    p = make_packing(4, 65, 0.5, seed=0)
    s = capi.System([p], [2])
    ref = capi.DeviceArray.from_numpy(compress.rest_shapes([p]))
    with pytest.raises(capi.PcpError, match="64"):
        s.ctx.call("pcp_shape_project", s.pos.ptr, ref.ptr, s.ta.ptr)


def test_c5_step_learned_warm_start_shake_and_shape_projection():
    """The compression configuration at reduced size (2 replicas of N=256 x 24, the committed
    compression network): minimized at phi0, then one phi step of all three pieces -- every
    polygon put to its rest shape at the new target area (2x2 polar decomposition), the learned
    warm start (with its SHAKE), FIRE with SHAKE to max|F| <= 1e-10.  From the warm-started
    positions, FIRE reaches the oracle's FIRE (lane-order solves) in the same steps, at the same
    energy and positions, replica by replica; a two-step ramp of the same pieces converges."""
    from oracle import oracle as O
    from pycudapolygon_amd import warmstart as W
    pk = [make_packing(256, 24, 0.80, seed=91 + k, min_center_frac=0.5) for k in range(2)]
    B = [box_size_for(p["diameter"] * np.sqrt(0.81 / 0.80)) for p in pk]
    s = capi.System(pk, B, model=compress.SQUARED)
    prm = s.fire_params(fTol=1e-10, dtMax=W.DT_MAX)
    s.fire_init(prm, 1e-3)
    assert all(q.converged for q in s.fire_run(prm, 5000, 8, True))
    ref = capi.DeviceArray.from_numpy(compress.rest_shapes(pk))
    phi0 = compress.packing_fractions(s)
    compress.set_phi(s, phi0 + 0.005, ref)
    xs = s.pos.get().copy()
    areas = np.array([shoelace(d) for d in unwrapped(xs, s.polyStart)])
    np.testing.assert_allclose(areas, s.ta.get(), rtol=1e-12)
    ell = np.array([float(np.mean(p["targetEdgeV"])) for p in pk]) * np.sqrt((phi0 + 0.005) / phi0)
    s.force_eval()
    W.WarmStart(W.load_weights(W.COMPRESS_WEIGHTS)).apply(s, ell)
    x0 = s.pos.get().copy()
    assert not np.array_equal(x0, xs)   # the network moved the vertices
    s.fire_init(prm, 1e-3)
    st = s.fire_run(prm, 5000, 8, True)
    xg = s.pos.get()
    ta, tev = s.ta.get(), s.tev.get()
    for k, p in enumerate(pk):
        a, b = s.vertex_range(k)
        pa, pb = int(s.packPoly[k]), int(s.packPoly[k + 1])
        print(f"[c5] replica {k}: GPU {st[k].steps} steps E {st[k].energy!r}", flush=True)
        sysO = O.System(p["startIndices"], B[k], O.NORMAL | 0x200, tev[a:b], ta[pa:pb],
                        solver="lanes")
        xo, fo, sto, steps = sysO.fire_minimize(x0[2 * a:2 * b], fTol=1e-10, dt=1e-3, maxSteps=5000,
                                                dtMax=W.DT_MAX)
        assert st[k].converged and st[k].maxForce <= 1e-10 and np.abs(fo).max() <= 1e-10
        assert st[k].steps == steps and st[k].energy == sto.energy
        assert np.array_equal(xg[2 * a:2 * b], xo)
    s.close()
    steps, s2 = compress.ramp(pk, 0.81, 5e-3, dict(fTol=1e-10, dtMax=W.DT_MAX), 5000,
                              warm=W.WarmStart(W.load_weights(W.COMPRESS_WEIGHTS)), shape_match=True)
    assert len(steps) == 3
    for r in steps:
        assert all(r["converged"]) and max(r["max_force"]) <= 1e-10
    s2.close()
