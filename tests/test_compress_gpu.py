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
