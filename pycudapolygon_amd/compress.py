"""Compression protocol: a packing-fraction ramp on several replicas in one context.

At each step phi -> phi + dphi, every packing's targets are scaled to the new packing
fraction (areas by phi'/phi, edge lengths by sqrt(phi'/phi), the reference's setPhi,
pyCudaPolygon.py:223-234), every polygon is rescaled about its centroid to its new target
area (resetAreas, kernels.cuh:1440-1455, through pcp_reset_areas) or, with shape matching,
replaced by its rest shape at the new target area in the least-squares rotation and
position (pcp_shape_project, the 2x2 polar decomposition per polygon), an optional learned warm
start is applied, and FP64 FIRE (rigid polygons: SHAKE onto edge lengths and area every
step, forces projected onto the constraint manifold) minimizes to max|F| <= fTol.  The box
stays the periodic unit square, as in the reference.  No counterpart of the loop itself in
the reference (its notebooks drive setPhi and minimizeFIRE by hand).
"""
import time

import numpy as np

from . import capi
from .generate import box_size_for

SQUARED = capi.NORMAL | 0x200


def packing_fractions(s):
    """sum of target areas of each packing of System s (the unit box has area 1)"""
    ta = s.ta.get()
    return np.array([ta[s.packPoly[p]:s.packPoly[p + 1]].sum() for p in range(s.P)])


def rest_shapes(packings):
    """Each polygon's vertices as offsets from their vertex mean (minimum image from the
    polygon's first vertex), concatenated over the packings: double[2V]."""
    out = []
    for p in packings:
        st, x = np.asarray(p["startIndices"]), np.asarray(p["positions"], np.float64).reshape(-1, 2)
        for a, b in zip(st[:-1], st[1:]):
            d = (x[a:b] - x[a] + 1.5) % 1.0 - 0.5
            out.append(d - d.mean(axis=0))
    return np.concatenate(out).reshape(-1)


def set_phi(s, phi, ref=None):
    """Scale every packing's targets to packing fraction phi[p] and rescale each polygon about
    its centroid to its new target area (setPhi + resetAreas), or, with ref (a device array
    of rest_shapes), put its rest shape at the new target area in place (pcp_shape_project)."""
    s.sync()
    phi = np.broadcast_to(np.asarray(phi, np.float64), (s.P,))
    r = np.ascontiguousarray(phi / packing_fractions(s))
    # areas times r and edge targets times sqrt(r), per packing, on the device
    s.ctx.call("pcp_scale_targets", capi._ptr(r), capi._ptr(np.sqrt(r)), s.ta.ptr, s.tev.ptr)
    if ref is not None:
        s.ctx.call("pcp_shape_project", s.pos.ptr, ref.ptr, s.ta.ptr)
    else:
        s.geometry(constraints=False)
        s.ctx.call("pcp_reset_areas", s.pos.ptr, s.areas.ptr, s.ta.ptr, s.comX.ptr, s.comY.ptr)
    s.sync()


def ramp(packings, phi_end, dphi, fire_kwargs, maxSteps, warm=None, block=8, graph=True, log=None,
         boxSizes=None, budget_s=None, dt0=1e-3, applications=1, shape_match=False):
    """Minimize the packings at their starting packing fraction, then step the packing fraction
    by dphi up to phi_end (or until budget_s seconds have passed), minimizing after each step.
    shape_match: after each step every polygon is put back to its initial shape (scaled to
    its new target area) by pcp_shape_project instead of rescaled by pcp_reset_areas.
    Returns (list of per-step dicts: phi, FIRE steps per replica, seconds, converged, energies;
    the System)."""
    if boxSizes is None:
        boxSizes = [box_size_for(p["diameter"] * np.sqrt(phi_end / float(np.sum(p["targetAreas"]))))
                    for p in packings]
    s = capi.System(packings, boxSizes, model=SQUARED)
    ref = capi.DeviceArray.from_numpy(rest_shapes(packings)) if shape_match else None
    prm = s.fire_params(**fire_kwargs)
    ell = np.array([float(np.mean(p["targetEdgeV"])) for p in packings])
    phi = packing_fractions(s)
    out = []
    k = 0
    tstart = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        if k > 0:
            prev = phi
            phi = np.minimum(phi + dphi, phi_end)
            set_phi(s, phi, ref)
            ell = ell * np.sqrt(phi / prev)
        tw = 0.0
        if warm is not None:
            for _ in range(applications):
                s.force_eval()
                tw += warm.apply(s, ell)
        s.fire_init(prm, dt0)
        st = s.fire_run(prm, maxSteps, block, graph)
        el = time.perf_counter() - t0
        rec = dict(step=k, phi=phi.tolist(), fire_steps=[q.steps for q in st],
                   converged=[bool(q.converged) for q in st], seconds=el, warm_seconds=tw,
                   energy=[q.energy for q in st], max_force=[q.maxForce for q in st])
        out.append(rec)
        if log:
            log(f"[compress] step {k}: phi {phi.mean():.4f}, FIRE steps {int(np.mean(rec['fire_steps']))} "
                f"(max {max(rec['fire_steps'])}), {sum(rec['converged'])}/{s.P} converged, {el:.2f} s")
        if np.all(phi >= phi_end - 1e-12):
            break
        if budget_s is not None and time.perf_counter() - tstart > budget_s:
            break
        k += 1
    return out, s
