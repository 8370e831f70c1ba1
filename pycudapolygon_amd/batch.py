"""Batched FIRE minimization of many independent packings with compaction.

One context holds P packings and every kernel covers all of them in one launch, but a
packing that has converged still costs its share of each launch until the whole batch is
done.  minimize_batch() therefore re-packs the batch when few packings remain: the
unconverged ones continue in a new, smaller context built from their current positions,
velocities and FIRE state (dt, alpha, nPos).  Every quantity of a packing is computed from
that packing alone, so the continued trajectories are bit-identical to the uncompacted
ones.  No counterpart in the reference (its Model holds one packing).
"""
import time

import numpy as np

from . import capi
from .generate import box_size_for


class BatchResult:
    def __init__(self, P):
        self.steps = np.full(P, -1)          # FIRE steps to convergence (-1: not converged)
        self.energy = np.full(P, np.nan)
        self.max_force = np.full(P, np.nan)
        self.positions = [None] * P
        self.seconds = 0.0
        self.contexts = []                   # packings per context used, in order
        self.schedule = None                 # lanes per polygon (pair, lens, exterior) when autotuned
        self.schedule_us = None              # the autotuner's microseconds per candidate launch
        self.warm_taken_back = None          # packings the warm start's guard set back (pcp_mpnn_guard)


def _with_positions(p, x):
    q = dict(p)
    q["positions"] = x
    return q


def minimize_batch(packings, model, fire_kwargs, maxSteps, dt0=1e-3, compact_below=0.5, min_context=8,
                   boxSizes=None, log=None, warm=None, block=8, graph=True, autotune_reps=0, small_context=16,
                   small_block=32):
    """FIRE on `packings` until each has max|F| <= fTol or maxSteps steps of its own.
    compact_below: re-pack when fewer than this fraction of the context's packings are
    still running (and the context holds more than min_context packings).
    warm: a warmstart.WarmStart applied once to the first context before FIRE starts
    (its time is included in the result's seconds).  FIRE runs in blocks of `block` steps
    (pcp_fire_run, replayed from a HIP graph when graph is set); the batch is checked for
    convergence and compaction between blocks (blocks of small_block steps once a context holds
    at most small_context packings).  autotune_reps > 0: the lanes-per-polygon
    schedule of the per-record kernels is picked by pcp_autotune on the first context
    (timed over that many launches per candidate; its time is included) and kept for the
    compacted ones.  Every schedule gives the same results bit for bit."""
    P = len(packings)
    res = BatchResult(P)
    if boxSizes is None:
        boxSizes = [box_size_for(p["diameter"]) for p in packings]
    idx = np.arange(P)
    cur = list(packings)
    offset = np.zeros(P, int)           # steps done in earlier contexts
    state = None                        # (vel per packing, dt, alpha, nPos) when continuing
    t0 = time.perf_counter()
    while len(idx):
        s = capi.System(cur, [boxSizes[i] for i in idx], model=model)
        prm = s.fire_params(**fire_kwargs)
        if autotune_reps > 0 and res.schedule is None:
            res.schedule, res.schedule_us = s.autotune(autotune_reps)
        elif res.schedule is not None:
            s.set_schedule(*res.schedule)
        if warm is not None and state is None:
            s.force_eval()
            warm.apply(s, np.array([float(np.mean(p["targetEdgeV"])) for p in cur]))
            res.warm_taken_back = warm.taken_back
        s.fire_init(prm, dt0)
        if state is not None:
            vel, dt, alpha, nPos = state
            s.vel.upload(np.concatenate(vel))
            s.ctx.call("pcp_fire_set", dt.ctypes.data_as(capi._p), alpha.ctypes.data_as(capi._p),
                       nPos.ctypes.data_as(capi._p))
        res.contexts.append(len(idx))
        done = np.zeros(len(idx), bool)
        local = 0
        start = int(offset[idx].max())    # every running packing has run the same number of steps
        # a small context's steps are short (a few microseconds per kernel), so it runs longer
        # blocks per status read: the read and the block-start snapshot cost as much as a step
        kb = block if len(idx) > small_context else small_block
        while True:
            K = min(kb, maxSteps - start - local)
            if K <= 0:
                break
            st = s.fire_run(prm, K, K, graph and K == kb)
            local += K
            for k in range(len(idx)):
                if not done[k] and st[k].converged:
                    done[k] = True
                    res.steps[idx[k]] = offset[idx[k]] + st[k].steps
            n_run = int((~done).sum())
            if n_run == 0 or start + local >= maxSteps:
                break
            if len(idx) > min_context and n_run < compact_below * len(idx):
                break
        x = s.pos.get()
        E = s.get_energy()
        F = s.max_abs_force()
        for k in range(len(idx)):
            a, b = s.vertex_range(k)
            g = idx[k]
            res.positions[g] = x[2 * a:2 * b].copy()
            res.energy[g] = E[k]
            res.max_force[g] = F[k]
        keep = np.where(~done & ~(start + local >= maxSteps))[0]
        if log:
            log(f"[batch] context of {len(idx)} packings: {local} steps, {int(done.sum())} converged, "
                f"{len(keep)} continue")
        if len(keep) == 0:
            break
        v = s.vel.get()
        st_k = [st[k] for k in keep]
        state = ([v[2 * s.vertex_range(k)[0]:2 * s.vertex_range(k)[1]].copy() for k in keep],
                 np.array([q.dt for q in st_k], np.float64), np.array([q.alpha for q in st_k], np.float64),
                 np.array([q.nPos for q in st_k], np.int32))
        for k in keep:
            offset[idx[k]] += st[k].steps
        cur = [_with_positions(packings[idx[k]], res.positions[idx[k]]) for k in keep]
        idx = idx[keep]
        s.close()
    res.seconds = time.perf_counter() - t0
    return res
