"""numpy front end of the CPU oracle (oracle/pcp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Each function
forwards to the C restatement of the reference kernel it names.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PCP_ORACLE_LIB names another build of the same source (the CPU baseline's timing build)
_LIB_PATH = os.environ.get("PCP_ORACLE_LIB") or os.path.join(_HERE, "build", "liboracle.so")

NORMAL, EDGE_ONLY, AREA_ONLY, SOFT_BODY, ABNORMAL, HYBRID = range(6)
CONSISTENT = 0x800   # ORC_CONSISTENT: side-consistent segment decision (OR into the model)
MODEL_IDS = {"normal": NORMAL, "edgeOnly": EDGE_ONLY, "areaOnly": AREA_ONLY,
             "softBody": SOFT_BODY, "abnormal": ABNORMAL, "hybrid": HYBRID}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p


class _System(ctypes.Structure):
    _fields_ = [("V", _i), ("N", _i), ("boxSize", _i), ("model", _i), ("maxNeighbors", _i),
                ("start", _p), ("targetEdgeV", _p), ("targetAreas", _p),
                ("stiffness", _d), ("compressibility", _d), ("solver", _i), ("perimStiffness", _d)]

SOLVERS = {"dense": 0, "lanes": 1}


class FireState(ctypes.Structure):
    _fields_ = [("dt", _d), ("alpha", _d), ("nPos", _i), ("energy", _d),
                ("lastShakeIters", _i), ("shakeFailed", _i), ("rolledBack", _i)]


def _declare(L):
    L.orc_wrap.restype = _d
    L.orc_wrap.argtypes = [_d]
    L.orc_topology.argtypes = [_i, _p, _i, _p, _p, _p]
    L.orc_geometry.argtypes = [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p]
    L.orc_cells.argtypes = [_i, _i, _p, _p, _i, _p, _p, _p]
    L.orc_neighbors.restype = _i
    L.orc_neighbors.argtypes = [_i, _i, _p, _p, _i, _p, _p, _p, _i, _p, _p, _p, _p]
    L.orc_neighbors_ex.argtypes = [_i, _i, _p, _p, _i, _p, _p, _p, _i, _p, _p, _p, _p, _i]
    L.orc_outersections.restype = _i
    L.orc_outersections.argtypes = [_i, _i, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p]
    L.orc_force_energy.restype = _d
    L.orc_force_energy.argtypes = [_i, _i, _i, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _d, _d, _p]
    L.orc_project_force.argtypes = [_i, _i, _p, _p, _p]
    L.orc_shake.restype = _i
    L.orc_shake.argtypes = [_i, _i, _p, _p, _p, _p, _i, _d]
    L.orc_project_force_lanes.restype = _i
    L.orc_project_force_lanes.argtypes = [_i, _i, _p, _p, _p]
    L.orc_shake_lanes.restype = _i
    L.orc_shake_lanes.argtypes = [_i, _i, _p, _p, _p, _p, _i, _d]
    L.orc_shape_project.restype = _i
    L.orc_shape_project.argtypes = [_i, _i, _p, _p, _p, _p]
    L.orc_xpbd.restype = _i
    L.orc_xpbd.argtypes = [_i, _i, _p, _p, _p, _p, _i, _d]
    L.orc_xpbd_lanes.restype = _i
    L.orc_xpbd_lanes.argtypes = [_i, _i, _p, _p, _p, _p, _p, _d, _d, _d, _d, _i, _d]
    L.orc_force_energy_ex.restype = _d
    L.orc_force_energy_ex.argtypes = [_i, _i, _i, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _d, _d, _d,
                                      _p, _p]
    L.orc_force_eval_ex.restype = _d
    L.orc_force_eval_ex.argtypes = [ctypes.POINTER(_System), _p, _p, _p, _p]
    L.orc_fire_step_ex.argtypes = [ctypes.POINTER(_System), _p, _p, _p, ctypes.POINTER(FireState),
                                   _d, _d, _d, _d, _d, _i, _i, _i, _d]
    L.orc_fire_minimize_ex.restype = _i
    L.orc_fire_minimize_ex.argtypes = [ctypes.POINTER(_System), _p, _p, _p, ctypes.POINTER(FireState),
                                       _d, _i, _d, _d, _d, _d, _d, _i, _i, _i, _d]
    L.orc_update_positions.argtypes = [_i, _p, _p, _d]
    L.orc_max_effective_force.restype = _d
    L.orc_max_effective_force.argtypes = [_i, _p, _p, _p, _d]
    L.orc_det_sum.restype = _d
    L.orc_det_sum.argtypes = [_p, _i]
    L.orc_overlap_count.restype = ctypes.c_longlong
    L.orc_overlap_count.argtypes = [_i, _i, _p, _p, _i, _p, _p, _i, _p]
    L.orc_force_eval.restype = _d
    L.orc_force_eval.argtypes = [ctypes.POINTER(_System), _p, _p, _p]
    L.orc_fire_init.argtypes = [ctypes.POINTER(_System), _p, _p, _p, ctypes.POINTER(FireState)]
    L.orc_fire_step.argtypes = [ctypes.POINTER(_System), _p, _p, _p, ctypes.POINTER(FireState),
                                _d, _d, _d, _d, _d, _i, _i]
    L.orc_fire_minimize.restype = _i
    L.orc_fire_minimize.argtypes = [ctypes.POINTER(_System), _p, _p, _p, ctypes.POINTER(FireState),
                                    _d, _i, _d, _d, _d, _d, _d, _i, _i]


def _ptr(a):
    return a.ctypes.data_as(_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def wrap(x):
    return lib().orc_wrap(float(x))


def topology(start):
    start = _i32(start)
    N = len(start) - 1
    V = int(start[-1])
    shapeId = np.zeros(V, np.int32); nxt = np.zeros(V, np.int32); prv = np.zeros(V, np.int32)
    lib().orc_topology(N, _ptr(start), V, _ptr(shapeId), _ptr(nxt), _ptr(prv))
    return shapeId, nxt, prv


def geometry(pos, start):
    """updatePolygonGeometry: dict of edgeLengths, areas, comX, comY, constraints, maxEdgeLength."""
    pos = _f64(pos); start = _i32(start)
    N = len(start) - 1; V = int(start[-1])
    out = dict(edgeLengths=np.zeros(V), areas=np.zeros(N), comX=np.zeros(N), comY=np.zeros(N),
               constraints=np.zeros(6 * V), constraintNormSq=np.zeros(3 * N))
    mx = np.zeros(1)
    lib().orc_geometry(V, N, _ptr(pos), _ptr(start), _ptr(out["edgeLengths"]), _ptr(out["areas"]),
                       _ptr(out["comX"]), _ptr(out["comY"]), _ptr(out["constraints"]),
                       _ptr(out["constraintNormSq"]), _ptr(mx))
    out["maxEdgeLength"] = float(mx[0])
    return out


def cells(pos, start, boxSize):
    pos = _f64(pos); start = _i32(start)
    N = len(start) - 1; V = int(start[-1])
    cl = np.zeros(V, np.int32); ni = np.zeros(V, np.int32); cpb = np.zeros(boxSize * boxSize, np.int32)
    lib().orc_cells(V, N, _ptr(pos), _ptr(start), int(boxSize), _ptr(cl), _ptr(ni), _ptr(cpb))
    return cl, ni, cpb


def neighbors(pos, start, boxSize, maxNeighbors=100, cell=None, consistent=False):
    """updateNeighbors incl. the overflow resize-and-retry of model.cpp:264-305; consistent:
    the side-consistent decision (ORC_CONSISTENT) instead of the parameter test."""
    pos = _f64(pos); start = _i32(start)
    N = len(start) - 1; V = int(start[-1])
    cl, ni, cpb = cell if cell is not None else cells(pos, start, boxSize)
    M = maxNeighbors
    while True:
        nb = np.full(V * M, -1, np.int32); ins = np.zeros(V * M, np.uint8)
        tu = np.zeros(2 * V * M); num = np.zeros(V, np.int32)
        mx = lib().orc_neighbors_ex(V, N, _ptr(pos), _ptr(start), int(boxSize), _ptr(cl), _ptr(ni), _ptr(cpb),
                                    M, _ptr(nb), _ptr(ins), _ptr(tu), _ptr(num), int(bool(consistent)))
        if mx <= M:
            break
        M = max(2 * M + 1, mx)
    return dict(neighbors=nb, inside=ins, tuSlots=tu, numNeighbors=num, maxNeighbors=M, maxCount=mx)


def outersections(start, nbr):
    start = _i32(start)
    N = len(start) - 1; V = int(start[-1])
    M = nbr["maxNeighbors"]
    cap = max(1, int(np.minimum(nbr["numNeighbors"], M).sum()))
    inter = np.zeros(cap, np.uint64); outer = np.zeros(cap, np.uint64)
    tu = np.zeros(2 * cap); ut = np.zeros(2 * cap); sc = np.zeros(max(N, 1), np.int32)
    I = lib().orc_outersections(V, N, _ptr(start), M, _ptr(nbr["neighbors"]), _ptr(nbr["inside"]),
                                _ptr(nbr["tuSlots"]), _ptr(nbr["numNeighbors"]), _ptr(inter), _ptr(tu),
                                _ptr(outer), _ptr(ut), _ptr(sc))
    return dict(intersections=inter[:I], tu=tu[:2 * I], outersections=outer[:I], ut=ut[:2 * I],
                shapeCounts=sc[:N], numIntersections=I)


def odd_crossing_records(intersections):
    """Records (keys [sj|si|i|j]) that belong to a pair of distinct polygons with an odd number
    of crossings: the check that pcp_odd_crossing_records makes on the device.  A closed
    overlap boundary crosses the other polygon's boundary an even number of times."""
    k = np.asarray(intersections, np.uint64)
    si = ((k >> np.uint64(32)) & np.uint64(0xFFFF)).astype(np.int64)
    sj = ((k >> np.uint64(48)) & np.uint64(0xFFFF)).astype(np.int64)
    lo, hi = np.minimum(si, sj), np.maximum(si, sj)
    pair = lo * 65536 + hi
    u, inv, cnt = np.unique(pair, return_inverse=True, return_counts=True)
    odd = (cnt[inv] % 2 == 1) & (si != sj)
    return int(odd.sum()), sorted({(int(a), int(b)) for a, b, o in zip(lo, hi, odd) if o})


def per_vertex_targets(start, targetEdgeLengths):
    start = np.asarray(start)
    return np.repeat(np.asarray(targetEdgeLengths, np.float64), np.diff(start))


def force_energy(model, pos, start, rec, geom, targetEdgeV=None, targetAreas=None,
                 stiffness=0.0, compressibility=0.0, perimStiffness=0.0, want_fext=False):
    """orc_force_energy_ex: (E, force) or, with want_fext, (E, force, fext)."""
    pos = _f64(pos); start = _i32(start)
    N = len(start) - 1; V = int(start[-1])
    mid = MODEL_IDS[model] if isinstance(model, str) else int(model)
    tev = _f64(targetEdgeV if targetEdgeV is not None else np.zeros(V))
    ta = _f64(targetAreas if targetAreas is not None else np.zeros(N))
    force = np.zeros(2 * V)
    I = rec["numIntersections"] if rec is not None else 0
    inter = _as_u64(rec["intersections"]) if rec is not None else np.zeros(1, np.uint64)
    outer = _as_u64(rec["outersections"]) if rec is not None else np.zeros(1, np.uint64)
    tu = _f64(rec["tu"]) if rec is not None else np.zeros(2)
    ut = _f64(rec["ut"]) if rec is not None else np.zeros(2)
    fext = np.zeros(2 * V) if want_fext else None
    E = lib().orc_force_energy_ex(mid, V, N, _ptr(start), _ptr(pos), I, _ptr(inter), _ptr(outer),
                                  _ptr(tu), _ptr(ut), _ptr(_f64(geom["edgeLengths"])),
                                  _ptr(_f64(geom["areas"])), _ptr(tev), _ptr(ta), float(stiffness),
                                  float(compressibility), float(perimStiffness), _ptr(force), _ptr(fext))
    return (E, force, fext) if want_fext else (E, force)


def _as_u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def project_force(pos, start, force):
    pos = _f64(pos); start = _i32(start); f = _f64(force).copy()
    lib().orc_project_force(int(start[-1]), len(start) - 1, _ptr(start), _ptr(pos), _ptr(f))
    return f


def shake(pos, start, targetEdgeV, targetAreas, maxIter, tol=1e-15):
    pos = _f64(pos).copy(); start = _i32(start)
    it = lib().orc_shake(int(start[-1]), len(start) - 1, _ptr(start), _ptr(pos), _ptr(_f64(targetEdgeV)),
                         _ptr(_f64(targetAreas)), int(maxIter), float(tol))
    return pos, it


def project_force_lanes(pos, start, force):
    pos = _f64(pos); start = _i32(start); f = _f64(force).copy()
    rc = lib().orc_project_force_lanes(int(start[-1]), len(start) - 1, _ptr(start), _ptr(pos), _ptr(f))
    if rc != 0:
        raise ValueError("lane-order projection takes polygons of at most 64 vertices")
    return f


def shake_lanes(pos, start, targetEdgeV, targetAreas, maxIter, tol=1e-15):
    pos = _f64(pos).copy(); start = _i32(start)
    it = lib().orc_shake_lanes(int(start[-1]), len(start) - 1, _ptr(start), _ptr(pos), _ptr(_f64(targetEdgeV)),
                               _ptr(_f64(targetAreas)), int(maxIter), float(tol))
    if it < 0:
        raise ValueError("lane-order SHAKE takes polygons of at most 64 vertices")
    return pos, it


def shape_project(pos, start, ref, targetAreas=None):
    pos = _f64(pos).copy(); start = _i32(start)
    ta = _f64(targetAreas) if targetAreas is not None else None
    rc = lib().orc_shape_project(int(start[-1]), len(start) - 1, _ptr(start), _ptr(pos), _ptr(_f64(ref)), _ptr(ta))
    if rc != 0:
        raise ValueError("shape matching takes polygons of at most 64 vertices")
    return pos


def xpbd(pos, start, targetEdgeLengths, targetAreas, nIter, tol=0.0):
    pos = _f64(pos).copy(); start = _i32(start)
    it = lib().orc_xpbd(int(start[-1]), len(start) - 1, _ptr(start), _ptr(pos),
                        _ptr(_f64(targetEdgeLengths)), _ptr(_f64(targetAreas)), int(nIter), float(tol))
    return pos, it


def xpbd_lanes(pos, start, kE, kA, kP, dt, nIter, tol=0.0, targetEdgeV=None, targetEdgeP=None,
               targetAreas=None):
    """Compliant XPBD in the kernel's lane order: (positions, iteration count)."""
    pos = _f64(pos).copy(); start = _i32(start)
    N = len(start) - 1
    tev = _f64(targetEdgeV) if targetEdgeV is not None else None
    tep = _f64(targetEdgeP) if targetEdgeP is not None else None
    ta = _f64(targetAreas) if targetAreas is not None else np.zeros(N)
    it = lib().orc_xpbd_lanes(int(start[-1]), N, _ptr(start), _ptr(pos), _ptr(tev), _ptr(tep), _ptr(ta),
                              float(kE), float(kA), float(kP), float(dt), int(nIter), float(tol))
    if it < 0:
        raise ValueError("XPBD takes polygons of at most 64 vertices")
    return pos, it


def update_positions(pos, force, dt):
    """updatePositionsKernel: frac(x + F dt)"""
    pos = _f64(pos).copy()
    lib().orc_update_positions(len(pos) // 2, _ptr(pos), _ptr(_f64(force)), float(dt))
    return pos


def max_effective_force(pos, tent, force, scale):
    pos = _f64(pos)
    return lib().orc_max_effective_force(len(pos) // 2, _ptr(pos), _ptr(_f64(tent)), _ptr(_f64(force)),
                                         float(scale))


def det_sum(x):
    x = _f64(x)
    return lib().orc_det_sum(_ptr(x), len(x))


def overlap_count(pos, start, boxSize, pointDensity):
    pos = _f64(pos); start = _i32(start)
    cl, ni, cpb = cells(pos, start, boxSize)
    cnt = np.zeros(pointDensity * pointDensity, np.int32)
    s = lib().orc_overlap_count(int(start[-1]), len(start) - 1, _ptr(start), _ptr(pos), int(boxSize),
                                _ptr(ni), _ptr(cpb), int(pointDensity), _ptr(cnt))
    return s, cnt


class System:
    """Holds the arrays behind an orc_system so they outlive the ctypes struct."""

    def __init__(self, start, boxSize, model="normal", targetEdgeV=None, targetAreas=None,
                 stiffness=0.0, compressibility=0.0, maxNeighbors=100, solver="dense", perimStiffness=0.0):
        """solver: "dense" (Householder projection, dense Gauss-Newton SHAKE) or "lanes"
        (the same solves in the HIP kernels' lane order, bit-identical to the library)."""
        self.start = _i32(start)
        N = len(self.start) - 1; V = int(self.start[-1])
        self.tev = _f64(targetEdgeV if targetEdgeV is not None else np.zeros(V))
        self.ta = _f64(targetAreas if targetAreas is not None else np.zeros(N))
        self.s = _System(V, N, int(boxSize), MODEL_IDS[model] if isinstance(model, str) else int(model),
                         int(maxNeighbors), _ptr(self.start), _ptr(self.tev), _ptr(self.ta),
                         float(stiffness), float(compressibility), SOLVERS[solver], float(perimStiffness))

    @property
    def V(self):
        return self.s.V

    def force_eval(self, pos):
        pos = _f64(pos); f = np.zeros(2 * self.V); nI = ctypes.c_int(0)
        E = lib().orc_force_eval(ctypes.byref(self.s), _ptr(pos), _ptr(f), ctypes.byref(nI))
        return E, f, nI.value

    def force_eval_ext(self, pos):
        """(E, total force, force before the springs, records)"""
        pos = _f64(pos); f = np.zeros(2 * self.V); fe = np.zeros(2 * self.V); nI = ctypes.c_int(0)
        E = lib().orc_force_eval_ex(ctypes.byref(self.s), _ptr(pos), _ptr(f), _ptr(fe), ctypes.byref(nI))
        return E, f, fe, nI.value

    def fire_init(self, pos, dt, alpha):
        st = FireState(dt, alpha, 0, 0.0, 0, 0, 0)
        vel = np.zeros(2 * self.V); f = np.zeros(2 * self.V)
        lib().orc_fire_init(ctypes.byref(self.s), _ptr(pos), _ptr(vel), _ptr(f), ctypes.byref(st))
        return st, vel, f

    def fire_step(self, pos, vel, force, st, dtMax=0.1, alphaStart=0.1, fAlpha=0.99, fInc=1.1,
                  fDec=0.5, nMin=5, shakeIter=5, xpbdIter=0, xpbdTol=0.0):
        lib().orc_fire_step_ex(ctypes.byref(self.s), _ptr(pos), _ptr(vel), _ptr(force), ctypes.byref(st),
                               dtMax, alphaStart, fAlpha, fInc, fDec, nMin, shakeIter, int(xpbdIter),
                               float(xpbdTol))

    def fire_minimize(self, pos, fTol=1e-10, dt=1e-3, maxSteps=100000, dtMax=0.1, alphaStart=0.1,
                      fAlpha=0.99, fInc=1.1, fDec=0.5, nMin=5, shakeIter=5, xpbdIter=0, xpbdTol=0.0):
        pos = _f64(pos).copy(); vel = np.zeros(2 * self.V); f = np.zeros(2 * self.V)
        st = FireState(dt, alphaStart, 0, 0.0, 0, 0, 0)
        steps = lib().orc_fire_minimize_ex(ctypes.byref(self.s), _ptr(pos), _ptr(vel), _ptr(f),
                                           ctypes.byref(st), fTol, maxSteps, dtMax, alphaStart, fAlpha,
                                           fInc, fDec, nMin, shakeIter, int(xpbdIter), float(xpbdTol))
        return pos, f, st, steps
