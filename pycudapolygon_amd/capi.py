"""ctypes binding of the C ABI in include/pcp.h (libpcp_hip.so).

This is the path the parity tests and the benchmark drive: numpy arrays in, device
buffers managed through the library's own allocator, HIP kernels underneath.  There
is no fallback -- if the library is missing, loading raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpcp_hip.so")

NORMAL, EDGE_ONLY, AREA_ONLY, SOFT_BODY, ABNORMAL, HYBRID = range(6)

_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p
_sz = ctypes.c_size_t


class PcpState(ctypes.Structure):
    _fields_ = [("pos", _p), ("vel", _p), ("force", _p), ("targetEdgeV", _p), ("targetAreas", _p),
                ("edgeLengths", _p), ("areas", _p), ("energy", _p), ("stiffness", _d),
                ("compressibility", _d), ("model", _i), ("perimStiffness", _d)]


class FireParams(ctypes.Structure):
    _fields_ = [("dtMax", _d), ("alphaStart", _d), ("fAlpha", _d), ("fInc", _d), ("fDec", _d),
                ("nMin", _i), ("shakeIter", _i), ("fTol", _d), ("xpbdIter", _i), ("xpbdTol", _d)]


class XpbdParams(ctypes.Structure):
    _fields_ = [("edgeStiffness", _d), ("areaStiffness", _d), ("perimStiffness", _d), ("dt", _d),
                ("nIter", _i), ("tol", _d)]


class FireStatus(ctypes.Structure):
    _fields_ = [("dt", _d), ("alpha", _d), ("energy", _d), ("maxForce", _d), ("nPos", _i),
                ("steps", _i), ("converged", _i), ("lastShakeIters", _i), ("shakeFailed", _i),
                ("rolledBack", _i)]


# name -> (restype, argtypes); every function declared in include/pcp.h
SIGNATURES = {
    "pcp_last_error": (ctypes.c_char_p, []),
    "pcp_device_count": (_i, [_p]),
    "pcp_malloc": (_i, [_p, _sz]),
    "pcp_free": (_i, [_p]),
    "pcp_memcpy_h2d": (_i, [_p, _p, _sz]),
    "pcp_memcpy_d2h": (_i, [_p, _p, _sz]),
    "pcp_memcpy_d2d": (_i, [_p, _p, _sz]),
    "pcp_memset": (_i, [_p, _i, _sz]),
    "pcp_create": (_i, [_p]),
    "pcp_destroy": (_i, [_p]),
    "pcp_synchronize": (_i, [_p]),
    "pcp_set_topology": (_i, [_p, _i, _p, _p]),
    "pcp_set_box_sizes": (_i, [_p, _p]),
    "pcp_set_max_neighbors": (_i, [_p, _i]),
    "pcp_get_max_neighbors": (_i, [_p, _p]),
    "pcp_update_geometry": (_i, [_p, _p, _p, _p, _p, _p, _p, _p]),
    "pcp_update_cells": (_i, [_p, _p]),
    "pcp_update_neighbors": (_i, [_p, _p, _p]),
    "pcp_update_records": (_i, [_p, _p, _p]),
    "pcp_update_force_energy": (_i, [_p, _i, _p, _p, _p, _p, _p, _d, _d, _p, _p]),
    "pcp_project_force": (_i, [_p, _p, _p]),
    "pcp_shake": (_i, [_p, _p, _p, _p, _i, _d, _p]),
    "pcp_xpbd": (_i, [_p, _p, _p, _p, _i, _d, _p]),
    "pcp_xpbd_ex": (_i, [_p, _p, _p, _p, ctypes.POINTER(XpbdParams), _p]),
    "pcp_max_abs": (_i, [_p, _p, _p]),
    "pcp_update_positions": (_i, [_p, _p, _p, _d]),
    "pcp_reset_areas": (_i, [_p, _p, _p, _p, _p, _p]),
    "pcp_shape_project": (_i, [_p, _p, _p, _p]),
    "pcp_scale_targets": (_i, [_p, _p, _p, _p, _p]),
    "pcp_max_effective_force": (_i, [_p, _p, _p, _p, _d, _p]),
    "pcp_overlap_count": (_i, [_p, _p, _i, _p, _p]),
    "pcp_overlap_count_noalloc": (_i, [_p, _p, _i, _p, _p]),
    "pcp_get_cells": (_i, [_p, _i, _p, _p, _p]),
    "pcp_get_num_intersections": (_i, [_p, _i, _p]),
    "pcp_get_records": (_i, [_p, _i, _p, _p, _p, _p, _p]),
    "pcp_get_neighbor_lists": (_i, [_p, _i, _p, _p, _p]),
    "pcp_get_compacted_records": (_i, [_p, _i, _p, _p]),
    "pcp_sort_keys": (_i, [_p, _i, _i, _p, _p]),
    "pcp_force_eval": (_i, [_p, ctypes.POINTER(PcpState)]),
    "pcp_fire_init": (_i, [_p, ctypes.POINTER(PcpState), ctypes.POINTER(FireParams), _d]),
    "pcp_fire_step": (_i, [_p, ctypes.POINTER(PcpState), ctypes.POINTER(FireParams), _p]),
    "pcp_fire_minimize": (_i, [_p, ctypes.POINTER(PcpState), ctypes.POINTER(FireParams), _i, _p]),
    "pcp_fire_run": (_i, [_p, ctypes.POINTER(PcpState), ctypes.POINTER(FireParams), _i, _i, _i, _p]),
    "pcp_fire_set": (_i, [_p, _p, _p, _p]),
    "pcp_fire_get": (_i, [_p, _p]),
    "pcp_profile_kernel": (_i, [_p, ctypes.c_char_p]),
    "pcp_profile_read": (_i, [_p, _p, _p]),
    "pcp_stream_copy": (_i, [_p, _p, _p, ctypes.c_longlong]),
    "pcp_detect_stats": (_i, [_p, _p, _p, _p]),
    "pcp_odd_crossing_records": (_i, [_p, _p]),
    "pcp_set_schedule": (_i, [_p, _i, _i, _i]),
    "pcp_get_schedule": (_i, [_p, _p, _p, _p]),
    "pcp_autotune": (_i, [_p, ctypes.POINTER(PcpState), _i, _p]),
    "pcp_set_detect_schedule": (_i, [_p, _i]),
    "pcp_get_detect_schedule": (_i, [_p, _p]),
    "pcp_mpnn_sizes": (_i, [_i, _p, _p]),
    "pcp_mpnn_graph": (_i, [_p, _p, _p]),
    "pcp_mpnn_features": (_i, [_p, ctypes.POINTER(PcpState), _p, _p, _i, _p, _p]),
    "pcp_mpnn_forward": (_i, [_p, _i, _p, _p, _p, _p, _p]),
    "pcp_mpnn_loss": (_i, [_p, _i, _p, _p, _p, _p, _p]),
    "pcp_mpnn_backward": (_i, [_p, _i, _p, _p, _p, _p, _p, _p, _p]),
    "pcp_mpnn_adam": (_i, [_p, _p, _p, _p, _p, _i, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                           ctypes.c_float]),
    "pcp_mpnn_targets": (_i, [_p, _p, _p, _p, _d, _p]),
    "pcp_mpnn_apply": (_i, [_p, _p, _p, _p, _d, _p]),
    "pcp_mpnn_guard": (_i, [_p, _p, _p, _d, _p, _p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m pycudapolygon_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class PcpError(RuntimeError):
    pass


def check(status):
    if status != 0:
        raise PcpError(lib().pcp_last_error().decode())


def device_count():
    n = ctypes.c_int(0)
    check(lib().pcp_device_count(ctypes.byref(n)))
    return n.value


def _ptr(a):
    return a.ctypes.data_as(_p)


class DeviceArray:
    """A device allocation mirroring a numpy array's dtype and shape."""

    def __init__(self, shape, dtype):
        self.dtype = np.dtype(dtype)
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _p()
        check(lib().pcp_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().pcp_memcpy_h2d(self.ptr, _ptr(a), self.nbytes))

    def get(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().pcp_memcpy_d2h(_ptr(out), self.ptr, self.nbytes))
        return out

    def zero(self):
        check(lib().pcp_memset(self.ptr, 0, self.nbytes))

    def __del__(self):
        try:
            if self.ptr:
                lib().pcp_free(self.ptr)
        except Exception:
            pass
        self.ptr = None


class Context:
    """One pcp context: P packings sharing one HIP stream."""

    def __init__(self):
        h = _p()
        check(lib().pcp_create(ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().pcp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        check(getattr(lib(), name)(self.h, *args))


class System:
    """Device-resident packings (one or many) + the state arrays of the C ABI.

    packings: list of dicts with startIndices, positions, targetEdgeV, targetAreas
    (as produced by pycudapolygon_amd.generate), boxSizes: cells per side per packing.
    """

    def __init__(self, packings, boxSizes, model=NORMAL, stiffness=0.0, compressibility=0.0,
                 maxNeighbors=100, perimStiffness=0.0):
        self.ctx = Context()
        self.P = len(packings)
        starts, packPoly, pos, tev, ta = [0], [0], [], [], []
        voff = 0
        for pk in packings:
            st = np.asarray(pk["startIndices"], np.int64)
            starts.extend((st[1:] + voff).tolist())
            packPoly.append(packPoly[-1] + len(st) - 1)
            voff += int(st[-1])
            pos.append(np.asarray(pk["positions"], np.float64))
            tev.append(np.asarray(pk["targetEdgeV"], np.float64))
            ta.append(np.asarray(pk["targetAreas"], np.float64))
        self.polyStart = np.asarray(starts, np.int32)
        self.packPoly = np.asarray(packPoly, np.int32)
        self.packVert = self.polyStart[self.packPoly]
        self.N = int(self.packPoly[-1])
        self.V = int(self.polyStart[-1])
        self.ctx.call("pcp_set_topology", self.P, _ptr(self.packPoly), _ptr(self.polyStart))
        self.boxSizes = np.asarray(boxSizes, np.int32).reshape(self.P)
        self.ctx.call("pcp_set_box_sizes", _ptr(self.boxSizes))
        self.ctx.call("pcp_set_max_neighbors", int(maxNeighbors))
        self.pos = DeviceArray.from_numpy(np.concatenate(pos))
        self.vel = DeviceArray((2 * self.V,), np.float64); self.vel.zero()
        self.force = DeviceArray((2 * self.V,), np.float64); self.force.zero()
        self.tev = DeviceArray.from_numpy(np.concatenate(tev))
        self.ta = DeviceArray.from_numpy(np.concatenate(ta))
        self.edgeLengths = DeviceArray((self.V,), np.float64)
        self.areas = DeviceArray((self.N,), np.float64)
        self.comX = DeviceArray((self.N,), np.float64)
        self.comY = DeviceArray((self.N,), np.float64)
        self.constraints = DeviceArray((6 * self.V,), np.float64)
        self.maxEdge = DeviceArray((self.P,), np.float64)
        self.energy = DeviceArray((self.P,), np.float64); self.energy.zero()
        self.scratchP = DeviceArray((self.P,), np.float64)
        self.model = model
        self.stiffness = stiffness
        self.compressibility = compressibility
        self._state = PcpState(self.pos.ptr, self.vel.ptr, self.force.ptr, self.tev.ptr, self.ta.ptr,
                               self.edgeLengths.ptr, self.areas.ptr, self.energy.ptr, stiffness,
                               compressibility, model, perimStiffness)

    # --- reference operations (single calls through the C ABI)
    def geometry(self, constraints=True):
        self.ctx.call("pcp_update_geometry", self.pos.ptr, self.edgeLengths.ptr, self.areas.ptr,
                      self.comX.ptr, self.comY.ptr, self.constraints.ptr if constraints else None,
                      self.maxEdge.ptr)

    def cells(self):
        self.ctx.call("pcp_update_cells", self.pos.ptr)

    def neighbors(self):
        mx = ctypes.c_int(0)
        self.ctx.call("pcp_update_neighbors", self.pos.ptr, ctypes.byref(mx))
        return mx.value

    def records(self, exact_pairing=False):
        n = np.zeros(self.P, np.int32)
        self.ctx.call("pcp_update_records", self.pos.ptr if exact_pairing else None, _ptr(n))
        return n

    def forces(self, model=None):
        m = self.model if model is None else model
        self.ctx.call("pcp_update_force_energy", int(m), self.pos.ptr, self.edgeLengths.ptr, self.areas.ptr,
                      self.tev.ptr, self.ta.ptr, float(self.stiffness), float(self.compressibility),
                      self.force.ptr, self.energy.ptr)

    def project(self):
        self.ctx.call("pcp_project_force", self.pos.ptr, self.force.ptr)

    def shake(self, maxIter, tol=1e-15):
        it = np.zeros(self.P, np.int32)
        self.ctx.call("pcp_shake", self.pos.ptr, self.tev.ptr, self.ta.ptr, int(maxIter), float(tol), _ptr(it))
        return it

    def xpbd(self, nIter, tol=0.0, targetEdgeP=None):
        """pcp_xpbd (the reference's hard edge + area projection); targetEdgeP: one edge target
        per polygon (default: the first edge target of each polygon)"""
        if targetEdgeP is None:
            targetEdgeP = self.tev.get()[self.polyStart[:-1]]
        tp = DeviceArray.from_numpy(np.asarray(targetEdgeP, np.float64))
        it = np.zeros(self.P, np.int32)
        self.ctx.call("pcp_xpbd", self.pos.ptr, tp.ptr, self.ta.ptr, int(nIter), float(tol), _ptr(it))
        return it

    def xpbd_ex(self, kE, kA, kP, dt, nIter, tol=0.0):
        """pcp_xpbd_ex: compliant XPBD with the context's per-edge and area targets"""
        prm = XpbdParams(float(kE), float(kA), float(kP), float(dt), int(nIter), float(tol))
        it = np.zeros(self.P, np.int32)
        self.ctx.call("pcp_xpbd_ex", self.pos.ptr, self.tev.ptr, self.ta.ptr, ctypes.byref(prm), _ptr(it))
        return it

    def overlap_count(self, D, counter=True):
        """pcp_overlap_count on the cells of the last cells() call: (sum per packing,
        counter [P, D*D] or None)"""
        cnt = DeviceArray((self.P * D * D,), np.int32) if counter else None
        sums = np.zeros(self.P, np.int64)
        self.ctx.call("pcp_overlap_count", self.pos.ptr, int(D), cnt.ptr if cnt is not None else None, _ptr(sums))
        return sums, (cnt.get().reshape(self.P, D * D) if cnt is not None else None)

    def max_abs_force(self):
        self.ctx.call("pcp_max_abs", self.force.ptr, self.scratchP.ptr)
        self.sync()
        return self.scratchP.get()

    def force_eval(self):
        self.ctx.call("pcp_force_eval", ctypes.byref(self._state))

    def sync(self):
        self.ctx.call("pcp_synchronize")

    # --- read-back (packing p, packing-local ids)
    def vertex_range(self, p):
        return int(self.packVert[p]), int(self.packVert[p + 1])

    def get_positions(self, p=None):
        x = self.pos.get()
        if p is None:
            return x
        a, b = self.vertex_range(p)
        return x[2 * a:2 * b]

    def get_forces(self, p=None):
        self.sync()
        f = self.force.get()
        if p is None:
            return f
        a, b = self.vertex_range(p)
        return f[2 * a:2 * b]

    def get_energy(self):
        self.sync()
        return self.energy.get()

    def get_cells(self, p):
        a, b = self.vertex_range(p)
        B = int(self.boxSizes[p])
        cl = np.zeros(b - a, np.int32); ni = np.zeros(b - a, np.int32); cpb = np.zeros(B * B, np.int32)
        self.ctx.call("pcp_get_cells", int(p), _ptr(cl), _ptr(ni), _ptr(cpb))
        return cl, ni, cpb

    def get_records(self, p):
        n = ctypes.c_int(0)
        self.ctx.call("pcp_get_num_intersections", int(p), ctypes.byref(n))
        I = n.value
        nP = int(self.packPoly[p + 1] - self.packPoly[p])
        inter = np.zeros(max(I, 1), np.uint64); outer = np.zeros(max(I, 1), np.uint64)
        tu = np.zeros(max(2 * I, 2)); ut = np.zeros(max(2 * I, 2)); sc = np.zeros(max(nP, 1), np.int32)
        self.ctx.call("pcp_get_records", int(p), _ptr(inter), _ptr(tu), _ptr(outer), _ptr(ut), _ptr(sc))
        return dict(intersections=inter[:I], tu=tu[:2 * I], outersections=outer[:I], ut=ut[:2 * I],
                    shapeCounts=sc[:nP], numIntersections=I)

    def get_compacted_records(self, p):
        """Crossings in compaction order (edge by edge, each edge's crossings in the
        order of the reference's stencil scan), before the record sort."""
        n = ctypes.c_int(0)
        self.ctx.call("pcp_get_num_intersections", int(p), ctypes.byref(n))
        I = n.value
        inter = np.zeros(max(I, 1), np.uint64)
        tu = np.zeros(max(2 * I, 2))
        self.ctx.call("pcp_get_compacted_records", int(p), _ptr(inter), _ptr(tu))
        return dict(intersections=inter[:I], tu=tu[:2 * I])

    def get_neighbor_lists(self, p):
        a, b = self.vertex_range(p)
        M = ctypes.c_int(0)
        self.ctx.call("pcp_get_max_neighbors", ctypes.byref(M))
        M = M.value
        nb = np.zeros((b - a) * M, np.int32); ins = np.zeros((b - a) * M, np.uint8); num = np.zeros(b - a, np.int32)
        self.ctx.call("pcp_get_neighbor_lists", int(p), _ptr(nb), _ptr(ins), _ptr(num))
        return dict(neighbors=nb, inside=ins, numNeighbors=num, maxNeighbors=M)

    # --- FIRE
    def fire_params(self, fTol=1e-10, dtMax=0.1, alphaStart=0.1, fAlpha=0.99, fInc=1.1, fDec=0.5, nMin=5,
                    shakeIter=5, xpbdIter=0, xpbdTol=0.0):
        return FireParams(dtMax, alphaStart, fAlpha, fInc, fDec, nMin, shakeIter, fTol, int(xpbdIter),
                          float(xpbdTol))

    def fire_init(self, prm, dt0=1e-3):
        check(lib().pcp_fire_init(self.ctx.h, ctypes.byref(self._state), ctypes.byref(prm), float(dt0)))

    def fire_step(self, prm):
        st = (FireStatus * self.P)()
        check(lib().pcp_fire_step(self.ctx.h, ctypes.byref(self._state), ctypes.byref(prm), st))
        return list(st)

    def fire_minimize(self, prm, maxSteps):
        st = (FireStatus * self.P)()
        check(lib().pcp_fire_minimize(self.ctx.h, ctypes.byref(self._state), ctypes.byref(prm),
                                      int(maxSteps), st))
        return list(st)

    def set_schedule(self, pair=16, lens=16, exterior=0):
        self.ctx.call("pcp_set_schedule", int(pair), int(lens), int(exterior))

    def get_schedule(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.ctx.call("pcp_get_schedule", ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def set_detect_schedule(self, mode):
        """0 = culled edge pairs, 1 = whole edge tiles, -1 = by the context's shape"""
        self.ctx.call("pcp_set_detect_schedule", int(mode))

    def odd_crossing_records(self):
        """records of the last pairing in polygon pairs with an odd crossing count (0 = healthy)"""
        n = ctypes.c_int()
        self.ctx.call("pcp_odd_crossing_records", ctypes.byref(n))
        return n.value

    def get_detect_schedule(self):
        m = ctypes.c_int()
        self.ctx.call("pcp_get_detect_schedule", ctypes.byref(m))
        return m.value

    def autotune(self, reps=20):
        """time every lanes-per-polygon candidate of the per-record kernels and both detection
        schedules, keep the fastest; returns (schedule, microseconds per launch of
        pair/lens/exterior x 16/32/64); the detection times are kept in self.detect_us"""
        t = np.zeros(11)
        check(lib().pcp_autotune(self.ctx.h, ctypes.byref(self._state), int(reps), _ptr(t)))
        self.detect_us = t[9:11].copy()
        return self.get_schedule(), t[:9].reshape(3, 3)

    def fire_run(self, prm, maxSteps, stepsPerBlock=16, graph=True):
        """FIRE in blocks of stepsPerBlock steps with one status read per block (pcp_fire_run)"""
        st = (FireStatus * self.P)()
        check(lib().pcp_fire_run(self.ctx.h, ctypes.byref(self._state), ctypes.byref(prm), int(maxSteps),
                                 int(stepsPerBlock), 1 if graph else 0, st))
        return list(st)

    def close(self):
        self.ctx.close()
