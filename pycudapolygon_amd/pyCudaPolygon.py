"""The `model` class: the reference's Python front end (camdennis/pyCudaPolygon
pyCudaPolygon/pyCudaPolygon.py, class model) over this build's libpyCudaPolygon.

Method names, arguments, defaults and driver behaviour follow the reference (cited
per method) so that existing calls keep working.  Out of scope and not provided: the
matplotlib drawing and random-polygon mixins (polygonMixins/*).
"""
import os

import numpy as np

from .pyCudaPolygonLink import libpyCudaPolygon as lpcp
from . import enums

_MODEL_NAMES = ("normal", "edgeOnly", "areaOnly", "softBody", "hybrid", "abnormal")


def _wrap(x):
    return (x + 1.5) % 1.0 - 0.5


class model(lpcp.Model):
    # ---------------------------------------------------------------- initializers
    def __init__(self, size=0, seed=None, modelType="normal", stiffness=0, compressibility=0):
        """pyCudaPolygon.py:39-48"""
        lpcp.Model.__init__(self, size)
        self.setModelEnum(modelType)
        self.setStiffness(stiffness)
        self.setCompressibility(compressibility)
        self.rng = np.random.default_rng(seed)
        self.size = size

    def initializeRandomSeed(self, seed):
        """pyCudaPolygon.py:50-52"""
        self.rng = np.random.default_rng(seed)
        lpcp.Model.initializeRandomSeed(self, seed)

    def initializeNeighborCells(self):
        """pyCudaPolygon.py:54-57: (re)allocate, then reset the cell side to the 0.5 default"""
        lpcp.Model.initializeNeighborCells(self)
        self.setMaxEdgeLength()

    def initForceEnergy(self):
        """pyCudaPolygon.py:59-65"""
        if self.getModelEnum() == "normal":
            self.initializeNeighborCells()
            self.updateNeighborCells()
            self.updateNeighbors()
            self.updateOutersections()

    # ---------------------------------------------------------------- setters
    def setModelEnum(self, modelType):
        """pyCudaPolygon.py:72-86"""
        if modelType not in _MODEL_NAMES:
            raise Exception("That Model type does not exist")
        lpcp.Model.setModelEnum(self, getattr(enums.modelEnum, modelType))

    def setNumVertices(self, n):
        self.size = n
        lpcp.Model.setNumVertices(self, n)

    def setnArray(self, nArray):
        """pyCudaPolygon.py:95-101"""
        nArray = np.asarray(nArray).astype(int)
        if nArray.sum() != self.getNumVertices():
            raise ValueError(f"Total vertices from nArray ({nArray.sum()}) "
                             f"does not match model size ({self.getNumVertices()})")
        lpcp.Model.setStartIndices(self, np.concatenate(([0], np.cumsum(nArray))).astype(int).tolist())

    def setPositions(self, positions):
        lpcp.Model.setPositions(self, np.asarray(positions, dtype=np.float64).tolist())

    def setAreas(self, targetAreas):
        """pyCudaPolygon.py:103-144: rescale each polygon about its centroid to its target area"""
        nArray = self.getnArray()
        pos = self.getPositions().reshape(self.getNumVertices(), 2)
        s = 0
        for i, n in enumerate(nArray):
            poly = pos[s:s + n].copy()
            poly[1:] = poly[0] + _wrap(poly[1:] - poly[0])
            x, y = poly[:, 0], poly[:, 1]
            area = 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(np.roll(x, -1), y))
            if area > 0:
                c = poly.mean(axis=0)
                pos[s:s + n] = np.mod((poly - c) * np.sqrt(targetAreas[i] / area) + c, 1.0)
            s += n
        self.setPositions(pos.reshape(-1))
        self.updatePolygonGeometry()

    def setMonoArea(self, phi=1):
        """pyCudaPolygon.py:146-156"""
        n = self.getNumPolygons()
        targetAreas = np.ones(n) * phi / n
        if np.max(targetAreas) > 1 / 9:
            raise Exception("The phi you have chosen has caused the shapes to be too large and "
                            "compromised the PBCs")
        self.updatePolygonGeometry()
        self.setAreas(targetAreas)

    def setPhi(self, phi):
        """pyCudaPolygon.py:223-234 (the definition in effect): scale targets to phi, then
        rescale each polygon about its COM (resetAreas)"""
        self.updatePolygonGeometry()
        targetAreas = self.getTargetAreas()
        ratio = phi / np.sum(targetAreas)
        self.setTargetAreas(targetAreas * ratio)
        self.setTargetEdgeLengths(self.getTargetEdgeLengths() * np.sqrt(ratio))
        self.updatePolygonGeometry()
        self.resetAreas()
        self.updatePolygonGeometry()

    def setMaxEdgeLength(self, maxEdgeLength=0.5):
        """pyCudaPolygon.py:165-180: None = the longest current edge"""
        if maxEdgeLength is None:
            start = self.getStartIndices()
            P = self.getPositions().reshape(-1, 2)
            maxEdgeLength = 0.0
            for a, b in zip(start[:-1], start[1:]):
                poly = P[a:b]
                d = _wrap(np.roll(poly, -1, axis=0) - poly)
                maxEdgeLength = max(maxEdgeLength, float(np.sqrt((d ** 2).sum(1)).max()))
        lpcp.Model.setMaxEdgeLength(self, maxEdgeLength)

    def setBiPerimeters(self, kappa, ratio=1.4):
        """pyCudaPolygon.py:182-203"""
        nArray = self.getnArray()
        self.setMaxEdgeLength()
        self.initializeNeighborCells()
        self.updateNeighborCells()
        self.updateNeighbors()
        mid = len(nArray) // 2
        self.updatePolygonGeometry()
        targetAreas = np.ones(len(nArray)) * np.min(self.getAreas())
        targetAreas[:mid] /= ratio ** 2
        self.setTargetAreas(targetAreas)
        self.setTargetEdgeLengths(kappa * np.sqrt(targetAreas) / nArray)
        self.updatePolygonGeometry()

    def setTargetEdgeLengths(self, targetEdgeLengths):
        lpcp.Model.setTargetEdgeLengths(self, np.asarray(targetEdgeLengths, np.float64).tolist())

    def setTargetAreas(self, targetAreas):
        lpcp.Model.setTargetAreas(self, np.asarray(targetAreas, np.float64).tolist())

    # ---------------------------------------------------------------- getters
    def getNumVertices(self):
        return int(lpcp.Model.getNumVertices(self))

    def getShapeId(self):
        return np.array(lpcp.Model.getShapeId(self))

    def getPositions(self):
        return np.array(lpcp.Model.getPositions(self))

    def getNumNeighbors(self):
        return np.array(lpcp.Model.getNumNeighbors(self))

    def getNeighbors(self):
        """pyCudaPolygon.py:265-277: {edge: neighbouring edges}"""
        v = self.getNumVertices()
        nb = np.array(lpcp.Model.getNeighbors(self)).reshape(v, -1)
        num = self.getNumNeighbors()
        return {i: nb[i, :num[i]] for i in range(v) if num[i] > 0}

    def getInsideFlag(self):
        return np.array(lpcp.Model.getInsideFlag(self))

    def getTU(self):
        return np.array(lpcp.Model.getTU(self))

    def getUT(self):
        return np.array(lpcp.Model.getUT(self))

    def getIntersections(self):
        return np.array(lpcp.Model.getIntersections(self), dtype=np.uint64)

    def getKeys(self):
        return np.array(lpcp.Model.getKeys(self))

    def getOutersections(self):
        return np.array(lpcp.Model.getOutersections(self), dtype=np.uint64)

    def getNeighborCells(self):
        return np.array(lpcp.Model.getNeighborCells(self))

    def getBoxCounts(self):
        return np.array(lpcp.Model.getBoxCounts(self))

    def getNeighborIndices(self):
        return np.array(lpcp.Model.getNeighborIndices(self))

    def getnArray(self):
        return np.diff(self.getStartIndices())

    def getForces(self):
        return np.array(lpcp.Model.getForces(self))

    def getCentersOfMass(self):
        """pyCudaPolygon.py:325-338: mean of the vertices unwrapped about vertex 0"""
        pos = self.getPositions().reshape(-1, 2)
        out = []
        for a, b in zip(self.getStartIndices()[:-1], self.getStartIndices()[1:]):
            p0 = pos[a]
            out.append(np.mean(_wrap(pos[a:b] - p0), axis=0) + p0)
        return np.concatenate(out) if out else np.zeros(0)

    def getShapeCounts(self):
        return np.array(lpcp.Model.getShapeCounts(self))

    def getIntersectionsCounter(self):
        c = np.array(lpcp.Model.getIntersectionsCounter(self))
        s = int(np.sqrt(len(c)))
        return c.reshape(s, s)

    def getConstraints(self):
        return np.array(lpcp.Model.getConstraints(self))

    def getAreas(self):
        return np.array(lpcp.Model.getAreas(self))

    def getTargetEdgeLengths(self):
        return np.array(lpcp.Model.getTargetEdgeLengths(self))

    def getTargetAreas(self):
        return np.array(lpcp.Model.getTargetAreas(self))

    def getEdgeLengths(self):
        return np.array(lpcp.Model.getEdgeLengths(self))

    def getCOM(self):
        return np.array(lpcp.Model.getCOM(self))

    def getPhi(self):
        return np.sum(self.getAreas())

    def getRestEdgeLengths(self):
        return np.array(lpcp.Model.getEdgeLengths(self))

    def getConstraintViolation(self):
        """pyCudaPolygon.py:395-408: rms relative area / edge / perimeter violations"""
        areas, el = self.getAreas(), self.getEdgeLengths()
        ta, tl = self.getTargetAreas(), self.getTargetEdgeLengths()
        nArray = self.getnArray()
        idx = np.repeat(np.arange(len(nArray)), nArray)
        per = np.bincount(idx, weights=el, minlength=len(nArray))
        return np.array([np.sqrt(np.mean((1 - areas / ta) ** 2)),
                         np.sqrt(np.mean((1 - el / np.repeat(tl, nArray)) ** 2)),
                         np.sqrt(np.mean((1 - per / (nArray * tl)) ** 2))])

    def getDf(self, vi, vzi, vj, vzj):
        """pyCudaPolygon.py:348-393: 2x8 Jacobian of the crossing point f = vi + u*(vzi - vi)
        of edges (vi, vzi) and (vj, vzj) with respect to (vi, vzi, vj, vzj), by the
        quotient rule on u = (dj x dij)/(dj x di)."""
        vi, vzi, vj, vzj = (np.asarray(a, np.float64) for a in (vi, vzi, vj, vzj))
        dj, di, dij = _wrap(vzj - vj), _wrap(vzi - vi), _wrap(vj - vi)
        w = dj[0] * di[1] - dj[1] * di[0]
        u = (dj[0] * dij[1] - dj[1] * dij[0]) / w
        # d(numerator)/d(vi, vzi, vj, vzj) and d(w)/d(...) as (2 x 4) [component, vertex]
        dk = np.array([[dj[1], 0.0, -dij[1] - dj[1], dij[1]],
                       [-dj[0], 0.0, dj[0] + dij[0], -dij[0]]])
        dw = np.array([[dj[1], -dj[1], -di[1], di[1]],
                       [-dj[0], dj[0], di[0], -di[0]]])
        du = dk / w - u * dw / w                      # (2 x 4): [beta, p]
        df = np.zeros((2, 8))
        for p in range(4):
            df[:, 2 * p:2 * p + 2] = np.outer(di, du[:, p])
        df[0, 0] += 1 - u; df[1, 1] += 1 - u
        df[0, 2] += u; df[1, 3] += u
        return df

    def getConstrainedForcePython(self, force):
        """pyCudaPolygon.py:441-487: remove from each polygon's force its components along
        the edge-length gradients and the area gradient (QR of the gradient rows)."""
        start, pos = self.getStartIndices(), self.getPositions()
        out = np.array(force, dtype=np.float64).copy()
        for a, b in zip(start[:-1], start[1:]):
            n = b - a
            P = pos[2 * a:2 * b].reshape(n, 2)
            rows = np.zeros((n + 1, 2 * n))
            t = _wrap(np.roll(P, -1, axis=0) - P)
            t /= np.linalg.norm(t, axis=1)[:, None]
            for e in range(n):
                f = (e + 1) % n
                rows[e, 2 * e:2 * e + 2] = -t[e]
                rows[e, 2 * f:2 * f + 2] = t[e]
            d = _wrap(np.roll(P, -1, axis=0) - np.roll(P, 1, axis=0))
            rows[n, 0::2], rows[n, 1::2] = d[:, 1] / 2, -d[:, 0] / 2
            Q, _ = np.linalg.qr(rows.T)
            seg = out[2 * a:2 * b]
            out[2 * a:2 * b] = seg - Q @ (Q.T @ seg)
        return out

    # ---------------------------------------------------------------- updaters / helpers
    def shakeProject(self, nIter=100, tol=1e-15):
        return lpcp.Model.shakeProject(self, nIter, tol)

    def getMaxEffectiveForce(self, dt, minimizerType="GD"):
        m = enums.minimizerEnum.GD if minimizerType == "GD" else enums.minimizerEnum.FIRE
        return lpcp.Model.getMaxEffectiveForce(self, dt, m)

    def z(self, i):
        start = self.getStartIndices()
        s = self.getShapeId()[i]
        return start[s] if i == start[s + 1] - 1 else i + 1

    def zp(self, i):
        start = self.getStartIndices()
        s = self.getShapeId()[i]
        return start[s + 1] - 1 if i == start[s] else i - 1

    @staticmethod
    def _unpack(keys):
        keys = np.asarray(keys, dtype=np.uint64)
        return np.vstack([(keys >> np.uint64(s)) & np.uint64(0xFFFF) for s in (48, 32, 16, 0)]).T

    def unpackIntersections(self):
        """pyCudaPolygon.py:552-558: columns (sj, si, i, j)"""
        return self._unpack(self.getIntersections())

    def unpackOutersections(self):
        return self._unpack(self.getOutersections())

    # ---------------------------------------------------------------- minimizers
    def minimizeGDStep(self, dt=1e-3, addedForce=None, dontMove=False, maxDisplacement=0.05, nIter=100,
                       tol=1e-15, minimizerType="GD"):
        """pyCudaPolygon.py:573-595"""
        if self.getModelEnum() not in ("edgeOnly", "areaOnly", "softBody"):
            self.updateNeighborCells()
            self.updateNeighbors()
            self.updateOutersections()
        self.updatePolygonGeometry()
        self.updateForceEnergy()
        if addedForce is not None:
            self.setForces(self.getForces() + addedForce)
        if self.getMaxUnbalancedForce() * dt > maxDisplacement:
            return -1, -1
        if dontMove:
            return self.getEnergy(), 0
        if self.getModelEnum() == "normal":
            self.projectForce()
        self.updatePositions(dt)
        if self.getModelEnum() == "normal" and nIter > 0:
            self.shakeProject(nIter, tol)
        self.updatePolygonGeometry()
        return self.getEnergy(), 0

    def minimizeGD(self, maxUnbalancedForceThreshold=1e-14, dt=1e-3, maxSteps=-1, addedForce=None,
                   progressBar=False, checkpointDir=None, checkpointFreq=1, overwriteCheckpoint=False,
                   maxDisplacement=0.5, maxConstraintViolationThreshold=1e-3, nIter=100, tol=1e-15,
                   minimizerType="GD"):
        """pyCudaPolygon.py:597-647: gradient descent with adaptive dt"""
        if maxSteps == -1 and maxUnbalancedForceThreshold is None:
            raise ValueError("maxSteps=-1 requires maxUnbalancedForceThreshold to be set")
        dontMove = maxSteps == 0
        if dontMove:
            maxSteps = 1
        successes, prevEnergy, step, energy = 0, None, 0, None
        while maxSteps == -1 or step < maxSteps:
            if checkpointDir is not None and step % checkpointFreq == 0:
                os.makedirs(checkpointDir, exist_ok=True)
                self.saveModel(os.path.join(checkpointDir, str(step)), overwrite=overwriteCheckpoint)
            energy, _ = self.minimizeGDStep(addedForce=addedForce, dt=dt, dontMove=dontMove,
                                            maxDisplacement=maxDisplacement, nIter=nIter, tol=tol,
                                            minimizerType=minimizerType)
            normal = self.getModelEnum() == "normal"
            fEff = self.getMaxEffectiveForce(dt, minimizerType) if normal else self.getMaxUnbalancedForce()
            if maxUnbalancedForceThreshold is not None and fEff <= maxUnbalancedForceThreshold:
                return energy, 0, np.array([0])
            if normal and np.max(self.getConstraintViolation()) > maxConstraintViolationThreshold:
                return energy, 0, np.array([0])
            if prevEnergy is not None and (energy > prevEnergy or energy == -1):
                dt, successes = max(1e-32, dt / 2.1), 0
            else:
                successes, prevEnergy = successes + 1, energy
            if successes > 5:
                successes, dt = 0, min(1e4, dt * 1.9)
            step += 1
        if energy is None:
            raise Exception("This is not an appropriate maximum step size. Reset maxSteps.")
        return energy, dt, np.array([0, 0, 0])

    def minimizeFIREStep(self, dt, alpha, nPos, dtMax=0.1, alphaStart=0.1, fAlpha=0.99, fInc=1.1, fDec=0.5,
                         nMin=5, shakeIter=5):
        """pyCudaPolygon.py:649-651 -> Model::minimizeFIREStep (model.cpp:405-486)"""
        return lpcp.Model.minimizeFIREStep(self, dt, alpha, nPos, dtMax, alphaStart, fAlpha, fInc, fDec, nMin,
                                           shakeIter)

    def minimizeFIRELoop(self, maxForceThreshold=1e-14, dt=1e-3, maxSteps=100000, dtMax=0.1, alphaStart=0.1,
                         fAlpha=0.99, fInc=1.1, fDec=0.5, nMin=5, shakeIter=5, progressBar=False):
        """pyCudaPolygon.py:653-672: generator of (energy, maxForce, dt, step)"""
        lpcp.Model.resetVelocities(self)
        self.updatePolygonGeometry()
        self.updateForceEnergy()
        if self.getModelEnum() == "normal":
            self.projectForce()
        alpha, nPos = alphaStart, 0
        for step in range(maxSteps):
            energy, dt, alpha, nPos = self.minimizeFIREStep(dt, alpha, nPos, dtMax, alphaStart, fAlpha, fInc,
                                                            fDec, nMin, shakeIter)
            maxForce = self.getMaxUnbalancedForce()
            yield energy, maxForce, dt, step + 1
            if maxForce <= maxForceThreshold:
                return

    def minimizeFIRE(self, maxForceThreshold=1e-14, dt=1e-3, maxSteps=100000000, dtMax=0.1, alphaStart=0.1,
                     fAlpha=0.99, fInc=1.1, fDec=0.5, nMin=5, shakeIter=5, progressBar=False, checkpointDir=None,
                     checkpointFreq=1, overwriteCheckpoint=False):
        """pyCudaPolygon.py:674-731: fresh records, one GD kick to expose containments, then
        FIRE steps; also stops once the energy stays within 1e-15 for more than 5 steps.
        Returns (energy, dt, steps, [min, mean, max] SHAKE iterations)."""
        lpcp.Model.resetVelocities(self)
        if self.getModelEnum() == "normal":
            self.initForceEnergy()
        self.updatePolygonGeometry()
        self.updateForceEnergy()
        if self.getModelEnum() == "normal":
            self.projectForce()
        if self.getModelEnum() == "normal" and maxSteps > 0:
            self.updatePositions(dt)
            self.updatePolygonGeometry()
            self.updateNeighborCells()
            self.updateNeighbors()
            self.updateOutersections()
            self.updateForceEnergy()
            self.projectForce()
            lpcp.Model.resetVelocities(self)
        if maxSteps == 0:
            return self.getEnergy(), dt, 0
        energy, steps = self.getEnergy(), 0
        lo, hi, tot, flat, prevEnergy = 1e9, 0, 0, 0, 1e9
        for energy, _, dt, steps in self.minimizeFIRELoop(maxForceThreshold=maxForceThreshold, dt=dt,
                                                          maxSteps=maxSteps, dtMax=dtMax, alphaStart=alphaStart,
                                                          fAlpha=fAlpha, fInc=fInc, fDec=fDec, nMin=nMin,
                                                          shakeIter=shakeIter, progressBar=progressBar):
            if checkpointDir is not None and steps % checkpointFreq == 0:
                os.makedirs(checkpointDir, exist_ok=True)
                self.saveModel(os.path.join(checkpointDir, str(steps)), overwrite=overwriteCheckpoint)
            it = self.getLastShakeIters()
            lo, hi, tot = min(lo, it), max(hi, it), tot + it
            if abs(energy - prevEnergy) < 1e-15:
                flat += 1
            if flat > 5:
                break
            prevEnergy = energy
        return energy, dt, steps, np.array([lo, tot / maxSteps, hi])

    # ---------------------------------------------------------------- persistence
    def saveModel(self, dirName, overwrite=False):
        """pyCudaPolygon.py:733-762: scalars.dat, state.npy (nArray ++ positions), arrays.npz"""
        if not overwrite and os.path.isdir(dirName):
            raise Exception("Packing exists. Not saving. To save over this file, set kwarg overwrite = True")
        os.makedirs(dirName, exist_ok=True)
        kv = {"modelEnum": str(self.getModelEnum()), "numVertices": self.getNumVertices(),
              "maxEdgeLength": self.getMaxEdgeLength(), "stiffness": self.getStiffness(),
              "compressibility": self.getCompressibility()}
        with open(os.path.join(dirName, "scalars.dat"), "w") as f:
            for k, v in kv.items():
                f.write(f"{k}:\t{v}\n")
        np.save(os.path.join(dirName, "state"), np.concatenate((self.getnArray(), self.getPositions())))
        np.savez(os.path.join(dirName, "arrays"), targetAreas=self.getTargetAreas(),
                 targetEdgeLengths=self.getTargetEdgeLengths())

    def loadModel(self, dirName):
        """pyCudaPolygon.py:775-817"""
        state = np.load(os.path.join(dirName, "state.npy"), allow_pickle=False)
        vals = {"modelEnum": "normal", "numVertices": "0", "maxEdgeLength": "0.5", "stiffness": "0.0",
                "compressibility": "0.0"}
        with open(os.path.join(dirName, "scalars.dat")) as f:
            for line in f:
                k, v = line.rstrip("\n").split("\t")
                vals[k.rstrip(":")] = v
        try:
            self.setModelEnum(vals["modelEnum"])
        except Exception:
            print("Warning: model enum not found. Setting to normal")
            self.setModelEnum("normal")
        self.setNumVertices(int(vals["numVertices"]))
        V = self.getNumVertices()
        self.setnArray(state[:-2 * V].astype(int))
        self.setPositions(state[-2 * V:].copy())
        self.setMaxEdgeLength(float(vals["maxEdgeLength"]))
        self.setStiffness(float(vals["stiffness"]))
        self.setCompressibility(float(vals["compressibility"]))
        # the reference calls its own initializeNeighborCells here, which resets the cell
        # side to the 0.5 default after allocating (pyCudaPolygon.py:810, :54-57)
        self.initializeNeighborCells()
        arrays = os.path.join(dirName, "arrays.npz")
        if os.path.exists(arrays):
            a = np.load(arrays, allow_pickle=False)
            self.setTargetAreas(a["targetAreas"])
            self.setTargetEdgeLengths(a["targetEdgeLengths"])

    def minimizeSoftBody(self):
        """pyCudaPolygon.py:764-773"""
        initial = self.getModelEnum()
        self.setStiffness(1)
        self.setCompressibility(1)
        self.updateForceEnergy()
        self.setModelEnum("softBody")
        self.minimizeFIRE(dt=0.1, maxForceThreshold=1e-14)
        self.resetAreas()
        self.updatePolygonGeometry()
        self.setModelEnum(initial)

    def makeSubModel(self, sub):
        """pyCudaPolygon.py:819-837: keep only the polygons listed in `sub`"""
        pos0, nArray, start = self.getPositions(), self.getnArray(), self.getStartIndices()
        parts = [pos0[2 * start[s]:2 * start[s + 1]] for s in sub]
        positions = np.concatenate(parts) if parts else np.zeros(0)
        self.setNumVertices(positions.size // 2)
        self.setPositions(positions)
        self.setnArray(nArray[sub])
