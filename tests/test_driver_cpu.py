"""The reference's Python front end as mirrored by pycudapolygon_amd.pyCudaPolygon, on the CPU.

* getDf (pyCudaPolygon.py:348-393): the 2x8 Jacobian of the crossing point of two segments,
  checked against central differences of the crossing point computed here.
* saveModel / loadModel (pyCudaPolygon.py:733-762, 775-817): the on-disk format (scalars.dat
  with "key:\\tvalue" lines, state.npy = nArray ++ positions, arrays.npz with the targets),
  written and read back through a recording stand-in for the native Model (the native one
  needs a GPU; tests/test_driver_gpu.py does the round trip through it), and the reference's
  own testSaver files (tests/golden/testsaver.json) read as the reference reads them: the
  old-format model enum "0" falls back to "normal", and the cell side ends at the 0.5 default
  because loadModel re-initializes the cells (pyCudaPolygon.py:810 -> :54-57).
"""
import json
import os

import numpy as np
import pytest

from pycudapolygon_amd.pyCudaPolygon import model as PM

HERE = os.path.dirname(os.path.abspath(__file__))


def _crossing(v):
    vi, vzi, vj, vzj = v[0:2], v[2:4], v[4:6], v[6:8]
    di, dj = vzi - vi, vzj - vj
    # vi + u di = vj + s dj  ->  solve the 2x2 system
    A = np.array([[di[0], -dj[0]], [di[1], -dj[1]]])
    u, _ = np.linalg.solve(A, vj - vi)
    return vi + u * di


@pytest.mark.parametrize("seed", range(5))
def test_getDf_is_the_jacobian_of_the_crossing_point(seed):
    rng = np.random.default_rng(seed)
    while True:
        v = rng.uniform(0.3, 0.7, 8)
        vi, vzi, vj, vzj = v[0:2], v[2:4], v[4:6], v[6:8]
        di, dj = vzi - vi, vzj - vj
        den = di[0] * dj[1] - di[1] * dj[0]
        if abs(den) > 0.02:
            break
    df = PM.getDf(None, vi, vzi, vj, vzj)
    h = 1e-6
    fd = np.zeros((2, 8))
    for k in range(8):
        a, b = v.copy(), v.copy()
        a[k] += h
        b[k] -= h
        fd[:, k] = (_crossing(a) - _crossing(b)) / (2 * h)
    np.testing.assert_allclose(df, fd, rtol=0, atol=1e-7)


class _Recorder:
    """Stand-in for the native Model: what saveModel reads, what loadModel sets."""

    def __init__(self, **kw):
        self.v = dict(modelEnum="normal", numVertices=0, maxEdgeLength=0.5, stiffness=0.0, compressibility=0.0,
                      nArray=np.zeros(0, int), positions=np.zeros(0), targetAreas=np.zeros(0),
                      targetEdgeLengths=np.zeros(0))
        self.v.update(kw)
        self.calls = []

    def getModelEnum(self): return self.v["modelEnum"]
    def getNumVertices(self): return self.v["numVertices"]
    def getMaxEdgeLength(self): return self.v["maxEdgeLength"]
    def getStiffness(self): return self.v["stiffness"]
    def getCompressibility(self): return self.v["compressibility"]
    def getnArray(self): return self.v["nArray"]
    def getPositions(self): return self.v["positions"]
    def getTargetAreas(self): return self.v["targetAreas"]
    def getTargetEdgeLengths(self): return self.v["targetEdgeLengths"]

    def setModelEnum(self, name):
        if name not in ("normal", "edgeOnly", "areaOnly", "softBody", "hybrid", "abnormal"):
            raise Exception("That Model type does not exist")
        self.v["modelEnum"] = name

    def setNumVertices(self, n): self.v["numVertices"] = n
    def setnArray(self, a): self.v["nArray"] = np.asarray(a)
    def setPositions(self, p): self.v["positions"] = np.asarray(p)
    def setMaxEdgeLength(self, m=0.5): self.v["maxEdgeLength"] = m
    def setStiffness(self, k): self.v["stiffness"] = k
    def setCompressibility(self, c): self.v["compressibility"] = c
    def setTargetAreas(self, a): self.v["targetAreas"] = np.asarray(a)
    def setTargetEdgeLengths(self, t): self.v["targetEdgeLengths"] = np.asarray(t)

    def initializeNeighborCells(self):
        self.calls.append("initializeNeighborCells")
        self.v["maxEdgeLength"] = 0.5


def test_save_load_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    src = _Recorder(modelEnum="softBody", numVertices=7, maxEdgeLength=0.125, stiffness=2.5, compressibility=0.75,
                    nArray=np.array([3, 4]), positions=rng.uniform(0, 1, 14), targetAreas=np.array([0.01, 0.02]),
                    targetEdgeLengths=np.array([0.05, 0.06]))
    d = str(tmp_path / "pack")
    PM.saveModel(src, d)
    with open(os.path.join(d, "scalars.dat")) as f:
        lines = f.read().splitlines()
    assert lines == ["modelEnum:\tsoftBody", "numVertices:\t7", "maxEdgeLength:\t0.125", "stiffness:\t2.5",
                     "compressibility:\t0.75"]
    state = np.load(os.path.join(d, "state.npy"), allow_pickle=False)
    assert np.array_equal(state, np.concatenate(([3, 4], src.v["positions"])))
    with pytest.raises(Exception, match="Packing exists"):
        PM.saveModel(src, d)
    PM.saveModel(src, d, overwrite=True)
    dst = _Recorder()
    PM.loadModel(dst, d)
    for k in ("modelEnum", "numVertices", "stiffness", "compressibility"):
        assert dst.v[k] == src.v[k], k
    for k in ("nArray", "positions", "targetAreas", "targetEdgeLengths"):
        assert np.array_equal(dst.v[k], src.v[k]), k
    assert dst.calls == ["initializeNeighborCells"] and dst.v["maxEdgeLength"] == 0.5


def test_load_reference_testsaver_files(tmp_path, capsys):
    g = json.load(open(os.path.join(HERE, "golden", "testsaver.json")))
    d = tmp_path / "testSaver"
    d.mkdir()
    (d / "scalars.dat").write_text(f"modelEnum:\t0\nnumVertices:\t{len(g['positions']) // 2}\n"
                                   f"maxEdgeLength:\t{g['maxEdgeLength_saved']}\n")
    np.save(d / "state", np.concatenate((np.asarray(g["nArray"], np.float64), g["positions"])))
    dst = _Recorder()
    PM.loadModel(dst, str(d))
    assert "model enum not found" in capsys.readouterr().out
    assert dst.v["modelEnum"] == "normal" and dst.v["numVertices"] == 10
    assert np.array_equal(dst.v["nArray"], g["nArray"]) and np.array_equal(dst.v["positions"], g["positions"])
    assert dst.v["maxEdgeLength"] == 0.5 and dst.calls == ["initializeNeighborCells"]
