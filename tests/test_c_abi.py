"""CPU-only checks of the drop-in boundary: the C ABI library loads and exports every
function include/pcp.h declares, the ctypes table covers them all, and the pybind11
module exposes every method name the reference binds (src/pybindWrapper.cpp:29-107)."""
import ctypes
import os
import re

from pycudapolygon_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

REFERENCE_BOUND = """initializeNeighborCells sortKeys getKeys setNumVertices setPositions setForces setModelEnum
setStartIndices setMaxEdgeLength setTargetEdgeLengths setTargetAreas setStiffness setCompressibility
getStiffness getCompressibility updatePolygonGeometry projectForce shakeProject getLastShakeIters
saveTentativePositions getMaxEffectiveForce updateNeighborCells updateNeighbors updateValidAndCounts
updateCompactedIntersections updateOutersections updateOverlapArea updateForceEnergy updatePositions
resetAreas getNumVertices getNumPolygons getShapeId getPositions getIntersectionsCounter getModelEnum
getStartIndices getMaxEdgeLength getAreas getNeighborCells getNeighborIndices getIntersections
getNumIntersections getOutersections getNeighbors getNumNeighbors getBoxCounts getInsideFlag
getPerimeters getForces getTU getUT getShapeCounts getEnergy getConstraints getTargetEdgeLengths
getTargetAreas getEdgeLengths getMaxUnbalancedForce getCOM getOverlapArea resetVelocities
minimizeFIREStep minimizeFIRE""".split()


def declared():
    with open(os.path.join(ROOT, "include", "pcp.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pcp_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    names = declared()
    assert len(names) >= 40
    assert "pcp_force_eval" in names and "pcp_fire_minimize" in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert sorted(capi.SIGNATURES) == declared()
    capi.lib()   # sets restype/argtypes for every entry


def test_error_reporting_without_a_gpu_call():
    # argument validation happens before any device work
    L = capi.lib()
    assert L.pcp_last_error() is not None


def test_module_binds_every_reference_method():
    from pycudapolygon_amd.pyCudaPolygonLink import libpyCudaPolygon as lpcp
    missing = [m for m in REFERENCE_BOUND if not hasattr(lpcp.Model, m)]
    assert not missing, missing
    assert [e for e in ("GD", "FIRE") if hasattr(lpcp.minimizerEnum, e)] == ["GD", "FIRE"]
    for name in ("normal", "edgeOnly", "areaOnly", "softBody", "hybrid", "abnormal"):
        assert hasattr(lpcp.modelEnum, name)
    # reference ordinal order (enumTypes.h:6)
    assert int(lpcp.modelEnum.abnormal) == 4 and int(lpcp.modelEnum.hybrid) == 5


def test_python_model_class_exposes_fire_driver():
    from pycudapolygon_amd import model
    for m in ("minimizeFIRE", "minimizeFIRELoop", "minimizeFIREStep", "initForceEnergy", "setnArray",
              "setMaxEdgeLength", "setPhi", "saveModel", "loadModel", "getConstrainedForcePython", "getDf"):
        assert hasattr(model, m)
