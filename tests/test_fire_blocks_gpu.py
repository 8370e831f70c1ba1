"""FIRE in blocks without per-step host round trips (pcp_fire_run), with and without HIP
graph capture, against the step-by-step entry point pcp_fire_step on the same packings:
device-side rollback restores the very values a fresh evaluation gives, so positions,
energies, step counts and FIRE scalars must agree bit for bit."""
import numpy as np
import pytest

from pycudapolygon_amd import capi
from pycudapolygon_amd.generate import make_batch, box_size_for

pytestmark = pytest.mark.gpu

SQUARED = capi.NORMAL | 0x200


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


def system(pk):
    return capi.System(pk, [box_size_for(p["diameter"]) for p in pk], model=SQUARED)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("block", [1, 8])
def test_fire_run_matches_fire_step(graph, block):
    pk = make_batch(6, 31, N_range=(48, 160), n_range=(8, 16), min_center_frac=0.5)
    maxSteps = 700
    print(f"[fire_run] graph={graph} block={block}", flush=True)
    a = system(pk)
    prm = a.fire_params(fTol=1e-10, dtMax=100.0)
    a.fire_init(prm, 1e-3)
    rolled = 0
    for _ in range(maxSteps):
        st = a.fire_step(prm)
        rolled += sum(q.rolledBack for q in st)
        if all(q.converged for q in st):
            break
    b = system(pk)
    b.fire_init(prm, 1e-3)
    sb = b.fire_run(prm, maxSteps, block, graph)
    assert rolled > 0   # the device-side rollback path was exercised
    np.testing.assert_array_equal(a.get_positions(), b.get_positions())
    np.testing.assert_array_equal(a.get_energy(), b.get_energy())
    np.testing.assert_array_equal(a.get_forces(), b.get_forces())
    for qa, qb in zip(st, sb):
        assert (qa.steps, qa.converged, qa.nPos) == (qb.steps, qb.converged, qb.nPos)
        assert (qa.dt, qa.alpha, qa.maxForce) == (qb.dt, qb.alpha, qb.maxForce)
