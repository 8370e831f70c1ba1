"""Batched FIRE with compaction (pycudapolygon_amd/batch.py) on the GPU: continuing the
unconverged packings in a smaller context must not change any trajectory, so steps,
energies and final positions equal those of the uncompacted batch bit for bit."""
import numpy as np
import pytest

from pycudapolygon_amd import capi, batch
from pycudapolygon_amd.generate import make_batch

pytestmark = pytest.mark.gpu

SQUARED = capi.NORMAL | 0x200


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


def test_compaction_keeps_trajectories_bit_identical():
    pk = make_batch(12, 21, N_range=(32, 160), n_range=(8, 20), min_center_frac=0.5)
    kw = dict(fTol=1e-10, dtMax=100.0)
    print("[batch] 12 packings, uncompacted", flush=True)
    a = batch.minimize_batch(pk, SQUARED, kw, 1500, compact_below=0.0)
    print("[batch] 12 packings, compacted", flush=True)
    b = batch.minimize_batch(pk, SQUARED, kw, 1500, compact_below=0.75, min_context=1,
                             log=lambda m: print(m, flush=True))
    assert len(a.contexts) == 1 and len(b.contexts) > 1
    assert (a.steps >= 0).sum() >= 4   # 6 of these 12 converge within 1500 steps
    np.testing.assert_array_equal(a.steps, b.steps)
    np.testing.assert_array_equal(a.energy, b.energy)
    for x, y in zip(a.positions, b.positions):
        np.testing.assert_array_equal(x, y)


def test_autotuned_batch_gives_the_same_trajectories():
    """minimize_batch with the schedule picked by pcp_autotune on its first context (kept for
    the compacted contexts): the same steps, energies and positions as the default schedule,
    with squared pair areas."""
    pk = make_batch(12, 23, N_range=(32, 160), n_range=(8, 20), min_center_frac=0.5)
    kw = dict(fTol=1e-10, dtMax=100.0)
    pair = capi.NORMAL | 0x400
    a = batch.minimize_batch(pk, pair, kw, 1500, compact_below=0.75, min_context=1)
    b = batch.minimize_batch(pk, pair, kw, 1500, compact_below=0.75, min_context=1, autotune_reps=3)
    print(f"[batch] autotuned schedule {b.schedule}", flush=True)
    assert b.schedule is not None and a.schedule is None and len(b.contexts) > 1
    np.testing.assert_array_equal(a.steps, b.steps)
    np.testing.assert_array_equal(a.energy, b.energy)
    for x, y in zip(a.positions, b.positions):
        np.testing.assert_array_equal(x, y)
