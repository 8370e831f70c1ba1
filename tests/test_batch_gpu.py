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


@pytest.mark.parametrize("model,n,phis,cap", [(SQUARED, 16, (0.82, 0.86), 3000),
                                              (capi.NORMAL | 0x400, 24, (0.82, 0.86), 3000),
                                              (capi.NORMAL, 16, (0.30, 0.86), 400)],
                         ids=["squared-lens-16", "squared-pair-24", "reference-16"])
def test_converged_packings_skipped_without_changing_trajectories(model, n, phis, cap):
    """A context of 2^17 vertices or more in several packings leaves the packings that have
    converged out of the detection and record kernels of its FIRE steps (their forces and
    energies stay those of their last evaluation, at the positions they keep).  Two packings
    in one context, run in graph-replayed blocks, give every packing the steps, energy and
    positions of its own one-packing run bit for bit; the packing that converges first is
    skipped while the other runs on.  The reference energy (whose FIRE stalls short of
    max|F| <= 1e-10) pairs a sparse packing that converges at once with a dense one run to the
    step cap."""
    from pycudapolygon_amd.generate import make_packing
    N = 131072 // (2 * n) + 1
    # the sparse packing of the reference case: circles apart, no overlap, converged at once
    mcf = (1.15, 0.5) if model == capi.NORMAL else (0.5, 0.5)
    pk = [make_packing(N, n, phi, seed=70 + k, min_center_frac=mcf[k]) for k, phi in enumerate(phis)]
    assert sum(int(p["startIndices"][-1]) for p in pk) >= 1 << 17
    kw = dict(fTol=1e-10, dtMax=100.0)
    print(f"[batch] 2 x N={N} x {n}, model {model:#x}, one context", flush=True)
    both = batch.minimize_batch(pk, model, kw, cap, compact_below=0.0)
    print(f"[batch] steps {both.steps}", flush=True)
    assert both.contexts == [2]
    assert (both.steps >= 0).any() and both.steps[0] != both.steps[1], both.steps   # one runs on alone
    for k in range(2):
        print(f"[batch] packing {k} alone", flush=True)
        one = batch.minimize_batch([pk[k]], model, kw, cap)
        assert one.steps[0] == both.steps[k]
        assert one.energy[0] == both.energy[k] and one.max_force[0] == both.max_force[k]
        np.testing.assert_array_equal(one.positions[0], both.positions[k])


def test_packings_of_a_full_context_equal_their_own_contexts():
    """The evaluation workload at the scale the chip is filled at: 32 packings of N=4096 x 16
    (reference energy, uniform centres) in one context.  Every packing's records, energy and
    forces equal those of the packing evaluated alone, bit for bit (the record rows by key,
    the forces as IEEE bit patterns), under the schedule the context's shape selects (the
    whole-tile detection)."""
    from pycudapolygon_amd.generate import make_packing, box_size_for
    P = 32
    ps = [make_packing(4096, 16, 0.85, seed=k) for k in range(P)]
    bs = [box_size_for(q["diameter"]) for q in ps]
    print(f"[batch] {P} x N=4096 x 16 in one context", flush=True)
    s = capi.System(ps, bs)
    for _ in range(4):   # the first evaluations may enlarge the record buckets
        try:
            s.force_eval()
            s.sync()
        except capi.PcpError:
            pass
    s.force_eval()
    s.sync()
    E = s.get_energy()
    for k in (0, 7, 19, 31):
        print(f"[batch] packing {k} alone", flush=True)
        one = capi.System([ps[k]], [bs[k]])
        for _ in range(3):
            try:
                one.force_eval()
                one.sync()
            except capi.PcpError:
                pass
        one.force_eval()
        one.sync()
        assert E[k] == one.get_energy()[0]
        assert np.array_equal(s.get_forces(k).view(np.uint64), one.get_forces(0).view(np.uint64))
        ra, rb = s.get_records(k), one.get_records(0)
        assert ra["numIntersections"] == rb["numIntersections"] > 0
        for key in ("intersections", "outersections"):
            assert np.array_equal(np.sort(np.asarray(ra[key])), np.sort(np.asarray(rb[key])))
        one.close()
    s.close()
