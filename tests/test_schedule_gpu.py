"""Schedules of the per-record kernels (lanes per polygon of k_pair_poly, k_lens_energy and
k_exterior_poly), the two detection schedules (culled edge pairs, k_detect_poly; whole edge
tiles, k_detect_tile) and the autotuner that picks them: every schedule must give the same
forces, energies and records bit for bit, and the autotuner keeps one of the candidates it
timed."""
import itertools

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


def evaluate(s):
    s.force_eval()
    s.sync()
    rec = s.get_records(0)
    return s.get_forces(), s.get_energy(), rec["intersections"], rec["outersections"]


@pytest.mark.parametrize("model", [capi.NORMAL, SQUARED])
def test_every_schedule_gives_identical_results(model):
    pk = make_batch(5, 41, N_range=(64, 256), n_range=(8, 32), min_center_frac=0.5)
    s = capi.System(pk, [box_size_for(p["diameter"]) for p in pk], model=model)
    assert s.get_schedule() == (16, 16, 0)
    ref = evaluate(s)
    for pair, lens, ext in itertools.product((16, 32, 64), (16, 32, 64), (0, 16, 32, 64)):
        s.set_schedule(pair, lens, ext)
        got = evaluate(s)
        for a, b in zip(ref, got):
            np.testing.assert_array_equal(a, b, err_msg=f"schedule {(pair, lens, ext)}")


@pytest.mark.parametrize("shape", ["c1", "n16x4", "uneven"])
def test_detection_schedules_give_identical_results(shape):
    from pycudapolygon_amd.generate import make_packing
    if shape == "c1":
        pk = [make_packing(64, 12, 0.80, seed=3)]
    elif shape == "n16x4":
        pk = [make_packing(1024, 16, 0.85, seed=k) for k in range(4)]
    else:
        pk = make_batch(6, 44, N_range=(64, 512), n_range=(5, 32))
    for model in (capi.NORMAL, SQUARED):
        s = capi.System(pk, [box_size_for(p["maxEdgeLength"]) for p in pk], model=model)
        out = []
        for mode in (0, 1):
            s.set_detect_schedule(mode)
            assert s.get_detect_schedule() == mode
            out.append(evaluate(s) + tuple(s.get_records(k)["intersections"] for k in range(len(pk))))
        for a, b in zip(*out):
            np.testing.assert_array_equal(a, b, err_msg=f"{shape} model {model}")
        s.close()


def test_autotune_keeps_a_timed_candidate():
    pk = make_batch(8, 43, N_range=(256, 1024), n_range=(8, 32), min_center_frac=0.5)
    s = capi.System(pk, [box_size_for(p["diameter"]) for p in pk], model=SQUARED)
    ref = evaluate(s)
    sched, t = s.autotune(reps=10)
    print(f"[schedule] chosen {sched}; us per launch pair/lens/exterior x 16/32/64:\n{t}", flush=True)
    assert (t[0] > 0).all() and (t[1] > 0).all() and (t[2] > 0).any()
    pair, lens, ext = sched
    assert t[0, [16, 32, 64].index(pair)] == t[0].min()
    assert t[1, [16, 32, 64].index(lens)] == t[1].min()
    assert t[2, [16, 32, 64].index(ext)] == t[2][t[2] > 0].min()
    d = s.detect_us
    assert (d > 0).all() and s.get_detect_schedule() == int(np.argmin(d))
    for a, b in zip(ref, evaluate(s)):
        np.testing.assert_array_equal(a, b)
