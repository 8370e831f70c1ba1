"""Host-side pieces that need no GPU: the benchmark's one-GPU guard and the rest shapes
used by shape matching."""
import os
import subprocess
import sys

import numpy as np

from pycudapolygon_amd import compress
from pycudapolygon_amd.generate import make_packing

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_rest_shapes_are_centred_minimum_image_offsets():
    """Each polygon's rest shape has zero mean, reproduces its edge vectors, and does not change
    when the packing is translated across the periodic boundary."""
    p = make_packing(32, 12, 0.8, seed=4)
    r = compress.rest_shapes([p]).reshape(-1, 2)
    x = p["positions"].reshape(-1, 2)
    st = p["startIndices"]
    shifted = dict(p, positions=((x + np.array([0.47, 0.83])) % 1.0).reshape(-1))
    r2 = compress.rest_shapes([shifted]).reshape(-1, 2)
    for a, b in zip(st[:-1], st[1:]):
        np.testing.assert_allclose(r[a:b].mean(axis=0), 0.0, atol=1e-16)
        e = (np.roll(x[a:b], -1, axis=0) - x[a:b] + 0.5) % 1.0 - 0.5
        np.testing.assert_allclose(np.roll(r[a:b], -1, axis=0) - r[a:b], e, atol=1e-15)
    np.testing.assert_allclose(r2, r, atol=1e-15)
