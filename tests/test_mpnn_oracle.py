"""CPU checks of the learned-warm-start oracle (oracle/mpnn_oracle.py).

The reference has no learned component, so nothing of it pins the network; the oracle is
pinned by mathematics instead:
  * bf16 rounding against torch's float32 -> bfloat16 conversion (nearest even) and
    hand-worked ties;
  * the backward pass of the unrounded network against central finite differences of
    its forward pass (a dropped term -- e.g. the transpose of the neighbour sum, the
    residual path or a ReLU mask -- fails this);
  * the GPU-order sums against plain float64 sums to float32 accuracy;
  * the contact graph of hand-built crossing records;
  * Adam against the textbook update in float64.
"""
import numpy as np
import pytest
import torch

from oracle import mpnn_oracle as M


def test_bf16_rounding_matches_torch_and_ties():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-20, 100), rng.normal(0, 1e20, 100),
                        [0.0, -0.0, 1.0, -3.5]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(M.bf(x), ref)
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1 (even mantissa);
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6 -> 1 + 2^-6
    t = np.array([1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -20], np.float32)
    np.testing.assert_array_equal(M.bf(t), np.array([1.0, 1 + 2.0 ** -6, 1 + 2.0 ** -7], np.float32))


def ring_graph(V, extra, seed):
    """A symmetric adjacency: polygons of 6 vertices (chains) plus random crossing pairs."""
    rng = np.random.default_rng(seed)
    start = np.arange(0, V + 1, 6)
    assert start[-1] == V
    recs = []
    for _ in range(extra):
        a, b = rng.integers(0, V, 2)
        sa, sb = a // 6, b // 6
        if sa == sb:
            continue
        recs.append((int(sb) << 48) | (int(sa) << 32) | (int(a - 6 * sa) << 16) | int(b - 6 * sb))
    adj, dropped = M.contact_graph(start, np.array(recs, np.uint64))
    return adj, dropped


def test_contact_graph_of_hand_built_records():
    # two triangles (vertices 0-2 and 3-5): edge 1 of polygon 0 crosses edge 2 of polygon 1
    # (record si=0 -> sj=1) and edge 2 of polygon 0 crosses edge 0 of polygon 1 (si=1 -> sj=0)
    start = np.array([0, 3, 6])
    rec = np.array([(1 << 48) | (0 << 32) | (1 << 16) | 2, (0 << 48) | (1 << 32) | (0 << 16) | 2], np.uint64)
    adj, dropped = M.contact_graph(start, rec)
    assert dropped == 0
    assert list(adj[1, :3]) == [0, 2, 5] and list(adj[5, :3]) == [4, 3, 1]
    assert list(adj[2, :3]) == [1, 0, 3] and list(adj[3, :3]) == [5, 4, 2]
    assert (adj[0, 2:] == -1).all() and (adj[4, 2:] == -1).all()


def test_contact_graph_is_symmetric_and_drops_beyond_width():
    adj, dropped = ring_graph(60, 200, 1)
    V = len(adj)
    A = np.zeros((V, V), int)
    for m in range(V):
        for u in adj[m]:
            if u >= 0:
                A[m, u] += 1
    if dropped == 0:
        np.testing.assert_array_equal(A, A.T)
    # one vertex with 20 partners keeps the 14 smallest ids
    start = np.arange(0, 6 * 22 + 1, 6)
    recs = [(int(k) << 48) | (0 << 32) | (0 << 16) | 1 for k in range(1, 21)]
    adj2, d2 = M.contact_graph(start, np.array(recs, np.uint64))
    assert d2 == 6
    assert list(adj2[0, 2:]) == [6 * k + 1 for k in range(1, 15)]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    adj, dropped = ring_graph(30, 40, 2)
    assert dropped == 0
    V = len(adj)
    feat = rng.normal(0, 1, (V, 4))
    params = M.init_params(5).astype(np.float64)
    params[M.OFF_B0:M.OFF_B0 + M.H] = rng.normal(0, 0.3, M.H)   # move ReLU kinks off zero
    for l in range(3):
        o = M.OFF_W1 + l * M.LAYER_NP + 2 * M.H * M.H
        params[o:o + M.H] = rng.normal(0, 0.3, M.H)
    dY = rng.normal(0, 1, (V, 2))

    def f(p):
        y, _ = M.forward(p, feat, adj, exact=True)
        return float((y * dY).sum())

    y, cache = M.forward(params, feat, adj, exact=True)
    g = M.backward(params, feat, adj, cache, dY, exact=True)
    idx = np.concatenate([rng.choice(np.arange(0, 320), 25, replace=False),
                          rng.choice(np.arange(320, 25088), 60, replace=False),
                          np.arange(25088, 25218)])
    h = 1e-6
    fd = []
    for i in idx:
        pp = params.copy(); pp[i] += h
        pm = params.copy(); pm[i] -= h
        fd.append((f(pp) - f(pm)) / (2 * h))
    fd = np.array(fd)
    err = np.abs(fd - g[idx])
    assert err.max() <= 1e-6 * max(1.0, np.abs(fd).max()), (err.max(), np.abs(fd).max())


def test_forward_exact_uses_float64_parameters():
    # the float64 path must not round the parameter vector to float32 (FD above relies on it)
    adj, _ = ring_graph(12, 4, 0)
    feat = np.ones((12, 4))
    p = np.zeros(M.NPARAM)
    p[M.OFF_BO] = 1e-9
    y, _ = M.forward(p, feat, adj, exact=True)
    assert y[0, 0] == 1e-9


def test_gpu_order_sums_agree_with_float64():
    rng = np.random.default_rng(4)
    dZ = rng.normal(0, 1, (5000, M.H)).astype(np.float32)
    b = M._bias_grad_gpu_order(dZ)
    np.testing.assert_allclose(b, dZ.astype(np.float64).sum(0), rtol=0, atol=1e-3)
    a = M._chunked(lambda x, y: M._seqsum(dZ[x:y]), len(dZ))
    np.testing.assert_allclose(a, dZ.astype(np.float64).sum(0), rtol=0, atol=1e-3)


def test_rounded_backward_close_to_exact():
    rng = np.random.default_rng(6)
    adj, _ = ring_graph(600, 400, 3)
    V = len(adj)
    feat = rng.normal(0, 1, (V, 4)).astype(np.float32)
    params = M.init_params(1)
    dY = rng.normal(0, 1, (V, 2)).astype(np.float32)
    y, cache = M.forward(params, feat, adj)
    g = M.backward(params, feat, adj, cache, dY)
    ye, ce = M.forward(params.astype(np.float64), feat.astype(np.float64), adj, exact=True)
    ge = M.backward(params.astype(np.float64), feat.astype(np.float64), adj, ce, dY.astype(np.float64), exact=True)
    assert np.linalg.norm(y - ye) <= 0.05 * np.linalg.norm(ye)
    assert np.linalg.norm(g - ge) <= 0.05 * np.linalg.norm(ge)


def test_adam_matches_textbook():
    rng = np.random.default_rng(7)
    p = rng.normal(0, 1, 100).astype(np.float32)
    m = np.zeros(100, np.float32); v = np.zeros(100, np.float32)
    P, Mm, Vv = p.astype(np.float64), np.zeros(100), np.zeros(100)
    for t in range(1, 6):
        g = rng.normal(0, 1, 100).astype(np.float32)
        p, m, v = M.adam(p, g, m, v, t, 1e-2)
        Mm = 0.9 * Mm + 0.1 * g
        Vv = 0.999 * Vv + 0.001 * g.astype(np.float64) ** 2
        P = P - 1e-2 * (Mm / (1 - 0.9 ** t)) / (np.sqrt(Vv / (1 - 0.999 ** t)) + 1e-8)
    np.testing.assert_allclose(p, P, rtol=0, atol=1e-5)


def test_features_closed_form():
    start = np.array([0, 4, 8])
    F = np.array([[1e-9, -2e-9]] * 8).reshape(-1)
    epoly = np.array([4e-12, 0.0])
    adj = -np.ones((8, M.DEG), np.int32)
    adj[0, 2] = 5
    adj[5, 2] = 0
    ell = 0.01
    x = M.features(start, F, epoly, adj, ell, 4, [1e-3, 1e-3, 1.0, 1e-4])
    np.testing.assert_allclose(x[0, 0], np.arcsinh(1e-9 / (1e-6 * 1e-3)), rtol=1e-6)
    np.testing.assert_allclose(x[0, 1], np.arcsinh(-2e-9 / (1e-6 * 1e-3)), rtol=1e-6)
    assert x[0, 2] == 1 and x[5, 2] == 1 and x[1, 2] == 0
    np.testing.assert_allclose(x[1, 3], np.arcsinh(4e-12 / (4 * 1e-8 * 1e-4)), rtol=1e-6)
    assert x[6, 3] == 0
