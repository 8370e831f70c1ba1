"""GPU parity of the learned warm start (csrc/pcp_mpnn.hip through the C ABI) against the
CPU oracle oracle/mpnn_oracle.py.

Bit-exact: the contact graph (integer), the neighbour sums and their bf16 rounding (X of
every layer), the 4x64 and 64x2 layers, the output-layer gradient and every
float32 sum done in a fixed order.  Matrix-core products (MFMA, float32 accumulation in
hardware order): |gpu - oracle| <= 1e-5 * sum |terms| + 1e-30 element by element, with
the oracle evaluated in float64 from the GPU's own bf16 operands of that layer.  End to
end (roundings may flip a bf16 ulp): relative norm error <= 1e-2 for outputs and each
parameter block of the gradient."""
import ctypes

import numpy as np
import pytest

from oracle import mpnn_oracle as M
from pycudapolygon_amd import capi, warmstart as W
from pycudapolygon_amd.generate import make_packing, box_size_for

pytestmark = pytest.mark.gpu

H, DEG = M.H, M.DEG


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


def evaluated_system(P=2, N=64, n=12, phi=0.85, seed=0):
    pk = [make_packing(N, n, phi, seed=seed + k, min_center_frac=0.5) for k in range(P)]
    s = capi.System(pk, [box_size_for(p["diameter"]) for p in pk], model=W.SQUARED)
    s.force_eval()
    s.sync()
    return s, pk


def gpu_graph(s):
    adj = capi.DeviceArray((s.V * DEG,), np.int32)
    d = ctypes.c_int(-1)
    s.ctx.call("pcp_mpnn_graph", adj.ptr, ctypes.byref(d))
    return adj, adj.get().reshape(s.V, DEG), d.value


def test_contact_graph_matches_records():
    s, pk = evaluated_system(P=3)
    print(f"[mpnn] graph V={s.V}", flush=True)
    _, adj, dropped = gpu_graph(s)
    assert dropped == 0
    for p in range(s.P):
        a, b = s.vertex_range(p)
        rec = s.get_records(p)["intersections"]
        assert len(rec) > 0
        oadj, od = M.contact_graph(pk[p]["startIndices"], rec)
        loc = adj[a:b].copy()
        loc[loc >= 0] -= a
        np.testing.assert_array_equal(loc, oadj)
    A = {}
    for m in range(s.V):
        for u in adj[m]:
            if u >= 0:
                A[(m, int(u))] = A.get((m, int(u)), 0) + 1
    assert all(A.get((u, m), 0) == c for (m, u), c in A.items())


def test_features_match_oracle():
    s, pk = evaluated_system(P=2)
    adjd, adj, _ = gpu_graph(s)
    ell = W.packing_length_units(pk)
    inScale = np.array([3e-3, 2e-3, 2.0, 5e-5], np.float32)
    gf = W.GraphFeatures(s, ell, inScale)
    gf.compute()
    feat = gf.feat.get().reshape(s.V, 4)
    F = s.get_forces().reshape(-1, 2)
    E = s.get_energy()
    for p in range(s.P):
        a, b = s.vertex_range(p)
        st = pk[p]["startIndices"]
        loc = adj[a:b].copy()
        # forces: same float64 arithmetic, asinh of the device library vs the host's
        # (one float32 ulp); crossing counts exact
        o = M.features(st, F[a:b].reshape(-1), np.zeros(len(st) - 1), loc, ell[p], 4, inScale)
        np.testing.assert_allclose(feat[a:b, :2], o[:, :2], rtol=2.4e-7, atol=0)
        np.testing.assert_array_equal(feat[a:b, 2], o[:, 2])
        # polygon energies are internal: their sum over the packing is the packing energy
        n = np.diff(st)
        epoly = np.sinh(feat[a:b, 3].astype(np.float64))[st[:-1]] * n * ell[p] ** 4 * float(np.float32(inScale[3]))
        assert abs(epoly.sum() - E[p]) <= 1e-5 * abs(E[p]) + 1e-30


def run_forward(s, params, feat_np, adjd):
    V = s.V
    net = W.Network(s.ctx, params, V)
    feat = capi.DeviceArray.from_numpy(feat_np.astype(np.float32))
    out = net.forward(V, feat.ptr, adjd.ptr)
    net.sync()
    y = out.get()[:2 * V].reshape(V, 2)
    w = net.work.get()
    S = V * H
    Hs = [w[l * S:(l + 1) * S].reshape(V, H) for l in range(4)]
    Zs = [w[(4 + l) * S:(5 + l) * S].reshape(V, H) for l in range(4)]
    Xs = [None] + [w[(8 + l) * S:(9 + l) * S].view(np.uint16).reshape(V, 2 * H) for l in range(3)]
    Xs = [None] + [(x.astype(np.uint32) << 16).view(np.float32) for x in Xs[1:]]
    return net, feat, y, dict(H=Hs, Z=Zs, X=Xs)


def mfma_close(gpu, X, Wt, b=None):
    """|gpu - X Wt (+ b)| <= 1e-5 * (|X| |Wt| (+ |b|)) + 1e-30 (float64 reference)"""
    ref = X.astype(np.float64) @ Wt.astype(np.float64)
    mag = np.abs(X.astype(np.float64)) @ np.abs(Wt.astype(np.float64))
    if b is not None:
        ref = ref + b
        mag = mag + np.abs(b)
    err = np.abs(gpu.astype(np.float64) - ref)
    assert (err <= 1e-5 * mag + 1e-30).all(), float((err / (mag + 1e-30)).max())


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_layers_match_oracle(seed):
    s, pk = evaluated_system(P=2, seed=10 * seed)
    adjd, adj, _ = gpu_graph(s)
    rng = np.random.default_rng(seed)
    feat = rng.normal(0, 1, (s.V, 4)).astype(np.float32)
    params = M.init_params(seed)
    params[M.OFF_B0:M.OFF_B0 + H] = rng.normal(0, 0.1, H)
    print(f"[mpnn] forward seed={seed} V={s.V}", flush=True)
    net, featd, y, c = run_forward(s, params, feat, adjd)
    W0, b0, Ws, bs, Wo, bo = M.unpack(params)
    Z0, H0 = M.embed(params, feat)
    np.testing.assert_array_equal(c["Z"][0], Z0)
    np.testing.assert_array_equal(c["H"][0], H0)
    for l in range(3):
        Hin = c["H"][l]
        Xo = M.bf(np.concatenate([Hin, M.neighbour_sum(Hin, adj)], 1))
        np.testing.assert_array_equal(c["X"][l + 1], Xo)
        mfma_close(c["Z"][l + 1], Xo, M.bf(Ws[l]), bs[l].astype(np.float64))
        np.testing.assert_array_equal(c["H"][l + 1], (Hin + np.maximum(c["Z"][l + 1], np.float32(0))).astype(np.float32))
    np.testing.assert_array_equal(y, M.output(params, c["H"][3]))
    yo, _ = M.forward(params, feat, adj)
    assert np.linalg.norm(y - yo) <= 1e-2 * np.linalg.norm(yo)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_backward_matches_oracle(seed):
    # seed 2: 32768 vertices -- 128 chunks of the VALU gradients in two groups
    s, pk = evaluated_system(P=2, seed=20 + seed) if seed < 2 else evaluated_system(P=2, N=1024, n=16, seed=22)
    adjd, adj, _ = gpu_graph(s)
    V = s.V
    rng = np.random.default_rng(100 + seed)
    feat = rng.normal(0, 1, (V, 4)).astype(np.float32)
    params = M.init_params(seed)
    params[M.OFF_B0:M.OFF_B0 + H] = rng.normal(0, 0.1, H)
    print(f"[mpnn] backward seed={seed} V={V}", flush=True)
    net = W.Network(s.ctx, params, V, train=True)
    featd = capi.DeviceArray.from_numpy(feat)
    target = rng.normal(0, 1, (V, 2)).astype(np.float32)
    tg = capi.DeviceArray.from_numpy(target)
    net.forward(V, featd.ptr, adjd.ptr)
    loss = ctypes.c_double(0)
    s.ctx.call("pcp_mpnn_loss", V, net.out.ptr, tg.ptr, net.dout.ptr, net.gwork.ptr, ctypes.byref(loss))
    s.ctx.call("pcp_mpnn_backward", V, net.params.ptr, featd.ptr, adjd.ptr, net.work.ptr, net.dout.ptr,
               net.gwork.ptr, net.grad.ptr)
    net.sync()
    g = net.grad.get()
    y = net.out.get()[:2 * V].reshape(V, 2)
    dY = net.dout.get()[:2 * V].reshape(V, 2)
    Lo, dYo = M.loss(y, target)
    assert abs(loss.value - Lo) <= 1e-12 * Lo
    np.testing.assert_array_equal(dY, dYo)
    w = net.work.get()
    S = V * H
    cache = dict(H=[w[l * S:(l + 1) * S].reshape(V, H) for l in range(4)],
                 Z=[w[(4 + l) * S:(5 + l) * S].reshape(V, H) for l in range(4)],
                 X=[None] + [(w[(8 + l) * S:(9 + l) * S].view(np.uint16).reshape(V, 2 * H).astype(np.uint32) << 16)
                             .view(np.float32) for l in range(3)])
    go = M.backward(params, feat, adj, cache, dY)
    # output layer: inputs identical on both sides, same order -> bit-exact
    np.testing.assert_array_equal(g[M.OFF_WO:M.OFF_WO + M.OUT_NP], go[M.OFF_WO:M.OFF_WO + M.OUT_NP])
    blocks = [(M.OFF_W0, M.EMBED_NP)] + [(M.OFF_W1 + l * M.LAYER_NP, M.LAYER_NP) for l in range(3)]
    for o, n in blocks:
        a, b = g[o:o + n].astype(np.float64), go[o:o + n].astype(np.float64)
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), (o, np.linalg.norm(a - b) / np.linalg.norm(b))
    # the top message layer's weight gradient: MFMA over the vertices from the GPU's
    # stored X and the (identical) dZ of that layer
    Wo_bf = M.bf(M.unpack(params)[4])
    dH3 = (dY[:, 0:1] * Wo_bf[:, 0] + dY[:, 1:2] * Wo_bf[:, 1]).astype(np.float32)
    dZ3 = np.where(cache["Z"][3] > 0, dH3, np.float32(0)).astype(np.float32)
    o = M.OFF_W1 + 2 * M.LAYER_NP
    mfma_close(g[o:o + 2 * H * H].reshape(2 * H, H), cache["X"][3].T, M.bf(dZ3))
    np.testing.assert_array_equal(g[o + 2 * H * H:o + M.LAYER_NP], M._bias_grad_gpu_order(dZ3))


def test_adam_matches_oracle():
    s, _ = evaluated_system(P=1)
    rng = np.random.default_rng(5)
    p = M.init_params(3)
    net = W.Network(s.ctx, p, 64, train=True)
    pm, mm, vm = p.copy(), np.zeros_like(p), np.zeros_like(p)
    for t in range(1, 4):
        g = rng.normal(0, 1e-2, M.NPARAM).astype(np.float32)
        net.grad.upload(g)
        s.ctx.call("pcp_mpnn_adam", net.params.ptr, net.grad.ptr, net.m.ptr, net.v.ptr, t, ctypes.c_float(1e-3),
                   ctypes.c_float(0.9), ctypes.c_float(0.999), ctypes.c_float(1e-8))
        pm, mm, vm = M.adam(pm, g, mm, vm, t, 1e-3)
    net.sync()
    np.testing.assert_allclose(net.params.get(), pm, rtol=2e-6, atol=1e-9)
    np.testing.assert_allclose(net.m.get(), mm, rtol=1e-6, atol=1e-12)


def test_training_reduces_loss_on_fire_trajectory_data():
    """a small data set from the method's own FIRE trajectories: the GPU-trained network
    fits the displacement to the minimized state better than predicting zero"""
    ds, runs, inScale = W.generate(1, 4, seed=11, capture_steps=(0, 50), maxSteps=6000, N=128, n=12)
    s = runs[0][0]
    outScale = W.label_all(ds, runs)
    assert np.isfinite(outScale) and outScale > 0
    params0 = W.init_params(0)
    net0 = W.Network(s.ctx, params0, s.V, train=True)
    d0 = ds.samples[0]
    L0 = net0.loss(d0["V"], d0["feat"].ptr, d0["adj"].ptr, d0["target"].ptr)
    tz = d0["target"].get().astype(np.float64)
    Lz = float((tz ** 2).sum() / d0["V"])   # loss of predicting no displacement
    params, losses, _ = W.train(ds.samples, params0, epochs=60, lr=2e-3, log=lambda *_: None)
    net1 = W.Network(s.ctx, params, s.V, train=True)
    L1 = net1.loss(d0["V"], d0["feat"].ptr, d0["adj"].ptr, d0["target"].ptr)
    print(f"[mpnn] loss {L0:.4f} -> {L1:.4f} (zero prediction: {Lz:.4f})", flush=True)
    assert np.isfinite(losses).all()
    assert L1 < L0 and L1 < 0.95 * Lz
    # training is reproducible bit for bit
    params2, _, _ = W.train(ds.samples, params0, epochs=3, lr=2e-3, log=lambda *_: None)
    params3, _, _ = W.train(ds.samples, params0, epochs=3, lr=2e-3, log=lambda *_: None)
    np.testing.assert_array_equal(params2, params3)


def test_training_data_capture_is_reproducible():
    """captured states (features, graph, positions) and targets are identical bit for bit
    when the data generation is repeated: the captures are ordered with the FIRE steps"""
    def run():
        ds, runs, _ = W.generate(1, 8, 5, capture_steps=(0, 10, 25), maxSteps=3000, N=256, n=16, dtMax=W.DT_MAX,
                                 log=lambda *_: None)
        W.label_all(ds, runs)
        return [(d["feat"].get(), d["adj"].get(), d["x"].get(), d["target"].get()) for d in ds.samples]
    a, b = run(), run()
    assert len(a) == 3
    for sa, sb in zip(a, b):
        for x, y in zip(sa, sb):
            np.testing.assert_array_equal(x, y)


def test_trained_warm_start_gain_over_plain_fire():
    """The committed network (trained on the GPU by these kernels on 960 FIRE trajectories
    of other seeds) against plain FIRE on 32 held-out packings of N=1024 x 16: both reach
    max|F| <= 1e-10 (FP64 FIRE judges the residual force) and the warm start needs fewer
    FIRE steps on average.  Deterministic end to end, so the numbers are exact for this
    seed: plain 297.8 mean steps, warm start 270.0 (one application)."""
    pk = W.make_c4_packings(32, 777777)
    res = W.ab_compare(W.load_weights(W.DEFAULT_WEIGHTS), pk, maxSteps=3000, dtMax=W.DT_MAX,
                       log=lambda m: print(m, flush=True))
    plain, warm = np.array(res["plain_steps"]), np.array(res["warm_steps"])
    assert (plain >= 0).all() and (warm >= 0).all()
    assert max(res["plain_maxF"]) <= 1e-10 and max(res["warm_maxF"]) <= 1e-10
    print(f"[mpnn] mean FIRE steps: plain {plain.mean():.1f}, warm start {warm.mean():.1f} "
          f"({1 - warm.mean() / plain.mean():+.1%} fewer)", flush=True)
    assert warm.mean() < 0.95 * plain.mean()


def _circles(x, start):
    """bounding circles as k_bounds builds them: chain of wrapped edge vectors from vertex 0,
    centre = p0 + mean of the chain, radius = largest distance of a chain point to it"""
    w = lambda d: (d + 0.5) % 1.0 - 0.5
    X = np.asarray(x).reshape(-1, 2)
    C, R = [], []
    for s in range(len(start) - 1):
        v = X[start[s]:start[s + 1]]
        u = np.vstack([np.zeros(2), np.cumsum(w(np.diff(v, axis=0)), axis=0)])
        m = u.mean(axis=0)
        C.append(v[0] + m)
        R.append(np.sqrt(((u - m) ** 2).sum(axis=1)).max())
    return np.array(C), np.array(R)


def _contained_pairs(x, start, kappa):
    """polygon pairs (a < b) whose smaller bounding circle lies inside the larger one:
    |c_a - c_b| + r_small < kappa r_large, minimum image (brute force over all pairs)"""
    C, R = _circles(x, start)
    d = C[:, None, :] - C[None, :, :]
    d = (d + 0.5) % 1.0 - 0.5
    D = np.sqrt((d ** 2).sum(axis=2))
    lim = kappa * np.maximum(R[:, None], R[None, :]) - np.minimum(R[:, None], R[None, :])
    a, b = np.nonzero(np.triu(D < lim, 1))
    return list(zip(a.tolist(), b.tolist())), (D + np.minimum(R[:, None], R[None, :])) / np.maximum(R[:, None], R[None, :])


def _guard_cpu(x, xprev, start, kappa):
    """the guard's rule on the CPU: the polygons of contained pairs take back xprev; if the
    packing still holds a contained pair, all of it does.  Returns (flag, positions)."""
    x = x.copy().reshape(-1, 2)
    pairs, _ = _contained_pairs(x, start, kappa)
    if not pairs:
        return 0, x.reshape(-1)
    for s in {s for pr in pairs for s in pr}:
        x[start[s]:start[s + 1]] = xprev.reshape(-1, 2)[start[s]:start[s + 1]]
    if _contained_pairs(x, start, kappa)[0]:
        return 2, xprev.copy()
    return 1, x.reshape(-1)


def test_warm_start_guard_takes_back_contained_polygons():
    """pcp_mpnn_guard against its rule restated on the CPU (brute force over polygon pairs):
    where one polygon's bounding circle lies inside another's (|c_A - c_B| + r_small <
    kappa r_large) both take back the earlier positions; a packing that still holds such a
    pair then takes back all of them"""
    pk = [make_packing(64, 12, 0.80, seed=s, min_center_frac=0.5) for s in range(5)]
    x_in = [p["positions"].copy() for p in pk]
    moved = [q.copy() for q in x_in]

    def carry(p, a, target):
        st = pk[p]["startIndices"]
        C, _ = _circles(moved[p], st)
        v = moved[p].reshape(-1, 2)
        v[st[a]:st[a + 1]] = (v[st[a]:st[a + 1]] + (target - C[a])) % 1.0

    # packings 1-3: the smallest polygon carried towards the largest, its centre at 0.0, 0.6
    # and 1.6 x (r_large - r_small) from the large one's
    for p, f in ((1, 0.0), (2, 0.6), (3, 1.6)):
        C, R = _circles(moved[p], pk[p]["startIndices"])
        a, b = int(np.argmin(R)), int(np.argmax(R))
        d = (C[a] - C[b] + 0.5) % 1.0 - 0.5
        carry(p, a, C[b] + f * (R[b] - R[a]) * d / np.linalg.norm(d))
    # packing 4: the smallest polygon onto the largest one's centre, and another large polygon
    # onto the smallest one's old centre (taking back the first pair makes the second)
    C, R = _circles(moved[4], pk[4]["startIndices"])
    order = np.argsort(R)
    a, b, c = int(order[0]), int(order[-1]), int(order[-2])
    carry(4, c, C[a].copy())
    carry(4, a, C[b].copy())
    expect = [_guard_cpu(moved[p], x_in[p], pk[p]["startIndices"], W.GUARD_KAPPA) for p in range(5)]
    assert [e[0] for e in expect] == [0, 1, 1, 0, 2]
    for p in range(5):   # every pair clear of the limit, so that rounding cannot decide
        q = _contained_pairs(moved[p], pk[p]["startIndices"], W.GUARD_KAPPA)[1]
        assert np.abs(q[np.triu_indices(len(q), 1)] - W.GUARD_KAPPA).min() > 0.05
    sysq = [dict(p, positions=m) for p, m in zip(pk, moved)]
    s = capi.System(sysq, [box_size_for(p["diameter"]) for p in pk], model=W.SQUARED)
    prev = capi.DeviceArray.from_numpy(np.concatenate(x_in))
    flags = capi.DeviceArray((5,), np.int32)
    n = ctypes.c_int(-1)
    print("[mpnn] guard, 5 packings of N=64 x 12", flush=True)
    s.ctx.call("pcp_mpnn_guard", s.pos.ptr, prev.ptr, W.GUARD_KAPPA, flags.ptr, ctypes.byref(n))
    np.testing.assert_array_equal(flags.get(), [e[0] for e in expect])
    assert n.value == 3
    for p in range(5):
        np.testing.assert_array_equal(s.get_positions(p), expect[p][1])
    # a second call on the result keeps every packing
    s.ctx.call("pcp_mpnn_guard", s.pos.ptr, prev.ptr, W.GUARD_KAPPA, flags.ptr, ctypes.byref(n))
    assert n.value == 0 and not flags.get().any()
    with pytest.raises(RuntimeError):
        s.ctx.call("pcp_mpnn_guard", s.pos.ptr, prev.ptr, 2.5, flags.ptr, None)
