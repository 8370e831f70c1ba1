"""CPU oracle of the learned warm start (test infrastructure only).

The reference (camdennis/pyCudaPolygon) has no learned component; this restates the
network of include/pcp.h (PCP_MPNN_*) and its gradients in numpy, so that the HIP
kernels (pycudapolygon_amd/csrc/pcp_mpnn.hip) can be checked against it:

  h0  = relu(bf(x) bf(W0) + b0)
  h_l = h_{l-1} + relu(bf([h_{l-1}, A h_{l-1}]) bf(W_l) + b_l),   l = 1, 2, 3
  y   = bf(h3) bf(Wo) + bo

A is the symmetric adjacency (fixed-width slots, -1 empty), bf() rounds float32 to
bfloat16 (nearest even).  Where the GPU sums in float32 in a fixed order (neighbour
sums, the 4x64 and 64x2 layers, bias and small-weight gradients) the oracle repeats
that order in float32 and agrees bit for bit; the matrix-core products (K = 128, 64
and the vertex sums of the weight gradients) are evaluated in float64 from the same
bf16 operands, so they agree to float32 accumulation error.

`exact=True` drops every rounding (pure float64); the gradient of that network is
pinned by central finite differences in tests/test_mpnn_oracle.py.

Nothing outside tests/, __graft_entry__.smoke() and bench.py's CPU baseline uses this.
"""
import numpy as np

H = 64
DEG = 16
OFF_W0, OFF_B0, EMBED_NP = 0, 256, 320
OFF_W1, LAYER_NP = 320, 8256
OFF_WO, OFF_BO, OUT_NP = 25088, 25216, 130
NPARAM = 25218
CH = 2048   # vertices per partial-sum chunk of the GPU's matrix-core weight gradients
CHV = 256   # vertices per partial-sum chunk of its VALU weight gradients
GRP = 64    # chunks per group when the chunk partials are summed


def bf(a):
    """float32 -> nearest-even bfloat16, returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(params):
    """views of the parameter vector (float64 stays float64, anything else float32)"""
    p = np.asarray(params)
    if p.dtype != np.float64:
        p = p.astype(np.float32)
    W0 = p[OFF_W0:OFF_W0 + 4 * H].reshape(4, H)
    b0 = p[OFF_B0:OFF_B0 + H]
    Ws, bs = [], []
    for l in range(3):
        o = OFF_W1 + l * LAYER_NP
        Ws.append(p[o:o + 2 * H * H].reshape(2 * H, H))
        bs.append(p[o + 2 * H * H:o + LAYER_NP])
    Wo = p[OFF_WO:OFF_WO + 2 * H].reshape(H, 2)
    bo = p[OFF_BO:OFF_BO + 2]
    return W0, b0, Ws, bs, Wo, bo


def init_params(seed, scale=1.0):
    """He-style initial parameters (float32), biases zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    p[OFF_W0:OFF_W0 + 4 * H] = rng.normal(0, scale * np.sqrt(2.0 / 4), 4 * H)
    for l in range(3):
        o = OFF_W1 + l * LAYER_NP
        p[o:o + 2 * H * H] = rng.normal(0, scale * 0.5 * np.sqrt(2.0 / (2 * H)), 2 * H * H)
    p[OFF_WO:OFF_WO + 2 * H] = rng.normal(0, scale * np.sqrt(1.0 / H), 2 * H)
    return p


# ---- graph and features -------------------------------------------------------------------
def contact_graph(startIndices, records):
    """Adjacency of one packing from its crossing records (packing-local ids).

    records: uint64 keys [sj|si|i|j] (kernels.cuh:860-874 layout), i and j local to their
    polygons.  Row m: prev, next, crossing partners ascending, -1 padding (as
    k_mpnn_graph).  Returns (adj int32[V, DEG], dropped)."""
    st = np.asarray(startIndices, np.int64)
    V = int(st[-1])
    N = len(st) - 1
    parts = [[] for _ in range(V)]
    for k in np.asarray(records, np.uint64):
        k = int(k)
        sj, si, i, j = (k >> 48) & 0xFFFF, (k >> 32) & 0xFFFF, (k >> 16) & 0xFFFF, k & 0xFFFF
        a, b = int(st[si]) + i, int(st[sj]) + j
        parts[a].append(b)
        parts[b].append(a)
    adj = -np.ones((V, DEG), np.int32)
    dropped = 0
    for s in range(N):
        v0, n = int(st[s]), int(st[s + 1] - st[s])
        for a in range(n):
            m = v0 + a
            adj[m, 0] = v0 + (a + n - 1) % n
            adj[m, 1] = v0 + (a + 1) % n
            p = sorted(parts[m])
            dropped += max(0, len(p) - (DEG - 2))
            p = p[:DEG - 2]
            adj[m, 2:2 + len(p)] = p
    return adj, dropped


def features(startIndices, force, epoly, adj, ell, energyPow, inScale):
    """k_mpnn_features for one packing (float64 arithmetic, float32 result)."""
    st = np.asarray(startIndices, np.int64)
    n = np.diff(st)
    shape = np.repeat(np.arange(len(n)), n)
    F = np.asarray(force, np.float64).reshape(-1, 2)
    Lf = 1.0
    for _ in range(energyPow - 1):   # repeated products, as the kernel
        Lf = Lf * ell
    Le = Lf * ell
    nc = (adj[:, 2:] >= 0).sum(1)
    out = np.empty((len(F), 4), np.float32)
    out[:, 0] = np.arcsinh(F[:, 0] / (Lf * float(np.float32(inScale[0]))))
    out[:, 1] = np.arcsinh(F[:, 1] / (Lf * float(np.float32(inScale[1]))))
    out[:, 2] = nc.astype(np.float32) / np.float32(inScale[2])
    out[:, 3] = np.arcsinh(np.asarray(epoly, np.float64)[shape] / (n[shape] * Le * float(np.float32(inScale[3]))))
    return out


# ---- forward --------------------------------------------------------------------------
def neighbour_sum(Hf, adj):
    """float32 sums over the slots in order (as k_mpnn_layer)."""
    Hf = np.asarray(Hf, np.float32)
    acc = np.zeros_like(Hf)
    for d in range(adj.shape[1]):
        u = adj[:, d]
        ok = u >= 0
        acc[ok] = acc[ok] + Hf[u[ok]]
    return acc


def embed(params, feat, exact=False):
    W0, b0, _, _, _, _ = unpack(params)
    if exact:
        Z = np.asarray(feat, np.float64) @ W0.astype(np.float64) + b0
        return Z, np.maximum(Z, 0.0)
    x, w = bf(feat), bf(W0)
    z = x[:, 0:1] * w[0]
    for f in range(1, 4):
        z = z + x[:, f:f + 1] * w[f]
    z = z + b0
    return z, np.maximum(z, np.float32(0))


def layer(W, b, Hin, adj, exact=False):
    """One message layer: returns X (the product's input), Z, Hout."""
    if exact:
        A = np.zeros_like(Hin)
        for d in range(adj.shape[1]):
            u = adj[:, d]
            ok = u >= 0
            A[ok] += Hin[u[ok]]
        X = np.concatenate([Hin, A], 1)
        Z = X @ W.astype(np.float64) + b
        return X, Z, Hin + np.maximum(Z, 0.0)
    X = bf(np.concatenate([Hin, neighbour_sum(Hin, adj)], 1))
    Z = (X.astype(np.float64) @ bf(W).astype(np.float64) + b.astype(np.float64)).astype(np.float32)
    return X, Z, (Hin + np.maximum(Z, np.float32(0))).astype(np.float32)


def output(params, H3, exact=False):
    _, _, _, _, Wo, bo = unpack(params)
    if exact:
        return H3 @ Wo.astype(np.float64) + bo
    h, w = bf(H3), bf(Wo)
    y = np.zeros((len(H3), 2), np.float32)
    for c in range(H):
        y = y + h[:, c:c + 1] * w[c]
    return y + bo


def forward(params, feat, adj, exact=False):
    """Returns (y, cache) with cache = dict(H=[H0..H3], Z=[Z0..Z3], X=[None, X1..X3])."""
    _, _, Ws, bs, _, _ = unpack(params)
    Z0, H0 = embed(params, feat, exact)
    Hs, Zs, Xs = [H0], [Z0], [None]
    for l in range(3):
        X, Z, Hn = layer(Ws[l], bs[l], Hs[-1], adj, exact)
        Hs.append(Hn); Zs.append(Z); Xs.append(X)
    return output(params, Hs[3], exact), dict(H=Hs, Z=Zs, X=Xs)


def loss(y, target):
    """k_mpnn_loss: float32 differences, squares summed in float64, dY = (2/V) d in float32"""
    d = (np.asarray(y, np.float32) - np.asarray(target, np.float32)).astype(np.float32)
    V = len(d)
    return float((d.astype(np.float64) ** 2).sum() / V), (np.float32(2.0) / np.float32(V)) * d


# ---- backward -------------------------------------------------------------------------
def _seqsum(a):
    """float32 sequential sum over axis 0 (numpy's cumsum accumulates in order)."""
    a = np.asarray(a, np.float32)
    if len(a) == 0:
        return np.zeros(a.shape[1:], np.float32)
    return np.cumsum(a, axis=0, dtype=np.float32)[-1]


def _chunked(fn, V, ch=CH):
    """GPU order of vertex sums: sequential inside chunks of `ch` vertices, then over groups of
    GRP consecutive chunks, then over the groups."""
    parts = np.stack([fn(m0, min(V, m0 + ch)) for m0 in range(0, V, ch)])
    groups = np.stack([_seqsum(parts[g:g + GRP]) for g in range(0, len(parts), GRP)])
    return _seqsum(groups)


def _bias_grad_gpu_order(dZ):
    """k_mpnn_layer_grad's bias sum: per chunk, four row groups ((row % 16) / 4 = q)
    summed block by block, rows in order, combined ((g0 + g1) + g2) + g3."""
    V = len(dZ)

    def chunk(m0, m1):
        rows = np.arange(m0, m1)
        rel = rows - m0
        gs = []
        for q in range(4):
            sel = rows[((rel % 16) // 4) == q]
            gs.append(_seqsum(dZ[sel]) if len(sel) else np.zeros(H, np.float32))
        return ((gs[0] + gs[1]) + gs[2]) + gs[3]

    return _chunked(chunk, V)


def backward(params, feat, adj, cache, dY, exact=False):
    """Gradient of sum(dY * y) with respect to the parameter vector (flat, float32 or
    float64 when exact).  Straight-through rounding: d bf(x)/dx = 1."""
    W0, b0, Ws, bs, Wo, bo = unpack(params)
    V = len(feat)
    Hs, Zs, Xs = cache["H"], cache["Z"], cache["X"]
    if exact:
        dY = np.asarray(dY, np.float64)
        g = np.zeros(NPARAM)
        g[OFF_WO:OFF_WO + 2 * H] = (Hs[3].T @ dY).reshape(-1)
        g[OFF_BO:OFF_BO + 2] = dY.sum(0)
        dH = dY @ Wo.T.astype(np.float64)
        for l in (2, 1, 0):
            dZ = dH * (Zs[l + 1] > 0)
            o = OFF_W1 + l * LAYER_NP
            g[o:o + 2 * H * H] = (Xs[l + 1].T @ dZ).reshape(-1)
            g[o + 2 * H * H:o + LAYER_NP] = dZ.sum(0)
            dX = dZ @ Ws[l].T.astype(np.float64)
            dA = dX[:, H:]
            dHn = dH + dX[:, :H]
            for d in range(adj.shape[1]):
                u = adj[:, d]
                ok = u >= 0
                dHn[ok] += dA[u[ok]]
            dH = dHn
        dZ0 = dH * (Zs[0] > 0)
        g[OFF_W0:OFF_W0 + 4 * H] = (np.asarray(feat, np.float64).T @ dZ0).reshape(-1)
        g[OFF_B0:OFF_B0 + H] = dZ0.sum(0)
        return g
    dY = np.asarray(dY, np.float32)
    g = np.zeros(NPARAM, np.float32)
    h3 = bf(Hs[3])
    gwo = _chunked(lambda a, b: np.concatenate([
        np.stack([_seqsum(h3[a:b, c] * dY[a:b, o]) for c in range(H) for o in range(2)]),
        _seqsum(dY[a:b])]), V, CHV)
    g[OFF_WO:OFF_WO + OUT_NP] = gwo
    wo = bf(Wo)
    dH = (dY[:, 0:1] * wo[:, 0] + dY[:, 1:2] * wo[:, 1]).astype(np.float32)
    for l in (2, 1, 0):
        dZ = np.where(Zs[l + 1] > 0, dH, np.float32(0)).astype(np.float32)
        o = OFF_W1 + l * LAYER_NP
        g[o:o + 2 * H * H] = (Xs[l + 1].astype(np.float64).T @ bf(dZ).astype(np.float64)).reshape(-1)
        g[o + 2 * H * H:o + LAYER_NP] = _bias_grad_gpu_order(dZ)
        dX = (bf(dZ).astype(np.float64) @ bf(Ws[l]).T.astype(np.float64)).astype(np.float32)
        dHself = (dH + dX[:, :H]).astype(np.float32)
        dA = dX[:, H:]
        acc = np.zeros_like(dHself)
        for d in range(adj.shape[1]):
            u = adj[:, d]
            ok = u >= 0
            acc[ok] = acc[ok] + dA[u[ok]]
        dH = (dHself + acc).astype(np.float32)
    dZ0 = np.where(Zs[0] > 0, dH, np.float32(0)).astype(np.float32)
    x = bf(feat)
    gw0 = _chunked(lambda a, b: np.concatenate([
        np.stack([_seqsum(x[a:b, f] * dZ0[a:b, c]) for f in range(4) for c in range(H)]),
        _seqsum(dZ0[a:b])]), V, CHV)
    g[OFF_W0:OFF_W0 + EMBED_NP] = gw0
    return g


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step in float32 (k_mpnn_adam)."""
    f = np.float32
    g = np.asarray(grad, f)
    m = f(b1) * m + (f(1) - f(b1)) * g
    v = f(b2) * v + (f(1) - f(b2)) * (g * g)
    c1 = f(1) - f(b1) ** f(step)
    c2 = f(1) - f(b2) ** f(step)
    p = params - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return p.astype(f), m.astype(f), v.astype(f)
