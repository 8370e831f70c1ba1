"""Seeded synthetic packings in the reference's unit-box convention.

The reference keeps every packing in the periodic unit box [0,1)^2 and sets the
packing fraction by scaling polygons (pyCudaPolygon.py:146-163, 223-234).  This
generator builds polygons in physical units -- a regular n-gon of circumradius
R in {1.0, 1.4} (exact 50:50, shuffled), orientation uniform in [0, 2pi),
per-vertex radial noise uniform in +-noise, centres uniform -- then divides by
L = sqrt(sum(A_i)/phi) and wraps into [0,1).  Vertices are counter-clockwise.

Targets follow the reference's per-polygon semantics (targetAreas[N],
targetEdgeLengths[N] = perimeter/n, src/model.cpp:630-636) plus a per-edge
target array (the exact initial edge lengths) for constraint solvers that keep
noisy polygons' shapes.
"""
import numpy as np


def polygon_areas(verts):
    x, y = verts[..., 0], verts[..., 1]
    return 0.5 * (x * np.roll(y, -1, axis=-1) - np.roll(x, -1, axis=-1) * y).sum(axis=-1)


def separated_centers(rng, R, L, frac, tries=10000):
    """Centres uniform in the unit box, placed one by one and re-drawn while closer than
    frac * (R_i + R_j) / L (minimum image) to a centre already placed.  Placed centres are
    kept in a grid of cells at least as wide as the largest exclusion distance, so each
    candidate is checked against the 3x3 cells around it only."""
    N = len(R)
    out = np.empty((N, 2))
    rmax = frac * 2.0 * float(np.max(R)) / L
    G = max(1, min(int(1.0 / rmax) if rmax > 0 else 1, 2048))
    cells = {}
    for i in range(N):
        for _ in range(tries):
            c = rng.uniform(0.0, 1.0, 2)
            cx, cy = int(c[0] * G) % G, int(c[1] * G) % G
            near = []
            for dx in (-1, 0, 1) if G > 2 else (0,):
                for dy in (-1, 0, 1) if G > 2 else (0,):
                    near.extend(cells.get(((cx + dx) % G, (cy + dy) % G), ()))
            if G <= 2:
                near = range(i)
            if near:
                idx = np.fromiter(near, int)
                d = out[idx] - c
                d -= np.round(d)
                if not np.all((d ** 2).sum(1) >= (frac * (R[idx] + R[i]) / L) ** 2):
                    continue
            break
        else:
            raise RuntimeError("separated_centers: no room; lower frac or phi")
        out[i] = c
        cells.setdefault((int(c[0] * G) % G, int(c[1] * G) % G), []).append(i)
    return out


def make_packing(N, n, phi, seed, radii=(1.0, 1.4), noise=0.05, min_center_frac=None):
    """One packing of N polygons with n vertices each.

    min_center_frac: None = centres uniform; f = centres uniform but no two closer than
    f * (R_i + R_j) (f = 0.5 keeps every polygon's centre outside the others, so no
    polygon starts inside another -- a state the edge-crossing overlap cannot see).

    Returns dict(nArray, startIndices, positions (AoS flat, 2V), targetAreas[N],
    targetEdgeLengths[N], targetEdgeV[V], boxLength, maxEdgeLength, diameter).
    """
    rng = np.random.default_rng(seed)
    R = np.where(np.arange(N) < N // 2, radii[0], radii[1]).astype(np.float64)
    rng.shuffle(R)
    theta0 = rng.uniform(0.0, 2.0 * np.pi, N)
    centers = rng.uniform(0.0, 1.0, (N, 2))
    eps = rng.uniform(-noise, noise, (N, n))
    ang = theta0[:, None] + 2.0 * np.pi * np.arange(n)[None, :] / n
    r = R[:, None] * (1.0 + eps)
    verts = np.stack([r * np.cos(ang), r * np.sin(ang)], axis=-1)      # (N, n, 2), CCW
    areas = polygon_areas(verts)
    L = np.sqrt(areas.sum() / phi)
    if min_center_frac is not None:
        centers = separated_centers(rng, R, L, min_center_frac)
    local = verts / L
    pos = (centers[:, None, :] + local) % 1.0
    pos = np.where(pos >= 1.0, 0.0, pos)
    edges = np.roll(local, -1, axis=1) - local
    elen = np.sqrt((edges ** 2).sum(-1))                               # (N, n)
    nArray = np.full(N, n, dtype=np.int64)
    start = np.concatenate([[0], np.cumsum(nArray)]).astype(np.int32)
    diam = 2.0 * r.max(axis=1) / L
    return dict(nArray=nArray, startIndices=start, positions=pos.reshape(-1).copy(),
                targetAreas=areas / L ** 2, targetEdgeLengths=elen.sum(1) / n,
                targetEdgeV=elen.reshape(-1).copy(), boxLength=L,
                maxEdgeLength=float(elen.max()), diameter=float(diam.max()))


def make_batch(P, seed, N_range=(64, 1024), n_range=(8, 32), phi_range=(0.80, 0.90), **kw):
    """P independent packings with N ~ U{N_range}, n ~ U{n_range}, phi ~ U[phi_range]."""
    rng = np.random.default_rng(seed)
    Ns = rng.integers(N_range[0], N_range[1] + 1, P)
    ns = rng.integers(n_range[0], n_range[1] + 1, P)
    phis = rng.uniform(phi_range[0], phi_range[1], P)
    seeds = rng.integers(0, 2 ** 31 - 1, P)
    return [make_packing(int(N), int(n), float(ph), int(sd), **kw)
            for N, n, ph, sd in zip(Ns, ns, phis, seeds)]


def box_size_for(cell_side):
    """Cells per side for a minimum cell side (reference Model::setMaxEdgeLength, model.cpp:513-516)."""
    return max(1, int(np.floor(1.0 / cell_side)))
