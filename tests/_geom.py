"""Independent plane geometry used to pin the oracle (no code shared with it):
convex polygon clipping (Sutherland-Hodgman), shoelace areas, segment crossings."""
import numpy as np


def shoelace(poly):
    x, y = poly[:, 0], poly[:, 1]
    return 0.5 * float(np.dot(x, np.roll(y, -1)) - np.dot(np.roll(x, -1), y))


def clip_convex(subject, clipper):
    """Intersection of two convex CCW polygons (arrays (n,2)), Sutherland-Hodgman."""
    out = [tuple(p) for p in subject]
    m = len(clipper)
    for k in range(m):
        a, b = clipper[k], clipper[(k + 1) % m]
        inp, out = out, []
        if not inp:
            break

        def inside(p):
            return (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0]) >= 0

        def cross_pt(p, q):
            x1, y1, x2, y2 = p[0], p[1], q[0], q[1]
            x3, y3, x4, y4 = a[0], a[1], b[0], b[1]
            den = (x1 - x2) * (y3 - y4) - (y1 - y2) * (x3 - x4)
            t = ((x1 - x3) * (y3 - y4) - (y1 - y3) * (x3 - x4)) / den
            return (x1 + t * (x2 - x1), y1 + t * (y2 - y1))

        for i in range(len(inp)):
            cur, prv = inp[i], inp[i - 1]
            if inside(cur):
                if not inside(prv):
                    out.append(cross_pt(prv, cur))
                out.append(cur)
            elif inside(prv):
                out.append(cross_pt(prv, cur))
    return np.array(out) if out else np.zeros((0, 2))


def overlap_area(p1, p2):
    c = clip_convex(p1, p2)
    return shoelace(c) if len(c) >= 3 else 0.0


def crossings(p1, p2):
    """All proper crossings between the boundaries of two polygons (list of points)."""
    pts = []
    n1, n2 = len(p1), len(p2)
    for i in range(n1):
        a, b = p1[i], p1[(i + 1) % n1]
        for j in range(n2):
            c, d = p2[j], p2[(j + 1) % n2]
            r, s = b - a, d - c
            den = r[0] * s[1] - r[1] * s[0]
            if abs(den) < 1e-300:
                continue
            g = c - a
            t = (g[0] * s[1] - g[1] * s[0]) / den
            u = (g[0] * r[1] - g[1] * r[0]) / den
            if 0 < t < 1 and 0 < u < 1:
                pts.append(a + t * r)
    return pts


def regular(n, R, center, theta=0.0):
    ang = theta + 2 * np.pi * np.arange(n) / n
    return np.stack([center[0] + R * np.cos(ang), center[1] + R * np.sin(ang)], axis=1)


def to_model(polys):
    """(startIndices, positions AoS) for a list of polygons already inside [0,1)^2."""
    n = [len(p) for p in polys]
    start = np.concatenate([[0], np.cumsum(n)]).astype(np.int32)
    pos = np.concatenate([p.reshape(-1) for p in polys]) % 1.0
    return start, pos
