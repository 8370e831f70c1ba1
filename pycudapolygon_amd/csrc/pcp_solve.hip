// Per-polygon constraint solves with one lane segment per polygon (part of the
// pcp_kernels.hip translation unit).
//
// Projection (projectForceCUDA model.cu:165-206) and every SHAKE iteration solve the
// normal equations of the edge-length and area constraints of one polygon:
//     [[G, b], [b^T, c]] [y; z] = [r; s],
// G cyclic tridiagonal (G_ee = 2|t_e|^2, G_{e,e+1} = -t_e.t_{e+1}), b_e = t_e.(a_{e+1}-a_e),
// c = |a|^2.  The leading (n-1)x(n-1) block of G is tridiagonal; it is solved for three
// right-hand sides (r, b and the border column of the cyclic corner) by parallel cyclic
// reduction across the lanes of the segment (lane e holds equation e; log2(n) steps of
// shuffles), then the 1x1 Schur complements of the corner and of the area row follow.
// Lane e of a segment owns vertex e and edge e (p_e -> p_{e+1}) of its polygon; segments
// are SEG = 4..64 lanes wide (the smallest power of two >= the largest polygon).

namespace pcp {

template <int SEG>
__device__ __forceinline__ double seg_sum(double v) {
#pragma unroll
    for (int o = SEG / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, SEG);
    return v;
}
template <int SEG>
__device__ __forceinline__ double seg_max(double v) {
#pragma unroll
    for (int o = SEG / 2; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, SEG));
    return v;
}
template <int SEG>
__device__ __forceinline__ double seg_get(double v, int src) {
    return __shfl(v, src, SEG);
}

// Solve for lane e < n: returns y_e (the edge multipliers), *zOut = area multiplier.
// t = unit edge tangent (0 for a degenerate edge), a = area gradient of vertex e,
// r = right-hand side of edge e, c/s = area-row entries (segment-uniform).
// zThresh: the area direction is dropped when its Schur complement is <= zThresh * c.
template <int SEG>
__device__ double seg_bordered_solve(int e, int n, double tx, double ty, double ax, double ay, double r,
                                     double c, double s, double zThresh, double* zOut) {
    int en = (e + 1 >= n) ? 0 : e + 1;   // padding lanes (e >= n) read lane 0
    bool live = e < n;
    double txn = seg_get<SEG>(tx, en), tyn = seg_get<SEG>(ty, en);
    double axn = seg_get<SEG>(ax, en), ayn = seg_get<SEG>(ay, en);
    double off = live ? -(tx * txn + ty * tyn) : 0.0;       // G_{e,e+1 mod n}
    double dsq = tx * tx + ty * ty;
    double diag = (dsq > 0.0) ? 2.0 * dsq : 1.0;
    double bv = live ? tx * (axn - ax) + ty * (ayn - ay) : 0.0;
    int m = n - 1;
    double offPrev = seg_get<SEG>(off, (e == 0) ? 0 : e - 1);
    double offM = seg_get<SEG>(off, m);        // G_{n-1,0}: cyclic corner
    double offM1 = seg_get<SEG>(off, m - 1);   // G_{n-2,n-1}
    double rLast = seg_get<SEG>(r, m), bLast = seg_get<SEG>(bv, m), diagLast = seg_get<SEG>(diag, m);
    // leading block, lanes e < m; other lanes carry x = 0
    bool inb = e < m;
    double A = (inb && e > 0) ? offPrev : 0.0;
    double C = (inb && e < m - 1) ? off : 0.0;
    double Bd = inb ? diag : 1.0;
    double d1 = inb ? r : 0.0, d2 = inb ? bv : 0.0;
    double d3 = 0.0;
    if (inb) {
        if (e == 0) d3 += offM;
        if (e == m - 1) d3 += offM1;
    }
    double w0 = offM + ((m - 1 == 0) ? offM1 : 0.0);
    double wl = (m - 1 == 0) ? 0.0 : offM1;
    for (int st = 1; st < m; st <<= 1) {
        double inv = 1.0 / Bd;
        int lo = e - st, hi = e + st;
        bool hasLo = inb && lo >= 0, hasHi = inb && hi < m;
        int srcLo = hasLo ? lo : e, srcHi = hasHi ? hi : e;
        double aL = seg_get<SEG>(A, srcLo), cL = seg_get<SEG>(C, srcLo), iL = seg_get<SEG>(inv, srcLo);
        double d1L = seg_get<SEG>(d1, srcLo), d2L = seg_get<SEG>(d2, srcLo), d3L = seg_get<SEG>(d3, srcLo);
        double aH = seg_get<SEG>(A, srcHi), cH = seg_get<SEG>(C, srcHi), iH = seg_get<SEG>(inv, srcHi);
        double d1H = seg_get<SEG>(d1, srcHi), d2H = seg_get<SEG>(d2, srcHi), d3H = seg_get<SEG>(d3, srcHi);
        double al = hasLo ? -A * iL : 0.0;
        double ga = hasHi ? -C * iH : 0.0;
        double nA = hasLo ? al * aL : 0.0;
        double nC = hasHi ? ga * cH : 0.0;
        double nB = Bd;
        double n1 = d1, n2 = d2, n3 = d3;
        if (hasLo) { nB += al * cL; n1 += al * d1L; n2 += al * d2L; n3 += al * d3L; }
        if (hasHi) { nB += ga * aH; n1 += ga * d1H; n2 += ga * d2H; n3 += ga * d3H; }
        if (inb) { A = nA; C = nC; Bd = nB; d1 = n1; d2 = n2; d3 = n3; }
    }
    double x1 = inb ? d1 / Bd : 0.0, x2 = inb ? d2 / Bd : 0.0, x3 = inb ? d3 / Bd : 0.0;
    double x10 = seg_get<SEG>(x1, 0), x20 = seg_get<SEG>(x2, 0), x30 = seg_get<SEG>(x3, 0);
    double x1l = seg_get<SEG>(x1, m - 1), x2l = seg_get<SEG>(x2, m - 1), x3l = seg_get<SEG>(x3, m - 1);
    double wX1 = w0 * x10 + wl * x1l;
    double wX2 = w0 * x20 + wl * x2l;
    double wX3 = w0 * x30 + wl * x3l;
    double schur = diagLast - wX3;
    double y1l = (rLast - wX1) / schur;
    double ul = (bLast - wX2) / schur;
    double y = inb ? x1 - x3 * y1l : ((e == m) ? y1l : 0.0);
    double u = inb ? x2 - x3 * ul : ((e == m) ? ul : 0.0);
    double bu = seg_sum<SEG>(live ? bv * u : 0.0);
    double by = seg_sum<SEG>(live ? bv * y : 0.0);
    double sigma = c - bu;
    double z = (sigma > zThresh * c) ? (s - by) / sigma : 0.0;
    *zOut = z;
    return live ? y - z * u : 0.0;
}

// Force of vertex e (lane e of the segment, e < n) with the components along the n
// edge-length gradients and the area gradient of its polygon removed.  Every lane of
// the segment calls it; padding lanes (e >= n) pass anything and get it back.
template <int SEG>
__device__ double2 seg_project_force(int e, int n, d2 pe, double2 fe) {
    bool live = e < n;
    int ef = (e + 1 >= n) ? 0 : e + 1, eg = (e == 0) ? n - 1 : e - 1;
    double pfx = seg_get<SEG>(pe.x, ef), pfy = seg_get<SEG>(pe.y, ef);
    double pgx = seg_get<SEG>(pe.x, eg), pgy = seg_get<SEG>(pe.y, eg);
    double tx = 0.0, ty = 0.0, ax = 0.0, ay = 0.0;
    if (live) {
        double dx = wrap(pfx - pe.x), dy = wrap(pfy - pe.y);
        double len = sqrt(dx * dx + dy * dy);
        if (!(len < 1e-14)) { tx = dx / len; ty = dy / len; }
        ax = wrap(pfy - pgy);
        ay = wrap(pgx - pfx);
    }
    double fx = live ? fe.x : 0.0, fy = live ? fe.y : 0.0;
    double ffx = seg_get<SEG>(fx, ef), ffy = seg_get<SEG>(fy, ef);
    double r = live ? tx * (ffx - fx) + ty * (ffy - fy) : 0.0;
    double c = seg_sum<SEG>(ax * ax + ay * ay);
    double sd = seg_sum<SEG>(ax * fx + ay * fy);
    double z;
    double y = seg_bordered_solve<SEG>(e, n, tx, ty, ax, ay, r, c, sd, 1e-28, &z);
    double yg = seg_get<SEG>(y, eg), tgx = seg_get<SEG>(tx, eg), tgy = seg_get<SEG>(ty, eg);
    if (!live) return fe;
    double px = -y * tx + yg * tgx + z * ax;
    double py = -y * ty + yg * tgy + z * ay;
    return make_double2(fe.x - px, fe.y - py);
}

// projection of the force of every polygon onto the tangent space of its constraints
template <int SEG>
__global__ __launch_bounds__(BLK) void k_project_seg(int N, const int* __restrict__ polyStart,
                                                    const double2* __restrict__ pos,
                                                    double2* __restrict__ force) {
    int e = threadIdx.x & (SEG - 1);
    int s = (blockIdx.x * BLK + threadIdx.x) / SEG;
    if (s >= N) return;   // whole segments leave together
    int st = polyStart[s], n = polyStart[s + 1] - st;
    if (n < 3 || n > SEG) return;
    bool live = e < n;
    d2 pe = live ? ldpos(pos, st + e) : mk(0.0, 0.0);
    double2 fe = live ? force[st + e] : make_double2(0.0, 0.0);
    double2 fo = seg_project_force<SEG>(e, n, pe, fe);
    if (live) force[st + e] = fo;
}

// SHAKE iterations of one polygon held in the lanes of its segment: Gauss-Newton
// projection onto {|edge e| = L_e, area = A0} until max|C| < tol (see k_shake).
template <int SEG>
__device__ int seg_shake(int e, int n, d2& p, double L, double A0, int maxIter, double tol) {
    bool live = e < n;
    int ef = (e + 1 >= n) ? 0 : e + 1, eg = (e == 0) ? n - 1 : e - 1;
    int it = 0;
    for (; it < maxIter; ++it) {
        double pfx = seg_get<SEG>(p.x, ef), pfy = seg_get<SEG>(p.y, ef);
        double pgx = seg_get<SEG>(p.x, eg), pgy = seg_get<SEG>(p.y, eg);
        double p0x = seg_get<SEG>(p.x, 0), p0y = seg_get<SEG>(p.y, 0);
        double ce = 0.0, tx = 0.0, ty = 0.0, ax = 0.0, ay = 0.0, part = 0.0;
        if (live) {
            double dx = wrap(pfx - p.x), dy = wrap(pfy - p.y);
            double len = sqrt(dx * dx + dy * dy);
            ce = len - L;
            if (len > 0.0) { tx = dx / len; ty = dy / len; }
            ax = wrap(pfy - pgy) / 2;
            ay = wrap(pgx - pfx) / 2;
            double x1 = wrap(p.x - p0x);
            double dy2 = wrap(pfy - p.y);
            double x2 = wrap(pfx - p0x);
            part = dy2 * (x2 + x1) / 2.0;   // shoelace term of edge e (poly_area_dev)
        }
        double cmax = seg_max<SEG>(fabs(ce));
        double c = seg_sum<SEG>(ax * ax + ay * ay);
        double cA = seg_sum<SEG>(part) - A0;
        cmax = fmax(cmax, fabs(cA));
        if (cmax < tol) break;
        double z;
        double y = seg_bordered_solve<SEG>(e, n, tx, ty, ax, ay, ce, c, cA, 1e-28, &z);
        double yg = seg_get<SEG>(y, eg), tgx = seg_get<SEG>(tx, eg), tgy = seg_get<SEG>(ty, eg);
        if (live) {
            double dx = -y * tx + yg * tgx + z * ax;
            double dy = -y * ty + yg * tgy + z * ay;
            double qx = p.x - dx, qy = p.y - dy;
            p.x = qx - floor(qx);
            p.y = qy - floor(qy);
        }
    }
    return it;
}

template <int SEG>
__global__ __launch_bounds__(BLK) void k_shake_seg(int N, int P, const int* __restrict__ polyStart,
                                                  const int* __restrict__ polyPack, double2* __restrict__ pos,
                                                  const double* __restrict__ targetEdgeV,
                                                  const double* __restrict__ targetAreas, int maxIter, double tol,
                                                  int retryIter, const int* __restrict__ active,
                                                  int* __restrict__ itOut) {
    int e = threadIdx.x & (SEG - 1);
    int s = (blockIdx.x * BLK + threadIdx.x) / SEG;
    if (s >= N) return;
    int pk = polyPack[s];
    if (active && !active[pk]) return;
    int st = polyStart[s], n = polyStart[s + 1] - st;
    if (n < 3 || n > SEG) return;
    bool live = e < n;
    d2 p = live ? ldpos(pos, st + e) : mk(0.0, 0.0);
    double L = live ? targetEdgeV[st + e] : 0.0;
    double A0 = targetAreas[s];
    int it1 = seg_shake<SEG>(e, n, p, L, A0, maxIter, tol);
    int it2 = -1;
    if (retryIter > 0 && it1 >= maxIter) it2 = seg_shake<SEG>(e, n, p, L, A0, retryIter, tol);
    if (live) pos[st + e] = make_double2(p.x, p.y);
    if (e == 0) {
        atomicMax(&itOut[pk], it1);
        if (it2 >= 0) atomicMax(&itOut[P + pk], it2);
    }
}

int run_project_seg(pcp_ctx* c, const double2* pos, double2* force) {
    int seg = c->seg;
    int grid = (c->N + BLK / seg - 1) / (BLK / seg);
    switch (seg) {
    case 4: PCP_LAUNCH(c, k_project_seg<4>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force); break;
    case 8: PCP_LAUNCH(c, k_project_seg<8>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force); break;
    case 16: PCP_LAUNCH(c, k_project_seg<16>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force); break;
    case 32: PCP_LAUNCH(c, k_project_seg<32>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force); break;
    default: PCP_LAUNCH(c, k_project_seg<64>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force); break;
    }
    return PCP_OK;
}

int run_shake_seg(pcp_ctx* c, double2* pos, const double* targetEdgeV, const double* targetAreas, int maxIter,
                  double tol, int retryIter, const int* active, int* itOut) {
    int seg = c->seg;
    int grid = (c->N + BLK / seg - 1) / (BLK / seg);
#define PCP_SHAKE_SEG(W)                                                                             \
    PCP_LAUNCH(c, k_shake_seg<W>, dim3(grid), dim3(BLK), c->N, c->P, c->dPolyStart, c->dPolyPack, pos, \
               targetEdgeV, targetAreas, maxIter, tol, retryIter, active, itOut)
    switch (seg) {
    case 4: PCP_SHAKE_SEG(4); break;
    case 8: PCP_SHAKE_SEG(8); break;
    case 16: PCP_SHAKE_SEG(16); break;
    case 32: PCP_SHAKE_SEG(32); break;
    default: PCP_SHAKE_SEG(64); break;
    }
#undef PCP_SHAKE_SEG
    return PCP_OK;
}

}  // namespace pcp
