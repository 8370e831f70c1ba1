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

// Row mode: when every polygon has exactly 16 vertices, a 16-lane segment is one DPP row
// and the lane moves of the solves below are VALU modifiers instead of shuffles (each
// shuffle an LDS round trip, and the solve is a chain of them).  With CTRL (DPP on gfx9):
// row_ror:15 = lane e+1 of the row (cyclic), row_ror:1 = lane e-1 (cyclic), row_shr:s =
// lane e-s, row_shl:s = lane e+s (lanes whose source leaves the row keep their own value),
// row_ror:8 = lane e^8, quad_perm [2,3,0,1] = e^2, [1,0,3,2] = e^1.  Values are moved, not
// combined, so every sum keeps its operands and order.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double row_next(double v) { return dpp_mov<0x12F>(v); }
__device__ __forceinline__ double row_prev(double v) { return dpp_mov<0x121>(v); }
template <int S> __device__ __forceinline__ double row_down(double v) { return dpp_mov<0x110 + S>(v); }
template <int S> __device__ __forceinline__ double row_up(double v) { return dpp_mov<0x100 + S>(v); }
__device__ __forceinline__ double row_xor(double v, int o) {
    if (o == 8) return dpp_mov<0x128>(v);
    if (o == 4) {   // lane e^4: e-4 where bit 2 is set, else e+4 (both inside the row)
        double a = dpp_mov<0x114>(v), b = dpp_mov<0x104>(v);
        return (threadIdx.x & 4) ? a : b;
    }
    if (o == 2) return dpp_mov<0x4E>(v);
    return dpp_mov<0xB1>(v);
}
// lane k of every row of the wavefront, each lane getting its own row's
__device__ __forceinline__ double row_bcast(double v, int k) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    int r = (threadIdx.x >> 4) & 3;
    int l0 = __builtin_amdgcn_readlane(lo, k), h0 = __builtin_amdgcn_readlane(hi, k);
    int l1 = __builtin_amdgcn_readlane(lo, 16 + k), h1 = __builtin_amdgcn_readlane(hi, 16 + k);
    int l2 = __builtin_amdgcn_readlane(lo, 32 + k), h2 = __builtin_amdgcn_readlane(hi, 32 + k);
    int l3 = __builtin_amdgcn_readlane(lo, 48 + k), h3 = __builtin_amdgcn_readlane(hi, 48 + k);
    int l = (r == 0) ? l0 : (r == 1) ? l1 : (r == 2) ? l2 : l3;
    int h = (r == 0) ? h0 : (r == 1) ? h1 : (r == 2) ? h2 : h3;
    return __hiloint2double(h, l);
}
// segment sum / max in the butterfly order of seg_sum / seg_max
template <int SEG, bool ROW>
__device__ __forceinline__ double segm_sum(double v) {
    if constexpr (ROW) {
#pragma unroll
        for (int o = 8; o >= 1; o >>= 1) v += row_xor(v, o);
        return v;
    } else {
        return seg_sum<SEG>(v);
    }
}
template <int SEG, bool ROW>
__device__ __forceinline__ double segm_max(double v) {
    if constexpr (ROW) {
#pragma unroll
        for (int o = 8; o >= 1; o >>= 1) v = fmax(v, row_xor(v, o));
        return v;
    } else {
        return seg_max<SEG>(v);
    }
}

// Solve for lane e < n: returns y_e (the edge multipliers), *zOut = area multiplier.
// t = unit edge tangent (0 for a degenerate edge), a = area gradient of vertex e,
// r = right-hand side of edge e, c/s = area-row entries (segment-uniform).
// zThresh: the area direction is dropped when its Schur complement is <= zThresh * c.
// One step of the parallel cyclic reduction (stride st); ROW: st is the template S.
template <int SEG, bool ROW, int S>
__device__ __forceinline__ void pcr_step(int e, int m, int st, bool inb, double& A, double& C, double& Bd, double& d1,
                                         double& d2, double& d3) {
    double inv = 1.0 / Bd;
    int lo = e - st, hi = e + st;
    bool hasLo = inb && lo >= 0, hasHi = inb && hi < m;
    double aL, cL, iL, d1L, d2L, d3L, aH, cH, iH, d1H, d2H, d3H;
    if constexpr (ROW) {   // lanes without a partner read anything and ignore it
        aL = row_down<S>(A); cL = row_down<S>(C); iL = row_down<S>(inv);
        d1L = row_down<S>(d1); d2L = row_down<S>(d2); d3L = row_down<S>(d3);
        aH = row_up<S>(A); cH = row_up<S>(C); iH = row_up<S>(inv);
        d1H = row_up<S>(d1); d2H = row_up<S>(d2); d3H = row_up<S>(d3);
    } else {
        int srcLo = hasLo ? lo : e, srcHi = hasHi ? hi : e;
        aL = seg_get<SEG>(A, srcLo); cL = seg_get<SEG>(C, srcLo); iL = seg_get<SEG>(inv, srcLo);
        d1L = seg_get<SEG>(d1, srcLo); d2L = seg_get<SEG>(d2, srcLo); d3L = seg_get<SEG>(d3, srcLo);
        aH = seg_get<SEG>(A, srcHi); cH = seg_get<SEG>(C, srcHi); iH = seg_get<SEG>(inv, srcHi);
        d1H = seg_get<SEG>(d1, srcHi); d2H = seg_get<SEG>(d2, srcHi); d3H = seg_get<SEG>(d3, srcHi);
    }
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

// ROW: the polygon has 16 vertices and its segment is a DPP row.
template <int SEG, bool ROW = false>
__device__ double seg_bordered_solve(int e, int n, double tx, double ty, double ax, double ay, double r,
                                     double c, double s, double zThresh, double* zOut) {
    int en = (e + 1 >= n) ? 0 : e + 1;   // padding lanes (e >= n) read lane 0
    bool live = e < n;
    double txn, tyn, axn, ayn;
    if constexpr (ROW) {
        txn = row_next(tx); tyn = row_next(ty); axn = row_next(ax); ayn = row_next(ay);
    } else {
        txn = seg_get<SEG>(tx, en); tyn = seg_get<SEG>(ty, en);
        axn = seg_get<SEG>(ax, en); ayn = seg_get<SEG>(ay, en);
    }
    double off = live ? -(tx * txn + ty * tyn) : 0.0;       // G_{e,e+1 mod n}
    double dsq = tx * tx + ty * ty;
    double diag = (dsq > 0.0) ? 2.0 * dsq : 1.0;
    double bv = live ? tx * (axn - ax) + ty * (ayn - ay) : 0.0;
    int m = n - 1;
    double offPrev, offM, offM1, rLast, bLast, diagLast;
    if constexpr (ROW) {   // m = 15
        offPrev = row_down<1>(off);   // lane 0 keeps its own, as e == 0 reads lane 0
        offM = row_bcast(off, 15);
        offM1 = row_bcast(off, 14);
        rLast = row_bcast(r, 15); bLast = row_bcast(bv, 15); diagLast = row_bcast(diag, 15);
    } else {
        offPrev = seg_get<SEG>(off, (e == 0) ? 0 : e - 1);
        offM = seg_get<SEG>(off, m);        // G_{n-1,0}: cyclic corner
        offM1 = seg_get<SEG>(off, m - 1);   // G_{n-2,n-1}
        rLast = seg_get<SEG>(r, m); bLast = seg_get<SEG>(bv, m); diagLast = seg_get<SEG>(diag, m);
    }
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
    if constexpr (ROW) {   // m = 15: strides 1, 2, 4, 8
        pcr_step<SEG, true, 1>(e, m, 1, inb, A, C, Bd, d1, d2, d3);
        pcr_step<SEG, true, 2>(e, m, 2, inb, A, C, Bd, d1, d2, d3);
        pcr_step<SEG, true, 4>(e, m, 4, inb, A, C, Bd, d1, d2, d3);
        pcr_step<SEG, true, 8>(e, m, 8, inb, A, C, Bd, d1, d2, d3);
    } else {
        for (int st = 1; st < m; st <<= 1) pcr_step<SEG, false, 1>(e, m, st, inb, A, C, Bd, d1, d2, d3);
    }
    double x1 = inb ? d1 / Bd : 0.0, x2 = inb ? d2 / Bd : 0.0, x3 = inb ? d3 / Bd : 0.0;
    double x10, x20, x30, x1l, x2l, x3l;
    if constexpr (ROW) {
        x10 = row_bcast(x1, 0); x20 = row_bcast(x2, 0); x30 = row_bcast(x3, 0);
        x1l = row_bcast(x1, 14); x2l = row_bcast(x2, 14); x3l = row_bcast(x3, 14);
    } else {
        x10 = seg_get<SEG>(x1, 0); x20 = seg_get<SEG>(x2, 0); x30 = seg_get<SEG>(x3, 0);
        x1l = seg_get<SEG>(x1, m - 1); x2l = seg_get<SEG>(x2, m - 1); x3l = seg_get<SEG>(x3, m - 1);
    }
    double wX1 = w0 * x10 + wl * x1l;
    double wX2 = w0 * x20 + wl * x2l;
    double wX3 = w0 * x30 + wl * x3l;
    double schur = diagLast - wX3;
    double y1l = (rLast - wX1) / schur;
    double ul = (bLast - wX2) / schur;
    double y = inb ? x1 - x3 * y1l : ((e == m) ? y1l : 0.0);
    double u = inb ? x2 - x3 * ul : ((e == m) ? ul : 0.0);
    double bu = segm_sum<SEG, ROW>(live ? bv * u : 0.0);
    double by = segm_sum<SEG, ROW>(live ? bv * y : 0.0);
    double sigma = c - bu;
    double z = (sigma > zThresh * c) ? (s - by) / sigma : 0.0;
    *zOut = z;
    return live ? y - z * u : 0.0;
}

// Force of vertex e (lane e of the segment, e < n) with the components along the n
// edge-length gradients and the area gradient of its polygon removed.  Every lane of
// the segment calls it; padding lanes (e >= n) pass anything and get it back.
template <int SEG, bool ROW = false>
__device__ double2 seg_project_force(int e, int n, d2 pe, double2 fe) {
    bool live = e < n;
    int ef = (e + 1 >= n) ? 0 : e + 1, eg = (e == 0) ? n - 1 : e - 1;
    double pfx, pfy, pgx, pgy;
    if constexpr (ROW) {
        pfx = row_next(pe.x); pfy = row_next(pe.y); pgx = row_prev(pe.x); pgy = row_prev(pe.y);
    } else {
        pfx = seg_get<SEG>(pe.x, ef); pfy = seg_get<SEG>(pe.y, ef);
        pgx = seg_get<SEG>(pe.x, eg); pgy = seg_get<SEG>(pe.y, eg);
    }
    double tx = 0.0, ty = 0.0, ax = 0.0, ay = 0.0;
    if (live) {
        double dx = wrap(pfx - pe.x), dy = wrap(pfy - pe.y);
        double len = sqrt(dx * dx + dy * dy);
        if (!(len < 1e-14)) { tx = dx / len; ty = dy / len; }
        ax = wrap(pfy - pgy);
        ay = wrap(pgx - pfx);
    }
    double fx = live ? fe.x : 0.0, fy = live ? fe.y : 0.0;
    double ffx, ffy;
    if constexpr (ROW) {
        ffx = row_next(fx); ffy = row_next(fy);
    } else {
        ffx = seg_get<SEG>(fx, ef); ffy = seg_get<SEG>(fy, ef);
    }
    double r = live ? tx * (ffx - fx) + ty * (ffy - fy) : 0.0;
    double c = segm_sum<SEG, ROW>(ax * ax + ay * ay);
    double sd = segm_sum<SEG, ROW>(ax * fx + ay * fy);
    double z;
    double y = seg_bordered_solve<SEG, ROW>(e, n, tx, ty, ax, ay, r, c, sd, 1e-28, &z);
    double yg, tgx, tgy;
    if constexpr (ROW) {
        yg = row_prev(y); tgx = row_prev(tx); tgy = row_prev(ty);
    } else {
        yg = seg_get<SEG>(y, eg); tgx = seg_get<SEG>(tx, eg); tgy = seg_get<SEG>(ty, eg);
    }
    if (!live) return fe;
    double px = -y * tx + yg * tgx + z * ax;
    double py = -y * ty + yg * tgy + z * ay;
    return make_double2(fe.x - px, fe.y - py);
}

// projection of the force of every polygon onto the tangent space of its constraints
template <int SEG, bool ROW = false>
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
    double2 fo = seg_project_force<SEG, ROW>(e, n, pe, fe);
    if (live) force[st + e] = fo;
}

// SHAKE iterations of one polygon held in the lanes of its segment: Gauss-Newton
// projection onto {|edge e| = L_e, area = A0} until max|C| < tol (see k_shake).
template <int SEG, bool ROW = false>
__device__ int seg_shake(int e, int n, d2& p, double L, double A0, int maxIter, double tol) {
    bool live = e < n;
    int ef = (e + 1 >= n) ? 0 : e + 1, eg = (e == 0) ? n - 1 : e - 1;
    int it = 0;
    for (; it < maxIter; ++it) {
        double pfx, pfy, pgx, pgy, p0x, p0y;
        if constexpr (ROW) {
            pfx = row_next(p.x); pfy = row_next(p.y); pgx = row_prev(p.x); pgy = row_prev(p.y);
            p0x = row_bcast(p.x, 0); p0y = row_bcast(p.y, 0);
        } else {
            pfx = seg_get<SEG>(p.x, ef); pfy = seg_get<SEG>(p.y, ef);
            pgx = seg_get<SEG>(p.x, eg); pgy = seg_get<SEG>(p.y, eg);
            p0x = seg_get<SEG>(p.x, 0); p0y = seg_get<SEG>(p.y, 0);
        }
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
        double cmax = segm_max<SEG, ROW>(fabs(ce));
        double c = segm_sum<SEG, ROW>(ax * ax + ay * ay);
        double cA = segm_sum<SEG, ROW>(part) - A0;
        cmax = fmax(cmax, fabs(cA));
        if (cmax < tol) break;
        double z;
        double y = seg_bordered_solve<SEG, ROW>(e, n, tx, ty, ax, ay, ce, c, cA, 1e-28, &z);
        double yg, tgx, tgy;
        if constexpr (ROW) {
            yg = row_prev(y); tgx = row_prev(tx); tgy = row_prev(ty);
        } else {
            yg = seg_get<SEG>(y, eg); tgx = seg_get<SEG>(tx, eg); tgy = seg_get<SEG>(ty, eg);
        }
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

// Iteration counts of the BLK / SEG polygons of a block into itOut (max per packing):
// itOut[pk] and itOut[P + pk] are read by every polygon's packing, and one device atomic
// per polygon on them serializes (thousands of polygons, one address); here a block that
// holds one packing issues at most one atomic per counter (max_if_above).  pk < 0: no polygon
// in this slot.
template <int Q>
__device__ __forceinline__ void block_pack_max(int q, bool leader, int pk, int it1, int it2, int P, int* itOut) {
    static_assert(Q <= 64, "one wavefront merges the slots");
    __shared__ int sPk[Q], sIt1[Q], sIt2[Q];
    if (leader) { sPk[q] = pk; sIt1[q] = it1; sIt2[q] = it2; }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    int l = threadIdx.x;
    int mpk = l < Q ? sPk[l] : -1, a1 = l < Q ? sIt1[l] : 0, a2 = l < Q ? sIt2[l] : -1;
    int pk0 = __shfl(mpk, 0);
    if (__all(l >= Q || mpk == pk0)) {
        a1 = wave_max_int(a1);
        a2 = wave_max_int(a2);
        if (l == 0 && pk0 >= 0) {
            max_if_above(&itOut[pk0], a1);
            if (a2 >= 0) max_if_above(&itOut[P + pk0], a2);
        }
    } else if (l < Q && mpk >= 0) {   // packings change inside the block
        max_if_above(&itOut[mpk], a1);
        if (a2 >= 0) max_if_above(&itOut[P + mpk], a2);
    }
}

template <int SEG, bool ROW = false>
__global__ __launch_bounds__(BLK) void k_shake_seg(int N, int P, const int* __restrict__ polyStart,
                                                  const int* __restrict__ polyPack, double2* __restrict__ pos,
                                                  const double* __restrict__ targetEdgeV,
                                                  const double* __restrict__ targetAreas, int maxIter, double tol,
                                                  int retryIter, const int* __restrict__ active,
                                                  int* __restrict__ itOut) {
    int e = threadIdx.x & (SEG - 1);
    int s = (blockIdx.x * BLK + threadIdx.x) / SEG;
    int pk = -1, it1 = 0, it2 = -1;
    if (s < N) {   // no early return: the whole block meets in block_pack_max
        int pkS = polyPack[s];
        int st = polyStart[s], n = polyStart[s + 1] - st;
        if ((!active || active[pkS]) && n >= 3 && n <= SEG) {
            pk = pkS;
            bool live = e < n;
            d2 p = live ? ldpos(pos, st + e) : mk(0.0, 0.0);
            double L = live ? targetEdgeV[st + e] : 0.0;
            double A0 = targetAreas[s];
            it1 = seg_shake<SEG, ROW>(e, n, p, L, A0, maxIter, tol);
            if (retryIter > 0 && it1 >= maxIter) it2 = seg_shake<SEG, ROW>(e, n, p, L, A0, retryIter, tol);
            if (live) pos[st + e] = make_double2(p.x, p.y);
        }
    }
    block_pack_max<BLK / SEG>(threadIdx.x / SEG, e == 0, pk, it1, it2, P, itOut);
}

// XPBD with compliance (xpbdProjectCUDA model.cu:209-278 / kernels.cuh:532-623, extended as
// in Macklin, Mueller & Chentanez, MIG 2016).  One lane segment per polygon, lane e owns
// vertex e and edge e (p_e -> p_{e+1}); positions, edge multipliers and the polygon's area
// and perimeter multipliers stay in registers for all iterations.  Per iteration: the edge
// constraints by colour (even edges -- for odd n the last one after the others, as it shares
// vertex 0 with edge 0 -- then odd edges; the edges of one pass share no vertex), then the
// area and the perimeter constraint together (their 2x2 block; the two gradients are nearly
// parallel, so one after the other converges slowly), or alone.  Every projection applies
//   dlambda = -(C + at lambda) / (sum |grad C|^2 + at),  x += grad C dlambda,
// at = 1/(k dt^2) (k = INFINITY: at = 0 and the edge and area steps are the reference's
// arithmetic, c = 0.5 (len - L)/len and alpha = (A - A0)/max(|grad A|^2, 1e-24)).  The
// perimeter target is the sum of the edge targets in vertex order.  tol > 0: the polygon
// stops after an iteration whose largest |C + at lambda| (edges after the iteration, area
// and perimeter before their step) is below tol; itOut receives the largest stopping
// index per packing.  vel != nullptr (FIRE): dt is the packing's, inactive packings are
// skipped, the result is wrapped into the box and v += wrap(x - x_start)/dt; otherwise
// positions are left unwrapped, as the reference leaves them.
template <int SEG>
__global__ __launch_bounds__(BLK) void k_xpbd_seg(int N, int P, const int* __restrict__ polyStart,
                                                 const int* __restrict__ polyPack, double2* __restrict__ pos,
                                                 const double* __restrict__ targetEdgeV,
                                                 const double* __restrict__ targetEdgeP,
                                                 const double* __restrict__ targetAreas, double kE, double kA,
                                                 double kP, double dtS, const double* __restrict__ dtPack,
                                                 int nIter, double tol, const int* __restrict__ active,
                                                 double2* __restrict__ vel, int* __restrict__ itOut) {
    int e = threadIdx.x & (SEG - 1);
    int s = (blockIdx.x * BLK + threadIdx.x) / SEG;
    int pk = -1, itDone = 0;
    if (s < N) {   // no early return: the whole block meets in block_pack_max
        int pkS = polyPack[s];
        int st = polyStart[s], n = polyStart[s + 1] - st;
        if ((!active || active[pkS]) && n >= 3 && n <= SEG) {
            pk = pkS;
            bool live = e < n;
            int en = (e + 1 >= n) ? 0 : e + 1, ep = (e == 0) ? n - 1 : e - 1;
            double2 p0 = live ? pos[st + e] : make_double2(0.0, 0.0);
            double x = p0.x, y = p0.y;
            double L = 0.0;
            if (live) L = targetEdgeV ? targetEdgeV[st + e] : (targetEdgeP ? targetEdgeP[s] : 0.0);
            double P0 = 0.0;
            for (int q = 0; q < n; ++q) P0 += seg_get<SEG>(L, q);
            bool edgeOn = kE != 0.0, areaOn = kA != 0.0, perimOn = kP != 0.0;
            double h = dtPack ? dtPack[pkS] : dtS;
            double atE = edgeOn ? 1.0 / (kE * (h * h)) : 0.0;
            double atA = areaOn ? 1.0 / (kA * (h * h)) : 0.0;
            double atP = perimOn ? 1.0 / (kP * (h * h)) : 0.0;
            double A0 = areaOn ? targetAreas[s] : 0.0;
            double lam = 0.0, lamA = 0.0, lamP = 0.0;
            bool odd = (n & 1) != 0;
            int it = 0;
            for (; it < nIter; ++it) {
                if (edgeOn) {
                    int nPass = odd ? 3 : 2;
                    for (int pass = 0; pass < nPass; ++pass) {
                        bool last = odd && e == n - 1;
                        bool act = live && ((pass == 0) ? (!(e & 1) && !last)
                                                        : (pass == 1 && nPass == 3) ? last : (e & 1) != 0);
                        double xn = seg_get<SEG>(x, en), yn = seg_get<SEG>(y, en);
                        double dx = 0.0, dy = 0.0;
                        if (act) {
                            double ex = wrap(xn - x), ey = wrap(yn - y);
                            double len = sqrt(ex * ex + ey * ey);
                            if (len < 1e-14) {
                                act = false;
                            } else {
                                double C = len - L;
                                double dl = (-C - atE * lam) / (2.0 + atE);
                                lam += dl;
                                double c = -dl / len;
                                dx = c * ex;
                                dy = c * ey;
                            }
                        }
                        double dxp = seg_get<SEG>(dx, ep), dyp = seg_get<SEG>(dy, ep);
                        int actp = __shfl((int)act, ep, SEG);
                        double nx = x, ny = y;
                        if (act) { nx = x + dx; ny = y + dy; }
                        if (actp) { nx = x - dxp; ny = y - dyp; }
                        x = nx;
                        y = ny;
                    }
                }
                double devA = 0.0, devP = 0.0;
                double gAx = 0.0, gAy = 0.0, gPx = 0.0, gPy = 0.0, rA = 0.0, rP = 0.0, aa = 0.0, pp = 0.0;
                if (areaOn) {
                    double xs = seg_get<SEG>(x, 0), ys = seg_get<SEG>(y, 0);
                    double xnk = seg_get<SEG>(x, en), ynk = seg_get<SEG>(y, en);
                    double xpk = seg_get<SEG>(x, ep), ypk = seg_get<SEG>(y, ep);
                    double term = 0.0, g2 = 0.0;
                    if (live) {
                        double xk = wrap(x - xs), yk = wrap(y - ys);
                        double xn2 = wrap(xnk - xs), yn2 = wrap(ynk - ys);
                        double xp2 = wrap(xpk - xs), yp2 = wrap(ypk - ys);
                        term = 0.5 * (xk * yn2 - xn2 * yk);
                        gAx = 0.5 * (yn2 - yp2);
                        gAy = 0.5 * (xp2 - xn2);
                        g2 = gAx * gAx + gAy * gAy;
                    }
                    double area = seg_sum<SEG>(term);
                    aa = seg_sum<SEG>(g2);
                    rA = (area - A0) + atA * lamA;
                    devA = fabs(rA);
                }
                if (perimOn) {
                    double xn = seg_get<SEG>(x, en), yn = seg_get<SEG>(y, en);
                    double len = 0.0, t1 = 0.0, t2 = 0.0;
                    if (live) {
                        double ex = wrap(xn - x), ey = wrap(yn - y);
                        len = sqrt(ex * ex + ey * ey);
                        t1 = (len >= 1e-14) ? ex / len : 0.0;
                        t2 = (len >= 1e-14) ? ey / len : 0.0;
                    }
                    double t1p = seg_get<SEG>(t1, ep), t2p = seg_get<SEG>(t2, ep);
                    double gp2 = 0.0;
                    if (live) {
                        gPx = t1p - t1;
                        gPy = t2p - t2;
                        gp2 = gPx * gPx + gPy * gPy;
                    }
                    double Pc = seg_sum<SEG>(len);
                    pp = seg_sum<SEG>(gp2);
                    rP = (Pc - P0) + atP * lamP;
                    devP = fabs(rP);
                }
                if (areaOn && perimOn) {
                    double ap = seg_sum<SEG>(live ? gAx * gPx + gAy * gPy : 0.0);
                    double m11 = aa + atA, m22 = pp + atP;
                    double det = m11 * m22 - ap * ap;
                    if (det > 1e-12 * (m11 * m22)) {
                        double dA = (ap * rP - m22 * rA) / det;
                        double dP = (ap * rA - m11 * rP) / det;
                        lamA += dA;
                        lamP += dP;
                        x += gAx * dA + gPx * dP;
                        y += gAy * dA + gPy * dP;
                    } else {
                        double alpha = rA / (fmax(aa, 1e-24) + atA);
                        lamA -= alpha;
                        x -= alpha * gAx;
                        y -= alpha * gAy;
                    }
                } else if (areaOn) {
                    double alpha = rA / (fmax(aa, 1e-24) + atA);
                    lamA -= alpha;
                    x -= alpha * gAx;
                    y -= alpha * gAy;
                } else if (perimOn) {
                    double alpha = rP / (fmax(pp, 1e-24) + atP);
                    lamP -= alpha;
                    x -= alpha * gPx;
                    y -= alpha * gPy;
                }
                if (tol > 0.0) {
                    double devE = 0.0;
                    if (edgeOn) {
                        double xn = seg_get<SEG>(x, en), yn = seg_get<SEG>(y, en);
                        double d = 0.0;
                        if (live) {
                            double ex = wrap(xn - x), ey = wrap(yn - y);
                            d = fabs(sqrt(ex * ex + ey * ey) - L + atE * lam);
                        }
                        devE = seg_max<SEG>(d);
                    }
                    if ((!edgeOn || devE < tol) && (!areaOn || devA < tol) && (!perimOn || devP < tol)) break;
                }
            }
            itDone = it;
            if (live) {
                if (vel) {
                    double px = x - floor(x), py = y - floor(y);
                    double2 v = vel[st + e];
                    vel[st + e] = make_double2(v.x + wrap(px - p0.x) / h, v.y + wrap(py - p0.y) / h);
                    pos[st + e] = make_double2(px, py);
                } else {
                    pos[st + e] = make_double2(x, y);
                }
            }
        }
    }
    block_pack_max<BLK / SEG>(threadIdx.x / SEG, e == 0, pk, itDone, -1, P, itOut);
}

int run_xpbd_seg(pcp_ctx* c, double2* pos, const double* targetEdgeV, const double* targetEdgeP,
                 const double* targetAreas, double kE, double kA, double kP, double dt, const double* dtPack,
                 int nIter, double tol, const int* active, double2* vel, int* itOut) {
    int seg = c->seg;
    int grid = (c->N + BLK / seg - 1) / (BLK / seg);
#define PCP_XPBD_SEG(W)                                                                                      \
    PCP_LAUNCH(c, k_xpbd_seg<W>, dim3(grid), dim3(BLK), c->N, c->P, c->dPolyStart, c->dPolyPack, pos,        \
               targetEdgeV, targetEdgeP, targetAreas, kE, kA, kP, dt, dtPack, nIter, tol, active, vel, itOut)
    switch (seg) {
    case 4: PCP_XPBD_SEG(4); break;
    case 8: PCP_XPBD_SEG(8); break;
    case 16: PCP_XPBD_SEG(16); break;
    case 32: PCP_XPBD_SEG(32); break;
    default: PCP_XPBD_SEG(64); break;
    }
#undef PCP_XPBD_SEG
    return PCP_OK;
}

// Shape matching of each polygon (lane e owns vertex e): its rest shape ref (offsets from
// the rest shape's vertex mean), scaled to the target area when targetAreas is given,
// turned by the rotation of the polar decomposition of M = sum_e q_e ref_e^T (q_e = the
// current minimum-image offsets from their vertex mean; in 2-D the rotation of the SVD
// polar factor is [[a, -b], [b, a]] / sqrt(a^2 + b^2), a = M00 + M11, b = M10 - M01) and
// placed at the current vertex mean: the rigid motion of the rest shape nearest to the
// current vertices in least squares.  Sums are segment butterflies (zero padding).
template <int SEG>
__global__ __launch_bounds__(BLK) void k_shape_project_seg(int N, const int* __restrict__ polyStart,
                                                          double2* __restrict__ pos,
                                                          const double2* __restrict__ ref,
                                                          const double* __restrict__ targetAreas) {
    int e = threadIdx.x & (SEG - 1);
    int s = (blockIdx.x * BLK + threadIdx.x) / SEG;
    if (s >= N) return;   // whole segments leave together
    int st = polyStart[s], n = polyStart[s + 1] - st;
    if (n < 3 || n > SEG) return;
    bool live = e < n;
    int en = (e + 1 >= n) ? 0 : e + 1;
    double2 p = live ? pos[st + e] : make_double2(0.0, 0.0);
    double2 r = live ? ref[st + e] : make_double2(0.0, 0.0);
    double p0x = seg_get<SEG>(p.x, 0), p0y = seg_get<SEG>(p.y, 0);
    double rnx = seg_get<SEG>(r.x, en), rny = seg_get<SEG>(r.y, en);
    double dx = live ? wrap(p.x - p0x) : 0.0, dy = live ? wrap(p.y - p0y) : 0.0;
    double mx = seg_sum<SEG>(dx) / n, my = seg_sum<SEG>(dy) / n;
    double qx = live ? dx - mx : 0.0, qy = live ? dy - my : 0.0;
    double m00 = seg_sum<SEG>(qx * r.x), m01 = seg_sum<SEG>(qx * r.y);
    double m10 = seg_sum<SEG>(qy * r.x), m11 = seg_sum<SEG>(qy * r.y);
    double twiceA = seg_sum<SEG>(live ? r.x * rny - rnx * r.y : 0.0);   // shoelace of the rest shape
    double a = m00 + m11, b = m10 - m01;
    double h = sqrt(a * a + b * b);
    double ca = (h > 0.0) ? a / h : 1.0, sb = (h > 0.0) ? b / h : 0.0;
    double sc = 1.0;
    if (targetAreas && twiceA > 0.0) sc = sqrt(targetAreas[s] / (twiceA / 2.0));
    if (!live) return;
    double ux = ca * r.x - sb * r.y, uy = sb * r.x + ca * r.y;
    double px = p0x + (mx + sc * ux), py = p0y + (my + sc * uy);
    pos[st + e] = make_double2(px - floor(px), py - floor(py));
}

int run_shape_project_seg(pcp_ctx* c, double2* pos, const double2* ref, const double* targetAreas) {
    int seg = c->seg;
    int grid = (c->N + BLK / seg - 1) / (BLK / seg);
    switch (seg) {
    case 4: PCP_LAUNCH(c, k_shape_project_seg<4>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, ref, targetAreas); break;
    case 8: PCP_LAUNCH(c, k_shape_project_seg<8>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, ref, targetAreas); break;
    case 16: PCP_LAUNCH(c, k_shape_project_seg<16>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, ref, targetAreas); break;
    case 32: PCP_LAUNCH(c, k_shape_project_seg<32>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, ref, targetAreas); break;
    default: PCP_LAUNCH(c, k_shape_project_seg<64>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, ref, targetAreas); break;
    }
    return PCP_OK;
}

int run_project_seg(pcp_ctx* c, const double2* pos, double2* force) {
    int seg = c->seg;
    int grid = (c->N + BLK / seg - 1) / (BLK / seg);
    switch (seg) {
    case 4: PCP_LAUNCH(c, k_project_seg<4>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force); break;
    case 8: PCP_LAUNCH(c, k_project_seg<8>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force); break;
    case 16:
        if (c->row16) PCP_LAUNCH(c, (k_project_seg<16, true>), dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force);
        else PCP_LAUNCH(c, k_project_seg<16>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force);
        break;
    case 32: PCP_LAUNCH(c, k_project_seg<32>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force); break;
    default: PCP_LAUNCH(c, k_project_seg<64>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, pos, force); break;
    }
    return PCP_OK;
}

int run_shake_seg(pcp_ctx* c, double2* pos, const double* targetEdgeV, const double* targetAreas, int maxIter,
                  double tol, int retryIter, const int* active, int* itOut) {
    int seg = c->seg;
    int grid = (c->N + BLK / seg - 1) / (BLK / seg);
#define PCP_SHAKE_SEG(K)                                                                             \
    PCP_LAUNCH(c, K, dim3(grid), dim3(BLK), c->N, c->P, c->dPolyStart, c->dPolyPack, pos, \
               targetEdgeV, targetAreas, maxIter, tol, retryIter, active, itOut)
    switch (seg) {
    case 4: PCP_SHAKE_SEG(k_shake_seg<4>); break;
    case 8: PCP_SHAKE_SEG(k_shake_seg<8>); break;
    case 16:
        if (c->row16) PCP_SHAKE_SEG((k_shake_seg<16, true>));
        else PCP_SHAKE_SEG(k_shake_seg<16>);
        break;
    case 32: PCP_SHAKE_SEG(k_shake_seg<32>); break;
    default: PCP_SHAKE_SEG(k_shake_seg<64>); break;
    }
#undef PCP_SHAKE_SEG
    return PCP_OK;
}

}  // namespace pcp
