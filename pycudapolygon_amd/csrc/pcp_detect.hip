// Edge-crossing detection by polygon pairs (part of the pcp_kernels.hip translation unit).
//
// updateNeighborsKernel (kernels.cuh:709-812) tests every edge against the edges of the
// 3x3 block of cells around its midpoint cell that have a larger vertex id, excluding
// edges that share a vertex.  The same tests are made here, organised by polygon pair:
//
//   k_bounds        bounding circle of every polygon (one lane segment per polygon);
//                   also clears the per-edge crossing counters
//   k_poly_cells    one workgroup per packing: counting sort of the polygons into an LDS
//                   cell grid whose side is at least the largest bounding diameter
//   k_detect_poly   one wavefront per polygon A: its edges staged in LDS; the polygons of
//                   the 3x3 cells around it gathered by the lanes in parallel, kept when
//                   their circle meets A's; for A itself and each kept polygon, its edges
//                   held in lanes and tested against A's.  An edge pair is tested when the
//                   midpoint cells of the two edges in the user's cell grid are neighbours
//                   (the reference's stencil), with the reference's arithmetic (edge of
//                   lower vertex id first).  Each pair of polygons is tested by both of
//                   their wavefronts, and each keeps the records that belong to its own
//                   buckets (si or sj), so every bucket has one writer and needs no atomics
//
// Two crossing edges lie in polygons whose bounding circles meet, so the candidates
// contain every crossing the reference can find, and the cell test drops exactly the
// edge pairs the reference never looks at: for every box size the crossings (keys,
// float parameters, inside flags) are the reference's.  Only the order inside a bucket
// differs from the reference's slot order; every consumer orders the crossings by key.

namespace pcp {

// radius beyond which a polygon makes its packing take all polygon pairs (the minimum
// image of the circle centres is only meaningful for circles well inside half a box)
constexpr double BIG_RADIUS = 0.2;
// cell grid of the polygon sort: at most PC_MAXB x PC_MAXB cells (LDS)
constexpr int PC_MAXB = 90;
constexpr int PC_CELLS = PC_MAXB * PC_MAXB;
constexpr int PP_BLK = 1024;

// ---------------------------------------------------------------- bounding circles
// Chain of the polygon in the frame of its own edge vectors, u_0 = 0,
// u_{k+1} = u_k + wrap(p_{k+1} - p_k) (the edge vectors the segment test uses);
// centre = p_0 + mean u_k, radius = max |u_k - mean u|.  bounds[s] = (cx, cy, r, big),
// big = 1 when the chain does not close (the polygon wraps the torus) or r is large.
template <int SEG>
__global__ __launch_bounds__(BLK) void k_bounds(int N, int P, const int* __restrict__ polyStart,
                                               const double2* __restrict__ pos, double4* __restrict__ bounds,
                                               int* __restrict__ num, int* __restrict__ bcnt, int* __restrict__ jcnt,
                                               unsigned long long* __restrict__ packMaxKey, int* __restrict__ scal) {
    if (blockIdx.x == 0 && threadIdx.x < SC_NSCAL) scal[threadIdx.x] = 0;
    for (int q = blockIdx.x * BLK + threadIdx.x; q < P; q += gridDim.x * BLK) packMaxKey[q] = 0ull;
    int sub = threadIdx.x & (SEG - 1);
    int s = (blockIdx.x * BLK + threadIdx.x) / SEG;
    int st = 0, n = 0;
    if (s < N) {
        st = polyStart[s];
        n = polyStart[s + 1] - st;
        if (sub == 0) { bcnt[s] = 0; jcnt[s] = 0; }
    }
    d2 p0 = (n > 0) ? ldpos(pos, st) : mk(0.0, 0.0);
    // pass 1: chain and its sum (chunks of SEG vertices, running carry)
    double carryX = 0.0, carryY = 0.0, sumX = 0.0, sumY = 0.0;
    for (int k0 = 0; k0 < n; k0 += SEG) {
        int k = k0 + sub;
        double dx = 0.0, dy = 0.0;
        if (k < n) {
            num[st + k] = 0;
            d2 a = ldpos(pos, st + k), b = ldpos(pos, st + ((k + 1 == n) ? 0 : k + 1));
            dx = wrap(b.x - a.x);
            dy = wrap(b.y - a.y);
        }
        double ix = dx, iy = dy;   // inclusive scan over the segment
#pragma unroll
        for (int o = 1; o < SEG; o <<= 1) {
            double tx = __shfl_up(ix, o, SEG), ty = __shfl_up(iy, o, SEG);
            if (sub >= o) { ix += tx; iy += ty; }
        }
        double ux = carryX + ix - dx, uy = carryY + iy - dy;   // u_k
        if (k < n) { sumX += ux; sumY += uy; }
        carryX += __shfl(ix, SEG - 1, SEG);
        carryY += __shfl(iy, SEG - 1, SEG);
    }
#pragma unroll
    for (int o = SEG / 2; o >= 1; o >>= 1) { sumX += __shfl_xor(sumX, o, SEG); sumY += __shfl_xor(sumY, o, SEG); }
    double mx = (n > 0) ? sumX / n : 0.0, my = (n > 0) ? sumY / n : 0.0;
    // pass 2: radius (the chain is rebuilt; it is cheap and keeps registers low)
    double r2 = 0.0;
    double cX = 0.0, cY = 0.0;
    for (int k0 = 0; k0 < n; k0 += SEG) {
        int k = k0 + sub;
        double dx = 0.0, dy = 0.0;
        if (k < n) {
            d2 a = ldpos(pos, st + k), b = ldpos(pos, st + ((k + 1 == n) ? 0 : k + 1));
            dx = wrap(b.x - a.x);
            dy = wrap(b.y - a.y);
        }
        double ix = dx, iy = dy;
#pragma unroll
        for (int o = 1; o < SEG; o <<= 1) {
            double tx = __shfl_up(ix, o, SEG), ty = __shfl_up(iy, o, SEG);
            if (sub >= o) { ix += tx; iy += ty; }
        }
        double ux = cX + ix - dx - mx, uy = cY + iy - dy - my;
        if (k < n) r2 = fmax(r2, ux * ux + uy * uy);
        cX += __shfl(ix, SEG - 1, SEG);
        cY += __shfl(iy, SEG - 1, SEG);
    }
#pragma unroll
    for (int o = SEG / 2; o >= 1; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, SEG));
    if (s < N && sub == 0) {
        double r = sqrt(r2);
        bool big = !(fabs(cX) < 1e-9 && fabs(cY) < 1e-9) || !(r < BIG_RADIUS);
        double cx = p0.x + mx, cy = p0.y + my;
        cx -= floor(cx);
        cy -= floor(cy);
        bounds[s] = make_double4(cx, cy, r, big ? 1.0 : 0.0);
    }
}

// ---------------------------------------------------------------- polygon cells
__device__ __forceinline__ int pc_coord(double x, int B) {
    double f = floor(x * B);
    int i = (f >= 0.0) ? (int)f : 0;
    return (i >= B) ? B - 1 : i;
}

// circles of A and B (minimum image of the centres) meet, with a relative margin so
// that rounding can only add candidates
__device__ __forceinline__ bool circles_meet(double4 a, double4 b) {
    double dx = wrap(b.x - a.x), dy = wrap(b.y - a.y);
    double rr = (a.z + b.z) * (1.0 + 1e-6) + 1e-12;
    return dx * dx + dy * dy <= rr * rr;
}

// One workgroup per packing: counting sort of the polygons by the cell of their circle
// centre in a B x B grid with side >= (1 + 1e-5) x the largest circle diameter, so that
// circles that meet have centres in neighbouring cells.  B = 1 (one cell) when B would be
// below 3 or a polygon is big.  Writes pcStart[p*(PC_CELLS+1) + 0..B*B], the sorted
// polygon ids into pcSorted[packPoly[p]..] (their circles into sBounds[..]), and
// packB[p] = B | big << 30.
__global__ __launch_bounds__(PP_BLK) void k_poly_cells(const int* __restrict__ packPoly,
                                                      const double4* __restrict__ bounds,
                                                      int* __restrict__ pcStart, int* __restrict__ pcSorted,
                                                      int* __restrict__ packB, double4* __restrict__ sBounds) {
    __shared__ int cStart[PC_CELLS + 1];
    __shared__ int cCur[PC_CELLS];
    __shared__ double redR[PP_BLK / 64];
    __shared__ int redBig[PP_BLK / 64];
    __shared__ int scanTmp[PP_BLK / 64];
    int p = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
    int s0 = packPoly[p], s1 = packPoly[p + 1], Np = s1 - s0;
    long long gbase = (long long)p * (PC_CELLS + 1);
    double r = 0.0;
    int big = 0;
    for (int s = s0 + t; s < s1; s += PP_BLK) {
        double4 b = bounds[s];
        r = fmax(r, b.z);
        big |= (b.w != 0.0);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) { r = fmax(r, __shfl_xor(r, o)); big |= __shfl_xor(big, o); }
    if (lane == 0) { redR[w] = r; redBig[w] = big; }
    __syncthreads();
    r = 0.0; big = 0;
    for (int q = 0; q < PP_BLK / 64; ++q) { r = fmax(r, redR[q]); big |= redBig[q]; }
    int B = 1;
    if (!big && Np > 1) {
        double side = 2.0 * r * (1.0 + 1e-5) + 1e-11;
        double q = floor(1.0 / side);
        B = (q >= (double)PC_MAXB) ? PC_MAXB : (int)q;
        if (B < 3) B = 1;   // every cell would neighbour every other
    }
    int C = B * B;
    for (int q = t; q < C; q += PP_BLK) cCur[q] = 0;
    __syncthreads();
    for (int s = s0 + t; s < s1; s += PP_BLK) {
        double4 b = bounds[s];
        atomicAdd(&cCur[pc_coord(b.y, B) * B + pc_coord(b.x, B)], 1);
    }
    __syncthreads();
    {   // exclusive scan of the counts: each thread a contiguous run of cells
        int per = (C + PP_BLK - 1) / PP_BLK;
        int q0 = min(t * per, C), q1 = min(q0 + per, C);
        int loc = 0;
        for (int q = q0; q < q1; ++q) loc += cCur[q];
        int x = loc;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) scanTmp[w] = x;
        __syncthreads();
        int off = 0;
        for (int q = 0; q < w; ++q) off += scanTmp[q];
        int run = off + x - loc;
        for (int q = q0; q < q1; ++q) {
            int cnt = cCur[q];
            cStart[q] = run;
            cCur[q] = run;
            run += cnt;
        }
    }
    __syncthreads();
    for (int q = t; q < C; q += PP_BLK) pcStart[gbase + q] = cStart[q];
    if (t == 0) {
        pcStart[gbase + C] = Np;
        packB[p] = B | (big ? (1 << 30) : 0);
    }
    for (int s = s0 + t; s < s1; s += PP_BLK) {
        double4 b = bounds[s];
        int slot = atomicAdd(&cCur[pc_coord(b.y, B) * B + pc_coord(b.x, B)], 1);
        pcSorted[s0 + slot] = s;
        sBounds[s0 + slot] = b;
    }
}

// ---------------------------------------------------------------- edge pairs
// LDS hand-off between the lanes of one wavefront: the LDS executes a wavefront's
// instructions in order, so only compiler reordering has to be prevented (no wait for
// outstanding global loads is needed here)
__device__ __forceinline__ void wave_lds_sync() {
    __asm__ __volatile__("" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

// cell of an edge in the user grid as x | y << 16
__device__ __forceinline__ int cell_xy(int flat, int B) { return (flat % B) | ((flat / B) << 16); }
__device__ __forceinline__ bool cells_adjacent(int ca, int cb, int B) {
    int dx = abs((ca & 0xFFFF) - (cb & 0xFFFF)), dy = abs((ca >> 16) - (cb >> 16));
    return (dx <= 1 || dx == B - 1) && (dy <= 1 || dy == B - 1);
}

struct EdgeStage {
    double2* P;
    double2* R;
    int* C;
};

// Crossings found by the wavefront of polygon A, appended to A's own buckets: a record
// whose si is A to the si bucket, one whose sj is A to the sj bucket (a crossing of A
// with itself to both).  The wavefront is the only writer of A's buckets, so slots come
// from ballots, without atomics.
struct Appender {
    long long base;   // A * Kp
    int Kp;
    int nSi, nSj;     // wavefront-uniform counts so far
    uint64_t* bKey;
    float2* bTU;
    uint64_t* jKey;
};
__device__ __forceinline__ void append_records(Appender& ap, int lane, bool toSi, bool toSj, uint64_t key,
                                               float2 tu) {
    unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    unsigned long long mi = __ballot(toSi), mj = __ballot(toSj);
    if (toSi) {
        int slot = ap.nSi + __popcll(mi & below);
        if (slot < ap.Kp) { ap.bKey[ap.base + slot] = key; ap.bTU[ap.base + slot] = tu; }
    }
    if (toSj) {
        int slot = ap.nSj + __popcll(mj & below);
        if (slot < ap.Kp) ap.jKey[ap.base + slot] = key;
    }
    ap.nSi += __popcll(mi);
    ap.nSj += __popcll(mj);
}

constexpr int TILE_UNROLL = 4;

// Test the edges of polygon A (staged in LDS, chunks of 64) against every edge of polygon
// Bq (held in lanes).  Each edge pair is computed with the edge of lower vertex id as the
// reference's edge 1 (updateNeighborsKernel tests id1 against nid > id1), so the tests
// from A's and from Bq's wavefront are the same arithmetic.  A == Bq tests each pair of
// non-adjacent edges once.  Every lane runs the same number of iterations (ballots).
template <int CELLS>
__device__ void edge_tile(int lane, int A, int stA, int nA, int Bq, int stB, int nB, int sAl, int sBl, int Bu,
                          const double2* __restrict__ pos, const int* __restrict__ cellOf, EdgeStage sg,
                          int& staged, Appender& ap, int* __restrict__ num, uint64_t& kmax) {
    const double eps = 1e-12;
    bool self = (A == Bq);
    bool aFirst = (A < Bq);
    for (int a0 = 0; a0 < nA; a0 += 64) {
        int na = min(64, nA - a0);
        if (staged != a0) {
            wave_lds_sync();
            if (lane < na) {
                int a = a0 + lane;
                d2 p0 = ldpos(pos, stA + a), p1 = ldpos(pos, stA + ((a + 1 == nA) ? 0 : a + 1));
                sg.P[lane] = make_double2(p0.x, p0.y);
                sg.R[lane] = make_double2(wrap(p1.x - p0.x), wrap(p1.y - p0.y));
                sg.C[lane] = cell_xy(CELLS ? cellOf[stA + a] : edge_cell(p0, p1, Bu), Bu);
            }
            wave_lds_sync();
            staged = a0;
        }
        for (int b0 = 0; b0 < nB; b0 += 64) {
            int nb = min(64, nB - b0);
            int G = 64 / nb;
            int g = lane / nb, bl = lane - g * nb;
            bool act = g < G;
            int b = b0 + bl;
            d2 Q = mk(0.0, 0.0), S = mk(0.0, 0.0);
            int cb = 0;
            if (act) {
                d2 q0 = ldpos(pos, stB + b), q1 = ldpos(pos, stB + ((b + 1 == nB) ? 0 : b + 1));
                Q = q0;
                S = mk(wrap(q1.x - q0.x), wrap(q1.y - q0.y));
                cb = cell_xy(CELLS ? cellOf[stB + b] : edge_cell(q0, q1, Bu), Bu);
            }
            int iters = (na + TILE_UNROLL * G - 1) / (TILE_UNROLL * G);
            for (int it = 0; it < iters; ++it) {
                double2 PA[TILE_UNROLL], RA[TILE_UNROLL];
                int CA[TILE_UNROLL];
                bool ok[TILE_UNROLL];
#pragma unroll
                for (int u = 0; u < TILE_UNROLL; ++u) {   // LDS loads of the unrolled pairs first
                    int al = g + (it * TILE_UNROLL + u) * G;
                    int a = a0 + al;
                    ok[u] = act && al < na;
                    if (self) ok[u] = ok[u] && b > a && b != a + 1 && !(a == 0 && b == nA - 1);
                    int ls = ok[u] ? al : 0;
                    PA[u] = sg.P[ls];
                    RA[u] = sg.R[ls];
                    CA[u] = sg.C[ls];
                }
#pragma unroll
                for (int u = 0; u < TILE_UNROLL; ++u) {
                    int al = g + (it * TILE_UNROLL + u) * G;
                    int a = a0 + al;
                    bool hit = ok[u] && cells_adjacent(CA[u], cb, Bu);
                    // edge 1 = the edge of lower vertex id
                    bool af = self || aFirst;
                    d2 P1 = af ? mk(PA[u].x, PA[u].y) : Q, R1 = af ? mk(RA[u].x, RA[u].y) : S;
                    d2 P2 = af ? Q : mk(PA[u].x, PA[u].y), R2 = af ? S : mk(RA[u].x, RA[u].y);
                    double denom = R1.x * R2.y - R1.y * R2.x;
                    hit = hit && !(fabs(denom) < eps);
                    double gx = wrap(P2.x - P1.x);
                    double gy = wrap(P2.y - P1.y);
                    double nt = gx * R2.y - gy * R2.x;
                    double nu = gx * R1.y - gy * R1.x;
                    // t = nt/denom and u = nu/denom lie in (0,1) only if nt and nu lie strictly
                    // between 0 and denom (a correctly rounded quotient keeps its sign and is
                    // >= 1 when |numerator| >= |denom|); only those pairs need the divisions
                    hit = hit && ((denom > 0.0) ? (nt > 0.0 && nt < denom && nu > 0.0 && nu < denom)
                                                : (nt < 0.0 && nt > denom && nu < 0.0 && nu > denom));
                    uint64_t key = 0;
                    float2 tu = make_float2(0.f, 0.f);
                    bool toSi = false, toSj = false;
                    if (hit) {
                        double tt = nt / denom;
                        double uu = nu / denom;
                        if (tt > 0.0 && tt < 1.0 && uu > 0.0 && uu < 1.0) {
                            int s1l = af ? sAl : sBl, s2l = af ? sBl : sAl;
                            int n1 = af ? a : b, l2 = af ? b : a;
                            // updateCompactedIntersectionsKernel kernels.cuh:857-876
                            if (denom > 0) {
                                key = ((uint64_t)(uint16_t)s1l << 48) | ((uint64_t)(uint16_t)s2l << 32) |
                                      ((uint64_t)(uint16_t)l2 << 16) | (uint64_t)(uint16_t)n1;
                                tu = make_float2((float)tt, (float)uu);
                            } else {
                                key = ((uint64_t)(uint16_t)s2l << 48) | ((uint64_t)(uint16_t)s1l << 32) |
                                      ((uint64_t)(uint16_t)n1 << 16) | (uint64_t)(uint16_t)l2;
                                tu = make_float2((float)uu, (float)tt);
                            }
                            toSi = (key_si(key) == sAl);
                            toSj = (key_sj(key) == sAl);
                            if (toSi) kmax = (key > kmax) ? key : kmax;
                            if (CELLS && af) atomicAdd(&num[stA + a], 1);   // per-edge count
                        }
                    }
                    append_records(ap, lane, toSi, toSj, key, tu);
                }
            }
        }
    }
}

// One wavefront per polygon A (in cell order): A's edges staged in LDS; the polygons of
// the 3x3 block of cells around A's cell gathered by the lanes in parallel, kept when
// their circle meets A's; A tested against itself and against every kept polygon, and
// every crossing involving A appended to A's buckets.
// CELLS = 0: the user-grid cell of every edge is computed from the live positions (the
// fused evaluation, where cells and crossings come from the same positions);
// CELLS = 1: the cells of the last cell-list build are used (updateNeighbors after
// updateNeighborCells, which may have been called at other positions).
template <int CELLS>
__global__ __launch_bounds__(BLK) void k_detect_poly(
    int N, const int* __restrict__ polyStart, const int* __restrict__ polyPack,
    const int* __restrict__ packPoly, const int* __restrict__ box, const double2* __restrict__ pos,
    const int* __restrict__ cellOf, const double4* __restrict__ bounds, const int* __restrict__ pcStart,
    const int* __restrict__ pcSorted, const double4* __restrict__ sBounds, const int* __restrict__ packB, int Kp,
    int* __restrict__ bcnt, uint64_t* __restrict__ bKey, float2* __restrict__ bTU, int* __restrict__ jcnt,
    uint64_t* __restrict__ jKey, unsigned long long* __restrict__ packMaxKey, int* __restrict__ num,
    int* __restrict__ scal) {
    __shared__ double2 sP[BLK / 64][64];
    __shared__ double2 sR[BLK / 64][64];
    __shared__ int sC[BLK / 64][64];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int idx = xcd_block(blockIdx.x, gridDim.x) * (BLK / 64) + w;
    if (idx >= N) return;   // whole wavefront
    int A = pcSorted[idx];   // polygons in cell order
    EdgeStage sg{sP[w], sR[w], sC[w]};
    int staged = -1;
    int stA = polyStart[A], nA = polyStart[A + 1] - stA;
    int pk = polyPack[A];
    int Bu = box[pk];
    int s0 = packPoly[pk];
    int pb = packB[pk];
    int Bp = pb & 0xFFFF;
    bool big = (pb >> 30) & 1;
    int sAl = A - s0;
    double4 bA = bounds[A];
    uint64_t kmax = 0;
    Appender ap{(long long)A * Kp, Kp, 0, 0, bKey, bTU, jKey};
    // candidate cells: lanes 0..8 hold the ranges of the 3x3 block (one cell when Bp == 1);
    // loaded before the polygon's own tile so that the two latencies overlap
    long long gbase = (long long)pk * (PC_CELLS + 1);
    int ncell = (Bp == 1) ? 1 : 9;
    int lo = 0, cnt = 0;
    if (lane < ncell) {
        int ax = pc_coord(bA.x, Bp), ay = pc_coord(bA.y, Bp);
        int cx = ax, cy = ay;
        if (Bp > 1) {
            cx = ax + lane / 3 - 1; cx += (cx < 0) ? Bp : 0; cx -= (cx >= Bp) ? Bp : 0;
            cy = ay + lane % 3 - 1; cy += (cy < 0) ? Bp : 0; cy -= (cy >= Bp) ? Bp : 0;
        }
        int cell = cy * Bp + cx;
        lo = pcStart[gbase + cell];
        cnt = pcStart[gbase + cell + 1] - lo;
    }
    if (nA >= 4) edge_tile<CELLS>(lane, A, stA, nA, A, stA, nA, sAl, sAl, Bu, pos, cellOf, sg, staged, ap, num, kmax);
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    int T = __shfl(incl, ncell - 1);
    for (int c0 = 0; c0 < T; c0 += 64) {
        int ci = c0 + lane;
        int q = 0;
        bool found = false;
        for (int k = 0; k < ncell; ++k) {   // every lane takes part in the shuffles
            int hiK = __shfl(incl, k), cK = __shfl(cnt, k), loK = __shfl(lo, k);
            if (!found && ci < hiK) { q = loK + ci - (hiK - cK); found = true; }
        }
        int Bq = -1, stq = 0, enq = 0;
        if (ci < T) {
            int cand = pcSorted[s0 + q];
            double4 bc = sBounds[s0 + q];
            if (cand != A && (big || circles_meet(bA, bc))) {
                Bq = cand;
                stq = polyStart[cand];
                enq = polyStart[cand + 1];
            }
        }
        unsigned long long keep = __ballot(Bq >= 0);
        while (keep) {
            int j = __ffsll((long long)keep) - 1;
            keep &= keep - 1;
            int Bsel = __shfl(Bq, j);
            int stB = __shfl(stq, j), nB = __shfl(enq, j) - stB;
            edge_tile<CELLS>(lane, A, stA, nA, Bsel, stB, nB, sAl, Bsel - s0, Bu, pos, cellOf, sg, staged, ap, num,
                             kmax);
        }
    }
    // A's bucket counts; the largest key of the packing (the pairing's last-record case)
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        unsigned long long y = __shfl_xor((unsigned long long)kmax, o);
        kmax = (y > kmax) ? y : kmax;
    }
    if (lane == 0) {
        bcnt[A] = ap.nSi;
        jcnt[A] = ap.nSj;
        int mx = max(ap.nSi, ap.nSj);
        if (mx > Kp) atomicMax(&scal[SC_BUCKET], mx);
        if (kmax > 0 && kmax > packMaxKey[pk]) atomicMax(&packMaxKey[pk], (unsigned long long)kmax);
    }
}

// exact maximum of the per-edge crossing counts (updateNeighbors' overflow check)
__global__ void k_num_max(int V, const int* __restrict__ num, int* __restrict__ out) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    int m = (v < V) ? num[v] : 0;
    m = wave_max_int(m);
    if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

static int max_poly_vertices(const pcp_ctx* c) {
    int m = 1;
    for (int s = 0; s < c->N; ++s) m = std::max(m, c->hPolyStart[s + 1] - c->hPolyStart[s]);
    return m;
}

int seg_width(int maxN) {
    int seg = 4;
    while (seg < maxN && seg < 64) seg <<= 1;
    return seg;
}

int launch_bounds(pcp_ctx* c, const double2* pos) {
    int seg = c->seg;
    int polysPerBlock = BLK / seg;
    int grid = (c->N + polysPerBlock - 1) / polysPerBlock;
    if (grid < 1) grid = 1;
    switch (seg) {
    case 4: PCP_LAUNCH(c, k_bounds<4>, dim3(grid), dim3(BLK), c->N, c->P, c->dPolyStart, pos, c->dBounds, c->dNum, c->dBcnt, c->dJcnt, c->dPackMaxKey, c->dScal); break;
    case 8: PCP_LAUNCH(c, k_bounds<8>, dim3(grid), dim3(BLK), c->N, c->P, c->dPolyStart, pos, c->dBounds, c->dNum, c->dBcnt, c->dJcnt, c->dPackMaxKey, c->dScal); break;
    case 16: PCP_LAUNCH(c, k_bounds<16>, dim3(grid), dim3(BLK), c->N, c->P, c->dPolyStart, pos, c->dBounds, c->dNum, c->dBcnt, c->dJcnt, c->dPackMaxKey, c->dScal); break;
    case 32: PCP_LAUNCH(c, k_bounds<32>, dim3(grid), dim3(BLK), c->N, c->P, c->dPolyStart, pos, c->dBounds, c->dNum, c->dBcnt, c->dJcnt, c->dPackMaxKey, c->dScal); break;
    default: PCP_LAUNCH(c, k_bounds<64>, dim3(grid), dim3(BLK), c->N, c->P, c->dPolyStart, pos, c->dBounds, c->dNum, c->dBcnt, c->dJcnt, c->dPackMaxKey, c->dScal); break;
    }
    return PCP_OK;
}

// cellsFromList: use the cells of the last run_cells (API updateNeighbors)
int run_detect(pcp_ctx* c, const double2* pos, int cellsFromList) {
    if (c->V == 0) return PCP_OK;
    PCP_TRY(launch_bounds(c, pos));
    PCP_LAUNCH(c, k_poly_cells, dim3(c->P), dim3(PP_BLK), c->dPackPoly, c->dBounds, c->dPcStart, c->dPcSorted,
               c->dPackB, c->dSBounds);
    int grid = (c->N + BLK / 64 - 1) / (BLK / 64);
    if (cellsFromList)
        PCP_LAUNCH(c, k_detect_poly<1>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack, c->dPackPoly,
                   c->dBox, pos, c->dCellOf, c->dBounds, c->dPcStart, c->dPcSorted, c->dSBounds, c->dPackB, c->Kp,
                   c->dBcnt,
                   c->dBKey, c->dBTU, c->dJcnt, c->dJKey, c->dPackMaxKey, c->dNum, c->dScal);
    else
        PCP_LAUNCH(c, k_detect_poly<0>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack, c->dPackPoly,
                   c->dBox, pos, c->dCellOf, c->dBounds, c->dPcStart, c->dPcSorted, c->dSBounds, c->dPackB, c->Kp,
                   c->dBcnt,
                   c->dBKey, c->dBTU, c->dJcnt, c->dJKey, c->dPackMaxKey, c->dNum, c->dScal);
    return PCP_OK;
}

int detect_alloc(pcp_ctx* c) {
    c->maxN = max_poly_vertices(c);
    c->seg = seg_width(c->maxN);
    PCP_TRY(dalloc(&c->dBounds, std::max(c->N, 1)));
    PCP_TRY(dalloc(&c->dSBounds, std::max(c->N, 1)));
    PCP_TRY(dalloc(&c->dPcSorted, std::max(c->N, 1)));
    {   // identity until the first detection sorts the polygons by cell (kernels that visit
        // polygons in cell order also run for models without overlap)
        std::vector<int> iota(std::max(c->N, 1));
        for (int s = 0; s < (int)iota.size(); ++s) iota[s] = s;
        PCP_HIP(hipMemcpy(c->dPcSorted, iota.data(), sizeof(int) * iota.size(), hipMemcpyHostToDevice));
    }
    PCP_TRY(dalloc(&c->dPcStart, (long long)c->P * (PC_CELLS + 1)));
    PCP_TRY(dalloc(&c->dPackB, std::max(c->P, 1)));
    return PCP_OK;
}

// Work of the last detection, recomputed on the host from the device's polygon cells
// with the kernel's candidate rule: tiles tested (a polygon with itself, and each pair of
// polygons once from each side) and the edge pairs they hold.
int detect_stats(pcp_ctx* c, long long* polygonPairs, long long* edgePairs) {
    PCP_HIP(hipStreamSynchronize(c->stream));
    std::vector<double4> bd(c->N);
    std::vector<int> sorted(c->N), packB(c->P), start((size_t)c->P * (PC_CELLS + 1));
    PCP_HIP(hipMemcpy(bd.data(), c->dBounds, sizeof(double4) * c->N, hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(sorted.data(), c->dPcSorted, sizeof(int) * c->N, hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(packB.data(), c->dPackB, sizeof(int) * c->P, hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(start.data(), c->dPcStart, sizeof(int) * start.size(), hipMemcpyDeviceToHost));
    auto wrapH = [](double x) { double y = x + 1.5; y = y - std::floor(y); return y - 0.5; };
    auto coord = [](double x, int B) { double f = std::floor(x * B); int i = (f >= 0.0) ? (int)f : 0; return (i >= B) ? B - 1 : i; };
    long long pp = 0, ep = 0;
    for (int p = 0; p < c->P; ++p) {
        int B = packB[p] & 0xFFFF;
        bool big = (packB[p] >> 30) & 1;
        int s0 = c->hPackPoly[p];
        const int* st = start.data() + (size_t)p * (PC_CELLS + 1);
        for (int A = s0; A < c->hPackPoly[p + 1]; ++A) {
            long long nA = c->hPolyStart[A + 1] - c->hPolyStart[A];
            if (nA >= 4) { ++pp; ep += nA * (nA - 3) / 2; }
            int ax = coord(bd[A].x, B), ay = coord(bd[A].y, B);
            int ncell = (B == 1) ? 1 : 9;
            for (int k = 0; k < ncell; ++k) {
                int cx = ax, cy = ay;
                if (B > 1) { cx = (ax + k / 3 - 1 + B) % B; cy = (ay + k % 3 - 1 + B) % B; }
                int cell = cy * B + cx;
                for (int q = st[cell]; q < st[cell + 1]; ++q) {
                    int Bq = sorted[s0 + q];
                    if (Bq == A) continue;
                    if (!big) {
                        double dx = wrapH(bd[Bq].x - bd[A].x), dy = wrapH(bd[Bq].y - bd[A].y);
                        double rr = (bd[A].z + bd[Bq].z) * (1.0 + 1e-6) + 1e-12;
                        if (!(dx * dx + dy * dy <= rr * rr)) continue;
                    }
                    ++pp;
                    ep += nA * (long long)(c->hPolyStart[Bq + 1] - c->hPolyStart[Bq]);
                }
            }
        }
    }
    *polygonPairs = pp;
    *edgePairs = ep;
    return PCP_OK;
}

}  // namespace pcp
