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
//                   their id is larger and their circle meets A's; for A itself and each
//                   kept polygon, its edges held in lanes and tested against A's.  An edge
//                   pair is tested when the midpoint cells of the two edges in the user's
//                   cell grid are neighbours (the reference's stencil), with the
//                   reference's arithmetic; a crossing is appended to the si bucket of the
//                   polygon its record belongs to and to the sj bucket of the other one
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
// polygon ids into pcSorted[packPoly[p]..], and packB[p] = B | big << 30.
__global__ __launch_bounds__(PP_BLK) void k_poly_cells(const int* __restrict__ packPoly,
                                                      const double4* __restrict__ bounds,
                                                      int* __restrict__ pcStart, int* __restrict__ pcSorted,
                                                      int* __restrict__ packB) {
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
    }
}

// ---------------------------------------------------------------- edge pairs
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bool cells_adjacent(int ca, int cb, int B) {
    int ax = ca % B, ay = ca / B, bx = cb % B, by = cb / B;
    int dx = abs(ax - bx), dy = abs(ay - by);
    return (dx <= 1 || dx == B - 1) && (dy <= 1 || dy == B - 1);
}

struct EdgeStage {
    double2* P;
    double2* R;
    int* C;
};

// Test the edges of polygon A (staged in LDS, chunk a0 of at most 64 edges) against
// every edge of polygon Bq (held in lanes).  A <= Bq; A == Bq tests each pair of
// non-adjacent edges once.
template <int CELLS>
__device__ void edge_tile(int lane, int A, int stA, int nA, int Bq, int sAl, int sBl, int Bu, int pbase,
                          const int* __restrict__ polyStart, const double2* __restrict__ pos,
                          const int* __restrict__ cellOf, EdgeStage sg, int& staged, int Kp,
                          int* __restrict__ bcnt, uint64_t* __restrict__ bKey, float2* __restrict__ bTU,
                          int* __restrict__ jcnt, uint64_t* __restrict__ jKey, int* __restrict__ num,
                          int* __restrict__ scal, uint64_t& kmax) {
    const double eps = 1e-12;
    bool self = (A == Bq);
    int stB = polyStart[Bq], nB = polyStart[Bq + 1] - stB;
    for (int a0 = 0; a0 < nA; a0 += 64) {
        int na = min(64, nA - a0);
        if (staged != a0) {
            wave_lds_sync();
            if (lane < na) {
                int a = a0 + lane;
                d2 p0 = ldpos(pos, stA + a), p1 = ldpos(pos, stA + ((a + 1 == nA) ? 0 : a + 1));
                sg.P[lane] = make_double2(p0.x, p0.y);
                sg.R[lane] = make_double2(wrap(p1.x - p0.x), wrap(p1.y - p0.y));
                sg.C[lane] = CELLS ? cellOf[stA + a] : edge_cell(p0, p1, Bu);
            }
            wave_lds_sync();
            staged = a0;
        }
        for (int b0 = 0; b0 < nB; b0 += 64) {
            int nb = min(64, nB - b0);
            int G = 64 / nb;
            int g = lane / nb, bl = lane - g * nb;
            if (g >= G) continue;
            int b = b0 + bl;
            d2 q0 = ldpos(pos, stB + b), q1 = ldpos(pos, stB + ((b + 1 == nB) ? 0 : b + 1));
            d2 Q = q0, S = mk(wrap(q1.x - q0.x), wrap(q1.y - q0.y));
            int cb = CELLS ? cellOf[stB + b] : edge_cell(q0, q1, Bu);
            for (int al = g; al < na; al += G) {
                int a = a0 + al;
                if (self) {
                    if (b <= a) continue;
                    if (b == a + 1 || (a == 0 && b == nA - 1)) continue;   // shared vertex
                }
                if (!cells_adjacent(sg.C[al], cb, Bu)) continue;
                double2 P1 = sg.P[al], R1 = sg.R[al];
                double gx = wrap(Q.x - P1.x);
                double gy = wrap(Q.y - P1.y);
                double denom = R1.x * S.y - R1.y * S.x;
                if (fabs(denom) < eps) continue;
                double tt = (gx * S.y - gy * S.x) / denom;
                double uu = (gx * R1.y - gy * R1.x) / denom;
                if (tt > 0.0 && tt < 1.0 && uu > 0.0 && uu < 1.0) {
                    uint64_t key;
                    float2 tu;
                    // updateCompactedIntersectionsKernel kernels.cuh:857-876
                    if (denom > 0) {
                        key = ((uint64_t)(uint16_t)sAl << 48) | ((uint64_t)(uint16_t)sBl << 32) |
                              ((uint64_t)(uint16_t)b << 16) | (uint64_t)(uint16_t)a;
                        tu = make_float2((float)tt, (float)uu);
                    } else {
                        key = ((uint64_t)(uint16_t)sBl << 48) | ((uint64_t)(uint16_t)sAl << 32) |
                              ((uint64_t)(uint16_t)a << 16) | (uint64_t)(uint16_t)b;
                        tu = make_float2((float)uu, (float)tt);
                    }
                    kmax = (key > kmax) ? key : kmax;
                    if (CELLS) atomicAdd(&num[stA + a], 1);   // per-edge count (updateNeighbors)
                    int sig = pbase + key_si(key), sjg = pbase + key_sj(key);
                    int b1 = atomicAdd(&bcnt[sig], 1);
                    if (b1 < Kp) {
                        bKey[(long long)sig * Kp + b1] = key;
                        bTU[(long long)sig * Kp + b1] = tu;
                    } else {
                        atomicMax(&scal[SC_BUCKET], b1 + 1);
                    }
                    int b2 = atomicAdd(&jcnt[sjg], 1);
                    if (b2 < Kp) jKey[(long long)sjg * Kp + b2] = key;
                    else atomicMax(&scal[SC_BUCKET], b2 + 1);
                }
            }
        }
    }
}

// One wavefront per polygon A: the polygons of the 3x3 block of cells around A's cell
// are loaded by the lanes in parallel and kept when their id is larger than A's and
// their circle meets A's; A's edges are staged in LDS once and tested against A itself
// and against every kept polygon.
// CELLS = 0: the user-grid cell of every edge is computed from the live positions (the
// fused evaluation, where cells and crossings come from the same positions);
// CELLS = 1: the cells of the last cell-list build are used (updateNeighbors after
// updateNeighborCells, which may have been called at other positions).
template <int CELLS>
__global__ __launch_bounds__(BLK) void k_detect_poly(
    int N, const int* __restrict__ polyStart, const int* __restrict__ polyPack,
    const int* __restrict__ packPoly, const int* __restrict__ box, const double2* __restrict__ pos,
    const int* __restrict__ cellOf, const double4* __restrict__ bounds, const int* __restrict__ pcStart,
    const int* __restrict__ pcSorted, const int* __restrict__ packB, int Kp, int* __restrict__ bcnt,
    uint64_t* __restrict__ bKey, float2* __restrict__ bTU, int* __restrict__ jcnt, uint64_t* __restrict__ jKey,
    unsigned long long* __restrict__ packMaxKey, int* __restrict__ num, int* __restrict__ scal) {
    __shared__ double2 sP[BLK / 64][64];
    __shared__ double2 sR[BLK / 64][64];
    __shared__ int sC[BLK / 64][64];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int A = blockIdx.x * (BLK / 64) + w;
    if (A >= N) return;   // whole wavefront
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
    if (nA >= 4)
        edge_tile<CELLS>(lane, A, stA, nA, A, sAl, sAl, Bu, s0, polyStart, pos, cellOf, sg, staged, Kp, bcnt,
                         bKey, bTU, jcnt, jKey, num, scal, kmax);
    // candidate cells: lanes 0..8 hold the ranges of the 3x3 block (one cell when Bp == 1)
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
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    int T = __shfl(incl, ncell - 1);
    for (int c0 = 0; c0 < T; c0 += 64) {
        int idx = c0 + lane;
        int q = 0;
        bool found = false;
        for (int k = 0; k < ncell; ++k) {   // every lane takes part in the shuffles
            int hiK = __shfl(incl, k), cK = __shfl(cnt, k), loK = __shfl(lo, k);
            if (!found && idx < hiK) { q = loK + idx - (hiK - cK); found = true; }
        }
        int Bq = -1;
        if (idx < T) {
            int cand = pcSorted[s0 + q];
            if (cand > A && (big || circles_meet(bA, bounds[cand]))) Bq = cand;
        }
        unsigned long long keep = __ballot(Bq >= 0);
        while (keep) {
            int j = __ffsll((long long)keep) - 1;
            keep &= keep - 1;
            int Bsel = __shfl(Bq, j);
            edge_tile<CELLS>(lane, A, stA, nA, Bsel, sAl, Bsel - s0, Bu, s0, polyStart, pos, cellOf, sg, staged,
                             Kp, bcnt, bKey, bTU, jcnt, jKey, num, scal, kmax);
        }
    }
    // largest key of the packing (the pairing's last-record case): one atomic per
    // wavefront, only when it raises the current value
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        unsigned long long y = __shfl_xor((unsigned long long)kmax, o);
        kmax = (y > kmax) ? y : kmax;
    }
    if (lane == 0 && kmax > 0 && kmax > packMaxKey[pk]) atomicMax(&packMaxKey[pk], (unsigned long long)kmax);
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
               c->dPackB);
    int grid = (c->N + BLK / 64 - 1) / (BLK / 64);
    if (cellsFromList)
        PCP_LAUNCH(c, k_detect_poly<1>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack, c->dPackPoly,
                   c->dBox, pos, c->dCellOf, c->dBounds, c->dPcStart, c->dPcSorted, c->dPackB, c->Kp, c->dBcnt,
                   c->dBKey, c->dBTU, c->dJcnt, c->dJKey, c->dPackMaxKey, c->dNum, c->dScal);
    else
        PCP_LAUNCH(c, k_detect_poly<0>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack, c->dPackPoly,
                   c->dBox, pos, c->dCellOf, c->dBounds, c->dPcStart, c->dPcSorted, c->dPackB, c->Kp, c->dBcnt,
                   c->dBKey, c->dBTU, c->dJcnt, c->dJKey, c->dPackMaxKey, c->dNum, c->dScal);
    return PCP_OK;
}

int detect_alloc(pcp_ctx* c) {
    c->maxN = max_poly_vertices(c);
    c->seg = seg_width(c->maxN);
    PCP_TRY(dalloc(&c->dBounds, std::max(c->N, 1)));
    PCP_TRY(dalloc(&c->dPcSorted, std::max(c->N, 1)));
    PCP_TRY(dalloc(&c->dPcStart, (long long)c->P * (PC_CELLS + 1)));
    PCP_TRY(dalloc(&c->dPackB, std::max(c->P, 1)));
    return PCP_OK;
}

// Work of the last detection, recomputed on the host from the device's polygon cells
// with the kernel's candidate rule: polygon pairs tested and the edge pairs they hold.
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
                    if (Bq <= A) continue;
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
