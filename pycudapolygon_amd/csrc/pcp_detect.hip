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
//
// With POST, FIRE's step after SHAKE (model.cpp:431-437) is done here first, where the
// detection follows SHAKE (area energies): a packing whose SHAKE failed (first call at
// shakeIter iterations and the retry at 10 * shakeIter) takes back its step-start
// positions and zero velocities, the others re-derive their velocities from the SHAKE
// displacement, v += wrap(x - xpre) / dt; the packing's first vertex records the outcome
// in flags.  The bounds are those of the resulting positions (read from xbak after a
// failure, so nothing written here is read back).
struct PostShake {
    int shakeIter, fnf, fFail, fLast;   // flags of packing p at fnf * p + fFail / fLast
    const int* polyPack;
    const int* packVert;
    const int* active;
    const double* dt;
    const int* shakeIt;                 // [P] first call, [P + p] retry iterations
    double2* x;
    double2* v;
    const double2* xbak;
    const double2* xpre;
    int* flags;
};
template <int SEG, bool POST = false>
__global__ __launch_bounds__(BLK) void k_bounds(int N, int P, const int* __restrict__ polyStart,
                                               const double2* pos, double4* __restrict__ bounds,
                                               int* __restrict__ num, int* __restrict__ bcnt, int* __restrict__ jcnt,
                                               unsigned long long* __restrict__ packMaxKey, long long pmkPitch,
                                               int* __restrict__ scal, PostShake ps) {
    if (blockIdx.x == 0 && threadIdx.x < SC_NSCAL) scal[threadIdx.x] = 0;
    for (int q = blockIdx.x * BLK + threadIdx.x; q < PMK_SLOTS * P; q += gridDim.x * BLK)
        packMaxKey[(q / P) * pmkPitch + q % P] = 0ull;
    int sub = threadIdx.x & (SEG - 1);
    int s = (blockIdx.x * BLK + threadIdx.x) / SEG;
    int st = 0, n = 0;
    if (s < N) {
        st = polyStart[s];
        n = polyStart[s + 1] - st;
        if (sub == 0) { bcnt[s] = 0; jcnt[s] = 0; }
    }
    if constexpr (POST) {
        if (s < N) {
            int p = ps.polyPack[s];
            if (ps.active[p]) {
                int it1 = ps.shakeIt[p], it2 = ps.shakeIt[P + p];
                bool retried = it1 >= ps.shakeIter;
                bool fail = retried && it2 >= ps.shakeIter * 10;
                double h = ps.dt[p];
                for (int k = sub; k < n; k += SEG) {
                    int i = st + k;
                    if (fail) {
                        ps.x[i] = ps.xbak[i];
                        ps.v[i] = make_double2(0.0, 0.0);
                    } else {
                        double2 xi = pos[i], xp = ps.xpre[i], vi = ps.v[i];
                        ps.v[i] = make_double2(vi.x + wrap(xi.x - xp.x) / h, vi.y + wrap(xi.y - xp.y) / h);
                    }
                }
                if (fail) pos = ps.xbak;
                if (sub == 0 && st == ps.packVert[p]) {
                    ps.flags[ps.fnf * p + ps.fFail] = fail ? 1 : 0;
                    ps.flags[ps.fnf * p + ps.fLast] = retried ? it2 : it1;
                }
            }
        }
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
// polygon ids into pcSorted[packPoly[p]..] and packB[p] = B | big << 30 (the detection
// reads a candidate's circle from bounds[id], in flight together with its vertex range).
// The bounds of up to PC_REG * PP_BLK polygons are read once, all loads in flight
// together, and kept in registers for the three passes; beyond that they are re-read.
// active (FIRE steps that skip converged packings; else NULL): the ids of the running packings'
// polygons also go to actOrder, packing after packing in cell order, and their number to
// *actCount; a converged packing's workgroup leaves its cell lists as they are.
constexpr int PC_REG = 8;
__global__ __launch_bounds__(PP_BLK) void k_poly_cells(const int* __restrict__ packPoly,
                                                      const double4* __restrict__ bounds,
                                                      int* __restrict__ pcStart, int* __restrict__ pcSorted,
                                                      int* __restrict__ packB, const int* __restrict__ active,
                                                      int* __restrict__ actOrder, int* __restrict__ actCount) {
    __shared__ int cStart[PC_CELLS + 1];
    __shared__ int cCur[PC_CELLS];
    __shared__ double redR[PP_BLK / 64];
    __shared__ int redBig[PP_BLK / 64];
    __shared__ int scanTmp[PP_BLK / 64];
    int p = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
    int s0 = packPoly[p], s1 = packPoly[p + 1], Np = s1 - s0;
    long long gbase = (long long)p * (PC_CELLS + 1);
    int off = 0;   // the packing's first place in actOrder
    if (active) {
        int below = 0, all = 0;
        for (int q = t; q < (int)gridDim.x; q += PP_BLK) {   // one workgroup per packing
            int sz = active[q] ? packPoly[q + 1] - packPoly[q] : 0;
            all += sz;
            below += (q < p) ? sz : 0;
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) { below += __shfl_xor(below, o); all += __shfl_xor(all, o); }
        if (lane == 0) { scanTmp[w] = below; redBig[w] = all; }
        __syncthreads();
        below = 0; all = 0;
        for (int q = 0; q < PP_BLK / 64; ++q) { below += scanTmp[q]; all += redBig[q]; }
        __syncthreads();   // scanTmp and redBig are reused below
        if (p == 0 && t == 0) *actCount = all;
        if (!active[p]) return;   // the whole workgroup
        off = below;
    }
    double4 bc[PC_REG];
#pragma unroll
    for (int k = 0; k < PC_REG; ++k) {
        int s = s0 + t + k * PP_BLK;
        bc[k] = (s < s1) ? bounds[s] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    double r = 0.0;
    int big = 0;
#pragma unroll
    for (int k = 0; k < PC_REG; ++k) {
        r = fmax(r, bc[k].z);
        big |= (bc[k].w != 0.0);
    }
    for (int s = s0 + t + PC_REG * PP_BLK; s < s1; s += PP_BLK) {
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
    int cellOf[PC_REG];
#pragma unroll
    for (int k = 0; k < PC_REG; ++k) {
        int s = s0 + t + k * PP_BLK;
        cellOf[k] = pc_coord(bc[k].y, B) * B + pc_coord(bc[k].x, B);
        if (s < s1) atomicAdd(&cCur[cellOf[k]], 1);
    }
    for (int s = s0 + t + PC_REG * PP_BLK; s < s1; s += PP_BLK) {
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
    // the ids in cell order: one workgroup's scattered stores are limited to about one
    // cache line per clock, so up to PC_CELLS + 1 polygons the ids are placed in LDS (over
    // cStart, which has been read out above) and written out in order
    bool inLds = Np <= PC_CELLS + 1;
    if (inLds) __syncthreads();
#pragma unroll
    for (int k = 0; k < PC_REG; ++k) {
        int s = s0 + t + k * PP_BLK;
        if (s < s1) {
            int slot = atomicAdd(&cCur[cellOf[k]], 1);
            if (inLds) {
                cStart[slot] = s;
            } else {
                pcSorted[s0 + slot] = s;
                if (active) actOrder[off + slot] = s;
            }
        }
    }
    for (int s = s0 + t + PC_REG * PP_BLK; s < s1; s += PP_BLK) {
        double4 b = bounds[s];
        int slot = atomicAdd(&cCur[pc_coord(b.y, B) * B + pc_coord(b.x, B)], 1);
        if (inLds) {
            cStart[slot] = s;
        } else {
            pcSorted[s0 + slot] = s;
            if (active) actOrder[off + slot] = s;
        }
    }
    if (inLds) {
        __syncthreads();
        for (int i = t; i < Np; i += PP_BLK) {
            pcSorted[s0 + i] = cStart[i];
            if (active) actOrder[off + i] = cStart[i];
        }
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

// The candidates of the 3x3 block of polygon cells as one list: lane k < ncell holds the
// range (lo, cnt) of cell k (0 elsewhere).  The ranges are read lane by lane as wave-uniform
// values (no LDS round trip per cell); candidate ci of the list is slot lo_k + ci - start_k of
// the sorted ids, k the first cell whose running count exceeds ci (start_k the count of the
// cells before k).
__device__ __forceinline__ int cell_list_length(int cnt) {
    int run = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) run += __builtin_amdgcn_readlane(cnt, k);
    return run;
}
// slot of candidate ci (< the list's length) in the sorted ids
__device__ __forceinline__ int cell_slot(int lo, int cnt, int ci) {
    int run = 0, off = 0;
    bool found = false;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        int l = __builtin_amdgcn_readlane(lo, k), c = __builtin_amdgcn_readlane(cnt, k);
        bool here = !found && ci < run + c;
        off = here ? l - run : off;
        found = found || here;
        run += c;
    }
    return ci + off;
}

// cell of an edge in the user grid as x | y << 16 (box sizes are below 2^16: at most 2e9 cells)
__device__ __forceinline__ int cell_xy(int flat, int B) { return (flat % B) | ((flat / B) << 16); }
__device__ __forceinline__ bool cells_adjacent(int ca, int cb, int B) {
    int dx = abs((ca & 0xFFFF) - (cb & 0xFFFF)), dy = abs((ca >> 16) - (cb >> 16));
    return (dx <= 1 || dx == B - 1) && (dy <= 1 || dy == B - 1);
}

// bits [0, n) and [lo, hi) of a 64-bit mask (bounds clamped to 0..64)
__device__ __forceinline__ unsigned long long bits_below(int n) {
    return n <= 0 ? 0ull : (n >= 64 ? ~0ull : ((1ull << n) - 1ull));
}
__device__ __forceinline__ unsigned long long bit_range(int lo, int hi) { return bits_below(hi) & ~bits_below(lo); }

// floor(x / d) for x >= 0, d >= 1 and x * d < 2^16 with a multiply and shift:
// m = ceil(2^16 / d) exceeds 2^16 / d by less than 1/d, which keeps x * m / 2^16 below
// the next integer while x * d < 2^16
__device__ __forceinline__ int small_div(int x, int d) {
    int m = (65536 + d - 1) / d;
    return (x * m) >> 16;
}

// Where the detection writes: per-polygon buckets of Kp records, the si bucket (keys and
// float parameters) and the sj bucket (keys), with their counters
struct DetOut {
    int Kp;
    int* bcnt;
    uint64_t* bKey;
    float2* bTU;
    int* jcnt;
    uint64_t* jKey;
    int* scal;
};

// A crossing record appended to the buckets of its two polygons (global ids s0 + si and
// s0 + sj; both the same for a self-crossing).  Slots come from atomic counters, since a
// bucket receives records from the wavefronts of several polygons; a slot beyond the
// capacity is not written and raises the overflow maximum, after which the caller grows
// the buckets and detects again.  The order inside a bucket is the order of arrival;
// every consumer orders the records by key.
// Both slots are taken before anything is stored: the stores would otherwise sit between
// the two atomics and the wait for the second one (the vector-memory counter also counts
// stores) would include their completion.
__device__ __forceinline__ void emit_record(uint64_t key, float2 tu, int s0, const DetOut& o) {
    int si = s0 + key_si(key), sj = s0 + key_sj(key);
    int a = atomicAdd(&o.bcnt[si], 1);
    int b = atomicAdd(&o.jcnt[sj], 1);
    int over = max(a, b);   // needs both slots: one wait, before the stores
    if (over >= o.Kp) atomicMax(&o.scal[SC_BUCKET], over + 1);
    if (a < o.Kp) {
        o.bKey[(long long)si * o.Kp + a] = key;
        o.bTU[(long long)si * o.Kp + a] = tu;
    }
    if (b < o.Kp) o.jKey[(long long)sj * o.Kp + b] = key;
}

// Segment test of updateNeighborsKernel (kernels.cuh:760-790) for edge 1 = (P1, P1 + R1)
// and edge 2 = (P2, P2 + R2): denom = R1 x R2, g = wrap(P2 - P1), t = (g x R2) / denom,
// u = (g x R1) / denom; a crossing when |denom| >= 1e-12 and t, u lie in (0, 1).
struct SegNum {
    double denom, nt, nu;
};
__device__ __forceinline__ SegNum seg_numerators(d2 P1, d2 R1, d2 P2, d2 R2) {
    SegNum s;
    s.denom = R1.x * R2.y - R1.y * R2.x;
    double gx = wrap(P2.x - P1.x);
    double gy = wrap(P2.y - P1.y);
    s.nt = gx * R2.y - gy * R2.x;
    s.nu = gx * R1.y - gy * R1.x;
    return s;
}
// Necessary condition for a crossing without the divisions: t = nt/denom and u = nu/denom
// lie in (0,1) only if nt and nu lie strictly between 0 and denom (a correctly rounded
// quotient keeps its sign and is >= 1 when |numerator| >= |denom|).  The sign of denom is
// moved onto the numerators by flipping their sign bits, which is exact, so one set of
// comparisons against |denom| covers both signs (NaNs fail every comparison, as in the
// two-sided form).
__device__ __forceinline__ bool seg_may_cross(const SegNum& s) {
    unsigned long long sg = (unsigned long long)__double_as_longlong(s.denom) & 0x8000000000000000ull;
    double at = __longlong_as_double((long long)((unsigned long long)__double_as_longlong(s.nt) ^ sg));
    double au = __longlong_as_double((long long)((unsigned long long)__double_as_longlong(s.nu) ^ sg));
    double ad = fabs(s.denom);
    return !(ad < 1e-12) && at > 0.0 && at < ad && au > 0.0 && au < ad;
}
// The exact test and the record of a pair that passed seg_may_cross: edge 1 is vertex n1
// of local polygon s1l, edge 2 vertex l2 of local polygon s2l.  The key and the inside
// order follow updateCompactedIntersectionsKernel (kernels.cuh:857-876); t and u are
// rounded to float as the reference stores them.
__device__ __forceinline__ bool exact_record(const SegNum& s, int s1l, int n1, int s2l, int l2, uint64_t& key,
                                             float2& tu) {
    double tt = s.nt / s.denom;
    double uu = s.nu / s.denom;
    if (!(tt > 0.0 && tt < 1.0 && uu > 0.0 && uu < 1.0)) return false;
    if (s.denom > 0) {
        key = ((uint64_t)(uint16_t)s1l << 48) | ((uint64_t)(uint16_t)s2l << 32) | ((uint64_t)(uint16_t)l2 << 16) |
              (uint64_t)(uint16_t)n1;
        tu = make_float2((float)tt, (float)uu);
    } else {
        key = ((uint64_t)(uint16_t)s2l << 48) | ((uint64_t)(uint16_t)s1l << 32) | ((uint64_t)(uint16_t)n1 << 16) |
              (uint64_t)(uint16_t)l2;
        tu = make_float2((float)uu, (float)tt);
    }
    return true;
}

// The side-consistent decision (PCP_CONSISTENT; no counterpart in the reference): the side of
// point X relative to the line of an edge (start E, wrapped edge vector ER) is the sign of
// ER x wrap(X - E), a function of that edge's two vertices and X only.  Edge 1 (P1, R1, next
// vertex P1n) and edge 2 (P2, R2, P2n) cross when each edge's endpoints lie on different sides
// of the other's line (a zero cross product counts with the positive side) and
// |denom| >= 1e-12 as in the reference.  The two edges at a vertex then agree about that
// vertex, and two polygons cross an even number of times, where the parameter test can count
// one of the two crossings near a vertex and not the other.  On configurations without such
// near-ties it decides exactly as the parameter test (t, u are still the reference's).
__device__ __forceinline__ bool side_neg(d2 E, d2 ER, d2 X) {
    double wx = wrap(X.x - E.x), wy = wrap(X.y - E.y);
    return (ER.x * wy - ER.y * wx) < 0.0;
}
__device__ __forceinline__ bool seg_cons_cross(const SegNum& s, d2 P1, d2 R1, d2 P1n, d2 P2, d2 R2, d2 P2n) {
    return !(fabs(s.denom) < 1e-12) && side_neg(P2, R2, P1) != side_neg(P2, R2, P1n) &&
           side_neg(P1, R1, P2) != side_neg(P1, R1, P2n);
}
// the record of a pair the side-consistent decision took (no parameter range test)
__device__ __forceinline__ void cons_record(const SegNum& s, int s1l, int n1, int s2l, int l2, uint64_t& key,
                                           float2& tu) {
    double tt = s.nt / s.denom;
    double uu = s.nu / s.denom;
    if (s.denom > 0) {
        key = ((uint64_t)(uint16_t)s1l << 48) | ((uint64_t)(uint16_t)s2l << 32) | ((uint64_t)(uint16_t)l2 << 16) |
              (uint64_t)(uint16_t)n1;
        tu = make_float2((float)tt, (float)uu);
    } else {
        key = ((uint64_t)(uint16_t)s2l << 48) | ((uint64_t)(uint16_t)s1l << 32) | ((uint64_t)(uint16_t)n1 << 16) |
              (uint64_t)(uint16_t)l2;
        tu = make_float2((float)uu, (float)tt);
    }
}

// An edge (start P, vector R) that crosses an edge of a polygon whose vertices lie in the
// circle (c, r) meets that circle (the crossing point lies on both edges, and the circle
// contains the polygon's edges).  Conservative test with the edge's own bounding circle:
// centre P + R/2, radius (|Rx| + |Ry|) / 2 >= |R| / 2, minimum image of the centres, a
// relative margin for rounding.  Valid while both polygons are small against the box
// (radii below CULL_RADIUS): then the minimum images of every vertex, centre and edge
// involved are consistent with the polygons' own chains.
constexpr double CULL_RADIUS = 0.1;
__device__ __forceinline__ bool edge_near_circle(d2 P, d2 R, double cx, double cy, double r) {
    double dx = wrap(cx - (P.x + 0.5 * R.x));
    double dy = wrap(cy - (P.y + 0.5 * R.y));
    double rr = (r + 0.5 * (fabs(R.x) + fabs(R.y))) * (1.0 + 1e-6) + 1e-12;
    return dx * dx + dy * dy <= rr * rr;
}
__device__ __forceinline__ bool cullable(double4 b) { return b.w == 0.0 && b.z < CULL_RADIUS; }

constexpr int TILE_UNROLL = 4;
// vertices and polygons of one batch of candidates in a wavefront's LDS
constexpr int DET_BATCH = 256;
constexpr int DET_KMAX = 16;
// pairs that pass the prefilter are collected (up to DET_MAY) and tested exactly 64 at a
// time, so a polygon pays the exact test's latency (divisions, the record's atomics) once
// or twice rather than once per pass that found any; pair ids of a batch are below
// 64 * DET_BATCH = 2^14
constexpr int DET_MAY = 128;

// A wavefront's LDS: A's edges (chunk of 64), the vertices of a batch of candidate
// polygons with the lane that loaded each, and the culled edge lists of the batch
struct DetLds {
    double2 P[64], R[64];         // A's edges: start and wrapped edge vector
    int C[64];                    // their user-grid cells (x | y << 16)
    double2 V[DET_BATCH];         // candidate vertices, polygon after polygon
    uint8_t own[DET_BATCH];       // lane of the candidate a vertex belongs to
    uint8_t listB[DET_BATCH];     // candidate edges near A's circle (local ids), per candidate
    int cellB[DET_BATCH];         // user-grid cell (x | y << 16) of each candidate edge
    uint8_t listA[DET_KMAX * 64]; // A's edges near each candidate's circle
    uint8_t klane[DET_KMAX];      // lane of the batch's k-th candidate
    int4 kinfo[DET_KMAX];         // per candidate: first pair, near-edge count, first vertex in V, vertices
    int2 kpoly[DET_KMAX];         // per candidate: polygon id, first vertex in pos
    uint16_t may[DET_MAY];        // batch pairs that passed the prefilter, awaiting the exact test
};

// Where the lanes take the vertices of polygon B from in edge_tile: global memory, or (A
// against itself, one chunk) A's own LDS stage
enum { SRC_GLOBAL = 0, SRC_STAGE = 2 };
// Tile roles: A against itself, or A against a polygon of larger id (A's edges are then
// the reference's edge 1: updateNeighborsKernel tests edge id1 against larger ids)
enum { TILE_SELF = 0, TILE_A_FIRST = 1 };

template <class LdsT>
__device__ __forceinline__ void stage_a(LdsT& L, int lane, int stA, int nA, int a0, int Bu,
                                        const double2* __restrict__ pos, const int* __restrict__ cellOf, bool cells) {
    int na = min(64, nA - a0);
    wave_lds_sync();
    if (lane < na) {
        int a = a0 + lane;
        d2 p0 = ldpos(pos, stA + a), p1 = ldpos(pos, stA + ((a + 1 == nA) ? 0 : a + 1));
        L.P[lane] = make_double2(p0.x, p0.y);
        L.R[lane] = make_double2(wrap(p1.x - p0.x), wrap(p1.y - p0.y));
        L.C[lane] = cells ? cell_xy(cellOf[stA + a], Bu) : edge_cell_xy(p0, p1, Bu);
    }
    wave_lds_sync();
}

// Test every edge of polygon A (staged in LDS, chunks of 64) against every edge of
// polygon B (held in lanes; ids wavefront-uniform; vertices from global memory at
// pos + stB, or for TILE_SELF with nA <= 64 from A's stage).  TILE_SELF tests each pair of
// non-adjacent edges of A once.  The hot loop evaluates the division-free necessary
// condition; iterations where some lane passes it take the exact test, the user-grid cell
// filter and the record.  Used for A itself and for polygons of more than 64 vertices.
// The start of the edge after edge a0 + al of polygon A (its next vertex), from A's stage
// when the chunk [a0, a0 + na) holds it, else from global memory
__device__ __forceinline__ d2 edge_next_start(const DetLds& L, const double2* __restrict__ pos, int stA, int nA,
                                              int a0, int na, int al) {
    int a = a0 + al, an = (a + 1 == nA) ? 0 : a + 1;
    if (an >= a0 && an < a0 + na) {
        double2 p = L.P[an - a0];
        return mk(p.x, p.y);
    }
    return ldpos(pos, stA + an);
}

template <int CELLS, int ROLE, int SRC, bool CONS>
__device__ void edge_tile(DetLds& L, int lane, int stA, int nA, int stB, int nB, int sAl, int sBl, int Bu, int s0,
                          const double2* __restrict__ pos, const int* __restrict__ cellOf, int& staged,
                          const DetOut& out, int* __restrict__ num, uint64_t& kmax) {
    for (int a0 = 0; a0 < nA; a0 += 64) {
        int na = min(64, nA - a0);
        if (staged != a0) {
            stage_a(L, lane, stA, nA, a0, Bu, pos, cellOf, CELLS != 0);
            staged = a0;
        }
        for (int b0 = 0; b0 < nB; b0 += 64) {
            int nb = min(64, nB - b0);
            int G = 64 / nb;
            int g = small_div(lane, nb), bl = lane - g * nb;
            bool act = g < G;
            int b = b0 + bl;
            int bn = (b + 1 == nB) ? 0 : b + 1;
            d2 Q = mk(0.0, 0.0), S = mk(0.0, 0.0), Qn = mk(0.0, 0.0);
            if (act) {
                if (SRC == SRC_STAGE) {   // A itself, all of it in the stage (nA <= 64)
                    double2 q = L.P[b], r = L.R[b];
                    Q = mk(q.x, q.y);
                    S = mk(r.x, r.y);
                    if (CONS) {
                        double2 qn = L.P[bn];
                        Qn = mk(qn.x, qn.y);
                    }
                } else {
                    d2 q0 = ldpos(pos, stB + b), q1 = ldpos(pos, stB + bn);
                    Q = q0;
                    S = mk(wrap(q1.x - q0.x), wrap(q1.y - q0.y));
                    Qn = q1;
                }
            }
            int iters = (na + TILE_UNROLL * G - 1) / (TILE_UNROLL * G);
            for (int it = 0; it < iters; ++it) {
                bool may[TILE_UNROLL];
                bool anyMay = false;
#pragma unroll
                for (int u = 0; u < TILE_UNROLL; ++u) {
                    int al = g + (it * TILE_UNROLL + u) * G;
                    int a = a0 + al;
                    bool ok = act && al < na;
                    if (ROLE == TILE_SELF) ok = ok && b > a && b != a + 1 && !(a == 0 && b == nA - 1);
                    double2 pa = L.P[ok ? al : 0], ra = L.R[ok ? al : 0];
                    SegNum s = seg_numerators(mk(pa.x, pa.y), mk(ra.x, ra.y), Q, S);
                    if (CONS) {
                        d2 pn = ok ? edge_next_start(L, pos, stA, nA, a0, na, al) : mk(0.0, 0.0);
                        may[u] = ok && seg_cons_cross(s, mk(pa.x, pa.y), mk(ra.x, ra.y), pn, Q, S, Qn);
                    } else {
                        may[u] = ok && seg_may_cross(s);
                    }
                    anyMay = anyMay || may[u];
                }
                if (__ballot(anyMay) == 0ull) continue;   // wavefront-uniform: no crossing here
                int cb = 0;   // the user-grid cell of B's edge, only needed here
                if (act) {
                    if (SRC == SRC_STAGE) cb = L.C[b];
                    else if (CELLS) cb = cell_xy(cellOf[stB + b], Bu);
                    else cb = edge_cell_xy(ldpos(pos, stB + b), ldpos(pos, stB + bn), Bu);
                }
#pragma unroll
                for (int u = 0; u < TILE_UNROLL; ++u) {
                    int al = g + (it * TILE_UNROLL + u) * G;
                    int a = a0 + al;
                    uint64_t key = 0;
                    float2 tu;
                    bool rec = false;
                    if (may[u] && cells_adjacent(L.C[al], cb, Bu)) {
                        double2 pa = L.P[al], ra = L.R[al];
                        SegNum s = seg_numerators(mk(pa.x, pa.y), mk(ra.x, ra.y), Q, S);
                        if (CONS) {
                            cons_record(s, sAl, a, sBl, b, key, tu);
                            rec = true;
                        } else {
                            rec = exact_record(s, sAl, a, sBl, b, key, tu);
                        }
                    }
                    if (rec) {
                        kmax = (key > kmax) ? key : kmax;
                        if (CELLS) atomicAdd(&num[stA + a], 1);   // per-edge count of edge 1
                        emit_record(key, tu, s0, out);
                    }
                }
            }
        }
    }
}

// Crossings of polygon A with itself and with every polygon of larger id in the 3x3 block
// of polygon cells around it whose circle meets A's (each pair of polygons is tested once,
// from its lower polygon, whose edges are then the reference's edge 1).
//   1. A's edges staged in LDS; A against itself (edge_tile).
//   2. Candidates gathered by the lanes in parallel, in chunks of 64.
//   3. Candidates of at most 64 vertices (A too) in batches of at most DET_KMAX polygons
//      and DET_BATCH vertices: all their vertices copied to LDS with one memory latency;
//      the candidate edges near A's circle and A's edges near each candidate's circle
//      listed (the cull, skipped for large polygons); the products of the lists, i.e. all
//      edge pairs that can cross, enumerated over the batch and tested 64 per pass, every
//      lane busy whatever the polygons' sizes.
//   4. Other candidates: one edge_tile each.
template <int CELLS, bool CONS>
__device__ void detect_polygon(
    int A, int lane, DetLds& L, const int* __restrict__ polyStart, const int* __restrict__ polyPack,
    const int* __restrict__ packPoly, const int* __restrict__ box, const double2* __restrict__ pos,
    const int* __restrict__ cellOf, const double4* __restrict__ bounds, const int* __restrict__ pcStart,
    const int* __restrict__ pcSorted, const int* __restrict__ packB,
    const DetOut& out, unsigned long long* __restrict__ packMaxKey, long long pmkPitch, int* __restrict__ num) {
    int stA = polyStart[A], nA = polyStart[A + 1] - stA;
    int pk = polyPack[A];
    int Bu = box[pk];
    int s0 = packPoly[pk];
    int pb = packB[pk];
    int Bp = pb & 0xFFFF;
    bool big = (pb >> 30) & 1;
    int sAl = A - s0;
    double4 bA = bounds[A];
    // the packing's largest key so far (read here, with the other loads of the polygon: it only
    // lets the final atomicMax be skipped, and a stale value can only be lower)
    unsigned long long* pmkSlot = packMaxKey + (A & (PMK_SLOTS - 1)) * pmkPitch + pk;
    unsigned long long pmk = *pmkSlot;
    bool cullA = cullable(bA);
    uint64_t kmax = 0;
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
    int staged = -1;
    if (nA <= 64) {
        stage_a(L, lane, stA, nA, 0, Bu, pos, cellOf, CELLS != 0);
        staged = 0;
        if (nA >= 4)
            edge_tile<CELLS, TILE_SELF, SRC_STAGE, CONS>(L, lane, stA, nA, stA, nA, sAl, sAl, Bu, s0, pos, cellOf, staged,
                                                   out, num, kmax);
    } else {
        edge_tile<CELLS, TILE_SELF, SRC_GLOBAL, CONS>(L, lane, stA, nA, stA, nA, sAl, sAl, Bu, s0, pos, cellOf, staged, out,
                                                num, kmax);
    }
    int T = cell_list_length(cnt);
    for (int c0 = 0; c0 < T; c0 += 64) {
        int ci = c0 + lane;
        int q = (ci < T) ? cell_slot(lo, cnt, ci) : 0;
        int Bq = -1, stq = 0, nq = 0;
        double4 bc = make_double4(0.0, 0.0, 0.0, 1.0);
        if (ci < T) {
            int cand = pcSorted[s0 + q];
            if (cand > A) {   // the candidate's circle and vertex range, loaded together
                double4 c4 = bounds[cand];
                int st4 = polyStart[cand], en4 = polyStart[cand + 1];
                if (big || circles_meet(bA, c4)) {
                    Bq = cand;
                    bc = c4;
                    stq = st4;
                    nq = en4 - st4;
                }
            }
        }
        bool packed = Bq >= 0 && nA <= 64 && nq <= 64;
        // candidates outside the packed scheme: one tile each
        unsigned long long tiles = __ballot(Bq >= 0 && !packed);
        while (tiles) {
            int j = __ffsll((long long)tiles) - 1;
            tiles &= tiles - 1;
            int Bsel = __builtin_amdgcn_readlane(Bq, j);
            int stB = __builtin_amdgcn_readlane(stq, j), nB = __builtin_amdgcn_readlane(nq, j);
            edge_tile<CELLS, TILE_A_FIRST, SRC_GLOBAL, CONS>(L, lane, stA, nA, stB, nB, sAl, Bsel - s0, Bu, s0, pos, cellOf,
                                                       staged, out, num, kmax);
        }
        // packed candidates, in lane order, cut into batches
        unsigned long long pmask = __ballot(packed);
        int nqp = packed ? nq : 0;
        int inc = nqp;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(inc, o);
            if (lane >= o) inc += y;
        }
        int exc = inc - nqp;
        int prank = __popcll(pmask & bits_below(lane));
        bool cullB = packed && cullA && cullable(bc);
        unsigned long long rest = pmask;
        while (rest) {
            int j0 = __ffsll((long long)rest) - 1;
            int base = __builtin_amdgcn_readlane(exc, j0), r0 = __builtin_amdgcn_readlane(prank, j0);
            bool inB = ((rest >> lane) & 1ull) && inc - base <= DET_BATCH && prank - r0 < DET_KMAX;
            unsigned long long batch = __ballot(inB);   // holds lane j0 at least (nq <= 64)
            rest &= ~batch;
            int K = __popcll(batch);
            int kr = inB ? prank - r0 : 0;   // the candidate's place in the batch
            int off = inB ? exc - base : 0;  // its first vertex in L.V
            int nqb = inB ? nq : 0;
            int Vb = nqb, m = nqb;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                Vb += __shfl_xor(Vb, o);
                m = max(m, __shfl_xor(m, o));
            }
            Vb = __builtin_amdgcn_readfirstlane(Vb);
            m = __builtin_amdgcn_readfirstlane(m);
            // the batch's vertices to LDS, each candidate by its own lane, all loads in flight
            wave_lds_sync();   // the previous batch's reads are issued
            for (int v0 = 0; v0 < m; v0 += 8) {
                double2 t[8];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    t[k] = (v0 + k < nqb) ? pos[stq + v0 + k] : make_double2(0.0, 0.0);
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (v0 + k < nqb) {
                        L.V[off + v0 + k] = t[k];
                        L.own[off + v0 + k] = (uint8_t)lane;
                    }
            }
            if (inB) L.klane[kr] = (uint8_t)lane;
            wave_lds_sync();
            // candidate edges near A's circle, listed per candidate in edge order
            int cntB = 0;
            for (int e0 = 0; e0 < Vb; e0 += 64) {
                int e = e0 + lane;
                bool valid = e < Vb;
                int k = valid ? L.own[e] : 0;
                int offk = __shfl(off, k), nk = __shfl(nqb, k), carry = __shfl(cntB, k);
                int stk = CELLS ? __shfl(stq, k) : 0;
                bool cu = __shfl((int)cullB, k) != 0;
                bool near = false;
                int v = e - offk;
                if (valid) {
                    double2 q0 = L.V[e], q1 = L.V[(v + 1 == nk) ? offk : e + 1];
                    d2 R = mk(wrap(q1.x - q0.x), wrap(q1.y - q0.y));
                    near = !cu || edge_near_circle(mk(q0.x, q0.y), R, bA.x, bA.y, bA.z);
                    // the edge's cell for the stencil filter, computed here for every edge at
                    // once rather than per crossing candidate in the passes below
                    L.cellB[e] = CELLS ? cell_xy(cellOf[stk + v], Bu) : edge_cell_xy(mk(q0.x, q0.y), mk(q1.x, q1.y), Bu);
                }
                unsigned long long nm = __ballot(near);
                if (near) L.listB[offk + carry + __popcll(nm & bit_range(offk - e0, lane))] = (uint8_t)v;
                cntB += __popcll(nm & bit_range(off - e0, off + nqb - e0));
            }
            // A's edges near each candidate's circle, listed per candidate in edge order
            int cntA = 0;
            int KA = K * nA;   // < DET_KMAX * 64 = 1024, so small_div applies (KA * nA < 2^16)
            for (int i0 = 0; i0 < KA; i0 += 64) {
                int i = i0 + lane;
                bool valid = i < KA;
                int kb = valid ? small_div(i, nA) : 0;
                int a = i - kb * nA;
                int kl = L.klane[kb];
                double cx = __shfl(bc.x, kl), cy = __shfl(bc.y, kl), cr = __shfl(bc.z, kl);
                int carry = __shfl(cntA, kl);
                bool cu = __shfl((int)cullB, kl) != 0;
                bool near = false;
                if (valid) {
                    double2 p = L.P[a], r = L.R[a];
                    near = !cu || edge_near_circle(mk(p.x, p.y), mk(r.x, r.y), cx, cy, cr);
                }
                unsigned long long nm = __ballot(near);
                if (near) L.listA[kb * 64 + carry + __popcll(nm & bit_range(kb * nA - i0, lane))] = (uint8_t)a;
                if (inB) cntA += __popcll(nm & bit_range(kr * nA - i0, (kr + 1) * nA - i0));
            }
            wave_lds_sync();
            // all pairs (A edge, candidate edge) of the lists, 64 per pass
            int np = inB ? cntA * cntB : 0;
            int pinc = np;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                int y = __shfl_up(pinc, o);
                if (lane >= o) pinc += y;
            }
            int pexc = pinc - np;
            int Tp = __builtin_amdgcn_readlane(pinc, 63);
            if (inB) {
                L.kinfo[kr] = make_int4(pexc, cntB, off, nqb);
                L.kpoly[kr] = make_int2(Bq, stq);
            }
            wave_lds_sync();
            // the batch's pair starts in registers: the owner of pair qq is the last candidate
            // whose first pair is <= qq (a candidate without pairs shares its successor's start
            // and so is never the last), found by counting, with no memory access per pass
            int pst[DET_KMAX];
#pragma unroll
            for (int j = 0; j < DET_KMAX; ++j) pst[j] = (j < K) ? L.kinfo[j].x : 0x7FFFFFFF;
            // pair qq of the batch: its candidate, its two edges, the division-free test
            struct PairT {
                SegNum s;
                int a, b, offk, Bk;
                bool may;
            };
            auto pair_at = [&](int qq) {
                PairT t;
                t.may = false;
                t.a = 0;
                t.b = 0;
                bool valid = qq < Tp;
                int krk = -1;
#pragma unroll
                for (int j = 0; j < DET_KMAX; ++j) krk += (qq >= pst[j]) ? 1 : 0;
                krk = valid ? krk : 0;
                int4 ki = L.kinfo[krk];
                int2 kp = L.kpoly[krk];
                int r = qq - ki.x;
                int nbS = ki.y, nk = ki.w;
                t.offk = ki.z;
                t.Bk = kp.x;
                if (valid) {
                    // r = ia * nbS + ib; the float quotient is exact here (r < 4096, nbS <= 64:
                    // (r + 0.5) / nbS stays 1/128 away from an integer)
                    int ia = (int)(((float)r + 0.5f) * __builtin_amdgcn_rcpf((float)nbS));
                    int ib = r - ia * nbS;
                    t.a = L.listA[krk * 64 + ia];
                    t.b = L.listB[t.offk + ib];
                    double2 q0 = L.V[t.offk + t.b], q1 = L.V[t.offk + ((t.b + 1 == nk) ? 0 : t.b + 1)];
                    double2 pa = L.P[t.a], ra = L.R[t.a];
                    d2 Sb = mk(wrap(q1.x - q0.x), wrap(q1.y - q0.y));
                    t.s = seg_numerators(mk(pa.x, pa.y), mk(ra.x, ra.y), mk(q0.x, q0.y), Sb);
                    if (CONS) {   // A is staged whole here (nA <= 64)
                        double2 pn = L.P[(t.a + 1 == nA) ? 0 : t.a + 1];
                        t.may = seg_cons_cross(t.s, mk(pa.x, pa.y), mk(ra.x, ra.y), mk(pn.x, pn.y), mk(q0.x, q0.y),
                                               Sb, mk(q1.x, q1.y));
                    } else {
                        t.may = seg_may_cross(t.s);
                    }
                }
                return t;
            };
            // the exact test, the stencil filter and the record of a pair that passed
            auto exact_pass = [&](const PairT& t) {
                uint64_t key = 0;
                float2 tu;
                bool rec = false;
                if (t.may && cells_adjacent(L.C[t.a], L.cellB[t.offk + t.b], Bu)) {
                    if (CONS) {
                        cons_record(t.s, sAl, t.a, t.Bk - s0, t.b, key, tu);
                        rec = true;
                    } else {
                        rec = exact_record(t.s, sAl, t.a, t.Bk - s0, t.b, key, tu);
                    }
                }
                if (rec) {
                    kmax = (key > kmax) ? key : kmax;
                    if (CELLS) atomicAdd(&num[stA + t.a], 1);   // per-edge count of edge 1 (A's)
                    emit_record(key, tu, s0, out);
                }
            };
            // the collected pairs, 64 per exact pass (their tests recomputed from the ids)
            int nmay = 0;
            auto flush = [&]() {
                wave_lds_sync();
                for (int f0 = 0; f0 < nmay; f0 += 64) {
                    PairT t = pair_at(f0 + lane < nmay ? (int)L.may[f0 + lane] : Tp);
                    exact_pass(t);
                }
                wave_lds_sync();   // the list's reads are issued before it is refilled
                nmay = 0;
            };
            // two pairs per lane per pass: two independent chains in flight
            for (int p0 = 0; p0 < Tp; p0 += 128) {
                PairT t0 = pair_at(p0 + lane), t1 = pair_at(p0 + 64 + lane);
                unsigned long long m0 = __ballot(t0.may), m1 = __ballot(t1.may);
                if (m0 | m1) {   // wavefront-uniform: most passes find nothing
                    int c0 = __popcll(m0), c1 = __popcll(m1);
                    if (nmay + c0 + c1 > DET_MAY) flush();
                    unsigned long long below = bits_below(lane);
                    if (t0.may) L.may[nmay + __popcll(m0 & below)] = (uint16_t)(p0 + lane);
                    if (t1.may) L.may[nmay + c0 + __popcll(m1 & below)] = (uint16_t)(p0 + 64 + lane);
                    nmay += c0 + c1;
                }
            }
            if (nmay) flush();
        }
    }
    // the largest key of the packing (the pairing's last-record case)
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        unsigned long long y = __shfl_xor((unsigned long long)kmax, o);
        kmax = (y > kmax) ? y : kmax;
    }
    if (lane == 0 && kmax > 0 && kmax > pmk) atomicMax(pmkSlot, (unsigned long long)kmax);
}

// ---------------------------------------------------------------- tiled polygon pairs
// k_detect_tile, used when every polygon of the context has at most 64 vertices (the
// segment width SEG is the smallest power of two >= the largest polygon, at least 8): one
// wavefront per polygon A, A's edges staged in LDS.  The polygon pairs A owns -- A with
// itself, and A with every polygon of larger id in the 3x3 block of polygon cells whose
// circle meets A's -- are laid out one per segment of SEG lanes, 64/SEG pairs per pass;
// lane b of a segment holds edge b of the pair's other polygon B in registers.  Every edge
// a of A is then read from LDS by all lanes at once (a broadcast) and tested against every
// lane's edge: all nA x nB edge pairs (the nA(nA-3)/2 non-adjacent ones for A itself) are
// screened by the FP32 bounding boxes of the two edges about A's circle centre (four
// comparisons, where the division-free FP64 prefilter of the segment test takes about 28
// FP64 operations); the pairs whose boxes meet are collected and take the segment test
// (prefilter, then the exact one), the user-grid stencil filter and the record 64 at a
// time.  There is no edge-circle cull: it removes only edge pairs that cannot cross (the
// records are the same), and the regular tile costs fewer instructions than listing and
// enumerating the culled pairs.
// edge pairs whose boxes met, awaiting the segment test: (a, owner lane) per entry
// (4 wavefronts x (LDS stage + list) stay within 20 KiB: eight workgroups per CU)
constexpr int TILE_MAY = 512;
struct TileLds {
    double2 P[64], R[64];   // A's edges: start and wrapped edge vector
    int C[64];              // their user-grid cells (x | y << 16)
    int uB[65], uSt[65], uN[65];   // the pairs of a pass group: other polygon, first vertex, vertices
    int uBox[65];                  // whether the pair is judged by bounding boxes (both polygons cullable)
    uint16_t may[TILE_MAY];        // a | lane << 6
};

// Bounding box of an edge (start at rel, vector R; rel relative to an origin within a few
// polygon radii) widened by TILE_BOX_EPS and rounded to float: an edge pair whose boxes do
// not meet has no crossing that the segment test can find (its t, u lie in (0, 1) only for
// segments that meet to within rounding, ~1e-16, far inside the widening).  Valid where the
// minimum images of both polygons' vertices about the origin follow their chains, i.e. for
// cullable polygons (radius < CULL_RADIUS, chain closed), as the edge-circle cull.
constexpr double TILE_BOX_EPS = 1e-6;
__device__ __forceinline__ float4 edge_box(double rx, double ry, d2 R) {
    double x0 = fmin(rx, rx + R.x), x1 = fmax(rx, rx + R.x);
    double y0 = fmin(ry, ry + R.y), y1 = fmax(ry, ry + R.y);
    return make_float4((float)(x0 - TILE_BOX_EPS), (float)(x1 + TILE_BOX_EPS), (float)(y0 - TILE_BOX_EPS),
                       (float)(y1 + TILE_BOX_EPS));
}
__device__ __forceinline__ bool boxes_meet(float4 a, float4 b) {
    return a.x <= b.y && b.x <= a.y && a.z <= b.w && b.z <= a.w;
}
// the box held by lane l (wavefront-uniform l): four scalar reads, no LDS round trip
__device__ __forceinline__ float4 lane_box(float4 v, int l) {
    return make_float4(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), l)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), l)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.z), l)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.w), l)));
}

template <int CELLS, int SEG, bool CONS>
__device__ void tile_pairs(TileLds& L, int lane, int nU, int A, int stA, int nA, int sAl, int s0, int Bu, double ox,
                           double oy, float4 aBox, const double2* __restrict__ pos, const int* __restrict__ cellOf,
                           const DetOut& out, int* __restrict__ num, uint64_t& kmax) {
    constexpr int S = 64 / SEG;
    const int j = lane / SEG, b = lane & (SEG - 1);
    const float4 anyBox = make_float4(-INFINITY, INFINITY, -INFINITY, INFINITY);
    for (int u0 = 0; u0 < nU; u0 += S) {
        int u = u0 + j;
        bool has = u < nU;
        int Bq = has ? L.uB[u] : -1, stB = has ? L.uSt[u] : 0, nB = has ? L.uN[u] : 0;
        bool self = Bq == A;
        bool act = has && b < nB;
        int bn = (b + 1 == nB) ? 0 : b + 1;
        d2 Q = mk(0.0, 0.0), Sv = mk(0.0, 0.0), Qn = mk(0.0, 0.0);
        int cb = 0;   // the user-grid cell of B's edge (for the stencil filter)
        float4 bb = anyBox;   // B's edge box (every box meets it when the pair is not judged by boxes)
        if (act) {
            d2 q0 = ldpos(pos, stB + b), q1 = ldpos(pos, stB + bn);
            Q = q0;
            Qn = q1;
            Sv = mk(wrap(q1.x - q0.x), wrap(q1.y - q0.y));
            cb = CELLS ? cell_xy(cellOf[stB + b], Bu) : edge_cell_xy(q0, q1, Bu);
            if (L.uBox[u]) bb = edge_box(wrap(q0.x - ox), wrap(q0.y - oy), Sv);
        }
        int nmay = 0;
        // the segment test, the stencil filter and the record of the collected pairs, 64 at a
        // time (one memory round trip for all their records): B's edge comes from the lane
        // that holds it
        auto flush = [&]() {
            wave_lds_sync();
            for (int f0 = 0; f0 < nmay; f0 += 64) {
                bool valid = f0 + lane < nmay;
                int e = valid ? (int)L.may[f0 + lane] : 0;
                int a = e & 63, ow = e >> 6;
                double qx = __shfl(Q.x, ow), qy = __shfl(Q.y, ow), sx = __shfl(Sv.x, ow), sy = __shfl(Sv.y, ow);
                double qnx = CONS ? __shfl(Qn.x, ow) : 0.0, qny = CONS ? __shfl(Qn.y, ow) : 0.0;
                int cbo = __shfl(cb, ow);
                int bo = ow & (SEG - 1), uo = u0 + ow / SEG;
                uint64_t key = 0;
                float2 tu;
                bool rec = false;
                if (valid && cells_adjacent(L.C[a], cbo, Bu)) {
                    double2 pa = L.P[a], ra = L.R[a];
                    SegNum sn = seg_numerators(mk(pa.x, pa.y), mk(ra.x, ra.y), mk(qx, qy), mk(sx, sy));
                    if (CONS) {   // A is staged whole (every polygon of the context has <= 64 vertices)
                        double2 pn = L.P[(a + 1 == nA) ? 0 : a + 1];
                        if (seg_cons_cross(sn, mk(pa.x, pa.y), mk(ra.x, ra.y), mk(pn.x, pn.y), mk(qx, qy), mk(sx, sy),
                                           mk(qnx, qny))) {
                            cons_record(sn, sAl, a, L.uB[uo] - s0, bo, key, tu);
                            rec = true;
                        }
                    } else if (seg_may_cross(sn)) {
                        rec = exact_record(sn, sAl, a, L.uB[uo] - s0, bo, key, tu);
                    }
                }
                if (rec) {
                    kmax = (key > kmax) ? key : kmax;
                    if (CELLS) atomicAdd(&num[stA + a], 1);   // per-edge count of edge 1 (A's)
                    emit_record(key, tu, s0, out);
                }
            }
            wave_lds_sync();   // the list's reads are issued before it is refilled
            nmay = 0;
        };
        // the edges a of A this lane's edge is tested against, aLo <= a < aLo + aSpan: every edge of
        // A, or for A itself the non-adjacent ones below b (a < b - 1, and not edge 0 with edge
        // nA - 1); one unsigned comparison per edge of A in the hot loop
        int aLo = 0, aSpan = 0;
        if (act) {
            aLo = (self && b == nA - 1) ? 1 : 0;
            aSpan = self ? max(b - 1 - aLo, 0) : nA;
        }
        // every edge a of A against every lane's edge: the pairs whose boxes meet are collected
        for (int a0 = 0; a0 < nA; a0 += TILE_UNROLL) {
            bool may[TILE_UNROLL];
            bool anyMay = false;
#pragma unroll
            for (int k = 0; k < TILE_UNROLL; ++k) {
                int a = a0 + k;
                float4 ab = lane_box(aBox, (a < nA) ? a : 0);   // A's edge a, from lane a
                may[k] = a < nA && (unsigned)(a - aLo) < (unsigned)aSpan && boxes_meet(ab, bb);
                anyMay = anyMay || may[k];
            }
            if (__ballot(anyMay) == 0ull) continue;   // wavefront-uniform: most blocks find nothing
            if (nmay > TILE_MAY - TILE_UNROLL * 64) flush();
            unsigned long long below = bits_below(lane);
#pragma unroll
            for (int k = 0; k < TILE_UNROLL; ++k) {
                unsigned long long m = __ballot(may[k]);
                if (may[k]) L.may[nmay + __popcll(m & below)] = (uint16_t)((a0 + k) | (lane << 6));
                nmay += __popcll(m);
            }
        }
        if (nmay) flush();
    }
}

template <int CELLS, int SEG, bool CONS = false>
__global__ __launch_bounds__(BLK) void k_detect_tile(
    int N, const int* __restrict__ polyStart, const int* __restrict__ polyPack,
    const int* __restrict__ packPoly, const int* __restrict__ box, const double2* __restrict__ pos,
    const int* __restrict__ cellOf, const double4* __restrict__ bounds, const int* __restrict__ pcStart,
    const int* __restrict__ pcSorted, const int* __restrict__ packB, int Kp,
    int* __restrict__ bcnt, uint64_t* __restrict__ bKey, float2* __restrict__ bTU, int* __restrict__ jcnt,
    uint64_t* __restrict__ jKey, unsigned long long* __restrict__ packMaxKey, long long pmkPitch,
    int* __restrict__ num, int* __restrict__ scal, const int* __restrict__ work, const int* __restrict__ nWork) {
    __shared__ TileLds lds[BLK / 64];
    int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int nw, wb = work_block(nWork, BLK / 64, nw, N);
    int idx = wb * (BLK / 64) + w;
    if (wb < 0 || idx >= nw) return;   // whole wavefront
    TileLds& L = lds[w];
    DetOut out{Kp, bcnt, bKey, bTU, jcnt, jKey, scal};
    int A = __builtin_amdgcn_readfirstlane(work[idx]);
    int stA = polyStart[A], nA = polyStart[A + 1] - stA;
    int pk = polyPack[A];
    int Bu = box[pk];
    int s0 = packPoly[pk];
    int pb = packB[pk];
    int Bp = pb & 0xFFFF;
    bool big = (pb >> 30) & 1;
    int sAl = A - s0;
    double4 bA = bounds[A];
    // the packing's largest key so far (read here, with the other loads of the polygon: it only
    // lets the final atomicMax be skipped, and a stale value can only be lower)
    unsigned long long* pmkSlot = packMaxKey + (A & (PMK_SLOTS - 1)) * pmkPitch + pk;
    unsigned long long pmk = *pmkSlot;
    uint64_t kmax = 0;
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
    stage_a(L, lane, stA, nA, 0, Bu, pos, cellOf, CELLS != 0);
    // A's edge boxes about A's circle centre, edge a in lane a (every box of a polygon that
    // is not cullable meets every other: its pairs take the segment test)
    bool cullA = cullable(bA);
    float4 aBox = make_float4(-INFINITY, INFINITY, -INFINITY, INFINITY);
    if (lane < nA && cullA) {
        double2 p = L.P[lane], r = L.R[lane];
        aBox = edge_box(wrap(p.x - bA.x), wrap(p.y - bA.y), mk(r.x, r.y));
    }
    int T = cell_list_length(cnt);
    bool selfPair = nA >= 4;   // a polygon of fewer vertices has no non-adjacent edge pair
    for (int c0 = 0; c0 < T || (c0 == 0 && selfPair); c0 += 64) {
        int ci = c0 + lane;
        int q = (ci < T) ? cell_slot(lo, cnt, ci) : 0;
        bool keep = false, boxed = false;
        int cand = -1;
        int stq = 0, nq = 0;
        if (ci < T) {
            cand = pcSorted[s0 + q];
            if (cand > A) {   // the candidate's circle and vertex range, loaded together
                double4 c4 = bounds[cand];
                int st4 = polyStart[cand], en4 = polyStart[cand + 1];
                keep = big || circles_meet(bA, c4);
                boxed = cullA && cullable(c4);
                if (keep) {
                    stq = st4;
                    nq = en4 - st4;
                }
            }
        }
        unsigned long long km = __ballot(keep);
        int first = (c0 == 0 && selfPair) ? 1 : 0;
        int nU = first + __popcll(km);
        wave_lds_sync();   // the previous group's unit reads are issued
        if (first && lane == 0) {
            L.uB[0] = A;
            L.uSt[0] = stA;
            L.uN[0] = nA;
            L.uBox[0] = cullA ? 1 : 0;
        }
        if (keep) {
            int u = first + __popcll(km & bits_below(lane));
            L.uB[u] = cand;
            L.uSt[u] = stq;
            L.uN[u] = nq;
            L.uBox[u] = boxed ? 1 : 0;
        }
        wave_lds_sync();
        tile_pairs<CELLS, SEG, CONS>(L, lane, nU, A, stA, nA, sAl, s0, Bu, bA.x, bA.y, aBox, pos, cellOf, out, num,
                                     kmax);
    }
    // the largest key of the packing (the pairing's last-record case)
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        unsigned long long y = __shfl_xor((unsigned long long)kmax, o);
        kmax = (y > kmax) ? y : kmax;
    }
    if (lane == 0 && kmax > 0 && kmax > pmk) atomicMax(pmkSlot, (unsigned long long)kmax);
}

// One wavefront per polygon A (in cell order, XCD-contiguous): detect_polygon.
// CELLS = 0: the user-grid cell of every edge is computed from the live positions (the
// fused evaluation, where cells and crossings come from the same positions);
// CELLS = 1: the cells of the last cell-list build are used (updateNeighbors after
// updateNeighborCells, which may have been called at other positions).
// The bucket counters bcnt/jcnt must be zero (k_bounds clears them).
template <int CELLS, bool CONS = false>
__global__ __launch_bounds__(BLK) void k_detect_poly(
    int N, const int* __restrict__ polyStart, const int* __restrict__ polyPack,
    const int* __restrict__ packPoly, const int* __restrict__ box, const double2* __restrict__ pos,
    const int* __restrict__ cellOf, const double4* __restrict__ bounds, const int* __restrict__ pcStart,
    const int* __restrict__ pcSorted, const int* __restrict__ packB, int Kp,
    int* __restrict__ bcnt, uint64_t* __restrict__ bKey, float2* __restrict__ bTU, int* __restrict__ jcnt,
    uint64_t* __restrict__ jKey, unsigned long long* __restrict__ packMaxKey, long long pmkPitch,
    int* __restrict__ num,
    int* __restrict__ scal, const int* __restrict__ work, const int* __restrict__ nWork) {
    __shared__ DetLds lds[BLK / 64];
    int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int nw, wb = work_block(nWork, BLK / 64, nw, N);
    int idx = wb * (BLK / 64) + w;
    if (wb < 0 || idx >= nw) return;   // whole wavefront
    DetOut out{Kp, bcnt, bKey, bTU, jcnt, jKey, scal};
    detect_polygon<CELLS, CONS>(__builtin_amdgcn_readfirstlane(work[idx]), lane, lds[w], polyStart, polyPack, packPoly,
                          box, pos, cellOf, bounds, pcStart, pcSorted, packB, out, packMaxKey, pmkPitch,
                          num);
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

template <bool POST>
static int launch_bounds_t(pcp_ctx* c, const double2* pos, const PostShake& ps) {
    int seg = c->seg;
    int polysPerBlock = BLK / seg;
    int grid = (c->N + polysPerBlock - 1) / polysPerBlock;
    if (grid < 1) grid = 1;
#define PCP_BOUNDS(W)                                                                                          \
    PCP_LAUNCH(c, (k_bounds<W, POST>), dim3(grid), dim3(BLK), c->N, c->P, c->dPolyStart, pos, c->dBounds,     \
               c->dNum, c->dBcnt, c->dJcnt, c->dPackMaxKey, c->pmkPitch, c->dScal, ps)
    switch (seg) {
    case 4: PCP_BOUNDS(4); break;
    case 8: PCP_BOUNDS(8); break;
    case 16: PCP_BOUNDS(16); break;
    case 32: PCP_BOUNDS(32); break;
    default: PCP_BOUNDS(64); break;
    }
#undef PCP_BOUNDS
    return PCP_OK;
}
int launch_bounds(pcp_ctx* c, const double2* pos, const PostShake* ps) {
    if (ps) return launch_bounds_t<true>(c, pos, *ps);
    PostShake none{};
    return launch_bounds_t<false>(c, pos, none);
}

// k_detect_tile<CELLS, SEG> on the context's stream; its launches are profiled under the
// name "k_detect_tile" (pcp_profile_kernel), with events attached to the dispatch
template <int CELLS, int SEG, bool CONS>
static int launch_tile_t(pcp_ctx* c, const double2* pos, int grid) {
    int slot = -1;
    PCP_TRY(prof_slot(c, "k_detect_tile", &slot));
    auto kern = k_detect_tile<CELLS, SEG, CONS>;
    if (slot >= 0)
        hipExtLaunchKernelGGL(kern, dim3(grid), dim3(BLK), 0, c->stream, c->profEvents[slot].first,
                              c->profEvents[slot].second, 0, c->N, c->dPolyStart, c->dPolyPack, c->dPackPoly,
                              c->dBox, pos, c->dCellOf, c->dBounds, c->dPcStart, c->dPcSorted, c->dPackB,
                              c->Kp, c->dBcnt, c->dBKey, c->dBTU, c->dJcnt, c->dJKey, c->dPackMaxKey, c->pmkPitch,
                              c->dNum, c->dScal, work_list(c), work_count(c));
    else
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BLK), 0, c->stream, c->N, c->dPolyStart, c->dPolyPack,
                           c->dPackPoly, c->dBox, pos, c->dCellOf, c->dBounds, c->dPcStart, c->dPcSorted,
                           c->dPackB, c->Kp, c->dBcnt, c->dBKey, c->dBTU, c->dJcnt, c->dJKey,
                           c->dPackMaxKey, c->pmkPitch, c->dNum, c->dScal, work_list(c), work_count(c));
    PCP_LAUNCHED();
    return PCP_OK;
}
template <int CELLS, bool CONS>
static int launch_tile_c(pcp_ctx* c, const double2* pos, int grid) {
    switch (c->seg) {
    case 4:
    case 8: return launch_tile_t<CELLS, 8, CONS>(c, pos, grid);
    case 16: return launch_tile_t<CELLS, 16, CONS>(c, pos, grid);
    case 32: return launch_tile_t<CELLS, 32, CONS>(c, pos, grid);
    default: return launch_tile_t<CELLS, 64, CONS>(c, pos, grid);
    }
}
static int launch_detect_tile(pcp_ctx* c, const double2* pos, bool cellsFromList, bool cons, int grid) {
    if (cons) return cellsFromList ? launch_tile_c<1, true>(c, pos, grid) : launch_tile_c<0, true>(c, pos, grid);
    return cellsFromList ? launch_tile_c<1, false>(c, pos, grid) : launch_tile_c<0, false>(c, pos, grid);
}

// the detection schedule in effect: the one set (pcp_set_detect_schedule / pcp_autotune), else
// tiles for polygons of at most 16 vertices (the box-screened tile needs fewer instructions and
// fewer dependent steps than listing and enumerating the culled pairs: 20.0 against 21.8 us at
// one N=4096 packing, 245 against 388 us at 32), the culled enumeration otherwise (with 32- or
// 64-lane segments small polygons leave tile lanes idle)
int detect_schedule(const pcp_ctx* c) {
    if (c->detSched >= 0) return (c->maxN <= 64) ? c->detSched : 0;
    return (c->maxN <= 64 && c->seg <= 16) ? 1 : 0;
}

// cellsFromList: use the cells of the last run_cells (API updateNeighbors).  ps: FIRE's
// post-SHAKE step, done by the bounds kernel (see k_bounds).  cons: the side-consistent
// decision (PCP_CONSISTENT) instead of the reference's parameter test.
int run_detect(pcp_ctx* c, const double2* pos, int cellsFromList, const PostShake* ps, bool cons) {
    if (c->V == 0) return PCP_OK;
    c->lastDetectPos = pos;
    PCP_TRY(launch_bounds(c, pos, ps));
    PCP_LAUNCH(c, k_poly_cells, dim3(c->P), dim3(PP_BLK), c->dPackPoly, c->dBounds, c->dPcStart, c->dPcSorted,
               c->dPackB, c->skipDone ? (const int*)c->dActive : nullptr, c->dActOrder, c->dActCount);
    int grid = (c->N + BLK / 64 - 1) / (BLK / 64);
    if (c->maxN <= 64 && detect_schedule(c) == 1) return launch_detect_tile(c, pos, cellsFromList != 0, cons, grid);
    if (cons) {
        if (cellsFromList)
            PCP_LAUNCH(c, (k_detect_poly<1, true>), dim3(grid), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack,
                       c->dPackPoly, c->dBox, pos, c->dCellOf, c->dBounds, c->dPcStart, c->dPcSorted,
                       c->dPackB, c->Kp, c->dBcnt, c->dBKey, c->dBTU, c->dJcnt, c->dJKey, c->dPackMaxKey,
                       c->pmkPitch, c->dNum, c->dScal, work_list(c), work_count(c));
        else
            PCP_LAUNCH(c, (k_detect_poly<0, true>), dim3(grid), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack,
                       c->dPackPoly, c->dBox, pos, c->dCellOf, c->dBounds, c->dPcStart, c->dPcSorted,
                       c->dPackB, c->Kp, c->dBcnt, c->dBKey, c->dBTU, c->dJcnt, c->dJKey, c->dPackMaxKey,
                       c->pmkPitch, c->dNum, c->dScal, work_list(c), work_count(c));
        return PCP_OK;
    }
    if (cellsFromList)
        PCP_LAUNCH(c, k_detect_poly<1>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack, c->dPackPoly,
                   c->dBox, pos, c->dCellOf, c->dBounds, c->dPcStart, c->dPcSorted, c->dPackB, c->Kp,
                   c->dBcnt,
                   c->dBKey, c->dBTU, c->dJcnt, c->dJKey, c->dPackMaxKey, c->pmkPitch, c->dNum, c->dScal, work_list(c), work_count(c));
    else
        PCP_LAUNCH(c, k_detect_poly<0>, dim3(grid), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack, c->dPackPoly,
                   c->dBox, pos, c->dCellOf, c->dBounds, c->dPcStart, c->dPcSorted, c->dPackB, c->Kp,
                   c->dBcnt,
                   c->dBKey, c->dBTU, c->dJcnt, c->dJKey, c->dPackMaxKey, c->pmkPitch, c->dNum, c->dScal, work_list(c), work_count(c));
    return PCP_OK;
}

int detect_alloc(pcp_ctx* c) {
    c->maxN = max_poly_vertices(c);
    c->seg = seg_width(c->maxN);
    c->row16 = c->N > 0;
    for (int s = 0; s < c->N; ++s) c->row16 = c->row16 && c->hPolyStart[s + 1] - c->hPolyStart[s] == 16;
    PCP_TRY(dalloc(&c->dBounds, std::max(c->N, 1)));
    PCP_TRY(dalloc(&c->dPcSorted, std::max(c->N, 1)));
    {   // identity until the first detection sorts the polygons by cell (kernels that visit
        // polygons in cell order also run for models without overlap)
        std::vector<int> iota(std::max(c->N, 1));
        for (int s = 0; s < (int)iota.size(); ++s) iota[s] = s;
        PCP_HIP(hipMemcpy(c->dPcSorted, iota.data(), sizeof(int) * iota.size(), hipMemcpyHostToDevice));
    }
    PCP_TRY(dalloc(&c->dPcStart, (long long)c->P * (PC_CELLS + 1)));
    PCP_TRY(dalloc(&c->dPackB, std::max(c->P, 1)));
    PCP_TRY(dalloc(&c->dActOrder, std::max(c->N, 1)));
    PCP_TRY(dalloc(&c->dActCount, 1));
    return PCP_OK;
}

// Work of the last detection, recomputed on the host from the device's polygon cells
// and positions with the kernel's rules: polygon pairs tested (each polygon with itself,
// each pair of polygons once) and the edge pairs whose segment test ran (A with itself:
// its non-adjacent pairs; A with B: the products of the edge lists left by the cull).
int detect_stats(pcp_ctx* c, long long* polygonPairs, long long* edgePairs, long long* screened) {
    PCP_HIP(hipStreamSynchronize(c->stream));
    long long scr = 0;
    if (!c->lastDetectPos) {
        *polygonPairs = 0;
        *edgePairs = 0;
        if (screened) *screened = 0;
        return PCP_OK;
    }
    bool tiles = detect_schedule(c) == 1;
    std::vector<double4> bd(c->N);
    std::vector<int> sorted(c->N), packB(c->P), start((size_t)c->P * (PC_CELLS + 1));
    std::vector<double2> X(c->V);
    PCP_HIP(hipMemcpy(bd.data(), c->dBounds, sizeof(double4) * c->N, hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(sorted.data(), c->dPcSorted, sizeof(int) * c->N, hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(packB.data(), c->dPackB, sizeof(int) * c->P, hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(start.data(), c->dPcStart, sizeof(int) * start.size(), hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(X.data(), c->lastDetectPos, sizeof(double2) * c->V, hipMemcpyDeviceToHost));
    auto wrapH = [](double x) { double y = x + 1.5; y = y - std::floor(y); return y - 0.5; };
    auto coord = [](double x, int B) { double f = std::floor(x * B); int i = (f >= 0.0) ? (int)f : 0; return (i >= B) ? B - 1 : i; };
    auto cullOk = [&](int q) { return bd[q].w == 0.0 && bd[q].z < CULL_RADIUS; };
    // edges of polygon q near the circle of polygon o (edge_near_circle)
    auto nearCount = [&](int q, int o) {
        int st = c->hPolyStart[q], n = c->hPolyStart[q + 1] - st;
        long long k = 0;
        for (int e = 0; e < n; ++e) {
            double2 p0 = X[st + e], p1 = X[st + ((e + 1 == n) ? 0 : e + 1)];
            double Rx = wrapH(p1.x - p0.x), Ry = wrapH(p1.y - p0.y);
            double dx = wrapH(bd[o].x - (p0.x + 0.5 * Rx)), dy = wrapH(bd[o].y - (p0.y + 0.5 * Ry));
            double rr = (bd[o].z + 0.5 * (std::fabs(Rx) + std::fabs(Ry))) * (1.0 + 1e-6) + 1e-12;
            k += (dx * dx + dy * dy <= rr * rr);
        }
        return k;
    };
    // the tile schedule's FP32 edge boxes about A's circle centre (edge_box, boxes_meet)
    auto boxOf = [&](int q, int e, const double4& o) {
        int st = c->hPolyStart[q], n = c->hPolyStart[q + 1] - st;
        double2 p0 = X[st + e], p1 = X[st + ((e + 1 == n) ? 0 : e + 1)];
        double Rx = wrapH(p1.x - p0.x), Ry = wrapH(p1.y - p0.y);
        double rx = wrapH(p0.x - o.x), ry = wrapH(p0.y - o.y);
        double x0 = std::fmin(rx, rx + Rx), x1 = std::fmax(rx, rx + Rx);
        double y0 = std::fmin(ry, ry + Ry), y1 = std::fmax(ry, ry + Ry);
        return std::array<float, 4>{(float)(x0 - TILE_BOX_EPS), (float)(x1 + TILE_BOX_EPS), (float)(y0 - TILE_BOX_EPS),
                                    (float)(y1 + TILE_BOX_EPS)};
    };
    auto meet = [](const std::array<float, 4>& a, const std::array<float, 4>& b) {
        return a[0] <= b[1] && b[0] <= a[1] && a[2] <= b[3] && b[2] <= a[3];
    };
    // edge pairs of A (edges ea) and B given the segment test by the tile: all of them where
    // either polygon is not cullable, else those whose boxes meet
    auto tilePairs = [&](int A, int Bq, bool self) {
        int nA = c->hPolyStart[A + 1] - c->hPolyStart[A], nB = c->hPolyStart[Bq + 1] - c->hPolyStart[Bq];
        long long k = 0, all = 0;
        bool boxed = cullOk(A) && cullOk(Bq);
        for (int a = 0; a < nA; ++a) {
            auto ba = boxOf(A, a, bd[A]);
            for (int b = 0; b < nB; ++b) {
                if (self && !(b > a + 1 && !(a == 0 && b == nA - 1))) continue;
                ++all;
                if (!boxed || meet(ba, boxOf(Bq, b, bd[A]))) ++k;
            }
        }
        scr += all;
        return k;
    };
    long long pp = 0, ep = 0;
    for (int p = 0; p < c->P; ++p) {
        int B = packB[p] & 0xFFFF;
        bool big = (packB[p] >> 30) & 1;
        int s0 = c->hPackPoly[p];
        const int* st = start.data() + (size_t)p * (PC_CELLS + 1);
        for (int A = s0; A < c->hPackPoly[p + 1]; ++A) {
            long long nA = c->hPolyStart[A + 1] - c->hPolyStart[A];
            if (nA >= 4) {
                ++pp;
                if (tiles) {
                    ep += tilePairs(A, A, true);
                } else {
                    ep += nA * (nA - 3) / 2;
                    scr += nA * (nA - 3) / 2;
                }
            }
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
                    long long nB = c->hPolyStart[Bq + 1] - c->hPolyStart[Bq];
                    if (tiles) {
                        ep += tilePairs(A, Bq, false);
                    } else {
                        long long k = (nA <= 64 && nB <= 64 && cullOk(A) && cullOk(Bq))
                                          ? nearCount(A, Bq) * nearCount(Bq, A) : nA * nB;
                        ep += k;
                        scr += k;
                    }
                }
            }
        }
    }
    *polygonPairs = pp;
    *edgePairs = ep;
    if (screened) *screened = scr;
    return PCP_OK;
}

}  // namespace pcp
