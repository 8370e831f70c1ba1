// Records, pairing and forces, one wavefront (or lane segment) per polygon (part of the
// pcp_kernels.hip translation unit).
//
// The detection appends every crossing to two per-polygon buckets of capacity Kp: the
// si bucket of the polygon the record belongs to (key, float t/u) and the sj bucket of
// the other polygon (key only).  Then:
//   k_pair_poly      wavefront / polygon s: rank of each record of s in the order
//                    (i, j, sj) -- the order left by the reference's 64-bit then stable
//                    48-bit radix sorts (model.cpp:354-360) -- and its exit record from
//                    the si bucket of sj (updateOutersectionsKernel kernels.cuh:981-1065);
//                    the sorted records of s are written to s's record rows
//   k_exterior_poly  wavefront / polygon s: the 8 exterior contributions and the energy of
//                    each record (updateForceEnergyExteriorKernel kernels.cuh:1067-1256);
//                    the 4 acting on s are summed per vertex in record order, the 4 acting
//                    on sj are written at the record's rank in sj's list (key order); the
//                    interior chain terms (updateForceEnergyInteriorKernel
//                    kernels.cuh:1354-1424) per vertex of s
//   k_gather_poly    lane segment / polygon: per vertex the sum over s's records, then the
//                    sj-side contributions in key order, the interior terms, the edge and
//                    area springs (kernels.cuh:1284-1352); the polygon energy; in the fused
//                    evaluation also edge lengths, area and the force projection
// The order of every floating-point sum is fixed (the CPU oracle uses the same), so
// forces and energies are bit-reproducible.

namespace pcp {

// k_exterior_poly keeps a polygon's records in LDS (Kp * 116 bytes, at most 64 KiB per
// workgroup); beyond that (Kp > 564) in a global staging buffer of the context
constexpr size_t EXT_LDS_MAX = 65536;

// ---------------------------------------------------------------- bucket scans
// Scans of a short bucket (a few to a few tens of keys) in groups of 8 with all the
// group's loads issued before any is used, so that a scan costs one memory latency per
// group instead of one per key.
template <class Below>
__device__ __forceinline__ int count_keys_below(const uint64_t* __restrict__ keys, int n, Below below) {
    int c = 0;
    for (int q0 = 0; q0 < n; q0 += 8) {
        uint64_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = (q0 + k < n) ? keys[q0 + k] : 0ull;
#pragma unroll
        for (int k = 0; k < 8; ++k) c += (q0 + k < n && below(v[k])) ? 1 : 0;
    }
    return c;
}

// ---------------------------------------------------------------- pairing
// The pairing of record f of polygon s's si bucket (k_pair_poly): the record, its place in
// the order (i, j, sj) of the bucket (rank), its exit record and that record's parameters
// (the record itself when it has no partner).  mxk: the packing's largest key.
struct PairRec {
    uint64_t key, outer;
    float2 tu, ut;
    int rank;
};
template <int EXACT>
__device__ __forceinline__ PairRec pair_record(int f, int cnt, long long base, int s, int pbase, int st1, int ni,
                                               uint64_t mxk, int Kp, const int* __restrict__ polyStart,
                                               const int* __restrict__ bcnt, const uint64_t* __restrict__ bKey,
                                               const float2* __restrict__ bTU, const double2* __restrict__ pairPos,
                                               int* __restrict__ scal) {
    uint64_t key = bKey[base + f];
    float2 tu = bTU[base + f];
    uint64_t ord = bucket_order(key);
    int si_l = key_si(key), sj_l = key_sj(key), i = key_i(key);
    // rank in the bucket, and the crossings of this polygon pair held in this bucket
    int rank = 0, nSame = 0;
    for (int q0 = 0; q0 < cnt; q0 += 8) {
        uint64_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = (q0 + k < cnt) ? bKey[base + q0 + k] : 0ull;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (q0 + k < cnt) {
                rank += (bucket_order(v[k]) < ord) ? 1 : 0;
                nSame += (key_sj(v[k]) == sj_l) ? 1 : 0;
            }
    }
    int nOther = 0;   // the crossings of the pair held in the other polygon's bucket
    int sjg = pbase + sj_l;
    int st2 = polyStart[sjg], nj = polyStart[sjg + 1] - st2;
    // exact parameter along s's edge ei of its crossing with sj's edge ej (area energy)
    auto exact = [&](int ei, int ej, double fb) -> double {
        int a = st1 + ei, b = st1 + ((ei + 1 == ni) ? 0 : ei + 1);
        int c2 = st2 + ej, d = st2 + ((ej + 1 == nj) ? 0 : ej + 1);
        return cross_param(ldpos(pairPos, a), ldpos(pairPos, b), ldpos(pairPos, c2), ldpos(pairPos, d), fb);
    };
    double tVal = (double)tu.y;
    if (EXACT) tVal = exact(i, key_j(key), tVal);
    uint64_t sij = ((uint64_t)si_l << 48) | ((uint64_t)sj_l << 32);
    int bestDist = -1;
    bool found = false;
    double bestU = 0.0;
    uint64_t bestOrd = 0, bestKey = 0;
    float2 bestTU = make_float2(0.f, 0.f);
    auto consider = [&](uint64_t k2, float2 tu2) {
        int l = key_j(k2);
        double uVal = (double)tu2.x;
        if (EXACT && mxk >= sij) uVal = exact(l, key_i(k2), uVal);
        int d = (l - i + ni) % ni;
        if (l == i && tVal >= uVal) return;
        if (bestDist == -1 || d < bestDist) { bestDist = d; found = false; }
        if (d == bestDist && uVal == uVal) {
            uint64_t o2 = bucket_order(k2);
            if (!found || uVal < bestU || (uVal == bestU && o2 < bestOrd)) {
                found = true; bestU = uVal; bestOrd = o2; bestKey = k2; bestTU = tu2;
            }
        }
    };
    if (mxk < sij) {
        int s3 = pbase + key_si(mxk);
        long long b3 = (long long)s3 * Kp;
        int c3 = min(bcnt[s3], Kp);
        for (int q = 0; q < c3; ++q)
            if (bKey[b3 + q] == mxk) { consider(mxk, bTU[b3 + q]); break; }
        // (no key of the packing reaches si<<48 | sj<<32, so sj's bucket holds no record of
        // this pair: nOther stays 0)
    } else {
        long long b2 = (long long)sjg * Kp;
        int c2 = min(bcnt[sjg], Kp);
        for (int q0 = 0; q0 < c2; q0 += 8) {   // the group's loads first, then in order
            uint64_t k8[8];
            float2 t8[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                k8[k] = (q0 + k < c2) ? bKey[b2 + q0 + k] : 0ull;
                t8[k] = (q0 + k < c2) ? bTU[b2 + q0 + k] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (q0 + k < c2 && key_sj(k8[k]) == si_l) {
                    consider(k8[k], t8[k]);
                    ++nOther;
                }
        }
    }
    // a polygon pair (not a polygon with itself) with an odd number of crossings
    if (sjg != s && ((nSame + nOther) & 1)) atomicAdd(&scal[SC_ODD], 1);
    // a partner whose edge indices fall outside the two polygons (possible only in the
    // last-record case above) is treated as no partner: the reference would read
    // vertices of other polygons there
    if (found && (key_i(bestKey) >= nj || key_j(bestKey) >= ni)) found = false;
    PairRec r;
    r.key = key;
    r.tu = tu;
    r.rank = rank;
    r.outer = found ? bestKey : key;
    r.ut = found ? bestTU : tu;
    return r;
}

// The reference's binary search starts with end = I-1, so when every key of the packing
// lies below (si<<48 | sj<<32) the last record of the packing is the only candidate;
// kept.  EXACT: crossing parameters recomputed from pairPos (area energy).
// SEGX lanes per polygon, 64 / SEGX polygons per wavefront (a polygon has a few records).
template <int EXACT, int SEGX>
__global__ __launch_bounds__(64) void k_pair_poly(int N, int Kp, const int* __restrict__ polyStart,
                                                 const int* __restrict__ polyPack, const int* __restrict__ packPoly,
                                                 const int* __restrict__ bcnt, const uint64_t* __restrict__ bKey,
                                                 const float2* __restrict__ bTU,
                                                 const unsigned long long* __restrict__ packMaxKey, long long pmkPitch,
                                                 const double2* __restrict__ pairPos, uint64_t* __restrict__ rKey,
                                                 float2* __restrict__ rTU, uint64_t* __restrict__ rOuter,
                                                 float2* __restrict__ rUT, const int* __restrict__ order,
                                                 const int* __restrict__ nWork, int* __restrict__ scal) {
    int lane = threadIdx.x & (SEGX - 1);
    int nw, wb = work_block(nWork, 64 / SEGX, nw, N);
    int slot = wb * (64 / SEGX) + threadIdx.x / SEGX;
    if (wb < 0 || slot >= nw) return;   // whole segments: the shuffles below stay inside a segment
    int s = order[slot];
    int cnt = min(bcnt[s], Kp);
    if (cnt == 0) return;
    int pk = polyPack[s];
    int pbase = packPoly[pk];
    int st1 = polyStart[s], ni = polyStart[s + 1] - st1;
    uint64_t mxk = 0;   // the maximum of the packing's partial maxima, across the segment
    for (int k = lane; k < PMK_SLOTS; k += SEGX) {
        uint64_t v = packMaxKey[k * pmkPitch + pk];
        mxk = (v > mxk) ? v : mxk;
    }
#pragma unroll
    for (int o = SEGX / 2; o >= 1; o >>= 1) {
        uint64_t y = __shfl_xor(mxk, o, SEGX);
        mxk = (y > mxk) ? y : mxk;
    }
    long long base = (long long)s * Kp;
    for (int f = lane; f < cnt; f += SEGX) {
        PairRec r = pair_record<EXACT>(f, cnt, base, s, pbase, st1, ni, mxk, Kp, polyStart, bcnt, bKey, bTU, pairPos,
                                       scal);
        rKey[base + r.rank] = r.key;
        rTU[base + r.rank] = r.tu;
        rOuter[base + r.rank] = r.outer;
        rUT[base + r.rank] = r.ut;
    }
}

// ---------------------------------------------------------------- exterior + interior
struct ExtLds {
    double2* c;    // Kp*4 contributions acting on s
    int* t;        // Kp*4 their vertex (local index in s, -1 none)
    double2* sp;   // Kp start points
    double* e;     // Kp record energies
    double* la;    // Kp lens areas (squared lens energy)
    int* il;       // Kp: iLoc | lLoc << 16, -1 for a record without partner
};
__device__ __forceinline__ ExtLds ext_lds(char* smem, int Kp) {
    ExtLds L;
    L.c = (double2*)smem;
    L.sp = L.c + 4 * Kp;
    L.e = (double*)(L.sp + Kp);
    L.la = L.e + Kp;
    L.t = (int*)(L.la + Kp);
    L.il = L.t + 4 * Kp;
    return L;
}
// one polygon's share of the LDS of k_exterior_poly, rounded to 16 bytes
__host__ __device__ __forceinline__ size_t ext_lds_stride(int Kp) { return ((size_t)Kp * 116 + 15) & ~(size_t)15; }

// the record of the packing whose key is `key`: its outer key and its row (false if absent)
__device__ __forceinline__ bool find_record(uint64_t key, int pbase, int Kp, const int* __restrict__ bcnt,
                                            const uint64_t* __restrict__ rKey,
                                            const uint64_t* __restrict__ rOuter, uint64_t* outer,
                                            long long* row = nullptr) {
    int s2 = pbase + key_si(key);
    long long b2 = (long long)s2 * Kp;
    int c2 = min(bcnt[s2], Kp);
    for (int q0 = 0; q0 < c2; q0 += 8) {   // the group's loads first, then the first match
        uint64_t k8[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) k8[k] = (q0 + k < c2) ? rKey[b2 + q0 + k] : ~key;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k8[k] == key) {
                *outer = rOuter[b2 + q0 + k];
                if (row) *row = b2 + q0 + k;
                return true;
            }
    }
    return false;
}

// The exit links from record (inter, outer) (row ownRow): whether they close back to it
// within 8 hops (a closed lens boundary).  Each link's record is searched for once; with
// KEEP the rows of the cycle's members (inter first, then in link order) go to rows,
// *cnt of them (at most CYCLE_MAX), and *k0 is the place of the smallest key among them
// (only the rows are kept, not the keys: registers set the kernels' occupancy).
constexpr int CYCLE_MAX = 10;
template <bool KEEP>
__device__ __forceinline__ bool cycle_walk(uint64_t inter, uint64_t outer, long long ownRow, int pbase, int Kp,
                                           const int* __restrict__ bcnt, const uint64_t* __restrict__ rKey,
                                           const uint64_t* __restrict__ rOuter, long long* rows, int* cnt, int* k0) {
    int n = 1;
    uint64_t kmin = inter;
    if (KEEP) { rows[0] = ownRow; *k0 = 0; }
    uint64_t cur = outer;
    for (int hop = 0;; ++hop) {
        if (cur == inter) {
            if (KEEP) *cnt = n;
            return true;
        }
        uint64_t nxt = 0;
        long long row = 0;
        if (!find_record(cur, pbase, Kp, bcnt, rKey, rOuter, &nxt, &row)) break;
        if (KEEP) {
            rows[n] = row;
            if (cur < kmin) { kmin = cur; *k0 = n; }
        }
        ++n;
        if (hop >= 8 || nxt == cur) break;
        cur = nxt;
    }
    if (KEEP) { *cnt = 1; *k0 = 0; }
    return false;
}
__device__ __forceinline__ bool closed_cycle(uint64_t inter, uint64_t outer, int pbase, int Kp,
                                             const int* __restrict__ bcnt, const uint64_t* __restrict__ rKey,
                                             const uint64_t* __restrict__ rOuter) {
    return cycle_walk<false>(inter, outer, 0, pbase, Kp, bcnt, rKey, rOuter, nullptr, nullptr, nullptr);
}

// Area of the lens of a record: half the h sums (Er) of the records of its closed cycle
// (rows/cnt/k0 from cycle_walk<true>), added in cycle order from the record with the
// smallest key; off a cycle, the record alone.  *rep: the record represents its lens
// (smallest key of the cycle, or off any).
__device__ __forceinline__ double lens_area(bool closed, const long long* rows, int cnt, int k0,
                                            const double* __restrict__ Er, int* lensN, bool* rep) {
    if (!closed) {
        *lensN = 1;
        *rep = true;
        return 0.5 * Er[rows[0]];
    }
    double sum = 0.0;
    for (int k = 0; k < cnt; ++k) {
        int q = k0 + k;
        sum += Er[rows[q >= cnt ? q - cnt : q]];
    }
    *lensN = cnt;
    *rep = (k0 == 0);
    return 0.5 * sum;
}

// Overlap area of the polygon pair (s, sj) (packing-global ids): laRep summed over the
// lower polygon's records whose sj is the higher one, then over the higher polygon's
// records whose sj is the lower one, each in record order.
__device__ __forceinline__ double pair_area(int s, int sj, int pbase, int Kp, const int* __restrict__ bcnt,
                                            const uint64_t* __restrict__ rKey, const double* __restrict__ laRep) {
    int a = min(s, sj), b = max(s, sj);
    double pa = 0.0;
    for (int pass = 0; pass < (a == b ? 1 : 2); ++pass) {
        int own = pass ? b : a, other = (pass ? a : b) - pbase;
        long long bo = (long long)own * Kp;
        int co = min(bcnt[own], Kp);
        for (int q = 0; q < co; ++q)
            if (key_sj(rKey[bo + q]) == other) pa += laRep[bo + q];
    }
    return pa;
}

// Squared pair areas, second pass (wavefront / polygon s, lane / record): each record's
// lens area if the record represents its lens, else 0.
template <int SEGX>
__global__ __launch_bounds__(64) void k_lens_rep(int N, int Kp, const int* __restrict__ polyPack,
                                                const int* __restrict__ packPoly, const int* __restrict__ bcnt,
                                                const uint64_t* __restrict__ rKey,
                                                const uint64_t* __restrict__ rOuter, const double* __restrict__ Er,
                                                double* __restrict__ laRep, const int* __restrict__ order,
                                                const int* __restrict__ nWork) {
    int lane = threadIdx.x & (SEGX - 1);
    int nw, wb = work_block(nWork, 64 / SEGX, nw, N);
    int slot = wb * (64 / SEGX) + threadIdx.x / SEGX;
    if (wb < 0 || slot >= nw) return;   // no barriers or shuffles below
    int s = order[slot];
    int cnt = min(bcnt[s], Kp);
    int pbase = packPoly[polyPack[s]];
    long long base = (long long)s * Kp;
    for (int f = lane; f < cnt; f += SEGX) {
        uint64_t inter = rKey[base + f], outer = rOuter[base + f];
        double la = 0.0;
        if (inter != outer) {
            long long cr[CYCLE_MAX];
            int cn = 1, k0 = 0;
            bool closed = cycle_walk<true>(inter, outer, base + f, pbase, Kp, bcnt, rKey, rOuter, cr, &cn, &k0);
            int lensN;
            bool rep;
            double a = lens_area(closed, cr, cn, k0, Er, &lensN, &rep);
            if (rep) la = a;
        }
        laRep[base + f] = la;
    }
}

// Squared lens areas, first pass (wavefront / polygon s, lane / record): the h sums of
// each record -- its exterior terms plus its chain of s's edges, measured from the same
// origin as in k_exterior_poly -- twice the area the record covers.
template <int SEGX>
__global__ __launch_bounds__(64) void k_lens_energy(int N, int Kp, const int* __restrict__ polyStart,
                                                   const int* __restrict__ polyPack,
                                                   const int* __restrict__ packPoly, const double2* __restrict__ pos,
                                                   const int* __restrict__ bcnt, const uint64_t* __restrict__ rKey,
                                                   const uint64_t* __restrict__ rOuter,
                                                   const float2* __restrict__ rTU, const float2* __restrict__ rUT,
                                                   double* __restrict__ Er, const int* __restrict__ order,
                                                   const int* __restrict__ nWork) {
    int lane = threadIdx.x & (SEGX - 1);
    int nw, wb = work_block(nWork, 64 / SEGX, nw, N);
    int slot = wb * (64 / SEGX) + threadIdx.x / SEGX;
    if (wb < 0 || slot >= nw) return;   // no barriers or shuffles below
    int s = order[slot];
    int cnt = min(bcnt[s], Kp);
    int pk = polyPack[s];
    int pbase = packPoly[pk];
    int st1 = polyStart[s], n1 = polyStart[s + 1] - st1;
    long long base = (long long)s * Kp;
    for (int f = lane; f < cnt; f += SEGX) {
        uint64_t inter = rKey[base + f], outer = rOuter[base + f];
        double E = 0.0;
        if (inter != outer) {
            int sjg = pbase + key_sj(inter);
            int st2 = polyStart[sjg], n2 = polyStart[sjg + 1] - st2;
            int iL = key_i(inter), jL = key_j(inter), kL = key_i(outer), lL = key_j(outer);
            int ziL = (iL + 1 == n1) ? 0 : iL + 1;
            int zjL = (jL + 1 == n2) ? 0 : jL + 1, zkL = (kL + 1 == n2) ? 0 : kL + 1;
            int zlL = (lL + 1 == n1) ? 0 : lL + 1;
            d2 pi = ldpos(pos, st1 + iL), pj = ldpos(pos, st2 + jL), pkk = ldpos(pos, st2 + kL);
            d2 pl = ldpos(pos, st1 + lL);
            d2 pzi = ldpos(pos, st1 + ziL), pzj = ldpos(pos, st2 + zjL), pzk = ldpos(pos, st2 + zkL);
            d2 pzl = ldpos(pos, st1 + zlL);
            d2 r1 = wrap2(mk(pzi.x - pi.x, pzi.y - pi.y));
            d2 r2 = wrap2(mk(pzk.x - pkk.x, pzk.y - pkk.y));
            double t1 = cross_param(pi, pzi, pj, pzj, (double)rTU[base + f].y);
            double t2 = cross_param(pkk, pzk, pl, pzl, (double)rUT[base + f].y);
            d2 fij = wrap2(mk(pi.x + t1 * r1.x, pi.y + t1 * r1.y));
            d2 fkl = wrap2(mk(pkk.x + t2 * r2.x, pkk.y + t2 * r2.y));
            d2 sp = fij;
            if (closed_cycle(inter, outer, pbase, Kp, bcnt, rKey, rOuter)) sp = ldpos(pos, polyStart[min(s, sjg)]);
            double e = 0.0;
            if (iL == lL) {
                e += hfun(fij, fkl, sp);
            } else {
                e += hfun(fij, pzi, sp);
                e += hfun(pl, fkl, sp);
            }
            double chain = 0.0;
            int lDist = (lL - iL + n1) % n1;
            for (int d = 1; d < lDist; ++d) {
                int m = (iL + d) % n1;
                int mn = (m + 1 == n1) ? 0 : m + 1;
                chain += hfun(ldpos(pos, st1 + m), ldpos(pos, st1 + mn), sp);
            }
            E = e + chain;
        }
        Er[base + f] = E;
    }
}

// One record (inter, outer, float parameters tuR/utR) of polygon s for k_exterior_poly: its 8
// exterior contributions and energy, the 4 acting on s staged in L at index idx (the record's
// place in key order), the 4 acting on sj written at the record's rank in sj's list.  ownRow:
// the record's row (the cycle walk of the area energies starts there).
template <int MODE>
__device__ __forceinline__ void exterior_record(uint64_t inter, uint64_t outer, float2 tuR, float2 utR, int idx,
                                                long long ownRow, int s, int pbase, int st1, int n1, int Kp,
                                                const int* __restrict__ polyStart, const double2* __restrict__ pos,
                                                const int* __restrict__ bcnt, const int* __restrict__ jcnt,
                                                const uint64_t* __restrict__ jKey, const uint64_t* __restrict__ rKey,
                                                const uint64_t* __restrict__ rOuter, double2* __restrict__ jC,
                                                int* __restrict__ jT, const double* __restrict__ Er,
                                                const double* __restrict__ laRep, const ExtLds& L) {
    constexpr int AREA = (MODE >= 1), SQ = (MODE >= 2), PAIR = (MODE == 3);
    const double sc = AREA ? 0.5 : 1.0;
    int sjg = pbase + key_sj(inter);
    int st2 = polyStart[sjg], n2 = polyStart[sjg + 1] - st2;
    d2 inc[8];
    int tg[8];
    double e = 0.0;
    d2 sp = mk(0.0, 0.0);
    double lensA = 0.0, scale = 0.0;   // scale: lens area, or with PAIR the pair's overlap area
    int lensN = 1;
    bool valid = (inter != outer);
    if (valid) {
        int iL = key_i(inter), jL = key_j(inter), kL = key_i(outer), lL = key_j(outer);
        int ziL = (iL + 1 == n1) ? 0 : iL + 1, zlL = (lL + 1 == n1) ? 0 : lL + 1;
        int zjL = (jL + 1 == n2) ? 0 : jL + 1, zkL = (kL + 1 == n2) ? 0 : kL + 1;
        d2 pi = ldpos(pos, st1 + iL), pj = ldpos(pos, st2 + jL), pkk = ldpos(pos, st2 + kL);
        d2 pl = ldpos(pos, st1 + lL);
        d2 pzi = ldpos(pos, st1 + ziL), pzj = ldpos(pos, st2 + zjL), pzk = ldpos(pos, st2 + zkL);
        d2 pzl = ldpos(pos, st1 + zlL);
        d2 r1 = wrap2(mk(pzi.x - pi.x, pzi.y - pi.y));
        d2 r2 = wrap2(mk(pzk.x - pkk.x, pzk.y - pkk.y));
        double t1 = (double)tuR.y, t2 = (double)utR.y;
        if (AREA) {   // exact crossing parameters at the current positions
            t1 = cross_param(pi, pzi, pj, pzj, t1);
            t2 = cross_param(pkk, pzk, pl, pzl, t2);
        }
        d2 fij = wrap2(mk(pi.x + t1 * r1.x, pi.y + t1 * r1.y));
        d2 fkl = wrap2(mk(pkk.x + t2 * r2.x, pkk.y + t2 * r2.y));
        sp = fij;
        bool closed = false;
        long long cr[SQ ? CYCLE_MAX : 1];
        int cn = 1, k0 = 0;
        if (AREA) {
            // a record on a closed cycle of exit links (a closed lens boundary) measures
            // from one origin per polygon pair: the first vertex of the lower polygon
            closed = cycle_walk<SQ>(inter, outer, ownRow, pbase, Kp, bcnt, rKey, rOuter, cr, &cn, &k0);
            if (closed) sp = ldpos(pos, polyStart[min(s, sjg)]);
        }
        if (SQ) {
            bool rep;
            lensA = lens_area(closed, cr, cn, k0, Er, &lensN, &rep);
            scale = PAIR ? pair_area(s, sjg, pbase, Kp, bcnt, rKey, laRep) : lensA;
        }
        double dfa[16], dfb[16];
        if (iL == lL) {
            e += hfun(fij, fkl, sp);
            d2 ga, gb;
            g12(fij, fkl, sp, ga, gb);
            getDf(pi, pzi, pj, pzj, dfa);
            getDf(pkk, pzk, pi, pzi, dfb);
            inc[0] = dfdot(dfa, 0, ga);
            inc[1] = dfdot(dfa, 2, ga);
            inc[4] = dfdot(dfa, 4, ga);
            inc[5] = dfdot(dfa, 6, ga);
            inc[6] = dfdot(dfb, 0, gb);
            inc[7] = dfdot(dfb, 2, gb);
            inc[2] = dfdot(dfb, 4, gb);
            inc[3] = dfdot(dfb, 6, gb);
        } else {
            e += hfun(fij, pzi, sp);
            e += hfun(pl, fkl, sp);
            d2 g1a, g2a;
            g12(fij, pzi, sp, g1a, g2a);
            getDf(pi, pzi, pj, pzj, dfa);
            getDf(pkk, pzk, pl, pzl, dfb);
            inc[0] = dfdot(dfa, 0, g1a);
            inc[4] = dfdot(dfa, 4, g1a);
            d2 dd = dfdot(dfa, 2, g1a);
            inc[1] = mk(dd.x + g2a.x, dd.y + g2a.y);
            inc[5] = dfdot(dfa, 6, g1a);
            d2 g1b, g2b;
            g12(pl, fkl, sp, g1b, g2b);
            d2 de = dfdot(dfb, 4, g2b);
            inc[2] = mk(g1b.x + de.x, g1b.y + de.y);
            inc[6] = dfdot(dfb, 0, g2b);
            inc[3] = dfdot(dfb, 6, g2b);
            inc[7] = dfdot(dfb, 2, g2b);
        }
        tg[0] = iL; tg[1] = ziL; tg[2] = lL; tg[3] = zlL;
        tg[4] = jL; tg[5] = zjL; tg[6] = kL; tg[7] = zkL;
        L.il[idx] = iL | (lL << 16);
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) { tg[q] = -1; inc[q] = mk(0.0, 0.0); }
        L.il[idx] = -1;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        double cx = -inc[q].x * sc, cy = -inc[q].y * sc;
        if (SQ) { cx *= scale; cy *= scale; }
        inc[q] = mk(cx, cy);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        L.c[4 * idx + q] = make_double2(inc[q].x, inc[q].y);
        L.t[4 * idx + q] = tg[q];
    }
    // energy of the record: its h terms (the polygon sums them with its interior
    // terms and halves), or with SQ its share scale*A/(2 lensN) of the lens's part of
    // the energy (A^2/2, or with PAIR A_pair*A/2), doubled so that the same halving
    // gives it exactly
    L.e[idx] = SQ ? 2.0 * ((scale * lensA) / (2.0 * lensN)) : e;
    L.la[idx] = scale;
    L.sp[idx] = make_double2(sp.x, sp.y);
    // sj side: at this record's rank among the records whose sj is sjg (key order)
    long long bj = (long long)sjg * Kp;
    int cj = min(jcnt[sjg], Kp);
    int r2 = count_keys_below(jKey + bj, cj, [&](uint64_t k2) { return k2 < inter; });
    if (r2 < Kp) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            jC[(bj + r2) * 4 + q] = make_double2(inc[4 + q].x, inc[4 + q].y);
            jT[(bj + r2) * 4 + q] = tg[4 + q];
        }
    }
}

// The vertex part of k_exterior_poly for polygon s (after its records are staged in L, in key
// order): per vertex the contributions acting on it in record order and the interior chain
// terms; the polygon's record and interior energy into accR[s].
template <int MODE, int SEGX>
__device__ __forceinline__ void exterior_vertices(int lane, int segBase, bool has, int s, int cnt, int st1, int n1,
                                                  const double2* __restrict__ pos, const ExtLds& L, double* leBuf,
                                                  double2* __restrict__ F1, double2* __restrict__ LMo,
                                                  double2* __restrict__ LNo, double* __restrict__ accR) {
    constexpr int AREA = (MODE >= 1), SQ = (MODE >= 2);
    __syncthreads();
    // polygon energy, first part: record energies in record order
    double acc = 0.0;
    if (lane == 0)
        for (int f = 0; f < cnt; ++f) acc += L.e[f];
    int nMax = n1;   // the wavefront's largest polygon sets the trip count
#pragma unroll
    for (int o = 32; o >= SEGX; o >>= 1) nMax = max(nMax, __shfl_xor(nMax, o));
    // per vertex: contributions acting on s (record order), interior chain terms
    for (int m0 = 0; m0 < nMax; m0 += SEGX) {
        int mLoc = m0 + lane;
        double le = 0.0;
        if (mLoc < n1) {
            d2 F = mk(0.0, 0.0);
            for (int f = 0; f < cnt; ++f) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (L.t[4 * f + q] == mLoc) {
                        double2 cc = L.c[4 * f + q];
                        F.x += cc.x; F.y += cc.y;
                    }
            }
            int nxL = (mLoc + 1 == n1) ? 0 : mLoc + 1, pmLoc = (mLoc == 0) ? n1 - 1 : mLoc - 1;
            d2 Pm = ldpos(pos, st1 + mLoc), Pn = ldpos(pos, st1 + nxL), Pp = ldpos(pos, st1 + pmLoc);
            d2 LM = mk(0.0, 0.0), LN = mk(0.0, 0.0);
            for (int f = 0; f < cnt; ++f) {
                int il = L.il[f];
                if (il < 0) continue;
                int iLoc = il & 0xFFFF, lLoc = il >> 16;
                int lDist = (lLoc - iLoc + n1) % n1;
                int dm = (mLoc - iLoc + n1) % n1;
                int dp = (pmLoc - iLoc + n1) % n1;
                bool cm = (dm > 0 && dm < lDist), cp = (dp > 0 && dp < lDist);
                if (cm || cp) {
                    double2 s2 = L.sp[f];
                    d2 spr = mk(s2.x, s2.y);
                    d2 g1, g2;
                    if (cm) {
                        if (!SQ) le += hfun(Pm, Pn, spr);
                        g12(Pm, Pn, spr, g1, g2);
                        if (AREA) { g1.x *= 0.5; g1.y *= 0.5; }
                        if (SQ) { g1.x *= L.la[f]; g1.y *= L.la[f]; }
                        LM.x -= g1.x; LM.y -= g1.y;
                    }
                    if (cp) {
                        g12(Pp, Pm, spr, g1, g2);
                        if (AREA) { g2.x *= 0.5; g2.y *= 0.5; }
                        if (SQ) { g2.x *= L.la[f]; g2.y *= L.la[f]; }
                        LN.x -= g2.x; LN.y -= g2.y;
                    }
                }
            }
            F1[st1 + mLoc] = make_double2(F.x, F.y);
            LMo[st1 + mLoc] = make_double2(LM.x, LM.y);
            LNo[st1 + mLoc] = make_double2(LN.x, LN.y);
        }
        leBuf[threadIdx.x] = le;
        __syncthreads();
        if (lane == 0) {
            int cl = min(SEGX, n1 - m0);
            for (int q = 0; q < cl; ++q) acc += leBuf[segBase + q];
        }
        __syncthreads();
    }
    if (has && lane == 0) accR[s] = acc;
}

// MODE 0: reference energy; 1: area energy; 2: squared lens areas; 3: squared pair areas.
// One wavefront holds 64 / SEGX polygons, SEGX lanes each: lanes take the polygon's
// records, then its vertices, SEGX at a time (a polygon has a few records, so a
// wavefront per polygon would leave most lanes idle).  Every sum keeps its order.
// GLOB: the per-polygon record staging lives in global memory (gstage, ext_lds_stride(Kp)
// bytes per polygon slot of the launch) instead of LDS, for buckets too large for LDS; the
// workgroup barriers order it as they order LDS (one wavefront per workgroup, one CU).
template <int MODE, int SEGX, bool GLOB = false>
__global__ __launch_bounds__(64) void k_exterior_poly(
    int N, int Kp, const int* __restrict__ polyStart, const int* __restrict__ polyPack,
    const int* __restrict__ packPoly, const double2* __restrict__ pos, const int* __restrict__ bcnt,
    const int* __restrict__ jcnt, const uint64_t* __restrict__ jKey, const uint64_t* __restrict__ rKey,
    const uint64_t* __restrict__ rOuter, const float2* __restrict__ rTU, const float2* __restrict__ rUT,
    double2* __restrict__ jC, int* __restrict__ jT, double2* __restrict__ F1, double2* __restrict__ LMo,
    double2* __restrict__ LNo, double* __restrict__ accR, const double* __restrict__ Er,
    const double* __restrict__ laRep, const int* __restrict__ order, const int* __restrict__ nWork,
    char* __restrict__ gstage) {
    constexpr int G = 64 / SEGX;
    extern __shared__ char smem[];
    __shared__ double leBuf[64];
    int lane = threadIdx.x & (SEGX - 1), g = threadIdx.x / SEGX, segBase = g * SEGX;
    ExtLds L = GLOB ? ext_lds(gstage + ((size_t)blockIdx.x * G + g) * ext_lds_stride(Kp), Kp)
                    : ext_lds(smem + g * ext_lds_stride(Kp), Kp);
    int nw, wb = work_block(nWork, G, nw, N);
    if (wb < 0) return;   // the whole workgroup
    int slot = wb * G + g;
    bool has = slot < nw;
    int s = has ? order[slot] : 0;
    int cnt = has ? min(bcnt[s], Kp) : 0;
    int pk = polyPack[s];
    int pbase = packPoly[pk];
    int st1 = polyStart[s], n1 = has ? polyStart[s + 1] - st1 : 0;
    long long base = (long long)s * Kp;
    for (int f = lane; f < cnt; f += SEGX)
        exterior_record<MODE>(rKey[base + f], rOuter[base + f], rTU[base + f], rUT[base + f], f, base + f, s, pbase,
                              st1, n1, Kp, polyStart, pos, bcnt, jcnt, jKey, rKey, rOuter, jC, jT, Er, laRep, L);
    exterior_vertices<MODE, SEGX>(lane, segBase, has, s, cnt, st1, n1, pos, L, leBuf, F1, LMo, LNo, accR);
}

// Pairing and exterior forces in one launch for the reference energy (MODE 0), where the
// exterior terms of polygon s need only the exits of its own records: the wavefront that pairs
// the records of s (pair_record, as k_pair_poly) stages their contributions at their ranks
// (exterior_record, as k_exterior_poly<0>) and goes on to the vertex part, with no kernel
// boundary and no round trip through the record rows (which are written all the same, for the
// getters).  The results are those of the two kernels bit for bit.
template <int SEGX>
__global__ __launch_bounds__(64) void k_pair_exterior(
    int N, int Kp, const int* __restrict__ polyStart, const int* __restrict__ polyPack,
    const int* __restrict__ packPoly, const double2* __restrict__ pos, const int* __restrict__ bcnt,
    const uint64_t* __restrict__ bKey, const float2* __restrict__ bTU,
    const unsigned long long* __restrict__ packMaxKey, long long pmkPitch, const int* __restrict__ jcnt,
    const uint64_t* __restrict__ jKey, uint64_t* __restrict__ rKey, float2* __restrict__ rTU,
    uint64_t* __restrict__ rOuter, float2* __restrict__ rUT, double2* __restrict__ jC, int* __restrict__ jT,
    double2* __restrict__ F1, double2* __restrict__ LMo, double2* __restrict__ LNo, double* __restrict__ accR,
    const int* __restrict__ order, const int* __restrict__ nWork, int* __restrict__ scal) {
    constexpr int G = 64 / SEGX;
    extern __shared__ char smem[];
    __shared__ double leBuf[64];
    int lane = threadIdx.x & (SEGX - 1), g = threadIdx.x / SEGX, segBase = g * SEGX;
    ExtLds L = ext_lds(smem + g * ext_lds_stride(Kp), Kp);
    int nw, wb = work_block(nWork, G, nw, N);
    if (wb < 0) return;   // the whole workgroup
    int slot = wb * G + g;
    bool has = slot < nw;
    int s = has ? order[slot] : 0;
    int cnt = has ? min(bcnt[s], Kp) : 0;
    int pk = polyPack[s];
    int pbase = packPoly[pk];
    int st1 = polyStart[s], n1 = has ? polyStart[s + 1] - st1 : 0;
    long long base = (long long)s * Kp;
    uint64_t mxk = 0;   // the maximum of the packing's partial maxima, across the segment
    for (int k = lane; k < PMK_SLOTS; k += SEGX) {
        uint64_t v = packMaxKey[k * pmkPitch + pk];
        mxk = (v > mxk) ? v : mxk;
    }
#pragma unroll
    for (int o = SEGX / 2; o >= 1; o >>= 1) {
        uint64_t y = __shfl_xor(mxk, o, SEGX);
        mxk = (y > mxk) ? y : mxk;
    }
    for (int f = lane; f < cnt; f += SEGX) {
        PairRec r = pair_record<0>(f, cnt, base, s, pbase, st1, n1, mxk, Kp, polyStart, bcnt, bKey, bTU, nullptr,
                                   scal);
        rKey[base + r.rank] = r.key;
        rTU[base + r.rank] = r.tu;
        rOuter[base + r.rank] = r.outer;
        rUT[base + r.rank] = r.ut;
        exterior_record<0>(r.key, r.outer, r.tu, r.ut, r.rank, base + r.rank, s, pbase, st1, n1, Kp, polyStart, pos,
                           bcnt, jcnt, jKey, rKey, rOuter, jC, jT, nullptr, nullptr, L);
    }
    exterior_vertices<0, SEGX>(lane, segBase, has, s, cnt, st1, n1, pos, L, leBuf, F1, LMo, LNo, accR);
}

// ---------------------------------------------------------------- per-vertex gather
// GEOM = 1: edge lengths and the area are computed here (reference summation order,
// kernels.cuh:295-333) and written out; GEOM = 0: they are read (updateForceEnergy uses
// the arrays of the last updatePolygonGeometry).  PROJ = 1: the force of each polygon is
// projected (n <= SEG) before it is stored.
template <int SEG, int GEOM, int PROJ>
__global__ __launch_bounds__(BLK) void k_gather_poly(
    int N, int Kp, int overlap, int edgeOn, int areaOn, const int* __restrict__ polyStart,
    const double2* __restrict__ pos, const int* __restrict__ jcnt, const double2* __restrict__ jC,
    const int* __restrict__ jT, const double2* __restrict__ F1, const double2* __restrict__ LMi,
    const double2* __restrict__ LNi, const double* __restrict__ accR, double* __restrict__ edgeLengths,
    double* __restrict__ areas, const double* __restrict__ targetEdgeV, const double* __restrict__ targetAreas,
    double stiffness, double compressibility, double2* __restrict__ force, double* __restrict__ epoly,
    const int* __restrict__ order, const int* __restrict__ nWork, double perimK, double2* __restrict__ fext) {
    int sub = threadIdx.x & (SEG - 1);
    int nw, wb = work_block(nWork, BLK / SEG, nw, N);
    int si = (wb * BLK + threadIdx.x) / SEG;
    if (wb < 0 || si >= nw) return;   // whole segments leave together
    int s = order[si];
    int st = polyStart[s], n = polyStart[s + 1] - st;
    // area in the reference's order (needed by the area spring and written when GEOM)
    double A = 0.0;
    if (GEOM) {
        d2 p0 = ldpos(pos, st);
        for (int k0 = 0; k0 < n; k0 += SEG) {
            int k = k0 + sub;
            double term = 0.0;
            if (k < n) {
                d2 pk = ldpos(pos, st + k), pn = ldpos(pos, st + ((k + 1 == n) ? 0 : k + 1));
                double x1a = wrap(pk.x - p0.x);
                double dy2 = wrap(pn.y - pk.y);
                double x2 = wrap(pn.x - p0.x);
                term = dy2 * (x2 + x1a) / 2.0;
            }
            int cl = min(SEG, n - k0);
            for (int q = 0; q < cl; ++q) A += __shfl(term, q, SEG);
        }
        if (sub == 0) areas[s] = A;
    } else if (areaOn) {
        A = areas[s];
    }
    // perimeter spring (no counterpart in the reference): perimeter and its target (the sum
    // of the edge targets), both summed in vertex order
    bool perimOn = perimK != 0.0;
    double Pp = 0.0, P0 = 0.0;
    if (perimOn) {
        for (int k0 = 0; k0 < n; k0 += SEG) {
            int k = k0 + sub;
            double lk = 0.0, tk = 0.0;
            if (k < n) {
                if (GEOM) {
                    d2 pk = ldpos(pos, st + k), pn = ldpos(pos, st + ((k + 1 == n) ? 0 : k + 1));
                    double dx = wrap(pn.x - pk.x), dy = wrap(pn.y - pk.y);
                    lk = sqrt(dx * dx + dy * dy);
                } else {
                    lk = edgeLengths[st + k];
                }
                tk = targetEdgeV[st + k];
            }
            int cl = min(SEG, n - k0);
            for (int q = 0; q < cl; ++q) {
                Pp += __shfl(lk, q, SEG);
                P0 += __shfl(tk, q, SEG);
            }
        }
    }
    double e = 0.5 * ((overlap) ? accR[s] : 0.0);
    int jn = overlap ? min(jcnt[s], Kp) : 0;
    long long bj = (long long)s * Kp;
    for (int m0 = 0; m0 < n; m0 += SEG) {
        int mLoc = m0 + sub;
        bool live = mLoc < n;
        double ee = 0.0;
        d2 F = mk(0.0, 0.0);
        d2 Pm = mk(0.0, 0.0);
        if (live) {
            int m = st + mLoc;
            int nx = st + ((mLoc + 1 == n) ? 0 : mLoc + 1), pm = st + ((mLoc == 0) ? n - 1 : mLoc - 1);
            Pm = ldpos(pos, m);
            if (overlap) {
                double2 f1 = F1[m];
                F = mk(f1.x, f1.y);
                for (int r0 = 0; r0 < jn; r0 += 4) {   // targets of 4 records loaded at once
                    int4 t4[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        t4[k] = (r0 + k < jn) ? ((const int4*)jT)[bj + r0 + k] : make_int4(-1, -1, -1, -1);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        long long e = (bj + r0 + k) * 4;
                        if (t4[k].x == mLoc) { double2 cc = jC[e + 0]; F.x += cc.x; F.y += cc.y; }
                        if (t4[k].y == mLoc) { double2 cc = jC[e + 1]; F.x += cc.x; F.y += cc.y; }
                        if (t4[k].z == mLoc) { double2 cc = jC[e + 2]; F.x += cc.x; F.y += cc.y; }
                        if (t4[k].w == mLoc) { double2 cc = jC[e + 3]; F.x += cc.x; F.y += cc.y; }
                    }
                }
                double2 LM = LMi[m], LN = LNi[m];
                if (LM.x != 0.0 || LM.y != 0.0) { F.x += LM.x; F.y += LM.y; }
                if (LN.x != 0.0 || LN.y != 0.0) { F.x += LN.x; F.y += LN.y; }
            }
            if (fext) fext[m] = make_double2(F.x, F.y);
            double l = 0.0, prvl = 0.0;
            if (GEOM || edgeOn || perimOn) {
                if (GEOM) {
                    d2 Pn = ldpos(pos, nx), Pp = ldpos(pos, pm);
                    double dx = wrap(Pn.x - Pm.x), dy = wrap(Pn.y - Pm.y);
                    l = sqrt(dx * dx + dy * dy);
                    double dxp = wrap(Pm.x - Pp.x), dyp = wrap(Pm.y - Pp.y);
                    prvl = sqrt(dxp * dxp + dyp * dyp);
                    edgeLengths[m] = l;
                } else {
                    l = edgeLengths[m];
                    prvl = edgeLengths[pm];
                }
            }
            if (edgeOn) {
                double l0 = targetEdgeV[m], l0prv = targetEdgeV[pm];
                d2 Pn = ldpos(pos, nx), Pp = ldpos(pos, pm);
                d2 a = mk(Pm.x - Pn.x, Pm.y - Pn.y), b = mk(Pp.x - Pm.x, Pp.y - Pm.y);
                a.x += 1.5; a.y += 1.5; b.x += 1.5; b.y += 1.5;
                a.x = shift_into_unit(a.x); a.y = shift_into_unit(a.y);
                b.x = shift_into_unit(b.x); b.y = shift_into_unit(b.y);
                a.x -= 0.5; a.y -= 0.5; b.x -= 0.5; b.y -= 0.5;
                double c1 = (l > 1e-14) ? (1.0 - l0 / l) : 0.0;
                double c2 = (prvl > 1e-14) ? (1.0 - l0prv / prvl) : 0.0;
                ee = stiffness * (l - l0) * (l - l0) / 2.0;
                d2 lf;
                lf.x = -(c1 * a.x - c2 * b.x) * stiffness;
                lf.y = -(c1 * a.y - c2 * b.y) * stiffness;
                if (lf.x != 0.0 || lf.y != 0.0) { F.x += lf.x; F.y += lf.y; }
            }
            if (areaOn) {
                double A0 = targetAreas[s];
                if (!(A0 < 1e-14)) {
                    double coeff = 0.5 * compressibility * (A - A0);
                    d2 Pn = ldpos(pos, nx), Pp = ldpos(pos, pm);
                    double fx = -coeff * wrap(Pn.y - Pp.y);
                    double fy = -coeff * wrap(Pp.x - Pn.x);
                    F.x += fx; F.y += fy;
                }
            }
            if (perimOn) {   // -kP (P - P0) (t_{m-1} - t_m), t the unit edge vectors
                d2 Pn = ldpos(pos, nx), Pq = ldpos(pos, pm);
                double dx = wrap(Pn.x - Pm.x), dy = wrap(Pn.y - Pm.y);
                double px = wrap(Pm.x - Pq.x), py = wrap(Pm.y - Pq.y);
                double tx = (l > 1e-14) ? dx / l : 0.0, ty = (l > 1e-14) ? dy / l : 0.0;
                double qx = (prvl > 1e-14) ? px / prvl : 0.0, qy = (prvl > 1e-14) ? py / prvl : 0.0;
                double coeff = perimK * (Pp - P0);
                F.x += -coeff * (qx - tx);
                F.y += -coeff * (qy - ty);
            }
        }
        if (edgeOn) {   // edge energies in vertex order
            int cl = min(SEG, n - m0);
            for (int q = 0; q < cl; ++q) e += __shfl(ee, q, SEG);
        }
        double2 fo = make_double2(F.x, F.y);
        if (PROJ && n <= SEG && n >= 3) fo = seg_project_force<SEG, PROJ == 2>(sub, n, Pm, fo);
        if (live) force[st + mLoc] = fo;
    }
    if (areaOn) {
        double A0 = targetAreas[s];
        if (!(A0 < 1e-14)) e += compressibility * (A - A0) * (A - A0) / 2.0;
    }
    if (perimOn) e += perimK * (Pp - P0) * (Pp - P0) / 2.0;
    if (sub == 0) epoly[s] = e;
}

// ---------------------------------------------------------------- host side
int grow_buckets(pcp_ctx* c, int need) {
    // record keys hold 16-bit indices; a bucket never needs more than 2^20 slots
    if (need > (1 << 20)) {
        set_error("more than 2^20 crossings on one polygon");
        return PCP_ERR_CAPACITY;
    }
    int Kp = std::min(std::max(need + need / 4, 2 * c->Kp), 1 << 20);
    return alloc_buckets(c, Kp);
}

int alloc_buckets(pcp_ctx* c, int Kp) {
    dfree(c->dBKey); dfree(c->dBTU); dfree(c->dJKey); dfree(c->dRKey); dfree(c->dROuter);
    dfree(c->dRTU); dfree(c->dRUT); dfree(c->dJC); dfree(c->dJT); dfree(c->dEr); dfree(c->dLaRep);
    long long n = (long long)std::max(c->N, 1) * Kp;
    PCP_TRY(dalloc(&c->dBKey, n)); PCP_TRY(dalloc(&c->dBTU, n)); PCP_TRY(dalloc(&c->dJKey, n));
    PCP_TRY(dalloc(&c->dRKey, n)); PCP_TRY(dalloc(&c->dROuter, n)); PCP_TRY(dalloc(&c->dRTU, n));
    PCP_TRY(dalloc(&c->dRUT, n)); PCP_TRY(dalloc(&c->dJC, 4 * n)); PCP_TRY(dalloc(&c->dJT, 4 * n));
    PCP_TRY(dalloc(&c->dEr, n));
    PCP_TRY(dalloc(&c->dLaRep, n));
    c->Kp = Kp;
    return PCP_OK;
}

template <int EXACT, int SEGX>
static int launch_pair_w(pcp_ctx* c, const double2* pairPos) {
    constexpr int G = 64 / SEGX;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_poly<EXACT, SEGX>), dim3((c->N + G - 1) / G), dim3(64), 0, c->stream,
                       c->N, c->Kp, c->dPolyStart, c->dPolyPack, c->dPackPoly, c->dBcnt, c->dBKey, c->dBTU,
                       c->dPackMaxKey, c->pmkPitch, pairPos, c->dRKey, c->dRTU, c->dROuter, c->dRUT, work_list(c),
                       work_count(c), c->dScal);
    PCP_LAUNCHED();
    return PCP_OK;
}
template <int EXACT>
static int launch_pair(pcp_ctx* c, const double2* pairPos) {
    int slot = -1;
    PCP_TRY(prof_begin(c, EXACT ? "k_pair_poly<1>" : "k_pair_poly<0>", &slot));
    if (c->segPair == 64) PCP_TRY((launch_pair_w<EXACT, 64>(c, pairPos)));
    else if (c->segPair == 32) PCP_TRY((launch_pair_w<EXACT, 32>(c, pairPos)));
    else PCP_TRY((launch_pair_w<EXACT, 16>(c, pairPos)));
    return prof_end(c, slot);
}

// deferOK: the caller evaluates the forces next (run_forces), so the reference pairing may be
// left to the fused launch k_pair_exterior (when a polygon's records fit the LDS stage)
int run_records(pcp_ctx* c, const double2* pairPos, bool deferOK) {
    c->pairDeferred = false;
    if (c->N == 0) return PCP_OK;
    if (pairPos) return launch_pair<1>(c, pairPos);
    if (deferOK && ext_lds_stride(c->Kp) <= EXT_LDS_MAX) {
        c->pairDeferred = true;
        return PCP_OK;
    }
    return launch_pair<0>(c, pairPos);
}

template <int SEG, int GEOM, int PROJ>
static int launch_gather(pcp_ctx* c, int overlap, int edgeOn, int areaOn, const double2* pos, double* edgeLengths,
                         double* areas, const double* targetEdgeV, const double* targetAreas, double stiffness,
                         double compressibility, double2* force, double perimK, double2* fext) {
    int grid = (c->N + BLK / SEG - 1) / (BLK / SEG);
    int slot = -1;
    PCP_TRY(prof_begin(c, "k_gather_poly", &slot));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gather_poly<SEG, GEOM, PROJ>), dim3(grid), dim3(BLK), 0, c->stream, c->N,
                       c->Kp, overlap, edgeOn, areaOn, c->dPolyStart, pos, c->dJcnt, c->dJC, c->dJT, c->dF1, c->dLM,
                       c->dLN, c->dAccR, edgeLengths, areas, targetEdgeV, targetAreas, stiffness, compressibility,
                       force, c->dEpoly, work_list(c), work_count(c), perimK, fext);
    PCP_LAUNCHED();
    PCP_TRY(prof_end(c, slot));
    return PCP_OK;
}

template <int GEOM, int PROJ>
static int dispatch_gather(pcp_ctx* c, int overlap, int edgeOn, int areaOn, const double2* pos,
                           double* edgeLengths, double* areas, const double* targetEdgeV,
                           const double* targetAreas, double stiffness, double compressibility, double2* force,
                           double perimK, double2* fext) {
#define PCP_GATHER(W) \
    return launch_gather<W, GEOM, PROJ>(c, overlap, edgeOn, areaOn, pos, edgeLengths, areas, targetEdgeV, \
                                        targetAreas, stiffness, compressibility, force, perimK, fext)
    switch (c->seg) {
    case 4: PCP_GATHER(4);
    case 8: PCP_GATHER(8);
    case 16:
        if (PROJ && c->row16)   // PROJ 2: the projection's lane moves as DPP row moves
            return launch_gather<16, GEOM, 2>(c, overlap, edgeOn, areaOn, pos, edgeLengths, areas, targetEdgeV,
                                              targetAreas, stiffness, compressibility, force, perimK, fext);
        PCP_GATHER(16);
    case 32: PCP_GATHER(32);
    default: PCP_GATHER(64);
    }
#undef PCP_GATHER
}

// Forces and per-packing energies.  geom: compute edge lengths / areas here and write
// them (fused evaluation); else read them.  project: project the force of every polygon.
// k_exterior_poly with as many polygons per wavefront as the LDS allows (64 KiB per
// workgroup): 4 (16 lanes each) for buckets up to 141 records, then 2, then 1
template <int MODE, int SEGX, bool GLOB = false>
static int launch_exterior_w(pcp_ctx* c, const double2* pos) {
    constexpr int G = 64 / SEGX;
    size_t lds = GLOB ? 0 : G * ext_lds_stride(c->Kp);
    int grid = (c->N + G - 1) / G;
    if (GLOB) {   // the staging buffer: one polygon slot per wavefront segment of the launch
        size_t need = (size_t)grid * G * ext_lds_stride(c->Kp);
        if (need > c->extStageBytes) {
            dfree(c->dExtStage);
            c->extStageBytes = 0;
            PCP_TRY(dalloc(&c->dExtStage, need));
            c->extStageBytes = need;
        }
    }
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_exterior_poly<MODE, SEGX, GLOB>), dim3(grid), dim3(64), lds, c->stream,
                       c->N, c->Kp, c->dPolyStart, c->dPolyPack, c->dPackPoly, pos, c->dBcnt, c->dJcnt, c->dJKey,
                       c->dRKey, c->dROuter, c->dRTU, c->dRUT, c->dJC, c->dJT, c->dF1, c->dLM, c->dLN, c->dAccR,
                       c->dEr, c->dLaRep, work_list(c), work_count(c), GLOB ? c->dExtStage : (char*)nullptr);
    PCP_LAUNCHED();
    return PCP_OK;
}
template <int MODE>
static int launch_exterior_m(pcp_ctx* c, const double2* pos) {
    int slot = -1;
    static const char* names[4] = {"k_exterior_poly<0>", "k_exterior_poly<1>", "k_exterior_poly<2>",
                                   "k_exterior_poly<3>"};
    PCP_TRY(prof_begin(c, names[MODE], &slot));
    size_t st = ext_lds_stride(c->Kp);
    int w = c->segExt;
    if (w == 16 && 4 * st > EXT_LDS_MAX) w = 0;   // a chosen width the LDS does not fit: the default rule
    if (w == 32 && 2 * st > EXT_LDS_MAX) w = 0;
    if (w == 0) w = (4 * st <= EXT_LDS_MAX) ? 16 : ((2 * st <= EXT_LDS_MAX) ? 32 : 64);
    if (st > EXT_LDS_MAX) PCP_TRY((launch_exterior_w<MODE, 64, true>(c, pos)));
    else if (w == 16) PCP_TRY((launch_exterior_w<MODE, 16>(c, pos)));
    else if (w == 32) PCP_TRY((launch_exterior_w<MODE, 32>(c, pos)));
    else PCP_TRY((launch_exterior_w<MODE, 64>(c, pos)));
    return prof_end(c, slot);
}
template <int SEGX>
static int launch_pair_exterior_w(pcp_ctx* c, const double2* pos) {
    constexpr int G = 64 / SEGX;
    size_t lds = G * ext_lds_stride(c->Kp);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_exterior<SEGX>), dim3((c->N + G - 1) / G), dim3(64), lds, c->stream,
                       c->N, c->Kp, c->dPolyStart, c->dPolyPack, c->dPackPoly, pos, c->dBcnt, c->dBKey, c->dBTU,
                       c->dPackMaxKey, c->pmkPitch, c->dJcnt, c->dJKey, c->dRKey, c->dRTU, c->dROuter, c->dRUT,
                       c->dJC, c->dJT, c->dF1, c->dLM, c->dLN, c->dAccR, work_list(c), work_count(c), c->dScal);
    PCP_LAUNCHED();
    return PCP_OK;
}
// k_pair_exterior with the exterior kernel's lanes per polygon (the records fit the LDS stage)
static int launch_pair_exterior(pcp_ctx* c, const double2* pos) {
    int slot = -1;
    PCP_TRY(prof_begin(c, "k_pair_exterior", &slot));
    size_t st = ext_lds_stride(c->Kp);
    int w = c->segExt;
    if (w == 16 && 4 * st > EXT_LDS_MAX) w = 0;
    if (w == 32 && 2 * st > EXT_LDS_MAX) w = 0;
    if (w == 0) w = (4 * st <= EXT_LDS_MAX) ? 16 : ((2 * st <= EXT_LDS_MAX) ? 32 : 64);
    if (w == 16) PCP_TRY(launch_pair_exterior_w<16>(c, pos));
    else if (w == 32) PCP_TRY(launch_pair_exterior_w<32>(c, pos));
    else PCP_TRY(launch_pair_exterior_w<64>(c, pos));
    return prof_end(c, slot);
}
static int launch_exterior(pcp_ctx* c, int mode, const double2* pos) {
    if (mode == 3) return launch_exterior_m<3>(c, pos);
    if (mode == 2) return launch_exterior_m<2>(c, pos);
    if (mode == 1) return launch_exterior_m<1>(c, pos);
    return launch_exterior_m<0>(c, pos);
}

static int launch_lens(pcp_ctx* c, const double2* pos) {
#define PCP_LENS(W)                                                                                           \
    PCP_LAUNCH(c, k_lens_energy<W>, dim3((c->N + 64 / W - 1) / (64 / W)), dim3(64), c->N, c->Kp, c->dPolyStart, \
               c->dPolyPack, c->dPackPoly, pos, c->dBcnt, c->dRKey, c->dROuter, c->dRTU, c->dRUT, c->dEr,       \
               work_list(c), work_count(c))
    if (c->segLens == 64) PCP_LENS(64);
    else if (c->segLens == 32) PCP_LENS(32);
    else PCP_LENS(16);
#undef PCP_LENS
    return PCP_OK;
}

// representative lens areas of the squared pair energy (the lens kernel's schedule)
static int launch_lens_rep(pcp_ctx* c) {
#define PCP_LENS_REP(W)                                                                                     \
    PCP_LAUNCH(c, k_lens_rep<W>, dim3((c->N + 64 / W - 1) / (64 / W)), dim3(64), c->N, c->Kp, c->dPolyPack, \
               c->dPackPoly, c->dBcnt, c->dRKey, c->dROuter, c->dEr, c->dLaRep, work_list(c),     \
               work_count(c))
    if (c->segLens == 64) PCP_LENS_REP(64);
    else if (c->segLens == 32) PCP_LENS_REP(32);
    else PCP_LENS_REP(16);
#undef PCP_LENS_REP
    return PCP_OK;
}

// Autotuner: time every candidate lane-segment width of the per-record kernels on the
// context's current buffers (filled by one evaluation of st), `reps` back-to-back launches
// each between two events, and keep the fastest.  All candidates give identical results.
static int time_launches(pcp_ctx* c, int reps, double* us, const std::function<int()>& launch) {
    hipEvent_t a, b;
    PCP_HIP(hipEventCreate(&a));
    PCP_HIP(hipEventCreate(&b));
    int rc = launch();   // warm-up
    if (rc == PCP_OK) {
        PCP_HIP(hipEventRecord(a, c->stream));
        for (int r = 0; r < reps && rc == PCP_OK; ++r) rc = launch();
        PCP_HIP(hipEventRecord(b, c->stream));
        PCP_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        PCP_HIP(hipEventElapsedTime(&ms, a, b));
        *us = 1e3 * ms / reps;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

int autotune(pcp_ctx* c, const pcp_state* st, int reps, double* timesUs) {
    const int W[3] = {16, 32, 64};
    double t[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int mt = st->model & 0xFF;
    bool overlap = (mt == PCP_NORMAL || mt == PCP_HYBRID);
    if (overlap && c->N > 0) {
        PCP_TRY(pcp_force_eval(c, st));
        PCP_TRY(pcp_synchronize(c));
        bool exact = (st->model & PCP_AREA_MODES) != 0;
        bool sq = (st->model & (PCP_AREA_SQUARED | PCP_PAIR_SQUARED)) != 0;
        int mode = (st->model & PCP_PAIR_SQUARED) ? 3 : (sq ? 2 : (exact ? 1 : 0));
        const double2* x = (const double2*)st->pos;
        int keepPair = c->segPair, keepLens = c->segLens, keepExt = c->segExt;
        size_t lst = ext_lds_stride(c->Kp);
        for (int k = 0; k < 3; ++k) {
            c->segPair = W[k];
            PCP_TRY(time_launches(c, reps, &t[k], [&]() { return exact ? launch_pair<1>(c, x) : launch_pair<0>(c, nullptr); }));
            if (sq) {
                c->segLens = W[k];
                PCP_TRY(time_launches(c, reps, &t[3 + k], [&]() { return launch_lens(c, x); }));
            }
            if ((size_t)(64 / W[k]) * lst <= 65536) {
                c->segExt = W[k];
                PCP_TRY(time_launches(c, reps, &t[6 + k], [&]() { return launch_exterior(c, mode, x); }));
            }
        }
        auto best = [&](int o, int keep) {
            int b = -1;
            for (int k = 0; k < 3; ++k)
                if (t[o + k] > 0.0 && (b < 0 || t[o + k] < t[o + b])) b = k;
            return b < 0 ? keep : W[b];
        };
        c->segPair = best(0, keepPair);
        c->segLens = sq ? best(3, keepLens) : keepLens;
        c->segExt = best(6, keepExt);
        // detection: the culled enumeration against whole edge tiles (bounds, cells and
        // detection timed together; each run starts from cleared buckets)
        int keepDet = c->detSched;
        const double2* xd = (const double2*)st->pos;
        for (int m = 0; m < (c->maxN <= 64 ? 2 : 1); ++m) {
            c->detSched = m;
            PCP_TRY(time_launches(c, reps, &t[9 + m], [&]() {
                return run_detect(c, xd, 0, nullptr, (st->model & PCP_CONSISTENT) != 0);
            }));
        }
        c->detSched = (t[10] > 0.0 && t[10] < t[9]) ? 1 : (t[9] > 0.0 ? 0 : keepDet);
        c->graphK = 0;   // a captured FIRE block holds the old schedule
        // leave the buffers as a plain evaluation with the chosen schedule leaves them
        PCP_TRY(pcp_force_eval(c, st));
        PCP_TRY(pcp_synchronize(c));
    }
    if (timesUs)
        for (int k = 0; k < 11; ++k) timesUs[k] = t[k];
    return PCP_OK;
}

int run_forces(pcp_ctx* c, int model, const double2* pos, double* edgeLengths, double* areas,
               const double* targetEdgeV, const double* targetAreas, double stiffness,
               double compressibility, double2* force, double* energy, int geom, int project, bool sticky,
               double perimK) {
    if (c->V == 0) return PCP_OK;
    int pair = (model & PCP_PAIR_SQUARED) ? 1 : 0;
    int sq = (model & (PCP_AREA_SQUARED | PCP_PAIR_SQUARED)) ? 1 : 0;
    int area = (model & PCP_AREA_MODES) ? 1 : 0;
    model &= 0xFF;
    int overlap = (model == PCP_NORMAL || model == PCP_HYBRID);
    int edgeOn = (model == PCP_SOFT_BODY || model == PCP_EDGE_ONLY || model == PCP_HYBRID);
    int areaOn = (model == PCP_SOFT_BODY || model == PCP_AREA_ONLY || model == PCP_HYBRID);
    // perimeter spring of the softBody and hybrid models; the force before the springs is
    // kept for FIRE with XPBD (every model with springs)
    if (!(model == PCP_SOFT_BODY || model == PCP_HYBRID)) perimK = 0.0;
    double2* fext = (model == PCP_NORMAL) ? nullptr : (double2*)c->dFext;
    if (model == PCP_ABNORMAL) {   // model.cpp:390-391: forces and energy zeroed only
        hipLaunchKernelGGL(k_fill_double, dim3(nblk(2LL * c->V)), dim3(BLK), 0, c->stream, (double*)force,
                           2LL * c->V, 0.0);
        PCP_LAUNCHED();
        if (energy) {
            hipLaunchKernelGGL(k_fill_double, dim3(nblk(c->P)), dim3(BLK), 0, c->stream, energy, (long long)c->P,
                               0.0);
            PCP_LAUNCHED();
        } else {
            hipLaunchKernelGGL(k_fill_double, dim3(nblk(c->N)), dim3(BLK), 0, c->stream, c->dEpoly,
                               (long long)c->N, 0.0);
            PCP_LAUNCHED();
        }
        if (geom) PCP_TRY(run_geometry(c, pos, edgeLengths, areas, c->dCom, c->dCom + c->N, nullptr, nullptr));
        return PCP_OK;
    }
    if (overlap) {
        int mode = pair ? 3 : (sq ? 2 : (area ? 1 : 0));
        if (c->pairDeferred) {   // the reference pairing left to this launch (run_records)
            c->pairDeferred = false;
            if (mode == 0) {
                PCP_TRY(launch_pair_exterior(c, pos));
                mode = -1;
            } else {
                PCP_TRY(launch_pair<0>(c, nullptr));
            }
        }
        if (sq) PCP_TRY(launch_lens(c, pos));
        if (pair) PCP_TRY(launch_lens_rep(c));
        if (mode >= 0) PCP_TRY(launch_exterior(c, mode, pos));
    } else {
        c->pairDeferred = false;
    }
    bool proj = project && c->maxN <= 64;
    if (geom && proj)
        PCP_TRY((dispatch_gather<1, 1>(c, overlap, edgeOn, areaOn, pos, edgeLengths, areas, targetEdgeV,
                                       targetAreas, stiffness, compressibility, force, perimK, fext)));
    else if (geom)
        PCP_TRY((dispatch_gather<1, 0>(c, overlap, edgeOn, areaOn, pos, edgeLengths, areas, targetEdgeV,
                                       targetAreas, stiffness, compressibility, force, perimK, fext)));
    else
        PCP_TRY((dispatch_gather<0, 0>(c, overlap, edgeOn, areaOn, pos, edgeLengths, areas, targetEdgeV,
                                       targetAreas, stiffness, compressibility, force, perimK, fext)));
    if (project && !proj) PCP_TRY(run_project(c, pos, force));
    if (!energy) return PCP_OK;   // the caller reduces the polygon energies (FIRE)
    return launch_det_reduce(c, c->dEpoly, 1, energy, 0, sticky);
}

// counts of the si buckets of packing p (host)
int packing_counts(pcp_ctx* c, int p, std::vector<int>& cnt) {
    int n0 = c->hPackPoly[p], n1 = c->hPackPoly[p + 1];
    cnt.assign(n1 - n0, 0);
    if (n1 > n0) PCP_HIP(hipMemcpy(cnt.data(), c->dBcnt + n0, sizeof(int) * (n1 - n0), hipMemcpyDeviceToHost));
    for (int& x : cnt) x = std::min(x, c->Kp);
    return PCP_OK;
}

}  // namespace pcp
