// HIP kernels + C ABI for the overlap-force path (gfx950).  See include/pcp.h.
//
// Pipeline of one force evaluation (one HIP stream, no host round trips):
//   geometry      thread / polygon     edge lengths, shoelace areas, COM, constraints
//   detect        polygon bounding circles (lane segment / polygon), polygon cell sort
//                 + candidate polygon pairs (workgroup / packing, LDS), edge-pair tiles
//                 (wavefront / polygon pair, LDS-staged edges) with the reference's
//                 cell-stencil test and segment test; crossings go to per-edge slots
//   cells         (cell-list getters and updateNeighborCells only) hash, exclusive scan,
//                 scatter, in-cell rank (ascending vertex id) + SoA edge staging
//   records       scan of bucket counts, scatter into si- and sj-buckets, one pass that
//                 ranks each record in its si bucket (final order (si,i,j,sj)) and pairs
//                 it with its exit record from the sj bucket
//   forces        exterior contributions per record -> deterministic per-vertex gather
//                 (exterior + interior chain + penalty terms), per-polygon energies,
//                 fixed 256-wide tree per packing
//   projection    lane segment / polygon: bordered cyclic-tridiagonal solve by parallel
//                 cyclic reduction over shuffles (thread / polygon beyond 64 vertices)
// Nothing is accumulated with floating-point atomics, so every result is
// bit-reproducible run to run.
#include "pcp_internal.hpp"
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>

using namespace pcp;

namespace pcp {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail_hip(hipError_t e, const char* what, const char* file, int line) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what,
             file, line);
    g_err = buf;
    return PCP_ERR_HIP;
}
static int fail_arg(const std::string& m) {
    g_err = m;
    return PCP_ERR_ARG;
}

// ------------------------------------------------------------------ wave helpers
__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int block_sum_int(int v, int* lds) {
    // BLK threads, wave64
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) lds[w] = v;
    __syncthreads();
    int s = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += lds[q];
    return s;
}
__device__ __forceinline__ int block_excl_scan_int(int v, int* lds, int* total) {
    int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o);
        if (l >= o) x += y;
    }
    __syncthreads();
    if (l == 63) lds[w] = x;
    __syncthreads();
    int off = 0, tot = 0;
    int nw = blockDim.x >> 6;
    for (int q = 0; q < nw; ++q) {
        if (q < w) off += lds[q];
        tot += lds[q];
    }
    *total = tot;
    return off + x - v;
}

// ------------------------------------------------------------------ small kernels
__global__ void k_fill_int(int* p, long long n, int v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_double(double* p, long long n, double v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// topology: shapeId/next/prev per vertex (updateShapeIdKernel kernels.cuh:652-659,
// computeNextPrevKernel :127-135), packing of each polygon
__global__ void k_topology(int V, int N, int P, const int* __restrict__ polyStart,
                           const int* __restrict__ packPoly, int* shapeId, int* next, int* prev,
                           int* polyPack) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) {
        int lo = 0, hi = N - 1;
        while (lo < hi) {   // last polygon with polyStart <= v
            int mid = (lo + hi + 1) >> 1;
            if (polyStart[mid] <= v) lo = mid; else hi = mid - 1;
        }
        int s = lo, st = polyStart[s], en = polyStart[s + 1];
        shapeId[v] = s;
        next[v] = (v + 1 == en) ? st : v + 1;
        prev[v] = (v == st) ? en - 1 : v - 1;
    }
    if (v < N) {
        int lo = 0, hi = P - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (packPoly[mid] <= v) lo = mid; else hi = mid - 1;
        }
        polyPack[v] = lo;
    }
}

// ------------------------------------------------------------------ scan
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = BLK * SCAN_ITEMS;

__global__ void k_scan_partials(const int* __restrict__ in, int n, int* __restrict__ part) {
    __shared__ int lds[16];
    long long base = (long long)blockIdx.x * SCAN_TILE;
    int s = 0;
    for (int q = threadIdx.x; q < SCAN_TILE; q += BLK) {
        long long i = base + q;
        if (i < n) s += in[i];
    }
    s = block_sum_int(s, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void k_scan_final(const int* __restrict__ in, int n, const int* __restrict__ part,
                             int* __restrict__ out) {
    __shared__ int lds[16];
    int b = blockIdx.x;
    int off = 0;
    for (int q = threadIdx.x; q < b; q += BLK) off += part[q];
    off = block_sum_int(off, lds);
    long long base = (long long)b * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int loc = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        loc += v[k];
    }
    int tot;
    int run = off + block_excl_scan_int(loc, lds, &tot);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (b == (int)gridDim.x - 1 && threadIdx.x == BLK - 1) out[n] = run;
}

int scan_exclusive(pcp_ctx* c, const int* in, int n, int* out) {
    int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb < 1) nb = 1;
    if (nb > c->scanPartCap) {
        dfree(c->dScanPart);
        PCP_TRY(dalloc(&c->dScanPart, nb));
        c->scanPartCap = nb;
    }
    PCP_LAUNCH(c, k_scan_partials, dim3(nb), dim3(BLK), in, n, c->dScanPart);
    PCP_LAUNCH(c, k_scan_final, dim3(nb), dim3(BLK), in, n, c->dScanPart, out);
    return PCP_OK;
}

// ------------------------------------------------------------------ geometry
// updatePolygonGeometryKernel kernels.cuh:295-333 + segmented sums model.cu:123-140
// (sequential per polygon) + normalizeKernel :635-650 + normalizeConstraintsKernel
// :335-345; max edge length per packing (model.cu:144-150).
__global__ void k_geometry(int N, const int* __restrict__ polyStart, const int* __restrict__ polyPack,
                           const double2* __restrict__ pos, double* __restrict__ edgeLengths,
                           double* __restrict__ areas, double* __restrict__ comX,
                           double* __restrict__ comY, double* __restrict__ constraints,
                           unsigned long long* __restrict__ maxEdgeBits) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    int st = polyStart[s], en = polyStart[s + 1], n = en - st;
    d2 p0 = ldpos(pos, st);
    double area = 0.0, cx = 0.0, cy = 0.0, ns0 = 0.0, ns1 = 0.0, ns2 = 0.0, mx = 0.0;
    d2 pprev = ldpos(pos, en - 1), pk = p0;
    for (int k = st; k < en; ++k) {
        int nx = (k + 1 == en) ? st : k + 1;
        d2 pn = ldpos(pos, nx);
        double dxPrev = wrap(pk.x - pprev.x);
        double dyPrev = wrap(pk.y - pprev.y);
        double dxNext = wrap(pk.x - pn.x);
        double dyNext = wrap(pk.y - pn.y);
        double lenPrev = sqrt(dxPrev * dxPrev + dyPrev * dyPrev);
        double lenNext = sqrt(dxNext * dxNext + dyNext * dyNext);
        // alpha = 0
        double x1a = wrap(pk.x - p0.x);
        double x4a = wrap(pprev.y - pn.y);
        cx += x1a;
        double a0 = (double)(2 * 0 - 1) * x4a;
        double e20 = dxNext + dxPrev;
        double e0 = (dxNext / lenNext) + (dxPrev / lenPrev);
        ns0 += a0 * a0;
        ns1 += e20 * e20;
        ns2 += e0 * e0;
        double dx = wrap(pn.x - pk.x);
        double dy = wrap(pn.y - pk.y);
        double el = sqrt(dx * dx + dy * dy);
        edgeLengths[k] = el;
        if (el > mx) mx = el;
        double dy2 = wrap(pn.y - pk.y);
        double x2 = wrap(pn.x - p0.x);
        area += dy2 * (x2 + x1a) / 2.0;
        // alpha = 1
        double x1b = wrap(pk.y - p0.y);
        double x4b = wrap(pprev.x - pn.x);
        cy += x1b;
        double a1 = (double)(2 * 1 - 1) * x4b;
        double e21 = dyNext + dyPrev;
        double e1 = (dyNext / lenNext) + (dyPrev / lenPrev);
        ns0 += a1 * a1;
        ns1 += e21 * e21;
        ns2 += e1 * e1;
        if (constraints) {
            double* cc = constraints + 6 * (long long)k;
            cc[0] = a0; cc[1] = a1; cc[2] = e20; cc[3] = e21; cc[4] = e0; cc[5] = e1;
        }
        pprev = pk;
        pk = pn;
    }
    areas[s] = area;
    double c = cx;
    c /= ((double)n);
    c += p0.x;
    c -= floor(c);
    comX[s] = c;
    c = cy;
    c /= ((double)n);
    c += p0.y;
    c -= floor(c);
    comY[s] = c;
    if (constraints) {
        double r0 = sqrt(ns0), r1 = sqrt(ns1), r2 = sqrt(ns2);
        for (int k = st; k < en; ++k) {
            double* cc = constraints + 6 * (long long)k;
            if (ns0 > 0.0) { cc[0] /= r0; cc[1] /= r0; }
            if (ns1 > 0.0) { cc[2] /= r1; cc[3] /= r1; }
            if (ns2 > 0.0) { cc[4] /= r2; cc[5] /= r2; }
        }
    }
    if (maxEdgeBits) atomicMax(&maxEdgeBits[polyPack[s]], (unsigned long long)__double_as_longlong(mx));
}

// k_geometry with one segment of SEG lanes per polygon (n <= SEG): lane k computes vertex k's
// terms -- the same expressions -- and stages them in LDS; six lanes then form the six sums
// (area, the two centroid sums, the three constraint norms) in vertex order, as the loop above
// adds them, so the results are k_geometry's bit for bit.  Edge lengths and constraint rows are
// written by their own lanes (contiguous within a segment).  polyMax (optional): each polygon's
// longest edge, reduced per packing afterwards (one atomic per polygon on the packing's maximum
// serializes: 262,144 polygons in 16 packings took 2.2 ms).
constexpr int GEO_T = 9;   // staged terms per vertex
template <int SEG>
__global__ __launch_bounds__(BLK) void k_geometry_seg(int N, const int* __restrict__ polyStart,
                                                      const int* __restrict__ polyPack,
                                                      const double2* __restrict__ pos,
                                                      double* __restrict__ edgeLengths, double* __restrict__ areas,
                                                      double* __restrict__ comX, double* __restrict__ comY,
                                                      double* __restrict__ constraints,
                                                      double* __restrict__ polyMax) {
    __shared__ double term[GEO_T][BLK];
    __shared__ double sums[BLK / SEG][6];
    int k = threadIdx.x & (SEG - 1), g = threadIdx.x / SEG, base = threadIdx.x - k;
    int s = blockIdx.x * (BLK / SEG) + g;
    bool has = s < N;
    int st = has ? polyStart[s] : 0, n = has ? polyStart[s + 1] - st : 0;
    bool live = k < n;
    double a0 = 0.0, a1 = 0.0, e20 = 0.0, e21 = 0.0, e0 = 0.0, e1 = 0.0, el = 0.0;
    if (live) {
        d2 p0 = ldpos(pos, st);
        d2 pk = ldpos(pos, st + k);
        d2 pprev = ldpos(pos, st + ((k == 0) ? n - 1 : k - 1));
        d2 pn = ldpos(pos, st + ((k + 1 == n) ? 0 : k + 1));
        double dxPrev = wrap(pk.x - pprev.x);
        double dyPrev = wrap(pk.y - pprev.y);
        double dxNext = wrap(pk.x - pn.x);
        double dyNext = wrap(pk.y - pn.y);
        double lenPrev = sqrt(dxPrev * dxPrev + dyPrev * dyPrev);
        double lenNext = sqrt(dxNext * dxNext + dyNext * dyNext);
        double x1a = wrap(pk.x - p0.x);
        double x4a = wrap(pprev.y - pn.y);
        a0 = (double)(2 * 0 - 1) * x4a;
        e20 = dxNext + dxPrev;
        e0 = (dxNext / lenNext) + (dxPrev / lenPrev);
        double dx = wrap(pn.x - pk.x);
        double dy = wrap(pn.y - pk.y);
        el = sqrt(dx * dx + dy * dy);
        edgeLengths[st + k] = el;
        double dy2 = wrap(pn.y - pk.y);
        double x2 = wrap(pn.x - p0.x);
        double x1b = wrap(pk.y - p0.y);
        double x4b = wrap(pprev.x - pn.x);
        a1 = (double)(2 * 1 - 1) * x4b;
        e21 = dyNext + dyPrev;
        e1 = (dyNext / lenNext) + (dyPrev / lenPrev);
        term[0][threadIdx.x] = dy2 * (x2 + x1a) / 2.0;   // area
        term[1][threadIdx.x] = x1a;                      // centroid x
        term[2][threadIdx.x] = x1b;                      // centroid y
        term[3][threadIdx.x] = a0 * a0;                  // norms: alpha = 0 then 1 of each vertex
        term[4][threadIdx.x] = a1 * a1;
        term[5][threadIdx.x] = e20 * e20;
        term[6][threadIdx.x] = e21 * e21;
        term[7][threadIdx.x] = e0 * e0;
        term[8][threadIdx.x] = e1 * e1;
    }
    __syncthreads();
    for (int j = k; j < 6; j += SEG) {   // sum j of the polygon, vertex after vertex
        double acc = 0.0;
        if (j < 3) {
            for (int m = 0; m < n; ++m) acc += term[j][base + m];
        } else {
            int t0 = 3 + 2 * (j - 3);
            for (int m = 0; m < n; ++m) {
                acc += term[t0][base + m];
                acc += term[t0 + 1][base + m];
            }
        }
        sums[g][j] = acc;
    }
    double mx = fmax(0.0, el);   // as "if (el > mx) mx = el" from 0 (a NaN length is passed over)
#pragma unroll
    for (int o = SEG / 2; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, SEG));
    __syncthreads();
    if (!has) return;
    if (k == 0) {
        areas[s] = sums[g][0];
        d2 p0 = ldpos(pos, st);
        double c = sums[g][1];
        c /= ((double)n);
        c += p0.x;
        c -= floor(c);
        comX[s] = c;
        c = sums[g][2];
        c /= ((double)n);
        c += p0.y;
        c -= floor(c);
        comY[s] = c;
        if (polyMax) polyMax[s] = mx;
    }
    if (constraints && live) {
        double ns0 = sums[g][3], ns1 = sums[g][4], ns2 = sums[g][5];
        double r0 = sqrt(ns0), r1 = sqrt(ns1), r2 = sqrt(ns2);
        double* cc = constraints + 6 * (long long)(st + k);
        cc[0] = (ns0 > 0.0) ? a0 / r0 : a0;
        cc[1] = (ns0 > 0.0) ? a1 / r0 : a1;
        cc[2] = (ns1 > 0.0) ? e20 / r1 : e20;
        cc[3] = (ns1 > 0.0) ? e21 / r1 : e21;
        cc[4] = (ns2 > 0.0) ? e0 / r2 : e0;
        cc[5] = (ns2 > 0.0) ? e1 / r2 : e1;
    }
}

int run_geometry(pcp_ctx* c, const double2* pos, double* edgeLengths, double* areas, double* comX,
                 double* comY, double* constraints, double* maxEdge) {
    if (c->N == 0) return PCP_OK;
    if (c->maxN > 64) {   // thread per polygon
        if (maxEdge) {
            hipLaunchKernelGGL(k_fill_double, dim3(nblk(c->P)), dim3(BLK), 0, c->stream, maxEdge,
                               (long long)c->P, 0.0);
            PCP_LAUNCHED();
        }
        PCP_LAUNCH(c, k_geometry, dim3(nblk(c->N)), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack, pos,
                   edgeLengths, areas, comX, comY, constraints, (unsigned long long*)maxEdge);
        return PCP_OK;
    }
    // the polygons' longest edges go to the reduction buffer (N of its 4N), then per packing
    double* polyMax = maxEdge ? c->dRed : nullptr;
#define PCP_GEOM(W)                                                                                           \
    PCP_LAUNCH(c, k_geometry_seg<W>, dim3((c->N + BLK / W - 1) / (BLK / W)), dim3(BLK), c->N, c->dPolyStart, \
               c->dPolyPack, pos, edgeLengths, areas, comX, comY, constraints, polyMax)
    switch (c->seg) {
    case 4: PCP_GEOM(4); break;
    case 8: PCP_GEOM(8); break;
    case 16: PCP_GEOM(16); break;
    case 32: PCP_GEOM(32); break;
    default: PCP_GEOM(64); break;
    }
#undef PCP_GEOM
    if (maxEdge) return launch_det_reduce(c, c->dRed, 1, maxEdge, 1, false);   // max from 0, as the fill
    return PCP_OK;
}

// ------------------------------------------------------------------ cell list
// updateNeighborCellsKernel kernels.cuh:661-707 (cell of the edge midpoint), as the
// coordinates ix | iy << 16
__device__ __forceinline__ int edge_cell_xy(d2 p, d2 q, int B) {
    double x = q.x - p.x + 0.5;
    double y = q.y - p.y + 0.5;
    x = shift_into_unit(x);
    y = shift_into_unit(y);
    x -= 0.5;
    y -= 0.5;
    x /= 2.0;
    y /= 2.0;
    x += p.x;
    y += p.y;
    x = shift_into_unit(x);
    y = shift_into_unit(y);
    double fx = floor(x * B), fy = floor(y * B);
    int ix = (fx == fx && fx >= 0.0) ? (fx < 2147483647.0 ? (int)fx : B - 1) : 0;
    int iy = (fy == fy && fy >= 0.0) ? (fy < 2147483647.0 ? (int)fy : B - 1) : 0;
    if (ix < 0) ix = 0; else if (ix >= B) ix = B - 1;
    if (iy < 0) iy = 0; else if (iy >= B) iy = B - 1;
    return ix | (iy << 16);
}
// the same cell as the flat index iy * B + ix
__device__ __forceinline__ int edge_cell(d2 p, d2 q, int B) {
    int xy = edge_cell_xy(p, q, B);
    return (xy >> 16) * B + (xy & 0xFFFF);
}

__global__ void k_cell_hash(int V, const double2* __restrict__ pos, const int* __restrict__ shapeId,
                            const int* __restrict__ next, const int* __restrict__ polyPack,
                            const int* __restrict__ box, const long long* __restrict__ cellOff,
                            int* __restrict__ cellOf, int* __restrict__ cellCount) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    int p = polyPack[shapeId[v]];
    int B = box[p];
    int cl = edge_cell(ldpos(pos, v), ldpos(pos, next[v]), B);
    cellOf[v] = cl;
    atomicAdd(&cellCount[cellOff[p] + cl], 1);
}

__global__ void k_cell_scatter(int V, const int* __restrict__ shapeId, const int* __restrict__ polyPack,
                               const long long* __restrict__ cellOff, const int* __restrict__ cellOf,
                               int* __restrict__ cellCount, const int* __restrict__ cellStart,
                               int* __restrict__ cellTmp) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    long long gc = cellOff[polyPack[shapeId[v]]] + cellOf[v];
    int slot = atomicSub(&cellCount[gc], 1) - 1;   // leaves the counts at zero
    cellTmp[cellStart[gc] + slot] = v;
}

// stable order inside each cell = ascending vertex id (CUB radix sort of iota values,
// model.cu:322-349); stage the edge data of every sorted position for the stencil scan
__global__ void k_cell_sort(int V, const double2* __restrict__ pos, const int* __restrict__ shapeId,
                            const int* __restrict__ next, const int* __restrict__ polyStart,
                            const int* __restrict__ polyPack, const long long* __restrict__ cellOff,
                            const int* __restrict__ cellOf, const int* __restrict__ cellStart,
                            const int* __restrict__ cellTmp, int* __restrict__ cellSorted,
                            int* __restrict__ cellInv, double2* __restrict__ SP, double2* __restrict__ SR,
                            int4* __restrict__ SI) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= V) return;
    int v = cellTmp[t];
    int s = shapeId[v];
    long long gc = cellOff[polyPack[s]] + cellOf[v];
    int lo = cellStart[gc], hi = cellStart[gc + 1];
    int rank = 0;
    for (int q = lo; q < hi; ++q) rank += (cellTmp[q] < v);
    int dst = lo + rank;
    cellSorted[dst] = v;
    cellInv[v] = dst;
    int nv = next[v];
    d2 p = ldpos(pos, v), pn = ldpos(pos, nv);
    SP[dst] = make_double2(p.x, p.y);
    SR[dst] = make_double2(wrap(pn.x - p.x), wrap(pn.y - p.y));
    SI[dst] = make_int4(v, nv, s, v - polyStart[s]);
}

int run_cells(pcp_ctx* c, const double2* pos) {
    if (c->V == 0) return PCP_OK;
    PCP_LAUNCH(c, k_cell_hash, dim3(nblk(c->V)), dim3(BLK), c->V, pos, c->dShapeId,
                       c->dNext, c->dPolyPack, c->dBox, c->dCellOff, c->dCellOf, c->dCellCount);
    PCP_TRY(scan_exclusive(c, c->dCellCount, (int)c->Ctot, c->dCellStart));
    PCP_LAUNCH(c, k_cell_scatter, dim3(nblk(c->V)), dim3(BLK), c->V, c->dShapeId,
                       c->dPolyPack, c->dCellOff, c->dCellOf, c->dCellCount, c->dCellStart, c->dCellTmp);
    PCP_LAUNCH(c, k_cell_sort, dim3(nblk(c->V)), dim3(BLK), c->V, pos, c->dShapeId,
                       c->dNext, c->dPolyStart, c->dPolyPack, c->dCellOff, c->dCellOf, c->dCellStart,
                       c->dCellTmp, c->dCellSorted, c->dCellInv, c->dSP, c->dSR, c->dSI);
    return PCP_OK;
}

// ------------------------------------------------------------------ edge intersections
}  // namespace pcp
#include "pcp_detect.hip"
namespace pcp {

// restage the sorted edge data from the live positions (updateNeighbors reads the
// positions at call time, kernels.cuh:721-725; the cells stay those of the last
// updateNeighborCells)
__global__ void k_restage(int V, const double2* __restrict__ pos, const int* __restrict__ next,
                          const int* __restrict__ cellSorted, double2* __restrict__ SP,
                          double2* __restrict__ SR) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= V) return;
    int v = cellSorted[t];
    d2 p = ldpos(pos, v), pn = ldpos(pos, next[v]);
    SP[t] = make_double2(p.x, p.y);
    SR[t] = make_double2(wrap(pn.x - p.x), wrap(pn.y - p.y));
}

// stable rank sort of keys by their low endBit bits (sortKeysCUDA model.cu:50-75)
__global__ void k_sort_keys(int n, uint64_t mask, const uint64_t* __restrict__ in,
                            uint64_t* __restrict__ out, uint32_t* __restrict__ perm) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint64_t k = in[r] & mask;
    int rank = 0;
    for (int q = 0; q < n; ++q) {
        uint64_t o = in[q] & mask;
        rank += (o < k) || (o == k && q < r);
    }
    out[rank] = in[r];
    perm[rank] = (uint32_t)r;
}

// Per-packing deterministic sum of n values (fixed order, the same as orc_det_sum):
// pairwise 256-wide trees over consecutive 256-value chunks (v[t] += v[t+s], s = 128..1),
// then one such tree over the chunk sums (n <= 65536).  One block of BLK threads per
// packing.  A chunk's tree is evaluated by one wavefront in registers (chunk_tree); the
// wavefronts take DR_BATCH chunks at a time, all their loads issued first.  Result valid
// in every thread.
constexpr int DR_BATCH = 4;
// block of the per-packing reduction kernels (k_det_reduce, k_fire_energy, k_fire_stats)
constexpr int REDUCE_BLK = 1024;
// the tree of one chunk: lane l holds v[l], v[l+64], v[l+128], v[l+192]; level 128
// (v[l] + v[l+128], v[l+64] + v[l+192]) and level 64 inside the lane, levels 32..1 by
// shuffles (lane t adds lane t + s); the same additions in the same order.  Sum in lane 0.
__device__ __forceinline__ double chunk_tree(double v0, double v1, double v2, double v3) {
    double a = v0 + v2;
    double b = v1 + v3;
    double x = a + b;
#pragma unroll
    for (int st = 32; st >= 1; st >>= 1) x += __shfl_down(x, st);
    return x;
}
// NB threads per block; with NB >= 1024 a wavefront takes one chunk per pass (a 4,096-polygon
// packing's 16 chunks are then evaluated by 16 wavefronts side by side)
template <int NB = BLK>
__device__ double det_sum_block(const double* __restrict__ xa, int n, double* part) {
    constexpr int DRB = (NB >= 1024) ? 1 : DR_BATCH;
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int chunks = (n + 255) >> 8;
    for (int c0 = w * DRB; c0 < chunks; c0 += (NB / 64) * DRB) {
        double v[DRB][4];
#pragma unroll
        for (int u = 0; u < DRB; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int q = (c0 + u) * 256 + lane + 64 * k;
                v[u][k] = (q < n) ? xa[q] : 0.0;
            }
#pragma unroll
        for (int u = 0; u < DRB; ++u)
            if (c0 + u < chunks) {   // wavefront-uniform
                double x = chunk_tree(v[u][0], v[u][1], v[u][2], v[u][3]);
                if (lane == 0) part[c0 + u] = x;
            }
    }
    __syncthreads();
    double res;
    if (chunks <= 1) {
        res = (chunks == 1) ? part[0] : 0.0;
    } else {
        if (w == 0) {
            double x = chunk_tree(lane < chunks ? part[lane] : 0.0, lane + 64 < chunks ? part[lane + 64] : 0.0,
                                  lane + 128 < chunks ? part[lane + 128] : 0.0,
                                  lane + 192 < chunks ? part[lane + 192] : 0.0);
            if (lane == 0) part[NB] = x;
        }
        __syncthreads();
        res = part[NB];
    }
    __syncthreads();
    return res;
}
// maximum of n values (exact in any order); valid in every thread
template <int NB = BLK>
__device__ double max_block(const double* __restrict__ xa, int n, double* L) {
    int t = threadIdx.x;
    double m = 0.0;
    for (int q = t; q < n; q += NB) m = fmax(m, xa[q]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((t & 63) == 0) L[t >> 6] = m;
    __syncthreads();
    double r = 0.0;
    for (int q = 0; q < NB / 64; ++q) r = fmax(r, L[q]);
    __syncthreads();
    return r;
}

// NS sums (det_sum_block's trees, bit-identical) of the arrays xa + a * stride, a < NS,
// and with MX the maximum of array NS, all from one pass whose loads are issued together
// (one memory latency instead of one per array).  part: NS * (BLK + 1) doubles, L: BLK / 64.
// res[a] = sum of array a; res[NS] = the maximum.  Valid in every thread.
template <int NS, bool MX, int NB = BLK>
__device__ void det_reduce_block(const double* __restrict__ xa, long long stride, int n, double* part, double* L,
                                 double* res) {
    constexpr int NA = NS + (MX ? 1 : 0);
    constexpr int DRB = (NB >= 1024) ? 1 : DR_BATCH;
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int chunks = (n + 255) >> 8;
    double mx = 0.0;
    for (int c0 = w * DRB; c0 < chunks; c0 += (NB / 64) * DRB) {
        double v[NA][DRB][4];
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int u = 0; u < DRB; ++u)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    int q = (c0 + u) * 256 + lane + 64 * k;
                    v[a][u][k] = (q < n) ? xa[a * stride + q] : 0.0;
                }
#pragma unroll
        for (int a = 0; a < NS; ++a)
#pragma unroll
            for (int u = 0; u < DRB; ++u)
                if (c0 + u < chunks) {   // wavefront-uniform
                    double x = chunk_tree(v[a][u][0], v[a][u][1], v[a][u][2], v[a][u][3]);
                    if (lane == 0) part[a * (NB + 1) + c0 + u] = x;
                }
        if constexpr (MX) {
#pragma unroll
            for (int u = 0; u < DRB; ++u)
#pragma unroll
                for (int k = 0; k < 4; ++k) mx = fmax(mx, v[NS][u][k]);
        }
    }
    if constexpr (MX) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
        if (lane == 0) L[w] = mx;
    }
    __syncthreads();
    if (chunks > 1 && w < NS) {   // wavefront a: the top tree of array a
        double* pa = part + w * (NB + 1);
        double x = chunk_tree(lane < chunks ? pa[lane] : 0.0, lane + 64 < chunks ? pa[lane + 64] : 0.0,
                              lane + 128 < chunks ? pa[lane + 128] : 0.0, lane + 192 < chunks ? pa[lane + 192] : 0.0);
        if (lane == 0) pa[NB] = x;
    }
    static_assert(NS <= NB / 64, "one wavefront per top tree");
    __syncthreads();
#pragma unroll
    for (int a = 0; a < NS; ++a)
        res[a] = (chunks > 1) ? part[a * (NB + 1) + NB] : (chunks == 1 ? part[a * (NB + 1)] : 0.0);
    if constexpr (MX) {
        double r = 0.0;
        for (int q = 0; q < NB / 64; ++q) r = fmax(r, L[q]);
        res[NS] = r;
    }
    __syncthreads();
}

// per-packing reductions over polygons: nArrays arrays of N values (stride N) are summed
// with det_sum_block; the last nMax of them are max-reduced instead.  Block per packing.
// With stickyScal, block 0 also records whether the evaluation overflowed a record bucket.
__global__ __launch_bounds__(REDUCE_BLK) void k_det_reduce(const int* __restrict__ packPoly,
                                                   const double* __restrict__ x, int N, int nArrays,
                                                   int nMax, double* __restrict__ out, int P,
                                                   const int* __restrict__ stickyScal, int Kp,
                                                   int* __restrict__ sticky) {
    __shared__ double L[REDUCE_BLK / 64];
    __shared__ double part[REDUCE_BLK + 1];
    int p = blockIdx.x;
    int a0 = packPoly[p], n = packPoly[p + 1] - a0;
    if (stickyScal && p == 0 && threadIdx.x == 0 && stickyScal[SC_BUCKET] > Kp) sticky[0] = 1;
    for (int a = 0; a < nArrays; ++a) {
        const double* xa = x + (long long)a * N + a0;
        double res = (a >= nArrays - nMax) ? max_block<REDUCE_BLK>(xa, n, L) : det_sum_block<REDUCE_BLK>(xa, n, part);
        if (threadIdx.x == 0) out[(long long)a * P + p] = res;
    }
}

int launch_det_reduce(pcp_ctx* c, const double* partials, int nArrays, double* outSums, int nMax, bool sticky) {
    if (c->P == 0) return PCP_OK;
    PCP_LAUNCH(c, k_det_reduce, dim3(c->P), dim3(REDUCE_BLK), c->dPackPoly, partials, c->N,
                       nArrays, nMax, outSums, c->P, sticky ? c->dScal : nullptr, c->Kp, c->dSticky);
    return PCP_OK;
}

// ------------------------------------------------------------------ projection
// Per polygon, remove from f its component in span{grad l_e (n edges), grad A}
// (projectForceCUDA model.cu:165-206, kernels.cuh:389-523; the reference builds an
// SVD basis of the edge gradients and Gram-Schmidts the normalized area gradient; the
// projector is basis independent).  Normal equations of [A a]: G = A^T A is cyclic
// tridiagonal (G_ee = 2|t_e|^2, G_{e,e+1} = -t_e.t_{e+1}), bordered by b = A^T a and
// c = a.a.  Solved by Thomas elimination on the leading (n-1)x(n-1) block with three
// right-hand sides, then the 1x1 Schur complements of the cyclic corner and of the
// area row.  The area direction is dropped when its residual after removing the edge
// directions is below 1e-14 (normalized), as kernels.cuh:464.  Degenerate edges
// (|e| < 1e-14, kernels.cuh:399) contribute no direction.
struct PolySolver {
    double* tx; double* ty; double* ax; double* ay; double* cp; double* x1; double* x2; double* x3;
};

// Solve [[G, b],[b^T, c]] [y; z] = [r; s]; y stored in S.x1[0..n), returns z.
// rhs r in S.x1 on entry.  S.tx/ty/ax/ay hold tangents and area gradient.
__device__ double bordered_solve(const PolySolver& S, int n, double c, double s, double cthresh) {
    // off_e = G_{e,e+1 mod n}
    auto off = [&](int e) -> double {
        int f = (e + 1 == n) ? 0 : e + 1;
        return -(S.tx[e] * S.tx[f] + S.ty[e] * S.ty[f]);
    };
    auto diag = [&](int e) -> double {
        double d = S.tx[e] * S.tx[e] + S.ty[e] * S.ty[e];
        return (d > 0.0) ? 2.0 * d : 1.0;
    };
    auto bvec = [&](int e) -> double {
        int f = (e + 1 == n) ? 0 : e + 1;
        return S.tx[e] * (S.ax[f] - S.ax[e]) + S.ty[e] * (S.ay[f] - S.ay[e]);
    };
    int m = n - 1;
    // right-hand sides: x1 = r (given), x2 = b, x3 = w (border column of index n-1)
    double r_last = S.x1[m];
    for (int j = 0; j < m; ++j) {
        S.x2[j] = bvec(j);
        S.x3[j] = 0.0;
    }
    double b_last = bvec(m);
    S.x3[0] = off(m);                       // G_{0,n-1} (cyclic corner)
    S.x3[m - 1] += off(m - 1);              // G_{n-2,n-1}
    double w0 = S.x3[0], wl = (m - 1 == 0) ? 0.0 : S.x3[m - 1];
    // forward sweep
    double cprev = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
    for (int j = 0; j < m; ++j) {
        double sub = (j > 0) ? off(j - 1) : 0.0;
        double den = diag(j) - sub * cprev;
        double sup = (j < m - 1) ? off(j) : 0.0;
        double cj = sup / den;
        d1 = (S.x1[j] - sub * d1) / den;
        d2 = (S.x2[j] - sub * d2) / den;
        d3 = (S.x3[j] - sub * d3) / den;
        S.cp[j] = cj;
        S.x1[j] = d1; S.x2[j] = d2; S.x3[j] = d3;
        cprev = cj;
    }
    for (int j = m - 2; j >= 0; --j) {
        S.x1[j] = S.x1[j] - S.cp[j] * S.x1[j + 1];
        S.x2[j] = S.x2[j] - S.cp[j] * S.x2[j + 1];
        S.x3[j] = S.x3[j] - S.cp[j] * S.x3[j + 1];
    }
    // w^T X for the three solutions (w nonzero at 0 and m-1 only)
    double wX1, wX2, wX3;
    if (m - 1 == 0) {
        wX1 = w0 * S.x1[0]; wX2 = w0 * S.x2[0]; wX3 = w0 * S.x3[0];
    } else {
        wX1 = w0 * S.x1[0] + wl * S.x1[m - 1];
        wX2 = w0 * S.x2[0] + wl * S.x2[m - 1];
        wX3 = w0 * S.x3[0] + wl * S.x3[m - 1];
    }
    double schur = diag(m) - wX3;
    double y1l = (r_last - wX1) / schur;
    double ul = (b_last - wX2) / schur;
    // y1 = G^{-1} r (into x1), u = G^{-1} b (into x2)
    for (int j = 0; j < m; ++j) {
        S.x1[j] = S.x1[j] - S.x3[j] * y1l;
        S.x2[j] = S.x2[j] - S.x3[j] * ul;
    }
    S.x1[m] = y1l;
    S.x2[m] = ul;
    double bu = 0.0, by = 0.0;
    for (int e = 0; e < n; ++e) {
        double be = bvec(e);
        bu += be * S.x2[e];
        by += be * S.x1[e];
    }
    double sigma = c - bu;
    double z = (sigma > cthresh * c) ? (s - by) / sigma : 0.0;
    for (int e = 0; e < n; ++e) S.x1[e] = S.x1[e] - z * S.x2[e];
    return z;
}

__global__ __launch_bounds__(BLK) void k_project(int N, int V, const int* __restrict__ polyStart,
                                                const double2* __restrict__ pos, double2* __restrict__ force,
                                                double* __restrict__ work) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    int st = polyStart[s], n = polyStart[s + 1] - st;
    if (n < 3) return;
    PolySolver S;
    S.tx = work + st; S.ty = work + V + st; S.ax = work + 2LL * V + st; S.ay = work + 3LL * V + st;
    S.cp = work + 4LL * V + st; S.x1 = work + 5LL * V + st; S.x2 = work + 6LL * V + st;
    S.x3 = work + 7LL * V + st;
    double c = 0.0, sdot = 0.0;
    for (int e = 0; e < n; ++e) {
        int f = (e + 1 == n) ? 0 : e + 1, g = (e == 0) ? n - 1 : e - 1;
        d2 pe = ldpos(pos, st + e), pf = ldpos(pos, st + f), pg = ldpos(pos, st + g);
        double dx = wrap(pf.x - pe.x), dy = wrap(pf.y - pe.y);
        double len = sqrt(dx * dx + dy * dy);
        if (len < 1e-14) { S.tx[e] = 0.0; S.ty[e] = 0.0; }
        else { S.tx[e] = dx / len; S.ty[e] = dy / len; }
        double ax = wrap(pf.y - pg.y), ay = wrap(pg.x - pf.x);
        S.ax[e] = ax; S.ay[e] = ay;
        double2 fe = force[st + e];
        c += ax * ax + ay * ay;
        sdot += ax * fe.x + ay * fe.y;
    }
    for (int e = 0; e < n; ++e) {
        int f = (e + 1 == n) ? 0 : e + 1;
        double2 fe = force[st + e], ff = force[st + f];
        S.x1[e] = S.tx[e] * (ff.x - fe.x) + S.ty[e] * (ff.y - fe.y);
    }
    double z = bordered_solve(S, n, c, sdot, 1e-28);
    for (int k = 0; k < n; ++k) {
        int g = (k == 0) ? n - 1 : k - 1;
        double px = -S.x1[k] * S.tx[k] + S.x1[g] * S.tx[g] + z * S.ax[k];
        double py = -S.x1[k] * S.ty[k] + S.x1[g] * S.ty[g] + z * S.ay[k];
        double2 fk = force[st + k];
        force[st + k] = make_double2(fk.x - px, fk.y - py);
    }
}

int run_project(pcp_ctx* c, const double2* pos, double2* force) {
    if (c->N == 0) return PCP_OK;
    if (c->maxN <= 64) return run_project_seg(c, pos, force);
    PCP_LAUNCH(c, k_project, dim3(nblk(c->N)), dim3(BLK), c->N, c->V, c->dPolyStart, pos,
                       force, c->dWork);
    return PCP_OK;
}

// ------------------------------------------------------------------ SHAKE
// shake.h is absent from the reference (model.cu:16); contract from shakeProjectCUDA
// model.cu:619-636 and Model::minimizeFIREStep model.cpp:418-438.  Per polygon,
// Gauss-Newton projection onto {|edge e| = L_e, area = A0}: J J^T lambda = C,
// x -= J^T lambda (same bordered cyclic-tridiagonal structure as the projection, with
// the exact area gradient), until max|C| < tol.  retryIter > 0 reproduces the second
// call with 10x iterations that minimizeFIREStep makes when the first call hit its
// limit: polygons that converged in the first call take 0 further iterations, so the
// per-polygon continuation equals the reference's packing-wide retry.
// itOut[2P]: [p] = max iterations of call 1, [P+p] = of call 2 (atomicMax).
__device__ double poly_area_dev(const double2* pos, int st, int n) {
    d2 p0 = ldpos(pos, st);
    double area = 0.0;
    for (int k = 0; k < n; ++k) {
        int kn = (k + 1 == n) ? 0 : k + 1;
        d2 pa = ldpos(pos, st + k), pb = ldpos(pos, st + kn);
        double x1 = wrap(pa.x - p0.x);
        double dy2 = wrap(pb.y - pa.y);
        double x2 = wrap(pb.x - p0.x);
        area += dy2 * (x2 + x1) / 2.0;
    }
    return area;
}

__device__ int shake_poly(double2* pos, int st, int n, const double* targetEdgeV, double A0, int maxIter,
                          double tol, const PolySolver& S) {
    int it = 0;
    for (; it < maxIter; ++it) {
        double cmax = 0.0, c = 0.0;
        for (int e = 0; e < n; ++e) {
            int f = (e + 1 == n) ? 0 : e + 1, g = (e == 0) ? n - 1 : e - 1;
            d2 pe = ldpos(pos, st + e), pf = ldpos(pos, st + f), pg = ldpos(pos, st + g);
            double dx = wrap(pf.x - pe.x), dy = wrap(pf.y - pe.y);
            double len = sqrt(dx * dx + dy * dy);
            double ce = len - targetEdgeV[st + e];
            if (fabs(ce) > cmax) cmax = fabs(ce);
            S.x1[e] = ce;
            if (len > 0.0) { S.tx[e] = dx / len; S.ty[e] = dy / len; }
            else { S.tx[e] = 0.0; S.ty[e] = 0.0; }
            double ax = wrap(pf.y - pg.y) / 2, ay = wrap(pg.x - pf.x) / 2;
            S.ax[e] = ax; S.ay[e] = ay;
            c += ax * ax + ay * ay;
        }
        double cA = poly_area_dev(pos, st, n) - 0.0;
        cA = cA - A0;
        if (fabs(cA) > cmax) cmax = fabs(cA);
        if (cmax < tol) break;
        double z = bordered_solve(S, n, c, cA, 1e-28);
        for (int k = 0; k < n; ++k) {
            int g = (k == 0) ? n - 1 : k - 1;
            double dx = -S.x1[k] * S.tx[k] + S.x1[g] * S.tx[g] + z * S.ax[k];
            double dy = -S.x1[k] * S.ty[k] + S.x1[g] * S.ty[g] + z * S.ay[k];
            double2 p = pos[st + k];
            double qx = p.x - dx, qy = p.y - dy;
            pos[st + k] = make_double2(qx - floor(qx), qy - floor(qy));
        }
    }
    return it;
}

__global__ __launch_bounds__(BLK) void k_shake(int N, int V, int P, const int* __restrict__ polyStart,
                                              const int* __restrict__ polyPack, double2* __restrict__ pos,
                                              const double* __restrict__ targetEdgeV,
                                              const double* __restrict__ targetAreas, int maxIter,
                                              double tol, int retryIter, const int* __restrict__ active,
                                              double* __restrict__ work, int* __restrict__ itOut) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    int p = polyPack[s];
    if (active && !active[p]) return;
    int st = polyStart[s], n = polyStart[s + 1] - st;
    if (n < 3) return;
    PolySolver S;
    S.tx = work + st; S.ty = work + V + st; S.ax = work + 2LL * V + st; S.ay = work + 3LL * V + st;
    S.cp = work + 4LL * V + st; S.x1 = work + 5LL * V + st; S.x2 = work + 6LL * V + st;
    S.x3 = work + 7LL * V + st;
    int it1 = shake_poly(pos, st, n, targetEdgeV, targetAreas[s], maxIter, tol, S);
    max_if_above(&itOut[p], it1);
    if (retryIter > 0 && it1 >= maxIter) {
        int it2 = shake_poly(pos, st, n, targetEdgeV, targetAreas[s], retryIter, tol, S);
        max_if_above(&itOut[P + p], it2);
    }
}

int run_shake(pcp_ctx* c, double2* pos, const double* targetEdgeV, const double* targetAreas, int maxIter,
              double tol, int retryIter, const int* active, int* itOut) {
    if (c->N == 0 || maxIter == 0) return PCP_OK;
    if (c->maxN <= 64)
        return run_shake_seg(c, pos, targetEdgeV, targetAreas, maxIter, tol, retryIter, active, itOut);
    PCP_LAUNCH(c, k_shake, dim3(nblk(c->N)), dim3(BLK), c->N, c->V, c->P, c->dPolyStart,
                       c->dPolyPack, (double2*)pos, targetEdgeV, targetAreas, maxIter, tol, retryIter,
                       active, c->dWork, itOut);
    return PCP_OK;
}

}  // namespace pcp
#include "pcp_solve.hip"
#include "pcp_records.hip"
namespace pcp {

// ------------------------------------------------------------------ misc kernels
__global__ void k_max_abs_poly(int N, const int* __restrict__ polyStart, const double2* __restrict__ x,
                               double* __restrict__ out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    double m = 0.0;
    for (int v = polyStart[s]; v < polyStart[s + 1]; ++v) {
        double2 a = x[v];
        m = fmax(m, fmax(fabs(a.x), fabs(a.y)));
    }
    out[s] = m;
}
// updatePositionsKernel kernels.cuh:1433-1438
__global__ void k_update_positions(long long n2, double* pos, const double* force, double dt) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2) return;
    double p = pos[i] + force[i] * dt;
    pos[i] = p - floor(p);
}
// resetAreasKernel kernels.cuh:1440-1455
__global__ void k_reset_areas(long long n2, const int* __restrict__ shapeId, double* pos,
                              const double* areas, const double* targetAreas, const double* comX,
                              const double* comY) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n2) return;
    double x = pos[idx];
    int s = shapeId[idx / 2];
    double currentArea = areas[s];
    double targetArea = targetAreas[s];
    if (!(currentArea > 0.0)) return;
    double rescale = sqrt(targetArea / currentArea);
    double u = (idx % 2) ? wrap(x - comY[s]) : wrap(x - comX[s]);
    u *= rescale;
    double p = (idx % 2) ? u + comY[s] : u + comX[s];
    pos[idx] = p - floor(p);
}
// the batched form of setPhi's target scaling (pyCudaPolygon.py:223-234: targetAreas *= r,
// targetEdgeLengths *= sqrt(r)): f[p] multiplies packing p's target areas, f[P + p] its edge
// targets.  The factors come from the host, so each product is that of the host's multiply.
__global__ void k_scale_targets(int V, int N, int P, const int* __restrict__ shapeId,
                                const int* __restrict__ polyPack, const double* __restrict__ f,
                                double* __restrict__ targetAreas, double* __restrict__ targetEdgeV) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (targetEdgeV && i < V) targetEdgeV[i] *= f[P + polyPack[shapeId[i]]];
    if (targetAreas && i < N) targetAreas[i] *= f[polyPack[i]];
}
// effectiveForceMagKernel kernels.cuh:625-633, max per polygon
__global__ void k_eff_force_poly(int N, const int* __restrict__ polyStart, const double2* __restrict__ pos,
                                 const double2* __restrict__ tent, const double2* __restrict__ force,
                                 double scale, double* __restrict__ out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    double m = 0.0;
    for (int k = polyStart[s]; k < polyStart[s + 1]; ++k) {
        double fx = force[k].x + scale * (pos[k].x - tent[k].x);
        double fy = force[k].y + scale * (pos[k].y - tent[k].y);
        m = fmax(m, sqrt(fx * fx + fy * fy));
    }
    out[s] = m;
}
// wrapPeriodic kernels.cuh:108-113
__device__ __forceinline__ double wrapPeriodic(double x) {
    x += 1.5;
    int guard = 0;
    while (x > 1.0 && guard < 64) { x -= 1.0; ++guard; }
    return x - 0.5;
}
// ordered neighbour lists for the getters (reference slot order: 3x3 stencil in
// (dx, dy) order with de-duplication, ascending vertex id inside a cell)
__global__ void k_neighbor_lists(int v0, int v1, int M, const int* __restrict__ shapeId,
                                 const int* __restrict__ polyPack, const int* __restrict__ box,
                                 const long long* __restrict__ cellOff, const int* __restrict__ cellOf,
                                 const int* __restrict__ cellStart, const int* __restrict__ next,
                                 const double2* __restrict__ SP, const double2* __restrict__ SR,
                                 const int4* __restrict__ SI, const int* __restrict__ cellInv,
                                 int* __restrict__ outN, unsigned char* __restrict__ outIn,
                                 int* __restrict__ outNum) {
    int id1 = v0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (id1 >= v1) return;
    int p = polyPack[shapeId[id1]];
    int B = box[p];
    long long cbase = cellOff[p];
    int id2 = next[id1];
    int me = cellInv[id1];
    double2 P1 = SP[me];
    double rx = SR[me].x, ry = SR[me].y;
    int cl = cellOf[id1];
    int bx = cl % B, by = cl / B;
    int count = 0;
    long long row = (long long)(id1 - v0) * M;
    for (int kc = 0; kc < 9; ++kc) {
        int nx = bx + kc / 3 - 1;
        if (nx < 0) nx += B; else if (nx >= B) nx -= B;
        int ny = by + kc % 3 - 1;
        if (ny < 0) ny += B; else if (ny >= B) ny -= B;
        int nb = ny * B + nx;
        bool dup = false;
        for (int kp = 0; kp < kc; ++kp) {
            int ex = bx + kp / 3 - 1;
            if (ex < 0) ex += B; else if (ex >= B) ex -= B;
            int ey = by + kp % 3 - 1;
            if (ey < 0) ey += B; else if (ey >= B) ey -= B;
            dup |= (ey * B + ex == nb);
        }
        if (dup) continue;
        for (int q = cellStart[cbase + nb]; q < cellStart[cbase + nb + 1]; ++q) {
            int4 ot = SI[q];
            int nid = ot.x, nid2 = ot.y;
            if (nid <= id1) continue;
            if (nid == id1 || nid == id2 || nid2 == id1 || nid2 == id2) continue;
            double2 Q = SP[q], S = SR[q];
            double gx = wrap(Q.x - P1.x), gy = wrap(Q.y - P1.y);
            double denom = rx * S.y - ry * S.x;
            if (fabs(denom) < 1e-12) continue;
            double tt = (gx * S.y - gy * S.x) / denom;
            double uu = (gx * ry - gy * rx) / denom;
            if (tt > 0.0 && tt < 1.0 && uu > 0.0 && uu < 1.0) {
                if (count < M) {
                    outN[row + count] = nid - (v0);
                    outIn[row + count] = (denom > 0) ? 1 : 0;
                }
                count++;
            }
        }
    }
    outNum[id1 - v0] = count;
}

int prof_begin(pcp_ctx* c, const char* name, int* slot) {
    *slot = -1;
    if (c->profName.empty() || c->profName != name) return PCP_OK;
    if (c->profUsed == c->profEvents.size()) {
        hipEvent_t a, b;
        PCP_HIP(hipEventCreate(&a));
        PCP_HIP(hipEventCreate(&b));
        c->profEvents.push_back({a, b});
    }
    *slot = (int)c->profUsed++;
    PCP_HIP(hipEventRecord(c->profEvents[*slot].first, c->stream));
    return PCP_OK;
}
int prof_end(pcp_ctx* c, int slot) {
    if (slot < 0) return PCP_OK;
    PCP_HIP(hipEventRecord(c->profEvents[slot].second, c->stream));
    return PCP_OK;
}
// an event pair for the next launch of the profiled kernel, recorded by the launch itself
int prof_slot(pcp_ctx* c, const char* name, int* slot) {
    *slot = -1;
    if (c->profName.empty() || c->profName != name) return PCP_OK;
    if (c->profUsed == c->profEvents.size()) {
        hipEvent_t a, b;
        PCP_HIP(hipEventCreate(&a));
        PCP_HIP(hipEventCreate(&b));
        c->profEvents.push_back({a, b});
    }
    *slot = (int)c->profUsed++;
    return PCP_OK;
}

// records whether the evaluation just enqueued overflowed a record bucket
__global__ void k_sticky(const int* __restrict__ scal, int Kp, int* __restrict__ sticky) {
    if (threadIdx.x == 0 && scal[SC_BUCKET] > Kp) sticky[0] = 1;
}

__global__ void k_copy(long long n, double* __restrict__ dst, const double* __restrict__ src) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace pcp

// =================================================================== C ABI
extern "C" {

const char* pcp_last_error(void) { return pcp::g_err.c_str(); }

int pcp_device_count(int* count) {
    int n = 0;
    PCP_HIP(hipGetDeviceCount(&n));
    *count = n;
    return PCP_OK;
}
int pcp_malloc(void** ptr, size_t bytes) {
    PCP_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return PCP_OK;
}
int pcp_free(void* ptr) {
    if (ptr) PCP_HIP(hipFree(ptr));
    return PCP_OK;
}
// The plumbing copies and memsets run on the null stream, which is not ordered with the
// contexts' non-blocking streams, and device-to-device copies and memsets may return
// before they are done.  Each call therefore waits for all queued device work first (so it
// sees what earlier kernels wrote and cannot overwrite a buffer they still read) and for
// itself afterwards (so work queued later sees its result).
int pcp_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    if (!bytes) return PCP_OK;
    PCP_HIP(hipDeviceSynchronize());
    PCP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    PCP_HIP(hipDeviceSynchronize());
    return PCP_OK;
}
int pcp_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    if (!bytes) return PCP_OK;
    PCP_HIP(hipDeviceSynchronize());
    PCP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PCP_OK;
}
int pcp_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    if (!bytes) return PCP_OK;
    PCP_HIP(hipDeviceSynchronize());
    PCP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    PCP_HIP(hipDeviceSynchronize());
    return PCP_OK;
}
int pcp_memset(void* ptr, int value, size_t bytes) {
    if (!bytes) return PCP_OK;
    PCP_HIP(hipDeviceSynchronize());
    PCP_HIP(hipMemset(ptr, value, bytes));
    PCP_HIP(hipDeviceSynchronize());
    return PCP_OK;
}

static void free_topology(pcp_ctx* c) {
    dfree(c->dOvScratch);
    c->ovScratchCap = 0;
    dfree(c->dPackPoly); dfree(c->dPackVert); dfree(c->dPolyStart); dfree(c->dShapeId);
    dfree(c->dNext); dfree(c->dPrev); dfree(c->dPolyPack); dfree(c->dBox); dfree(c->dCellOff);
    dfree(c->dCellOf); dfree(c->dCellCount); dfree(c->dCellStart); dfree(c->dCellTmp);
    dfree(c->dCellSorted); dfree(c->dSP); dfree(c->dSR); dfree(c->dSI);
    dfree(c->dNum); dfree(c->dPackMaxKey); dfree(c->dScal);
    dfree(c->dBcnt); dfree(c->dJcnt); dfree(c->dBKey); dfree(c->dBTU); dfree(c->dJKey); dfree(c->dRKey);
    dfree(c->dROuter); dfree(c->dRTU); dfree(c->dRUT); dfree(c->dJC); dfree(c->dJT);
    dfree(c->dF1); dfree(c->dLM); dfree(c->dLN); dfree(c->dAccR); dfree(c->dEr); dfree(c->dLaRep);
    dfree(c->dEpoly); dfree(c->dWork); dfree(c->dRed); dfree(c->dGuard);
    dfree(c->dRedOut); dfree(c->dScanPart); c->scanPartCap = 0;
    dfree(c->dExtStage); c->extStageBytes = 0;
    dfree(c->dXbak); dfree(c->dXpre); dfree(c->dVbak); dfree(c->dDt); dfree(c->dAlpha);
    dfree(c->dFbak); dfree(c->dFext); dfree(c->dFextBak); dfree(c->dSnapX); dfree(c->dSnapV); dfree(c->dSnapF); dfree(c->dSnapD); dfree(c->dSnapI);
    dfree(c->dBlockOver);
    if (c->fireGraph) (void)hipGraphExecDestroy(c->fireGraph);
    c->fireGraph = nullptr;
    dfree(c->dEpre); dfree(c->dNPos); dfree(c->dSteps); dfree(c->dActive); dfree(c->dFlags);
    dfree(c->dFireScal); dfree(c->dShakeIt); dfree(c->dCellInv); dfree(c->dCom); dfree(c->dStatusBlock);
    c->dStatus = nullptr; c->dSummary = nullptr;
    dfree(c->dBounds); dfree(c->dPcSorted); dfree(c->dActOrder); dfree(c->dActCount); dfree(c->dPcStart); dfree(c->dPackB); dfree(c->dSticky);
    c->haveBoxes = false;
}

int pcp_create(pcp_ctx** out) {
    pcp_ctx* c = new pcp_ctx();
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return pcp::fail_hip(e, "hipStreamCreate", __FILE__, __LINE__);
    }
    // the device address of the no-allocation entry point's scratch array, once per device
    int rc = pcp::scratch_setup(c->device);
    if (rc != PCP_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return PCP_OK;
}
int pcp_destroy(pcp_ctx* c) {
    if (!c) return PCP_OK;
    (void)hipStreamSynchronize(c->stream);
    free_topology(c);
    for (auto& e : c->profEvents) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    (void)hipStreamDestroy(c->stream);
    delete c;
    return PCP_OK;
}
int pcp_synchronize(pcp_ctx* c) {
    PCP_HIP(hipStreamSynchronize(c->stream));
    if (!c->dSticky) return PCP_OK;
    int sticky = 0;
    PCP_HIP(hipMemcpy(&sticky, c->dSticky, sizeof(int), hipMemcpyDeviceToHost));
    if (!sticky) return PCP_OK;
    // a fused evaluation since the last check overflowed a record bucket: enlarge the
    // buckets and report it, so that the caller evaluates again
    int scal[SC_NSCAL];
    PCP_HIP(hipMemcpy(scal, c->dScal, sizeof(int) * SC_NSCAL, hipMemcpyDeviceToHost));
    PCP_TRY(pcp::grow_buckets(c, std::max(scal[SC_BUCKET], c->Kp + 1)));
    sticky = 0;
    PCP_HIP(hipMemcpy(c->dSticky, &sticky, sizeof(int), hipMemcpyHostToDevice));
    pcp::set_error("crossing buffers overflowed in an evaluation since the last synchronisation; "
                   "they were enlarged: evaluate again");
    return PCP_ERR_CAPACITY;
}

}  // extern "C"


extern "C" {

int pcp_set_topology(pcp_ctx* c, int P, const int* packPolyStart, const int* polyStart) {
    if (P < 1) return fail_arg("pcp_set_topology: need at least one packing");
    int N = packPolyStart[P];
    if (packPolyStart[0] != 0) return fail_arg("pcp_set_topology: packPolyStart[0] must be 0");
    for (int p = 0; p < P; ++p) {
        if (packPolyStart[p + 1] < packPolyStart[p])
            return fail_arg("pcp_set_topology: packPolyStart must be non-decreasing");
        if (packPolyStart[p + 1] - packPolyStart[p] > 65535)
            return fail_arg("pcp_set_topology: more than 65535 polygons in one packing (16-bit record ids)");
    }
    if (polyStart[0] != 0) return fail_arg("pcp_set_topology: polyStart[0] must be 0");
    for (int s = 0; s < N; ++s) {
        int n = polyStart[s + 1] - polyStart[s];
        if (n < 1) return fail_arg("pcp_set_topology: every polygon needs at least one vertex");
        if (n > 65535) return fail_arg("pcp_set_topology: more than 65535 vertices in one polygon");
    }
    PCP_HIP(hipStreamSynchronize(c->stream));
    free_topology(c);
    int V = polyStart[N];
    c->P = P; c->N = N; c->V = V;
    c->hPackPoly.assign(packPolyStart, packPolyStart + P + 1);
    c->hPolyStart.assign(polyStart, polyStart + N + 1);
    c->hPackVert.resize(P + 1);
    for (int p = 0; p <= P; ++p) c->hPackVert[p] = polyStart[packPolyStart[p]];
    PCP_TRY(dalloc(&c->dPackPoly, P + 1)); PCP_TRY(dalloc(&c->dPackVert, P + 1));
    PCP_TRY(dalloc(&c->dPolyStart, N + 1)); PCP_TRY(dalloc(&c->dShapeId, V));
    PCP_TRY(dalloc(&c->dNext, V)); PCP_TRY(dalloc(&c->dPrev, V)); PCP_TRY(dalloc(&c->dPolyPack, N));
    PCP_TRY(dalloc(&c->dBox, P)); PCP_TRY(dalloc(&c->dCellOff, P + 1));
    PCP_HIP(hipMemcpy(c->dPackPoly, packPolyStart, sizeof(int) * (P + 1), hipMemcpyHostToDevice));
    PCP_HIP(hipMemcpy(c->dPackVert, c->hPackVert.data(), sizeof(int) * (P + 1), hipMemcpyHostToDevice));
    PCP_HIP(hipMemcpy(c->dPolyStart, polyStart, sizeof(int) * (N + 1), hipMemcpyHostToDevice));
    int nt = std::max(V, N);
    if (nt > 0) {
        hipLaunchKernelGGL(k_topology, dim3(nblk(nt)), dim3(BLK), 0, c->stream, V, N, P, c->dPolyStart,
                           c->dPackPoly, c->dShapeId, c->dNext, c->dPrev, c->dPolyPack);
        PCP_LAUNCHED();
    }
    PCP_TRY(dalloc(&c->dCellOf, V)); PCP_TRY(dalloc(&c->dCellTmp, V)); PCP_TRY(dalloc(&c->dCellSorted, V));
    PCP_TRY(dalloc(&c->dSP, V)); PCP_TRY(dalloc(&c->dSR, V)); PCP_TRY(dalloc(&c->dSI, V));
    PCP_TRY(dalloc(&c->dNum, V));
    c->pmkPitch = pmk_pitch(P);
    PCP_TRY(dalloc(&c->dPackMaxKey, PMK_SLOTS * c->pmkPitch)); PCP_TRY(dalloc(&c->dScal, SC_NSCAL));
    PCP_HIP(hipMemsetAsync(c->dScal, 0, sizeof(int) * SC_NSCAL, c->stream));
    PCP_HIP(hipMemsetAsync(c->dNum, 0, sizeof(int) * (size_t)std::max(V, 1), c->stream));
    PCP_TRY(dalloc(&c->dBcnt, N)); PCP_TRY(dalloc(&c->dJcnt, N));
    PCP_HIP(hipMemsetAsync(c->dBcnt, 0, sizeof(int) * (size_t)std::max(N, 1), c->stream));
    PCP_HIP(hipMemsetAsync(c->dJcnt, 0, sizeof(int) * (size_t)std::max(N, 1), c->stream));
    // a polygon of a packing rarely has more than ~10 records; buckets start at 16 and grow on
    // overflow (the exterior kernel stages Kp records per polygon in LDS, so a small Kp keeps
    // more of its workgroups resident)
    PCP_TRY(alloc_buckets(c, 16));
    PCP_TRY(dalloc(&c->dF1, V)); PCP_TRY(dalloc(&c->dLM, V)); PCP_TRY(dalloc(&c->dLN, V));
    PCP_TRY(dalloc(&c->dAccR, N)); PCP_TRY(dalloc(&c->dEpoly, N));
    PCP_TRY(dalloc(&c->dWork, 8LL * V)); PCP_TRY(dalloc(&c->dRed, 4LL * N)); PCP_TRY(dalloc(&c->dRedOut, 4LL * P));
    PCP_TRY(dalloc(&c->dXbak, 2LL * V)); PCP_TRY(dalloc(&c->dXpre, 2LL * V)); PCP_TRY(dalloc(&c->dVbak, 2LL * V));
    PCP_TRY(dalloc(&c->dFbak, 2LL * V)); PCP_TRY(dalloc(&c->dFext, 2LL * V)); PCP_TRY(dalloc(&c->dFextBak, 2LL * V)); PCP_TRY(dalloc(&c->dSnapX, 2LL * V)); PCP_TRY(dalloc(&c->dSnapV, 2LL * V));
    PCP_TRY(dalloc(&c->dSnapF, 2LL * V)); PCP_TRY(dalloc(&c->dSnapD, 4LL * P)); PCP_TRY(dalloc(&c->dSnapI, 3LL * P));
    PCP_TRY(dalloc(&c->dBlockOver, 1));
    PCP_TRY(dalloc(&c->dDt, P)); PCP_TRY(dalloc(&c->dAlpha, P)); PCP_TRY(dalloc(&c->dEpre, P));
    PCP_TRY(dalloc(&c->dNPos, P)); PCP_TRY(dalloc(&c->dSteps, P)); PCP_TRY(dalloc(&c->dActive, P));
    PCP_TRY(dalloc(&c->dFlags, 8LL * P)); PCP_TRY(dalloc(&c->dFireScal, 16LL * P));
    PCP_TRY(dalloc(&c->dShakeIt, 2LL * P));
    PCP_TRY(dalloc(&c->dCellInv, V)); PCP_TRY(dalloc(&c->dCom, 2LL * N));
    PCP_TRY(dalloc(&c->dStatusBlock, 16 + sizeof(pcp_fire_status) * (size_t)P));
    c->dSummary = (int*)c->dStatusBlock;
    c->dStatus = (pcp_fire_status*)(c->dStatusBlock + 16);
    PCP_TRY(detect_alloc(c));
    PCP_TRY(dalloc(&c->dSticky, 1));
    PCP_HIP(hipMemsetAsync(c->dSticky, 0, sizeof(int), c->stream));
    PCP_HIP(hipMemsetAsync(c->dFlags, 0, sizeof(int) * 8 * (size_t)P, c->stream));
    PCP_HIP(hipMemsetAsync(c->dActive, 0, sizeof(int) * (size_t)P, c->stream));
    PCP_HIP(hipStreamSynchronize(c->stream));
    return PCP_OK;
}

int pcp_set_box_sizes(pcp_ctx* c, const int* box) {
    if (c->P == 0) return fail_arg("pcp_set_box_sizes: set the topology first");
    for (int p = 0; p < c->P; ++p)
        if (box[p] < 1) return fail_arg("pcp_set_box_sizes: box size must be >= 1");
    PCP_HIP(hipStreamSynchronize(c->stream));
    c->hBox.assign(box, box + c->P);
    c->hCellOff.assign(c->P + 1, 0);
    for (int p = 0; p < c->P; ++p) c->hCellOff[p + 1] = c->hCellOff[p] + (long long)box[p] * box[p];
    long long Ctot = c->hCellOff[c->P];
    if (Ctot > 2000000000LL) return fail_arg("pcp_set_box_sizes: too many cells");
    dfree(c->dCellCount); dfree(c->dCellStart);
    c->Ctot = Ctot;
    PCP_TRY(dalloc(&c->dCellCount, Ctot)); PCP_TRY(dalloc(&c->dCellStart, Ctot + 1));
    PCP_HIP(hipMemsetAsync(c->dCellCount, 0, sizeof(int) * (size_t)Ctot, c->stream));
    PCP_HIP(hipMemcpy(c->dBox, box, sizeof(int) * c->P, hipMemcpyHostToDevice));
    PCP_HIP(hipMemcpy(c->dCellOff, c->hCellOff.data(), sizeof(long long) * (c->P + 1), hipMemcpyHostToDevice));
    PCP_HIP(hipStreamSynchronize(c->stream));
    c->haveBoxes = true;
    return PCP_OK;
}
int pcp_set_max_neighbors(pcp_ctx* c, int M) {
    if (M < 1) return fail_arg("maxNeighbors must be >= 1");
    c->M = M;
    return PCP_OK;
}
int pcp_get_max_neighbors(pcp_ctx* c, int* M) {
    *M = c->M;
    return PCP_OK;
}

int pcp_update_geometry(pcp_ctx* c, const double* pos, double* edgeLengths, double* areas, double* comX,
                        double* comY, double* constraints, double* maxEdgeLength) {
    return run_geometry(c, (const double2*)pos, edgeLengths, areas, comX, comY, constraints, maxEdgeLength);
}
int pcp_update_cells(pcp_ctx* c, const double* pos) {
    if (!c->haveBoxes) return fail_arg("pcp_update_cells: set the box sizes first");
    return run_cells(c, (const double2*)pos);
}

// detect (cells of the last cell-list build); grow the record buckets when they
// overflowed and repeat.  *maxCount = largest crossing count of one edge.
static int detect_checked(pcp_ctx* c, const double2* pos, int* maxCount) {
    for (int attempt = 0; attempt < 6; ++attempt) {
        PCP_TRY(run_detect(c, pos, 1));
        if (c->V > 0) {
            hipLaunchKernelGGL(k_num_max, dim3(nblk(c->V)), dim3(BLK), 0, c->stream, c->V, c->dNum,
                               c->dScal + SC_MAXCOUNT);
            PCP_LAUNCHED();
        }
        int scal[SC_NSCAL];
        PCP_HIP(hipMemcpyAsync(scal, c->dScal, sizeof(int) * SC_NSCAL, hipMemcpyDeviceToHost, c->stream));
        PCP_HIP(hipStreamSynchronize(c->stream));
        if (scal[SC_BUCKET] > c->Kp) {
            PCP_TRY(grow_buckets(c, scal[SC_BUCKET]));
            continue;
        }
        if (maxCount) *maxCount = scal[SC_MAXCOUNT];
        return PCP_OK;
    }
    set_error("pcp_update_neighbors: record buckets kept overflowing");
    return PCP_ERR_CAPACITY;
}
int pcp_update_neighbors(pcp_ctx* c, const double* pos, int* maxCount) {
    if (!c->haveBoxes) return fail_arg("pcp_update_neighbors: set the box sizes first");
    if (c->V > 0) {   // staged edges for the neighbour-list getter
        hipLaunchKernelGGL(k_restage, dim3(nblk(c->V)), dim3(BLK), 0, c->stream, c->V, (const double2*)pos,
                           c->dNext, c->dCellSorted, c->dSP, c->dSR);
        PCP_LAUNCHED();
    }
    return detect_checked(c, (const double2*)pos, maxCount);
}

int pcp_update_records(pcp_ctx* c, const double* pairPos, int* numIntersections) {
    PCP_TRY(run_records(c, (const double2*)pairPos));
    if (numIntersections) {
        PCP_HIP(hipStreamSynchronize(c->stream));
        std::vector<int> cnt;
        for (int p = 0; p < c->P; ++p) {
            PCP_TRY(packing_counts(c, p, cnt));
            long long I = 0;
            for (int x : cnt) I += x;
            numIntersections[p] = (int)I;
        }
    }
    return PCP_OK;
}

int pcp_update_force_energy(pcp_ctx* c, int model, const double* pos, const double* edgeLengths,
                            const double* areas, const double* targetEdgeV, const double* targetAreas,
                            double stiffness, double compressibility, double* force, double* energy) {
    return run_forces(c, model, (const double2*)pos, (double*)edgeLengths, (double*)areas, targetEdgeV,
                      targetAreas, stiffness, compressibility, (double2*)force, energy, 0, 0);
}
int pcp_project_force(pcp_ctx* c, const double* pos, double* force) {
    return run_project(c, (const double2*)pos, (double2*)force);
}
int pcp_shake(pcp_ctx* c, double* pos, const double* targetEdgeV, const double* targetAreas, int maxIter,
              double tol, int* iters) {
    if (c->N == 0 || maxIter == 0) {
        if (iters) for (int p = 0; p < c->P; ++p) iters[p] = 0;
        return PCP_OK;
    }
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(2LL * c->P)), dim3(BLK), 0, c->stream, c->dShakeIt,
                       2LL * c->P, 0);
    PCP_LAUNCHED();
    PCP_TRY(run_shake(c, (double2*)pos, targetEdgeV, targetAreas, maxIter, tol, 0, nullptr, c->dShakeIt));
    if (iters) {
        PCP_HIP(hipMemcpyAsync(iters, c->dShakeIt, sizeof(int) * c->P, hipMemcpyDeviceToHost, c->stream));
        PCP_HIP(hipStreamSynchronize(c->stream));
    }
    return PCP_OK;
}
static int xpbd_common(pcp_ctx* c, double* pos, const double* targetEdgeV, const double* targetEdgeP,
                       const double* targetAreas, double kE, double kA, double kP, double dt, int nIter, double tol,
                       int* iters) {
    if (c->maxN > 64) return fail_arg("pcp_xpbd: polygons of at most 64 vertices");
    if (nIter < 0) return fail_arg("pcp_xpbd: nIter < 0");
    if ((kE != 0.0 || kA != 0.0 || kP != 0.0) && !(dt > 0.0)) return fail_arg("pcp_xpbd: dt must be > 0");
    if ((kE != 0.0 || kP != 0.0) && !targetEdgeV && !targetEdgeP) return fail_arg("pcp_xpbd: no edge targets");
    if (kA != 0.0 && !targetAreas) return fail_arg("pcp_xpbd: no area targets");
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(2LL * c->P)), dim3(BLK), 0, c->stream, c->dShakeIt, 2LL * c->P, 0);
    PCP_LAUNCHED();
    if (c->N > 0 && nIter > 0)
        PCP_TRY(run_xpbd_seg(c, (double2*)pos, targetEdgeV, targetEdgeP, targetAreas, kE, kA, kP, dt, nullptr,
                             nIter, tol, nullptr, nullptr, c->dShakeIt));
    if (iters) {
        PCP_HIP(hipMemcpyAsync(iters, c->dShakeIt, sizeof(int) * c->P, hipMemcpyDeviceToHost, c->stream));
        PCP_HIP(hipStreamSynchronize(c->stream));
        if (tol <= 0.0)
            for (int p = 0; p < c->P; ++p) iters[p] = nIter;   // the reference's loop runs to nIter
    }
    return PCP_OK;
}
int pcp_xpbd(pcp_ctx* c, double* pos, const double* targetEdgeP, const double* targetAreas, int nIter,
             double tol, int* iters) {
    return xpbd_common(c, pos, nullptr, targetEdgeP, targetAreas, INFINITY, INFINITY, 0.0, 1.0, nIter, tol, iters);
}
int pcp_xpbd_ex(pcp_ctx* c, double* pos, const double* targetEdgeV, const double* targetAreas,
                const pcp_xpbd_params* prm, int* iters) {
    if (!prm) return fail_arg("pcp_xpbd_ex: no parameters");
    return xpbd_common(c, pos, targetEdgeV, nullptr, targetAreas, prm->edgeStiffness, prm->areaStiffness,
                       prm->perimStiffness, prm->dt, prm->nIter, prm->tol, iters);
}
int pcp_max_abs(pcp_ctx* c, const double* x, double* out) {
    if (c->N == 0) return PCP_OK;
    hipLaunchKernelGGL(k_max_abs_poly, dim3(nblk(c->N)), dim3(BLK), 0, c->stream, c->N, c->dPolyStart,
                       (const double2*)x, c->dRed);
    PCP_LAUNCHED();
    return launch_det_reduce(c, c->dRed, 1, out, 1);
}
int pcp_update_positions(pcp_ctx* c, double* pos, const double* force, double dt) {
    if (c->V == 0) return PCP_OK;
    hipLaunchKernelGGL(k_update_positions, dim3(nblk(2LL * c->V)), dim3(BLK), 0, c->stream, 2LL * c->V, pos,
                       force, dt);
    PCP_LAUNCHED();
    return PCP_OK;
}
int pcp_reset_areas(pcp_ctx* c, double* pos, const double* areas, const double* targetAreas,
                    const double* comX, const double* comY) {
    if (c->V == 0) return PCP_OK;
    hipLaunchKernelGGL(k_reset_areas, dim3(nblk(2LL * c->V)), dim3(BLK), 0, c->stream, 2LL * c->V,
                       c->dShapeId, pos, areas, targetAreas, comX, comY);
    PCP_LAUNCHED();
    return PCP_OK;
}
int pcp_scale_targets(pcp_ctx* c, const double* areaFactor, const double* edgeFactor, double* targetAreas,
                      double* targetEdgeV) {
    if ((targetAreas && !areaFactor) || (targetEdgeV && !edgeFactor))
        return fail_arg("pcp_scale_targets: a target array without its factors");
    if (c->N == 0 || (!targetAreas && !targetEdgeV)) return PCP_OK;
    // the factors go to the per-packing reduction buffer (4P doubles), free between calls
    if (targetAreas)
        PCP_HIP(hipMemcpyAsync(c->dRedOut, areaFactor, sizeof(double) * c->P, hipMemcpyHostToDevice, c->stream));
    if (targetEdgeV)
        PCP_HIP(hipMemcpyAsync(c->dRedOut + c->P, edgeFactor, sizeof(double) * c->P, hipMemcpyHostToDevice,
                               c->stream));
    hipLaunchKernelGGL(k_scale_targets, dim3(nblk(std::max(c->V, c->N))), dim3(BLK), 0, c->stream, c->V, c->N, c->P,
                       c->dShapeId, c->dPolyPack, c->dRedOut, targetAreas, targetEdgeV);
    PCP_LAUNCHED();
    return PCP_OK;
}
int pcp_shape_project(pcp_ctx* c, double* pos, const double* refShape, const double* targetAreas) {
    if (c->N == 0) return PCP_OK;
    if (c->maxN > 64) {
        set_error("pcp_shape_project: polygons of at most 64 vertices");
        return PCP_ERR_ARG;
    }
    return run_shape_project_seg(c, (double2*)pos, (const double2*)refShape, targetAreas);
}
int pcp_max_effective_force(pcp_ctx* c, const double* pos, const double* tentPos, const double* force,
                            double scale, double* out) {
    if (c->N == 0) return PCP_OK;
    hipLaunchKernelGGL(k_eff_force_poly, dim3(nblk(c->N)), dim3(BLK), 0, c->stream, c->N, c->dPolyStart,
                       (const double2*)pos, (const double2*)tentPos, (const double2*)force, scale, c->dRed);
    PCP_LAUNCHED();
    PCP_TRY(launch_det_reduce(c, c->dRed, 1, c->dRedOut, 1));
    PCP_HIP(hipMemcpyAsync(out, c->dRedOut, sizeof(double) * c->P, hipMemcpyDeviceToHost, c->stream));
    PCP_HIP(hipStreamSynchronize(c->stream));
    return PCP_OK;
}
// ---- getters
int pcp_get_cells(pcp_ctx* c, int p, int* cellLocation, int* neighborIndices, int* countPerBox) {
    if (p < 0 || p >= c->P) return fail_arg("packing out of range");
    PCP_HIP(hipStreamSynchronize(c->stream));
    int v0 = c->hPackVert[p], v1 = c->hPackVert[p + 1];
    if (cellLocation) PCP_HIP(hipMemcpy(cellLocation, c->dCellOf + v0, sizeof(int) * (v1 - v0), hipMemcpyDeviceToHost));
    if (neighborIndices) {
        PCP_HIP(hipMemcpy(neighborIndices, c->dCellSorted + v0, sizeof(int) * (v1 - v0), hipMemcpyDeviceToHost));
        for (int i = 0; i < v1 - v0; ++i) neighborIndices[i] -= v0;
    }
    if (countPerBox) {
        long long nb = (long long)c->hBox[p] * c->hBox[p];
        PCP_HIP(hipMemcpy(countPerBox, c->dCellStart + c->hCellOff[p], sizeof(int) * nb, hipMemcpyDeviceToHost));
        for (long long i = 0; i < nb; ++i) countPerBox[i] -= v0;
    }
    return PCP_OK;
}
int pcp_get_num_intersections(pcp_ctx* c, int p, int* count) {
    if (p < 0 || p >= c->P) return fail_arg("packing out of range");
    PCP_HIP(hipStreamSynchronize(c->stream));
    std::vector<int> cnt;
    PCP_TRY(packing_counts(c, p, cnt));
    long long I = 0;
    for (int x : cnt) I += x;
    *count = (int)I;
    return PCP_OK;
}
// records of packing p in the reference's final order: polygon by polygon (si), each
// polygon's records by (i, j, sj)
int pcp_get_records(pcp_ctx* c, int p, uint64_t* inter, double* tu, uint64_t* outer, double* ut,
                    int* shapeCounts) {
    if (p < 0 || p >= c->P) return fail_arg("packing out of range");
    PCP_HIP(hipStreamSynchronize(c->stream));
    int n0 = c->hPackPoly[p], n1 = c->hPackPoly[p + 1];
    std::vector<int> cnt;
    PCP_TRY(packing_counts(c, p, cnt));
    if (shapeCounts) for (int s = 0; s < n1 - n0; ++s) shapeCounts[s] = cnt[s];
    if (n1 == n0) return PCP_OK;
    size_t rows = (size_t)(n1 - n0) * c->Kp, off = (size_t)n0 * c->Kp;
    std::vector<uint64_t> k1(rows), k2(rows);
    std::vector<float2> t1(rows), t2(rows);
    PCP_HIP(hipMemcpy(k1.data(), c->dRKey + off, sizeof(uint64_t) * rows, hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(k2.data(), c->dROuter + off, sizeof(uint64_t) * rows, hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(t1.data(), c->dRTU + off, sizeof(float2) * rows, hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(t2.data(), c->dRUT + off, sizeof(float2) * rows, hipMemcpyDeviceToHost));
    long long w = 0;
    for (int s = 0; s < n1 - n0; ++s)
        for (int f = 0; f < cnt[s]; ++f, ++w) {
            size_t r = (size_t)s * c->Kp + f;
            if (inter) inter[w] = k1[r];
            if (outer) outer[w] = k2[r];
            if (tu) { tu[2 * w] = t1[r].x; tu[2 * w + 1] = t1[r].y; }
            if (ut) { ut[2 * w] = t2[r].x; ut[2 * w + 1] = t2[r].y; }
        }
    return PCP_OK;
}
int pcp_get_neighbor_lists(pcp_ctx* c, int p, int* neighbors, unsigned char* inside, int* numNeighbors) {
    if (p < 0 || p >= c->P) return fail_arg("packing out of range");
    int v0 = c->hPackVert[p], v1 = c->hPackVert[p + 1];
    long long n = (long long)(v1 - v0) * c->M;
    int* dN = nullptr; unsigned char* dI = nullptr; int* dC = nullptr;
    PCP_TRY(dalloc(&dN, n)); PCP_TRY(dalloc(&dI, n)); PCP_TRY(dalloc(&dC, v1 - v0));
    PCP_HIP(hipMemsetAsync(dN, 0xFF, sizeof(int) * n, c->stream));
    PCP_HIP(hipMemsetAsync(dI, 0, n, c->stream));
    if (v1 > v0) {
        hipLaunchKernelGGL(k_neighbor_lists, dim3(nblk(v1 - v0)), dim3(BLK), 0, c->stream, v0, v1, c->M,
                           c->dShapeId, c->dPolyPack, c->dBox, c->dCellOff, c->dCellOf, c->dCellStart,
                           c->dNext, c->dSP, c->dSR, c->dSI, c->dCellInv, dN, dI, dC);
        PCP_LAUNCHED();
    }
    PCP_HIP(hipStreamSynchronize(c->stream));
    if (neighbors) PCP_HIP(hipMemcpy(neighbors, dN, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (inside) PCP_HIP(hipMemcpy(inside, dI, n, hipMemcpyDeviceToHost));
    if (numNeighbors) PCP_HIP(hipMemcpy(numNeighbors, dC, sizeof(int) * (v1 - v0), hipMemcpyDeviceToHost));
    dfree(dN); dfree(dI); dfree(dC);
    return PCP_OK;
}
int pcp_get_compacted_records(pcp_ctx* c, int p, uint64_t* inter, double* tu) {
    if (p < 0 || p >= c->P) return fail_arg("packing out of range");
    PCP_HIP(hipStreamSynchronize(c->stream));
    int v0 = c->hPackVert[p], v1 = c->hPackVert[p + 1], nv = v1 - v0;
    int n0 = c->hPackPoly[p], n1 = c->hPackPoly[p + 1];
    int B = c->hBox[p];
    std::vector<int> cnt, cell(nv);
    PCP_TRY(packing_counts(c, p, cnt));
    if (nv) PCP_HIP(hipMemcpy(cell.data(), c->dCellOf + v0, sizeof(int) * nv, hipMemcpyDeviceToHost));
    size_t rows = (size_t)(n1 - n0) * c->Kp, off = (size_t)n0 * c->Kp;
    std::vector<uint64_t> keys(rows);
    std::vector<float2> tus(rows);
    if (rows) {
        PCP_HIP(hipMemcpy(keys.data(), c->dBKey + off, sizeof(uint64_t) * rows, hipMemcpyDeviceToHost));
        PCP_HIP(hipMemcpy(tus.data(), c->dBTU + off, sizeof(float2) * rows, hipMemcpyDeviceToHost));
    }
    // compaction order of updateCompactedIntersectionsKernel (kernels.cuh:838-877): edge by
    // edge (the edge with the lower vertex id of each crossing), each edge's crossings in
    // the order updateNeighborsKernel found them -- 3x3 stencil cells in (dx, dy) order, a
    // cell counted once, ascending vertex id inside a cell.
    struct Item { int id1; long long ord; uint64_t key; float2 tu; };
    std::vector<Item> items;
    const int* ps = c->hPolyStart.data();
    for (int s = 0; s < n1 - n0; ++s)
        for (int f = 0; f < cnt[s]; ++f) {
            uint64_t k = keys[(size_t)s * c->Kp + f];
            int sj = (int)(k >> 48), si = (int)((k >> 32) & 0xFFFF);
            int vi = ps[n0 + si] + (int)((k >> 16) & 0xFFFF) - v0;   // edge i of si
            int vj = ps[n0 + sj] + (int)(k & 0xFFFF) - v0;           // edge j of sj
            int id1 = std::min(vi, vj), nid = std::max(vi, vj);
            int bx = cell[id1] % B, by = cell[id1] / B, oc = cell[nid];
            int first = 9;
            for (int kc = 0; kc < 9 && first == 9; ++kc) {
                int nx = (bx + kc / 3 - 1 + B) % B, ny = (by + kc % 3 - 1 + B) % B;
                if (ny * B + nx == oc) first = kc;
            }
            items.push_back({id1, (long long)first * (1LL << 32) + nid, k, tus[(size_t)s * c->Kp + f]});
        }
    std::sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
        return a.id1 != b.id1 ? a.id1 < b.id1 : a.ord < b.ord;
    });
    for (size_t w = 0; w < items.size(); ++w) {
        if (inter) inter[w] = items[w].key;
        if (tu) { tu[2 * w] = items[w].tu.x; tu[2 * w + 1] = items[w].tu.y; }
    }
    return PCP_OK;
}

int pcp_profile_kernel(pcp_ctx* c, const char* name) {
    PCP_HIP(hipStreamSynchronize(c->stream));
    c->profName = name ? name : "";
    c->profUsed = 0;
    return PCP_OK;
}
int pcp_profile_read(pcp_ctx* c, double* totalMs, int* launches) {
    PCP_HIP(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (size_t k = 0; k < c->profUsed; ++k) {
        float ms = 0.f;
        PCP_HIP(hipEventElapsedTime(&ms, c->profEvents[k].first, c->profEvents[k].second));
        tot += ms;
    }
    *totalMs = tot;
    *launches = (int)c->profUsed;
    return PCP_OK;
}
int pcp_set_detect_schedule(pcp_ctx* c, int mode) {
    if (mode < -1 || mode > 1) return fail_arg("pcp_set_detect_schedule: mode must be -1, 0 or 1");
    c->detSched = mode;
    c->graphK = 0;   // a captured FIRE block holds the old kernels
    return PCP_OK;
}
int pcp_get_detect_schedule(pcp_ctx* c, int* mode) {
    *mode = detect_schedule(c);
    return PCP_OK;
}
int pcp_set_schedule(pcp_ctx* c, int pairLanes, int lensLanes, int exteriorLanes) {
    auto ok = [](int w, bool zero) { return w == 16 || w == 32 || w == 64 || (zero && w == 0); };
    if (!ok(pairLanes, false) || !ok(lensLanes, false) || !ok(exteriorLanes, true)) {
        set_error("pcp_set_schedule: lanes per polygon must be 16, 32 or 64 (exterior also 0 = by LDS)");
        return PCP_ERR_ARG;
    }
    PCP_HIP(hipStreamSynchronize(c->stream));
    c->segPair = pairLanes;
    c->segLens = lensLanes;
    c->segExt = exteriorLanes;
    c->graphK = 0;
    return PCP_OK;
}
int pcp_get_schedule(pcp_ctx* c, int* pairLanes, int* lensLanes, int* exteriorLanes) {
    if (pairLanes) *pairLanes = c->segPair;
    if (lensLanes) *lensLanes = c->segLens;
    if (exteriorLanes) *exteriorLanes = c->segExt;
    return PCP_OK;
}
int pcp_autotune(pcp_ctx* c, const pcp_state* st, int reps, double* timesUs) {
    if (reps < 1 || reps > 10000) {
        set_error("pcp_autotune: reps must be in 1..10000");
        return PCP_ERR_ARG;
    }
    return autotune(c, st, reps, timesUs);
}
int pcp_stream_copy(pcp_ctx* c, double* dst, const double* src, long long n) {
    int g = (int)std::min<long long>(nblk(n), 8192);
    PCP_LAUNCH(c, k_copy, dim3(std::max(g, 1)), dim3(BLK), n, dst, src);
    return PCP_OK;
}

int pcp_detect_stats(pcp_ctx* c, long long* polygonPairs, long long* edgePairs, long long* screened) {
    if (!polygonPairs || !edgePairs) return fail_arg("pcp_detect_stats: polygonPairs and edgePairs are required");
    return detect_stats(c, polygonPairs, edgePairs, screened);
}
int pcp_odd_crossing_records(pcp_ctx* c, int* count) {
    if (!count) return fail_arg("pcp_odd_crossing_records: count is null");
    *count = 0;
    if (!c->dScal) return PCP_OK;
    PCP_HIP(hipStreamSynchronize(c->stream));
    PCP_HIP(hipMemcpy(count, c->dScal + SC_ODD, sizeof(int), hipMemcpyDeviceToHost));
    return PCP_OK;
}

int pcp_sort_keys(const uint64_t* keys, int n, int endBit, uint64_t* sorted, uint32_t* perm) {
    if (n <= 0) return PCP_OK;
    uint64_t mask = (endBit >= 64) ? ~0ull : ((endBit <= 0) ? 0ull : ((1ull << endBit) - 1));
    uint64_t *din = nullptr, *dout = nullptr;
    uint32_t* dp = nullptr;
    PCP_TRY(dalloc(&din, n)); PCP_TRY(dalloc(&dout, n)); PCP_TRY(dalloc(&dp, n));
    PCP_HIP(hipMemcpy(din, keys, sizeof(uint64_t) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_sort_keys, dim3(nblk(n)), dim3(BLK), 0, 0, n, mask, din, dout, dp);
    PCP_LAUNCHED();
    PCP_HIP(hipMemcpy(sorted, dout, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    PCP_HIP(hipMemcpy(perm, dp, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    dfree(din); dfree(dout); dfree(dp);
    return PCP_OK;
}

int pcp_force_eval(pcp_ctx* c, const pcp_state* st) {
    const double2* pos = (const double2*)st->pos;
    if (!c->haveBoxes) return fail_arg("pcp_force_eval: set the box sizes first");
    int mt = st->model & 0xFF;
    int overlap = (mt == PCP_NORMAL || mt == PCP_HYBRID);
    if (overlap) {
        PCP_TRY(run_detect(c, pos, 0, nullptr, (st->model & PCP_CONSISTENT) != 0));
        PCP_TRY(run_records(c, (st->model & PCP_AREA_MODES) ? pos : nullptr, true));
    }
    // the overflow check rides on the energy reduction when there is one
    bool fused = overlap && st->energy;
    PCP_TRY(run_forces(c, st->model, pos, st->edgeLengths, st->areas, st->targetEdgeV, st->targetAreas,
                       st->stiffness, st->compressibility, (double2*)st->force, st->energy, 1,
                       mt == PCP_NORMAL, fused, st->perimStiffness));
    if (overlap && !fused) {
        hipLaunchKernelGGL(k_sticky, dim3(1), dim3(64), 0, c->stream, c->dScal, c->Kp, c->dSticky);
        PCP_LAUNCHED();
    }
    return PCP_OK;
}

}  // extern "C"

#include "pcp_overlap.hip"
#include "pcp_fire.hip"
#include "pcp_mpnn.hip"
