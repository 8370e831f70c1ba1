/*
 * CPU oracle: plain-C restatement of the reference overlap-force-minimization
 * path (camdennis/pyCudaPolygon: src/kernels.cuh, src/model.cu, src/model.cpp,
 * pyCudaPolygon/pyCudaPolygon.py).
 *
 * TEST INFRASTRUCTURE ONLY -- this is the checker, never the product.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off -fno-fast-math -fopenmp).
 * No FMA contraction: every expression is evaluated operation by operation in
 * the order written in the reference source, so crossing decisions and
 * float-rounded intersection parameters are reproduced exactly.
 *
 * Where the reference accumulates with double atomicAdd (order unspecified,
 * kernels.cuh:1143-1144,1271,1330-1350,1415-1422) this oracle fixes ONE valid
 * serialization, the same one the HIP kernels use:
 *   force[m] = 0
 *            + exterior contributions of records whose si owns m (record order,
 *              slots i, zi, l, zl)
 *            + exterior contributions of records whose sj owns m (record order,
 *              slots j, zj, k, zk)
 *            + interior local sum of thread m  (kernels.cuh:1416-1419)
 *            + interior local sum of thread prev(m) for next(prev(m)) = m
 *            + edge spring term, + area spring term.
 * Energies: per polygon, sequential; packings reduce polygon energies with a
 * fixed 256-wide pairwise tree (orc_det_sum).  Parity of these sums against the
 * reference is therefore exact up to the reference's own atomic ordering.
 *
 * Functions whose reference kernels are missing from the reference tree
 * (FIRE.h, shake.h; src/model.cu:15-16) are written from the host contract in
 * src/model.cpp:405-504 and src/model.cu:576-636; see the notes at each.
 */
#include "pcp_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

typedef struct { double x, y; } d2;

/* kernels.cuh:115-119 */
double orc_wrap(double x) {
    double y = x + 1.5;
    y = y - floor(y);
    return y - 0.5;
}
static d2 wrap2(d2 v) { v.x = orc_wrap(v.x); v.y = orc_wrap(v.y); return v; }
static d2 P(const double* pos, int v) { d2 r = { pos[2 * v], pos[2 * v + 1] }; return r; }

/* while-loop minimum image used by kernels.cuh:670-699 and :1303-1316.  Loops are
 * bounded; for |x| beyond 2^20 (never produced by positions in the box) floor is used. */
static double shift_into_unit(double x) {   /* while(x<0)x+=1; while(x>=1)x-=1 */
    if (!(fabs(x) < 1048576.0)) return (x == x) ? x - floor(x) : x;
    while (x < 0.0) x += 1.0;
    while (x >= 1.0) x -= 1.0;
    return x;
}

/* kernels.cuh:127-135, 652-659 */
void orc_topology(int N, const int* start, int V, int* shapeId, int* next, int* prev) {
    for (int s = 0; s < N; ++s)
        for (int i = start[s]; i < start[s + 1]; ++i) shapeId[i] = s;
    for (int i = 0; i < V; ++i) {
        int s = shapeId[i];
        next[i] = (i + 1 == start[s + 1]) ? start[s] : i + 1;
        prev[i] = (i == start[s]) ? start[s + 1] - 1 : i - 1;
    }
}

/* updatePolygonGeometryKernel kernels.cuh:295-333, segmented sums model.cu:123-140
 * (sequential per polygon), normalizeKernel :635-650, normalizeConstraintsKernel
 * :335-345, max edge length model.cu:144-150. */
void orc_geometry(int V, int N, const double* pos, const int* start,
                  double* edgeLengths, double* areas, double* comX, double* comY,
                  double* constraints, double* constraintNormSq, double* maxEdgeLength) {
    double mx = -DBL_MAX;
    for (int s = 0; s < N; ++s) {
        int st = start[s], en = start[s + 1], n = en - st;
        double area = 0.0, cx = 0.0, cy = 0.0, ns0 = 0.0, ns1 = 0.0, ns2 = 0.0;
        for (int k = st; k < en; ++k) {
            int nx = (k + 1 == en) ? st : k + 1;
            int pv = (k == st) ? en - 1 : k - 1;
            for (int alpha = 0; alpha < 2; ++alpha) {
                int idx = 2 * k + alpha;
                double x1 = orc_wrap(pos[idx] - pos[st * 2 + alpha]);
                double x4 = orc_wrap(pos[pv * 2 + 1 - alpha] - pos[nx * 2 + 1 - alpha]);
                double dxPrev = orc_wrap(pos[2 * k] - pos[pv * 2]);
                double dyPrev = orc_wrap(pos[2 * k + 1] - pos[pv * 2 + 1]);
                double dxNext = orc_wrap(pos[2 * k] - pos[nx * 2]);
                double dyNext = orc_wrap(pos[2 * k + 1] - pos[nx * 2 + 1]);
                if (alpha == 0) cx += x1; else cy += x1;
                double lenPrev = sqrt(dxPrev * dxPrev + dyPrev * dyPrev);
                double lenNext = sqrt(dxNext * dxNext + dyNext * dyNext);
                double areaComp = (2 * alpha - 1) * x4;
                double edge2Comp = alpha ? (dyNext + dyPrev) : (dxNext + dxPrev);
                double edgeComp = alpha ? (dyNext / lenNext) + (dyPrev / lenPrev)
                                        : (dxNext / lenNext) + (dxPrev / lenPrev);
                constraints[6 * k + alpha] = areaComp;
                constraints[6 * k + 2 + alpha] = edge2Comp;
                constraints[6 * k + 4 + alpha] = edgeComp;
                ns0 += areaComp * areaComp;
                ns1 += edge2Comp * edge2Comp;
                ns2 += edgeComp * edgeComp;
                if (alpha == 0) {
                    double dx = orc_wrap(pos[2 * nx] - pos[2 * k]);
                    double dy = orc_wrap(pos[2 * nx + 1] - pos[2 * k + 1]);
                    double el = sqrt(dx * dx + dy * dy);
                    edgeLengths[k] = el;
                    if (el > mx) mx = el;
                    double dy2 = orc_wrap(pos[nx * 2 + 1] - pos[idx + 1]);
                    double x2 = orc_wrap(pos[nx * 2 + alpha] - pos[st * 2 + alpha]);
                    area += dy2 * (x2 + x1) / 2.0;
                }
            }
        }
        areas[s] = area;
        double c = cx;
        c /= ((double)n);
        c += pos[2 * st];
        c -= floor(c);
        comX[s] = c;
        c = cy;
        c /= ((double)n);
        c += pos[2 * st + 1];
        c -= floor(c);
        comY[s] = c;
        if (constraintNormSq) {
            constraintNormSq[3 * s] = ns0; constraintNormSq[3 * s + 1] = ns1; constraintNormSq[3 * s + 2] = ns2;
        }
        double nsv[3] = { ns0, ns1, ns2 };
        for (int k = st; k < en; ++k)
            for (int alpha = 0; alpha < 2; ++alpha)
                for (int cc = 0; cc < 3; ++cc)
                    if (nsv[cc] > 0.0) constraints[6 * k + 2 * cc + alpha] /= sqrt(nsv[cc]);
    }
    if (maxEdgeLength) *maxEdgeLength = (V > 0) ? mx : 0.0;
}

/* cell of an edge midpoint: updateNeighborCellsKernel kernels.cuh:661-707 */
static int cell_of_edge(const double* pos, int idx, int v2, int boxSize) {
    double x = pos[v2 * 2] - pos[idx * 2] + 0.5;
    double y = pos[v2 * 2 + 1] - pos[idx * 2 + 1] + 0.5;
    x = shift_into_unit(x);
    y = shift_into_unit(y);
    x -= 0.5; y -= 0.5;
    x /= 2.0; y /= 2.0;
    x += pos[idx * 2];
    y += pos[idx * 2 + 1];
    x = shift_into_unit(x);
    y = shift_into_unit(y);
    double fx = floor(x * boxSize), fy = floor(y * boxSize);
    int ix = (fx == fx && fx >= 0.0) ? (fx < 2147483647.0 ? (int)fx : boxSize - 1) : 0;
    int iy = (fy == fy && fy >= 0.0) ? (fy < 2147483647.0 ? (int)fy : boxSize - 1) : 0;
    if (ix < 0) ix = 0; else if (ix >= boxSize) ix = boxSize - 1;
    if (iy < 0) iy = 0; else if (iy >= boxSize) iy = boxSize - 1;
    return iy * boxSize + ix;
}

/* updateNeighborCellsCUDA model.cu:308-385: hash, stable sort of vertex ids by cell
 * (CUB radix SortPairs over iota values), lowerBoundKernel kernels.cuh:45-63. */
void orc_cells(int V, int N, const double* pos, const int* start, int boxSize,
               int* cellLocation, int* neighborIndices, int* countPerBox) {
    int nb = boxSize * boxSize;
    int* shapeId = (int*)malloc(sizeof(int) * (V > 0 ? V : 1));
    for (int s = 0; s < N; ++s) for (int i = start[s]; i < start[s + 1]; ++i) shapeId[i] = s;
    for (int idx = 0; idx < V; ++idx) {
        int v2 = idx + 1;
        if (idx == V - 1 || shapeId[idx] != shapeId[idx + 1]) v2 = start[shapeId[idx]];
        cellLocation[idx] = cell_of_edge(pos, idx, v2, boxSize);
    }
    int* cnt = (int*)calloc(nb + 1, sizeof(int));
    for (int i = 0; i < V; ++i) cnt[cellLocation[i] + 1]++;
    for (int c = 0; c < nb; ++c) cnt[c + 1] += cnt[c];
    for (int c = 0; c < nb; ++c) countPerBox[c] = cnt[c];   /* lower bound of c */
    for (int i = 0; i < V; ++i) neighborIndices[cnt[cellLocation[i]]++] = i;  /* stable */
    free(cnt);
    free(shapeId);
}

/* updateNeighborsKernel kernels.cuh:709-812.  tu holds double2 per slot [2*V*M].
 * Returns max(numNeighbors) (maxReduceKernel :65-92). */
/* Side of point X relative to the line of the edge from vertex e0 to vertex e1: the sign of
 * wrap(e1 - e0) x wrap(X - e0), as 1 (negative) or 0 (zero or positive).  A function of the
 * edge's two vertices and X only, so every test that involves the same edge and point agrees. */
static int orc_side(const double* pos, int e0, int e1, int X) {
    double ex = orc_wrap(pos[2 * e1] - pos[2 * e0]), ey = orc_wrap(pos[2 * e1 + 1] - pos[2 * e0 + 1]);
    double wx = orc_wrap(pos[2 * X] - pos[2 * e0]), wy = orc_wrap(pos[2 * X + 1] - pos[2 * e0 + 1]);
    return (ex * wy - ey * wx) < 0.0;
}

int orc_neighbors(int V, int N, const double* pos, const int* start, int boxSize,
                  const int* cellLocation, const int* neighborIndices, const int* countPerBox,
                  int maxNeighbors, int* neighbors, unsigned char* inside, double* tu,
                  int* numNeighbors) {
    return orc_neighbors_ex(V, N, pos, start, boxSize, cellLocation, neighborIndices, countPerBox,
                            maxNeighbors, neighbors, inside, tu, numNeighbors, 0);
}

int orc_neighbors_ex(int V, int N, const double* pos, const int* start, int boxSize,
                     const int* cellLocation, const int* neighborIndices, const int* countPerBox,
                     int maxNeighbors, int* neighbors, unsigned char* inside, double* tu,
                     int* numNeighbors, int consistent) {
    const double eps = 1e-12;
    int* shapeId = (int*)malloc(sizeof(int) * (V > 0 ? V : 1));
    for (int s = 0; s < N; ++s) for (int i = start[s]; i < start[s + 1]; ++i) shapeId[i] = s;
    int boxCount = boxSize * boxSize;
    int mx = (V > 0) ? -2147483647 - 1 : 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : mx)
    for (int id1 = 0; id1 < V; ++id1) {
        int shape = shapeId[id1];
        int st = start[shape + 1] - 1;
        int id2 = (id1 == st) ? start[shape] : id1 + 1;
        const double px = pos[2 * id1], py = pos[2 * id1 + 1];
        double rx = orc_wrap(pos[2 * id2] - px);
        double ry = orc_wrap(pos[2 * id2 + 1] - py);
        int box = cellLocation[id1];
        int bx = box % boxSize, by = box / boxSize;
        int count = 0;
        int processed[16], np = 0;
        for (int dx = -1; dx <= 1; dx++) {
            int nx = bx + dx;
            if (nx < 0) nx += boxSize; else if (nx >= boxSize) nx -= boxSize;
            for (int dy = -1; dy <= 1; dy++) {
                int ny = by + dy;
                if (ny < 0) ny += boxSize; else if (ny >= boxSize) ny -= boxSize;
                int nbx = ny * boxSize + nx;
                int seen = 0;
                for (int b = 0; b < np; b++) if (processed[b] == nbx) { seen = 1; break; }
                if (seen) continue;
                if (np < 16) processed[np++] = nbx;
                int si = countPerBox[nbx];
                int sf = (nbx + 1 < boxCount) ? countPerBox[nbx + 1] : V;
                if (si >= sf) continue;
                for (int q = si; q < sf; q++) {
                    int nid = neighborIndices[q];
                    if (nid <= id1) continue;
                    int shape2 = shapeId[nid];
                    int st2 = start[shape2 + 1] - 1;
                    int nid2 = (nid == st2) ? start[shape2] : nid + 1;
                    if (nid == id1 || nid == id2 || nid2 == id1 || nid2 == id2) continue;
                    double qx = pos[2 * nid], qy = pos[2 * nid + 1];
                    double sx = orc_wrap(pos[2 * nid2] - qx);
                    double sy = orc_wrap(pos[2 * nid2 + 1] - qy);
                    double gx = orc_wrap(qx - px);
                    double gy = orc_wrap(qy - py);
                    double denom = rx * sy - ry * sx;
                    if (fabs(denom) < eps) continue;
                    double tt = (gx * sy - gy * sx) / denom;
                    double uu = (gx * ry - gy * rx) / denom;
                    int hit = consistent
                        ? (orc_side(pos, nid, nid2, id1) != orc_side(pos, nid, nid2, id2) &&
                           orc_side(pos, id1, id2, nid) != orc_side(pos, id1, id2, nid2))
                        : (tt > 0.0 && tt < 1.0 && uu > 0.0 && uu < 1.0);
                    if (hit) {
                        if (count < maxNeighbors) {
                            long slot = (long)id1 * maxNeighbors + count;
                            neighbors[slot] = nid;
                            if (denom > 0) { inside[slot] = 1; tu[2 * slot] = tt; tu[2 * slot + 1] = uu; }
                            else { inside[slot] = 0; tu[2 * slot] = uu; tu[2 * slot + 1] = tt; }
                        }
                        count++;
                    }
                }
            }
        }
        numNeighbors[id1] = count;
        if (count > mx) mx = count;
    }
    free(shapeId);
    return mx;
}

typedef struct { uint64_t key; int idx; } keyidx;
static int cmp_keyidx(const void* a, const void* b) {
    const keyidx* x = (const keyidx*)a; const keyidx* y = (const keyidx*)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx);   /* stable */
}

/* Model::updateOutersections model.cpp:351-361:
 *   updateValidAndCountsKernel kernels.cuh:814-836 + exclusive scan model.cu:413-450,
 *   updateCompactedIntersectionsKernel :838-877 (tu rounded to float),
 *   sortKeysCUDA(0,64) model.cu:50-75 + gather, updateOutersectionsKernel :981-1065,
 *   sortKeysCUDA(0,48) + gathers.
 * tuSlots is the V*M double2 array written by orc_neighbors.  Returns I. */
static double cross_param(d2 vi, d2 vzi, d2 vj, d2 vzj, double fallback);
/* exact parameter along si's edge `ei` of its crossing with sj's edge `ej` */
static double exact_param(const double* pos, const int* start, int si, int ei, int sj, int ej,
                          double fallback) {
    int ni = start[si + 1] - start[si], nj = start[sj + 1] - start[sj];
    int a = start[si] + ei, b = start[si] + (ei + 1) % ni;
    int c = start[sj] + ej, d = start[sj] + (ej + 1) % nj;
    return cross_param(P(pos, a), P(pos, b), P(pos, c), P(pos, d), fallback);
}

int orc_outersections(int V, int N, const int* start, int maxNeighbors,
                      const int* neighbors, const unsigned char* inside, const double* tuSlots,
                      const int* numNeighbors, uint64_t* intersections, double* tu,
                      uint64_t* outersections, double* ut, int* shapeCounts) {
    return orc_outersections_ex(V, N, start, maxNeighbors, neighbors, inside, tuSlots, numNeighbors,
                                intersections, tu, outersections, ut, shapeCounts, NULL);
}

/* pairExactPos != NULL (area energy): the pairing compares exact double parameters
 * recomputed at these positions instead of the float-rounded ones. */
int orc_outersections_ex(int V, int N, const int* start, int maxNeighbors,
                         const int* neighbors, const unsigned char* inside, const double* tuSlots,
                         const int* numNeighbors, uint64_t* intersections, double* tu,
                         uint64_t* outersections, double* ut, int* shapeCounts,
                         const double* pairExactPos) {
    int* shapeId = (int*)malloc(sizeof(int) * (V > 0 ? V : 1));
    for (int s = 0; s < N; ++s) for (int i = start[s]; i < start[s + 1]; ++i) shapeId[i] = s;
    for (int s = 0; s < N; ++s) shapeCounts[s] = 0;
    /* compaction in (id1, slot) order */
    int I = 0;
    for (int n1 = 0; n1 < V; ++n1) {
        int c = numNeighbors[n1] < maxNeighbors ? numNeighbors[n1] : maxNeighbors;
        for (int k = 0; k < c; ++k) {
            long slot = (long)n1 * maxNeighbors + k;
            int n2 = neighbors[slot];
            int s1 = shapeId[n1], s2 = shapeId[n2];
            int in = inside[slot];
            shapeCounts[in ? s2 : s1]++;
            int l1 = n1 - start[s1], l2 = n2 - start[s2];
            float tVal = (float)tuSlots[2 * slot];
            float uVal = (float)tuSlots[2 * slot + 1];
            uint64_t packed;
            if (in) packed = ((uint64_t)(uint16_t)s1 << 48) | ((uint64_t)(uint16_t)s2 << 32) |
                             ((uint64_t)(uint16_t)l2 << 16) | (uint64_t)(uint16_t)l1;
            else packed = ((uint64_t)(uint16_t)s2 << 48) | ((uint64_t)(uint16_t)s1 << 32) |
                          ((uint64_t)(uint16_t)l1 << 16) | (uint64_t)(uint16_t)l2;
            intersections[I] = packed;
            tu[2 * I] = (double)tVal;
            tu[2 * I + 1] = (double)uVal;
            I++;
        }
    }
    if (I == 0) { free(shapeId); return 0; }
    /* sort by 64 bits (order A), permute tu */
    keyidx* ka = (keyidx*)malloc(sizeof(keyidx) * I);
    for (int r = 0; r < I; ++r) { ka[r].key = intersections[r]; ka[r].idx = r; }
    qsort(ka, I, sizeof(keyidx), cmp_keyidx);
    uint64_t* keyA = (uint64_t*)malloc(sizeof(uint64_t) * I);
    double* tuA = (double*)malloc(sizeof(double) * 2 * I);
    for (int r = 0; r < I; ++r) {
        keyA[r] = ka[r].key;
        tuA[2 * r] = tu[2 * ka[r].idx];
        tuA[2 * r + 1] = tu[2 * ka[r].idx + 1];
    }
    /* pairing in order A */
    uint64_t* outA = (uint64_t*)malloc(sizeof(uint64_t) * I);
    double* utA = (double*)malloc(sizeof(double) * 2 * I);
    for (int idx = 0; idx < I; ++idx) {
        uint64_t inter = keyA[idx];
        int sj = (int)((inter >> 48) & 0xFFFF);
        int si = (int)((inter >> 32) & 0xFFFF);
        int i = (int)((inter >> 16) & 0xFFFF);
        int ni = start[si + 1] - start[si];
        uint64_t sij = ((uint64_t)si << 48) | ((uint64_t)sj << 32);
        double tValD = (float)tuA[2 * idx + 1];
        if (pairExactPos)
            tValD = exact_param(pairExactPos, start, si, i, sj, (int)(inter & 0xFFFF), tValD);
        int lo = 0, hi = I - 1, mid;
        while (hi > lo) {
            mid = (hi + lo) / 2;
            if (keyA[mid] < sij) lo = mid + 1; else hi = mid;
        }
        uint64_t ub = ((uint64_t)si << 48) | ((uint64_t)(sj + 1) << 32);
        int bestDist = -1; double bestU = FLT_MAX; int bestIdx = -1;
        double fallbackU = FLT_MAX; int fallbackIdx = -1;
        int k = lo;
        while (k < I && keyA[k] < ub) {
            int l = (int)(keyA[k] & 0xFFFF);
            double uVal = (float)tuA[2 * k];
            if (pairExactPos && (keyA[k] >> 32) == (sij >> 32))   /* candidate of the proper group */
                uVal = exact_param(pairExactPos, start, si, l, sj, (int)((keyA[k] >> 16) & 0xFFFF), uVal);
            double tVal = tValD;
            int d = (l - i + ni) % ni;
            if (l == i && tVal >= uVal) { k++; continue; }
            if (bestDist == -1 || d < bestDist) {
                bestDist = d; bestU = FLT_MAX; bestIdx = -1; fallbackU = FLT_MAX; fallbackIdx = -1;
            }
            if (d == bestDist) {
                if (d == 0) {
                    if (uVal >= tVal && uVal < bestU) { bestU = uVal; bestIdx = k; }
                    if (uVal < fallbackU) { fallbackU = uVal; fallbackIdx = k; }
                } else {
                    if (uVal < bestU) { bestU = uVal; bestIdx = k; }
                }
            }
            k++;
        }
        int player = (bestIdx != -1) ? bestIdx : fallbackIdx;
        if (player != -1) {   /* partner edges outside the two polygons: no partner (the
                                 reference would read vertices of other polygons) */
            int nj = start[sj + 1] - start[sj];
            if ((int)((keyA[player] >> 16) & 0xFFFF) >= nj || (int)(keyA[player] & 0xFFFF) >= ni) player = -1;
        }
        if (player == -1) {
            outA[idx] = keyA[idx];
            utA[2 * idx] = tuA[2 * idx]; utA[2 * idx + 1] = tuA[2 * idx + 1];
        } else {
            outA[idx] = keyA[player];
            utA[2 * idx] = tuA[2 * player]; utA[2 * idx + 1] = tuA[2 * player + 1];
        }
    }
    /* stable sort by bits [0,48) (order B) */
    const uint64_t m48 = (((uint64_t)1) << 48) - 1;
    for (int r = 0; r < I; ++r) { ka[r].key = keyA[r] & m48; ka[r].idx = r; }
    qsort(ka, I, sizeof(keyidx), cmp_keyidx);
    for (int r = 0; r < I; ++r) {
        int a = ka[r].idx;
        intersections[r] = keyA[a];
        tu[2 * r] = tuA[2 * a]; tu[2 * r + 1] = tuA[2 * a + 1];
        ut[2 * r] = utA[2 * a]; ut[2 * r + 1] = utA[2 * a + 1];
        outersections[r] = outA[a];
    }
    free(ka); free(keyA); free(tuA); free(outA); free(utA); free(shapeId);
    return I;
}

/* h, g12 kernels.cuh:139-159 */
static double hfun(d2 p1, d2 p2, d2 sp) {
    d2 r1 = { p1.x - sp.x, p1.y - sp.y }, r2 = { p2.x - sp.x, p2.y - sp.y };
    r1 = wrap2(r1); r2 = wrap2(r2);
    d2 dd = { r2.x - r1.x, r2.y - r1.y };
    d2 d = wrap2(dd);
    return (r1.x + r2.x) * d.y;
}
static void g12(d2 p1, d2 p2, d2 sp, d2* g1, d2* g2) {
    d2 r1 = { p1.x - sp.x, p1.y - sp.y }, r2 = { p2.x - sp.x, p2.y - sp.y };
    r1 = wrap2(r1); r2 = wrap2(r2);
    d2 d = { p2.x - p1.x, p2.y - p1.y };
    d = wrap2(d);
    g1->x = d.y; g1->y = -(r1.x + r2.x);
    g2->x = d.y; g2->y = (r1.x + r2.x);
}
/* getDf kernels.cuh:161-222 (statement for statement) */
static void getDf(d2 vi, d2 vzi, d2 vj, d2 vzj, double* df) {
    d2 t;
    t.x = vzj.x - vj.x; t.y = vzj.y - vj.y; d2 dj = wrap2(t);
    t.x = vzi.x - vi.x; t.y = vzi.y - vi.y; d2 di = wrap2(t);
    t.x = vj.x - vi.x; t.y = vj.y - vi.y; d2 dij = wrap2(t);
    double w = dj.x * di.y - dj.y * di.x;
    double diNorm2 = di.x * di.x + di.y * di.y;
    double djNorm2 = dj.x * dj.x + dj.y * dj.y;
    if (w * w < 1e-12 * diNorm2 * djNorm2 || diNorm2 < 1e-28 || djNorm2 < 1e-28) {
        for (int i = 0; i < 16; i++) df[i] = 0.0;
        return;
    }
    double k = dj.x * dij.y - dj.y * dij.x;
    double u = k / w;
    double dk[8] = { 0 };
    dk[0] = dj.y; dk[1] = -dj.x;
    dk[4] = -dij.y - dj.y; dk[5] = dj.x + dij.x;
    dk[6] = dij.y; dk[7] = -dij.x;
    double dw[8];
    dw[0] = dj.y; dw[1] = -dj.x; dw[2] = -dj.y; dw[3] = dj.x;
    dw[4] = -di.y; dw[5] = di.x; dw[6] = di.y; dw[7] = -di.x;
    for (int i = 0; i < 16; ++i) df[i] = 0.0;
    for (int alpha = 0; alpha < 2; ++alpha) {
        for (int beta = 0; beta < 2; ++beta) {
            for (int p = 0; p < 4; ++p) {
                double du = dk[p * 2 + beta] / w - u * dw[p * 2 + beta] / w;
                int col = 2 * p + beta;
                df[col * 2 + alpha] += (alpha == 0 ? di.x : di.y) * du;
            }
            if (alpha == beta) {
                df[beta * 2 + alpha] += 1.0 - u;
                df[(2 + beta) * 2 + alpha] += u;
            }
        }
    }
}
/* parameter along edge (vi -> vzi) of its crossing with the line (vj -> vzj): the u of
 * getDf kernels.cuh:163-175; fallback for (near) parallel edges */
static double cross_param(d2 vi, d2 vzi, d2 vj, d2 vzj, double fallback) {
    d2 t;
    t.x = vzj.x - vj.x; t.y = vzj.y - vj.y; d2 dj = wrap2(t);
    t.x = vzi.x - vi.x; t.y = vzi.y - vi.y; d2 di = wrap2(t);
    t.x = vj.x - vi.x; t.y = vj.y - vi.y; d2 dij = wrap2(t);
    double w = dj.x * di.y - dj.y * di.x;
    double diNorm2 = di.x * di.x + di.y * di.y;
    double djNorm2 = dj.x * dj.x + dj.y * dj.y;
    if (w * w < 1e-12 * diNorm2 * djNorm2 || diNorm2 < 1e-28 || djNorm2 < 1e-28) return fallback;
    double k = dj.x * dij.y - dj.y * dij.x;
    return k / w;
}
/* negated g . df[cols c0, c0+1] */
static d2 contrib(const double* df, int c0, d2 g) {
    d2 inc;
    inc.x = g.x * df[c0 * 2 + 0] + g.y * df[c0 * 2 + 1];
    inc.y = g.x * df[(c0 + 1) * 2 + 0] + g.y * df[(c0 + 1) * 2 + 1];
    return inc;
}

/* Exterior record contributions: updateForceEnergyExteriorKernel kernels.cuh:1067-1256.
 * Slots: 0:i 1:zi 2:l 3:zl (owned by si), 4:j 5:zj 6:k 7:zk (owned by sj).
 * c[slot] is the value the reference atomically adds (-inc). */
static double exterior_record(uint64_t inter, uint64_t outer, const double* tuR, const double* utR,
                              const double* pos, const int* start, const int* next,
                              int tgt[8], d2 c[8], int areaMode, d2* spOut) {
    int s1 = (int)((inter >> 32) & 0xFFFF), s2 = (int)((inter >> 48) & 0xFFFF);
    int iLoc = (int)((inter >> 16) & 0xFFFF), jLoc = (int)(inter & 0xFFFF);
    int kLoc = (int)((outer >> 16) & 0xFFFF), lLoc = (int)(outer & 0xFFFF);
    int i = start[s1] + iLoc, j = start[s2] + jLoc, k = start[s2] + kLoc, l = start[s1] + lLoc;
    d2 pi = P(pos, i), pj = P(pos, j), pk = P(pos, k), pl = P(pos, l);
    int zi = next[i], zj = next[j], zk = next[k], zl = next[l];
    d2 pzi = P(pos, zi), pzj = P(pos, zj), pzk = P(pos, zk), pzl = P(pos, zl);
    d2 t;
    t.x = pzi.x - pi.x; t.y = pzi.y - pi.y; d2 r1 = wrap2(t);
    t.x = pzk.x - pk.x; t.y = pzk.y - pk.y; d2 r2 = wrap2(t);
    double t1 = tuR[1], t2 = utR[1];
    if (areaMode) {   /* exact crossing parameters at the current positions */
        t1 = cross_param(pi, pzi, pj, pzj, t1);
        t2 = cross_param(pk, pzk, pl, pzl, t2);
    }
    t.x = pi.x + t1 * r1.x; t.y = pi.y + t1 * r1.y; d2 fij = wrap2(t);
    t.x = pk.x + t2 * r2.x; t.y = pk.y + t2 * r2.y; d2 fkl = wrap2(t);
    d2 sp = fij;
    if (areaMode == 1) {   /* closed loop: one origin per polygon pair (first vertex of the
                                 lower-numbered polygon) */
        int so = s1 < s2 ? s1 : s2;
        sp = P(pos, start[so]);
    }
    *spOut = sp;
    double e = 0.0;
    double dfa[16], dfb[16];
    tgt[0] = i; tgt[1] = zi; tgt[2] = l; tgt[3] = zl; tgt[4] = j; tgt[5] = zj; tgt[6] = k; tgt[7] = zk;
    d2 inc[8];
    if (i == l) {
        e += hfun(fij, fkl, sp);
        d2 ga, gb;
        g12(fij, fkl, sp, &ga, &gb);
        getDf(pi, pzi, pj, pzj, dfa);
        getDf(pk, pzk, pi, pzi, dfb);
        inc[0] = contrib(dfa, 0, ga);
        inc[1] = contrib(dfa, 2, ga);
        inc[4] = contrib(dfa, 4, ga);
        inc[5] = contrib(dfa, 6, ga);
        inc[6] = contrib(dfb, 0, gb);
        inc[7] = contrib(dfb, 2, gb);
        inc[2] = contrib(dfb, 4, gb);
        inc[3] = contrib(dfb, 6, gb);
    } else {
        e += hfun(fij, pzi, sp);
        e += hfun(pl, fkl, sp);
        d2 g1a, g2a;
        g12(fij, pzi, sp, &g1a, &g2a);
        getDf(pi, pzi, pj, pzj, dfa);
        getDf(pk, pzk, pl, pzl, dfb);
        inc[0] = contrib(dfa, 0, g1a);
        inc[4] = contrib(dfa, 4, g1a);
        {
            d2 dd = contrib(dfa, 2, g1a);
            inc[1].x = dd.x + g2a.x;
            inc[1].y = dd.y + g2a.y;
        }
        inc[5] = contrib(dfa, 6, g1a);
        d2 g1b, g2b;
        g12(pl, fkl, sp, &g1b, &g2b);
        {
            d2 dd = contrib(dfb, 4, g2b);
            inc[2].x = g1b.x; inc[2].y = g1b.y;
            inc[2].x += dd.x; inc[2].y += dd.y;
        }
        inc[6] = contrib(dfb, 0, g2b);
        inc[3] = contrib(dfb, 6, g2b);
        inc[7] = contrib(dfb, 2, g2b);
    }
    double sc = areaMode ? 0.5 : 1.0;   /* areaMode 2: broken loop, own entry point */
    for (int s = 0; s < 8; ++s) { c[s].x = -inc[s].x * sc; c[s].y = -inc[s].y * sc; }
    return e;
}

/* Pairwise 256-wide tree sum (fixed order), recursive over partials. */
static double tree256(const double* x, int n) {
    double v[256];
    for (int t = 0; t < 256; ++t) v[t] = (t < n) ? x[t] : 0.0;
    for (int s = 128; s >= 1; s >>= 1)
        for (int t = 0; t < s; ++t) v[t] += v[t + s];
    return v[0];
}
double orc_det_sum(const double* x, int n) {
    if (n <= 0) return 0.0;
    if (n <= 256) return tree256(x, n);
    int m = (n + 255) / 256;
    double* p = (double*)malloc(sizeof(double) * m);
    for (int b = 0; b < m; ++b) {
        int len = n - 256 * b; if (len > 256) len = 256;
        p[b] = tree256(x + 256 * b, len);
    }
    double r = orc_det_sum(p, m);
    free(p);
    return r;
}
double orc_max_abs(const double* x, int n) {
    double m = 0.0;
    for (int i = 0; i < n; ++i) { double a = fabs(x[i]); if (a > m) m = a; }
    return m;
}

/* Model::updateForceEnergy model.cpp:363-395 with kernels
 *   exterior kernels.cuh:1067-1273, updateShapeRangesKernel :1275-1282,
 *   interior :1354-1424, edge :1284-1335, area :1337-1352.
 * targetEdgeV: per-vertex target edge length (reference: targetEdgeLengths[shapeId[m]]).
 * Returns the energy; force[2V] is overwritten. */
double orc_force_energy(int model, int V, int N, const int* start, const double* pos,
                        int I, const uint64_t* intersections, const uint64_t* outersections,
                        const double* tu, const double* ut, const double* edgeLengths,
                        const double* areas, const double* targetEdgeV, const double* targetAreas,
                        double stiffness, double compressibility, double* force) {
    return orc_force_energy_ex(model, V, N, start, pos, I, intersections, outersections, tu, ut,
                               edgeLengths, areas, targetEdgeV, targetAreas, stiffness, compressibility,
                               0.0, force, NULL);
}

/* perimeter of polygon s and its target (sum of the edge targets), both summed in
 * vertex order (no counterpart in the reference: the perimeter spring below) */
static void perimeter_and_target(int s, const int* start, const double* edgeLengths,
                                 const double* targetEdgeV, double* P, double* P0) {
    double a = 0.0, b = 0.0;
    for (int m = start[s]; m < start[s + 1]; ++m) { a += edgeLengths[m]; b += targetEdgeV[m]; }
    *P = a; *P0 = b;
}

/* As orc_force_energy, plus (no counterpart in the reference) a perimeter spring
 * E = kP (P - P0)^2 / 2 per polygon of the softBody and hybrid models (P0 = sum of the
 * polygon's edge targets; force -kP (P - P0) grad P, grad_m P = t_{m-1} - t_m with t the
 * unit edge vectors), added after the area spring; and fext (optional, [2V]): each
 * vertex's force before any spring term (the overlap force; zero without overlap). */
double orc_force_energy_ex(int model, int V, int N, const int* start, const double* pos,
                           int I, const uint64_t* intersections, const uint64_t* outersections,
                           const double* tu, const double* ut, const double* edgeLengths,
                           const double* areas, const double* targetEdgeV, const double* targetAreas,
                           double stiffness, double compressibility, double perimStiffness,
                           double* force, double* fext) {
    int* shapeId = (int*)malloc(sizeof(int) * (V > 0 ? V : 1));
    int* next = (int*)malloc(sizeof(int) * (V > 0 ? V : 1));
    int* prev = (int*)malloc(sizeof(int) * (V > 0 ? V : 1));
    orc_topology(N, start, V, shapeId, next, prev);
    int pairMode = (model & ORC_PAIR_SQUARED) != 0;
    int sqMode = (model & (ORC_AREA_SQUARED | ORC_PAIR_SQUARED)) != 0;
    int areaMode = (model & ORC_AREA_MODES) != 0;
    model &= 0xFF;
    int overlap = (model == ORC_NORMAL || model == ORC_HYBRID);
    int edgeOn = (model == ORC_SOFT_BODY || model == ORC_EDGE_ONLY || model == ORC_HYBRID);
    int areaOn = (model == ORC_SOFT_BODY || model == ORC_AREA_ONLY || model == ORC_HYBRID);
    int perimOn = (model == ORC_SOFT_BODY || model == ORC_HYBRID) && perimStiffness != 0.0;
    if (!overlap) I = 0;
    int Ialloc = I > 0 ? I : 1;
    int* tgt = (int*)malloc(sizeof(int) * 8 * Ialloc);
    d2* cc = (d2*)malloc(sizeof(d2) * 8 * Ialloc);
    double* eext = (double*)malloc(sizeof(double) * Ialloc);
    unsigned char* self = (unsigned char*)malloc(Ialloc);
    d2* spr = (d2*)malloc(sizeof(d2) * Ialloc);
    int* shapeStart = (int*)malloc(sizeof(int) * (N > 0 ? N : 1));
    int* shapeEnd = (int*)malloc(sizeof(int) * (N > 0 ? N : 1));
    for (int s = 0; s < N; ++s) { shapeStart[s] = I; shapeEnd[s] = -1; }
    /* area energy: records on a closed partner cycle (<= 8 hops) share the pair origin */
    int* loopMode = (int*)malloc(sizeof(int) * Ialloc);
    int* partner = (int*)malloc(sizeof(int) * Ialloc);
    for (int r = 0; r < I; ++r) loopMode[r] = areaMode ? 2 : 0;
    if (areaMode && I > 0) {
        for (int r = 0; r < I; ++r) {
            partner[r] = -1;
            if (intersections[r] == outersections[r]) continue;
            int lo = 0, hi = I;   /* records are sorted by (si,i,j,sj): binary search */
            uint64_t o = outersections[r];
            uint64_t ko = ((o & 0xFFFFFFFFull) << 16) | (o >> 48);
            uint64_t kp = (o >> 32) & 0xFFFF;
            while (lo < hi) {
                int mid = (lo + hi) / 2;
                uint64_t km = intersections[mid];
                uint64_t sm = (km >> 32) & 0xFFFF;
                uint64_t om = ((km & 0xFFFFFFFFull) << 16) | (km >> 48);
                if (sm < kp || (sm == kp && om < ko)) lo = mid + 1; else hi = mid;
            }
            if (lo < I && intersections[lo] == o) partner[r] = lo;
        }
        for (int r = 0; r < I; ++r) {
            int q = partner[r];
            for (int hop = 0; hop < 8 && q >= 0 && q != r; ++hop) q = partner[q];
            if (q == r) loopMode[r] = 1;
        }
    }
#pragma omp parallel for schedule(static)
    for (int r = 0; r < I; ++r) {
        self[r] = (intersections[r] == outersections[r]);
        if (self[r]) { eext[r] = 0.0; continue; }
        eext[r] = exterior_record(intersections[r], outersections[r], tu + 2 * r, ut + 2 * r,
                                  pos, start, next, tgt + 8 * r, cc + 8 * r, loopMode[r], spr + r);
    }
    for (int r = 0; r < I; ++r) {
        int s = (int)((intersections[r] >> 32) & 0xFFFF);
        if (r < shapeStart[s]) shapeStart[s] = r;
        if (r > shapeEnd[s]) shapeEnd[s] = r;
    }
    /* squared lens areas: per record E_r = exterior + its chain (h sums, twice the area it
     * covers); a lens is a closed partner cycle, A = 0.5 * sum of E_r over the cycle starting
     * at its smallest key; a record off a cycle is its own lens.  Contributions scale by A. */
    double* lensA = (double*)malloc(sizeof(double) * Ialloc);
    int* lensN = (int*)malloc(sizeof(int) * Ialloc);
    /* with ORC_PAIR_SQUARED: the scale of a record's contributions is the overlap area of
     * its polygon pair, pairA = sum of the pair's lens areas (one representative record
     * per lens: the smallest key of its cycle, or a record off any cycle; every other
     * record adds 0), added over the lower polygon's records with sj = the higher one,
     * then the higher polygon's records with sj = the lower one, each in record order */
    double* scaleR = (double*)malloc(sizeof(double) * Ialloc);
    unsigned char* isRep = (unsigned char*)calloc(Ialloc, 1);
    if (sqMode && I > 0) {
        double* Er = (double*)malloc(sizeof(double) * I);
        for (int r = 0; r < I; ++r) {
            Er[r] = 0.0;
            if (self[r]) continue;
            uint64_t inter = intersections[r], outer = outersections[r];
            int si = (int)((inter >> 32) & 0xFFFF), n = start[si + 1] - start[si];
            int iLoc = (int)((inter >> 16) & 0xFFFF), lLoc = (int)(outer & 0xFFFF);
            int lDist = (lLoc - iLoc + n) % n;
            double chain = 0.0;
            for (int d = 1; d < lDist; ++d) {
                int m = start[si] + (iLoc + d) % n;
                chain += hfun(P(pos, m), P(pos, next[m]), spr[r]);
            }
            Er[r] = eext[r] + chain;
        }
        for (int r = 0; r < I; ++r) {
            lensA[r] = 0.0; lensN[r] = 1;
            if (self[r]) continue;
            if (loopMode[r] != 1) { lensA[r] = 0.5 * Er[r]; isRep[r] = 1; continue; }
            int mem[16], cnt = 0, q = r;
            do { mem[cnt++] = q; q = partner[q]; } while (q != r && cnt < 16);
            int k0 = 0;
            for (int k = 1; k < cnt; ++k) if (intersections[mem[k]] < intersections[mem[k0]]) k0 = k;
            double sum = 0.0;
            for (int k = 0; k < cnt; ++k) sum += Er[mem[(k0 + k) % cnt]];
            lensA[r] = 0.5 * sum;
            lensN[r] = cnt;
            isRep[r] = (mem[k0] == r);
        }
        for (int r = 0; r < I; ++r) {
            scaleR[r] = lensA[r];
            if (!pairMode || self[r]) continue;
            int si = (int)((intersections[r] >> 32) & 0xFFFF), sj = (int)((intersections[r] >> 48) & 0xFFFF);
            int a = si < sj ? si : sj, b = si < sj ? sj : si;
            double pa = 0.0;
            for (int pass = 0; pass < (a == b ? 1 : 2); ++pass) {
                int own = pass ? b : a, other = pass ? a : b;
                if (shapeStart[own] > shapeEnd[own]) continue;
                for (int q = shapeStart[own]; q <= shapeEnd[own]; ++q)
                    if ((int)((intersections[q] >> 48) & 0xFFFF) == other) pa += isRep[q] ? lensA[q] : 0.0;
            }
            scaleR[r] = pa;
        }
        for (int r = 0; r < I; ++r)
            for (int q = 0; q < 8; ++q) { cc[8 * r + q].x *= scaleR[r]; cc[8 * r + q].y *= scaleR[r]; }
        free(Er);
    }
    /* records owned through sj, in record order */
    int* sjCount = (int*)calloc(N + 1, sizeof(int));
    for (int r = 0; r < I; ++r) sjCount[((intersections[r] >> 48) & 0xFFFF) + 1]++;
    for (int s = 0; s < N; ++s) sjCount[s + 1] += sjCount[s];
    int* sjList = (int*)malloc(sizeof(int) * Ialloc);
    int* fill = (int*)malloc(sizeof(int) * (N > 0 ? N : 1));
    for (int s = 0; s < N; ++s) fill[s] = sjCount[s];
    for (int r = 0; r < I; ++r) sjList[fill[(intersections[r] >> 48) & 0xFFFF]++] = r;
    /* interior: thread-local sums (kernels.cuh:1354-1411) */
    d2* LM = (d2*)malloc(sizeof(d2) * (V > 0 ? V : 1));
    d2* LN = (d2*)malloc(sizeof(d2) * (V > 0 ? V : 1));
    double* eint = (double*)malloc(sizeof(double) * (V > 0 ? V : 1));
#pragma omp parallel for schedule(static)
    for (int m = 0; m < V; ++m) {
        d2 lm = { 0.0, 0.0 }, ln = { 0.0, 0.0 };
        double le = 0.0;
        int s = shapeId[m];
        if (overlap && shapeStart[s] <= shapeEnd[s]) {
            int n = start[s + 1] - start[s];
            int mLoc = m - start[s];
            d2 pm = P(pos, m);
            int nxt = next[m];
            d2 pn = P(pos, nxt);
            for (int k = shapeStart[s]; k <= shapeEnd[s]; ++k) {
                uint64_t inter = intersections[k], outer = outersections[k];
                if (inter == outer) continue;
                int si = (int)((inter >> 32) & 0xFFFF);
                int iLoc = (int)((inter >> 16) & 0xFFFF);
                int lLoc = (int)(outer & 0xFFFF);
                int iDist = (mLoc - iLoc + n) % n;
                int lDist = (lLoc - iLoc + n) % n;
                if (iDist > 0 && iDist < lDist) {
                    int ik = start[si] + iLoc;
                    d2 pik = P(pos, ik), pzik = P(pos, next[ik]);
                    d2 t = { pzik.x - pik.x, pzik.y - pik.y };
                    d2 rik = wrap2(t);
                    double t1k = tu[2 * k + 1];
                    t.x = pik.x + t1k * rik.x; t.y = pik.y + t1k * rik.y;
                    d2 sp = wrap2(t);
                    if (areaMode) sp = spr[k];
                    le += hfun(pm, pn, sp);
                    d2 g1, g2;
                    g12(pm, pn, sp, &g1, &g2);
                    if (areaMode) { g1.x *= 0.5; g1.y *= 0.5; g2.x *= 0.5; g2.y *= 0.5; }
                    if (sqMode) { g1.x *= scaleR[k]; g1.y *= scaleR[k]; g2.x *= scaleR[k]; g2.y *= scaleR[k]; }
                    lm.x -= g1.x; lm.y -= g1.y;
                    ln.x -= g2.x; ln.y -= g2.y;
                }
            }
        }
        LM[m] = lm; LN[m] = ln; eint[m] = le;
    }
    /* canonical per-vertex accumulation */
    double* eedge = (double*)calloc(V > 0 ? V : 1, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int m = 0; m < V; ++m) {
        int s = shapeId[m];
        d2 F = { 0.0, 0.0 };
        if (overlap) {
            if (shapeStart[s] <= shapeEnd[s])
                for (int r = shapeStart[s]; r <= shapeEnd[s]; ++r) {
                    if (self[r]) continue;
                    for (int q = 0; q < 4; ++q)
                        if (tgt[8 * r + q] == m) { F.x += cc[8 * r + q].x; F.y += cc[8 * r + q].y; }
                }
            for (int e = sjCount[s]; e < sjCount[s + 1]; ++e) {
                int r = sjList[e];
                if (self[r]) continue;
                for (int q = 4; q < 8; ++q)
                    if (tgt[8 * r + q] == m) { F.x += cc[8 * r + q].x; F.y += cc[8 * r + q].y; }
            }
            if (LM[m].x != 0.0 || LM[m].y != 0.0) { F.x += LM[m].x; F.y += LM[m].y; }
            int pm = prev[m];
            if (LN[pm].x != 0.0 || LN[pm].y != 0.0) { F.x += LN[pm].x; F.y += LN[pm].y; }
        }
        if (fext) { fext[2 * m] = F.x; fext[2 * m + 1] = F.y; }
        if (edgeOn) {
            int prv = prev[m], nxt = next[m];
            double l0 = targetEdgeV[m], l0prv = targetEdgeV[prv];
            double l = edgeLengths[m], prvl = edgeLengths[prv];
            d2 a = { pos[2 * m] - pos[2 * nxt], pos[2 * m + 1] - pos[2 * nxt + 1] };
            d2 b = { pos[2 * prv] - pos[2 * m], pos[2 * prv + 1] - pos[2 * m + 1] };
            a.x += 1.5; a.y += 1.5; b.x += 1.5; b.y += 1.5;
            a.x = shift_into_unit(a.x); a.y = shift_into_unit(a.y);
            b.x = shift_into_unit(b.x); b.y = shift_into_unit(b.y);
            a.x -= 0.5; a.y -= 0.5; b.x -= 0.5; b.y -= 0.5;
            double c1 = (l > 1e-14) ? (1.0 - l0 / l) : 0.0;
            double c2 = (prvl > 1e-14) ? (1.0 - l0prv / prvl) : 0.0;
            eedge[m] = stiffness * (l - l0) * (l - l0) / 2.0;
            d2 lf;
            lf.x = -(c1 * a.x - c2 * b.x) * stiffness;
            lf.y = -(c1 * a.y - c2 * b.y) * stiffness;
            if (lf.x != 0.0 || lf.y != 0.0) { F.x += lf.x; F.y += lf.y; }
        }
        if (areaOn) {
            double A = areas[s], A0 = targetAreas[s];
            if (!(A0 < 1e-14)) {
                double coeff = 0.5 * compressibility * (A - A0);
                double fx = -coeff * orc_wrap(pos[2 * next[m] + 1] - pos[2 * prev[m] + 1]);
                double fy = -coeff * orc_wrap(pos[2 * prev[m]] - pos[2 * next[m]]);
                F.x += fx; F.y += fy;
            }
        }
        if (perimOn) {
            int prv = prev[m], nxt = next[m];
            double P, P0;
            perimeter_and_target(s, start, edgeLengths, targetEdgeV, &P, &P0);
            double l = edgeLengths[m], lp = edgeLengths[prv];
            double dx = orc_wrap(pos[2 * nxt] - pos[2 * m]), dy = orc_wrap(pos[2 * nxt + 1] - pos[2 * m + 1]);
            double px = orc_wrap(pos[2 * m] - pos[2 * prv]), py = orc_wrap(pos[2 * m + 1] - pos[2 * prv + 1]);
            double tx = (l > 1e-14) ? dx / l : 0.0, ty = (l > 1e-14) ? dy / l : 0.0;
            double qx = (lp > 1e-14) ? px / lp : 0.0, qy = (lp > 1e-14) ? py / lp : 0.0;
            double coeff = perimStiffness * (P - P0);
            F.x += -coeff * (qx - tx);
            F.y += -coeff * (qy - ty);
        }
        force[2 * m] = F.x; force[2 * m + 1] = F.y;
    }
    /* energies per polygon, then tree over polygons */
    double* ep = (double*)malloc(sizeof(double) * (N > 0 ? N : 1));
    for (int s = 0; s < N; ++s) {
        double acc = 0.0;
        double e;
        if (sqMode) {   /* each lens's A^2/2 split evenly over its records */
            if (overlap && shapeStart[s] <= shapeEnd[s])
                for (int r = shapeStart[s]; r <= shapeEnd[s]; ++r)
                    acc += (scaleR[r] * lensA[r]) / (2.0 * lensN[r]);
            e = acc;
        } else {
            if (overlap) {
                if (shapeStart[s] <= shapeEnd[s])
                    for (int r = shapeStart[s]; r <= shapeEnd[s]; ++r) acc += eext[r];
                for (int m = start[s]; m < start[s + 1]; ++m) acc += eint[m];
            }
            e = 0.5 * acc;
        }
        if (edgeOn) for (int m = start[s]; m < start[s + 1]; ++m) e += eedge[m];
        if (areaOn) {
            double A = areas[s], A0 = targetAreas[s];
            if (!(A0 < 1e-14)) e += compressibility * (A - A0) * (A - A0) / 2.0;
        }
        if (perimOn) {
            double P, P0;
            perimeter_and_target(s, start, edgeLengths, targetEdgeV, &P, &P0);
            e += perimStiffness * (P - P0) * (P - P0) / 2.0;
        }
        ep[s] = e;
    }
    double E = orc_det_sum(ep, N);
    free(ep); free(eedge); free(LM); free(LN); free(eint); free(sjCount); free(sjList); free(fill);
    free(shapeStart); free(shapeEnd); free(tgt); free(cc); free(eext); free(self); free(spr);
    free(loopMode); free(partner); free(lensA); free(lensN); free(scaleR); free(isRep);
    free(shapeId); free(next); free(prev);
    return E;
}

/* ---------- force projection: getConstrainedForcePython pyCudaPolygon.py:441-487 ----------
 * Per polygon: Householder QR of M = [edge-length gradients | area gradient]
 * (2n x (n+1)), f <- f - Q Q^T f.  Equal (up to rounding) to the reference's
 * projectForceCUDA model.cu:165-206 (SVD basis of the edge gradients + Gram-Schmidt
 * of the normalized area gradient, kernels.cuh:389-523), which is basis-independent. */
void orc_project_force(int V, int N, const int* start, const double* pos, double* force) {
    (void)V;
#pragma omp parallel for schedule(dynamic, 16)
    for (int s = 0; s < N; ++s) {
        int st = start[s], n = start[s + 1] - st;
        int rows = 2 * n, cols = n + 1;
        double* M = (double*)calloc((size_t)rows * cols, sizeof(double));
        for (int e = 0; e < n; ++e) {
            int a = st + e, b = st + (e + 1) % n;
            double dx = orc_wrap(pos[2 * b] - pos[2 * a]);
            double dy = orc_wrap(pos[2 * b + 1] - pos[2 * a + 1]);
            double len = sqrt(dx * dx + dy * dy);
            if (len < 1e-14) continue;
            double tx = dx / len, ty = dy / len;
            M[(2 * e) * cols + e] = -tx;
            M[(2 * e + 1) * cols + e] = -ty;
            M[(2 * ((e + 1) % n)) * cols + e] = tx;
            M[(2 * ((e + 1) % n) + 1) * cols + e] = ty;
        }
        for (int k = 0; k < n; ++k) {
            int nx = st + (k + 1) % n, pv = st + (k + n - 1) % n;
            M[(2 * k) * cols + n] = orc_wrap(pos[2 * nx + 1] - pos[2 * pv + 1]) / 2;
            M[(2 * k + 1) * cols + n] = -orc_wrap(pos[2 * nx] - pos[2 * pv]) / 2;
        }
        /* Householder: store reflectors, apply Q^T to f, zero the first cols entries, apply Q */
        double* f = (double*)malloc(sizeof(double) * rows);
        for (int r = 0; r < rows; ++r) f[r] = force[2 * st + r];
        double* vk = (double*)malloc(sizeof(double) * rows * cols);
        double* beta = (double*)malloc(sizeof(double) * cols);
        for (int k = 0; k < cols; ++k) {
            double nrm = 0.0;
            for (int r = k; r < rows; ++r) nrm += M[r * cols + k] * M[r * cols + k];
            nrm = sqrt(nrm);
            double* v = vk + (size_t)k * rows;
            for (int r = 0; r < rows; ++r) v[r] = 0.0;
            if (nrm == 0.0) { beta[k] = 0.0; continue; }
            double alpha = (M[k * cols + k] > 0) ? -nrm : nrm;
            for (int r = k; r < rows; ++r) v[r] = M[r * cols + k];
            v[k] -= alpha;
            double vv = 0.0;
            for (int r = k; r < rows; ++r) vv += v[r] * v[r];
            beta[k] = (vv > 0.0) ? 2.0 / vv : 0.0;
            for (int c = k; c < cols; ++c) {
                double d = 0.0;
                for (int r = k; r < rows; ++r) d += v[r] * M[r * cols + c];
                d *= beta[k];
                for (int r = k; r < rows; ++r) M[r * cols + c] -= d * v[r];
            }
        }
        /* g = Q^T f */
        for (int k = 0; k < cols; ++k) {
            double* v = vk + (size_t)k * rows;
            double d = 0.0;
            for (int r = k; r < rows; ++r) d += v[r] * f[r];
            d *= beta[k];
            for (int r = k; r < rows; ++r) f[r] -= d * v[r];
        }
        for (int k = 0; k < cols; ++k) f[k] = 0.0;   /* drop the span components */
        for (int k = cols - 1; k >= 0; --k) {       /* back to the original basis */
            double* v = vk + (size_t)k * rows;
            double d = 0.0;
            for (int r = k; r < rows; ++r) d += v[r] * f[r];
            d *= beta[k];
            for (int r = k; r < rows; ++r) f[r] -= d * v[r];
        }
        for (int r = 0; r < rows; ++r) force[2 * st + r] = f[r];
        free(M); free(f); free(vk); free(beta);
    }
}

/* ---------- SHAKE (shake.h is absent: src/model.cu:16) ----------
 * Contract from shakeProjectCUDA model.cu:619-636 and Model::minimizeFIREStep
 * model.cpp:418-438: per polygon, project positions onto {|edge e| = L_e for all e,
 * area = A0} (n+1 constraints), Newton iterations with the (n+1)x(n+1) system
 * J J^T lambda = C, x -= J^T lambda, until max|C| < tol; returns the largest
 * iteration count (maxIter when some polygon did not converge).
 * This oracle solves the system densely (Gaussian elimination, partial pivoting),
 * independently of the structured solver in the HIP kernel. */
static double poly_area(const double* pos, int st, int n) {   /* kernels.cuh:302-332 */
    double area = 0.0;
    for (int k = 0; k < n; ++k) {
        int a = st + k, b = st + (k + 1) % n;
        double x1 = orc_wrap(pos[2 * a] - pos[2 * st]);
        double dy2 = orc_wrap(pos[2 * b + 1] - pos[2 * a + 1]);
        double x2 = orc_wrap(pos[2 * b] - pos[2 * st]);
        area += dy2 * (x2 + x1) / 2.0;
    }
    return area;
}
static int solve_dense(double* A, double* b, int n) {
    for (int c = 0; c < n; ++c) {
        int p = c; double best = fabs(A[c * n + c]);
        for (int r = c + 1; r < n; ++r) if (fabs(A[r * n + c]) > best) { best = fabs(A[r * n + c]); p = r; }
        if (best == 0.0) return -1;
        if (p != c) {
            for (int k = 0; k < n; ++k) { double t = A[c * n + k]; A[c * n + k] = A[p * n + k]; A[p * n + k] = t; }
            double t = b[c]; b[c] = b[p]; b[p] = t;
        }
        for (int r = c + 1; r < n; ++r) {
            double f = A[r * n + c] / A[c * n + c];
            for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        double s = b[c];
        for (int k = c + 1; k < n; ++k) s -= A[c * n + k] * b[k];
        b[c] = s / A[c * n + c];
    }
    return 0;
}
int orc_shake(int V, int N, const int* start, double* pos, const double* targetEdgeV,
              const double* targetAreas, int maxIter, double tol) {
    (void)V;
    if (N == 0 || maxIter == 0) return 0;
    int worst = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(max : worst)
    for (int s = 0; s < N; ++s) {
        int st = start[s], n = start[s + 1] - st, nc = n + 1, rows = 2 * n;
        double* J = (double*)malloc(sizeof(double) * nc * rows);
        double* A = (double*)malloc(sizeof(double) * nc * nc);
        double* C = (double*)malloc(sizeof(double) * nc);
        int it = 0;
        for (; it < maxIter; ++it) {
            for (int q = 0; q < nc * rows; ++q) J[q] = 0.0;
            double cmax = 0.0;
            for (int e = 0; e < n; ++e) {
                int a = st + e, b = st + (e + 1) % n;
                double dx = orc_wrap(pos[2 * b] - pos[2 * a]);
                double dy = orc_wrap(pos[2 * b + 1] - pos[2 * a + 1]);
                double len = sqrt(dx * dx + dy * dy);
                C[e] = len - targetEdgeV[a];
                double tx = len > 0 ? dx / len : 0.0, ty = len > 0 ? dy / len : 0.0;
                J[e * rows + 2 * e] -= tx; J[e * rows + 2 * e + 1] -= ty;
                J[e * rows + 2 * ((e + 1) % n)] += tx; J[e * rows + 2 * ((e + 1) % n) + 1] += ty;
                if (fabs(C[e]) > cmax) cmax = fabs(C[e]);
            }
            C[n] = poly_area(pos, st, n) - targetAreas[s];
            if (fabs(C[n]) > cmax) cmax = fabs(C[n]);
            for (int k = 0; k < n; ++k) {
                int nx = st + (k + 1) % n, pv = st + (k + n - 1) % n;
                J[n * rows + 2 * k] = orc_wrap(pos[2 * nx + 1] - pos[2 * pv + 1]) / 2;
                J[n * rows + 2 * k + 1] = orc_wrap(pos[2 * pv] - pos[2 * nx]) / 2;
            }
            if (cmax < tol) break;
            for (int a = 0; a < nc; ++a)
                for (int b = 0; b < nc; ++b) {
                    double d = 0.0;
                    for (int r = 0; r < rows; ++r) d += J[a * rows + r] * J[b * rows + r];
                    A[a * nc + b] = d;
                }
            if (solve_dense(A, C, nc) != 0) { it = maxIter; break; }
            for (int r = 0; r < rows; ++r) {
                double d = 0.0;
                for (int a = 0; a < nc; ++a) d += J[a * rows + r] * C[a];
                double p = pos[2 * st + r] - d;
                pos[2 * st + r] = p - floor(p);
            }
        }
        if (it > worst) worst = it;
        free(J); free(A); free(C);
    }
    return worst;
}

/* ---------- the same projection and SHAKE in the HIP kernels' lane order ----------
 * projectForceCUDA model.cu:165-206 and the SHAKE contract above, solved the way the
 * library's segment kernels solve them (one lane e per vertex/edge of the polygon):
 * the normal equations [[G, b], [b^T, c]] [y; z] = [r; s] of the n edge-length
 * constraints and the area constraint, G cyclic tridiagonal.  The leading (n-1)x(n-1)
 * block is solved by parallel cyclic reduction for three right-hand sides at once, then
 * the cyclic corner and the area row by their 1x1 Schur complements.  Sums over the
 * polygon are butterfly all-reductions over the next power of two >= n lanes (zero
 * padding).  Every operation is written in the kernels' order, so FIRE with these
 * solves follows the GPU's trajectory bit for bit; the mathematics is pinned by the
 * dense solvers above (orc_project_force: Householder QR; orc_shake: Gaussian
 * elimination), which these must equal to rounding.  Polygons of at most 64 vertices
 * (-1 otherwise); polygons of fewer than 3 are left as they are, as in the kernels. */
#define LANES_MAX 64
static double lanes_sum(const double* v, int n) {
    double a[LANES_MAX], b[LANES_MAX];
    int P = 1;
    while (P < n) P <<= 1;
    for (int e = 0; e < P; ++e) a[e] = (e < n) ? v[e] : 0.0;
    for (int o = P / 2; o >= 1; o >>= 1) {
        for (int e = 0; e < P; ++e) b[e] = a[e] + a[e ^ o];
        for (int e = 0; e < P; ++e) a[e] = b[e];
    }
    return a[0];
}

/* y[e] (edge multipliers) and *zOut (area multiplier) for edges e < n.  t = unit edge
 * tangents, a = area gradient, r = edge right-hand side, c = |a|^2, s = area
 * right-hand side; the area direction is dropped when its Schur complement is
 * <= zThresh * c. */
static void lanes_bordered_solve(int n, const double* tx, const double* ty, const double* ax,
                                 const double* ay, const double* r, double c, double s,
                                 double zThresh, double* y, double* zOut) {
    double off[LANES_MAX] = {0}, diag[LANES_MAX] = {0}, bv[LANES_MAX] = {0};
    double A[LANES_MAX] = {0}, C[LANES_MAX] = {0}, Bd[LANES_MAX] = {0}, d1[LANES_MAX] = {0},
           d2[LANES_MAX] = {0}, d3[LANES_MAX] = {0};
    double inv[LANES_MAX] = {0}, nA[LANES_MAX] = {0}, nC[LANES_MAX] = {0}, nB[LANES_MAX] = {0},
           n1[LANES_MAX] = {0}, n2[LANES_MAX] = {0}, n3[LANES_MAX] = {0};
    double x1[LANES_MAX] = {0}, x2[LANES_MAX] = {0}, x3[LANES_MAX] = {0}, u[LANES_MAX] = {0},
           tmp[LANES_MAX] = {0};
    int m = n - 1;
    for (int e = 0; e < n; ++e) {
        int en = (e + 1 >= n) ? 0 : e + 1;
        off[e] = -(tx[e] * tx[en] + ty[e] * ty[en]);      /* G_{e,e+1 mod n} */
        double dsq = tx[e] * tx[e] + ty[e] * ty[e];
        diag[e] = (dsq > 0.0) ? 2.0 * dsq : 1.0;
        bv[e] = tx[e] * (ax[en] - ax[e]) + ty[e] * (ay[en] - ay[e]);
    }
    double offM = off[m], offM1 = off[m - 1];
    double rLast = r[m], bLast = bv[m], diagLast = diag[m];
    for (int e = 0; e < n; ++e) {
        int inb = e < m;
        double offPrev = off[(e == 0) ? 0 : e - 1];
        A[e] = (inb && e > 0) ? offPrev : 0.0;
        C[e] = (inb && e < m - 1) ? off[e] : 0.0;
        Bd[e] = inb ? diag[e] : 1.0;
        d1[e] = inb ? r[e] : 0.0;
        d2[e] = inb ? bv[e] : 0.0;
        d3[e] = 0.0;
        if (inb) {
            if (e == 0) d3[e] += offM;
            if (e == m - 1) d3[e] += offM1;
        }
    }
    double w0 = offM + ((m - 1 == 0) ? offM1 : 0.0);
    double wl = (m - 1 == 0) ? 0.0 : offM1;
    for (int st = 1; st < m; st <<= 1) {                  /* one PCR step, all lanes at once */
        for (int e = 0; e < n; ++e) inv[e] = 1.0 / Bd[e];
        for (int e = 0; e < m; ++e) {
            int lo = e - st, hi = e + st;
            int hasLo = lo >= 0, hasHi = hi < m;
            double al = hasLo ? -A[e] * inv[lo] : 0.0;
            double ga = hasHi ? -C[e] * inv[hi] : 0.0;
            nA[e] = hasLo ? al * A[lo] : 0.0;
            nC[e] = hasHi ? ga * C[hi] : 0.0;
            nB[e] = Bd[e]; n1[e] = d1[e]; n2[e] = d2[e]; n3[e] = d3[e];
            if (hasLo) { nB[e] += al * C[lo]; n1[e] += al * d1[lo]; n2[e] += al * d2[lo]; n3[e] += al * d3[lo]; }
            if (hasHi) { nB[e] += ga * A[hi]; n1[e] += ga * d1[hi]; n2[e] += ga * d2[hi]; n3[e] += ga * d3[hi]; }
        }
        for (int e = 0; e < m; ++e) {
            A[e] = nA[e]; C[e] = nC[e]; Bd[e] = nB[e]; d1[e] = n1[e]; d2[e] = n2[e]; d3[e] = n3[e];
        }
    }
    for (int e = 0; e < n; ++e) {
        int inb = e < m;
        x1[e] = inb ? d1[e] / Bd[e] : 0.0;
        x2[e] = inb ? d2[e] / Bd[e] : 0.0;
        x3[e] = inb ? d3[e] / Bd[e] : 0.0;
    }
    double wX1 = w0 * x1[0] + wl * x1[m - 1];
    double wX2 = w0 * x2[0] + wl * x2[m - 1];
    double wX3 = w0 * x3[0] + wl * x3[m - 1];
    double schur = diagLast - wX3;
    double y1l = (rLast - wX1) / schur;
    double ul = (bLast - wX2) / schur;
    for (int e = 0; e < n; ++e) {
        int inb = e < m;
        y[e] = inb ? x1[e] - x3[e] * y1l : ((e == m) ? y1l : 0.0);
        u[e] = inb ? x2[e] - x3[e] * ul : ((e == m) ? ul : 0.0);
    }
    for (int e = 0; e < n; ++e) tmp[e] = bv[e] * u[e];
    double bu = lanes_sum(tmp, n);
    for (int e = 0; e < n; ++e) tmp[e] = bv[e] * y[e];
    double by = lanes_sum(tmp, n);
    double sigma = c - bu;
    double z = (sigma > zThresh * c) ? (s - by) / sigma : 0.0;
    *zOut = z;
    for (int e = 0; e < n; ++e) y[e] = y[e] - z * u[e];
}

/* projection of each polygon's force (getConstrainedForcePython pyCudaPolygon.py:441-487) */
int orc_project_force_lanes(int V, int N, const int* start, const double* pos, double* force) {
    (void)V;
    for (int s = 0; s < N; ++s)
        if (start[s + 1] - start[s] > LANES_MAX) return -1;
#pragma omp parallel for schedule(dynamic, 16)
    for (int s = 0; s < N; ++s) {
        int st = start[s], n = start[s + 1] - st;
        if (n < 3) continue;   /* as the kernel: a degenerate polygon is left as it is */
        double tx[LANES_MAX], ty[LANES_MAX], ax[LANES_MAX], ay[LANES_MAX], r[LANES_MAX], y[LANES_MAX];
        double sq[LANES_MAX], af[LANES_MAX], z;
        for (int e = 0; e < n; ++e) {
            int ef = (e + 1 >= n) ? 0 : e + 1, eg = (e == 0) ? n - 1 : e - 1;
            const double* pe = pos + 2 * (st + e);
            const double* pf = pos + 2 * (st + ef);
            const double* pg = pos + 2 * (st + eg);
            double dx = orc_wrap(pf[0] - pe[0]), dy = orc_wrap(pf[1] - pe[1]);
            double len = sqrt(dx * dx + dy * dy);
            tx[e] = 0.0; ty[e] = 0.0;
            if (!(len < 1e-14)) { tx[e] = dx / len; ty[e] = dy / len; }
            ax[e] = orc_wrap(pf[1] - pg[1]);
            ay[e] = orc_wrap(pg[0] - pf[0]);
            double fx = force[2 * (st + e)], fy = force[2 * (st + e) + 1];
            double ffx = force[2 * (st + ef)], ffy = force[2 * (st + ef) + 1];
            r[e] = tx[e] * (ffx - fx) + ty[e] * (ffy - fy);
            sq[e] = ax[e] * ax[e] + ay[e] * ay[e];
            af[e] = ax[e] * fx + ay[e] * fy;
        }
        double c = lanes_sum(sq, n), sd = lanes_sum(af, n);
        lanes_bordered_solve(n, tx, ty, ax, ay, r, c, sd, 1e-28, y, &z);
        double fo[2 * LANES_MAX];
        for (int e = 0; e < n; ++e) {
            int eg = (e == 0) ? n - 1 : e - 1;
            double px = -y[e] * tx[e] + y[eg] * tx[eg] + z * ax[e];
            double py = -y[e] * ty[e] + y[eg] * ty[eg] + z * ay[e];
            fo[2 * e] = force[2 * (st + e)] - px;
            fo[2 * e + 1] = force[2 * (st + e) + 1] - py;
        }
        for (int q = 0; q < 2 * n; ++q) force[2 * st + q] = fo[q];
    }
    return 0;
}

/* Gauss-Newton iterations of one polygon (positions p[2n]) until max|C| < tol or maxIter */
static int lanes_shake_poly(int n, double* p, const double* L, double A0, int maxIter, double tol) {
    double tx[LANES_MAX], ty[LANES_MAX], ax[LANES_MAX], ay[LANES_MAX], ce[LANES_MAX], y[LANES_MAX];
    double sq[LANES_MAX], part[LANES_MAX], z;
    int it = 0;
    for (; it < maxIter; ++it) {
        double cmax = 0.0;
        for (int e = 0; e < n; ++e) {
            int ef = (e + 1 >= n) ? 0 : e + 1, eg = (e == 0) ? n - 1 : e - 1;
            double pfx = p[2 * ef], pfy = p[2 * ef + 1], pgx = p[2 * eg], pgy = p[2 * eg + 1];
            double dx = orc_wrap(pfx - p[2 * e]), dy = orc_wrap(pfy - p[2 * e + 1]);
            double len = sqrt(dx * dx + dy * dy);
            ce[e] = len - L[e];
            tx[e] = 0.0; ty[e] = 0.0;
            if (len > 0.0) { tx[e] = dx / len; ty[e] = dy / len; }
            ax[e] = orc_wrap(pfy - pgy) / 2;
            ay[e] = orc_wrap(pgx - pfx) / 2;
            double x1 = orc_wrap(p[2 * e] - p[0]);
            double dy2 = orc_wrap(pfy - p[2 * e + 1]);
            double x2 = orc_wrap(pfx - p[0]);
            part[e] = dy2 * (x2 + x1) / 2.0;
            sq[e] = ax[e] * ax[e] + ay[e] * ay[e];
            cmax = fmax(cmax, fabs(ce[e]));
        }
        double c = lanes_sum(sq, n);
        double cA = lanes_sum(part, n) - A0;
        cmax = fmax(cmax, fabs(cA));
        if (cmax < tol) break;
        lanes_bordered_solve(n, tx, ty, ax, ay, ce, c, cA, 1e-28, y, &z);
        double q[2 * LANES_MAX];
        for (int e = 0; e < n; ++e) {
            int eg = (e == 0) ? n - 1 : e - 1;
            double dx = -y[e] * tx[e] + y[eg] * tx[eg] + z * ax[e];
            double dy = -y[e] * ty[e] + y[eg] * ty[eg] + z * ay[e];
            double qx = p[2 * e] - dx, qy = p[2 * e + 1] - dy;
            q[2 * e] = qx - floor(qx);
            q[2 * e + 1] = qy - floor(qy);
        }
        for (int k = 0; k < 2 * n; ++k) p[k] = q[k];
    }
    return it;
}

int orc_shake_lanes(int V, int N, const int* start, double* pos, const double* targetEdgeV,
                    const double* targetAreas, int maxIter, double tol) {
    (void)V;
    if (N == 0 || maxIter == 0) return 0;
    for (int s = 0; s < N; ++s)
        if (start[s + 1] - start[s] > LANES_MAX) return -1;
    int worst = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(max : worst)
    for (int s = 0; s < N; ++s) {
        int st = start[s], n = start[s + 1] - st;
        if (n < 3) continue;   /* as the kernel: a degenerate polygon is left as it is */
        int it = lanes_shake_poly(n, pos + 2 * st, targetEdgeV + st, targetAreas[s], maxIter, tol);
        if (it > worst) worst = it;
    }
    return worst;
}

/* ---------- shape matching (no counterpart in the reference) ----------
 * Every polygon becomes its rest shape ref (offsets from the rest shape's vertex mean),
 * scaled to its target area (when targetAreas is given), turned by the rotation that
 * best aligns ref with the current offsets q (minimum image from vertex 0, minus their
 * mean) and placed at the current vertex mean.  The rotation maximizes
 * trace(R^T M), M = sum q r^T: in 2-D the polar factor of the SVD of M, R = [[a, -b],
 * [b, a]] / sqrt(a^2 + b^2), a = M00 + M11, b = M10 - M01.  Sums in lane order
 * (lanes_sum), so the library's kernel reproduces it bit for bit; the least-squares
 * property is pinned by a brute-force search over the angle in the tests. */
int orc_shape_project(int V, int N, const int* start, double* pos, const double* ref,
                      const double* targetAreas) {
    (void)V;
    for (int s = 0; s < N; ++s)
        if (start[s + 1] - start[s] > LANES_MAX) return -1;
#pragma omp parallel for schedule(dynamic, 16)
    for (int s = 0; s < N; ++s) {
        int st = start[s], n = start[s + 1] - st;
        if (n < 3) continue;   /* as the kernel: a degenerate polygon is left as it is */
        double dx[LANES_MAX], dy[LANES_MAX], t[LANES_MAX];
        double p0x = pos[2 * st], p0y = pos[2 * st + 1];
        for (int e = 0; e < n; ++e) {
            dx[e] = orc_wrap(pos[2 * (st + e)] - p0x);
            dy[e] = orc_wrap(pos[2 * (st + e) + 1] - p0y);
        }
        double mx = lanes_sum(dx, n) / n, my = lanes_sum(dy, n) / n;
        double qx[LANES_MAX], qy[LANES_MAX];
        for (int e = 0; e < n; ++e) { qx[e] = dx[e] - mx; qy[e] = dy[e] - my; }
        const double* r = ref + 2 * st;
        for (int e = 0; e < n; ++e) t[e] = qx[e] * r[2 * e];
        double m00 = lanes_sum(t, n);
        for (int e = 0; e < n; ++e) t[e] = qx[e] * r[2 * e + 1];
        double m01 = lanes_sum(t, n);
        for (int e = 0; e < n; ++e) t[e] = qy[e] * r[2 * e];
        double m10 = lanes_sum(t, n);
        for (int e = 0; e < n; ++e) t[e] = qy[e] * r[2 * e + 1];
        double m11 = lanes_sum(t, n);
        for (int e = 0; e < n; ++e) {
            int en = (e + 1 >= n) ? 0 : e + 1;
            t[e] = r[2 * e] * r[2 * en + 1] - r[2 * en] * r[2 * e + 1];
        }
        double twiceA = lanes_sum(t, n);
        double a = m00 + m11, b = m10 - m01;
        double h = sqrt(a * a + b * b);
        double ca = (h > 0.0) ? a / h : 1.0, sb = (h > 0.0) ? b / h : 0.0;
        double sc = 1.0;
        if (targetAreas && twiceA > 0.0) sc = sqrt(targetAreas[s] / (twiceA / 2.0));
        for (int e = 0; e < n; ++e) {
            double ux = ca * r[2 * e] - sb * r[2 * e + 1], uy = sb * r[2 * e] + ca * r[2 * e + 1];
            double px = p0x + (mx + sc * ux), py = p0y + (my + sc * uy);
            pos[2 * (st + e)] = px - floor(px);
            pos[2 * (st + e) + 1] = py - floor(py);
        }
    }
    return 0;
}

/* ---------- XPBD: xpbdProjectCUDA model.cu:209-278, kernels.cuh:532-623 ----------
 * Colour 0 then colour 1 edge projections (disjoint within a colour for even n;
 * here applied in ascending edge order, which is the reference's result for even n),
 * then the area correction from the area/gradient reductions. */
int orc_xpbd(int V, int N, const int* start, double* pos, const double* targetEdgeP,
             const double* targetAreas, int nIter, double tol) {
    (void)V;
    int iter = 0;
    for (; iter < nIter; ++iter) {
        for (int color = 0; color < 2; ++color)
            for (int s = 0; s < N; ++s) {
                int st = start[s], n = start[s + 1] - st;
                for (int e = color; e < n; e += 2) {
                    int k = st + e, nk = st + (e + 1) % n;
                    double ex = orc_wrap(pos[2 * nk] - pos[2 * k]);
                    double ey = orc_wrap(pos[2 * nk + 1] - pos[2 * k + 1]);
                    double len = sqrt(ex * ex + ey * ey);
                    if (len < 1e-14) continue;
                    double c = 0.5 * (len - targetEdgeP[s]) / len;
                    pos[2 * k] += c * ex; pos[2 * k + 1] += c * ey;
                    pos[2 * nk] -= c * ex; pos[2 * nk + 1] -= c * ey;
                }
            }
        double maxA = 0.0, maxE = 0.0;
        for (int s = 0; s < N; ++s) {
            int st = start[s], n = start[s + 1] - st;
            double xs = pos[2 * st], ys = pos[2 * st + 1];
            double area = 0.0, g2 = 0.0;
            for (int e = 0; e < n; ++e) {
                int k = st + e, nk = st + (e + 1) % n, pk = st + (e + n - 1) % n;
                double xk = orc_wrap(pos[2 * k] - xs), yk = orc_wrap(pos[2 * k + 1] - ys);
                double xnk = orc_wrap(pos[2 * nk] - xs), ynk = orc_wrap(pos[2 * nk + 1] - ys);
                double xpk = orc_wrap(pos[2 * pk] - xs), ypk = orc_wrap(pos[2 * pk + 1] - ys);
                area += 0.5 * (xk * ynk - xnk * yk);
                double gx = 0.5 * (ynk - ypk), gy = 0.5 * (xpk - xnk);
                g2 += gx * gx + gy * gy;
            }
            double alpha = (area - targetAreas[s]) / fmax(g2, 1e-24);
            double* old = (double*)malloc(sizeof(double) * 2 * n);
            for (int q = 0; q < 2 * n; ++q) old[q] = pos[2 * st + q];
            for (int e = 0; e < n; ++e) {
                int nk = (e + 1) % n, pk = (e + n - 1) % n;
                double xnk = orc_wrap(old[2 * nk] - xs), ynk = orc_wrap(old[2 * nk + 1] - ys);
                double xpk = orc_wrap(old[2 * pk] - xs), ypk = orc_wrap(old[2 * pk + 1] - ys);
                double gx = 0.5 * (ynk - ypk), gy = 0.5 * (xpk - xnk);
                pos[2 * (st + e)] -= alpha * gx;
                pos[2 * (st + e) + 1] -= alpha * gy;
            }
            free(old);
            double dA = fabs(area - targetAreas[s]);
            if (dA > maxA) maxA = dA;
        }
        if (tol > 0) {
            for (int s = 0; s < N; ++s) {
                int st = start[s], n = start[s + 1] - st;
                for (int e = 0; e < n; ++e) {
                    int k = st + e, nk = st + (e + 1) % n;
                    double ex = orc_wrap(pos[2 * nk] - pos[2 * k]);
                    double ey = orc_wrap(pos[2 * nk + 1] - pos[2 * k + 1]);
                    double d = fabs(sqrt(ex * ex + ey * ey) - targetEdgeP[s]);
                    if (d > maxE) maxE = d;
                }
            }
            if (maxE < tol && maxA < tol) break;
        }
    }
    return iter;
}

/* ---------- XPBD with compliance, in the HIP kernel's lane order ----------
 * Extends xpbdProjectCUDA model.cu:209-278 / kernels.cuh:532-623 (no compliance, edges +
 * area) to compliant constraints (Macklin, Mueller & Chentanez, "XPBD: position-based
 * simulation of compliant constrained dynamics", MIG 2016): every constraint j keeps a
 * multiplier lambda_j, zero at the start of the call, and a projection of j applies
 *     dlambda = (-C_j - at_j lambda_j) / (sum_k w_k |grad_k C_j|^2 + at_j),
 *     lambda_j += dlambda,  x_k += w_k grad_k C_j dlambda        (unit inverse masses)
 * with at_j = 1 / (k_j dt^2) for stiffness k_j (k = INFINITY: at = 0, the reference's
 * hard projection; k = 0: the constraint is off).  Constraints per polygon: its n edge
 * lengths (target per vertex, or one per polygon), its area (shoelace relative to the
 * start vertex, kernels.cuh:567-577) and its perimeter (target: the sum of the edge
 * targets in vertex order).  One iteration: edges of colour 0 (even index; for odd n the
 * last edge after the others, as it shares vertex 0 with edge 0), colour 1, then the area,
 * then the perimeter.  With at = 0 each edge and area step is the reference's arithmetic
 * (c = 0.5 (len - L)/len, alpha = (A - A0)/max(|grad A|^2, 1e-24)).  Sums over a polygon
 * are butterfly sums (lanes_sum).  Convergence (tol > 0) is per polygon: after an
 * iteration the largest |C + at lambda| of the edges (at the positions after the
 * iteration), of the area and of the perimeter (each before its own correction) is
 * compared with tol, as the reference's deviation kernels (kernels.cuh:603-623) do with
 * at = 0; the returned count is the largest index at which a polygon stopped (nIter if
 * one did not).  Positions are not wrapped into the box (the reference does not either).
 * Polygons of at most 64 vertices (-1 otherwise); fewer than 3 are left as they are. */
int orc_xpbd_lanes(int V, int N, const int* start, double* pos, const double* targetEdgeV,
                   const double* targetEdgeP, const double* targetAreas, double kE, double kA,
                   double kP, double dt, int nIter, double tol) {
    (void)V;
    for (int s = 0; s < N; ++s)
        if (start[s + 1] - start[s] > LANES_MAX) return -1;
    int edgeOn = kE != 0.0, areaOn = kA != 0.0, perimOn = kP != 0.0;
    double atE = edgeOn ? 1.0 / (kE * (dt * dt)) : 0.0;
    double atA = areaOn ? 1.0 / (kA * (dt * dt)) : 0.0;
    double atP = perimOn ? 1.0 / (kP * (dt * dt)) : 0.0;
    int worst = 0;
    for (int s = 0; s < N; ++s) {
        int st = start[s], n = start[s + 1] - st;
        if (n < 3) continue;
        double x[LANES_MAX], y[LANES_MAX], L[LANES_MAX], lam[LANES_MAX];
        double t1[LANES_MAX], t2[LANES_MAX];
        double P0 = 0.0;
        for (int e = 0; e < n; ++e) {
            x[e] = pos[2 * (st + e)]; y[e] = pos[2 * (st + e) + 1];
            L[e] = targetEdgeV ? targetEdgeV[st + e] : (targetEdgeP ? targetEdgeP[s] : 0.0);
            lam[e] = 0.0;
        }
        for (int e = 0; e < n; ++e) P0 += L[e];
        double A0 = areaOn ? targetAreas[s] : 0.0;
        double lamA = 0.0, lamP = 0.0;
        int it = 0;
        for (; it < nIter; ++it) {
            if (edgeOn) {
                int nPass = (n & 1) ? 3 : 2;
                for (int pass = 0; pass < nPass; ++pass) {
                    double dx[LANES_MAX], dy[LANES_MAX];
                    int act[LANES_MAX];
                    for (int e = 0; e < n; ++e) {
                        int en = (e + 1 == n) ? 0 : e + 1;
                        int last = (n & 1) && e == n - 1;
                        act[e] = (pass == 0) ? (!(e & 1) && !last) : (pass == 1 && nPass == 3) ? last : (e & 1);
                        dx[e] = 0.0; dy[e] = 0.0;
                        if (!act[e]) continue;
                        double ex = orc_wrap(x[en] - x[e]), ey = orc_wrap(y[en] - y[e]);
                        double len = sqrt(ex * ex + ey * ey);
                        if (len < 1e-14) { act[e] = 0; continue; }
                        double C = len - L[e];
                        double dl = (-C - atE * lam[e]) / (2.0 + atE);
                        lam[e] += dl;
                        double c = -dl / len;
                        dx[e] = c * ex; dy[e] = c * ey;
                    }
                    double nx[LANES_MAX], ny[LANES_MAX];
                    for (int e = 0; e < n; ++e) {
                        int ep = (e == 0) ? n - 1 : e - 1;
                        nx[e] = x[e]; ny[e] = y[e];
                        if (act[e]) { nx[e] = x[e] + dx[e]; ny[e] = y[e] + dy[e]; }
                        if (act[ep]) { nx[e] = x[e] - dx[ep]; ny[e] = y[e] - dy[ep]; }
                    }
                    for (int e = 0; e < n; ++e) { x[e] = nx[e]; y[e] = ny[e]; }
                }
            }
            double devA = 0.0, devP = 0.0, devE = 0.0;
            double gAx[LANES_MAX], gAy[LANES_MAX], gPx[LANES_MAX], gPy[LANES_MAX];
            double rA = 0.0, rP = 0.0, aa = 0.0, pp = 0.0, ap = 0.0;
            if (areaOn) {
                double xs = x[0], ys = y[0];
                double term[LANES_MAX], g2v[LANES_MAX];
                for (int e = 0; e < n; ++e) {
                    int en = (e + 1 == n) ? 0 : e + 1, ep = (e == 0) ? n - 1 : e - 1;
                    double xk = orc_wrap(x[e] - xs), yk = orc_wrap(y[e] - ys);
                    double xnk = orc_wrap(x[en] - xs), ynk = orc_wrap(y[en] - ys);
                    double xpk = orc_wrap(x[ep] - xs), ypk = orc_wrap(y[ep] - ys);
                    term[e] = 0.5 * (xk * ynk - xnk * yk);
                    gAx[e] = 0.5 * (ynk - ypk);
                    gAy[e] = 0.5 * (xpk - xnk);
                    g2v[e] = gAx[e] * gAx[e] + gAy[e] * gAy[e];
                }
                double area = lanes_sum(term, n);
                aa = lanes_sum(g2v, n);
                rA = (area - A0) + atA * lamA;
                devA = fabs(rA);
            }
            if (perimOn) {
                double len[LANES_MAX], gp2v[LANES_MAX];
                for (int e = 0; e < n; ++e) {
                    int en = (e + 1 == n) ? 0 : e + 1;
                    double ex = orc_wrap(x[en] - x[e]), ey = orc_wrap(y[en] - y[e]);
                    len[e] = sqrt(ex * ex + ey * ey);
                    t1[e] = (len[e] >= 1e-14) ? ex / len[e] : 0.0;
                    t2[e] = (len[e] >= 1e-14) ? ey / len[e] : 0.0;
                }
                for (int e = 0; e < n; ++e) {
                    int ep = (e == 0) ? n - 1 : e - 1;
                    gPx[e] = t1[ep] - t1[e];
                    gPy[e] = t2[ep] - t2[e];
                    gp2v[e] = gPx[e] * gPx[e] + gPy[e] * gPy[e];
                }
                double P = lanes_sum(len, n);
                pp = lanes_sum(gp2v, n);
                rP = (P - P0) + atP * lamP;
                devP = fabs(rP);
            }
            if (areaOn && perimOn) {   /* the 2x2 block of the two polygon-wide constraints */
                double apv[LANES_MAX];
                for (int e = 0; e < n; ++e) apv[e] = gAx[e] * gPx[e] + gAy[e] * gPy[e];
                ap = lanes_sum(apv, n);
                double m11 = aa + atA, m22 = pp + atP;
                double det = m11 * m22 - ap * ap;
                if (det > 1e-12 * (m11 * m22)) {
                    double dA = (ap * rP - m22 * rA) / det;
                    double dP = (ap * rA - m11 * rP) / det;
                    lamA += dA;
                    lamP += dP;
                    for (int e = 0; e < n; ++e) {
                        x[e] += gAx[e] * dA + gPx[e] * dP;
                        y[e] += gAy[e] * dA + gPy[e] * dP;
                    }
                } else {   /* (nearly) parallel gradients: area step only */
                    double alpha = rA / (fmax(aa, 1e-24) + atA);
                    lamA -= alpha;
                    for (int e = 0; e < n; ++e) { x[e] -= alpha * gAx[e]; y[e] -= alpha * gAy[e]; }
                }
            } else if (areaOn) {
                double alpha = rA / (fmax(aa, 1e-24) + atA);
                lamA -= alpha;
                for (int e = 0; e < n; ++e) { x[e] -= alpha * gAx[e]; y[e] -= alpha * gAy[e]; }
            } else if (perimOn) {
                double alpha = rP / (fmax(pp, 1e-24) + atP);
                lamP -= alpha;
                for (int e = 0; e < n; ++e) { x[e] -= alpha * gPx[e]; y[e] -= alpha * gPy[e]; }
            }
            if (tol > 0) {
                if (edgeOn)
                    for (int e = 0; e < n; ++e) {
                        int en = (e + 1 == n) ? 0 : e + 1;
                        double ex = orc_wrap(x[en] - x[e]), ey = orc_wrap(y[en] - y[e]);
                        double d = fabs(sqrt(ex * ex + ey * ey) - L[e] + atE * lam[e]);
                        if (d > devE) devE = d;
                    }
                if ((!edgeOn || devE < tol) && (!areaOn || devA < tol) && (!perimOn || devP < tol)) break;
            }
        }
        for (int e = 0; e < n; ++e) { pos[2 * (st + e)] = x[e]; pos[2 * (st + e) + 1] = y[e]; }
        if (it > worst) worst = it;
    }
    return worst;
}

/* ---------- GD pieces ----------
 * updatePositionsKernel kernels.cuh:1433-1438: every coordinate x <- frac(x + F dt). */
void orc_update_positions(int V, double* pos, const double* force, double dt) {
    for (int q = 0; q < 2 * V; ++q) {
        double p = pos[q] + force[q] * dt;
        pos[q] = p - floor(p);
    }
}
/* effectiveForceMagKernel kernels.cuh:625-633 with the max of getMaxEffectiveForceCUDA
 * model.cu:286-306: max over vertices of |F_k + scale (x_k - tent_k)| (no wrap; scale = 1/dt for
 * GD and 2/dt^2 for FIRE, Model::getMaxEffectiveForce model.cpp:332-335). */
double orc_max_effective_force(int V, const double* pos, const double* tent, const double* force,
                               double scale) {
    double m = 0.0;
    for (int k = 0; k < V; ++k) {
        double fx = force[2 * k] + scale * (pos[2 * k] - tent[2 * k]);
        double fy = force[2 * k + 1] + scale * (pos[2 * k + 1] - tent[2 * k + 1]);
        m = fmax(m, sqrt(fx * fx + fy * fy));
    }
    return m;
}

/* ---------- Monte-Carlo overlap counter: updateOverlapAreaKernel kernels.cuh:879-979 ---------- */
static double wrapPeriodic(double x) {   /* kernels.cuh:108-113 */
    x += 1.5;
    while (x > 1.0) x -= 1.0;
    return x - 0.5;
}
long long orc_overlap_count(int V, int N, const int* start, const double* pos, int boxSize,
                            const int* neighborIndices, const int* countPerBox,
                            int pointDensity, int* counter) {
    const double eps = 1e-12, pi = 3.141592653589793238462643383279;
    int* shapeId = (int*)malloc(sizeof(int) * (V > 0 ? V : 1));
    for (int s = 0; s < N; ++s) for (int i = start[s]; i < start[s + 1]; ++i) shapeId[i] = s;
    int total = pointDensity * pointDensity, boxCount = boxSize * boxSize;
    long long sum = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : sum)
    for (int idx = 0; idx < total; ++idx) {
        int ix = idx % pointDensity, iy = idx / pointDensity;
        double px = (ix + 0.5) / pointDensity, py = (iy + 0.5) / pointDensity;
        int num = 0;
        int cellX = (int)(px * boxSize), cellY = (int)(py * boxSize);
        int tracked[512], nt = 0, full = 0;
        for (int dx = -1; dx <= 1 && !full; dx++) {
            int nx = cellX + dx;
            if (nx < 0) nx += boxSize; else if (nx >= boxSize) nx -= boxSize;
            for (int dy = -1; dy <= 1 && !full; dy++) {
                int ny = cellY + dy;
                if (ny < 0) ny += boxSize; else if (ny >= boxSize) ny -= boxSize;
                int boxId = ny * boxSize + nx;
                int si = countPerBox[boxId];
                int sf = (boxId + 1 < boxCount) ? countPerBox[boxId + 1] : V;
                for (int q = si; q < sf && !full; q++) {
                    int poly = shapeId[neighborIndices[q]];
                    int seen = 0;
                    for (int t = 0; t < nt; t++) if (tracked[t] == poly) { seen = 1; break; }
                    if (seen) continue;
                    int st = start[poly], en = start[poly + 1];
                    if (en - st < 3) continue;
                    double angleSum = 0.0;
                    double vxPrev = wrapPeriodic(pos[2 * st] - px);
                    double vyPrev = wrapPeriodic(pos[2 * st + 1] - py);
                    double vx0 = vxPrev + 0.0, vy0 = vyPrev + 0.0;
                    for (int e = st + 1; e < en; e++) {
                        double vx = vxPrev + wrapPeriodic(pos[2 * e] - pos[2 * e - 2]);
                        double vy = vyPrev + wrapPeriodic(pos[2 * e + 1] - pos[2 * e - 1]);
                        double cross = vxPrev * vy - vyPrev * vx;
                        double dot = vxPrev * vx + vyPrev * vy;
                        angleSum += atan2(cross, dot);
                        vxPrev = vx; vyPrev = vy;
                    }
                    angleSum += atan2(vxPrev * vy0 - vyPrev * vx0, vxPrev * vx0 + vyPrev * vy0);
                    if (fabs(angleSum - 2.0 * pi) < eps) num++;
                    if (nt < 512) tracked[nt++] = poly; else full = 1;
                }
            }
        }
        int c = num * (num - 1) / 2;
        if (counter) counter[idx] = c;
        sum += c;
    }
    free(shapeId);
    return sum;
}

/* ---------- one force evaluation + FIRE ---------- */
typedef struct {
    int* cellLocation; int* neighborIndices; int* countPerBox;
    int M; int* neighbors; unsigned char* inside; double* tuSlots; int* numNeighbors;
    uint64_t* inter; uint64_t* outer; double* tu; double* ut; int* shapeCounts;
    double* edgeLengths; double* areas; double* comX; double* comY; double* constraints;
    int I;
} orc_ws;

static void ws_free(orc_ws* w) {
    free(w->cellLocation); free(w->neighborIndices); free(w->countPerBox);
    free(w->neighbors); free(w->inside); free(w->tuSlots); free(w->numNeighbors);
    free(w->inter); free(w->outer); free(w->tu); free(w->ut); free(w->shapeCounts);
    free(w->edgeLengths); free(w->areas); free(w->comX); free(w->comY); free(w->constraints);
}
static void ws_alloc(orc_ws* w, int V, int N, int B, int M) {
    memset(w, 0, sizeof(*w));
    int Vp = V > 0 ? V : 1, Np = N > 0 ? N : 1;
    w->cellLocation = (int*)malloc(sizeof(int) * Vp);
    w->neighborIndices = (int*)malloc(sizeof(int) * Vp);
    w->countPerBox = (int*)malloc(sizeof(int) * B * B);
    w->M = M;
    w->neighbors = (int*)malloc(sizeof(int) * (size_t)Vp * M);
    w->inside = (unsigned char*)malloc((size_t)Vp * M);
    w->tuSlots = (double*)malloc(sizeof(double) * 2 * (size_t)Vp * M);
    w->numNeighbors = (int*)malloc(sizeof(int) * Vp);
    w->inter = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)Vp * M);
    w->outer = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)Vp * M);
    w->tu = (double*)malloc(sizeof(double) * 2 * (size_t)Vp * M);
    w->ut = (double*)malloc(sizeof(double) * 2 * (size_t)Vp * M);
    w->shapeCounts = (int*)malloc(sizeof(int) * Np);
    w->edgeLengths = (double*)malloc(sizeof(double) * Vp);
    w->areas = (double*)malloc(sizeof(double) * Np);
    w->comX = (double*)malloc(sizeof(double) * Np);
    w->comY = (double*)malloc(sizeof(double) * Np);
    w->constraints = (double*)malloc(sizeof(double) * 6 * Vp);
}

static void detect(const orc_system* s, const double* pos, orc_ws* w) {
    orc_cells(s->V, s->N, pos, s->start, s->boxSize, w->cellLocation, w->neighborIndices, w->countPerBox);
    int cons = (s->model & ORC_CONSISTENT) != 0;
    int mx = orc_neighbors_ex(s->V, s->N, pos, s->start, s->boxSize, w->cellLocation, w->neighborIndices,
                              w->countPerBox, w->M, w->neighbors, w->inside, w->tuSlots, w->numNeighbors, cons);
    if (mx > w->M) {   /* model.cpp:264-305: grow and retry once */
        int newM = (2 * w->M + 1 > mx) ? 2 * w->M + 1 : mx;
        ws_free(w);
        ws_alloc(w, s->V, s->N, s->boxSize, newM);
        orc_cells(s->V, s->N, pos, s->start, s->boxSize, w->cellLocation, w->neighborIndices, w->countPerBox);
        orc_neighbors_ex(s->V, s->N, pos, s->start, s->boxSize, w->cellLocation, w->neighborIndices,
                         w->countPerBox, w->M, w->neighbors, w->inside, w->tuSlots, w->numNeighbors, cons);
    }
}
static void records(const orc_system* s, orc_ws* w, const double* pos) {
    const double* ex = (s->model & ORC_AREA_MODES) ? pos : NULL;
    w->I = orc_outersections_ex(s->V, s->N, s->start, w->M, w->neighbors, w->inside, w->tuSlots,
                                w->numNeighbors, w->inter, w->tu, w->outer, w->ut, w->shapeCounts, ex);
}
static void geometry(const orc_system* s, const double* pos, orc_ws* w) {
    orc_geometry(s->V, s->N, pos, s->start, w->edgeLengths, w->areas, w->comX, w->comY,
                 w->constraints, NULL, NULL);
}
static double forces_ex(const orc_system* s, const double* pos, orc_ws* w, double* force, double* fext) {
    double E = orc_force_energy_ex(s->model, s->V, s->N, s->start, pos, w->I, w->inter, w->outer,
                                   w->tu, w->ut, w->edgeLengths, w->areas, s->targetEdgeV, s->targetAreas,
                                   s->stiffness, s->compressibility, s->perimStiffness, force, fext);
    if ((s->model & 0xFF) == ORC_NORMAL) {
        if (s->solver == ORC_SOLVER_LANES) orc_project_force_lanes(s->V, s->N, s->start, pos, force);
        else orc_project_force(s->V, s->N, s->start, pos, force);
    }
    return E;
}
static double forces(const orc_system* s, const double* pos, orc_ws* w, double* force) {
    return forces_ex(s, pos, w, force, NULL);
}
static int fire_shake(const orc_system* s, double* pos, int maxIter) {
    if (s->solver == ORC_SOLVER_LANES)
        return orc_shake_lanes(s->V, s->N, s->start, pos, s->targetEdgeV, s->targetAreas, maxIter, 1e-15);
    return orc_shake(s->V, s->N, s->start, pos, s->targetEdgeV, s->targetAreas, maxIter, 1e-15);
}

double orc_force_eval(const orc_system* s, const double* pos, double* force, int* numIntersections) {
    return orc_force_eval_ex(s, pos, force, NULL, numIntersections);
}
/* as orc_force_eval, with fext [2V] (optional): the force before the spring terms */
double orc_force_eval_ex(const orc_system* s, const double* pos, double* force, double* fext,
                         int* numIntersections) {
    orc_ws w;
    ws_alloc(&w, s->V, s->N, s->boxSize, s->maxNeighbors > 0 ? s->maxNeighbors : 100);
    geometry(s, pos, &w);
    int overlap = ((s->model & 0xFF) == ORC_NORMAL || (s->model & 0xFF) == ORC_HYBRID);
    if (overlap) { detect(s, pos, &w); records(s, &w, pos); }
    double E = forces_ex(s, pos, &w, force, fext);
    if (numIntersections) *numIntersections = w.I;
    ws_free(&w);
    return E;
}

/* per-polygon sequential dot, then tree over polygons */
static double fire_dot(const orc_system* s, const double* a, const double* b) {
    double* pp = (double*)malloc(sizeof(double) * (s->N > 0 ? s->N : 1));
    for (int q = 0; q < s->N; ++q) {
        double acc = 0.0;
        for (int m = s->start[q]; m < s->start[q + 1]; ++m) {
            acc += a[2 * m] * b[2 * m];
            acc += a[2 * m + 1] * b[2 * m + 1];
        }
        pp[q] = acc;
    }
    double r = orc_det_sum(pp, s->N);
    free(pp);
    return r;
}

/* minimizeFIRE's prologue model.cpp:489-496 */
void orc_fire_init(const orc_system* s, double* pos, double* vel, double* force, orc_fire_state* st) {
    for (int q = 0; q < 2 * s->V; ++q) vel[q] = 0.0;
    int nI = 0;
    st->energy = orc_force_eval(s, pos, force, &nI);
    st->lastShakeIters = 0;
    st->shakeFailed = 0;
    st->rolledBack = 0;
}

/* Model::minimizeFIREStep model.cpp:405-486.  The FIRE kernels (FIRE.h) are absent
 * from the reference; their semantics are fixed here from the launch signatures
 * model.cu:576-608 and Bitzek et al., PRL 97, 170201 (2006):
 *   updatePositionAndVelocity:  vh = v + (0.5*dt)*F; x = frac(x + dt*vh); v = vh
 *   updateVelocity:             v = v + (0.5*dt)*F
 *   bendVelocityTowardsForce:   P = F.v, |v|, |F|; v = (1-a)*v + ((a*|v|)/|F|)*F
 *   rederiveVelocity:           v = v + wrap(x_shaken - x_unshaken)/dt            */
void orc_fire_step(const orc_system* s, double* pos, double* vel, double* force, orc_fire_state* st,
                   double dtMax, double alphaStart, double fAlpha, double fInc, double fDec,
                   int nMin, int shakeIter) {
    orc_fire_step_ex(s, pos, vel, force, st, dtMax, alphaStart, fAlpha, fInc, fDec, nMin, shakeIter, 0, 0.0);
}

/* orc_fire_step; with xpbdIter > 0 (models with springs; no counterpart in the reference,
 * whose XPBD launcher is never called, K10) the springs of the model (edge: stiffness,
 * area: compressibility, perimeter: perimStiffness) are integrated implicitly by XPBD and
 * only the rest of the force (fext, the overlap force) explicitly -- XPBD's own time
 * step (Macklin et al. 2016, Algorithm 1) inside FIRE:
 *   v' = v + dt fext;  x~ = frac(x + dt v');  x' = XPBD(x~; at = 1/(k dt^2)), wrapped;
 *   v = v' + wrap(x' - x~)/dt;  F = total force at x' (springs included)
 * then the energy test and the mixing of model.cpp:444-485 on the total energy and force,
 * with no second half-kick.  A fixed point (v = 0) has fext = sum k C grad C, i.e. a zero
 * total force. */
void orc_fire_step_ex(const orc_system* s, double* pos, double* vel, double* force, orc_fire_state* st,
                      double dtMax, double alphaStart, double fAlpha, double fInc, double fDec,
                      int nMin, int shakeIter, int xpbdIter, double xpbdTol) {
    int V2 = 2 * s->V;
    int mt = s->model & 0xFF;
    int needsInter = (mt == ORC_NORMAL || mt == ORC_HYBRID);
    int rigid = (mt == ORC_NORMAL);
    int xpbd = xpbdIter > 0 && !rigid && mt != ORC_ABNORMAL;
    double dt = st->dt;
    double* bak = (double*)malloc(sizeof(double) * V2);
    double* tmp = (double*)malloc(sizeof(double) * V2);
    memcpy(bak, pos, sizeof(double) * V2);
    double ePre = st->energy;
    if (xpbd) {
        double* fext = (double*)malloc(sizeof(double) * V2);
        double* ftmp = (double*)malloc(sizeof(double) * V2);
        orc_force_eval_ex(s, pos, ftmp, fext, NULL);   /* the step's start state, as stored */
        for (int q = 0; q < V2; ++q) {
            double vh = vel[q] + dt * fext[q];
            double p = pos[q] + dt * vh;
            pos[q] = p - floor(p);
            vel[q] = vh;
        }
        memcpy(tmp, pos, sizeof(double) * V2);
        int edgeOn = (mt == ORC_SOFT_BODY || mt == ORC_EDGE_ONLY || mt == ORC_HYBRID);
        int areaOn = (mt == ORC_SOFT_BODY || mt == ORC_AREA_ONLY || mt == ORC_HYBRID);
        int perimOn = (mt == ORC_SOFT_BODY || mt == ORC_HYBRID);
        int it = orc_xpbd_lanes(s->V, s->N, s->start, pos, s->targetEdgeV, NULL, s->targetAreas,
                                edgeOn ? s->stiffness : 0.0, areaOn ? s->compressibility : 0.0,
                                perimOn ? s->perimStiffness : 0.0, dt, xpbdIter, xpbdTol);
        st->lastShakeIters = it;
        for (int q = 0; q < V2; ++q) {
            double p = pos[q] - floor(pos[q]);
            pos[q] = p;
            vel[q] = vel[q] + orc_wrap(p - tmp[q]) / dt;
        }
        free(fext); free(ftmp);
    } else {
        for (int q = 0; q < V2; ++q) {
            double vh = vel[q] + (0.5 * dt) * force[q];
            double p = pos[q] + dt * vh;
            pos[q] = p - floor(p);
            vel[q] = vh;
        }
    }
    orc_ws w;
    ws_alloc(&w, s->V, s->N, s->boxSize, s->maxNeighbors > 0 ? s->maxNeighbors : 100);
    int shakeFailed = 0;
    st->rolledBack = 0;
    int areaMode = (s->model & ORC_AREA_MODES) != 0;
    if (needsInter && xpbd) {
        detect(s, pos, &w);
        records(s, &w, pos);
    } else if (needsInter) {
        if (!areaMode) detect(s, pos, &w);   /* reference order: detect, SHAKE, records */
        if (rigid && shakeIter > 0) {
            memcpy(tmp, pos, sizeof(double) * V2);
            int it = fire_shake(s, pos, shakeIter);
            st->lastShakeIters = it;
            if (it >= shakeIter) {
                it = fire_shake(s, pos, shakeIter * 10);
                st->lastShakeIters = it;
                if (it >= shakeIter * 10) shakeFailed = 1;
            }
            if (shakeFailed) {
                memcpy(pos, bak, sizeof(double) * V2);
                for (int q = 0; q < V2; ++q) vel[q] = 0.0;
            } else {
                for (int q = 0; q < V2; ++q) vel[q] = vel[q] + orc_wrap(pos[q] - tmp[q]) / dt;
            }
        }
        if (areaMode) detect(s, pos, &w);    /* area energy: crossings of the final positions */
        records(s, &w, pos);
    }
    geometry(s, pos, &w);
    double ePost = forces(s, pos, &w, force);
    st->energy = ePost;
    st->shakeFailed = shakeFailed;
    if (shakeFailed) {
        st->dt = dt * fDec; st->alpha = alphaStart; st->nPos = 0;
    } else if (ePost > ePre + fmax(fabs(ePre) * 1e-10, 1e-14)) {
        memcpy(pos, bak, sizeof(double) * V2);
        geometry(s, pos, &w);
        if (needsInter) { detect(s, pos, &w); records(s, &w, pos); }
        st->energy = forces(s, pos, &w, force);
        for (int q = 0; q < V2; ++q) vel[q] = 0.0;
        st->dt = dt * fDec; st->alpha = alphaStart; st->nPos = 0;
        st->rolledBack = 1;
    } else {
        if (!xpbd)
            for (int q = 0; q < V2; ++q) vel[q] = vel[q] + (0.5 * dt) * force[q];
        double Pw = fire_dot(s, force, vel);
        double vnorm = sqrt(fire_dot(s, vel, vel));
        double fnorm = sqrt(fire_dot(s, force, force));
        double a = st->alpha;
        double scale = (fnorm > 0.0) ? (a * vnorm) / fnorm : 0.0;
        for (int q = 0; q < V2; ++q) vel[q] = (1.0 - a) * vel[q] + scale * force[q];
        if (Pw > 0.0) {
            if (++st->nPos >= nMin) {
                double newDt = fmin(dt * fInc, dtMax);
                if (newDt > dt) st->alpha = st->alpha * fAlpha;
                dt = newDt;
            }
            st->dt = dt;
        } else {
            for (int q = 0; q < V2; ++q) vel[q] = 0.0;
            st->dt = dt * fDec; st->alpha = alphaStart; st->nPos = 0;
        }
    }
    ws_free(&w);
    free(bak); free(tmp);
}

/* Model::minimizeFIRE model.cpp:488-504 (convergence on max |force component|,
 * model.cu:612-614).  Returns the number of steps taken. */
int orc_fire_minimize(const orc_system* s, double* pos, double* vel, double* force,
                      orc_fire_state* st, double fTol, int maxSteps, double dtMax,
                      double alphaStart, double fAlpha, double fInc, double fDec, int nMin,
                      int shakeIter) {
    return orc_fire_minimize_ex(s, pos, vel, force, st, fTol, maxSteps, dtMax, alphaStart, fAlpha, fInc,
                                fDec, nMin, shakeIter, 0, 0.0);
}
int orc_fire_minimize_ex(const orc_system* s, double* pos, double* vel, double* force,
                         orc_fire_state* st, double fTol, int maxSteps, double dtMax,
                         double alphaStart, double fAlpha, double fInc, double fDec, int nMin,
                         int shakeIter, int xpbdIter, double xpbdTol) {
    st->alpha = alphaStart;
    st->nPos = 0;
    orc_fire_init(s, pos, vel, force, st);
    for (int step = 0; step < maxSteps; ++step) {
        orc_fire_step_ex(s, pos, vel, force, st, dtMax, alphaStart, fAlpha, fInc, fDec, nMin, shakeIter,
                         xpbdIter, xpbdTol);
        if (orc_max_abs(force, 2 * s->V) <= fTol) return step + 1;
    }
    return maxSteps;
}
