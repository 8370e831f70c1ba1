// Device-side geometry primitives for the overlap-force path (gfx950, FP64 VALU).
//
// Every helper evaluates its expression operation by operation in the order of
// the reference source (camdennis/pyCudaPolygon src/kernels.cuh, cited per
// function); the library is compiled with -ffp-contract=off so no FMA changes a
// rounding.  This is what makes crossing decisions, float-rounded intersection
// parameters and force contributions reproducible bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcp {

struct d2 { double x, y; };

__device__ __forceinline__ d2 mk(double x, double y) { d2 r; r.x = x; r.y = y; return r; }

// kernels.cuh:115-119
__device__ __forceinline__ double wrap(double x) {
    double y = x + 1.5;
    y = y - floor(y);
    return y - 0.5;
}
__device__ __forceinline__ d2 wrap2(d2 v) { return mk(wrap(v.x), wrap(v.y)); }

// while(x<0)x+=1; while(x>=1)x-=1 (kernels.cuh:670-699, 1303-1316), bounded:
// beyond |x| = 2^20 (never reached from positions in the box) floor is used.
__device__ __forceinline__ double shift_into_unit(double x) {
    if (!(fabs(x) < 1048576.0)) return (x == x) ? x - floor(x) : x;
    while (x < 0.0) x += 1.0;
    while (x >= 1.0) x -= 1.0;
    return x;
}

__device__ __forceinline__ d2 ldpos(const double2* __restrict__ pos, int v) {
    double2 p = pos[v];
    return mk(p.x, p.y);
}

// h, g12: kernels.cuh:139-159
__device__ __forceinline__ double hfun(d2 p1, d2 p2, d2 sp) {
    d2 r1 = wrap2(mk(p1.x - sp.x, p1.y - sp.y));
    d2 r2 = wrap2(mk(p2.x - sp.x, p2.y - sp.y));
    d2 d = wrap2(mk(r2.x - r1.x, r2.y - r1.y));
    return (r1.x + r2.x) * d.y;
}
__device__ __forceinline__ void g12(d2 p1, d2 p2, d2 sp, d2& g1, d2& g2) {
    d2 r1 = wrap2(mk(p1.x - sp.x, p1.y - sp.y));
    d2 r2 = wrap2(mk(p2.x - sp.x, p2.y - sp.y));
    d2 d = wrap2(mk(p2.x - p1.x, p2.y - p1.y));
    g1.x = d.y; g1.y = -(r1.x + r2.x);
    g2.x = d.y; g2.y = (r1.x + r2.x);
}

// getDf: kernels.cuh:161-222.  df[col*2+row], 8 columns (vi, vzi, vj, vzj) x 2.
__device__ __forceinline__ void getDf(d2 vi, d2 vzi, d2 vj, d2 vzj, double* df) {
    d2 dj = wrap2(mk(vzj.x - vj.x, vzj.y - vj.y));
    d2 di = wrap2(mk(vzi.x - vi.x, vzi.y - vi.y));
    d2 dij = wrap2(mk(vj.x - vi.x, vj.y - vi.y));
    double w = dj.x * di.y - dj.y * di.x;
    double diNorm2 = di.x * di.x + di.y * di.y;
    double djNorm2 = dj.x * dj.x + dj.y * dj.y;
#pragma unroll
    for (int i = 0; i < 16; ++i) df[i] = 0.0;
    if (w * w < 1e-12 * diNorm2 * djNorm2 || diNorm2 < 1e-28 || djNorm2 < 1e-28) return;
    double k = dj.x * dij.y - dj.y * dij.x;
    double u = k / w;
    double dk[8];
    dk[0] = dj.y; dk[1] = -dj.x; dk[2] = 0.0; dk[3] = 0.0;
    dk[4] = -dij.y - dj.y; dk[5] = dj.x + dij.x;
    dk[6] = dij.y; dk[7] = -dij.x;
    double dw[8];
    dw[0] = dj.y; dw[1] = -dj.x; dw[2] = -dj.y; dw[3] = dj.x;
    dw[4] = -di.y; dw[5] = di.x; dw[6] = di.y; dw[7] = -di.x;
#pragma unroll
    for (int alpha = 0; alpha < 2; ++alpha) {
#pragma unroll
        for (int beta = 0; beta < 2; ++beta) {
#pragma unroll
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

// parameter along edge (vi -> vzi) of its crossing with the line (vj -> vzj): the u
// of getDf (kernels.cuh:163-175); `fallback` for (near) parallel edges
__device__ __forceinline__ double cross_param(d2 vi, d2 vzi, d2 vj, d2 vzj, double fallback) {
    d2 dj = wrap2(mk(vzj.x - vj.x, vzj.y - vj.y));
    d2 di = wrap2(mk(vzi.x - vi.x, vzi.y - vi.y));
    d2 dij = wrap2(mk(vj.x - vi.x, vj.y - vi.y));
    double w = dj.x * di.y - dj.y * di.x;
    double diNorm2 = di.x * di.x + di.y * di.y;
    double djNorm2 = dj.x * dj.x + dj.y * dj.y;
    if (w * w < 1e-12 * diNorm2 * djNorm2 || diNorm2 < 1e-28 || djNorm2 < 1e-28) return fallback;
    double k = dj.x * dij.y - dj.y * dij.x;
    return k / w;
}

// g . df[cols c0, c0+1]  (the "inc" of kernels.cuh:1134-1142)
__device__ __forceinline__ d2 dfdot(const double* df, int c0, d2 g) {
    d2 inc;
    inc.x = g.x * df[c0 * 2 + 0] + g.y * df[c0 * 2 + 1];
    inc.y = g.x * df[(c0 + 1) * 2 + 0] + g.y * df[(c0 + 1) * 2 + 1];
    return inc;
}

// record key fields (kernels.cuh:860-874): [63:48]=sj [47:32]=si [31:16]=i [15:0]=j
__device__ __forceinline__ int key_sj(uint64_t k) { return (int)((k >> 48) & 0xFFFF); }
__device__ __forceinline__ int key_si(uint64_t k) { return (int)((k >> 32) & 0xFFFF); }
__device__ __forceinline__ int key_i(uint64_t k) { return (int)((k >> 16) & 0xFFFF); }
__device__ __forceinline__ int key_j(uint64_t k) { return (int)(k & 0xFFFF); }

// order inside one si bucket: (i, j, sj) -- the order left by the reference's
// 64-bit then 48-bit stable radix sorts (model.cpp:354-360)
__device__ __forceinline__ uint64_t bucket_order(uint64_t k) {
    return ((k & 0xFFFFFFFFull) << 16) | (k >> 48);
}

// Workgroups are dispatched to the 8 XCDs in turn (workgroup b on XCD b % 8), each XCD
// with its own L2.  Renumber the G workgroups of a launch so that XCD x runs a contiguous
// range of logical indices; with polygons in cell order, each XCD then works on one
// region of the box and finds its neighbours' data in its own L2.
__device__ __forceinline__ int xcd_block(int b, int G) {
    int q = G / 8, r = G % 8, x = b % 8, k = b / 8;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// atomicMax on a maximum that only grows during the launch (zeroed by an earlier launch): the
// value is read first at device scope (past the XCD's L2), and a value not above what was read
// cannot raise it, so its atomic is skipped.  Device atomics on one cache line are carried out
// one after another (about 11 ns apart on MI355X), so one atomic per workgroup on a packing's
// counter queued tens of thousands of them at large packings.
__device__ __forceinline__ void max_if_above(int* a, int v) {
    if (v > __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(a, v);
}

// The polygons of a launch over a work list: list[0..n), n = *count when count is given (in
// FIRE, the polygons of the packings still running: k_poly_cells writes the list), else all N
// (the launch's grid).  Workgroups beyond the ceil(n / per) that the list needs (per = list
// entries per workgroup) get -1 and leave; the others are renumbered over the XCDs as a launch
// of exactly that many.  n is wavefront-uniform.
__device__ __forceinline__ int work_block(const int* count, int per, int& n, int N) {
    int G = gridDim.x;
    n = N;
    if (count) {
        n = __builtin_amdgcn_readfirstlane(*count);
        G = (n + per - 1) / per;
        if ((int)blockIdx.x >= G) return -1;
    }
    return xcd_block(blockIdx.x, G);
}

}  // namespace pcp
