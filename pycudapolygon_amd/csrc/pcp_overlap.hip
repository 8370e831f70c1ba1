// Grid estimator of pairwise overlap (updateOverlapAreaCUDA model.cu:466-507,
// updateOverlapAreaKernel kernels.cuh:879-979), part of the pcp_kernels.hip translation unit.
//
// For every point ((ix + 0.5)/D, (iy + 0.5)/D) of a D x D grid the reference counts the
// polygons that contain it among those with an edge midpoint in the 3x3 block of cells
// around the point's cell, stores k(k-1)/2 for k such polygons and sums it.  The
// candidate set depends only on the point's cell, so here one workgroup takes one cell:
//   1. it lists the distinct candidate polygons of the cell's 3x3 block once, in the
//      reference's visiting order (dx, then dy, each cell in ascending vertex order),
//      de-duplicated through a bitmap of the packing's polygons in LDS, and keeps the first
//      513 of them with at least 3 vertices (the reference tests a polygon before it tracks
//      it and stops after 512 tracked ones; polygons of fewer vertices are skipped untracked);
//   2. its lanes take (polygon, point) pairs of the cell's points, polygon-major so that a
//      wavefront reads one polygon's vertices together, decide containment with the
//      reference's winding sum (sum of atan2 of consecutive vertex vectors, inside when
//      |sum - 2 pi| < 1e-12) and count per point in LDS;
//   3. it writes k(k-1)/2 per point and one partial sum per cell; a second kernel adds the
//      partials of each packing in a fixed order (integers: exact).
// The result equals the reference's per point; only the work is organized differently (the
// reference runs a thread per point with a private 512-entry list of visited polygons).

namespace pcp {

constexpr int OV_LIST = 513;          // polygons the reference tests at most per point
constexpr int OV_BITS = 65536;        // polygons per packing (16-bit record ids)
constexpr int OV_PTS = 1024;          // points per counting pass

// block-wide exclusive rank of `flag` in thread order and the block's total
__device__ __forceinline__ int block_rank(bool flag, int* waveTot, int* total) {
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long m = __ballot(flag);
    int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) waveTot[w] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int k = 0; k < BLK / 64; ++k) {
        if (k < w) off += waveTot[k];
        tot += waveTot[k];
    }
    *total = tot;
    return off + below;
}

__device__ __forceinline__ bool winding_inside(const double2* __restrict__ pos, int st, int en, double px,
                                               double py) {
    const double eps = 1e-12, pi = 3.141592653589793238462643383279;
    double angleSum = 0.0;
    double vxPrev = wrapPeriodic(pos[st].x - px);
    double vyPrev = wrapPeriodic(pos[st].y - py);
    double vx0 = vxPrev + 0.0, vy0 = vyPrev + 0.0;
    for (int e = st + 1; e < en; e++) {
        double vx = vxPrev + wrapPeriodic(pos[e].x - pos[e - 1].x);
        double vy = vyPrev + wrapPeriodic(pos[e].y - pos[e - 1].y);
        double cross = vxPrev * vy - vyPrev * vx;
        double dot = vxPrev * vx + vyPrev * vy;
        angleSum += atan2(cross, dot);
        vxPrev = vx;
        vyPrev = vy;
    }
    angleSum += atan2(vxPrev * vy0 - vyPrev * vx0, vxPrev * vx0 + vyPrev * vy0);
    return fabs(angleSum - 2.0 * pi) < eps;
}

// one workgroup per cell of every packing (cell g of the context: packing p with
// cellOff[p] <= g < cellOff[p+1]); counter: [P][D*D] or null; partial: [cells]
__global__ __launch_bounds__(BLK) void k_overlap_cells(int P, int D, const int* __restrict__ box,
                                                      const long long* __restrict__ cellOff,
                                                      const int* __restrict__ packPoly,
                                                      const int* __restrict__ polyStart,
                                                      const int* __restrict__ shapeId,
                                                      const int* __restrict__ cellStart,
                                                      const int* __restrict__ cellSorted,
                                                      const double2* __restrict__ pos, int* __restrict__ counter,
                                                      long long* __restrict__ partial) {
    __shared__ unsigned seen[OV_BITS / 32];
    __shared__ int list[OV_LIST];
    __shared__ int cnt[OV_PTS];
    __shared__ int waveTot[BLK / 64];
    __shared__ long long waveSum[BLK / 64];
    long long g = blockIdx.x;
    int lo = 0, hi = P - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cellOff[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    const int p = lo;
    const int B = box[p];
    const long long off = cellOff[p];
    const int cell = (int)(g - off);
    const int cellX = cell % B, cellY = cell / B;
    const int n0 = packPoly[p], np = packPoly[p + 1] - n0;
    for (int w = threadIdx.x; w < (np + 31) / 32; w += BLK) seen[w] = 0u;
    __syncthreads();

    // 1. distinct candidates of the 3x3 block in the reference's order
    int nList = 0;
    for (int dx = -1; dx <= 1 && nList < OV_LIST; dx++) {
        int nx = cellX + dx;
        if (nx < 0) nx += B; else if (nx >= B) nx -= B;
        for (int dy = -1; dy <= 1 && nList < OV_LIST; dy++) {
            int ny = cellY + dy;
            if (ny < 0) ny += B; else if (ny >= B) ny -= B;
            long long boxId = off + (long long)ny * B + nx;
            int q0 = cellStart[boxId], q1 = cellStart[boxId + 1];
            for (int base = q0; base < q1 && nList < OV_LIST; base += BLK) {
                int q = base + (int)threadIdx.x;
                bool cand = false;
                int gp = -1;
                if (q < q1) {
                    gp = shapeId[cellSorted[q]];
                    // a polygon's vertices in one cell are consecutive (ascending vertex ids),
                    // so its first entry there stands for all of them
                    bool head = (q == q0) || shapeId[cellSorted[q - 1]] != gp;
                    int lp = gp - n0;
                    cand = head && (polyStart[gp + 1] - polyStart[gp] >= 3) &&
                           !((seen[lp >> 5] >> (lp & 31)) & 1u);
                }
                int total;
                int r = block_rank(cand, waveTot, &total);   // synchronizes: seen is read before set
                if (cand) {
                    if (nList + r < OV_LIST) list[nList + r] = gp;
                    int lp = gp - n0;
                    atomicOr(&seen[lp >> 5], 1u << (lp & 31));
                }
                nList = min(nList + total, OV_LIST);
                __syncthreads();
            }
        }
    }

    // 2./3. the cell's points: ix with (int)(((ix + 0.5) / D) * B) == cellX, as the reference
    // assigns points to cells
    long long D2 = (long long)D * D;
    int x0 = max(0, (int)(((long long)cellX * D) / B) - 1), x1 = min(D - 1, (int)(((long long)(cellX + 1) * D) / B) + 1);
    int y0 = max(0, (int)(((long long)cellY * D) / B) - 1), y1 = min(D - 1, (int)(((long long)(cellY + 1) * D) / B) + 1);
    int wx = x1 - x0 + 1, wy = y1 - y0 + 1;
    long long nWin = (x1 >= x0 && y1 >= y0) ? (long long)wx * wy : 0;
    long long mySum = 0;
    for (long long c0 = 0; c0 < nWin; c0 += OV_PTS) {
        int m = (int)min((long long)OV_PTS, nWin - c0);
        for (int k = threadIdx.x; k < m; k += BLK) cnt[k] = 0;
        __syncthreads();
        long long pairs = (long long)m * nList;
        for (long long j = threadIdx.x; j < pairs; j += BLK) {
            int k = (int)(j % m), li = (int)(j / m);
            long long w = c0 + k;
            int ix = x0 + (int)(w % wx), iy = y0 + (int)(w / wx);
            double px = (ix + 0.5) / D, py = (iy + 0.5) / D;
            if ((int)(px * B) != cellX || (int)(py * B) != cellY) continue;
            int gp = list[li];
            if (winding_inside(pos, polyStart[gp], polyStart[gp + 1], px, py)) atomicAdd(&cnt[k], 1);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += BLK) {
            long long w = c0 + k;
            int ix = x0 + (int)(w % wx), iy = y0 + (int)(w / wx);
            double px = (ix + 0.5) / D, py = (iy + 0.5) / D;
            if ((int)(px * B) != cellX || (int)(py * B) != cellY) continue;
            int num = cnt[k];
            int cc = num * (num - 1) / 2;
            if (counter) counter[(long long)p * D2 + (long long)iy * D + ix] = cc;
            mySum += cc;
        }
        __syncthreads();
    }
    for (int o = 32; o >= 1; o >>= 1) mySum += __shfl_xor(mySum, o, 64);
    if ((threadIdx.x & 63) == 0) waveSum[threadIdx.x >> 6] = mySum;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long s = 0;
        for (int k = 0; k < BLK / 64; ++k) s += waveSum[k];
        partial[g] = s;
    }
}

// per packing: the sum of its cells' partials
__global__ __launch_bounds__(BLK) void k_overlap_sum(const long long* __restrict__ cellOff,
                                                    const long long* __restrict__ partial,
                                                    long long* __restrict__ out) {
    __shared__ long long waveSum[BLK / 64];
    int p = blockIdx.x;
    long long s = 0;
    for (long long g = cellOff[p] + threadIdx.x; g < cellOff[p + 1]; g += BLK) s += partial[g];
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) waveSum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int k = 0; k < BLK / 64; ++k) t += waveSum[k];
        out[p] = t;
    }
}

// both kernels on the context's stream; scratch: Ctot partials then P sums; the sums are
// copied to sum host[P] and the stream synchronized
static int overlap_run(pcp_ctx* c, const double* pos, int D, int* counter, long long* sum, long long* scratch) {
    long long* partial = scratch;
    long long* dsum = scratch + c->Ctot;
    PCP_LAUNCH(c, k_overlap_cells, dim3((unsigned)c->Ctot), dim3(BLK), c->P, D, c->dBox, c->dCellOff,
               c->dPackPoly, c->dPolyStart, c->dShapeId, c->dCellStart, c->dCellSorted, (const double2*)pos,
               counter, partial);
    PCP_LAUNCH(c, k_overlap_sum, dim3(c->P), dim3(BLK), c->dCellOff, (const long long*)partial, dsum);
    PCP_HIP(hipMemcpyAsync(sum, dsum, sizeof(long long) * c->P, hipMemcpyDeviceToHost, c->stream));
    PCP_HIP(hipStreamSynchronize(c->stream));
    return PCP_OK;
}

static int overlap_check(pcp_ctx* c, int D, long long* sum) {
    if (!c->haveBoxes) return fail_arg("pcp_overlap_count: set the box sizes first");
    if (D < 1) return fail_arg("pcp_overlap_count: pointDensity must be >= 1");
    if (!sum) return fail_arg("pcp_overlap_count: sum is required");
    for (int p = 0; p < c->P; ++p)
        if (c->hPackPoly[p + 1] - c->hPackPoly[p] > OV_BITS)
            return fail_arg("pcp_overlap_count: more than 65536 polygons in a packing");
    return PCP_OK;
}

// ---- the variant that allocates nothing (pcp_overlap_count_noalloc)
// Its partial sums live in a fixed device array of this library.  In host code `scratch`
// only names the symbol: its device address is looked up once per device with that device
// current (scratch_setup, called when a context is created) and kept per device; every use
// on the host goes through that address.
constexpr long long OV_SCRATCH = (long long)PCP_OVERLAP_NOALLOC_MAX_CELLS + PCP_OVERLAP_NOALLOC_MAX_PACKINGS;
static __device__ __align__(256) long long scratch[OV_SCRATCH];
constexpr int OV_MAX_DEVICES = 64;
static std::mutex g_scratchMutex;
static long long* g_scratchPtr[OV_MAX_DEVICES] = {};

int scratch_setup(int device) {
    if (device < 0 || device >= OV_MAX_DEVICES) {
        set_error("scratch_setup: device index out of range");
        return PCP_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchPtr[device]) return PCP_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    g_scratchPtr[device] = (long long*)ptr;
    return PCP_OK;
}

static int scratch_of(int device, long long** out) {
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (device < 0 || device >= OV_MAX_DEVICES || !g_scratchPtr[device]) {
        set_error("pcp_overlap_count_noalloc: the device's scratch array was not set up");
        return PCP_ERR_ARG;
    }
    *out = g_scratchPtr[device];
    return PCP_OK;
}

}  // namespace pcp

extern "C" int pcp_overlap_count_noalloc(pcp_ctx* c, const double* pos, int D, int* counter, long long* sum) {
    using namespace pcp;
    PCP_TRY(overlap_check(c, D, sum));
    if (c->Ctot > PCP_OVERLAP_NOALLOC_MAX_CELLS || c->P > PCP_OVERLAP_NOALLOC_MAX_PACKINGS) {
        set_error("pcp_overlap_count_noalloc: the context has " + std::to_string(c->Ctot) + " cells in " +
                  std::to_string(c->P) + " packings; at most " + std::to_string(PCP_OVERLAP_NOALLOC_MAX_CELLS) +
                  " cells and " + std::to_string(PCP_OVERLAP_NOALLOC_MAX_PACKINGS) + " packings");
        return PCP_ERR_ARG;
    }
    long long* ptr = nullptr;
    PCP_TRY(scratch_of(c->device, &ptr));
    return overlap_run(c, pos, D, counter, sum, ptr);
}

extern "C" int pcp_overlap_count(pcp_ctx* c, const double* pos, int D, int* counter, long long* sum) {
    using namespace pcp;
    PCP_TRY(overlap_check(c, D, sum));
    long long need = c->Ctot + c->P;
    if (c->ovScratchCap < need) {
        dfree(c->dOvScratch);
        c->ovScratchCap = 0;
        PCP_TRY(dalloc(&c->dOvScratch, (size_t)need));
        c->ovScratchCap = need;
    }
    return overlap_run(c, pos, D, counter, sum, c->dOvScratch);
}
