// Internal context of libpcp_hip.so (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "pcp_device.hpp"
#include "../../include/pcp.h"

// every energy flag that recomputes exact crossings and measures closed cycles from the
// pair origin
#define PCP_AREA_MODES (PCP_AREA_ENERGY | PCP_AREA_SQUARED | PCP_PAIR_SQUARED)

namespace pcp {

void set_error(const std::string& msg);
int fail_hip(hipError_t e, const char* what, const char* file, int line);

#define PCP_HIP(call)                                                         \
    do {                                                                      \
        hipError_t _e = (call);                                               \
        if (_e != hipSuccess) return ::pcp::fail_hip(_e, #call, __FILE__, __LINE__); \
    } while (0)
#define PCP_LAUNCHED()                                                        \
    do {                                                                      \
        hipError_t _e = hipGetLastError();                                    \
        if (_e != hipSuccess) return ::pcp::fail_hip(_e, "kernel launch", __FILE__, __LINE__); \
    } while (0)
#define PCP_TRY(call)                                                         \
    do {                                                                      \
        int _s = (call);                                                      \
        if (_s != PCP_OK) return _s;                                          \
    } while (0)

constexpr int BLK = 256;
// The packing's largest record key is kept as PMK_SLOTS partial maxima, each polygon's
// detection updating slot (polygon & (PMK_SLOTS - 1)): a single counter updated by one
// atomic per polygon queues thousands of atomics on one address, and the memory channel
// that holds it then stalls every other access for microseconds.  Slot rows are
// 4 KiB + 256 B apart (and at least P entries long), so the slots fall on different
// channels whether addresses interleave at 256 B or at 4 KiB.
constexpr int PMK_SLOTS = 16;
inline long long pmk_pitch(int P) { return ((long long)(P > 1 ? P : 1) + 511) / 512 * 512 + 32; }
inline int nblk(long long n, int b = BLK) { return (int)((n + b - 1) / b); }

// scalar slots in ctx->dScal
// SC_MAXCOUNT: largest per-edge crossing count (updateNeighbors); SC_BUCKET: largest
// bucket count, set only when a bucket overflowed
// SC_ODD: records of the last pairing whose polygon pair has an odd number of crossings
// (a closed overlap boundary crosses the other polygon's boundary an even number of times;
// an odd count means a crossing was lost or gained in rounding, see k_pair_poly)
// SC_MPNN_DROP: adjacency entries dropped by the last contact-graph build
enum { SC_MAXCOUNT = 0, SC_BUCKET = 1, SC_ODD = 2, SC_MPNN_DROP = 7, SC_NSCAL = 8 };

template <class T> int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    PCP_HIP(hipMalloc((void**)p, count * sizeof(T)));
    return PCP_OK;
}
template <class T> void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace pcp

struct pcp_ctx {
    hipStream_t stream = nullptr;
    int device = 0;      // the device current at pcp_create (the stream's)
    int P = 0, N = 0, V = 0;
    int Kp = 16;         // records per polygon bucket (grown on demand)
    int M = 100;         // reference maxNeighbors for the neighbour-list getters
    long long Ctot = 0;  // total cells over all packings
    bool haveBoxes = false;

    std::vector<int> hPackPoly, hPackVert, hPolyStart, hBox;
    std::vector<long long> hCellOff;

    // topology
    int *dPackPoly = nullptr, *dPackVert = nullptr, *dPolyStart = nullptr, *dShapeId = nullptr;
    int *dNext = nullptr, *dPrev = nullptr, *dPolyPack = nullptr, *dBox = nullptr;
    long long* dCellOff = nullptr;

    // cell list
    int *dCellOf = nullptr, *dCellCount = nullptr, *dCellStart = nullptr;
    int *dCellTmp = nullptr, *dCellSorted = nullptr, *dCellInv = nullptr;
    double2 *dSP = nullptr, *dSR = nullptr;
    int4* dSI = nullptr;

    // detection: polygon bounding circles and their per-packing cell sort
    int seg = 16;                  // lanes per polygon in the segment kernels
    int maxN = 1;                  // largest polygon (vertices)
    bool row16 = false;            // every polygon has exactly 16 vertices (DPP row solves)
    double4* dBounds = nullptr;    // N: centre x, y, radius, big flag
    int* dPcSorted = nullptr;      // N: polygons sorted by cell, per packing
    int* dPcStart = nullptr;       // P x (PC_CELLS + 1): cell starts of each packing
    int* dPackB = nullptr;         // P: polygon grid size | big << 30
    const double2* lastDetectPos = nullptr;   // positions of the last detection (statistics)
    int* dSticky = nullptr;        // [0] a fused evaluation overflowed a buffer
    int* dNum = nullptr;           // V: crossings per edge (updateNeighbors' count)
    // largest record key of each packing, as the maximum of PMK_SLOTS partial maxima (slot k
    // of packing p at k * pmkPitch + p)
    unsigned long long* dPackMaxKey = nullptr;
    long long pmkPitch = 0;
    int* dScal = nullptr;

    // records, per polygon s at rows [s*Kp, s*Kp + count)
    int *dBcnt = nullptr, *dJcnt = nullptr;      // N: si / sj bucket counts
    uint64_t* dBKey = nullptr;     // si bucket as detected (key, float t/u)
    float2* dBTU = nullptr;
    uint64_t* dJKey = nullptr;     // sj bucket keys as detected
    uint64_t *dRKey = nullptr, *dROuter = nullptr;   // si bucket sorted (i, j, sj) + exit record
    float2 *dRTU = nullptr, *dRUT = nullptr;
    double2* dJC = nullptr;        // 4 per sj-list entry: contributions acting on the polygon
    int* dJT = nullptr;            // their vertex (local index, -1 none)
    double2 *dF1 = nullptr, *dLM = nullptr, *dLN = nullptr;   // V: per-vertex partial sums
    double* dAccR = nullptr;       // N: record + interior energy sums
    double* dEr = nullptr;         // per record: its h sums (squared lens energy)
    double* dLaRep = nullptr;      // per record: its lens area if it represents the lens, else 0
    char* dExtStage = nullptr;     // k_exterior_poly's record staging when a bucket outgrows LDS
    size_t extStageBytes = 0;

    // per-vertex / per-polygon scratch
    double* dEpoly = nullptr;
    double* dWork = nullptr;       // 8*V doubles (projection / SHAKE solver)
    double* dRed = nullptr;        // 4*N per-polygon partials
    double* dRedOut = nullptr;     // 4*P per-packing reductions
    int* dScanPart = nullptr;
    int scanPartCap = 0;

    // FIRE (per packing unless noted)
    double *dXbak = nullptr, *dXpre = nullptr, *dVbak = nullptr;   // 2V
    double *dDt = nullptr, *dAlpha = nullptr, *dEpre = nullptr;
    int *dNPos = nullptr, *dSteps = nullptr, *dActive = nullptr;
    int* dFlags = nullptr;         // per packing: see FL_* in pcp_fire.hip
    double* dFireScal = nullptr;   // per packing: bend scale, 1-alpha, maxF, backups
    int* dShakeIt = nullptr;       // per packing: [0] call-1 max, [1] call-2 max
    double* dCom = nullptr;        // 2N scratch (COM of internal geometry passes)
    int* dGuard = nullptr;         // N: polygons taken back by pcp_mpnn_guard (allocated there)
    char* dStatusBlock = nullptr;  // summary (16 bytes) then status[P], read back in one copy
    std::vector<char> hStatusBlock;
    pcp_fire_status* dStatus = nullptr;   // inside dStatusBlock
    int* dSummary = nullptr;       // inside dStatusBlock: [0] any rollback, [1] any unconverged,
                                   // [2] largest record bucket count of the step
    // lanes per polygon of the per-record kernels (pcp_set_schedule / pcp_autotune):
    // k_pair_poly, k_lens_energy (16, 32 or 64) and k_exterior_poly (0 = the widest packing
    // of polygons per wavefront the LDS allows)
    int segPair = 16, segLens = 16, segExt = 0;
    // detection schedule: 0 = culled edge pairs (k_detect_poly), 1 = whole edge tiles of polygon
    // pairs (k_detect_tile, polygons of at most 64 vertices), -1 = by the context's shape
    int detSched = -1;
    // the reference pairing of the last run_records is left to the next run_forces (k_pair_exterior)
    bool pairDeferred = false;
    // FIRE steps of a context large enough to be throughput-bound skip the packings that have
    // converged: k_poly_cells lists the polygons of the running packings in cell order
    // (dActOrder[0..*dActCount)) and the detection and record kernels work on that list.  A
    // converged packing keeps the forces, energies and records of its last evaluation, which
    // were computed at the positions it keeps.
    bool skipDone = false;
    int* dActOrder = nullptr;      // N
    int* dActCount = nullptr;      // 1
    // multi-step FIRE (pcp_fire_run): force backups for device-side rollback, a block-start
    // snapshot (x, v, F; per packing dt, alpha, energy, max|F| and nPos, steps, active), the
    // largest bucket count seen in the block, and the captured graph of one block of steps
    double* dFbak = nullptr;
    // FIRE with XPBD springs: the force before the springs (written by every evaluation of a
    // model with springs) and its backup for device-side rollback
    double *dFext = nullptr, *dFextBak = nullptr;   // 2V
    // grid overlap estimator: per-cell partial sums, then per-packing sums (grown on demand)
    long long* dOvScratch = nullptr;
    long long ovScratchCap = 0;
    double *dSnapX = nullptr, *dSnapV = nullptr, *dSnapF = nullptr, *dSnapD = nullptr;
    int* dSnapI = nullptr;
    int* dBlockOver = nullptr;
    hipGraphExec_t fireGraph = nullptr;
    int graphK = 0, graphKp = 0;
    pcp_fire_params graphPrm{};
    pcp_state graphSt{};
    std::string profName;          // kernel bracketed with events (measurement)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> profEvents;
    size_t profUsed = 0;
};

namespace pcp {
int scan_exclusive(pcp_ctx* c, const int* in, int n, int* out);
int launch_det_reduce(pcp_ctx* c, const double* partials, int nArrays, double* outSums, int nMax,
                      bool sticky = false);
int run_cells(pcp_ctx* c, const double2* pos);
struct PostShake;   // FIRE's post-SHAKE step, done by the detection's bounds kernel (pcp_detect.hip)
int run_detect(pcp_ctx* c, const double2* pos, int cellsFromList, const PostShake* ps = nullptr,
               bool cons = false);
int detect_schedule(const pcp_ctx* c);
// the polygons the detection and record kernels work on: all of them in cell order, or, in a
// FIRE step that skips converged packings, those of the running packings (with their count)
inline const int* work_list(const pcp_ctx* c) { return c->skipDone ? c->dActOrder : c->dPcSorted; }
inline const int* work_count(const pcp_ctx* c) { return c->skipDone ? c->dActCount : nullptr; }
int detect_alloc(pcp_ctx* c);
int detect_stats(pcp_ctx* c, long long* polygonPairs, long long* edgePairs, long long* screened);
int seg_width(int maxN);
int alloc_buckets(pcp_ctx* c, int Kp);
int grow_buckets(pcp_ctx* c, int need);
int packing_counts(pcp_ctx* c, int p, std::vector<int>& cnt);
int run_records(pcp_ctx* c, const double2* pairPos, bool deferOK = false);
int run_forces(pcp_ctx* c, int model, const double2* pos, double* edgeLengths, double* areas,
               const double* targetEdgeV, const double* targetAreas, double stiffness,
               double compressibility, double2* force, double* energy, int geom, int project,
               bool sticky = false, double perimK = 0.0);
int run_xpbd_seg(pcp_ctx* c, double2* pos, const double* targetEdgeV, const double* targetEdgeP,
                 const double* targetAreas, double kE, double kA, double kP, double dt, const double* dtPack,
                 int nIter, double tol, const int* active, double2* vel, int* itOut);
int run_project(pcp_ctx* c, const double2* pos, double2* force);
int run_project_seg(pcp_ctx* c, const double2* pos, double2* force);
int run_shape_project_seg(pcp_ctx* c, double2* pos, const double2* ref, const double* targetAreas);
int run_geometry(pcp_ctx* c, const double2* pos, double* edgeLengths, double* areas, double* comX,
                 double* comY, double* constraints, double* maxEdge);
int run_shake(pcp_ctx* c, double2* pos, const double* targetEdgeV, const double* targetAreas,
              int maxIter, double tol, int retryIter, const int* active, int* itOut);
int run_shake_seg(pcp_ctx* c, double2* pos, const double* targetEdgeV, const double* targetAreas, int maxIter,
                  double tol, int retryIter, const int* active, int* itOut);
int prof_begin(pcp_ctx* c, const char* name, int* slot);
int prof_end(pcp_ctx* c, int slot);
int prof_slot(pcp_ctx* c, const char* name, int* slot);
int autotune(pcp_ctx* c, const pcp_state* st, int reps, double* timesUs);
int scratch_setup(int device);
}  // namespace pcp

// launch `kern` on the context stream; when it is the profiled kernel, the launch carries a
// start and a stop event that the dispatch itself records (hipExtLaunchKernelGGL), so their
// elapsed time is the kernel's own duration, without the dispatch latency an event recorded
// before the launch would add
#define PCP_LAUNCH(c, kern, grid, block, ...) PCP_LAUNCH_LDS(c, kern, grid, block, 0, __VA_ARGS__)
#define PCP_LAUNCH_LDS(c, kern, grid, block, lds, ...)                                          \
    do {                                                                                        \
        int _slot = -1;                                                                         \
        PCP_TRY(::pcp::prof_slot((c), #kern, &_slot));                                          \
        if (_slot >= 0)                                                                         \
            hipExtLaunchKernelGGL(kern, grid, block, lds, (c)->stream, (c)->profEvents[_slot].first, \
                                  (c)->profEvents[_slot].second, 0, __VA_ARGS__);               \
        else                                                                                    \
            hipLaunchKernelGGL(kern, grid, block, lds, (c)->stream, __VA_ARGS__);               \
        PCP_LAUNCHED();                                                                         \
    } while (0)
