/*
 * pcp.h -- C ABI of the MI355X (gfx950) overlap-force-minimization library
 * (libpcp_hip.so).  Plain pointers and sizes; no torch or HIP types in the
 * signatures.  Device pointers are marked "device", host pointers "host".
 *
 * Each entry point replaces one or more of the reference's extern "C"
 * launchers (camdennis/pyCudaPolygon src/model.cpp:17-56 <-> src/model.cu);
 * the replaced interface is cited at each declaration.  All entry points
 * return 0 on success and a nonzero status otherwise (text from
 * pcp_last_error()); the reference instead prints and exit()s
 * (src/cuda_check.h:9-38).
 *
 * Batching: one context holds P independent packings, each in its own
 * periodic unit box.  Polygons are numbered globally (packing-major) and
 * vertices globally (polygon-major); record keys use packing-local ids exactly
 * like the reference's single packing (src/kernels.cuh:860-874), so a context
 * with P = 1 is the reference's Model.  Positions and forces are AoS
 * double[2V] ([x0,y0,x1,y1,...], src/model.cpp:522-528).
 *
 * Ordering/threading: all work of a context is queued on that context's own
 * HIP stream; entry points that return host values synchronize it.  A
 * context is not thread-safe (as the reference Model, src/model.hpp:18).
 */
#ifndef PCP_H
#define PCP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCP_OK 0
#define PCP_ERR_HIP 1
#define PCP_ERR_ARG 2
#define PCP_ERR_CAPACITY 3

/* model types, ordinal order of src/enumTypes.h:6 */
enum { PCP_NORMAL = 0, PCP_EDGE_ONLY = 1, PCP_AREA_ONLY = 2, PCP_SOFT_BODY = 3,
       PCP_ABNORMAL = 4, PCP_HYBRID = 5 };
/* Flag OR-ed into a model type (no counterpart in the reference): "area energy".
 * The reference measures each record's chain from its own entry point with float-
 * rounded crossing parameters (kernels.cuh:1117-1120, 857-858), so its energy is
 * overlap area + sum (Qx-Px)(Qy-Py) over lenses and its force is not the gradient of
 * that energy; FIRE's energy test (model.cpp:453) then stalls.  With this flag the
 * crossing parameters are recomputed exactly at the current positions (pairing
 * included), every record of a closed partner cycle measures from one origin per
 * polygon pair, and forces are halved: energy = overlap area, force = -grad(energy). */
#define PCP_AREA_ENERGY 0x100
/* Flag (no counterpart in the reference): squared lens areas.  Implies PCP_AREA_ENERGY's
 * exact crossings and cycle origins; every overlap lens (a closed cycle of exit links, or
 * a record off a cycle on its own) has area A, the energy is sum A^2/2 and the force
 * -sum A grad A.  The area itself has kinks where polygons touch edge to edge, at which
 * no force tolerance is reachable; the squared energy is smooth there and FIRE converges
 * to max|F| <= 1e-10. */
#define PCP_AREA_SQUARED 0x200
/* Flag (no counterpart in the reference): squared overlap area per polygon pair.  As
 * PCP_AREA_SQUARED, but the energy is sum over polygon pairs of A_pair^2/2 with A_pair the
 * sum of the pair's lens areas, and the force -sum A_pair grad A_pair.  Where a vertex
 * crosses an edge, two lenses of one pair merge or split; sum A^2/2 jumps there by
 * A1*A2 and FIRE can stall against the jump, while A_pair is continuous. */
#define PCP_PAIR_SQUARED 0x400
/* Flag (no counterpart in the reference): side-consistent crossings.  The reference decides
 * each edge pair by its parameters t, u in (0, 1) (updateNeighborsKernel kernels.cuh:760-790);
 * where a vertex lies within rounding of another polygon's edge, that test can count the
 * crossing of one of the vertex's two edges and not of the other, and the pair then has an odd
 * number of crossings (pcp_odd_crossing_records), wrong exits and an energy jump that FIRE can
 * stall against.  With this flag an edge pair crosses when each edge's endpoints lie on
 * different sides of the other's line, every side being a function of one edge and one point
 * (so the two edges at a vertex agree about it); |denom| >= 1e-12 as in the reference, and
 * t, u are still the reference's.  Away from such near-ties the crossings are the reference's
 * bit for bit.  Applies to pcp_force_eval, FIRE and pcp_autotune. */
#define PCP_CONSISTENT 0x800

typedef struct pcp_ctx pcp_ctx;

const char* pcp_last_error(void);
int pcp_device_count(int* count);

/* ---- device memory plumbing (replaces the cudaMalloc/cudaMemcpy calls in
 *      src/model.cpp:60-101, 522-628 and the getters :680-889).  Each copy / memset is
 *      ordered after all device work queued before the call (any context) and has
 *      completed when it returns (as the reference's synchronous cudaMemcpy calls). ---- */
int pcp_malloc(void** ptr, size_t bytes);
int pcp_free(void* ptr);
int pcp_memcpy_h2d(void* dst_device, const void* src_host, size_t bytes);
int pcp_memcpy_d2h(void* dst_host, const void* src_device, size_t bytes);
int pcp_memcpy_d2d(void* dst_device, const void* src_device, size_t bytes);
int pcp_memset(void* ptr_device, int value, size_t bytes);

/* ---- context ---- */
int pcp_create(pcp_ctx** ctx);
int pcp_destroy(pcp_ctx* ctx);
/* waits for the context's stream; PCP_ERR_CAPACITY if an asynchronous evaluation since
 * the last call overflowed a crossing buffer (the buffers are enlarged by this call) */
int pcp_synchronize(pcp_ctx* ctx);

/* Topology.  packPolyStart host[P+1] (polygon offsets per packing),
 * polyStart host[N+1] (global vertex offsets per polygon; packing p owns
 * vertices polyStart[packPolyStart[p]] .. polyStart[packPolyStart[p+1]]).
 * Replaces setStartIndices model.cpp:537-598 + updateShapeIdCUDA model.cu:387-391
 * + computeNextPrevCUDA model.cu:77-81 (shapeId/next/prev built on the device). */
int pcp_set_topology(pcp_ctx* ctx, int numPackings, const int* packPolyStart, const int* polyStart);
/* Cells per side for each packing, host[P] (setMaxEdgeLength model.cpp:513-516 +
 * the cell-list allocations of initializeNeighborCells model.cpp:176-229). */
int pcp_set_box_sizes(pcp_ctx* ctx, const int* boxSize);
/* Reference neighbour-slot count M used by the neighbour-list getters
 * (model.hpp:129 maxNeighbors; grown like model.cpp:264-305). */
int pcp_set_max_neighbors(pcp_ctx* ctx, int maxNeighbors);
int pcp_get_max_neighbors(pcp_ctx* ctx, int* maxNeighbors);

/* ---- geometry: updatePolygonGeometryCUDA model.cu:113-163 ----
 * edgeLengths device[V], areas/comX/comY device[N], constraints device[6V] or NULL,
 * maxEdgeLength device[P] or NULL. */
int pcp_update_geometry(pcp_ctx* ctx, const double* pos, double* edgeLengths, double* areas,
                        double* comX, double* comY, double* constraints, double* maxEdgeLength);

/* ---- cell list: updateNeighborCellsCUDA model.cu:308-385 (counting sort by the
 *      cell hash of each edge midpoint; ascending vertex order inside a cell) ---- */
int pcp_update_cells(pcp_ctx* ctx, const double* pos);

/* ---- edge intersections: updateNeighborsCUDA model.cu:393-411 + maxReduceKernel.
 *      Writes per-edge crossing counts and the crossing records.  maxCount host[1]
 *      (optional; synchronizes) receives the largest per-edge count. ---- */
int pcp_update_neighbors(pcp_ctx* ctx, const double* pos, int* maxCount);

/* ---- records: updateValidAndCountsCUDA model.cu:413-450,
 *      updateCompactedIntersectionsCUDA :452-464, sortKeysCUDA :50-75 (64 then 48
 *      bits), applyPermutationCUDA_* :32-48, updateOutersectionsCUDA :509-517,
 *      as orchestrated by Model::updateOutersections model.cpp:351-361.
 *      numIntersections host[P] (optional; synchronizes). ---- */
int pcp_update_records(pcp_ctx* ctx, const double* pairPos, int* numIntersections);
/* pairPos: NULL = the reference's float-rounded pairing; device[2V] positions = exact
 * crossing parameters at those positions (area energy) */

/* ---- forces: Model::updateForceEnergy model.cpp:363-395 (exterior model.cu:519-527,
 *      shape ranges :529-541, interior :543-551, edge :553-559, area :561-567).
 *      targetEdgeV device[V]: target length of edge m -> next(m);
 *      force device[2V] (overwritten), energy device[P]. ---- */
int pcp_update_force_energy(pcp_ctx* ctx, int model, const double* pos, const double* edgeLengths,
                            const double* areas, const double* targetEdgeV,
                            const double* targetAreas, double stiffness, double compressibility,
                            double* force, double* energy);

/* ---- projection: projectForceCUDA model.cu:165-206 (removes from each polygon's
 *      force the components along its n edge-length gradients and area gradient) ---- */
int pcp_project_force(pcp_ctx* ctx, const double* pos, double* force);

/* ---- SHAKE: shakeProjectCUDA model.cu:619-636 (kernel absent from the reference).
 *      iters host[P] (optional; synchronizes): largest iteration count. ---- */
int pcp_shake(pcp_ctx* ctx, double* pos, const double* targetEdgeV, const double* targetAreas,
              int maxIter, double tol, int* iters);

/* ---- XPBD: xpbdProjectCUDA model.cu:209-278 (kernels.cuh:532-623): per iteration the edge
 *      projections of colour 0 then colour 1, then the area correction, with the reference's
 *      arithmetic; tol > 0 stops a polygon once its edge deviations (after the iteration) and
 *      area deviation (before the correction) are below tol.  targetEdgeP device[N] per
 *      polygon.  iters host[P] (optional; synchronizes): per packing the largest iteration
 *      index at which a polygon stopped (nIter when tol <= 0 or one did not), which is the
 *      reference's returned count.  Deviations from the reference: one lane segment per
 *      polygon, so edges of one colour never race (the reference races for odd n; here the
 *      last edge of an odd polygon follows the other even edges) and the area correction
 *      reads consistent positions; convergence is decided per polygon; polygons of at most
 *      64 vertices (PCP_ERR_ARG otherwise), fewer than 3 are left as they are. ---- */
int pcp_xpbd(pcp_ctx* ctx, double* pos, const double* targetEdgeP, const double* targetAreas,
             int nIter, double tol, int* iters);
/* XPBD with compliance (an extension of xpbdProjectCUDA; Macklin, Mueller & Chentanez,
 * MIG 2016): every constraint j keeps a multiplier lambda_j (zero at the call's start) and
 * each projection applies dlambda = -(C_j + at_j lambda_j) / (sum |grad C_j|^2 + at_j),
 * x += grad C_j dlambda, with at_j = 1/(k_j dt^2).  Constraints per polygon: edge lengths
 * (targetEdgeV device[V]), area (targetAreas device[N]) and perimeter (target: the sum of the
 * polygon's edge targets); each type with its stiffness k (0 = off, INFINITY = hard: the
 * reference's projection).  The area and perimeter constraints of a polygon are solved
 * together (their 2x2 block) when both are on.  Same per-polygon stopping rule (residual
 * |C + at lambda|), iteration counts and limits as pcp_xpbd. */
typedef struct {
    double edgeStiffness, areaStiffness, perimStiffness;
    double dt;        /* time step of the compliance 1/(k dt^2); > 0 */
    int nIter;
    double tol;       /* <= 0: run all nIter iterations */
} pcp_xpbd_params;
int pcp_xpbd_ex(pcp_ctx* ctx, double* pos, const double* targetEdgeV, const double* targetAreas,
                const pcp_xpbd_params* prm, int* iters);

/* ---- reductions / small updates ---- */
/* getMaxUnbalancedForceCUDA model.cu:612-614: max |component| per packing, out device[P] */
int pcp_max_abs(pcp_ctx* ctx, const double* x, double* out);
/* updatePositionsCUDA model.cu:569-574: x <- frac(x + dt*F) */
int pcp_update_positions(pcp_ctx* ctx, double* pos, const double* force, double dt);
/* resetAreasCUDA model.cu:639-645 */
int pcp_reset_areas(pcp_ctx* ctx, double* pos, const double* areas, const double* targetAreas,
                    const double* comX, const double* comY);
/* No counterpart in the reference, whose setPhi (pyCudaPolygon.py:223-234) scales the target
 * arrays on the host: targetAreas (device[N]) of packing p multiplied by areaFactor[p] and
 * targetEdgeV (device[V]) by edgeFactor[p] (host[P] each; setPhi's r and sqrt(r)), on the
 * device, so that a compression step moves no per-vertex array over PCIe.  Either target
 * array may be NULL (then its factors are not read). */
int pcp_scale_targets(pcp_ctx* ctx, const double* areaFactor, const double* edgeFactor, double* targetAreas,
                      double* targetEdgeV);
/* Shape matching, the "batched 2x2 SVD shape projection" of the compression protocol (no
 * counterpart in the reference, whose projection acts on forces only): every polygon is
 * replaced by its rest shape refShape (device[2V]: each vertex's offset from the rest
 * shape's vertex mean), scaled to its target area (targetAreas device[N], or NULL for no
 * scaling), turned by the rotation of the polar decomposition of the 2x2 matrix
 * M = sum_i q_i refShape_i^T (q_i: current minimum-image offsets from their vertex mean;
 * the SVD polar factor is [[a, -b], [b, a]] / sqrt(a^2 + b^2), a = M00 + M11,
 * b = M10 - M01) and placed at the current vertex mean: the rigid motion of the rest shape
 * closest to the current vertices in least squares.  Polygons of at most 64 vertices
 * (PCP_ERR_ARG otherwise); polygons of fewer than 3 are left as they are. */
int pcp_shape_project(pcp_ctx* ctx, double* pos, const double* refShape, const double* targetAreas);
/* saveTentativePositionsCUDA + getMaxEffectiveForceCUDA model.cu:280-306: out host[P] */
int pcp_max_effective_force(pcp_ctx* ctx, const double* pos, const double* tentPos,
                            const double* force, double scale, double* out);
/* updateOverlapAreaCUDA model.cu:466-507 (grid estimator, kernels.cuh:879-979): for every point
 * ((ix + 0.5)/D, (iy + 0.5)/D) of a D x D grid (D = pointDensity) in each packing, k = the number
 * of polygons containing it among those with an edge midpoint in the 3x3 block of cells of the
 * last pcp_update_cells around the point's cell (the first 513 such polygons with >= 3 vertices,
 * in the reference's visiting order); counter device[P*D*D] (packing-major, then iy*D + ix;
 * 64-bit offsets) or NULL receives k(k-1)/2, sum host[P] its sum per packing.  At most 65536
 * polygons per packing. */
int pcp_overlap_count(pcp_ctx* ctx, const double* pos, int pointDensity, int* counter, long long* sum);
/* No counterpart in the reference: pcp_overlap_count for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results; the estimator's per-cell partial sums go to a
 * fixed device array of the library (one per GPU, held until the process ends) instead of a
 * buffer the context allocates, so the call allocates nothing.  Limits: the context's cells
 * (the sum of boxSize^2 over its packings) at most PCP_OVERLAP_NOALLOC_MAX_CELLS and its
 * packings at most PCP_OVERLAP_NOALLOC_MAX_PACKINGS, else PCP_ERR_ARG.  Every call on one GPU
 * shares that array: calls that overlap in time on one GPU (other threads, other contexts'
 * streams) race and must be serialized by the caller. */
#define PCP_OVERLAP_NOALLOC_MAX_CELLS (1 << 19)
#define PCP_OVERLAP_NOALLOC_MAX_PACKINGS 4096
int pcp_overlap_count_noalloc(pcp_ctx* ctx, const double* pos, int pointDensity, int* counter, long long* sum);

/* ---- read-back of the context's internal arrays for one packing (the getters
 *      getNeighborCells/getNeighborIndices/getBoxCounts/getNumNeighbors/
 *      getNeighbors/getInsideFlag/getIntersections/getTU/getOutersections/getUT/
 *      getShapeCounts, model.cpp:704-873).  All outputs are host arrays; any may
 *      be NULL.  Vertex/polygon ids are packing-local, as in the reference. ---- */
int pcp_get_cells(pcp_ctx* ctx, int packing, int* cellLocation, int* neighborIndices, int* countPerBox);
int pcp_get_num_intersections(pcp_ctx* ctx, int packing, int* count);
int pcp_get_records(pcp_ctx* ctx, int packing, uint64_t* intersections, double* tu,
                    uint64_t* outersections, double* ut, int* shapeCounts);
/* neighbors host[V_p*M], inside host[V_p*M], numNeighbors host[V_p] in the
 * reference's slot order (unused slots: -1 / 0) */
int pcp_get_neighbor_lists(pcp_ctx* ctx, int packing, int* neighbors, unsigned char* inside,
                           int* numNeighbors);
/* records in compaction order (edge-major, slot order) as left by
 * updateCompactedIntersections before sorting (model.cpp:347-349) */
int pcp_get_compacted_records(pcp_ctx* ctx, int packing, uint64_t* intersections, double* tu);
/* stable sort of n keys by bits [0, endBit) on the device: sortKeysCUDA model.cu:50-75
 * (host in/out: keys host[n], sorted host[n], perm host[n]) */
int pcp_sort_keys(const uint64_t* keys, int n, int endBit, uint64_t* sorted, uint32_t* perm);

/* ---- one force evaluation (the unit of "force evals/s"): geometry, intersections,
 *      records, forces, projection (normal model) -- the sequence of
 *      Model::minimizeFIREStep model.cpp:414-443 without the integrator.  The edge cells
 *      of the user grid are computed inline; the cell-list getters keep reflecting the
 *      last pcp_update_cells.  Asynchronous: a buffer overflow is reported by the next
 *      pcp_synchronize (PCP_ERR_CAPACITY, buffers enlarged; evaluate again). ---- */
typedef struct {
    double* pos;                /* device [2V] */
    double* vel;                /* device [2V] */
    double* force;              /* device [2V] */
    const double* targetEdgeV;  /* device [V]  */
    const double* targetAreas;  /* device [N]  */
    double* edgeLengths;        /* device [V]  */
    double* areas;              /* device [N]  */
    double* energy;             /* device [P]  */
    double stiffness, compressibility;
    int model;
    /* perimeter spring of the softBody and hybrid models (no counterpart in the reference):
     * E += perimStiffness (P - P0)^2 / 2 per polygon, P0 = the sum of its edge targets;
     * 0 = none */
    double perimStiffness;
} pcp_state;
int pcp_force_eval(pcp_ctx* ctx, const pcp_state* st);

/* ---- FIRE: Model::minimizeFIREStep / Model::minimizeFIRE model.cpp:405-504 and the
 *      launchers model.cu:576-608 (kernels absent from the reference).  State per
 *      packing (dt, alpha, nPos, energy, steps, convergence) stays on the device. ---- */
typedef struct {
    double dtMax, alphaStart, fAlpha, fInc, fDec;
    int nMin, shakeIter;
    double fTol;        /* convergence: max |force component| <= fTol */
    /* xpbdIter > 0 (no counterpart in the reference, whose XPBD launcher is never called):
     * the springs of a model that has them (edge: stiffness, area: compressibility,
     * perimeter: perimStiffness) are integrated implicitly by XPBD (pcp_xpbd_ex with the
     * packing's dt, xpbdIter iterations, tolerance xpbdTol) and the rest of the force (the
     * overlap force) explicitly, XPBD's own time step inside FIRE:
     *   v' = v + dt Fext;  x~ = frac(x + dt v');  x = XPBD(x~), wrapped;  v = v' + wrap(x - x~)/dt
     * then the energy test and mixing on the total energy and force (no second half-kick).
     * status.lastShakeIters reports the step's XPBD iteration count.  0: velocity Verlet. */
    int xpbdIter;
    double xpbdTol;
} pcp_fire_params;
typedef struct {
    double dt, alpha, energy, maxForce;
    int nPos, steps, converged, lastShakeIters, shakeFailed, rolledBack;
} pcp_fire_status;
/* resetVelocities + fresh force evaluation (minimizeFIRE prologue model.cpp:489-496);
 * dt0 host[P] or NULL (all = dt0s), alpha = alphaStart */
int pcp_fire_init(pcp_ctx* ctx, const pcp_state* st, const pcp_fire_params* prm, double dt0);
/* one FIRE step for every not-yet-converged packing; status host[P] (optional) */
int pcp_fire_step(pcp_ctx* ctx, const pcp_state* st, const pcp_fire_params* prm, pcp_fire_status* status);
/* FIRE until every packing has max|F| <= fTol or maxSteps steps; status host[P] */
int pcp_fire_minimize(pcp_ctx* ctx, const pcp_state* st, const pcp_fire_params* prm, int maxSteps,
                      pcp_fire_status* status);
/* FIRE until every packing has max|F| <= fTol or maxSteps steps, in blocks of stepsPerBlock
 * steps queued without host round trips (the reference's minimizeFIRE loop, model.cpp:488-504,
 * reads back after every step).  A rolled-back packing takes the step's start x, F and energy
 * from the step's backups instead of a fresh evaluation (model.cpp:454-464; the same values,
 * evaluations being deterministic).  A block in which a record bucket overflowed is undone
 * from a block-start snapshot and run again with larger buckets.  useGraph != 0: the launches
 * of one block are captured once in a HIP graph and replayed (captured again when the block
 * length, bucket capacity, parameters or state buffers change).  status host[P] (optional). */
int pcp_fire_run(pcp_ctx* ctx, const pcp_state* st, const pcp_fire_params* prm, int maxSteps, int stepsPerBlock,
                 int useGraph, pcp_fire_status* status);
/* set the per-packing FIRE scalars (host[P] arrays, NULL = leave) and mark every packing
 * active with a zero step count; read them back */
int pcp_fire_set(pcp_ctx* ctx, const double* dt, const double* alpha, const int* nPos);
int pcp_fire_get(pcp_ctx* ctx, pcp_fire_status* status);

/* ---- measurement support (no counterpart in the reference) ----
 * pcp_profile_kernel: bracket every later launch of the kernel called `name` (e.g.
 * "k_detect_poly<0>") with HIP events on the context's stream; NULL stops.  pcp_profile_read:
 * total milliseconds and launch count since the last pcp_profile_kernel call.
 * pcp_stream_copy: plain device copy kernel of n doubles (HBM counter calibration). */
int pcp_profile_kernel(pcp_ctx* ctx, const char* name);
int pcp_profile_read(pcp_ctx* ctx, double* totalMs, int* launches);
int pcp_stream_copy(pcp_ctx* ctx, double* dst, const double* src, long long n);
/* Schedules of the per-record kernels (no counterpart in the reference): lanes per polygon
 * of the pairing (k_pair_poly), lens-area (k_lens_energy) and exterior-force (k_exterior_poly)
 * kernels -- 16, 32 or 64, i.e. 4, 2 or 1 polygons per wavefront; exterior 0 = as many
 * polygons per wavefront as the LDS allows.  Every schedule gives identical results (each
 * lane keeps its sums in order); only the speed differs. */
int pcp_set_schedule(pcp_ctx* ctx, int pairLanes, int lensLanes, int exteriorLanes);
int pcp_get_schedule(pcp_ctx* ctx, int* pairLanes, int* lensLanes, int* exteriorLanes);
/* Detection schedule (no counterpart in the reference): 0 = one wavefront per polygon A lists,
 * for each candidate polygon, the edges of each near the other's bounding circle and tests the
 * products of the lists, candidates batched so that lanes stay busy; 1 = one wavefront per
 * polygon A tests whole edge tiles, one candidate polygon per segment of lanes (polygons of at
 * most 64 vertices); -1 = by the context's shape (tiles for 4 or more packings of polygons of
 * at most 16 vertices).  Both give identical crossings; only the speed differs. */
int pcp_set_detect_schedule(pcp_ctx* ctx, int mode);
int pcp_get_detect_schedule(pcp_ctx* ctx, int* mode);
/* autotuner: one evaluation of st fills the context's buffers, then every candidate width of
 * each kernel, and each detection schedule, is timed over `reps` back-to-back launches and the
 * fastest is kept.  timesUs host[11] (optional): microseconds per launch of pair 16/32/64,
 * lens 16/32/64, exterior 16/32/64, detection 0/1 (0 = not a candidate). */
int pcp_autotune(pcp_ctx* ctx, const pcp_state* st, int reps, double* timesUs);
/* work of the last detection under the schedule in effect: polygon pairs tested (each polygon
 * with itself, each pair of distinct polygons whose bounding circles meet once), the edge
 * pairs whose segment test ran, and (screened, optional) the edge pairs considered.  Culled
 * enumeration: n(n-3)/2 per polygon with itself; per pair of distinct polygons the product
 * of the counts of edges of each near the other's circle, or nA*nB where the cull does not
 * apply (screened = edgePairs).  Whole tiles: every non-adjacent pair of a polygon and every
 * nA*nB pair are screened, those whose FP32 edge boxes meet (all of them where a polygon is
 * not cullable) take the segment test.  The last detection's position buffer must still be
 * valid. */
int pcp_detect_stats(pcp_ctx* ctx, long long* polygonPairs, long long* edgePairs, long long* screened);
/* records of the last pairing (pcp_force_eval, pcp_update_outersections) that belong to a
 * pair of distinct polygons with an odd number of crossings.  The boundary of every overlap
 * crosses the other polygon's boundary an even number of times; an odd count arises where a
 * vertex lies within rounding of the other polygon's edge and the segment test
 * (updateNeighborsKernel kernels.cuh:760-790, kept bit for bit) counts the crossing of one of
 * the vertex's two edges and not of the other.  Such a pair's records are then paired with
 * wrong exits and its overlap energy jumps.  A diagnostic with no counterpart in the
 * reference (which has the same window): 0 in a healthy state. */
int pcp_odd_crossing_records(pcp_ctx* ctx, int* count);

/* ---- learned warm start (no counterpart in the reference, which has no learned
 *      component: an addition in front of the plain FIRE method) ----
 * A message-passing network over the contact graph of the packings of a context.
 * Nodes are vertices; vertex m is linked to prev(m), next(m) and to the start vertex of
 * every edge crossing edge m (from the crossing records of the last evaluation).
 *   h0 = relu(x W0 + b0)                                          x: 4 features / vertex
 *   h_l = h_{l-1} + relu([h_l-1, sum over neighbours of h_l-1] W_l + b_l),  l = 1..3
 *   y = h3 Wo + bo                                                2 outputs / vertex
 * Parameters: float32 vector of PCP_MPNN_NPARAM entries (offsets below, matrices
 * row-major [in][out]).  Products take bf16-rounded operands and accumulate in float32
 * (the 128x64 layers on the matrix cores). */
#define PCP_MPNN_H 64
#define PCP_MPNN_DEG 16
#define PCP_MPNN_OFF_W0 0          /* W0[4][64]   */
#define PCP_MPNN_OFF_B0 256        /* b0[64]      */
#define PCP_MPNN_EMBED_NP 320
#define PCP_MPNN_OFF_W1 320        /* layer l: W_l[128][64] then b_l[64] at OFF_W1 + (l-1)*LAYER_NP */
#define PCP_MPNN_LAYER_NP 8256
#define PCP_MPNN_OFF_WO 25088      /* Wo[64][2]   */
#define PCP_MPNN_OFF_BO 25216      /* bo[2]       */
#define PCP_MPNN_OUT_NP 130
#define PCP_MPNN_NPARAM 25218
/* floats of the forward work buffer (activations kept for the backward pass) and of the
 * backward work buffer, for V vertices */
int pcp_mpnn_sizes(int V, long long* workFloats, long long* gradWorkFloats);
/* contact graph of the context's last evaluation: adj device[V*PCP_MPNN_DEG] (global
 * vertex ids: prev, next, crossing partners ascending, -1 empty); dropped host[1]
 * (optional; synchronizes): partners beyond the width */
int pcp_mpnn_graph(pcp_ctx* ctx, int* adj, int* dropped);
/* input features of every vertex after an evaluation of st (force, per-polygon energy,
 * crossings counted in adj): feat device[V*4] =
 *   asinh(Fx/(ell^(e-1) s0)), asinh(Fy/(ell^(e-1) s1)), crossings/s2, asinh(E_poly/(n ell^e s3))
 * ell device[P] = each packing's length unit, e = energyPow, s = inScale host[4] */
int pcp_mpnn_features(pcp_ctx* ctx, const pcp_state* st, const int* adj, const double* ell, int energyPow,
                      const float* inScale, float* feat);
/* forward pass over V vertices: params/feat/adj/work/out device; out[V*2] */
int pcp_mpnn_forward(pcp_ctx* ctx, int V, const float* params, const float* feat, const int* adj, float* work,
                     float* out);
/* mean squared error L = (1/V) sum |out - target|^2 and dOut = dL/dout (device[V*2]);
 * loss host[1] optional (synchronizes) */
int pcp_mpnn_loss(pcp_ctx* ctx, int V, const float* out, const float* target, float* dOut, float* gradWork,
                  double* loss);
/* backward pass of the last forward (same work buffer): grad device[PCP_MPNN_NPARAM] =
 * dL/dparams for the output gradient dOut (overwritten, not accumulated) */
int pcp_mpnn_backward(pcp_ctx* ctx, int V, const float* params, const float* feat, const int* adj,
                      const float* work, const float* dOut, float* gradWork, float* grad);
/* one Adam step (bias-corrected, step counts from 1) on params with moments m, v */
int pcp_mpnn_adam(pcp_ctx* ctx, float* params, const float* grad, float* m, float* v, int step, float lr,
                  float beta1, float beta2, float eps);
/* training targets: target device[V*2] = minimum image (xTo - xFrom) / (ell * outScale) */
int pcp_mpnn_targets(pcp_ctx* ctx, const double* xFrom, const double* xTo, const double* ell, double outScale,
                     float* target);
/* warm start: pos <- frac(pos + ell * outScale * out) for the packings with active != 0
 * (active device[P] or NULL = all) */
int pcp_mpnn_apply(pcp_ctx* ctx, double* pos, const float* out, const double* ell, double outScale,
                   const int* active);
/* guard of the warm start: an edge-crossing energy cannot see a polygon that lies inside
 * another.  Where the bounding circle of one polygon (of pos) lies inside another's,
 * |c_A - c_B| + min(r_A, r_B) < kappa * max(r_A, r_B), both polygons take back their
 * positions xPrev (device[2V], e.g. pos before pcp_mpnn_apply); a packing that still holds
 * such a pair afterwards takes back all of its positions.  flags device[P] (out): 0 kept,
 * 1 polygons taken back, 2 the packing taken back.  nTaken host[1] (optional; synchronizes):
 * the number of packings with flags != 0.  kappa in [0, 2], else PCP_ERR_ARG.  Resets the
 * context's records (a move makes them stale). */
int pcp_mpnn_guard(pcp_ctx* ctx, double* pos, const double* xPrev, double kappa, int* flags, int* nTaken);

#ifdef __cplusplus
}
#endif
#endif
