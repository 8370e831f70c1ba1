/*
 * CPU oracle for the overlap-force-minimization path of camdennis/pyCudaPolygon.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may load it; the product path
 * (pycudapolygon_amd) never does.  Every function restates the reference's
 * algorithm (file:line cited at each definition in pcp_oracle.c) in plain C,
 * compiled with -ffp-contract=off so that each IEEE operation happens exactly
 * as written.
 *
 * Layouts follow the reference: positions/forces are AoS double[2V]
 * ([x0,y0,x1,y1,...]), startIndices int[N+1], one packing per call.
 */
#ifndef PCP_ORACLE_H
#define PCP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference enumTypes.h:6 ordinal order */
enum { ORC_NORMAL = 0, ORC_EDGE_ONLY = 1, ORC_AREA_ONLY = 2, ORC_SOFT_BODY = 3,
       ORC_ABNORMAL = 4, ORC_HYBRID = 5 };
/* model flag: overlap energy measured from one origin per polygon pair (the first
 * vertex of the lower-numbered polygon) instead of each record's entry point, with
 * forces halved -- energy = exact overlap area and force = -grad(energy). */
#define ORC_AREA_ENERGY 0x100
/* squared lens areas (implies ORC_AREA_ENERGY): energy sum over overlap lenses of A^2/2,
 * force -sum A grad A (smooth where a lens closes, unlike the area itself) */
#define ORC_AREA_SQUARED 0x200
/* squared overlap area per polygon pair (implies ORC_AREA_ENERGY): energy sum over
 * polygon pairs of (sum of the pair's lens areas)^2/2, force -A_pair grad A_pair;
 * continuous where two lenses of one pair merge or split, unlike ORC_AREA_SQUARED */
#define ORC_PAIR_SQUARED 0x400
#define ORC_AREA_MODES (ORC_AREA_ENERGY | ORC_AREA_SQUARED | ORC_PAIR_SQUARED)
/* OR-ed into the model: decide each segment test by the sides of the four endpoints (each side
 * a function of one edge and one point, so the two edges at a vertex agree about that vertex:
 * a polygon pair always crosses an even number of times) instead of the reference's parameter
 * test; the crossing parameters are still the reference's */
#define ORC_CONSISTENT 0x800

double orc_wrap(double x);
void orc_topology(int N, const int* start, int V, int* shapeId, int* next, int* prev);
void orc_geometry(int V, int N, const double* pos, const int* start,
                  double* edgeLengths, double* areas, double* comX, double* comY,
                  double* constraints, double* constraintNormSq, double* maxEdgeLength);
void orc_cells(int V, int N, const double* pos, const int* start, int boxSize,
               int* cellLocation, int* neighborIndices, int* countPerBox);
int orc_neighbors(int V, int N, const double* pos, const int* start, int boxSize,
                  const int* cellLocation, const int* neighborIndices, const int* countPerBox,
                  int maxNeighbors, int* neighbors, unsigned char* inside, double* tu,
                  int* numNeighbors);
/* orc_neighbors with the side-consistent decision when consistent != 0 (ORC_CONSISTENT) */
int orc_neighbors_ex(int V, int N, const double* pos, const int* start, int boxSize,
                  const int* cellLocation, const int* neighborIndices, const int* countPerBox,
                  int maxNeighbors, int* neighbors, unsigned char* inside, double* tu,
                  int* numNeighbors, int consistent);
int orc_outersections(int V, int N, const int* start, int maxNeighbors,
                      const int* neighbors, const unsigned char* inside, const double* tuSlots,
                      const int* numNeighbors, uint64_t* intersections, double* tu,
                      uint64_t* outersections, double* ut, int* shapeCounts);
int orc_outersections_ex(int V, int N, const int* start, int maxNeighbors,
                         const int* neighbors, const unsigned char* inside, const double* tuSlots,
                         const int* numNeighbors, uint64_t* intersections, double* tu,
                         uint64_t* outersections, double* ut, int* shapeCounts,
                         const double* pairExactPos);
double orc_force_energy_ex(int model, int V, int N, const int* start, const double* pos,
                           int I, const uint64_t* intersections, const uint64_t* outersections,
                           const double* tu, const double* ut, const double* edgeLengths,
                           const double* areas, const double* targetEdgeV, const double* targetAreas,
                           double stiffness, double compressibility, double perimStiffness,
                           double* force, double* fext);
double orc_force_energy(int model, int V, int N, const int* start, const double* pos,
                        int I, const uint64_t* intersections, const uint64_t* outersections,
                        const double* tu, const double* ut, const double* edgeLengths,
                        const double* areas, const double* targetEdgeV, const double* targetAreas,
                        double stiffness, double compressibility, double* force);
void orc_project_force(int V, int N, const int* start, const double* pos, double* force);
int orc_shake(int V, int N, const int* start, double* pos, const double* targetEdgeV,
              const double* targetAreas, int maxIter, double tol);
/* the same two solves in the HIP kernels' lane order (polygons of at most 64 vertices,
 * -1 otherwise; fewer than 3 are left as they are): bit-identical to the library, equal
 * to the two above to rounding */
int orc_project_force_lanes(int V, int N, const int* start, const double* pos, double* force);
int orc_shake_lanes(int V, int N, const int* start, double* pos, const double* targetEdgeV,
                    const double* targetAreas, int maxIter, double tol);
/* shape matching: rest shape ref (offsets from its vertex mean), scaled to targetAreas
 * (or not, when NULL), at the least-squares rotation and translation (at most 64
 * vertices, -1 otherwise; fewer than 3 are left as they are) */
int orc_shape_project(int V, int N, const int* start, double* pos, const double* ref,
                      const double* targetAreas);
int orc_xpbd(int V, int N, const int* start, double* pos, const double* targetEdgeP,
             const double* targetAreas, int nIter, double tol);
/* compliant XPBD in the kernel's lane order (see pcp_oracle.c) */
int orc_xpbd_lanes(int V, int N, const int* start, double* pos, const double* targetEdgeV,
                   const double* targetEdgeP, const double* targetAreas, double kE, double kA,
                   double kP, double dt, int nIter, double tol);
void orc_update_positions(int V, double* pos, const double* force, double dt);
double orc_max_effective_force(int V, const double* pos, const double* tent, const double* force,
                               double scale);
double orc_det_sum(const double* x, int n);
double orc_max_abs(const double* x, int n);
long long orc_overlap_count(int V, int N, const int* start, const double* pos, int boxSize,
                            const int* neighborIndices, const int* countPerBox,
                            int pointDensity, int* counter);

/* One force evaluation as Model::minimizeFIREStep performs it: geometry, cell list,
 * neighbours, records, force/energy (+projection for the normal model).
 * Workspace is allocated internally.  Returns the energy. */
typedef struct {
    int V, N, boxSize, model, maxNeighbors;
    const int* start;
    const double* targetEdgeV;   /* [V] per-edge target length (edge m -> next(m)) */
    const double* targetAreas;   /* [N] */
    double stiffness, compressibility;
    int solver;                  /* projection/SHAKE: ORC_SOLVER_DENSE or ORC_SOLVER_LANES */
    double perimStiffness;       /* perimeter spring of softBody / hybrid (0 = none) */
} orc_system;
enum { ORC_SOLVER_DENSE = 0, ORC_SOLVER_LANES = 1 };
double orc_force_eval(const orc_system* s, const double* pos, double* force, int* numIntersections);
double orc_force_eval_ex(const orc_system* s, const double* pos, double* force, double* fext,
                         int* numIntersections);

/* FIRE (host contract model.cpp:405-504) */
typedef struct {
    double dt, alpha;
    int nPos;
    double energy;
    int lastShakeIters;
    int shakeFailed, rolledBack;
} orc_fire_state;
void orc_fire_init(const orc_system* s, double* pos, double* vel, double* force, orc_fire_state* st);
void orc_fire_step(const orc_system* s, double* pos, double* vel, double* force, orc_fire_state* st,
                   double dtMax, double alphaStart, double fAlpha, double fInc, double fDec,
                   int nMin, int shakeIter);
int orc_fire_minimize(const orc_system* s, double* pos, double* vel, double* force,
                      orc_fire_state* st, double fTol, int maxSteps, double dtMax,
                      double alphaStart, double fAlpha, double fInc, double fDec, int nMin,
                      int shakeIter);
/* with XPBD integration of the springs (xpbdIter > 0): see pcp_oracle.c */
void orc_fire_step_ex(const orc_system* s, double* pos, double* vel, double* force, orc_fire_state* st,
                      double dtMax, double alphaStart, double fAlpha, double fInc, double fDec,
                      int nMin, int shakeIter, int xpbdIter, double xpbdTol);
int orc_fire_minimize_ex(const orc_system* s, double* pos, double* vel, double* force,
                         orc_fire_state* st, double fTol, int maxSteps, double dtMax,
                         double alphaStart, double fAlpha, double fInc, double fDec, int nMin,
                         int shakeIter, int xpbdIter, double xpbdTol);

#ifdef __cplusplus
}
#endif
#endif
