// Model over the C ABI (include/pcp.h).  Method-by-method counterpart of
// camdennis/pyCudaPolygon src/model.cpp; each method cites the reference method whose
// observable behaviour it keeps.  Errors raise std::runtime_error (pybind11 turns them
// into Python RuntimeError) where the reference prints and exit()s (cuda_check.h:9-38).
#include "model.hpp"
#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>

using std::vector;

void Model::check(int status) const {
    if (status != PCP_OK) throw std::runtime_error(std::string("pyCudaPolygon (HIP): ") + pcp_last_error());
}

template <class T> vector<T> Model::d2h(const T* d, size_t n) const {
    vector<T> h(n);
    if (n && d) {
        check(pcp_synchronize(ctx));
        check(pcp_memcpy_d2h(h.data(), d, n * sizeof(T)));
    }
    return h;
}

template <class T> static void dalloc(T** p, size_t n) {
    void* q = nullptr;
    if (pcp_malloc(&q, (n ? n : 1) * sizeof(T)) != PCP_OK)
        throw std::runtime_error(std::string("pyCudaPolygon (HIP): ") + pcp_last_error());
    *p = (T*)q;
}
template <class T> static void dfree(T*& p) {
    if (p) pcp_free(p);
    p = nullptr;
}

// Model::Model model.cpp:60-101
Model::Model(int size_) : size(size_) {
    check(pcp_create(&ctx));
    dalloc(&energy, 1);
    dalloc(&maxEdgeLength, 1);
    check(pcp_memset(energy, 0, sizeof(double)));
    check(pcp_memset(maxEdgeLength, 0, sizeof(double)));
    setNumVertices(size_);
}

Model::~Model() {
    freeVertexBuffers();
    freePolygonBuffers();
    dfree(energy);
    dfree(maxEdgeLength);
    dfree(intersectionsCounter);
    if (ctx) pcp_destroy(ctx);
}

void Model::freeVertexBuffers() {
    dfree(positions); dfree(force); dfree(velocities); dfree(positionsTMP);
    dfree(edgeLengths); dfree(constraints); dfree(targetEdgeV); dfree(scratch);
}
void Model::freePolygonBuffers() {
    dfree(areas); dfree(comX); dfree(comY); dfree(targetAreas);
}

void Model::ensureTopology() const {
    if (!topoDirty) return;
    if (hStart.empty()) throw std::runtime_error("pyCudaPolygon: setStartIndices before using the polygons");
    if (hStart.back() != size)
        throw std::runtime_error("pyCudaPolygon: startIndices do not match the number of vertices");
    int pack[2] = {0, numPolygons};
    check(pcp_set_topology(ctx, 1, pack, hStart.data()));
    check(pcp_set_max_neighbors(ctx, maxNeighbors));
    topoDirty = false;
    boxDirty = true;
}
void Model::ensureCells() const {
    ensureTopology();
    if (!boxDirty) return;
    check(pcp_set_box_sizes(ctx, &boxSize));
    boxDirty = false;
}

pcp_state Model::state() const {
    pcp_state s;
    s.pos = positions; s.vel = velocities; s.force = force;
    s.targetEdgeV = targetEdgeV; s.targetAreas = targetAreas;
    s.edgeLengths = edgeLengths; s.areas = areas; s.energy = energy;
    s.stiffness = stiffness; s.compressibility = compressibility;
    s.model = (int)simControl.modelType | energyFlags;
    s.perimStiffness = 0.0;
    return s;
}

// initializeNeighborCells model.cpp:176-229 (buffers for boxSize^2 cells, shapeId, next/prev)
void Model::initializeNeighborCells() {
    topoDirty = true;
    ensureCells();
    numIntersections = 0;
}
void Model::initializeRandomSeed(unsigned long long seed_) { seed = seed_; }

// sortKeys model.cpp:238-240 / sortKeysCUDA model.cu:50-75: stable sort of the current
// intersections by bits [0, endBit); the permutation is kept for getKeys
void Model::sortKeys(int endBit) {
    vector<uint64_t> cur = getIntersections();
    vector<uint64_t> sorted(cur.size());
    keys.assign(cur.size(), 0);
    if (!cur.empty()) check(pcp_sort_keys(cur.data(), (int)cur.size(), endBit, sorted.data(), keys.data()));
    vector<double> tu = getTU();
    viewKeys = sorted;
    viewTU = tu;   // tu is not permuted by sortKeys (model.cpp:238-240 sorts keys only)
    keysView = true;
}
vector<uint32_t> Model::getKeys() const { return keys; }

// setNumVertices model.cpp:600-628
void Model::setNumVertices(int n) {
    size = n;
    freeVertexBuffers();
    size_t m = (size_t)std::max(n, 1);
    dalloc(&positions, 2 * m); dalloc(&force, 2 * m); dalloc(&velocities, 2 * m);
    dalloc(&positionsTMP, 2 * m); dalloc(&edgeLengths, m); dalloc(&constraints, 6 * m);
    dalloc(&targetEdgeV, m); dalloc(&scratch, 2 * m);
    check(pcp_memset(force, 0, 16 * (size_t)std::max(n, 1)));
    check(pcp_memset(velocities, 0, 16 * (size_t)std::max(n, 1)));
    check(pcp_memset(targetEdgeV, 0, 8 * (size_t)std::max(n, 1)));
    topoDirty = true;
}

// setPositions model.cpp:522-528 (size mismatch: message and return)
void Model::setPositions(const vector<double>& p) {
    if ((long long)p.size() != 2LL * size) {
        std::cout << "Update numVertices before setting the positions" << std::endl;
        return;
    }
    check(pcp_synchronize(ctx));
    check(pcp_memcpy_h2d(positions, p.data(), p.size() * sizeof(double)));
}
// setForces model.cpp:530-535
void Model::setForces(const vector<double>& f) {
    check(pcp_synchronize(ctx));
    size_t n = std::min(f.size(), 2 * (size_t)size);
    check(pcp_memcpy_h2d(force, f.data(), n * sizeof(double)));
}

// setStartIndices model.cpp:537-598
void Model::setStartIndices(const vector<int>& s) {
    if (s.empty()) throw std::invalid_argument("startIndices must have at least one entry");
    numPolygons = (int)s.size() - 1;
    hStart = s;
    freePolygonBuffers();
    size_t m = (size_t)std::max(numPolygons, 1);
    dalloc(&areas, m); dalloc(&comX, m); dalloc(&comY, m);
    dalloc(&targetAreas, m);
    check(pcp_memset(targetAreas, 0, 8 * (size_t)std::max(numPolygons, 1)));
    check(pcp_memset(areas, 0, 8 * (size_t)std::max(numPolygons, 1)));
    hTargetEdgeP.assign(numPolygons, 0.0);
    topoDirty = true;
}

void Model::setModelEnum(simControlStruct::modelEnum m) { simControl.modelType = m; }
void Model::setOverlapEnergy(const std::string& name) {
    if (name == "reference") energyFlags = 0;
    else if (name == "area") energyFlags = PCP_AREA_ENERGY;
    else if (name == "squaredLens") energyFlags = PCP_AREA_SQUARED;
    else if (name == "squaredPair") energyFlags = PCP_PAIR_SQUARED;
    else throw std::invalid_argument("setOverlapEnergy: expected reference, area, squaredLens or squaredPair, got " + name);
}
std::string Model::getOverlapEnergy() const {
    switch (energyFlags) {
        case PCP_AREA_ENERGY: return "area";
        case PCP_AREA_SQUARED: return "squaredLens";
        case PCP_PAIR_SQUARED: return "squaredPair";
        default: return "reference";
    }
}

// setMaxEdgeLength model.cpp:513-516: boxSize = floor(1/L); the device copy of L is
// overwritten by updatePolygonGeometry, as in the reference
void Model::setMaxEdgeLength(double L) {
    double b = std::floor(1.0 / L);
    if (!(b >= 1.0)) b = 1.0;
    if (b > 46340.0) b = 46340.0;
    boxSize = (int)b;
    boxDirty = true;
    check(pcp_synchronize(ctx));
    check(pcp_memcpy_h2d(maxEdgeLength, &L, sizeof(double)));
}

// setTargetEdgeLengths model.cpp:630-632 (one target per polygon, stored per edge)
void Model::setTargetEdgeLengths(const vector<double>& t) {
    if ((int)t.size() < numPolygons) throw std::invalid_argument("need one target edge length per polygon");
    hTargetEdgeP.assign(t.begin(), t.begin() + numPolygons);
    vector<double> perEdge(size, 0.0);
    for (int s = 0; s < numPolygons; ++s)
        for (int v = hStart[s]; v < hStart[s + 1] && v < size; ++v) perEdge[v] = t[s];
    check(pcp_synchronize(ctx));
    check(pcp_memcpy_h2d(targetEdgeV, perEdge.data(), sizeof(double) * size));
}
void Model::setTargetEdgeLengthsPerEdge(const vector<double>& t) {
    if ((int)t.size() != size) throw std::invalid_argument("need one target length per edge");
    for (int s = 0; s < numPolygons; ++s) {
        double acc = 0.0;
        for (int v = hStart[s]; v < hStart[s + 1]; ++v) acc += t[v];
        hTargetEdgeP[s] = acc / std::max(1, hStart[s + 1] - hStart[s]);
    }
    check(pcp_synchronize(ctx));
    check(pcp_memcpy_h2d(targetEdgeV, t.data(), sizeof(double) * size));
}
vector<double> Model::getTargetEdgeLengthsPerEdge() const { return d2h(targetEdgeV, size); }
// setTargetAreas model.cpp:634-636
void Model::setTargetAreas(const vector<double>& t) {
    if ((int)t.size() < numPolygons) throw std::invalid_argument("need one target area per polygon");
    check(pcp_synchronize(ctx));
    check(pcp_memcpy_h2d(targetAreas, t.data(), sizeof(double) * numPolygons));
}
void Model::setStiffness(double s) { stiffness = s; }
void Model::setCompressibility(double c) { compressibility = c; }
double Model::getStiffness() const { return stiffness; }
double Model::getCompressibility() const { return compressibility; }
int Model::getMaxNeighbors() const { return maxNeighbors; }
int Model::getBoxSize() const { return boxSize; }

// getters model.cpp:656-893
std::string Model::getModelEnum() const {
    switch (simControl.modelType) {
        case simControlStruct::modelEnum::normal: return "normal";
        case simControlStruct::modelEnum::edgeOnly: return "edgeOnly";
        case simControlStruct::modelEnum::areaOnly: return "areaOnly";
        case simControlStruct::modelEnum::softBody: return "softBody";
        case simControlStruct::modelEnum::hybrid: return "hybrid";
        case simControlStruct::modelEnum::abnormal: return "abnormal";
    }
    return std::to_string((int)simControl.modelType);
}
int Model::getNumVertices() const { return size; }
int Model::getNumPolygons() const { return numPolygons; }
unsigned long long Model::getRandomSeed() const { return seed; }
double Model::getEnergy() const { return d2h(energy, 1)[0]; }
vector<int> Model::getShapeId() const {
    vector<int> r(size);
    for (int s = 0; s < numPolygons; ++s)
        for (int v = hStart[s]; v < hStart[s + 1] && v < size; ++v) r[v] = s;
    return r;
}
vector<double> Model::getPositions() const { return d2h(positions, 2 * (size_t)size); }
vector<double> Model::getForces() const { return d2h(force, 2 * (size_t)size); }
vector<int> Model::getIntersectionsCounter() const {
    if (pointDensity <= 0) return {};
    return d2h(intersectionsCounter, (size_t)pointDensity * pointDensity);
}
double Model::getMaxEdgeLength() const { return d2h(maxEdgeLength, 1)[0]; }
vector<int> Model::getNeighbors() const {
    ensureCells();
    vector<int> n((size_t)maxNeighbors * size, -1);
    vector<unsigned char> in((size_t)maxNeighbors * size);
    check(pcp_get_neighbor_lists(ctx, 0, n.data(), in.data(), nullptr));
    return n;
}
vector<bool> Model::getInsideFlag() const {
    ensureCells();
    vector<unsigned char> in((size_t)maxNeighbors * size);
    check(pcp_get_neighbor_lists(ctx, 0, nullptr, in.data(), nullptr));
    vector<bool> r(in.size());
    for (size_t i = 0; i < in.size(); ++i) r[i] = in[i] != 0;
    return r;
}
vector<int> Model::getNumNeighbors() const {
    ensureCells();
    vector<int> n(size);
    check(pcp_get_neighbor_lists(ctx, 0, nullptr, nullptr, n.data()));
    return n;
}
vector<double> Model::getTU() const {
    if (keysView || compactedView) return viewTU;
    if (numIntersections == 0) return {};
    vector<double> tu(2 * (size_t)numIntersections);
    check(pcp_get_records(ctx, 0, nullptr, tu.data(), nullptr, nullptr, nullptr));
    return tu;
}
vector<double> Model::getUT() const {
    if (numIntersections == 0) return {};
    vector<double> ut(2 * (size_t)numIntersections);
    check(pcp_get_records(ctx, 0, nullptr, nullptr, nullptr, ut.data(), nullptr));
    return ut;
}
vector<uint64_t> Model::getIntersections() const {
    if (keysView || compactedView) return viewKeys;
    vector<uint64_t> r(numIntersections);
    if (numIntersections) check(pcp_get_records(ctx, 0, r.data(), nullptr, nullptr, nullptr, nullptr));
    return r;
}
int Model::getNumIntersections() const { return numIntersections; }
vector<uint64_t> Model::getOutersections() const {
    vector<uint64_t> r(numIntersections);
    if (numIntersections) check(pcp_get_records(ctx, 0, nullptr, nullptr, r.data(), nullptr, nullptr));
    return r;
}
vector<int> Model::getShapeCounts() const {
    vector<int> r(numPolygons, 0);
    if (!topoDirty) check(pcp_get_records(ctx, 0, nullptr, nullptr, nullptr, nullptr, r.data()));
    return r;
}
vector<int> Model::getStartIndices() const { return hStart; }
vector<double> Model::getAreas() const { return d2h(areas, numPolygons); }
vector<double> Model::getTargetAreas() const { return d2h(targetAreas, numPolygons); }
vector<double> Model::getTargetEdgeLengths() const { return hTargetEdgeP; }
vector<double> Model::getConstraints() const { return d2h(constraints, 6 * (size_t)size); }
vector<double> Model::getCOM() const {
    vector<double> x = d2h(comX, numPolygons), y = d2h(comY, numPolygons);
    x.insert(x.end(), y.begin(), y.end());
    return x;
}
// the reference allocates perimeters but never writes them (model.cpp:556); here they
// are the sums of the current edge lengths
vector<double> Model::getPerimeters() const {
    vector<double> el = getEdgeLengths(), r(numPolygons, 0.0);
    for (int s = 0; s < numPolygons; ++s)
        for (int v = hStart[s]; v < hStart[s + 1]; ++v) r[s] += el[v];
    return r;
}
vector<int> Model::getNeighborCells() const {
    ensureCells();
    vector<int> r(size);
    check(pcp_get_cells(ctx, 0, r.data(), nullptr, nullptr));
    return r;
}
vector<int> Model::getNeighborIndices() const {
    ensureCells();
    vector<int> r(size);
    check(pcp_get_cells(ctx, 0, nullptr, r.data(), nullptr));
    return r;
}
vector<int> Model::getBoxCounts() const {
    ensureCells();
    vector<int> r((size_t)boxSize * boxSize);
    check(pcp_get_cells(ctx, 0, nullptr, nullptr, r.data()));
    return r;
}
vector<double> Model::getEdgeLengths() const { return d2h(edgeLengths, size); }
double Model::getOverlapArea() const { return overlapArea; }
// getMaxUnbalancedForceCUDA model.cu:612-614
double Model::getMaxUnbalancedForce() const {
    ensureTopology();
    check(pcp_max_abs(ctx, force, scratch));
    return d2h(scratch, 1)[0];
}
void Model::resetMaxActualNeighbors() {}

// updatePolygonGeometry model.cpp:313-315
void Model::updatePolygonGeometry() {
    ensureTopology();
    check(pcp_update_geometry(ctx, positions, edgeLengths, areas, comX, comY, constraints, maxEdgeLength));
}
// projectForce model.cpp:337-344
void Model::projectForce() {
    ensureTopology();
    check(pcp_project_force(ctx, positions, force));
}
// shakeProject model.cpp:317-322
int Model::shakeProject(int nIter, double tol) {
    ensureTopology();
    int it = 0;
    check(pcp_shake(ctx, positions, targetEdgeV, targetAreas, nIter, tol, &it));
    lastShakeIters = it;
    return it;
}
// xpbdProjectCUDA model.cu:209-278 (launcher present but unwired in the reference)
int Model::xpbdProject(int nIter, double tol) {
    ensureTopology();
    double* tp = nullptr;
    dalloc(&tp, numPolygons);
    check(pcp_memcpy_h2d(tp, hTargetEdgeP.data(), sizeof(double) * numPolygons));
    int it = 0;
    int st = pcp_xpbd(ctx, positions, tp, targetAreas, nIter, tol, &it);
    dfree(tp);
    check(st);
    return it;
}
int Model::getLastShakeIters() const { return lastShakeIters; }
// saveTentativePositions model.cpp:328-330
void Model::saveTentativePositions() {
    check(pcp_synchronize(ctx));
    check(pcp_memcpy_d2d(positionsTMP, positions, 16 * (size_t)size));
}
// getMaxEffectiveForce model.cpp:332-335
double Model::getMaxEffectiveForce(double dt, minimizerEnum type) const {
    ensureTopology();
    double scale = (type == minimizerEnum::GD) ? (1.0 / dt) : (2.0 / (dt * dt));
    double out = 0.0;
    check(pcp_max_effective_force(ctx, positions, positionsTMP, force, scale, &out));
    return out;
}
// updateNeighborCells model.cpp:252-254
void Model::updateNeighborCells() {
    ensureCells();
    check(pcp_update_cells(ctx, positions));
}
// updateNeighbors model.cpp:256-307 (overflow: warn, grow maxNeighbors)
void Model::updateNeighbors() {
    ensureCells();
    int mx = 0;
    check(pcp_update_neighbors(ctx, positions, &mx));
    if (mx > maxNeighbors) {
        std::cerr << "Warning: neighbor buffer overflow: maxNeighbors=" << maxNeighbors
                  << " required=" << mx << ". Resizing and retrying.\n";
        maxNeighbors = std::max(maxNeighbors * 2 + 1, mx);
        check(pcp_set_max_neighbors(ctx, maxNeighbors));
    }
}
// updateValidAndCounts model.cpp:309-311: numIntersections and shapeCounts
void Model::updateValidAndCounts() {
    ensureCells();
    int I = 0;
    check(pcp_update_records(ctx, nullptr, &I));
    numIntersections = I;
}
// updateCompactedIntersections model.cpp:347-349: records in compaction order
void Model::updateCompactedIntersections() {
    ensureCells();
    int I = 0;
    check(pcp_update_records(ctx, energyFlags ? positions : nullptr, &I));   // exact pairing for the area energies
    numIntersections = I;
    viewKeys.assign(I, 0);
    viewTU.assign(2 * (size_t)I, 0.0);
    if (I) check(pcp_get_compacted_records(ctx, 0, viewKeys.data(), viewTU.data()));
    compactedView = true;
    keysView = false;
}
// updateOutersections model.cpp:351-361
void Model::updateOutersections() {
    ensureCells();
    int I = 0;
    check(pcp_update_records(ctx, energyFlags ? positions : nullptr, &I));   // exact pairing for the area energies
    numIntersections = I;
    compactedView = false;
    keysView = false;
}
// updateOverlapArea model.cpp:824-847
void Model::updateOverlapArea(int D) {
    ensureCells();
    if (pointDensity != D) {
        dfree(intersectionsCounter);
        dalloc(&intersectionsCounter, (size_t)D * D);
        pointDensity = D;
    }
    long long sum = 0;
    check(pcp_overlap_count(ctx, positions, D, intersectionsCounter, &sum));
    overlapArea = (double)sum / (double)((long long)D * D);
}
// updateForceEnergy model.cpp:363-395
void Model::updateForceEnergy() {
    ensureTopology();
    bool overlap = simControl.modelType == simControlStruct::modelEnum::normal ||
                   simControl.modelType == simControlStruct::modelEnum::hybrid;
    if (overlap) ensureCells();
    check(pcp_update_force_energy(ctx, (int)simControl.modelType | energyFlags, positions, edgeLengths, areas, targetEdgeV,
                                  targetAreas, stiffness, compressibility, force, energy));
}
// updatePositions model.cpp:397-399
void Model::updatePositions(double dt) {
    ensureTopology();
    check(pcp_update_positions(ctx, positions, force, dt));
}
// resetVelocities model.cpp:401-403
void Model::resetVelocities() {
    check(pcp_synchronize(ctx));
    check(pcp_memset(velocities, 0, 16 * (size_t)size));
}
// resetAreas model.cpp:507-509
void Model::resetAreas() {
    ensureTopology();
    check(pcp_reset_areas(ctx, positions, areas, targetAreas, comX, comY));
}

// minimizeFIREStep model.cpp:405-486.  Returns (energy, dt, alpha, nPos); on an energy
// rollback the returned energy is E_pre (model.cpp:469).
std::tuple<double, double, double, int> Model::minimizeFIREStep(double dt, double alpha, int nPos, double dtMax,
                                                                double alphaStart, double fAlpha, double fInc,
                                                                double fDec, int nMin, int shakeIter) {
    bool overlap = simControl.modelType == simControlStruct::modelEnum::normal ||
                   simControl.modelType == simControlStruct::modelEnum::hybrid;
    if (overlap) ensureCells(); else ensureTopology();
    double ePre = getEnergy();
    check(pcp_fire_set(ctx, &dt, &alpha, &nPos));
    pcp_fire_params prm{dtMax, alphaStart, fAlpha, fInc, fDec, nMin, shakeIter, -1.0};
    pcp_state st = state();
    pcp_fire_status s;
    check(pcp_fire_step(ctx, &st, &prm, &s));
    if (simControl.modelType == simControlStruct::modelEnum::normal && shakeIter > 0)
        lastShakeIters = s.lastShakeIters;
    if (overlap) check(pcp_get_num_intersections(ctx, 0, &numIntersections));
    compactedView = keysView = false;
    return {s.rolledBack ? ePre : s.energy, s.dt, s.alpha, s.nPos};
}

// minimizeFIRE model.cpp:488-504: fresh forces from the current records, then FIRE
// steps until max |force component| <= maxForceThreshold
std::tuple<double, double, int> Model::minimizeFIRE(double thr, double dtInit, int maxSteps, double dtMax,
                                                    double alphaStart, double fAlpha, double fInc, double fDec,
                                                    int nMin, int shakeIter) {
    bool overlap = simControl.modelType == simControlStruct::modelEnum::normal ||
                   simControl.modelType == simControlStruct::modelEnum::hybrid;
    if (overlap) ensureCells(); else ensureTopology();
    resetVelocities();
    if (energyFlags && overlap) {
        // the area energies need crossings with exact parameters: one full evaluation
        pcp_state st0 = state();
        int rc = PCP_OK;
        for (int attempt = 0; attempt < 4; ++attempt) {   // the buckets grow on overflow: evaluate again
            rc = pcp_force_eval(ctx, &st0);
            if (rc == PCP_OK) rc = pcp_synchronize(ctx);
            if (rc != PCP_ERR_CAPACITY) break;
        }
        check(rc);
    } else {
        updatePolygonGeometry();
        updateForceEnergy();
        if (simControl.modelType == simControlStruct::modelEnum::normal) projectForce();
    }
    int nPos = 0;
    check(pcp_fire_set(ctx, &dtInit, &alphaStart, &nPos));
    pcp_fire_params prm{dtMax, alphaStart, fAlpha, fInc, fDec, nMin, shakeIter, thr};
    pcp_state st = state();
    pcp_fire_status s;
    check(pcp_fire_minimize(ctx, &st, &prm, maxSteps, &s));
    if (simControl.modelType == simControlStruct::modelEnum::normal && shakeIter > 0)
        lastShakeIters = s.lastShakeIters;
    if (overlap) check(pcp_get_num_intersections(ctx, 0, &numIntersections));
    compactedView = keysView = false;
    if (s.converged) return {s.energy, s.dt, s.steps};
    return {s.energy, s.dt, maxSteps};
}
