// Model: the reference's C++ class (camdennis/pyCudaPolygon src/model.hpp:18-185) with
// the same public methods and semantics, implemented over the C ABI of libpcp_hip.so
// (include/pcp.h).  One Model = one packing (a pcp context with P = 1).
#pragma once
#include <cstdint>
#include <string>
#include <tuple>
#include <vector>
#include "pcp.h"

struct simControlStruct {
    enum class modelEnum { normal, edgeOnly, areaOnly, softBody, abnormal, hybrid } modelType;
};
enum class minimizerEnum { GD, FIRE };

class Model {
public:
    explicit Model(int size_);
    ~Model();
    Model(const Model&) = delete;
    Model& operator=(const Model&) = delete;

    // initializers
    void initializeNeighborCells();
    void initializeRandomSeed(unsigned long long seed_);

    // helpers
    void sortKeys(int endBit);
    std::vector<uint32_t> getKeys() const;

    // setters
    void setNumVertices(int size);
    void setPositions(const std::vector<double>& positions_);
    void setForces(const std::vector<double>& forces_);
    void setStartIndices(const std::vector<int>& startIndices_);
    void setModelEnum(simControlStruct::modelEnum modelType_);
    // overlap energy of the overlap models (no counterpart in the reference): "reference"
    // (its convention, the default), "area" (overlap area, force = -gradient), "squaredLens"
    // (sum over lenses of A^2/2) or "squaredPair" (sum over polygon pairs of A_pair^2/2);
    // the last two are what FIRE can minimize to |F|max <= 1e-10
    void setOverlapEnergy(const std::string& name);
    std::string getOverlapEnergy() const;
    void setMaxEdgeLength(double maxEdgeLength);
    void setTargetEdgeLengths(const std::vector<double>& targetEdgeLengths_);
    void setTargetAreas(const std::vector<double>& targetAreas_);
    void setStiffness(double stiffness_);
    void setCompressibility(double compressibility_);
    double getStiffness() const;
    double getCompressibility() const;
    // extensions: per-edge target lengths (edge m -> next(m)), reference M
    void setTargetEdgeLengthsPerEdge(const std::vector<double>& t);
    std::vector<double> getTargetEdgeLengthsPerEdge() const;
    int getMaxNeighbors() const;
    int getBoxSize() const;

    // getters
    int getNumVertices() const;
    int getNumPolygons() const;
    std::vector<int> getShapeId() const;
    std::vector<double> getPositions() const;
    std::vector<int> getStartIndices() const;
    std::vector<int> getNeighborCells() const;
    std::vector<int> getBoxCounts() const;
    std::vector<int> getNeighborIndices() const;
    std::vector<int> getNumNeighbors() const;
    std::vector<int> getNeighbors() const;
    double getMaxEdgeLength() const;
    std::string getModelEnum() const;
    std::vector<double> getForces() const;
    void resetMaxActualNeighbors();
    std::vector<bool> getInsideFlag() const;
    std::vector<double> getPerimeters() const;
    std::vector<int> getIntersectionsCounter() const;
    std::vector<double> getTU() const;
    std::vector<double> getUT() const;
    std::vector<int> getShapeCounts() const;
    std::vector<uint64_t> getIntersections() const;
    int getNumIntersections() const;
    std::vector<uint64_t> getOutersections() const;
    unsigned long long getRandomSeed() const;
    std::vector<double> getAreas() const;
    double getEnergy() const;
    std::vector<double> getTargetEdgeLengths() const;
    std::vector<double> getTargetAreas() const;
    std::vector<double> getCOM() const;
    double getMaxUnbalancedForce() const;
    std::vector<double> getEdgeLengths() const;
    double getOverlapArea() const;
    std::vector<double> getConstraints() const;

    // updaters
    void updatePolygonGeometry();
    void projectForce();
    int shakeProject(int nIter, double tol = 1e-15);
    int xpbdProject(int nIter, double tol);
    int getLastShakeIters() const;
    void saveTentativePositions();
    double getMaxEffectiveForce(double dt, minimizerEnum minimizerType) const;
    void updateNeighborCells();
    void updateNeighbors();
    void updateOverlapArea(int pointDensity_);
    void updateValidAndCounts();
    void updateOutersections();
    void updateCompactedIntersections();
    void updateForceEnergy();
    void updatePositions(double dt);
    void resetVelocities();
    std::tuple<double, double, double, int> minimizeFIREStep(double dt, double alpha, int nPos,
                                                             double dtMax = 0.1, double alphaStart = 0.1,
                                                             double fAlpha = 0.99, double fInc = 1.1,
                                                             double fDec = 0.5, int nMin = 5,
                                                             int shakeIter = 5);
    std::tuple<double, double, int> minimizeFIRE(double maxForceThreshold, double dtInit, int maxSteps,
                                                 double dtMax = 0.1, double alphaStart = 0.1,
                                                 double fAlpha = 0.99, double fInc = 1.1, double fDec = 0.5,
                                                 int nMin = 5, int shakeIter = 5);

    // misc
    void resetAreas();

private:
    void check(int status) const;
    void ensureTopology() const;
    void ensureCells() const;
    void freeVertexBuffers();
    void freePolygonBuffers();
    pcp_state state() const;
    template <class T> std::vector<T> d2h(const T* d, size_t n) const;

    simControlStruct simControl{simControlStruct::modelEnum::normal};
    int energyFlags = 0;   // PCP_AREA_ENERGY / PCP_AREA_SQUARED / PCP_PAIR_SQUARED or 0
    mutable pcp_ctx* ctx = nullptr;
    mutable bool topoDirty = true, boxDirty = true;
    int size = 0, numPolygons = 0;
    std::vector<int> hStart;
    std::vector<double> hTargetEdgeP;
    unsigned long long seed = 0;
    int boxSize = 1;
    int maxNeighbors = 100;
    double stiffness = 0.0, compressibility = 0.0;
    int numIntersections = 0;
    int lastShakeIters = 0;
    int pointDensity = -1;
    double overlapArea = 0.0;
    bool compactedView = false;
    bool keysView = false;
    std::vector<uint64_t> viewKeys;
    std::vector<double> viewTU;
    std::vector<uint32_t> keys;

    // device buffers
    double *positions = nullptr, *force = nullptr, *velocities = nullptr, *positionsTMP = nullptr;
    double *edgeLengths = nullptr, *constraints = nullptr, *targetEdgeV = nullptr;
    double *areas = nullptr, *comX = nullptr, *comY = nullptr, *targetAreas = nullptr;
    double *energy = nullptr, *maxEdgeLength = nullptr, *scratch = nullptr;
    int* intersectionsCounter = nullptr;
};
