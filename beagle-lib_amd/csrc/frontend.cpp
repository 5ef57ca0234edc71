// C ABI of libhmsbeagle_amd.so: the beagle.h entry points, the instance
// table, the resource list and the reference's flag-scoring selection.
//
// Mirrors the reference front end (libhmsbeagle/beagle.cpp): resources are
// scored by counting preference bits 0-31 they support (beagle.cpp:341-351),
// filtered by requirement bits (:353-389), then (resource, implementation)
// pairs are ranked and tried in order (:391-460, 632-742).  Differences:
// the instance-number check is `>=` (the reference's `>` reads past the
// table, beagle.cpp:94), every entry point catches every C++ exception, and
// the only implementations are the two gfx950 ones (double and single).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <list>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "beagle.h"
#include "beagle_amd_ext.h"
#include "instance.h"

using hipbeagle::Instance;

namespace {

struct Factory {
    const char* name;
    long flags;
    bool dp;
};

std::mutex gMutex;
std::vector<Instance*> gInstances;
std::vector<BeagleResource> gRes;
std::vector<std::string> gResNames, gResDesc;
std::vector<int> gResDevice;  // -1 for the host entry
BeagleResourceList gResList = {nullptr, 0};
bool gResInit = false;

const long kGpuSupport =
    BEAGLE_FLAG_COMPUTATION_SYNCH | BEAGLE_FLAG_COMPUTATION_ASYNCH | BEAGLE_FLAG_PRECISION_SINGLE |
    BEAGLE_FLAG_PRECISION_DOUBLE | BEAGLE_FLAG_SCALING_MANUAL | BEAGLE_FLAG_SCALING_ALWAYS |
    BEAGLE_FLAG_SCALING_AUTO |
    BEAGLE_FLAG_THREADING_NONE | BEAGLE_FLAG_VECTOR_NONE | BEAGLE_FLAG_PROCESSOR_GPU |
    BEAGLE_FLAG_SCALERS_LOG | BEAGLE_FLAG_SCALERS_RAW | BEAGLE_FLAG_EIGEN_REAL |
    BEAGLE_FLAG_EIGEN_COMPLEX |
    BEAGLE_FLAG_INVEVEC_STANDARD | BEAGLE_FLAG_INVEVEC_TRANSPOSED |
    BEAGLE_FLAG_PREORDER_TRANSPOSE_MANUAL | (long)BEAGLE_FLAG_PREORDER_TRANSPOSE_AUTO |
    BEAGLE_FLAG_PARALLELOPS_GRID | BEAGLE_FLAG_PARALLELOPS_STREAMS | BEAGLE_FLAG_FRAMEWORK_HIP;

std::vector<Factory> factories() {
    return {{"MI355X-HIP-Double", kGpuSupport & ~(long)BEAGLE_FLAG_PRECISION_SINGLE, true},
            {"MI355X-HIP-Single", kGpuSupport & ~(long)BEAGLE_FLAG_PRECISION_DOUBLE, false}};
}

void initResources() {
    if (gResInit) return;
    gResInit = true;
    gResNames.push_back("CPU (host)");
    gResDesc.push_back("host processor; no implementation of this library runs on it");
    gResDevice.push_back(-1);
    std::vector<long> support = {BEAGLE_FLAG_PROCESSOR_CPU | BEAGLE_FLAG_FRAMEWORK_CPU |
                                 BEAGLE_FLAG_COMPUTATION_SYNCH};
    std::vector<long> required = {BEAGLE_FLAG_FRAMEWORK_CPU};
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) != hipSuccess) continue;
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) continue;
        gResNames.push_back(prop.name[0] ? std::string(prop.name) : std::string("AMD gfx950"));
        gResDesc.push_back(std::string("gfx950, ") + std::to_string(prop.multiProcessorCount) +
                           " CUs, " + std::to_string(prop.totalGlobalMem >> 30) + " GiB HBM");
        gResDevice.push_back(d);
        support.push_back(kGpuSupport);
        required.push_back(BEAGLE_FLAG_FRAMEWORK_HIP);
    }
    gRes.resize(gResNames.size());
    for (size_t i = 0; i < gRes.size(); ++i) {
        gRes[i].name = const_cast<char*>(gResNames[i].c_str());
        gRes[i].description = const_cast<char*>(gResDesc[i].c_str());
        gRes[i].supportFlags = support[i];
        gRes[i].requiredFlags = required[i];
    }
    gResList.list = gRes.data();
    gResList.length = (int)gRes.size();
}

// Negative count of preference bits (0-31 only) present in flags
// (beagle.cpp:341-351).
int scoreFlags(long a, long b) {
    int score = 0;
    long trait = 1;
    for (int bit = 0; bit < 32; ++bit, trait <<= 1)
        if ((a & trait) && (b & trait)) ++score;
    return -score;
}

Instance* instanceAt(int i) {
    std::lock_guard<std::mutex> lock(gMutex);
    if (i < 0 || i >= (int)gInstances.size()) return nullptr;
    return gInstances[i];
}

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (std::bad_alloc&) {
        return BEAGLE_ERROR_OUT_OF_MEMORY;
    } catch (std::out_of_range&) {
        return BEAGLE_ERROR_OUT_OF_RANGE;
    } catch (std::runtime_error&) {
        return BEAGLE_ERROR_GENERAL;
    } catch (...) {
        return BEAGLE_ERROR_UNIDENTIFIED_EXCEPTION;
    }
}

template <typename F>
int withInstance(int instance, F&& f) {
    Instance* in = instanceAt(instance);
    if (!in) return BEAGLE_ERROR_UNINITIALIZED_INSTANCE;
    return guarded([&]() { return f(in); });
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

const char* beagleGetVersion(void) { return "4.0.1 (MI355X HIP)"; }

const char* beagleGetCitation(void) {
    return "Using the MI355X (gfx950) BEAGLE plugin, API-compatible with BEAGLE v4.0.1\n"
           "--- BEAGLE RESOURCES ---\n"
           "Ayres et al. (2019) BEAGLE 3: improved performance, scaling, and usability for a "
           "high-performance computing library for statistical phylogenetics. Syst Biol 68: "
           "1052-1061.\n";
}

BeagleResourceList* beagleGetResourceList(void) {
    std::lock_guard<std::mutex> lock(gMutex);
    initResources();
    return &gResList;
}

BeagleBenchmarkedResourceList* beagleGetBenchmarkedResourceList(int, int, int, int, int, int*,
                                                                int, long, long, int, int, int,
                                                                long) {
    return nullptr;
}

int beagleCreateInstance(int tipCount, int partialsBufferCount, int compactBufferCount,
                         int stateCount, int patternCount, int eigenBufferCount,
                         int matrixBufferCount, int categoryCount, int scaleBufferCount,
                         int* resourceList, int resourceCount, long preferenceFlags,
                         long requirementFlags, BeagleInstanceDetails* returnInfo) {
    return guarded([&]() -> int {
        {
            std::lock_guard<std::mutex> lock(gMutex);
            initResources();
        }
        // filterResources (beagle.cpp:353-389)
        std::list<std::pair<int, int>> possible;
        if (resourceList == nullptr || resourceCount == 0) {
            for (int i = 0; i < gResList.length; ++i)
                possible.push_back({scoreFlags(preferenceFlags, gRes[i].supportFlags), i});
        } else {
            for (int i = 0; i < resourceCount; ++i) {
                const int r = resourceList[i];
                if (r < 0 || r >= gResList.length) return BEAGLE_ERROR_OUT_OF_RANGE;
                possible.push_back({scoreFlags(preferenceFlags, gRes[r].supportFlags), r});
            }
        }
        if (requirementFlags != 0)
            possible.remove_if([&](const std::pair<int, int>& p) {
                return (gRes[p.second].supportFlags & requirementFlags) < requirementFlags;
            });
        if (possible.empty()) return BEAGLE_ERROR_NO_RESOURCE;
        possible.sort([](const std::pair<int, int>& a, const std::pair<int, int>& b) {
            return a.first < b.first;
        });
        // rankResourceImplementationPairs (beagle.cpp:391-460)
        struct Cand {
            int score, resource;
            Factory f;
        };
        std::list<Cand> cands;
        for (auto& p : possible) {
            const BeagleResource& R = gRes[p.second];
            for (const Factory& f : factories()) {
                if ((requirementFlags & f.flags) >= requirementFlags &&
                    (R.requiredFlags & f.flags) >= R.requiredFlags &&
                    (requirementFlags & R.supportFlags) >= requirementFlags)
                    cands.push_back({p.first + scoreFlags(preferenceFlags, f.flags), p.second, f});
            }
        }
        cands.sort([](const Cand& a, const Cand& b) { return a.score < b.score; });
        int err = BEAGLE_ERROR_NO_RESOURCE;
        for (const Cand& c : cands) {
            const long both = preferenceFlags | requirementFlags;
            // complex eigen-systems stage V and V^-1 in LDS: up to 62 states
            if ((both & BEAGLE_FLAG_SCALING_DYNAMIC) ||
                ((both & BEAGLE_FLAG_EIGEN_COMPLEX) && stateCount > 62)) {
                err = BEAGLE_ERROR_NO_IMPLEMENTATION;
                continue;
            }
            Instance* inst = new Instance(tipCount, partialsBufferCount, compactBufferCount,
                                          stateCount, patternCount, eigenBufferCount,
                                          matrixBufferCount, categoryCount, scaleBufferCount,
                                          gResDevice[c.resource], c.resource, preferenceFlags,
                                          requirementFlags, c.f.dp);
            int id;
            {
                std::lock_guard<std::mutex> lock(gMutex);
                id = (int)gInstances.size();
                gInstances.push_back(inst);
            }
            if (returnInfo) {
                returnInfo->resourceNumber = c.resource;
                returnInfo->resourceName = gRes[c.resource].name;
                returnInfo->implName = const_cast<char*>(inst->implName());
                returnInfo->implDescription = const_cast<char*>("none");
                returnInfo->flags = inst->flags();
            }
            return id;
        }
        return err;
    });
}

int beagleFinalizeInstance(int instance) {
    return guarded([&]() -> int {
        Instance* in;
        {
            std::lock_guard<std::mutex> lock(gMutex);
            if (instance < 0 || instance >= (int)gInstances.size() || !gInstances[instance])
                return BEAGLE_ERROR_UNINITIALIZED_INSTANCE;
            in = gInstances[instance];
            gInstances[instance] = nullptr;
        }
        delete in;
        return BEAGLE_SUCCESS;
    });
}

int beagleFinalize(void) {
    return guarded([&]() -> int {
        std::vector<Instance*> all;
        {
            std::lock_guard<std::mutex> lock(gMutex);
            all.swap(gInstances);
        }
        for (Instance* in : all) delete in;
        return BEAGLE_SUCCESS;
    });
}

int beagleSetCPUThreadCount(int instance, int) {
    return withInstance(instance, [&](Instance*) { return BEAGLE_SUCCESS; });
}

int beagleSetTipStates(int instance, int tipIndex, const int* inStates) {
    return withInstance(instance, [&](Instance* in) { return in->setTipStates(tipIndex, inStates); });
}

int beagleSetTipPartials(int instance, int tipIndex, const double* inPartials) {
    return withInstance(instance,
                        [&](Instance* in) { return in->setTipPartials(tipIndex, inPartials); });
}

int beagleSetRootPrePartials(const int instance, const int* bufferIndices,
                             const int* stateFrequenciesIndices, int count) {
    return withInstance(instance, [&](Instance* in) {
        return in->setRootPrePartials(bufferIndices, stateFrequenciesIndices, count);
    });
}

int beagleSetPartials(int instance, int bufferIndex, const double* inPartials) {
    return withInstance(instance, [&](Instance* in) { return in->setPartials(bufferIndex, inPartials); });
}

int beagleGetPartials(int instance, int bufferIndex, int scaleIndex, double* outPartials) {
    return withInstance(instance, [&](Instance* in) {
        return in->getPartials(bufferIndex, scaleIndex, outPartials);
    });
}

int beagleSetEigenDecomposition(int instance, int eigenIndex, const double* inEigenVectors,
                                const double* inInverseEigenVectors, const double* inEigenValues) {
    return withInstance(instance, [&](Instance* in) {
        return in->setEigenDecomposition(eigenIndex, inEigenVectors, inInverseEigenVectors,
                                         inEigenValues);
    });
}

int beagleSetStateFrequencies(int instance, int idx, const double* f) {
    return withInstance(instance, [&](Instance* in) { return in->setStateFrequencies(idx, f); });
}

int beagleSetCategoryWeights(int instance, int idx, const double* w) {
    return withInstance(instance, [&](Instance* in) { return in->setCategoryWeights(idx, w); });
}

int beagleSetCategoryRates(int instance, const double* r) {
    return withInstance(instance, [&](Instance* in) { return in->setCategoryRates(r); });
}

int beagleSetCategoryRatesWithIndex(int instance, int idx, const double* r) {
    return withInstance(instance, [&](Instance* in) { return in->setCategoryRatesWithIndex(idx, r); });
}

int beagleSetPatternWeights(int instance, const double* w) {
    return withInstance(instance, [&](Instance* in) { return in->setPatternWeights(w); });
}

int beagleSetPatternPartitions(int instance, int partitionCount, const int* parts) {
    return withInstance(instance,
                        [&](Instance* in) { return in->setPatternPartitions(partitionCount, parts); });
}

int beagleConvolveTransitionMatrices(int instance, const int* a, const int* b, const int* out,
                                     int count) {
    return withInstance(instance,
                        [&](Instance* in) { return in->convolveTransitionMatrices(a, b, out, count); });
}

int beagleAddTransitionMatrices(int instance, const int* a, const int* b, const int* out,
                                int count) {
    return withInstance(instance,
                        [&](Instance* in) { return in->addTransitionMatrices(a, b, out, count); });
}

int beagleTransposeTransitionMatrices(int instance, const int* inIdx, const int* outIdx, int count) {
    return withInstance(instance, [&](Instance* in) {
        return in->transposeTransitionMatrices(inIdx, outIdx, count);
    });
}

int beagleUpdateTransitionMatrices(int instance, int eigenIndex, const int* probabilityIndices,
                                   const int* firstDerivativeIndices,
                                   const int* secondDerivativeIndices, const double* edgeLengths,
                                   int count) {
    return withInstance(instance, [&](Instance* in) {
        return in->updateTransitionMatrices(eigenIndex, probabilityIndices, firstDerivativeIndices,
                                            secondDerivativeIndices, edgeLengths, count);
    });
}

int beagleUpdateTransitionMatricesWithModelCategories(int instance, int* eigenIndices,
                                                      const int* probabilityIndices,
                                                      const int* firstDerivativeIndices,
                                                      const int* secondDerivativeIndices,
                                                      const double* edgeLengths, int count) {
    return withInstance(instance, [&](Instance* in) {
        return in->updateTransitionMatricesWithModelCategories(
            eigenIndices, probabilityIndices, firstDerivativeIndices, secondDerivativeIndices,
            edgeLengths, count);
    });
}

int beagleUpdateTransitionMatricesWithMultipleModels(int instance, const int* eigenIndices,
                                                     const int* categoryRateIndices,
                                                     const int* probabilityIndices,
                                                     const int* firstDerivativeIndices,
                                                     const int* secondDerivativeIndices,
                                                     const double* edgeLengths, int count) {
    return withInstance(instance, [&](Instance* in) {
        return in->updateTransitionMatricesWithMultipleModels(
            eigenIndices, categoryRateIndices, probabilityIndices, firstDerivativeIndices,
            secondDerivativeIndices, edgeLengths, count);
    });
}

int beagleSetTransitionMatrix(int instance, int matrixIndex, const double* inMatrix,
                              double paddedValue) {
    return withInstance(instance, [&](Instance* in) {
        return in->setTransitionMatrix(matrixIndex, inMatrix, paddedValue);
    });
}

int beagleSetDifferentialMatrix(int instance, int matrixIndex, const double* inMatrix) {
    return withInstance(instance,
                        [&](Instance* in) { return in->setDifferentialMatrix(matrixIndex, inMatrix); });
}

int beagleGetTransitionMatrix(int instance, int matrixIndex, double* outMatrix) {
    return withInstance(instance,
                        [&](Instance* in) { return in->getTransitionMatrix(matrixIndex, outMatrix); });
}

int beagleSetTransitionMatrices(int instance, const int* matrixIndices, const double* inMatrices,
                                const double* paddedValues, int count) {
    return withInstance(instance, [&](Instance* in) {
        return in->setTransitionMatrices(matrixIndices, inMatrices, paddedValues, count);
    });
}

int beagleUpdatePartials(const int instance, const BeagleOperation* operations, int operationCount,
                         int cumulativeScaleIndex) {
    return withInstance(instance, [&](Instance* in) {
        return in->updatePartials(reinterpret_cast<const int*>(operations), operationCount,
                                  cumulativeScaleIndex, false);
    });
}

int beagleUpdatePrePartials(const int instance, const BeagleOperation* operations,
                            int operationCount, int cumulativeScaleIndex) {
    return withInstance(instance, [&](Instance* in) {
        return in->updatePrePartials(reinterpret_cast<const int*>(operations), operationCount,
                                     cumulativeScaleIndex, false);
    });
}

int beagleUpdatePartialsByPartition(const int instance, const BeagleOperationByPartition* operations,
                                    int operationCount) {
    return withInstance(instance, [&](Instance* in) {
        return in->updatePartials(reinterpret_cast<const int*>(operations), operationCount,
                                  BEAGLE_OP_NONE, true);
    });
}

int beagleUpdatePrePartialsByPartition(const int instance,
                                       const BeagleOperationByPartition* operations,
                                       int operationCount) {
    return withInstance(instance, [&](Instance* in) {
        return in->updatePrePartials(reinterpret_cast<const int*>(operations), operationCount,
                                     BEAGLE_OP_NONE, true);
    });
}

int beagleWaitForPartials(const int instance, const int*, int) {
    return withInstance(instance, [&](Instance* in) { return in->waitForPartials(); });
}

int beagleAccumulateScaleFactors(int instance, const int* scaleIndices, int count,
                                 int cumulativeScaleIndex) {
    return withInstance(instance, [&](Instance* in) {
        return in->accumulateScaleFactors(scaleIndices, count, cumulativeScaleIndex, 1.0, -1);
    });
}

int beagleAccumulateScaleFactorsByPartition(int instance, const int* scaleIndices, int count,
                                            int cumulativeScaleIndex, int partitionIndex) {
    return withInstance(instance, [&](Instance* in) {
        return in->accumulateScaleFactors(scaleIndices, count, cumulativeScaleIndex, 1.0,
                                          partitionIndex);
    });
}

int beagleRemoveScaleFactors(int instance, const int* scaleIndices, int count,
                             int cumulativeScaleIndex) {
    return withInstance(instance, [&](Instance* in) {
        return in->accumulateScaleFactors(scaleIndices, count, cumulativeScaleIndex, -1.0, -1);
    });
}

int beagleRemoveScaleFactorsByPartition(int instance, const int* scaleIndices, int count,
                                        int cumulativeScaleIndex, int partitionIndex) {
    return withInstance(instance, [&](Instance* in) {
        return in->accumulateScaleFactors(scaleIndices, count, cumulativeScaleIndex, -1.0,
                                          partitionIndex);
    });
}

int beagleResetScaleFactors(int instance, int cumulativeScaleIndex) {
    return withInstance(instance,
                        [&](Instance* in) { return in->resetScaleFactors(cumulativeScaleIndex, -1); });
}

int beagleResetScaleFactorsByPartition(int instance, int cumulativeScaleIndex, int partitionIndex) {
    return withInstance(instance, [&](Instance* in) {
        return in->resetScaleFactors(cumulativeScaleIndex, partitionIndex);
    });
}

int beagleCopyScaleFactors(int instance, int dst, int src) {
    return withInstance(instance, [&](Instance* in) { return in->copyScaleFactors(dst, src); });
}

int beagleGetScaleFactors(int instance, int src, double* out) {
    return withInstance(instance, [&](Instance* in) { return in->getScaleFactors(src, out); });
}

int beagleExtCalculateRootLogLikelihoodsNoAlloc(int instance, const int* bufferIndices,
                                                const int* categoryWeightsIndices,
                                                const int* stateFrequenciesIndices,
                                                const int* cumulativeScaleIndices, int count,
                                                double* outSumLogLikelihood) {
    return withInstance(instance, [&](Instance* in) {
        return in->calculateRootLogLikelihoodsNoAlloc(bufferIndices, categoryWeightsIndices,
                                                      stateFrequenciesIndices,
                                                      cumulativeScaleIndices, count,
                                                      outSumLogLikelihood);
    });
}

int beagleCalculateRootLogLikelihoods(int instance, const int* bufferIndices,
                                      const int* categoryWeightsIndices,
                                      const int* stateFrequenciesIndices,
                                      const int* cumulativeScaleIndices, int count,
                                      double* outSumLogLikelihood) {
    return withInstance(instance, [&](Instance* in) {
        return in->calculateRootLogLikelihoods(bufferIndices, categoryWeightsIndices,
                                               stateFrequenciesIndices, cumulativeScaleIndices,
                                               count, outSumLogLikelihood);
    });
}

int beagleCalculateRootLogLikelihoodsByPartition(
    int instance, const int* bufferIndices, const int* categoryWeightsIndices,
    const int* stateFrequenciesIndices, const int* cumulativeScaleIndices,
    const int* partitionIndices, int partitionCount, int count,
    double* outSumLogLikelihoodByPartition, double* outSumLogLikelihood) {
    return withInstance(instance, [&](Instance* in) {
        return in->calculateRootLogLikelihoodsByPartition(
            bufferIndices, categoryWeightsIndices, stateFrequenciesIndices, cumulativeScaleIndices,
            partitionIndices, partitionCount, count, outSumLogLikelihoodByPartition,
            outSumLogLikelihood);
    });
}

int beagleCalculateEdgeDerivative(int, const int*, const int*, const int, const int*, const int*,
                                  const int, const int, const int, const int*, int, double*,
                                  double*) {
    return BEAGLE_ERROR_NO_IMPLEMENTATION;
}

int beagleCalculateEdgeDerivatives(int instance, const int* postBufferIndices,
                                   const int* preBufferIndices, const int* derivativeMatrixIndices,
                                   const int* categoryWeightsIndices, int count,
                                   double* outDerivatives, double* outSumDerivatives,
                                   double* outSumSquaredDerivatives) {
    return withInstance(instance, [&](Instance* in) {
        return in->calculateEdgeDerivatives(postBufferIndices, preBufferIndices,
                                            derivativeMatrixIndices, categoryWeightsIndices, count,
                                            outDerivatives, outSumDerivatives,
                                            outSumSquaredDerivatives);
    });
}

int beagleCalculateCrossProductDerivative(int instance, const int* postBufferIndices,
                                          const int* preBufferIndices,
                                          const int* categoryRatesIndices,
                                          const int* categoryWeightsIndices,
                                          const double* edgeLengths, int count,
                                          double* outSumDerivatives,
                                          double* outSumSquaredDerivatives) {
    (void)outSumSquaredDerivatives;
    return withInstance(instance, [&](Instance* in) {
        return in->calculateCrossProducts(postBufferIndices, preBufferIndices, categoryRatesIndices,
                                          categoryWeightsIndices, edgeLengths, count,
                                          outSumDerivatives);
    });
}

int beagleCalculateEdgeLogLikelihoods(int instance, const int* parentBufferIndices,
                                      const int* childBufferIndices, const int* probabilityIndices,
                                      const int* firstDerivativeIndices,
                                      const int* secondDerivativeIndices,
                                      const int* categoryWeightsIndices,
                                      const int* stateFrequenciesIndices,
                                      const int* cumulativeScaleIndices, int count,
                                      double* outSumLogLikelihood, double* outSumFirstDerivative,
                                      double* outSumSecondDerivative) {
    return withInstance(instance, [&](Instance* in) {
        return in->calculateEdgeLogLikelihoods(
            parentBufferIndices, childBufferIndices, probabilityIndices, firstDerivativeIndices,
            secondDerivativeIndices, categoryWeightsIndices, stateFrequenciesIndices,
            cumulativeScaleIndices, count, outSumLogLikelihood, outSumFirstDerivative,
            outSumSecondDerivative);
    });
}

int beagleCalculateEdgeLogLikelihoodsByPartition(
    int instance, const int* parentBufferIndices, const int* childBufferIndices,
    const int* probabilityIndices, const int* firstDerivativeIndices,
    const int* secondDerivativeIndices, const int* categoryWeightsIndices,
    const int* stateFrequenciesIndices, const int* cumulativeScaleIndices,
    const int* partitionIndices, int partitionCount, int count,
    double* outSumLogLikelihoodByPartition, double* outSumLogLikelihood,
    double* outSumFirstDerivativeByPartition, double* outSumFirstDerivative,
    double* outSumSecondDerivativeByPartition, double* outSumSecondDerivative) {
    return withInstance(instance, [&](Instance* in) {
        return in->calculateEdgeLogLikelihoodsByPartition(
            parentBufferIndices, childBufferIndices, probabilityIndices, firstDerivativeIndices,
            secondDerivativeIndices, categoryWeightsIndices, stateFrequenciesIndices,
            cumulativeScaleIndices, partitionIndices, partitionCount, count,
            outSumLogLikelihoodByPartition, outSumLogLikelihood, outSumFirstDerivativeByPartition,
            outSumFirstDerivative, outSumSecondDerivativeByPartition, outSumSecondDerivative);
    });
}

int beagleGetLogLikelihood(int instance, double* out) {
    return withInstance(instance, [&](Instance* in) { return in->getLogLikelihood(out); });
}

int beagleGetDerivatives(int instance, double* d1, double* d2) {
    return withInstance(instance, [&](Instance* in) { return in->getDerivatives(d1, d2); });
}

int beagleGetSiteLogLikelihoods(int instance, double* out) {
    return withInstance(instance, [&](Instance* in) { return in->getSiteLogLikelihoods(out); });
}

int beagleGetSiteDerivatives(int instance, double* d1, double* d2) {
    return withInstance(instance, [&](Instance* in) { return in->getSiteDerivatives(d1, d2); });
}

int beagleExtUpdatePrePartialsWithDerivatives(int instance, const BeagleOperation* operations,
                                              int operationCount, const int* postBufferIndices,
                                              const int* firstDerivativeIndices,
                                              const int* secondDerivativeIndices,
                                              int categoryWeightsIndex,
                                              double* outSumFirstDerivatives,
                                              double* outSumSquaredFirstDerivatives,
                                              double* outSumSecondDerivatives) {
    return withInstance(instance, [&](Instance* in) {
        return in->updatePrePartialsWithDerivatives(
            reinterpret_cast<const int*>(operations), operationCount, postBufferIndices,
            firstDerivativeIndices, secondDerivativeIndices, categoryWeightsIndex,
            outSumFirstDerivatives, outSumSquaredFirstDerivatives, outSumSecondDerivatives);
    });
}

int beagleExtCalculateLogLikelihoodAndDerivatives(
    int instance, int rootBufferIndex, int categoryWeightsIndex, int stateFrequenciesIndex,
    int cumulativeScaleIndex, int rootPreBufferIndex, const BeagleOperation* operations,
    int operationCount, const int* postBufferIndices, const int* firstDerivativeIndices,
    const int* secondDerivativeIndices, double* outLogLikelihood, double* outSumFirstDerivatives,
    double* outSumSquaredFirstDerivatives, double* outSumSecondDerivatives) {
    if (!outLogLikelihood) return BEAGLE_ERROR_OUT_OF_RANGE;
    return withInstance(instance, [&](Instance* in) {
        return in->logLikelihoodAndDerivatives(
            rootBufferIndex, categoryWeightsIndex, stateFrequenciesIndex, cumulativeScaleIndex,
            rootPreBufferIndex, reinterpret_cast<const int*>(operations), operationCount,
            postBufferIndices, firstDerivativeIndices, secondDerivativeIndices, outLogLikelihood,
            outSumFirstDerivatives, outSumSquaredFirstDerivatives, outSumSecondDerivatives);
    });
}

int beagleExtCalculateLogLikelihoodAndGradient(
    int instance, int rootBufferIndex, int categoryWeightsIndex, int stateFrequenciesIndex,
    int cumulativeScaleIndex, int rootPreBufferIndex, const BeagleOperation* operations,
    int operationCount, const int* postBufferIndices, const int* firstDerivativeIndices,
    const int* secondDerivativeIndices, int eigenIndex, double* outLogLikelihood,
    double* outSumFirstDerivatives, double* outSumSquaredFirstDerivatives,
    double* outSumSecondDerivatives, double* outGradientRateMatrix,
    double* outGradientFrequencies) {
    // every output NULL: enqueue only (checked by the instance)
    const bool none = !outLogLikelihood && !outSumFirstDerivatives &&
                      !outSumSquaredFirstDerivatives && !outSumSecondDerivatives &&
                      !outGradientRateMatrix && !outGradientFrequencies;
    if (!outLogLikelihood && !none) return BEAGLE_ERROR_OUT_OF_RANGE;
    return withInstance(instance, [&](Instance* in) {
        return in->logLikelihoodAndGradient(
            rootBufferIndex, categoryWeightsIndex, stateFrequenciesIndex, cumulativeScaleIndex,
            rootPreBufferIndex, reinterpret_cast<const int*>(operations), operationCount,
            postBufferIndices, firstDerivativeIndices, secondDerivativeIndices, eigenIndex,
            outLogLikelihood, outSumFirstDerivatives, outSumSquaredFirstDerivatives,
            outSumSecondDerivatives, outGradientRateMatrix, outGradientFrequencies);
    });
}

int beagleExtCalculateRateMatrixGradient(int instance, int eigenIndex,
                                          const BeagleOperation* operations, int operationCount,
                                          const int* postBufferIndices, int categoryWeightsIndex,
                                          int rootBufferIndex, int stateFrequenciesIndex,
                                          double* outGradientRateMatrix,
                                          double* outGradientFrequencies) {
    return withInstance(instance, [&](Instance* in) {
        return in->rateMatrixGradient(eigenIndex, reinterpret_cast<const int*>(operations),
                                      operationCount, postBufferIndices, categoryWeightsIndex,
                                      rootBufferIndex, stateFrequenciesIndex,
                                      outGradientRateMatrix, outGradientFrequencies);
    });
}

int beagleExtGetStream(int instance, void** outStream) {
    return withInstance(instance, [&](Instance* in) {
        *outStream = reinterpret_cast<void*>(in->stream());
        return BEAGLE_SUCCESS;
    });
}

int beagleExtSetSplitBF16(int instance, int terms) {
    return withInstance(instance, [&](Instance* in) { return in->setSplitBf16(terms); });
}

int beagleExtSetAutotune(int instance, int on) {
    return withInstance(instance, [&](Instance* in) { return in->setAutotune(on); });
}

int beagleExtGetAutotuneStats(int instance, int* outTuned, int* outPending) {
    return withInstance(instance, [&](Instance* in) { return in->autotuneStats(outTuned, outPending); });
}

int beagleExtPlanCheckpoints(int tipCount, const BeagleOperation* operations, int count,
                             int bufferCount, int interval, int tipPreorderBuffers,
                             int* outInterval, int* outPeakBuffers, int* outCheckpoints,
                             int* outRecomputed) {
    return guarded([&]() {
        return hipbeagle::planCheckpoints(tipCount, reinterpret_cast<const int*>(operations), count,
                               bufferCount, interval, tipPreorderBuffers, outInterval,
                               outPeakBuffers, outCheckpoints, outRecomputed);
    });
}

int beagleExtCheckpointedGradient(int instance, const BeagleOperation* operations, int count,
                                  const int* firstDerivativeIndices,
                                  const int* secondDerivativeIndices,
                                  int categoryWeightsIndex, int stateFrequenciesIndex,
                                  const int* bufferIndices, int bufferCount, int interval,
                                  int* outInterval, int* outPeakBuffers, double* outLogL,
                                  double* outSumFirstDerivatives,
                                  double* outSumSquaredFirstDerivatives,
                                  double* outSumSecondDerivatives) {
    return withInstance(instance, [&](Instance* in) {
        return in->checkpointedGradient(reinterpret_cast<const int*>(operations), count,
                                        firstDerivativeIndices, secondDerivativeIndices,
                                        categoryWeightsIndex, stateFrequenciesIndex,
                                        bufferIndices, bufferCount, interval, outInterval,
                                        outPeakBuffers, outLogL, outSumFirstDerivatives,
                                        outSumSquaredFirstDerivatives, outSumSecondDerivatives);
    });
}

int beagleExtBeginCapture(int instance) {
    return withInstance(instance, [&](Instance* in) { return in->beginCapture(); });
}

int beagleExtEndCapture(int instance, int* outGraph) {
    if (!outGraph) return BEAGLE_ERROR_OUT_OF_RANGE;
    return withInstance(instance, [&](Instance* in) { return in->endCapture(outGraph); });
}

int beagleExtLaunchGraph(int instance, int graph) {
    return withInstance(instance, [&](Instance* in) { return in->launchGraph(graph); });
}

int beagleExtWaitGraph(int instance, int graph) {
    return withInstance(instance, [&](Instance* in) { return in->waitGraph(graph); });
}

int beagleExtSetGraphBranchLengths(int instance, int graph, const double* lengths, int count) {
    if (!lengths && count) return BEAGLE_ERROR_OUT_OF_RANGE;
    return withInstance(instance, [&](Instance* in) {
        return in->setGraphBranchLengths(graph, lengths, count);
    });
}

int beagleExtDestroyGraph(int instance, int graph) {
    return withInstance(instance, [&](Instance* in) { return in->destroyGraph(graph); });
}

}  // extern "C"
#pragma GCC visibility pop
