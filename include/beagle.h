/*
 * beagle.h -- C ABI of the MI355X (gfx950) BEAGLE plugin library.
 *
 * The library libhmsbeagle_amd.so is a drop-in replacement for the
 * reference's libhmsbeagle.so on the likelihood hot path: every function
 * below has the reference's name, signature, argument meaning and error
 * convention, so a client compiled against the reference header links
 * against this library unchanged.  Enumerator values and struct layouts are
 * identical because they are part of that binary interface (reference:
 * libhmsbeagle/beagle.h).  Each declaration cites the reference declaration
 * it replaces as  [ref beagle.h:LINE].
 *
 * Arrays passed in or out are owned by the caller and copied synchronously
 * (host memory, always double precision at the boundary whatever the
 * internal precision).  Return values >= 0 are success (beagleCreateInstance
 * returns the instance number); negative values are BeagleReturnCodes.
 * No C++ exception crosses this boundary.
 */
#ifndef BEAGLE_AMD_BEAGLE_H
#define BEAGLE_AMD_BEAGLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes [ref beagle.h:110-122] */
enum BeagleReturnCodes {
    BEAGLE_SUCCESS                      =  0,
    BEAGLE_ERROR_GENERAL                = -1,
    BEAGLE_ERROR_OUT_OF_MEMORY          = -2,
    BEAGLE_ERROR_UNIDENTIFIED_EXCEPTION = -3,
    BEAGLE_ERROR_UNINITIALIZED_INSTANCE = -4,
    BEAGLE_ERROR_OUT_OF_RANGE           = -5,
    BEAGLE_ERROR_NO_RESOURCE            = -6,
    BEAGLE_ERROR_NO_IMPLEMENTATION      = -7,
    BEAGLE_ERROR_FLOATING_POINT         = -8
};

/* Capability flags, bits of a 64-bit long [ref beagle.h:132-177].
 * THREADING_CPP and PREORDER_TRANSPOSE_MANUAL share bit 30 in the
 * reference, and PREORDER_TRANSPOSE_AUTO is bit 31 (negative as an int);
 * both quirks are kept because clients pass these exact values. */
enum BeagleFlags {
    BEAGLE_FLAG_PRECISION_SINGLE    = 1 << 0,
    BEAGLE_FLAG_PRECISION_DOUBLE    = 1 << 1,
    BEAGLE_FLAG_COMPUTATION_SYNCH   = 1 << 2,
    BEAGLE_FLAG_COMPUTATION_ASYNCH  = 1 << 3,
    BEAGLE_FLAG_EIGEN_REAL          = 1 << 4,
    BEAGLE_FLAG_EIGEN_COMPLEX       = 1 << 5,
    BEAGLE_FLAG_SCALING_MANUAL      = 1 << 6,
    BEAGLE_FLAG_SCALING_AUTO        = 1 << 7,
    BEAGLE_FLAG_SCALING_ALWAYS      = 1 << 8,
    BEAGLE_FLAG_SCALING_DYNAMIC     = 1 << 25,
    BEAGLE_FLAG_SCALERS_RAW         = 1 << 9,
    BEAGLE_FLAG_SCALERS_LOG         = 1 << 10,
    BEAGLE_FLAG_INVEVEC_STANDARD    = 1 << 20,
    BEAGLE_FLAG_INVEVEC_TRANSPOSED  = 1 << 21,
    BEAGLE_FLAG_VECTOR_SSE          = 1 << 11,
    BEAGLE_FLAG_VECTOR_AVX          = 1 << 24,
    BEAGLE_FLAG_VECTOR_NONE         = 1 << 12,
    BEAGLE_FLAG_THREADING_CPP       = 1 << 30,
    BEAGLE_FLAG_THREADING_OPENMP    = 1 << 13,
    BEAGLE_FLAG_THREADING_NONE      = 1 << 14,
    BEAGLE_FLAG_PROCESSOR_CPU       = 1 << 15,
    BEAGLE_FLAG_PROCESSOR_GPU       = 1 << 16,
    BEAGLE_FLAG_PROCESSOR_FPGA      = 1 << 17,
    BEAGLE_FLAG_PROCESSOR_CELL      = 1 << 18,
    BEAGLE_FLAG_PROCESSOR_PHI       = 1 << 19,
    BEAGLE_FLAG_PROCESSOR_OTHER     = 1 << 26,
    BEAGLE_FLAG_FRAMEWORK_CUDA      = 1 << 22,
    BEAGLE_FLAG_FRAMEWORK_OPENCL    = 1 << 23,
    BEAGLE_FLAG_FRAMEWORK_CPU       = 1 << 27,
    BEAGLE_FLAG_PARALLELOPS_STREAMS = 1 << 28,
    BEAGLE_FLAG_PARALLELOPS_GRID    = 1 << 29,
    BEAGLE_FLAG_PREORDER_TRANSPOSE_MANUAL = 1 << 30,
    BEAGLE_FLAG_PREORDER_TRANSPOSE_AUTO   = 1 << 31
};

/* Resource-type bit of this library's HIP resources.  Bits 0-31 are all
 * taken by the reference, so it sits above them; the reference's scoring
 * (beagle.cpp:341-351) only counts bits 0-31, so it filters but never
 * scores.  No counterpart in the reference. */
#define BEAGLE_FLAG_FRAMEWORK_HIP (1L << 32)

/* Benchmark flags [ref beagle.h:187-191] */
enum BeagleBenchmarkFlags {
    BEAGLE_BENCHFLAG_SCALING_NONE    = 1 << 0,
    BEAGLE_BENCHFLAG_SCALING_ALWAYS  = 1 << 1,
    BEAGLE_BENCHFLAG_SCALING_DYNAMIC = 1 << 2
};

/* Operation codes [ref beagle.h:200-205] */
enum BeagleOpCodes {
    BEAGLE_OP_COUNT           = 7,
    BEAGLE_PARTITION_OP_COUNT = 9,
    BEAGLE_OP_NONE            = -1
};

/* [ref beagle.h:210-219] */
typedef struct {
    int   resourceNumber;
    char* resourceName;
    char* implName;
    char* implDescription;
    long  flags;
} BeagleInstanceDetails;

/* [ref beagle.h:224-229] */
typedef struct {
    char* name;
    char* description;
    long  supportFlags;
    long  requiredFlags;
} BeagleResource;

/* [ref beagle.h:234-237] */
typedef struct {
    BeagleResource* list;
    int length;
} BeagleResourceList;

/* [ref beagle.h:242-255] */
typedef struct {
    int    number;
    char*  name;
    char*  description;
    long   supportFlags;
    long   requiredFlags;
    int    returnCode;
    char*  implName;
    long   benchedFlags;
    double benchmarkResult;
    double performanceRatio;
} BeagleBenchmarkedResource;

/* [ref beagle.h:260-263] */
typedef struct {
    BeagleBenchmarkedResource* list;
    int length;
} BeagleBenchmarkedResourceList;

/* Post-order / pre-order operation: 7 ints [ref beagle.h:870-878].
 * For pre-order the fields mean [dest pre-order, write scale, read scale,
 * parent pre-order, P(t) of dest's branch, sibling post-order, P(t) of the
 * sibling's branch] (reference caller examples/hmctest/hmctest.cpp:497-508). */
typedef struct {
    int destinationPartials;
    int destinationScaleWrite;
    int destinationScaleRead;
    int child1Partials;
    int child1TransitionMatrix;
    int child2Partials;
    int child2TransitionMatrix;
} BeagleOperation;

/* Partitioned operation: 9 ints [ref beagle.h:922-932] */
typedef struct {
    int destinationPartials;
    int destinationScaleWrite;
    int destinationScaleRead;
    int child1Partials;
    int child1TransitionMatrix;
    int child2Partials;
    int child2TransitionMatrix;
    int partition;
    int cumulativeScaleIndex;
} BeagleOperationByPartition;

/* ---- library and instances ------------------------------------------- */

/* [ref beagle.h:280] */
const char* beagleGetVersion(void);
/* [ref beagle.h:290] */
const char* beagleGetCitation(void);
/* Resource 0 is the host CPU (listed so resource numbers match the
 * reference's, where 0 is the CPU; no implementation here runs on it);
 * resources 1..n are the visible gfx950 GPUs. [ref beagle.h:300] */
BeagleResourceList* beagleGetResourceList(void);
/* Returns NULL: resource benchmarking is not provided. [ref beagle.h:333] */
BeagleBenchmarkedResourceList* beagleGetBenchmarkedResourceList(
    int tipCount, int compactBufferCount, int stateCount, int patternCount,
    int categoryCount, int* resourceList, int resourceCount,
    long preferenceFlags, long requirementFlags, int eigenModelCount,
    int partitionCount, int calculateDerivatives, long benchmarkFlags);
/* [ref beagle.h:378-391] */
int beagleCreateInstance(int tipCount, int partialsBufferCount,
                         int compactBufferCount, int stateCount,
                         int patternCount, int eigenBufferCount,
                         int matrixBufferCount, int categoryCount,
                         int scaleBufferCount, int* resourceList,
                         int resourceCount, long preferenceFlags,
                         long requirementFlags,
                         BeagleInstanceDetails* returnInfo);
/* [ref beagle.h:402] */
int beagleFinalizeInstance(int instance);
/* [ref beagle.h:412] */
int beagleFinalize(void);
/* No effect on GPU instances (as in the reference). [ref beagle.h:429] */
int beagleSetCPUThreadCount(int instance, int threadCount);

/* ---- data in and out ------------------------------------------------- */

/* [ref beagle.h:446] */
int beagleSetTipStates(int instance, int tipIndex, const int* inStates);
/* [ref beagle.h:464] */
int beagleSetTipPartials(int instance, int tipIndex, const double* inPartials);
/* [ref beagle.h:481] */
int beagleSetRootPrePartials(const int instance, const int* bufferIndices,
                             const int* stateFrequenciesIndices, int count);
/* [ref beagle.h:498] */
int beagleSetPartials(int instance, int bufferIndex, const double* inPartials);
/* [ref beagle.h:515] */
int beagleGetPartials(int instance, int bufferIndex, int scaleIndex,
                      double* outPartials);
/* [ref beagle.h:534] */
int beagleSetEigenDecomposition(int instance, int eigenIndex,
                                const double* inEigenVectors,
                                const double* inInverseEigenVectors,
                                const double* inEigenValues);
/* [ref beagle.h:551] */
int beagleSetStateFrequencies(int instance, int stateFrequenciesIndex,
                              const double* inStateFrequencies);
/* [ref beagle.h:566] */
int beagleSetCategoryWeights(int instance, int categoryWeightsIndex,
                             const double* inCategoryWeights);
/* [ref beagle.h:580] */
int beagleSetCategoryRates(int instance, const double* inCategoryRates);
/* [ref beagle.h:594] */
int beagleSetCategoryRatesWithIndex(int instance, int categoryRatesIndex,
                                    const double* inCategoryRates);
/* [ref beagle.h:608] */
int beagleSetPatternWeights(int instance, const double* inPatternWeights);
/* [ref beagle.h:623] */
int beagleSetPatternPartitions(int instance, int partitionCount,
                               const int* inPatternPartitions);

/* ---- transition matrices --------------------------------------------- */

/* [ref beagle.h:661] */
int beagleConvolveTransitionMatrices(int instance, const int* firstIndices,
                                     const int* secondIndices,
                                     const int* resultIndices, int matrixCount);
/* [ref beagle.h:681] */
int beagleAddTransitionMatrices(int instance, const int* firstIndices,
                                const int* secondIndices,
                                const int* resultIndices, int matrixCount);
/* [ref beagle.h:699] */
int beagleTransposeTransitionMatrices(int instance, const int* inputIndices,
                                      const int* resultIndices, int matrixCount);
/* [ref beagle.h:723-729] */
int beagleUpdateTransitionMatrices(int instance, int eigenIndex,
                                   const int* probabilityIndices,
                                   const int* firstDerivativeIndices,
                                   const int* secondDerivativeIndices,
                                   const double* edgeLengths, int count);
/* [ref beagle.h:753] */
int beagleUpdateTransitionMatricesWithModelCategories(
    int instance, int* eigenIndices, const int* probabilityIndices,
    const int* firstDerivativeIndices, const int* secondDerivativeIndices,
    const double* edgeLengths, int count);
/* [ref beagle.h:782] */
int beagleUpdateTransitionMatricesWithMultipleModels(
    int instance, const int* eigenIndices, const int* categoryRateIndices,
    const int* probabilityIndices, const int* firstDerivativeIndices,
    const int* secondDerivativeIndices, const double* edgeLengths, int count);
/* [ref beagle.h:806] */
int beagleSetTransitionMatrix(int instance, int matrixIndex,
                              const double* inMatrix, double paddedValue);
/* [ref beagle.h:823] */
int beagleSetDifferentialMatrix(int instance, int matrixIndex,
                                const double* inMatrix);
/* [ref beagle.h:840] */
int beagleGetTransitionMatrix(int instance, int matrixIndex, double* outMatrix);
/* [ref beagle.h:860] */
int beagleSetTransitionMatrices(int instance, const int* matrixIndices,
                                const double* inMatrices,
                                const double* paddedValues, int count);

/* ---- traversals ------------------------------------------------------ */

/* [ref beagle.h:894-897] */
int beagleUpdatePartials(const int instance, const BeagleOperation* operations,
                         int operationCount, int cumulativeScaleIndex);
/* [ref beagle.h:914-917] */
int beagleUpdatePrePartials(const int instance,
                            const BeagleOperation* operations,
                            int operationCount, int cumulativeScaleIndex);
/* [ref beagle.h:947] */
int beagleUpdatePartialsByPartition(const int instance,
                                    const BeagleOperationByPartition* operations,
                                    int operationCount);
/* [ref beagle.h:964] */
int beagleUpdatePrePartialsByPartition(
    const int instance, const BeagleOperationByPartition* operations,
    int operationCount);
/* [ref beagle.h:984] */
int beagleWaitForPartials(const int instance, const int* destinationPartials,
                          int destinationPartialsCount);

/* ---- scale buffers --------------------------------------------------- */

/* [ref beagle.h:999] */
int beagleAccumulateScaleFactors(int instance, const int* scaleIndices,
                                 int count, int cumulativeScaleIndex);
/* [ref beagle.h:1016] */
int beagleAccumulateScaleFactorsByPartition(int instance,
                                            const int* scaleIndices, int count,
                                            int cumulativeScaleIndex,
                                            int partitionIndex);
/* [ref beagle.h:1033] */
int beagleRemoveScaleFactors(int instance, const int* scaleIndices, int count,
                             int cumulativeScaleIndex);
/* [ref beagle.h:1052] */
int beagleRemoveScaleFactorsByPartition(int instance, const int* scaleIndices,
                                        int count, int cumulativeScaleIndex,
                                        int partitionIndex);
/* [ref beagle.h:1067] */
int beagleResetScaleFactors(int instance, int cumulativeScaleIndex);
/* [ref beagle.h:1080] */
int beagleResetScaleFactorsByPartition(int instance, int cumulativeScaleIndex,
                                       int partitionIndex);
/* [ref beagle.h:1094] */
int beagleCopyScaleFactors(int instance, int destScalingIndex,
                           int srcScalingIndex);
/* Copies the scale buffer out (the reference CPU path leaves the output
 * untouched; the GPU path copies it). [ref beagle.h:1107] */
int beagleGetScaleFactors(int instance, int srcScalingIndex,
                          double* outScaleFactors);

/* ---- likelihoods and derivatives ------------------------------------- */

/* [ref beagle.h:1133-1139] */
int beagleCalculateRootLogLikelihoods(int instance, const int* bufferIndices,
                                      const int* categoryWeightsIndices,
                                      const int* stateFrequenciesIndices,
                                      const int* cumulativeScaleIndices,
                                      int count, double* outSumLogLikelihood);
/* [ref beagle.h:1167] */
int beagleCalculateRootLogLikelihoodsByPartition(
    int instance, const int* bufferIndices, const int* categoryWeightsIndices,
    const int* stateFrequenciesIndices, const int* cumulativeScaleIndices,
    const int* partitionIndices, int partitionCount, int count,
    double* outSumLogLikelihoodByPartition, double* outSumLogLikelihood);
/* Deprecated in the reference; returns BEAGLE_ERROR_NO_IMPLEMENTATION
 * there and here. [ref beagle.h:1204, beagle.cpp:1856-1864] */
int beagleCalculateEdgeDerivative(int instance, const int* postBufferIndices,
                                  const int* preBufferIndices,
                                  const int rootBufferIndex,
                                  const int* firstDerivativeIndices,
                                  const int* secondDerivativeIndices,
                                  const int categoryWeightsIndex,
                                  const int categoryRatesIndex,
                                  const int stateFrequenciesIndex,
                                  const int* cumulativeScaleIndices, int count,
                                  double* outFirstDerivative,
                                  double* outDiagonalSecondDerivative);
/* Per node n and pattern p: ratio = (pre . D post) / (pre . post), mixing
 * categories with their weights inside numerator and denominator.
 * outDerivatives[n*patternCount+p] = ratio (if non-NULL),
 * outSumDerivatives[n] = sum_p w_p ratio, outSumSquaredDerivatives[n] =
 * sum_p w_p ratio^2.  [ref beagle.h:1218-1226, beagle.cpp:1794-1824] */
int beagleCalculateEdgeDerivatives(int instance, const int* postBufferIndices,
                                   const int* preBufferIndices,
                                   const int* derivativeMatrixIndices,
                                   const int* categoryWeightsIndices, int count,
                                   double* outDerivatives,
                                   double* outSumDerivatives,
                                   double* outSumSquaredDerivatives);
/* Adds into outSumDerivatives[k*S+j] the stateCount^2 cross products
 * sum_n sum_p (w_p / L_p) sum_c w_c r_c t_n pre[c,p,k] post[c,p,j].
 * [ref beagle.h:1228-1236, beagle.cpp:1826-1854] */
int beagleCalculateCrossProductDerivative(int instance,
                                          const int* postBufferIndices,
                                          const int* preBufferIndices,
                                          const int* categoryRatesIndices,
                                          const int* categoryWeightsIndices,
                                          const double* edgeLengths, int count,
                                          double* outSumDerivatives,
                                          double* outSumSquaredDerivatives);
/* [ref beagle.h:1265-1277] */
int beagleCalculateEdgeLogLikelihoods(int instance,
                                      const int* parentBufferIndices,
                                      const int* childBufferIndices,
                                      const int* probabilityIndices,
                                      const int* firstDerivativeIndices,
                                      const int* secondDerivativeIndices,
                                      const int* categoryWeightsIndices,
                                      const int* stateFrequenciesIndices,
                                      const int* cumulativeScaleIndices,
                                      int count, double* outSumLogLikelihood,
                                      double* outSumFirstDerivative,
                                      double* outSumSecondDerivative);
/* [ref beagle.h:1314] */
int beagleCalculateEdgeLogLikelihoodsByPartition(
    int instance, const int* parentBufferIndices, const int* childBufferIndices,
    const int* probabilityIndices, const int* firstDerivativeIndices,
    const int* secondDerivativeIndices, const int* categoryWeightsIndices,
    const int* stateFrequenciesIndices, const int* cumulativeScaleIndices,
    const int* partitionIndices, int partitionCount, int count,
    double* outSumLogLikelihoodByPartition, double* outSumLogLikelihood,
    double* outSumFirstDerivativeByPartition, double* outSumFirstDerivative,
    double* outSumSecondDerivativeByPartition, double* outSumSecondDerivative);
/* [ref beagle.h:1350] */
int beagleGetLogLikelihood(int instance, double* outSumLogLikelihood);
/* [ref beagle.h:1368] */
int beagleGetDerivatives(int instance, double* outSumFirstDerivative,
                         double* outSumSecondDerivative);
/* [ref beagle.h:1383] */
int beagleGetSiteLogLikelihoods(int instance, double* outLogLikelihoods);
/* [ref beagle.h:1397] */
int beagleGetSiteDerivatives(int instance, double* outFirstDerivatives,
                             double* outSecondDerivatives);

#ifdef __cplusplus
}
#endif

#endif /* BEAGLE_AMD_BEAGLE_H */
