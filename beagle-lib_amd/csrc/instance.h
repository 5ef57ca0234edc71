// One BEAGLE instance on one MI355X: owns its HBM buffers and HIP stream,
// turns API calls into job lists and launches the gfx950 kernels.
//
// Mirrors the per-instance operator interface of the reference
// (libhmsbeagle/BeagleImpl.h:31-268); each method returns a BEAGLE return
// code and never throws past the front end.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.h"

namespace hipbeagle {

// Buffers an operation reads and writes (partials buffer indices, scale
// buffers offset by the buffer count; -1 unused), for dependency levels.
struct Access {
    int reads[3];
    int writes[2];
};

// Eigen-system kept on the host (for the rate-matrix gradient).
struct HostEigen {
    std::vector<double> V, Vi, lam;   // row-major V, V^-1; eigenvalues
};

// How a transition matrix was last produced: from eigen-system `eigen`
// with category-rate buffer `rates` and length `t` (eigen < 0: set by the
// client, transposed, convolved, or from several eigen-systems).
struct MatrixOrigin {
    int eigen = -1;
    int rates = 0;
    double t = 0.0;
};

// Operations launched together: batched (one dependency level, grid over
// pattern tiles x operations) or fused (operations walked in list order by
// each pattern tile).
struct Segment {
    bool batched;
    std::vector<int> ops;
};

// The checkpoint schedule of beagleExtCheckpointedGradient for a tree,
// without an instance (beagleExtPlanCheckpoints).
int planCheckpoints(int tipCount, const int* ops, int count, int bufferCount, int interval,
                    int tipPreorder, int* outInterval, int* outPeak, int* outCheckpoints,
                    int* outRecomputed);

class Instance {
public:
    Instance(int tipCount, int partialsBufferCount, int compactBufferCount, int stateCount,
             int patternCount, int eigenBufferCount, int matrixBufferCount, int categoryCount,
             int scaleBufferCount, int device, int resourceNumber, long preferenceFlags,
             long requirementFlags, bool doublePrecision);
    ~Instance();

    long flags() const { return flags_; }
    const char* implName() const;
    int resourceNumber() const { return resource_; }

    int setTipStates(int tipIndex, const int* inStates);
    int setTipPartials(int tipIndex, const double* inPartials);
    int setRootPrePartials(const int* bufferIndices, const int* freqIndices, int count);
    int setPartials(int bufferIndex, const double* inPartials);
    int getPartials(int bufferIndex, int scaleIndex, double* outPartials);
    int setEigenDecomposition(int eigenIndex, const double* V, const double* Vinv,
                              const double* evals);
    int setStateFrequencies(int index, const double* f);
    int setCategoryWeights(int index, const double* w);
    int setCategoryRates(const double* r);
    int setCategoryRatesWithIndex(int index, const double* r);
    int setPatternWeights(const double* w);
    int setPatternPartitions(int partitionCount, const int* parts);

    int updateTransitionMatrices(int eigenIndex, const int* probIdx, const int* d1Idx,
                                 const int* d2Idx, const double* lengths, int count);
    int updateTransitionMatricesWithModelCategories(const int* eigenIndices, const int* probIdx,
                                                    const int* d1Idx, const int* d2Idx,
                                                    const double* lengths, int count);
    int updateTransitionMatricesWithMultipleModels(const int* eigenIdx, const int* rateIdx,
                                                   const int* probIdx, const int* d1Idx,
                                                   const int* d2Idx, const double* lengths,
                                                   int count);
    int setTransitionMatrix(int index, const double* in, double padValue);
    int setDifferentialMatrix(int index, const double* in);
    int setTransitionMatrices(const int* idx, const double* in, const double* pads, int count);
    int getTransitionMatrix(int index, double* out);
    int transposeTransitionMatrices(const int* in, const int* out, int count);
    int convolveTransitionMatrices(const int* a, const int* b, const int* out, int count);
    int addTransitionMatrices(const int* a, const int* b, const int* out, int count);

    int updatePartials(const int* ops, int count, int cumulativeIndex, bool byPartition);
    int updatePrePartials(const int* ops, int count, int cumulativeIndex, bool byPartition);
    int updatePrePartialsWithDerivatives(const int* ops, int count, const int* postIdx,
                                         const int* d1Idx, const int* d2Idx, int wIdx,
                                         double* outD1, double* outD1Sq, double* outD2);
    int rateMatrixGradient(int eigenIdx, const int* ops, int count, const int* postIdx, int wIdx,
                           int rootIdx, int fIdx, double* outQ, double* outFreqs);
    int rootLogLikelihoodsMulti(const int* bufIdx, const int* wIdx, const int* fIdx,
                                const int* cumIdx, int count, double* out);
    int rootLogLikelihoodsPerCategory(int b, int f, int cum, double* out);
    int rateMatrixGradientGeneric(int eigenIdx, const int* ops, int count, const int* postIdx,
                                  int wIdx, int rootIdx, int fIdx, double* outQ,
                                  double* outFreqs);
    int logLikelihoodAndDerivatives(int rootIdx, int wIdx, int fIdx, int cumIdx, int rootPreIdx,
                                    const int* ops, int count, const int* postIdx,
                                    const int* d1Idx, const int* d2Idx, double* outLogL,
                                    double* outD1, double* outD1Sq, double* outD2);
    int logLikelihoodAndGradient(int rootIdx, int wIdx, int fIdx, int cumIdx, int rootPreIdx,
                                 const int* ops, int count, const int* postIdx, const int* d1Idx,
                                 const int* d2Idx, int eigenIdx, double* outLogL, double* outD1,
                                 double* outD1Sq, double* outD2, double* outQ,
                                 double* outFreqs);
    int checkpointedGradient(const int* ops, int count, const int* d1Idx, const int* d2Idx,
                             int wIdx, int fIdx, const int* pool, int poolCount, int interval,
                             int* outInterval, int* outPeak, double* outLogL, double* outD1,
                             double* outD1Sq, double* outD2);
    int waitForPartials();

    int accumulateScaleFactors(const int* idx, int count, int cumIdx, double sign, int part);
    int resetScaleFactors(int cumIdx, int part);
    int copyScaleFactors(int dst, int src);
    int getScaleFactors(int src, double* out);

    int calculateRootLogLikelihoods(const int* bufIdx, const int* wIdx, const int* fIdx,
                                    const int* cumIdx, int count, double* out);
    // calculateRootLogLikelihoods (one root) with its workgroup sums and
    // result in the GPU's fixed scratch array (noalloc.hip): no allocation
    int calculateRootLogLikelihoodsNoAlloc(const int* bufIdx, const int* wIdx, const int* fIdx,
                                           const int* cumIdx, int count, double* out);
    int calculateRootLogLikelihoodsByPartition(const int* bufIdx, const int* wIdx,
                                               const int* fIdx, const int* cumIdx,
                                               const int* partIdx, int partitionCount, int count,
                                               double* outByPart, double* outSum);
    int calculateEdgeDerivatives(const int* postIdx, const int* preIdx, const int* dIdx,
                                 const int* wIdx, int count, double* outD, double* outSum,
                                 double* outSumSq);
    int calculateCrossProducts(const int* postIdx, const int* preIdx, const int* rIdx,
                               const int* wIdx, const double* lengths, int count, double* out);
    int calculateEdgeLogLikelihoods(const int* parIdx, const int* childIdx, const int* pIdx,
                                    const int* d1Idx, const int* d2Idx, const int* wIdx,
                                    const int* fIdx, const int* cumIdx, int count, double* outL,
                                    double* outD1, double* outD2);
    int calculateEdgeLogLikelihoodsByPartition(
        const int* parIdx, const int* childIdx, const int* pIdx, const int* d1Idx,
        const int* d2Idx, const int* wIdx, const int* fIdx, const int* cumIdx,
        const int* partIdx, int partitionCount, int count, double* outLByPart, double* outL,
        double* outD1ByPart, double* outD1, double* outD2ByPart, double* outD2);
    int getLogLikelihood(double* out);
    int getDerivatives(double* d1, double* d2);
    int getSiteLogLikelihoods(double* out);
    int getSiteDerivatives(double* d1, double* d2);

    // Device pointers for the extension entry points (Q-parameter
    // gradients, fused evaluations).
    // HIP-graph capture of this instance's calls (beagle_amd_ext.h)
    int beginCapture();
    int endCapture(int* outGraph);
    int launchGraph(int graph);
    int waitGraph(int graph);
    int setGraphBranchLengths(int graph, const double* lengths, int count);
    int destroyGraph(int graph);

    const Dims& dims() const { return dims_; }
    hipStream_t stream() const { return stream_; }
    bool doublePrecision() const { return dp_; }

private:
    struct Slot {
        void* host = nullptr;
        void* dev = nullptr;
        size_t cap = 0;
        hipEvent_t ev = nullptr;
        bool used = false;
    };

    void release();
    void* upload(const void* data, size_t bytes);
    // Job lists that depend only on the operations (and the buffers they
    // name): an identical list reuses its device copy, no transfer.
    void* uploadJobs(const void* data, size_t bytes);
    // SYNCH instances wait at the end of an update call (not while capturing)
    void finishCall();

    // A captured graph: its uploads and results live in its own pinned host
    // and device arena (memcpy nodes read and write them at every replay);
    // `outs` are copied to the client's arrays after each replay.
    struct CapturedOut {
        double* dst;
        size_t src;      // byte offset of the first value in the host arena
        int n, stride;   // dst[k] = value k * stride
        bool logL;       // a log-likelihood: NaN gives BEAGLE_ERROR_FLOATING_POINT
    };
    struct Graph {
        hipGraphExec_t exec = nullptr;
        char* host = nullptr;      // pinned, coherent: the replays' outputs and lengths
        char* hostDev = nullptr;   // the same memory as the device addresses it
        char* dev = nullptr;       // the replays' job lists (copied once, at the end)
        size_t used = 0;
        std::vector<CapturedOut> outs;
        std::vector<std::pair<size_t, int>> matJobs;   // (host offset, count) of matrix jobs
        bool live = false;
    };
    static constexpr size_t kCaptureArena = size_t(4) << 20;
    size_t captureAlloc(size_t bytes);
    size_t captureResult(const double* dsrc, size_t n);
    void captureBreak(const char* why);
    std::vector<Graph> graphs_;
    std::vector<void*> retired_;      // grown-out buffers a live graph may still use
    Graph cap_;
    bool capturing_ = false;
    std::string captureError_;
    size_t lastUploadOff_ = 0;        // arena offset of the last upload while capturing
    void* grow(void*& buf, size_t& cap, size_t bytes);
    void ensurePartials(int b);
    void clearExponent(int b);
    void* partials(int b) const { return dPartials_[b]; }
    void* matrix(int m) const { return (char*)dMatrices_ + (size_t)m * matrixBytes_; }
    double* scale(int s) const { return dScale_ + (size_t)s * dims_.Ppad; }
    size_t realSize() const { return dp_ ? sizeof(double) : sizeof(float); }
    void toDevice(void* dst, const double* src, size_t n);   // double -> instance precision
    void toHost(double* dst, const void* src, size_t n);
    void copyIn(void* dst, const void* src, size_t bytes);
    void copyOut(void* dst, const void* src, size_t bytes);
    void sync();
    int matrixUpdate(const std::vector<MatrixJob>& jobs);
    int buildPreJobs(const int* ops, int count, bool byPartition, bool allowNoDest);
    int attachOwnOperands(int count, const int* postIdx, const int* d1Idx, const int* d2Idx);
    void launchPre(const int* ops, int stride, int count, double* partial);
    void launchPostPlan(const std::vector<Segment>& plan);
    // S > 4, FP32: matrix repacking, then one MFMA launch per dependency level
    void launchPostMfma(const int* ops, int stride, int count);
    bool mfmaAdjoint() const;
    size_t packFloats() const;
    // tip branches of the fused gradient on the MFMA path: their derivative
    // sums formed by the pre-order kernel (rows of `partial`)
    struct TipDeriv {
        const int* postIdx;
        const int* d1;
        const int* d2;
        int wIdx;
        double* partial;
    };
    void launchPreMfma(const int* ops, int stride, int count, const TipDeriv* td = nullptr);
    int enqueueDerivMfma(const int* postIdx, const int* preIdx, const int* d1, const int* d2,
                         int wIdx, int count, double* sites, double** dSums,
                         const char* skip = nullptr, double* work = nullptr);
    int deliverDerivativeSums(const double* sums, int count, double* outD1, double* outD1Sq,
                              double* outD2);

public:
    // post-order matrix-core engine for S > 4 in FP32: 0 FP32, 3 or 4 split BF16
    int setSplitBf16(int terms);
    // measure-and-keep choice of launch plans per shape class (on by default)
    int setAutotune(int on);
    // shape classes tuned so far, and those whose timings are still pending
    int autotuneStats(int* outTuned, int* outPending);

private:
    // dependency levels of a post-order list (reads children and a read
    // scale buffer; writes the destination and a write scale buffer)
    int postLevels(const int* ops, int stride, int count, std::vector<int>& level);
    bool launchPreNodes(const int* ops, int count, const int* postIdx, double* partial,
                        NodeQGrad* q = nullptr);
    bool makeNodeJobs(const int* ops, int count, const int* postIdx, std::vector<NodeJob>& nodes,
                      std::vector<Access>& acc);
    int enqueueRoot(int b, int w, int f, int cum);
    // The rate-matrix gradient requested from the derivative pass: eigen
    // system in, device pointers out (fused false: the pass could not form
    // it, it must be computed separately)
    struct QGradPass {
        int eigen;
        bool fused;
        double *hacc, *slotPart, *rootRows, *sums, *out;
        long slots;   // H slots the pass wrote
    };
    int enqueuePreDerivatives(const int* ops, int count, const int* postIdx, const int* d1Idx,
                              const int* d2Idx, int wIdx, bool reduce, const double** dSums,
                              QGradPass* qg = nullptr);
    int checkQGradOperands(int e, const int* ops, int count, int wIdx);
    void reserveStage(size_t n);
    const double* stageOut(const double* dsrc, size_t n, size_t at);
    static constexpr int kNarrowLevel = 8;   // levels with fewer ops run fused
    bool validBuffer(int b) const { return b >= 0 && b < bufferCount_; }
    bool validMatrix(int m) const { return m >= 0 && m < matrixCount_; }
    bool validScale(int s) const { return s >= 0 && s < scaleCount_; }
    bool validEigen(int e) const { return e >= 0 && e < eigenCount_; }

    int tipCount_, bufferCount_, eigenCount_, matrixCount_, scaleCount_;
    int device_, resource_;
    long flags_;
    bool dp_;
    bool preTransposed_;   // S > 4 without PREORDER_TRANSPOSE_AUTO
    bool complex_ = false; // EIGEN_COMPLEX: 2 S eigenvalues, conjugate pairs as 2 x 2 blocks
    bool always_;          // SCALING_ALWAYS
    bool auto_;            // SCALING_AUTO: power-of-two exponents per buffer
    bool logScalers_;
    int preWeights_ = 0;      // category-weights buffer of the fused derivative pass
    int traversalMode_ = 0;   // 0 automatic, 1 always per dependency level, 2 always fused
    bool splitBf16_ = false;  // S > 4 FP32 post-order on the BF16 matrix cores (hi/lo split)
    int bf16Terms_ = 3;       // products per split pair: 3, or 4 (adds lo x lo)
    int mfmaVariant_ = 0;
    bool mfmaVariantForced_ = false;
    // autotuning: chosen plan candidate per shape class (shapeKey); a
    // class being tuned holds the events around each candidate's timed run
    // until they have completed (read on a later call, never waited for)
    struct PlanChoice {
        bool fused;
        int narrow;
    };
    static constexpr int kNumPlans = 5;
    static const PlanChoice kPlans[kNumPlans];
    static constexpr size_t kMaxTuned = 64;   // shape classes kept per instance
    struct Tuning {
        int choice = -1;                                   // -1: timings pending
        std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;  // per candidate
    };
    bool autotune_ = true;
    bool tileLists_ = true;   // work lists for level launches over pattern ranges
    std::unordered_map<uint64_t, Tuning> tuned_;
    int tunings_ = 0;   // shape classes tuned so far
    static uint64_t shapeKey(uint64_t tag, int count, int levels);
    bool resolveTuning(Tuning& t);
    void clearTuning();
    int forcedPlan() const;
    template <typename F>
    int tunedChoice(uint64_t key, int ncand, int forced, F&& run);
    void launchNodePlan(const std::vector<NodeJob>& nodes, const std::vector<Segment>& plan,
                        double* partial, NodeQGrad* q = nullptr);
    void launchPrePlan(const std::vector<Segment>& plan, const double* cw, double* partial);     // MFMA post-order: 0 partials pipelined, 1 whole tile up front
    int levelChunk_ = 1;      // batched level: 1 one op per workgroup, k > 1 chunks of k ops
    Dims dims_;
    size_t partialsBytes_, matrixBytes_;
    hipStream_t stream_ = nullptr;

    std::vector<void*> dPartials_;
    std::vector<int*> dStates_;
    std::vector<short*> dExp_;        // SCALING_AUTO exponents per buffer
    void* dMatrices_ = nullptr;
    double* dScale_ = nullptr;
    double* dCube_ = nullptr;
    double* dVVi_ = nullptr;          // per eigen-system [V | V^-1] (S x S each)
    double* dEvals_ = nullptr;
    std::vector<bool> eigenSet_;
    std::vector<HostEigen> hEigen_;
    std::vector<MatrixOrigin> matOrigin_;
    double* dFreqs_ = nullptr;
    double* dCatWeights_ = nullptr;
    double* dCatRates_ = nullptr;
    std::vector<std::vector<double>> hFreqs_, hCatWeights_, hCatRates_;
    double* dPatternWeights_ = nullptr;
    std::vector<double> hPatternWeights_;
    double* dSiteL_ = nullptr;
    double* dSiteD1_ = nullptr;
    double* dSiteD2_ = nullptr;
    double* dResult_ = nullptr;       // small device result vector
    double* dF_ = nullptr;            // 4 real states: Frechet factors per matrix [C][4][4]
    double* noAllocScratch_ = nullptr;  // this device's scratch array (noalloc.hip)
    char* dDummy_ = nullptr;          // scratch tile for the pipelined traversal
    double* hResult_ = nullptr;       // pinned mirror
    void* dWork_ = nullptr;           // grow-only reduction scratch
    size_t workCap_ = 0;
    double* hStage_ = nullptr;        // grow-only pinned staging for results
    size_t stageCap_ = 0;             // doubles
    void* dJobMats_ = nullptr;        // grow-only gathered matrices of node jobs
    size_t jobMatsCap_ = 0;
    int maxLdsPerBlock_ = 0;          // the device's LDS limit per workgroup (bytes)
    void* dSites_ = nullptr;          // grow-only per-site derivative ratios (MFMA path)
    size_t sitesCap_ = 0;
    float* dPack_ = nullptr;          // grow-only MFMA-packed matrices of a traversal
    size_t packCap_ = 0;
    void* dScratch_ = nullptr;        // C*Ppad*Sp reals (generic pre-order)
    std::vector<int> partStart_;
    std::vector<int> newOrder_;       // after a pattern reorder: internal index of pattern i
    int partitionCount_ = 1;
    std::vector<Slot> ring_;
    int ringPos_ = 0;
    // device copies of recent job lists (least recently used replaced);
    // an entry's event, recorded when the next list is asked for, covers
    // the kernels launched with it
    struct JobEntry {
        void* host = nullptr;   // pinned copy of the bytes (the transfer's source)
        void* dev = nullptr;
        size_t cap = 0, bytes = 0;
        hipEvent_t ev = nullptr;
        unsigned long long last = 0;
        bool pending = false;
    };
    std::vector<JobEntry> jobs_;
    int jobsLast_ = -1;
    unsigned long long jobsTick_ = 0;
    std::vector<PostJob> hPost_;
    std::vector<PreJob> hPre_;
    double lastLogL_ = 0.0, lastD1_ = 0.0, lastD2_ = 0.0;
    bool haveDerivs_ = false;
};

}  // namespace hipbeagle
