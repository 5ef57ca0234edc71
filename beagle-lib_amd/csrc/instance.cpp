// Host side of one MI355X BEAGLE instance.  See instance.h.
#include "instance.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <unordered_map>

#include "beagle.h"

namespace hipbeagle {

namespace {

void check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) throw std::bad_alloc();
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHECK(x) check((x), #x)

template <typename T>
T* dalloc(size_t n) {
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return reinterpret_cast<T*>(p);
}

int roundUp(int x, int m) { return (x + m - 1) / m * m; }

// Dependency level of each operation of a list run in order: one more than
// the latest earlier operation that writes a resource it reads or writes,
// or reads a resource it writes.  Resources are buffer indices (partials
// buffers, then scale buffers offset by the buffer count); -1 = unused.

int dependencyLevels(const std::vector<Access>& acc, int nres, std::vector<int>& level) {
    const int count = (int)acc.size();
    level.assign(count, 0);
    std::vector<int> lastWrite(nres, -1), lastRead(nres, -1);
    int nl = 0;
    for (int o = 0; o < count; ++o) {
        int lv = 0;
        for (int b : acc[o].reads)
            if (b >= 0 && lastWrite[b] >= 0) lv = std::max(lv, lastWrite[b] + 1);
        for (int b : acc[o].writes) {
            if (b < 0) continue;
            if (lastWrite[b] >= 0) lv = std::max(lv, lastWrite[b] + 1);
            if (lastRead[b] >= 0) lv = std::max(lv, lastRead[b] + 1);
        }
        level[o] = lv;
        for (int b : acc[o].writes)
            if (b >= 0) lastWrite[b] = lv;
        for (int b : acc[o].reads)
            if (b >= 0) lastRead[b] = std::max(lastRead[b], lv);
        nl = std::max(nl, lv + 1);
    }
    return nl;
}

// Launch plan: levels in order; a level with at least `narrow` operations
// becomes one batched launch over (tiles x ops); runs of consecutive
// narrower levels are merged into one fused launch that walks their
// operations in list order.
std::vector<Segment> planSegments(const std::vector<int>& level, int nl, int narrow) {
    std::vector<std::vector<int>> byLevel(nl);
    for (int o = 0; o < (int)level.size(); ++o) byLevel[level[o]].push_back(o);
    std::vector<Segment> plan;
    std::vector<int> pending;  // levels merged into the next fused segment
    auto flush = [&]() {
        if (pending.empty()) return;
        Segment s{false, {}};
        for (int o = 0; o < (int)level.size(); ++o)
            if (level[o] >= pending.front() && level[o] <= pending.back()) s.ops.push_back(o);
        plan.push_back(std::move(s));
        pending.clear();
    };
    for (int l = 0; l < nl; ++l) {
        if ((int)byLevel[l].size() >= narrow) {
            flush();
            plan.push_back(Segment{true, byLevel[l]});
        } else {
            pending.push_back(l);
        }
    }
    flush();
    return plan;
}

}  // namespace

// Scaling-mode precedence follows the reference CPU implementation:
// AUTO > ALWAYS > DYNAMIC > SCALERS_LOG (manual) > manual raw
// (libhmsbeagle/CPU/BeagleCPUImpl.hpp:294-311).  DYNAMIC is refused by the
// factory before this constructor runs.  AUTO is this plugin's own exact
// scheme: every operation divides its output by a power of two per pattern
// and stores the exponent with the buffer; exponents add up the tree, so
// the root (and any edge) log-likelihood adds ln2 * exponent and no client
// scale-buffer bookkeeping is needed.
Instance::Instance(int tipCount, int partialsBufferCount, int compactBufferCount, int stateCount,
                   int patternCount, int eigenBufferCount, int matrixBufferCount,
                   int categoryCount, int scaleBufferCount, int device, int resourceNumber,
                   long preferenceFlags, long requirementFlags, bool doublePrecision)
    : tipCount_(tipCount),
      bufferCount_(partialsBufferCount + compactBufferCount),
      eigenCount_(eigenBufferCount),
      matrixCount_(matrixBufferCount),
      scaleCount_(scaleBufferCount),
      device_(device),
      resource_(resourceNumber),
      dp_(doublePrecision) {
    if (stateCount < 2 || patternCount < 1 || categoryCount < 1 || tipCount < 1 ||
        bufferCount_ <= tipCount || eigenBufferCount < 1 || matrixBufferCount < 1 ||
        scaleBufferCount < 0)
        throw std::out_of_range("instance dimensions");
    // a failed allocation (or any other error) part way through releases
    // what was made so far before the error reaches the caller
    try {
        HIPCHECK(hipSetDevice(device_));
        const long both = preferenceFlags | requirementFlags;
        complex_ = (both & BEAGLE_FLAG_EIGEN_COMPLEX) != 0;
        flags_ = BEAGLE_FLAG_PROCESSOR_GPU | BEAGLE_FLAG_VECTOR_NONE | BEAGLE_FLAG_THREADING_NONE |
                 (complex_ ? BEAGLE_FLAG_EIGEN_COMPLEX : BEAGLE_FLAG_EIGEN_REAL) |
                 BEAGLE_FLAG_FRAMEWORK_HIP | BEAGLE_FLAG_PARALLELOPS_GRID;
        flags_ |= dp_ ? BEAGLE_FLAG_PRECISION_DOUBLE : BEAGLE_FLAG_PRECISION_SINGLE;
        flags_ |= (both & BEAGLE_FLAG_COMPUTATION_ASYNCH) ? BEAGLE_FLAG_COMPUTATION_ASYNCH
                                                          : BEAGLE_FLAG_COMPUTATION_SYNCH;
        flags_ |= (both & BEAGLE_FLAG_INVEVEC_TRANSPOSED) ? BEAGLE_FLAG_INVEVEC_TRANSPOSED
                                                          : BEAGLE_FLAG_INVEVEC_STANDARD;
        const int internal = bufferCount_ - tipCount_;
        always_ = false;
        auto_ = false;
        if (both & BEAGLE_FLAG_SCALING_AUTO) {
            auto_ = true;
            flags_ |= BEAGLE_FLAG_SCALING_AUTO | BEAGLE_FLAG_SCALERS_LOG;
        } else if (both & BEAGLE_FLAG_SCALING_ALWAYS) {
            always_ = true;
            flags_ |= BEAGLE_FLAG_SCALING_ALWAYS | BEAGLE_FLAG_SCALERS_LOG;
            scaleCount_ = internal + 1;
        } else if (both & BEAGLE_FLAG_SCALERS_LOG) {
            flags_ |= BEAGLE_FLAG_SCALING_MANUAL | BEAGLE_FLAG_SCALERS_LOG;
        } else {
            flags_ |= BEAGLE_FLAG_SCALING_MANUAL | BEAGLE_FLAG_SCALERS_RAW;
        }
        logScalers_ = (flags_ & BEAGLE_FLAG_SCALERS_LOG) != 0;
        // Pre-order "self" matrices and derivative matrices arrive transposed
        // for S > 4 unless the client asked the library to transpose
        // (the reference GPU convention, GPU/BeagleGPUImpl.hpp:483-486, 787-788).
        const bool autoT = (both & (long)BEAGLE_FLAG_PREORDER_TRANSPOSE_AUTO) != 0;
        flags_ |= autoT ? (long)BEAGLE_FLAG_PREORDER_TRANSPOSE_AUTO
                        : (long)BEAGLE_FLAG_PREORDER_TRANSPOSE_MANUAL;
        preTransposed_ = stateCount > 4 && !autoT;

        dims_.S = stateCount;
        dims_.P = patternCount;
        dims_.C = categoryCount;
        dims_.Sp = stateCount;
        dims_.Ppad = roundUp(patternCount, 64);
        dims_.catStride = (long)dims_.Ppad * dims_.Sp;
        dims_.autoScale = auto_ ? 1 : 0;
        dims_.sStride = stateCount > 4 ? PTILE : 1;
        partialsBytes_ = (size_t)categoryCount * dims_.catStride * realSize();
        matrixBytes_ = (size_t)categoryCount * stateCount * (stateCount + 1) * realSize();
        matrixBytes_ = (matrixBytes_ + 255) / 256 * 256;

        HIPCHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        dPartials_.assign(bufferCount_, nullptr);
        dExp_.assign(bufferCount_, nullptr);
        dStates_.assign(bufferCount_, nullptr);
        for (int b = tipCount_; b < bufferCount_; ++b) ensurePartials(b);
        dMatrices_ = dalloc<char>(matrixBytes_ * matrixCount_);
        HIPCHECK(hipMemsetAsync(dMatrices_, 0, matrixBytes_ * matrixCount_, stream_));
        dScale_ = dalloc<double>((size_t)std::max(scaleCount_, 1) * dims_.Ppad);
        launch_fill(dScale_, 0.0, (long)std::max(scaleCount_, 1) * dims_.Ppad, stream_);
        const size_t S = stateCount;
        dCube_ = dalloc<double>(eigenCount_ * S * S * S);
        dVVi_ = dalloc<double>(eigenCount_ * 2 * S * S);
        dEvals_ = dalloc<double>(eigenCount_ * 2 * S);   // real parts, then imaginary parts
        eigenSet_.assign(eigenCount_, false);
        hEigen_.resize(eigenCount_);
        matOrigin_.resize(matrixCount_);
        dFreqs_ = dalloc<double>(eigenCount_ * S);
        dCatWeights_ = dalloc<double>(eigenCount_ * categoryCount);
        dCatRates_ = dalloc<double>(eigenCount_ * categoryCount);
        hFreqs_.assign(eigenCount_, {});
        hCatWeights_.assign(eigenCount_, {});
        hCatRates_.assign(eigenCount_, {});
        dPatternWeights_ = dalloc<double>(dims_.Ppad);
        launch_fill(dPatternWeights_, 0.0, dims_.Ppad, stream_);
        hPatternWeights_.assign(patternCount, 0.0);
        dSiteL_ = dalloc<double>(dims_.Ppad);
        dSiteD1_ = dalloc<double>(dims_.Ppad);
        dSiteD2_ = dalloc<double>(dims_.Ppad);
        dResult_ = dalloc<double>(64);
        // 4-state real eigen-systems: every matrix computed from one also
        // gets the Frechet factors of its exponential (rate-matrix gradient)
        if (stateCount == 4 && !complex_ && categoryCount <= 8) {
            const long nF = (long)matrixCount_ * categoryCount * 16;
            dF_ = dalloc<double>(nF);
            launch_fill(dF_, 0.0, nF, stream_);
        }
        // one-tile stand-ins read by unused operand slots of the pipelined
        // traversal (in bounds for every lane and category, never stored to)
        dDummy_ = dalloc<char>(dummy_bytes());
        HIPCHECK(hipMemsetAsync(dDummy_, 0, dummy_bytes(), stream_));
        HIPCHECK(hipHostMalloc((void**)&hResult_, 64 * sizeof(double), hipHostMallocDefault));
        // device address of the allocation-free entry point's scratch array,
        // looked up once per device with that device current (null: the entry
        // point reports BEAGLE_ERROR_GENERAL)
        noAllocScratch_ = noalloc_scratch(device_);
        HIPCHECK(hipDeviceGetAttribute(&maxLdsPerBlock_, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                       device_));
        partStart_ = {0, patternCount};
        if (const char* m = std::getenv("BEAGLE_AMD_TRAVERSAL")) traversalMode_ = std::atoi(m);
        if (const char* m = std::getenv("BEAGLE_AMD_LEVEL_CHUNK")) levelChunk_ = std::max(1, std::atoi(m));
        if (const char* m = std::getenv("BEAGLE_AMD_MFMA_VARIANT")) {
            mfmaVariant_ = std::atoi(m);
            mfmaVariantForced_ = true;
        }
        if (const char* m = std::getenv("BEAGLE_AMD_AUTOTUNE")) autotune_ = std::atoi(m) != 0;
        if (const char* m = std::getenv("BEAGLE_AMD_TILE_LISTS")) tileLists_ = std::atoi(m) != 0;
        ring_.resize(8);
        for (auto& s : ring_) HIPCHECK(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
        jobs_.resize(8);
        for (auto& e : jobs_) HIPCHECK(hipEventCreateWithFlags(&e.ev, hipEventDisableTiming));
        sync();
    } catch (...) {
        release();
        throw;
    }
}

Instance::~Instance() { release(); }

// Frees everything the instance holds (also a partly built one: every
// handle starts null).
void Instance::release() {
    hipSetDevice(device_);
    if (capturing_) {
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(stream_, &g) == hipSuccess && g) hipGraphDestroy(g);
        (void)hipGetLastError();
        hipHostFree(cap_.host);
        hipFree(cap_.dev);
        capturing_ = false;
    }
    if (stream_) hipStreamSynchronize(stream_);
    for (Graph& G : graphs_)
        if (G.live) {
            hipGraphExecDestroy(G.exec);
            hipHostFree(G.host);
            hipFree(G.dev);
        }
    for (void* p : retired_) hipFree(p);
    for (void* p : dPartials_) hipFree(p);
    for (short* p : dExp_) hipFree(p);
    for (int* p : dStates_) hipFree(p);
    hipFree(dMatrices_);
    hipFree(dScale_);
    hipFree(dCube_);
    hipFree(dVVi_);
    hipFree(dEvals_);
    hipFree(dFreqs_);
    hipFree(dCatWeights_);
    hipFree(dCatRates_);
    hipFree(dPatternWeights_);
    hipFree(dSiteL_);
    hipFree(dSiteD1_);
    hipFree(dSiteD2_);
    hipFree(dResult_);
    hipFree(dF_);
    hipFree(dDummy_);
    hipHostFree(hResult_);
    hipFree(dWork_);
    hipFree(dJobMats_);
    hipFree(dPack_);
    hipFree(dSites_);
    if (hStage_) hipHostFree(hStage_);
    hipFree(dScratch_);
    for (auto& s : ring_) {
        if (s.host) hipHostFree(s.host);
        if (s.dev) hipFree(s.dev);
        if (s.ev) hipEventDestroy(s.ev);
    }
    for (auto& e : jobs_) {
        if (e.host) hipHostFree(e.host);
        if (e.dev) hipFree(e.dev);
        if (e.ev) hipEventDestroy(e.ev);
    }
    clearTuning();
    if (stream_) hipStreamDestroy(stream_);
    stream_ = nullptr;
    // a failed allocation leaves its error for hipGetLastError: clear it so
    // it is not reported by a later, unrelated call
    (void)hipGetLastError();
}

const char* Instance::implName() const {
    return dp_ ? "MI355X-HIP-Double" : "MI355X-HIP-Single";
}

// ---- transfers --------------------------------------------------------------

void Instance::sync() {
    if (capturing_) captureBreak("a call that waits for the device cannot be captured");
    HIPCHECK(hipStreamSynchronize(stream_));
}

void Instance::finishCall() {
    if (!capturing_ && !(flags_ & BEAGLE_FLAG_COMPUTATION_ASYNCH)) sync();
}

void Instance::copyIn(void* dst, const void* src, size_t bytes) {
    if (capturing_) captureBreak("a host copy cannot be captured");
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
    sync();
}

void Instance::copyOut(void* dst, const void* src, size_t bytes) {
    if (capturing_) captureBreak("a host copy cannot be captured");
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_));
    sync();
}

void Instance::toDevice(void* dst, const double* src, size_t n) {
    if (dp_) {
        copyIn(dst, src, n * sizeof(double));
    } else {
        std::vector<float> tmp(src, src + n);
        copyIn(dst, tmp.data(), n * sizeof(float));
    }
}

void Instance::toHost(double* dst, const void* src, size_t n) {
    if (dp_) {
        copyOut(dst, src, n * sizeof(double));
    } else {
        std::vector<float> tmp(n);
        copyOut(tmp.data(), src, n * sizeof(float));
        std::copy(tmp.begin(), tmp.end(), dst);
    }
}

// Small per-call uploads (job lists) go through a ring of pinned staging
// slots so the host never waits on the device unless a slot is reused
// before its previous copy finished.
void* Instance::upload(const void* data, size_t bytes) {
    if (capturing_) {
        // the graph's own copy, in its device arena: the whole arena is
        // copied once when the capture ends, so replays copy nothing
        const size_t off = captureAlloc(bytes);
        std::memcpy(cap_.host + off, data, bytes);
        lastUploadOff_ = off;
        return cap_.dev + off;
    }
    Slot& s = ring_[ringPos_];
    ringPos_ = (ringPos_ + 1) % (int)ring_.size();
    if (s.used) HIPCHECK(hipEventSynchronize(s.ev));
    if (s.cap < bytes) {
        size_t cap = std::max<size_t>(bytes, 4096);
        cap = (cap + 4095) / 4096 * 4096;
        if (s.host) HIPCHECK(hipHostFree(s.host));
        if (s.dev) {
            sync();
            HIPCHECK(hipFree(s.dev));
        }
        HIPCHECK(hipHostMalloc(&s.host, cap, hipHostMallocDefault));
        HIPCHECK(hipMalloc(&s.dev, cap));
        s.cap = cap;
    }
    std::memcpy(s.host, data, bytes);
    HIPCHECK(hipMemcpyAsync(s.dev, s.host, bytes, hipMemcpyHostToDevice, stream_));
    HIPCHECK(hipEventRecord(s.ev, stream_));
    s.used = true;
    return s.dev;
}

// Repeated traversals of one tree upload the same job lists: the list is
// compared with the recent ones (bytes, so the same operations on the same
// buffers) and a match is launched from its device copy.  Under graph
// capture the graph keeps its own copy (upload).
void* Instance::uploadJobs(const void* data, size_t bytes) {
    if (capturing_) return upload(data, bytes);
    if (jobsLast_ >= 0) {
        HIPCHECK(hipEventRecord(jobs_[jobsLast_].ev, stream_));
        jobs_[jobsLast_].pending = true;
    }
    ++jobsTick_;
    int pick = 0;
    for (int i = 0; i < (int)jobs_.size(); ++i) {
        JobEntry& e = jobs_[i];
        if (e.bytes == bytes && e.host && std::memcmp(e.host, data, bytes) == 0) {
            e.last = jobsTick_;
            jobsLast_ = i;
            return e.dev;
        }
        if (e.last < jobs_[pick].last) pick = i;
    }
    JobEntry& e = jobs_[pick];
    // its earlier kernels and its last transfer read these buffers
    if (e.pending) HIPCHECK(hipEventSynchronize(e.ev));
    e.pending = false;
    if (e.cap < bytes) {
        size_t cap = (std::max<size_t>(bytes, 4096) + 4095) / 4096 * 4096;
        if (e.host) HIPCHECK(hipHostFree(e.host));
        if (e.dev) HIPCHECK(hipFree(e.dev));
        e.host = e.dev = nullptr;
        HIPCHECK(hipHostMalloc(&e.host, cap, hipHostMallocDefault));
        HIPCHECK(hipMalloc(&e.dev, cap));
        e.cap = cap;
    }
    std::memcpy(e.host, data, bytes);
    e.bytes = bytes;
    HIPCHECK(hipMemcpyAsync(e.dev, e.host, bytes, hipMemcpyHostToDevice, stream_));
    e.last = jobsTick_;
    jobsLast_ = pick;
    return e.dev;
}

void* Instance::grow(void*& buf, size_t& cap, size_t bytes) {
    if (cap < bytes) {
        if (capturing_) captureBreak("a work buffer must grow: run the calls once before capturing");
        sync();
        // a live graph may hold the old buffer's address: keep it until the end
        bool held = false;
        for (const Graph& G : graphs_) held |= G.live;
        if (buf && held) retired_.push_back(buf);
        else if (buf) HIPCHECK(hipFree(buf));
        size_t c = std::max(bytes, cap * 2);
        HIPCHECK(hipMalloc(&buf, c));
        cap = c;
    }
    return buf;
}

void Instance::ensurePartials(int b) {
    if (dPartials_[b]) return;
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, partialsBytes_));
    HIPCHECK(hipMemsetAsync(p, 0, partialsBytes_, stream_));
    dPartials_[b] = p;
    if (auto_) {
        short* e = nullptr;
        HIPCHECK(hipMalloc((void**)&e, dims_.Ppad * sizeof(short)));
        HIPCHECK(hipMemsetAsync(e, 0, dims_.Ppad * sizeof(short), stream_));
        dExp_[b] = e;
    }
}

// a fill kernel, not a memset: the call may be captured into a graph
void Instance::clearExponent(int b) {
    if (auto_ && dExp_[b]) launch_fill_short(dExp_[b], 0, dims_.Ppad, stream_);
}

// ---- data in and out ---------------------------------------------------------

// States >= S mean missing and are stored as S; padded patterns are missing
// (libhmsbeagle/CPU/BeagleCPUImpl.hpp:570-584).
int Instance::setTipStates(int tip, const int* in) {
    if (tip < 0 || tip >= tipCount_) return BEAGLE_ERROR_OUT_OF_RANGE;
    if (!dStates_[tip]) {
        int* p = nullptr;
        HIPCHECK(hipMalloc((void**)&p, dims_.Ppad * sizeof(int)));
        dStates_[tip] = p;
    }
    std::vector<int> tmp(dims_.Ppad, dims_.S);
    for (int k = 0; k < dims_.P; ++k) tmp[k] = (in[k] >= 0 && in[k] < dims_.S) ? in[k] : dims_.S;
    copyIn(dStates_[tip], tmp.data(), tmp.size() * sizeof(int));
    return BEAGLE_SUCCESS;
}

int Instance::setTipPartials(int tip, const double* in) {
    if (tip < 0 || tip >= tipCount_) return BEAGLE_ERROR_OUT_OF_RANGE;
    ensurePartials(tip);
    const size_t n = (size_t)dims_.C * dims_.catStride;
    std::vector<double> tmp(n, 0.0);
    for (int c = 0; c < dims_.C; ++c)
        for (int k = 0; k < dims_.P; ++k)
            for (int i = 0; i < dims_.S; ++i)
                tmp[pbase(dims_, c, k) + (size_t)i * dims_.sStride] = in[(size_t)k * dims_.S + i];
    toDevice(dPartials_[tip], tmp.data(), n);
    clearExponent(tip);
    return BEAGLE_SUCCESS;
}

int Instance::setPartials(int b, const double* in) {
    if (!validBuffer(b)) return BEAGLE_ERROR_OUT_OF_RANGE;
    ensurePartials(b);
    const size_t n = (size_t)dims_.C * dims_.catStride;
    std::vector<double> tmp(n, 0.0);
    for (int c = 0; c < dims_.C; ++c)
        for (int k = 0; k < dims_.P; ++k)
            for (int i = 0; i < dims_.S; ++i)
                tmp[pbase(dims_, c, k) + (size_t)i * dims_.sStride] =
                    in[((size_t)c * dims_.P + k) * dims_.S + i];
    toDevice(dPartials_[b], tmp.data(), n);
    clearExponent(b);
    return BEAGLE_SUCCESS;
}

// Root pre-order partials = state frequencies replicated over categories
// and patterns (BeagleCPUImpl.hpp:623-677).
int Instance::setRootPrePartials(const int* bufIdx, const int* fIdx, int count) {
    if (count != 1) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    const int b = bufIdx[0], f = fIdx[0];
    if (!validBuffer(b) || !validEigen(f) || hFreqs_[f].empty()) return BEAGLE_ERROR_OUT_OF_RANGE;
    ensurePartials(b);
    // the same launch zeroes the buffer's scaling exponents
    short* ex = auto_ ? dExp_[b] : nullptr;
    if (dp_) launch_tile_vector<double>(dPartials_[b], dFreqs_ + (size_t)f * dims_.S, dims_, stream_, ex);
    else launch_tile_vector<float>(dPartials_[b], dFreqs_ + (size_t)f * dims_.S, dims_, stream_, ex);
    HIPCHECK(hipGetLastError());
    return BEAGLE_SUCCESS;
}

// With a cumulative scale buffer the reference multiplies the first
// patternCount*stateCount outputs (category 0) by exp(scale[pattern])
// (BeagleCPUImpl.hpp:743-754); kept as is.
int Instance::getPartials(int b, int scaleIndex, double* out) {
    if (!validBuffer(b) || !dPartials_[b]) return BEAGLE_ERROR_OUT_OF_RANGE;
    if (scaleIndex != BEAGLE_OP_NONE && !validScale(scaleIndex)) return BEAGLE_ERROR_OUT_OF_RANGE;
    const size_t n = (size_t)dims_.C * dims_.catStride;
    std::vector<double> tmp(n);
    toHost(tmp.data(), dPartials_[b], n);
    for (int c = 0; c < dims_.C; ++c)
        for (int k = 0; k < dims_.P; ++k)
            for (int i = 0; i < dims_.S; ++i)
                out[((size_t)c * dims_.P + k) * dims_.S + i] =
                    tmp[pbase(dims_, c, k) + (size_t)i * dims_.sStride];
    if (scaleIndex != BEAGLE_OP_NONE) {
        std::vector<double> sc(dims_.P);
        copyOut(sc.data(), scale(scaleIndex), dims_.P * sizeof(double));
        size_t idx = 0;
        for (int k = 0; k < dims_.P; ++k) {
            const double f = std::exp(sc[k]);
            for (int i = 0; i < dims_.S; ++i) out[idx++] *= f;
        }
    }
    return BEAGLE_SUCCESS;
}

// C_ijk = V_ik * Vinv_kj, kept in FP64 on the device
// (EigenDecompositionCube.hpp:69-100).
int Instance::setEigenDecomposition(int e, const double* V, const double* Vi, const double* lam) {
    if (!validEigen(e)) return BEAGLE_ERROR_OUT_OF_RANGE;
    const int S = dims_.S;
    const bool trans = (flags_ & BEAGLE_FLAG_INVEVEC_TRANSPOSED) != 0;
    std::vector<double> cube((size_t)S * S * S);
    size_t l = 0;
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j)
            for (int k = 0; k < S; ++k)
                cube[l++] = V[i * S + k] * (trans ? Vi[k + j * S] : Vi[k * S + j]);
    copyIn(dCube_ + (size_t)e * S * S * S, cube.data(), cube.size() * sizeof(double));
    std::vector<double> vv(2 * (size_t)S * S);
    for (int i = 0; i < S; ++i)
        for (int k = 0; k < S; ++k) {
            vv[(size_t)i * S + k] = V[i * S + k];
            vv[(size_t)S * S + (size_t)i * S + k] = trans ? Vi[i + k * S] : Vi[i * S + k];
        }
    copyIn(dVVi_ + (size_t)e * 2 * S * S, vv.data(), vv.size() * sizeof(double));
    // a complex system passes S real then S imaginary parts
    // (EigenDecompositionSquare.hpp:100, 121)
    copyIn(dEvals_ + (size_t)e * 2 * S, lam, (complex_ ? 2 : 1) * S * sizeof(double));
    eigenSet_[e] = true;
    // host copy for the rate-matrix gradient: V, V^-1 (row-major), lambda
    HostEigen& h = hEigen_[e];
    h.V.assign(V, V + (size_t)S * S);
    h.Vi.resize((size_t)S * S);
    for (int k = 0; k < S; ++k)
        for (int j = 0; j < S; ++j) h.Vi[(size_t)k * S + j] = trans ? Vi[k + j * S] : Vi[k * S + j];
    h.lam.assign(lam, lam + S);
    return BEAGLE_SUCCESS;
}

int Instance::setStateFrequencies(int i, const double* f) {
    if (!validEigen(i)) return BEAGLE_ERROR_OUT_OF_RANGE;
    hFreqs_[i].assign(f, f + dims_.S);
    copyIn(dFreqs_ + (size_t)i * dims_.S, f, dims_.S * sizeof(double));
    return BEAGLE_SUCCESS;
}

int Instance::setCategoryWeights(int i, const double* w) {
    if (!validEigen(i)) return BEAGLE_ERROR_OUT_OF_RANGE;
    hCatWeights_[i].assign(w, w + dims_.C);
    copyIn(dCatWeights_ + (size_t)i * dims_.C, w, dims_.C * sizeof(double));
    return BEAGLE_SUCCESS;
}

int Instance::setCategoryRates(const double* r) { return setCategoryRatesWithIndex(0, r); }

int Instance::setCategoryRatesWithIndex(int i, const double* r) {
    if (!validEigen(i)) return BEAGLE_ERROR_OUT_OF_RANGE;
    hCatRates_[i].assign(r, r + dims_.C);
    copyIn(dCatRates_ + (size_t)i * dims_.C, r, dims_.C * sizeof(double));
    return BEAGLE_SUCCESS;
}

int Instance::setPatternWeights(const double* w) {
    hPatternWeights_.assign(w, w + dims_.P);
    copyIn(dPatternWeights_, w, dims_.P * sizeof(double));
    return BEAGLE_SUCCESS;
}

// Pattern k of partition q lies in [partStart[q], partStart[q+1]).  When the
// partition indices are not already 0, 1, 2, ... in runs, the patterns are
// reordered so that each partition is contiguous, keeping their order within
// a partition: pattern i moves to partStart[part[i]] + its rank in its
// partition.  The tips' states or partials and the pattern weights are
// permuted (other buffers are not), site log-likelihoods are returned in the
// original order, and a second reorder is refused (BeagleCPUImpl.hpp:803-920,
// 4140-4219, 1009-1021).
int Instance::setPatternPartitions(int count, const int* parts) {
    if (count < 1) return BEAGLE_ERROR_OUT_OF_RANGE;
    const int P = dims_.P;
    for (int k = 0; k < P; ++k)
        if (parts[k] < 0 || parts[k] >= count) return BEAGLE_ERROR_OUT_OF_RANGE;
    bool reorder = false;
    for (int k = 0, run = 0; k < P; ++k) {
        if (k > 0 && parts[k] != parts[k - 1]) ++run;
        if (parts[k] != run) {
            reorder = true;
            break;
        }
    }
    if (!reorder) {
        partStart_.assign(count + 1, P);
        partStart_[0] = 0;
        for (int k = 1; k < P; ++k)
            if (parts[k] != parts[k - 1]) partStart_[parts[k]] = k;
        partitionCount_ = count;
        return BEAGLE_SUCCESS;
    }
    if (!newOrder_.empty()) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    std::vector<int> size(count, 0), order(P);
    for (int k = 0; k < P; ++k) order[k] = size[parts[k]]++;
    partStart_.assign(count + 1, 0);
    for (int q = 0; q < count; ++q) partStart_[q + 1] = partStart_[q] + size[q];
    for (int k = 0; k < P; ++k) order[k] += partStart_[parts[k]];
    // pattern weights
    std::vector<double> w(P);
    for (int k = 0; k < P; ++k) w[order[k]] = hPatternWeights_.empty() ? 0.0 : hPatternWeights_[k];
    if (!hPatternWeights_.empty()) setPatternWeights(w.data());
    // tips: states, or partials in the device layout
    for (int t = 0; t < tipCount_; ++t) {
        if (dStates_[t]) {
            std::vector<int> a(dims_.Ppad), b;
            copyOut(a.data(), dStates_[t], a.size() * sizeof(int));
            b = a;
            for (int k = 0; k < P; ++k) b[order[k]] = a[k];
            copyIn(dStates_[t], b.data(), b.size() * sizeof(int));
        } else if (dPartials_[t]) {
            const size_t rs = realSize();
            std::vector<char> a(partialsBytes_), b;
            copyOut(a.data(), dPartials_[t], partialsBytes_);
            b = a;
            for (int c = 0; c < dims_.C; ++c)
                for (int k = 0; k < P; ++k)
                    for (int i = 0; i < dims_.S; ++i)
                        std::memcpy(&b[(pbase(dims_, c, order[k]) + (long)i * dims_.sStride) * rs],
                                    &a[(pbase(dims_, c, k) + (long)i * dims_.sStride) * rs], rs);
            copyIn(dPartials_[t], b.data(), partialsBytes_);
        }
    }
    newOrder_ = std::move(order);
    partitionCount_ = count;
    return BEAGLE_SUCCESS;
}

// ---- transition matrices -------------------------------------------------------

int Instance::matrixUpdate(const std::vector<MatrixJob>& jobs) {
    if (jobs.empty()) return BEAGLE_SUCCESS;
    const MatrixJob* d = reinterpret_cast<const MatrixJob*>(
        upload(jobs.data(), jobs.size() * sizeof(MatrixJob)));
    if (capturing_) {
        // replays read each job's length from the graph's host memory,
        // where setGraphBranchLengths writes it
        cap_.matJobs.push_back({lastUploadOff_, (int)jobs.size()});
        MatrixJob* h = reinterpret_cast<MatrixJob*>(cap_.host + lastUploadOff_);
        for (size_t j = 0; j < jobs.size(); ++j)
            h[j].tHost = reinterpret_cast<const double*>(
                cap_.hostDev + ((char*)&h[j].t - cap_.host));
    }
    if (dp_) launch_matrices<double>(d, (int)jobs.size(), dims_.S, dims_.C, stream_);
    else launch_matrices<float>(d, (int)jobs.size(), dims_.S, dims_.C, stream_);
    HIPCHECK(hipGetLastError());
    return BEAGLE_SUCCESS;
}

// Category rates come from rate buffer 0 (BeagleCPUImpl.hpp:1225-1238).
int Instance::updateTransitionMatrices(int e, const int* probIdx, const int* d1Idx,
                                       const int* d2Idx, const double* t, int count) {
    if (!validEigen(e) || !eigenSet_[e] || hCatRates_[0].empty()) return BEAGLE_ERROR_OUT_OF_RANGE;
    const int S = dims_.S;
    std::vector<MatrixJob> jobs(count);
    for (int u = 0; u < count; ++u) {
        if (!validMatrix(probIdx[u])) return BEAGLE_ERROR_OUT_OF_RANGE;
        MatrixJob& J = jobs[u];
        J.P = matrix(probIdx[u]);
        J.D1 = nullptr;
        J.D2 = nullptr;
        if (d1Idx) {
            if (!validMatrix(d1Idx[u])) return BEAGLE_ERROR_OUT_OF_RANGE;
            J.D1 = matrix(d1Idx[u]);
            if (d2Idx) {
                if (!validMatrix(d2Idx[u])) return BEAGLE_ERROR_OUT_OF_RANGE;
                J.D2 = matrix(d2Idx[u]);
            }
        }
        J.cube = dCube_ + (size_t)e * S * S * S;
        J.vv = dVVi_ + (size_t)e * 2 * S * S;
        J.evals = dEvals_ + (size_t)e * 2 * S;
        J.cplx = complex_ ? 1 : 0;
        J.rates = dCatRates_;
        J.t = t[u];
        J.cat = -1;
        J.F = dF_ ? dF_ + (size_t)probIdx[u] * dims_.C * 16 : nullptr;
        J.tHost = nullptr;
    }
    for (int u = 0; u < count; ++u) {
        matOrigin_[probIdx[u]] = MatrixOrigin{e, 0, t[u]};
        if (d1Idx) matOrigin_[d1Idx[u]] = MatrixOrigin{};
        if (d1Idx && d2Idx) matOrigin_[d2Idx[u]] = MatrixOrigin{};
    }
    return matrixUpdate(jobs);
}

// Category c uses eigen-system eigenIndices[c] (its own rate from buffer 0):
// expressed as one single-category job per (matrix, category).
int Instance::updateTransitionMatricesWithModelCategories(const int* eig, const int* probIdx,
                                                          const int* d1Idx, const int* d2Idx,
                                                          const double* t, int count) {
    const int S = dims_.S, C = dims_.C;
    for (int c = 0; c < C; ++c)
        if (!validEigen(eig[c]) || !eigenSet_[eig[c]]) return BEAGLE_ERROR_OUT_OF_RANGE;
    std::vector<MatrixJob> jobs;
    jobs.reserve((size_t)count * C);
    for (int u = 0; u < count; ++u) {
        if (!validMatrix(probIdx[u]) || (d1Idx && !validMatrix(d1Idx[u])) ||
            (d1Idx && d2Idx && !validMatrix(d2Idx[u])))
            return BEAGLE_ERROR_OUT_OF_RANGE;
        for (int c = 0; c < C; ++c) {
            MatrixJob J;
            J.P = matrix(probIdx[u]);
            J.D1 = d1Idx ? matrix(d1Idx[u]) : nullptr;
            J.D2 = (d1Idx && d2Idx) ? matrix(d2Idx[u]) : nullptr;
            J.cube = dCube_ + (size_t)eig[c] * S * S * S;
            J.vv = dVVi_ + (size_t)eig[c] * 2 * S * S;
            J.evals = dEvals_ + (size_t)eig[c] * 2 * S;
            J.cplx = complex_ ? 1 : 0;
            J.rates = nullptr;
            J.t = t[u];
            J.cat = c;
            J.F = nullptr;
            J.tHost = nullptr;
            jobs.push_back(J);
        }
        matOrigin_[probIdx[u]] = MatrixOrigin{};   // several eigen-systems
        if (d1Idx) matOrigin_[d1Idx[u]] = MatrixOrigin{};
        if (d1Idx && d2Idx) matOrigin_[d2Idx[u]] = MatrixOrigin{};
    }
    return matrixUpdate(jobs);
}

int Instance::updateTransitionMatricesWithMultipleModels(const int* eig, const int* rateIdx,
                                                         const int* probIdx, const int* d1Idx,
                                                         const int* d2Idx, const double* t,
                                                         int count) {
    const int S = dims_.S;
    std::vector<MatrixJob> jobs(count);
    for (int u = 0; u < count; ++u) {
        const int e = eig[u], r = rateIdx[u];
        if (!validEigen(e) || !eigenSet_[e] || !validEigen(r) || hCatRates_[r].empty() ||
            !validMatrix(probIdx[u]))
            return BEAGLE_ERROR_OUT_OF_RANGE;
        MatrixJob& J = jobs[u];
        J.P = matrix(probIdx[u]);
        J.D1 = (d1Idx && validMatrix(d1Idx[u])) ? matrix(d1Idx[u]) : nullptr;
        J.D2 = (d1Idx && d2Idx && validMatrix(d2Idx[u])) ? matrix(d2Idx[u]) : nullptr;
        J.cube = dCube_ + (size_t)e * S * S * S;
        J.vv = dVVi_ + (size_t)e * 2 * S * S;
        J.evals = dEvals_ + (size_t)e * 2 * S;
        J.cplx = complex_ ? 1 : 0;
        J.rates = dCatRates_ + (size_t)r * dims_.C;
        J.t = t[u];
        J.cat = -1;
        J.F = dF_ ? dF_ + (size_t)probIdx[u] * dims_.C * 16 : nullptr;
        J.tHost = nullptr;
    }
    for (int u = 0; u < count; ++u) {
        matOrigin_[probIdx[u]] = MatrixOrigin{eig[u], rateIdx[u], t[u]};
        if (jobs[u].D1) matOrigin_[d1Idx[u]] = MatrixOrigin{};
        if (jobs[u].D2) matOrigin_[d2Idx[u]] = MatrixOrigin{};
    }
    return matrixUpdate(jobs);
}

// Client matrices [C][S][S] plus the pad value (BeagleCPUImpl.hpp:1035-1091).
int Instance::setTransitionMatrix(int m, const double* in, double pad) {
    if (!validMatrix(m)) return BEAGLE_ERROR_OUT_OF_RANGE;
    matOrigin_[m] = MatrixOrigin{};
    const int S = dims_.S, W = S + 1;
    std::vector<double> tmp((size_t)dims_.C * S * W);
    for (int c = 0; c < dims_.C; ++c)
        for (int i = 0; i < S; ++i) {
            for (int j = 0; j < S; ++j)
                tmp[((size_t)c * S + i) * W + j] = in[((size_t)c * S + i) * S + j];
            tmp[((size_t)c * S + i) * W + S] = pad;
        }
    toDevice(matrix(m), tmp.data(), tmp.size());
    return BEAGLE_SUCCESS;
}

int Instance::setDifferentialMatrix(int m, const double* in) { return setTransitionMatrix(m, in, 0.0); }

int Instance::setTransitionMatrices(const int* idx, const double* in, const double* pads, int count) {
    const size_t n = (size_t)dims_.C * dims_.S * dims_.S;
    for (int k = 0; k < count; ++k) {
        const int rc = setTransitionMatrix(idx[k], in + k * n, pads[k]);
        if (rc != BEAGLE_SUCCESS) return rc;
    }
    return BEAGLE_SUCCESS;
}

int Instance::getTransitionMatrix(int m, double* out) {
    if (!validMatrix(m)) return BEAGLE_ERROR_OUT_OF_RANGE;
    const int S = dims_.S, W = S + 1;
    std::vector<double> tmp((size_t)dims_.C * S * W);
    toHost(tmp.data(), matrix(m), tmp.size());
    for (int c = 0; c < dims_.C; ++c)
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j)
                out[((size_t)c * S + i) * S + j] = tmp[((size_t)c * S + i) * W + j];
    return BEAGLE_SUCCESS;
}

int Instance::transposeTransitionMatrices(const int* in, const int* out, int count) {
    std::vector<const void*> src(count);
    std::vector<void*> dst(count);
    for (int u = 0; u < count; ++u) {
        if (!validMatrix(in[u]) || !validMatrix(out[u])) return BEAGLE_ERROR_OUT_OF_RANGE;
        src[u] = matrix(in[u]);
        dst[u] = matrix(out[u]);
        matOrigin_[out[u]] = MatrixOrigin{};
    }
    if (count == 0) return BEAGLE_SUCCESS;
    auto ds = reinterpret_cast<const void* const*>(upload(src.data(), count * sizeof(void*)));
    auto dd = reinterpret_cast<void* const*>(upload(dst.data(), count * sizeof(void*)));
    if (dp_) launch_transpose<double>(ds, dd, count, dims_.S, dims_.C, stream_);
    else launch_transpose<float>(ds, dd, count, dims_.S, dims_.C, stream_);
    HIPCHECK(hipGetLastError());
    return BEAGLE_SUCCESS;
}

// result = first x second per category; in-place results are refused
// (BeagleCPUImpl.hpp:1096-1162).
int Instance::convolveTransitionMatrices(const int* a, const int* b, const int* out, int count) {
    std::vector<ConvolveJob> jobs(count);
    for (int u = 0; u < count; ++u) {
        if (!validMatrix(a[u]) || !validMatrix(b[u]) || !validMatrix(out[u]) || a[u] == out[u] ||
            b[u] == out[u])
            return BEAGLE_ERROR_OUT_OF_RANGE;
        jobs[u] = ConvolveJob{matrix(a[u]), matrix(b[u]), matrix(out[u])};
        matOrigin_[out[u]] = MatrixOrigin{};
    }
    if (count == 0) return BEAGLE_SUCCESS;
    auto d = reinterpret_cast<const ConvolveJob*>(upload(jobs.data(), count * sizeof(ConvolveJob)));
    if (dp_) launch_convolve<double>(d, count, dims_.S, dims_.C, stream_);
    else launch_convolve<float>(d, count, dims_.S, dims_.C, stream_);
    HIPCHECK(hipGetLastError());
    finishCall();
    return BEAGLE_SUCCESS;
}

int Instance::addTransitionMatrices(const int*, const int*, const int*, int) {
    return BEAGLE_ERROR_NO_IMPLEMENTATION;
}

// ---- traversals ------------------------------------------------------------------

// Post-order operations (BeagleCPUImpl.hpp:1497-1686): the rescale mode of
// each op is resolved here, the arithmetic on the device.
int Instance::updatePartials(const int* ops, int count, int cumulativeIndex, bool byPartition) {
    const int stride = byPartition ? BEAGLE_PARTITION_OP_COUNT : BEAGLE_OP_COUNT;
    hPost_.resize(count);
    for (int o = 0; o < count; ++o) {
        const int* q = ops + (size_t)o * stride;
        const int dest = q[0], wS = q[1], rS = q[2], c1 = q[3], m1 = q[4], c2 = q[5], m2 = q[6];
        int cum = cumulativeIndex;
        int p0 = 0, p1 = dims_.P;
        if (byPartition) {
            const int part = q[7];
            if (part < 0 || part >= partitionCount_) return BEAGLE_ERROR_OUT_OF_RANGE;
            p0 = partStart_[part];
            p1 = partStart_[part + 1];
            cum = q[8];
        }
        if (dest < tipCount_ || !validBuffer(dest) || !validBuffer(c1) || !validBuffer(c2) ||
            !validMatrix(m1) || !validMatrix(m2))
            return BEAGLE_ERROR_OUT_OF_RANGE;
        if ((!dStates_[c1] && !dPartials_[c1]) || (!dStates_[c2] && !dPartials_[c2]))
            return BEAGLE_ERROR_OUT_OF_RANGE;
        PostJob& J = hPost_[o];
        J.x1 = dStates_[c1] ? nullptr : dPartials_[c1];
        J.s1 = dStates_[c1];
        J.m1 = matrix(m1);
        J.x2 = dStates_[c2] ? nullptr : dPartials_[c2];
        J.s2 = dStates_[c2];
        J.m2 = matrix(m2);
        J.mo1 = (int)(m1 * (matrixBytes_ / realSize()));
        J.mo2 = (int)(m2 * (matrixBytes_ / realSize()));
        J.k1 = dStates_[c1] ? 1 : 0;
        J.k2 = dStates_[c2] ? 1 : 0;
        J.dest = dPartials_[dest];
        J.scaleWrite = nullptr;
        J.scaleRead = nullptr;
        J.cumulative = nullptr;
        J.accA = J.accB = nullptr;
        J.e1 = dStates_[c1] ? nullptr : dExp_[c1];
        J.e2 = dStates_[c2] ? nullptr : dExp_[c2];
        J.ed = dExp_[dest];
        J.mode = SCALE_NONE;
        J.logScalers = logScalers_ ? 1 : 0;
        J.p0 = p0;
        J.p1 = p1;
        if (auto_) continue;  // every op rescales by powers of two
        if (cum != BEAGLE_OP_NONE && !validScale(cum)) return BEAGLE_ERROR_OUT_OF_RANGE;
        if (always_) {
            J.mode = SCALE_WRITE;
            J.scaleWrite = scale(dest - tipCount_);
            if (cum != BEAGLE_OP_NONE) J.cumulative = scale(cum);
            if (c1 >= tipCount_) J.accA = scale(c1 - tipCount_);
            if (c2 >= tipCount_) J.accB = scale(c2 - tipCount_);
        } else if (wS >= 0) {
            if (!validScale(wS)) return BEAGLE_ERROR_OUT_OF_RANGE;
            J.mode = SCALE_WRITE;
            J.scaleWrite = scale(wS);
            if (cum != BEAGLE_OP_NONE) J.cumulative = scale(cum);
        } else if (rS >= 0) {
            if (!validScale(rS)) return BEAGLE_ERROR_OUT_OF_RANGE;
            J.mode = SCALE_READ;
            J.scaleRead = scale(rS);
        }
    }
    if (count == 0) return BEAGLE_SUCCESS;
    const bool fourState = dims_.S == 4 && dims_.C <= 8;
    if (!fourState && mfma_post_supported(dims_.S, dims_.C)) {
        launchPostMfma(ops, stride, count);
    } else if (!fourState) {
        auto d = reinterpret_cast<const PostJob*>(uploadJobs(hPost_.data(), count * sizeof(PostJob)));
        if (dp_) launch_post<double>(d, hPost_.data(), count, 0, dims_, dMatrices_, dDummy_, stream_);
        else launch_post<float>(d, hPost_.data(), count, 0, dims_, dMatrices_, dDummy_, stream_);
    } else {
        // A shared cumulative scale buffer is updated by every rescaling op of
        // a pattern, so such lists stay in one fused launch (ops of a pattern
        // in order); otherwise wide levels are batched.
        bool shared = false;
        for (int o = 0; o < count; ++o) shared |= hPost_[o].cumulative != nullptr;
        std::vector<int> level;
        int nl = 0;
        if (!shared) nl = postLevels(ops, stride, count, level);
        // plan candidate k: kPlans[k] = {fused list, narrow-level threshold}
        auto plan = [&](int k) {
            std::vector<Segment> pl;
            if (shared || kPlans[k].fused) {
                Segment sg{false, {}};
                for (int o = 0; o < count; ++o) sg.ops.push_back(o);
                pl.push_back(std::move(sg));
            } else {
                pl = planSegments(level, nl, kPlans[k].narrow);
            }
            return pl;
        };
        if (shared) {
            launchPostPlan(plan(0));
        } else {
            const int k = tunedChoice(shapeKey(1, count, nl), kNumPlans, forcedPlan(),
                                      [&](int c) { launchPostPlan(plan(c)); });
            if (k >= 0) launchPostPlan(plan(k));
        }
    }
    HIPCHECK(hipGetLastError());
    finishCall();
    return BEAGLE_SUCCESS;
}

int Instance::setSplitBf16(int terms) {
    if (terms != 0 && terms != 3 && terms != 4) return BEAGLE_ERROR_OUT_OF_RANGE;
    if (dp_ || !mfma_post_supported(dims_.S, dims_.C)) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    splitBf16_ = terms != 0;
    if (terms) bf16Terms_ = terms;
    return BEAGLE_SUCCESS;
}

// ---- autotuning ----------------------------------------------------------------
// A traversal's launch plan (4-state: one fused launch, or levels batched
// above a width threshold; MFMA: partials pipelined or a whole tile up
// front) does not change its results, only its speed, and which plan is
// fastest depends on the instance's sizes (fixed), the call kind, the
// number of operations and how they spread over dependency levels.  Calls
// are keyed by that shape class (shapeKey), not by the exact operation
// list, so a sampler proposing new topologies of one tree size tunes once.
// The first call of a class runs every candidate twice and records events
// around the second run; the timings are read on a later call of the class
// once the events have completed (hipEventQuery, no wait: an ASYNCH caller
// is never blocked), and until then the default plan runs.  The last
// candidate's outputs are left in place (the operations are idempotent:
// lists that add to a cumulative scale buffer are not tuned).  At most
// kMaxTuned classes are kept; beyond that the default plan runs untuned.

const Instance::PlanChoice Instance::kPlans[Instance::kNumPlans] = {
    {false, kNarrowLevel}, {true, 0}, {false, 4}, {false, 16}, {false, 1}};

// Shape class of a call: kind (tag), operation count and the mean width of
// its dependency levels in three classes (chains, below 16, 16 or more).
uint64_t Instance::shapeKey(uint64_t tag, int count, int levels) {
    const int width = count / std::max(levels, 1);
    const uint64_t cls = width < 2 ? 0 : (width < 16 ? 1 : 2);
    return (tag << 48) ^ ((uint64_t)(uint32_t)count << 4) ^ cls;
}

int Instance::forcedPlan() const {
    if (traversalMode_ == 1) return 4;   // every level batched
    if (traversalMode_ == 2) return 1;   // the whole list fused
    return -1;
}

// True once every candidate's timed run has completed; then picks the
// fastest and frees the events.
bool Instance::resolveTuning(Tuning& t) {
    if (t.choice >= 0) return true;
    for (auto& e : t.ev) {
        const hipError_t q = hipEventQuery(e.second);
        if (q == hipErrorNotReady) return false;
        HIPCHECK(q);
    }
    int best = 0;
    float bestMs = 0.f;
    for (size_t k = 0; k < t.ev.size(); ++k) {
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, t.ev[k].first, t.ev[k].second));
        if (k == 0 || ms < bestMs) {
            best = (int)k;
            bestMs = ms;
        }
    }
    for (auto& e : t.ev) {
        hipEventDestroy(e.first);
        hipEventDestroy(e.second);
    }
    t.ev.clear();
    t.choice = best;
    return true;
}

void Instance::clearTuning() {
    for (auto& kv : tuned_)
        for (auto& e : kv.second.ev) {
            hipEventDestroy(e.first);
            hipEventDestroy(e.second);
        }
    tuned_.clear();
}

template <typename F>
int Instance::tunedChoice(uint64_t key, int ncand, int forced, F&& run) {
    if (forced >= 0) return forced;
    const auto it = tuned_.find(key);
    if (it != tuned_.end()) {
        if (it->second.choice >= 0) return it->second.choice;
        return !capturing_ && resolveTuning(it->second) ? it->second.choice : 0;
    }
    if (!autotune_ || capturing_ || tuned_.size() >= kMaxTuned) return 0;
    Tuning t;
    for (int k = 0; k < ncand; ++k) {
        hipEvent_t a, b;
        HIPCHECK(hipEventCreate(&a));
        HIPCHECK(hipEventCreate(&b));
        t.ev.emplace_back(a, b);
        run(k);
        HIPCHECK(hipEventRecord(a, stream_));
        run(k);
        HIPCHECK(hipEventRecord(b, stream_));
    }
    tuned_.emplace(key, std::move(t));
    ++tunings_;
    return -1;
}

int Instance::setAutotune(int on) {
    autotune_ = on != 0;
    if (!autotune_) clearTuning();
    return BEAGLE_SUCCESS;
}

int Instance::autotuneStats(int* outTuned, int* outPending) {
    int pending = 0;
    for (auto& kv : tuned_) pending += resolveTuning(kv.second) ? 0 : 1;
    if (outTuned) *outTuned = tunings_;
    if (outPending) *outPending = pending;
    return BEAGLE_SUCCESS;
}

int Instance::postLevels(const int* ops, int stride, int count, std::vector<int>& level) {
    std::vector<Access> acc(count);
    for (int o = 0; o < count; ++o) {
        const int* q = ops + (size_t)o * stride;
        const PostJob& J = hPost_[o];
        const int sw = J.scaleWrite ? (int)((J.scaleWrite - dScale_) / dims_.Ppad) : -1;
        const int sr = J.scaleRead ? (int)((J.scaleRead - dScale_) / dims_.Ppad) : -1;
        acc[o] = Access{{q[3], q[5], sr >= 0 ? bufferCount_ + sr : -1},
                        {q[0], sw >= 0 ? bufferCount_ + sw : -1}};
    }
    return dependencyLevels(acc, bufferCount_ + scaleCount_, level);
}

// Work lists (TileList) of the levels [start[l], start[l + 1]) of `jobs`:
// nothing when every job covers all patterns; otherwise, per level, first[]
// over its jobs' tiles of tileP patterns (from the tile that holds the job's
// first pattern) appended to `ints`, off[l] its offset there (-1: the level
// keeps the (tiles, jobs) grid) and items[l] its tile count.
constexpr int kTile4 = 64;   // patterns per tile of the 4-state level kernels

template <typename Job>
static void levelTileLists(const Job* jobs, const int* start, int nl, int P, int tileP,
                           std::vector<int>& ints, std::vector<int>& off,
                           std::vector<int>& items) {
    off.assign(nl, -1);
    items.assign(nl, 0);
    for (int l = 0; l < nl; ++l) {
        bool whole = true;
        for (int o = start[l]; o < start[l + 1]; ++o) whole &= jobs[o].p0 == 0 && jobs[o].p1 == P;
        if (whole) continue;
        off[l] = (int)ints.size();
        int acc = 0;
        for (int o = start[l]; o < start[l + 1]; ++o) {
            ints.push_back(acc);
            const int t0 = jobs[o].p0 / tileP, t1 = (jobs[o].p1 + tileP - 1) / tileP;
            acc += std::max(0, t1 - t0);
        }
        ints.push_back(acc);
        items[l] = acc;
    }
}

// MFMA post-order (S > 4; FP64, FP32 or split-BF16 matrix cores): every
// child matrix of the list is repacked once into the pack slab, then each
// dependency level is one launch over (pattern tiles x ops).  A list whose
// ops share a cumulative scale buffer runs one op per launch in list order
// (each rescaling op adds to the same per-pattern sums).
void Instance::launchPostMfma(const int* ops, int stride, int count) {
    const size_t per = packFloats();
    std::vector<PackJob> pj(2 * (size_t)count);
    for (int o = 0; o < count; ++o) {
        PostJob& J = hPost_[o];
        J.pk1 = (int)(2 * o * per);
        J.pk2 = (int)((2 * o + 1) * per);
        const int operand = splitBf16_ ? PACK_OPERAND_BF16 : PACK_OPERAND;
        pj[2 * o] = PackJob{J.mo1, J.pk1, J.s1 ? PACK_ROWS : operand};
        pj[2 * o + 1] = PackJob{J.mo2, J.pk2, J.s2 ? PACK_ROWS : operand};
    }
    float* pack = static_cast<float*>(
        grow(reinterpret_cast<void*&>(dPack_), packCap_, pj.size() * per * sizeof(float)));
    std::vector<int> level(count);
    bool shared = false;
    for (int o = 0; o < count; ++o) shared |= hPost_[o].cumulative != nullptr;
    int nl;
    if (shared) {
        for (int o = 0; o < count; ++o) level[o] = o;
        nl = count;
    } else {
        nl = postLevels(ops, stride, count, level);
    }
    std::vector<int> start(nl + 1, 0);
    for (int o = 0; o < count; ++o) ++start[level[o] + 1];
    for (int l = 0; l < nl; ++l) start[l + 1] += start[l];
    std::vector<PostJob> sorted(count);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int o = 0; o < count; ++o) sorted[fill[level[o]]++] = hPost_[o];
    auto dpj = reinterpret_cast<const PackJob*>(uploadJobs(pj.data(), pj.size() * sizeof(PackJob)));
    if (dp_) launch_pack_matrices64(dpj, (int)pj.size(), dims_.S, dims_.C, dMatrices_, pack, stream_);
    else launch_pack_matrices<float>(dpj, (int)pj.size(), dims_.S, dims_.C, dMatrices_, pack, stream_);
    auto d = reinterpret_cast<const PostJob*>(uploadJobs(sorted.data(), count * sizeof(PostJob)));
    // operations by partition: each level's (operation, 32-pattern tile) list
    std::vector<int> lints, loff, litems;
    if (!dp_ && tileLists_) levelTileLists(sorted.data(), start.data(), nl, dims_.P, PTILE, lints, loff, litems);
    const int* dl = lints.empty() ? nullptr
                                  : static_cast<const int*>(uploadJobs(lints.data(), lints.size() * sizeof(int)));
    auto run = [&](int variant) {
        for (int l = 0; l < nl; ++l) {
            const int nj = start[l + 1] - start[l];
            const TileList w{dl && loff[l] >= 0 ? dl + loff[l] : nullptr, nj};
            const int it = dl ? litems[l] : 0;
            if (dp_) launch_post_mfma64(d + start[l], nj, dims_, pack, stream_);
            else if (splitBf16_) launch_post_bf16(d + start[l], nj, dims_, pack, bf16Terms_, stream_, w, it);
            else launch_post_mfma(d + start[l], nj, dims_, pack, variant, stream_, w, it);
        }
    };
    if (dp_ || splitBf16_ || shared) {
        run(mfmaVariant_);
    } else {
        // candidates: partials pipelined (0) or the whole tile up front (1)
        const int k = tunedChoice(shapeKey(3, count, nl), 2,
                                  mfmaVariantForced_ ? mfmaVariant_ : -1, run);
        if (k >= 0) run(k);
    }
}

// Launch a 4-state post-order plan with one upload of all its jobs.
// Batched segment: one launch over (tiles x ops).  Fused segment: the
// previous op's output is forwarded from registers when it is an operand
// over the same pattern range; a read of it over a different range
// (partitions) needs the stored values, so such a segment runs one
// operation per launch.
void Instance::launchPostPlan(const std::vector<Segment>& plan) {
    std::vector<PostJob> all;
    struct Launch { int first, n; bool batched; };
    std::vector<Launch> launches;
    for (const Segment& s : plan) {
        const int first = (int)all.size(), n = (int)s.ops.size();
        if (n == 0) continue;
        for (int o : s.ops) all.push_back(hPost_[o]);
        PostJob* jobs = all.data() + first;
        if (s.batched) {
            launches.push_back({first, n, true});
            continue;
        }
        bool late = false;
        for (int o = 1; o < n; ++o) {
            PostJob& J = jobs[o];
            const PostJob& Pv = jobs[o - 1];
            const bool same = J.p0 == Pv.p0 && J.p1 == Pv.p1;
            if (J.x1 && J.x1 == Pv.dest) { if (same) J.k1 = 2; else late = true; }
            if (J.x2 && J.x2 == Pv.dest) { if (same) J.k2 = 2; else late = true; }
        }
        if (late) {
            for (int o = 0; o < n; ++o) {
                jobs[o].k1 = jobs[o].s1 ? 1 : 0;
                jobs[o].k2 = jobs[o].s2 ? 1 : 0;
                launches.push_back({first + o, 1, false});
            }
        } else {
            launches.push_back({first, n, false});
        }
    }
    if (all.empty()) return;
    auto d = reinterpret_cast<const PostJob*>(uploadJobs(all.data(), all.size() * sizeof(PostJob)));
    // batched levels of operations by partition: (operation, tile) lists
    std::vector<int> lints, loff(launches.size(), -1), litems(launches.size(), 0);
    for (size_t i = 0; i < launches.size(); ++i) {
        const Launch& L = launches[i];
        if (!L.batched || levelChunk_ > 1 || !tileLists_) continue;
        std::vector<int> o1, i1;
        const int start[2] = {L.first, L.first + L.n};
        levelTileLists(all.data(), start, 1, dims_.P, kTile4, lints, o1, i1);
        loff[i] = o1[0];
        litems[i] = i1[0];
    }
    const int* dl = lints.empty() ? nullptr
                                  : static_cast<const int*>(uploadJobs(lints.data(), lints.size() * sizeof(int)));
    for (size_t i = 0; i < launches.size(); ++i) {
        const Launch& L = launches[i];
        PostJob* h = all.data() + L.first;
        if (L.batched && levelChunk_ <= 1) {
            const int start[2] = {0, L.n};
            const TileList w{dl && loff[i] >= 0 ? dl + loff[i] : nullptr, L.n};
            if (dp_) launch_post_levels<double>(d + L.first, h, start, 1, dims_, dMatrices_, stream_, w, litems[i]);
            else launch_post_levels<float>(d + L.first, h, start, 1, dims_, dMatrices_, stream_, w, litems[i]);
        } else {
            // fused list, or a batched level walked in chunks of independent ops
            const int chunk = L.batched ? levelChunk_ : 0;
            if (dp_) launch_post<double>(d + L.first, h, L.n, chunk, dims_, dMatrices_, dDummy_, stream_);
            else launch_post<float>(d + L.first, h, L.n, chunk, dims_, dMatrices_, dDummy_, stream_);
        }
    }
}

// Pre-order operations [dest, wScale, rScale, parent, P_self, sibling,
// P_sib] (BeagleCPUImpl.hpp:1689-1862); rescaling never feeds a cumulative
// buffer and there is no read-scale path, as in the reference.
int Instance::buildPreJobs(const int* ops, int count, bool byPartition, bool allowNoDest) {
    const int stride = byPartition ? BEAGLE_PARTITION_OP_COUNT : BEAGLE_OP_COUNT;
    hPre_.resize(count);
    const size_t mstride = matrixBytes_ / realSize();
    for (int o = 0; o < count; ++o) {
        const int* q = ops + (size_t)o * stride;
        const int dest = q[0], wS = q[1], par = q[3], mSelf = q[4], sib = q[5], mSib = q[6];
        int p0 = 0, p1 = dims_.P;
        if (byPartition) {
            const int part = q[7];
            if (part < 0 || part >= partitionCount_) return BEAGLE_ERROR_OUT_OF_RANGE;
            p0 = partStart_[part];
            p1 = partStart_[part + 1];
        }
        const bool noDest = allowNoDest && dest == BEAGLE_OP_NONE;
        if ((!noDest && !validBuffer(dest)) || !validBuffer(par) || !validBuffer(sib) ||
            !validMatrix(mSelf) || !validMatrix(mSib) || !dPartials_[par])
            return BEAGLE_ERROR_OUT_OF_RANGE;
        if (!dStates_[sib] && !dPartials_[sib]) return BEAGLE_ERROR_OUT_OF_RANGE;
        if (!noDest) ensurePartials(dest);
        PreJob& J = hPre_[o];
        std::memset(&J, 0, sizeof(PreJob));
        J.oidx = o;
        J.parent = dPartials_[par];
        J.mSelf = matrix(mSelf);
        J.xs = dStates_[sib] ? nullptr : dPartials_[sib];
        J.ss = dStates_[sib];
        J.mSib = matrix(mSib);
        J.moSelf = (int)(mSelf * mstride);
        J.moSib = (int)(mSib * mstride);
        J.kSib = dStates_[sib] ? 1 : 0;
        J.kPar = 0;
        J.dest = noDest ? nullptr : dPartials_[dest];
        J.e1 = dExp_[par];
        J.e2 = dStates_[sib] ? nullptr : dExp_[sib];
        J.ed = noDest ? nullptr : dExp_[dest];
        J.mode = SCALE_NONE;
        J.logScalers = logScalers_ ? 1 : 0;
        J.p0 = p0;
        J.p1 = p1;
        if (auto_ || noDest) continue;  // nothing kept: no scale factors either
        if (always_) {
            if (!validScale(dest - tipCount_)) return BEAGLE_ERROR_OUT_OF_RANGE;
            J.mode = SCALE_WRITE;
            J.scaleWrite = scale(dest - tipCount_);
        } else if (wS >= 0) {
            if (!validScale(wS)) return BEAGLE_ERROR_OUT_OF_RANGE;
            J.mode = SCALE_WRITE;
            J.scaleWrite = scale(wS);
        }
    }
    return BEAGLE_SUCCESS;
}

// Launch the pre-order jobs in hPre_ (built from `ops`): the generic-S
// path one launch per op; for 4 states a level plan of batched and fused
// segments (see planSegments).  partial != null adds the derivative sums.
// S > 4: pre-order and edge derivatives on the matrix cores of the
// instance's engine: FP32, split BF16 (beagleExtSetSplitBF16) or FP64.
bool Instance::mfmaAdjoint() const { return mfma_post_supported(dims_.S, dims_.C); }

// floats per packed child matrix (all categories) of the MFMA kernels
size_t Instance::packFloats() const {
    return (size_t)(dp_ ? 2 * mfma_pack_doubles(dims_.S) : mfma_pack_reals(dims_.S)) * dims_.C;
}

// MFMA pre-order: P_self^T packed for the register-resident second product
// (from the stored transpose when the client passes transposed self
// matrices), P_sib as a post-order child; one launch per dependency level.
void Instance::launchPreMfma(const int* ops, int stride, int count, const TipDeriv* td) {
    const size_t per = packFloats();
    std::vector<PackJob> pj(2 * (size_t)count);
    const int tin = preTransposed_ ? PACK_TRANSPOSED_INPUT : 0;
    const size_t mstride = matrixBytes_ / realSize();
    for (int o = 0; o < count; ++o) {
        PreJob& J = hPre_[o];
        J.pkT = J.pkM1 = J.pkM2 = -1;
        const bool bf = splitBf16_ && !dp_;
        if (td && !dp_ && dStates_[td->postIdx[o]]) {
            // the tip's columns of P_self and of P_self D1, P_self D2
            J.sself = dStates_[td->postIdx[o]];
            const int base = (int)pj.size();
            J.pkT = (int)(base * per);
            J.pkM1 = (int)((base + 1) * per);
            J.pkM2 = (int)((base + 2) * per);
            pj.push_back(PackJob{J.moSelf, J.pkT, PACK_TIP_COLUMN | tin, 0});
            pj.push_back(PackJob{J.moSelf, J.pkM1, PACK_TIP_PRODUCT | tin, (int)(td->d1[o] * mstride)});
            pj.push_back(PackJob{J.moSelf, J.pkM2, PACK_TIP_PRODUCT | tin, (int)(td->d2[o] * mstride)});
        } else if (td && !dp_) {
            // an internal branch: v's post-order partials, D1 and D2 packed
            // as P_self is (their transposes applied to pre(v) in registers)
            J.xself = dPartials_[td->postIdx[o]];
            const int base = (int)pj.size();
            J.pkM1 = (int)(base * per);
            J.pkM2 = (int)((base + 1) * per);
            const int kd = (bf ? PACK_PRE_SELF_BF16 : PACK_PRE_SELF) | tin;
            pj.push_back(PackJob{(int)(td->d1[o] * mstride), J.pkM1, kd, 0});
            pj.push_back(PackJob{(int)(td->d2[o] * mstride), J.pkM2, kd, 0});
        }
        J.pkSelf = (int)(2 * o * per);
        J.pkSib = (int)((2 * o + 1) * per);
        pj[2 * o] = PackJob{J.moSelf, J.pkSelf, (bf ? PACK_PRE_SELF_BF16 : PACK_PRE_SELF) | tin};
        pj[2 * o + 1] =
            PackJob{J.moSib, J.pkSib, J.ss ? PACK_ROWS : bf ? PACK_OPERAND_BF16 : PACK_OPERAND};
    }
    float* pack = static_cast<float*>(
        grow(reinterpret_cast<void*&>(dPack_), packCap_, pj.size() * per * sizeof(float)));
    std::vector<Access> acc(count);
    std::vector<int> level;
    for (int o = 0; o < count; ++o) {
        const int* q = ops + (size_t)o * stride;
        const PreJob& J = hPre_[o];
        const int sw = J.scaleWrite ? (int)((J.scaleWrite - dScale_) / dims_.Ppad) : -1;
        acc[o] = Access{{q[3], q[5], -1}, {q[0], sw >= 0 ? bufferCount_ + sw : -1}};
    }
    const int nl = dependencyLevels(acc, bufferCount_ + scaleCount_, level);
    std::vector<int> start(nl + 1, 0);
    for (int o = 0; o < count; ++o) ++start[level[o] + 1];
    for (int l = 0; l < nl; ++l) start[l + 1] += start[l];
    std::vector<PreJob> sorted(count);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int o = 0; o < count; ++o) sorted[fill[level[o]]++] = hPre_[o];
    auto dpj = reinterpret_cast<const PackJob*>(uploadJobs(pj.data(), pj.size() * sizeof(PackJob)));
    if (dp_) launch_pack_matrices64(dpj, (int)pj.size(), dims_.S, dims_.C, dMatrices_, pack, stream_);
    else launch_pack_matrices<float>(dpj, (int)pj.size(), dims_.S, dims_.C, dMatrices_, pack, stream_);
    if (td && !dp_) {
        std::vector<PackJob> prod;
        for (const PackJob& q : pj)
            if ((q.kind & ~PACK_TRANSPOSED_INPUT) == PACK_TIP_PRODUCT) prod.push_back(q);
        if (!prod.empty()) {
            auto dpp = reinterpret_cast<const PackJob*>(
                uploadJobs(prod.data(), prod.size() * sizeof(PackJob)));
            launch_pack_tip_products<float>(dpp, (int)prod.size(), dims_.S, dims_.C, dMatrices_,
                                            pack, stream_);
        }
    }
    auto d = reinterpret_cast<const PreJob*>(uploadJobs(sorted.data(), count * sizeof(PreJob)));
    // operations by partition: each level's (operation, 32-pattern tile) list
    std::vector<int> lints, loff, litems;
    if (!dp_ && tileLists_) levelTileLists(sorted.data(), start.data(), nl, dims_.P, PTILE, lints, loff, litems);
    const int* dl = lints.empty() ? nullptr
                                  : static_cast<const int*>(uploadJobs(lints.data(), lints.size() * sizeof(int)));
    const double* cw = td ? dCatWeights_ + (size_t)td->wIdx * dims_.C : nullptr;
    for (int l = 0; l < nl; ++l) {
        const int nj = start[l + 1] - start[l];
        const TileList w{dl && loff[l] >= 0 ? dl + loff[l] : nullptr, nj};
        if (dp_) launch_pre_mfma64(d + start[l], nj, dims_, pack, stream_);
        else launch_pre_mfma(d + start[l], nj, dims_, pack, splitBf16_ ? bf16Terms_ : 0, stream_,
                             w, dl ? litems[l] : 0, cw, dPatternWeights_,
                             td ? td->partial : nullptr);
    }
}

// MFMA edge derivatives of `count` nodes with one (d2 null) or two
// derivative matrices, enqueued: *dSums receives per node NR = 2 or 3 sums
// (sum w r1, sum w r1^2, sum w r2), interleaved; sites (one matrix only, may
// be null) the per-pattern ratios.
// skip[n] != 0: entry n's sums are already in `work` (tip branches of the
// fused gradient, written by the pre-order kernel); work: that buffer, sized
// for all entries, else one is taken here.
int Instance::enqueueDerivMfma(const int* postIdx, const int* preIdx, const int* d1,
                               const int* d2, int wIdx, int count, double* sites,
                               double** dSums, const char* skip, double* work) {
    const size_t per = packFloats();
    const int nd = d2 ? 2 : 1, NR = d2 ? 3 : 2;
    const size_t mstride = matrixBytes_ / realSize();
    const int tin = preTransposed_ ? PACK_TRANSPOSED_INPUT : 0;
    std::vector<DerivJob> jobs;
    std::vector<PackJob> pj;
    jobs.reserve(count);
    pj.reserve((size_t)nd * count);
    for (int n = 0; n < count; ++n) {
        if (skip && skip[n]) continue;
        const int po = postIdx[n], pr = preIdx[n];
        if (!validBuffer(po) || !validBuffer(pr) || !dPartials_[pr] || !validMatrix(d1[n]) ||
            (d2 && !validMatrix(d2[n])) || (!dStates_[po] && !dPartials_[po]))
            return BEAGLE_ERROR_OUT_OF_RANGE;
        jobs.push_back(DerivJob{});
        DerivJob& J = jobs.back();
        J.pre = dPartials_[pr];
        J.states = dStates_[po];
        J.post = dStates_[po] ? nullptr : dPartials_[po];
        J.row = n;
        const int kind =
            (J.states ? PACK_ROWS : splitBf16_ && !dp_ ? PACK_OPERAND_BF16 : PACK_OPERAND) | tin;
        J.pk1 = (int)(pj.size() * per);
        pj.push_back(PackJob{(int)(d1[n] * mstride), J.pk1, kind, 0});
        J.pk2 = J.pk1;
        if (d2) {
            J.pk2 = (int)(pj.size() * per);
            pj.push_back(PackJob{(int)(d2[n] * mstride), J.pk2, kind, 0});
        }
    }
    const int njobs = (int)jobs.size();
    float* pack = static_cast<float*>(
        grow(reinterpret_cast<void*&>(dPack_), packCap_, pj.size() * per * sizeof(float)));
    const int tileP = dp_ ? 16 : PTILE;          // patterns per derivative tile
    const int ntiles = (dims_.P + tileP - 1) / tileP;
    const size_t nPart = (size_t)NR * count * ntiles;
    const size_t need = (nPart + (size_t)NR * count) * sizeof(double);
    if (work && (work != dWork_ || need > workCap_)) return BEAGLE_ERROR_GENERAL;
    double* wk = work ? work : reinterpret_cast<double*>(grow(dWork_, workCap_, need));
    if (njobs > 0) {
        auto dpj = reinterpret_cast<const PackJob*>(uploadJobs(pj.data(), pj.size() * sizeof(PackJob)));
        if (dp_) launch_pack_matrices64(dpj, (int)pj.size(), dims_.S, dims_.C, dMatrices_, pack, stream_);
        else launch_pack_matrices<float>(dpj, (int)pj.size(), dims_.S, dims_.C, dMatrices_, pack, stream_);
        auto dj = reinterpret_cast<const DerivJob*>(uploadJobs(jobs.data(), njobs * sizeof(DerivJob)));
        if (dp_)
            launch_deriv_mfma64(dj, njobs, nd, dims_, pack, dCatWeights_ + (size_t)wIdx * dims_.C,
                                dPatternWeights_, sites, wk, stream_);
        else
            launch_deriv_mfma(dj, njobs, nd, dims_, pack, dCatWeights_ + (size_t)wIdx * dims_.C,
                              dPatternWeights_, sites, wk, splitBf16_ ? bf16Terms_ : 0, stream_);
    }
    launch_reduce_rows(wk, ntiles, NR * count, wk + nPart, stream_);
    HIPCHECK(hipGetLastError());
    *dSums = wk + nPart;
    return BEAGLE_SUCCESS;
}

void Instance::launchPre(const int* ops, int stride, int count, double* partial) {
    const bool fourState = dims_.S == 4 && dims_.C <= 8 && !preTransposed_;
    const double* cw = partial ? dCatWeights_ + (size_t)preWeights_ * dims_.C : nullptr;
    if (!fourState && !partial && mfmaAdjoint()) {
        launchPreMfma(ops, stride, count);
        return;
    }
    if (!fourState) {
        if (!dScratch_) HIPCHECK(hipMalloc(&dScratch_, partialsBytes_));
        auto d = reinterpret_cast<const PreJob*>(uploadJobs(hPre_.data(), count * sizeof(PreJob)));
        const int t = preTransposed_ ? 1 : 0;
        if (dp_) launch_pre<double>(d, hPre_.data(), count, dims_, t, dScratch_, dMatrices_, dDummy_, cw, dPatternWeights_, partial, stream_);
        else launch_pre<float>(d, hPre_.data(), count, dims_, t, dScratch_, dMatrices_, dDummy_, cw, dPatternWeights_, partial, stream_);
        return;
    }
    std::vector<Access> acc(count);
    std::vector<int> level;
    for (int o = 0; o < count; ++o) {
        const int* q = ops + (size_t)o * stride;
        const PreJob& J = hPre_[o];
        const int sw = J.scaleWrite ? (int)((J.scaleWrite - dScale_) / dims_.Ppad) : -1;
        acc[o] = Access{{q[3], q[5], -1}, {q[0], sw >= 0 ? bufferCount_ + sw : -1}};
    }
    const int nl = dependencyLevels(acc, bufferCount_ + scaleCount_, level);
    auto run = [&](int k) {
        std::vector<Segment> plan;
        if (kPlans[k].fused) {
            Segment sg{false, {}};
            for (int o = 0; o < count; ++o) sg.ops.push_back(o);
            plan.push_back(std::move(sg));
        } else {
            plan = planSegments(level, nl, kPlans[k].narrow);
        }
        launchPrePlan(plan, cw, partial);
    };
    const int k = tunedChoice(shapeKey(partial ? 5 : 4, count, nl), kNumPlans, forcedPlan(),
                              run);
    if (k >= 0) run(k);
}

// Launch a 4-state pre-order plan with one upload of all its jobs (the
// batched / fused rules of launchPostPlan).
void Instance::launchPrePlan(const std::vector<Segment>& plan, const double* cw, double* partial) {
    std::vector<PreJob> all;
    struct Launch { int first, n; bool batched; };
    std::vector<Launch> launches;
    for (const Segment& s : plan) {
        const int first = (int)all.size(), n = (int)s.ops.size();
        if (n == 0) continue;
        for (int o : s.ops) all.push_back(hPre_[o]);
        PreJob* jobs = all.data() + first;
        if (s.batched) {
            launches.push_back({first, n, true});
            continue;
        }
        // parent operand kinds: 2 = previous op's output (same pattern
        // range), 3 = previous op's parent (same range); a read of the
        // previous output over a different range runs one op per launch
        bool late = false;
        for (int o = 1; o < n; ++o) {
            PreJob& J = jobs[o];
            const PreJob& Pv = jobs[o - 1];
            const bool same = J.p0 == Pv.p0 && J.p1 == Pv.p1;
            if (J.parent == Pv.dest) {
                if (same) J.kPar = 2;
                else late = true;
            } else if (J.parent == Pv.parent && same) {
                J.kPar = 3;
            }
        }
        if (late) {
            for (int o = 0; o < n; ++o) {
                jobs[o].kPar = 0;
                launches.push_back({first + o, 1, false});
            }
        } else {
            launches.push_back({first, n, false});
        }
    }
    if (all.empty()) return;
    auto d = reinterpret_cast<const PreJob*>(uploadJobs(all.data(), all.size() * sizeof(PreJob)));
    // batched levels of operations by partition (no derivative sums, which
    // need every tile's slot): (operation, tile) lists
    std::vector<int> lints, loff(launches.size(), -1), litems(launches.size(), 0);
    for (size_t i = 0; i < launches.size() && !partial; ++i) {
        const Launch& L = launches[i];
        if (!L.batched || !tileLists_) continue;
        std::vector<int> o1, i1;
        const int start[2] = {L.first, L.first + L.n};
        levelTileLists(all.data(), start, 1, dims_.P, kTile4, lints, o1, i1);
        loff[i] = o1[0];
        litems[i] = i1[0];
    }
    const int* dl = lints.empty() ? nullptr
                                  : static_cast<const int*>(uploadJobs(lints.data(), lints.size() * sizeof(int)));
    for (size_t i = 0; i < launches.size(); ++i) {
        const Launch& L = launches[i];
        if (L.batched) {
            const int start[2] = {0, L.n};
            const TileList w{dl && loff[i] >= 0 ? dl + loff[i] : nullptr, L.n};
            if (dp_) launch_pre_levels<double>(d + L.first, start, 1, dims_, dMatrices_, dDummy_, cw, dPatternWeights_, partial, stream_, w, litems[i]);
            else launch_pre_levels<float>(d + L.first, start, 1, dims_, dMatrices_, dDummy_, cw, dPatternWeights_, partial, stream_, w, litems[i]);
        } else {
            PreJob* h = all.data() + L.first;
            if (dp_) launch_pre<double>(d + L.first, h, L.n, dims_, 0, dScratch_, dMatrices_, dDummy_, cw, dPatternWeights_, partial, stream_);
            else launch_pre<float>(d + L.first, h, L.n, dims_, 0, dScratch_, dMatrices_, dDummy_, cw, dPatternWeights_, partial, stream_);
        }
    }
}

int Instance::updatePrePartials(const int* ops, int count, int cumulativeIndex, bool byPartition) {
    (void)cumulativeIndex;
    const int rc = buildPreJobs(ops, count, byPartition, false);
    if (rc != BEAGLE_SUCCESS || count == 0) return rc;
    launchPre(ops, byPartition ? BEAGLE_PARTITION_OP_COUNT : BEAGLE_OP_COUNT, count, nullptr);
    HIPCHECK(hipGetLastError());
    finishCall();
    return BEAGLE_SUCCESS;
}

// Extension: pre-order operations fused with the derivative sums of each
// operation's branch (see beagle_amd_ext.h).  Where the fused 4-state
// kernel does not apply it runs the equivalent API sequence.
int Instance::updatePrePartialsWithDerivatives(const int* ops, int count, const int* postIdx,
                                               const int* d1Idx, const int* d2Idx, int wIdx,
                                               double* outD1, double* outD1Sq, double* outD2) {
    if (!validEigen(wIdx) || hCatWeights_[wIdx].empty()) return BEAGLE_ERROR_OUT_OF_RANGE;
    const bool fused = dims_.S == 4 && dims_.C <= 8 && !preTransposed_ && d2Idx != nullptr;
    if (!fused && mfmaAdjoint() && d2Idx != nullptr) {
        // S > 4: MFMA pre-order, then both derivative matrices in one pass.
        // FP32 / split BF16: every branch's sums come from the pre-order
        // kernel (a tip branch's from the operands of the tip's pre-order
        // partials, which a NONE destination leaves unwritten; an internal
        // branch's from pre(v) in registers), the derivative pass only
        // reduces them.  FP64: the derivative pass.
        std::vector<char> inPre(count, 0);
        for (int o = 0; o < count; ++o) {
            const int po = postIdx[o];
            if (!validBuffer(po) || (!dStates_[po] && !dPartials_[po]) ||
                !validMatrix(d1Idx[o]) || !validMatrix(d2Idx[o]))
                return BEAGLE_ERROR_OUT_OF_RANGE;
            inPre[o] = !dp_;
            if (ops[(size_t)o * BEAGLE_OP_COUNT] == BEAGLE_OP_NONE && (dp_ || !dStates_[po]))
                return BEAGLE_ERROR_OUT_OF_RANGE;
        }
        int rc = buildPreJobs(ops, count, false, true);
        if (rc != BEAGLE_SUCCESS || count == 0) return rc;
        const bool want = outD1 || outD1Sq || outD2;
        // the derivative kernels' tiles: 32 patterns (FP32), 16 (FP64)
        const int tileP = dp_ ? 16 : PTILE;
        const size_t nPart = (size_t)3 * count * ((dims_.P + tileP - 1) / tileP);
        double* wk = want ? reinterpret_cast<double*>(grow(dWork_, workCap_,
                                                           (nPart + (size_t)3 * count) * sizeof(double)))
                          : nullptr;
        const TipDeriv td{postIdx, d1Idx, d2Idx, wIdx, wk};
        launchPreMfma(ops, BEAGLE_OP_COUNT, count, want && !dp_ ? &td : nullptr);
        HIPCHECK(hipGetLastError());
        if (!want) {
            finishCall();
            return BEAGLE_SUCCESS;
        }
        std::vector<int> pre(count);
        for (int o = 0; o < count; ++o) pre[o] = ops[(size_t)o * BEAGLE_OP_COUNT];
        double* sums = nullptr;
        rc = enqueueDerivMfma(postIdx, pre.data(), d1Idx, d2Idx, wIdx, count, nullptr, &sums,
                              inPre.data(), wk);
        if (rc != BEAGLE_SUCCESS) return rc;
        return deliverDerivativeSums(sums, count, outD1, outD1Sq, outD2);
    }
    if (!fused) {
        for (int o = 0; o < count; ++o)
            if (ops[(size_t)o * BEAGLE_OP_COUNT] == BEAGLE_OP_NONE) return BEAGLE_ERROR_OUT_OF_RANGE;
        int rc = updatePrePartials(ops, count, BEAGLE_OP_NONE, false);
        if (rc != BEAGLE_SUCCESS) return rc;
        std::vector<int> pre(count), w(1, wIdx);
        for (int o = 0; o < count; ++o) pre[o] = ops[(size_t)o * BEAGLE_OP_COUNT];
        rc = calculateEdgeDerivatives(postIdx, pre.data(), d1Idx, w.data(), count, nullptr, outD1,
                                      outD1Sq);
        if (rc != BEAGLE_SUCCESS || !d2Idx) return rc;
        std::vector<double> s2(count);
        rc = calculateEdgeDerivatives(postIdx, pre.data(), d2Idx, w.data(), count, nullptr,
                                      outD2 ? outD2 : s2.data(), nullptr);
        return rc;
    }
    const double* sums = nullptr;
    const bool want = outD1 || outD1Sq || outD2;
    int rc = enqueuePreDerivatives(ops, count, postIdx, d1Idx, d2Idx, wIdx, want, &sums);
    if (rc != BEAGLE_SUCCESS || count == 0) return rc;
    if (!want) {
        // nothing to return: no reduction, no copy; an asynchronous
        // instance returns without waiting
        finishCall();
        return BEAGLE_SUCCESS;
    }
    return deliverDerivativeSums(sums, count, outD1, outD1Sq, outD2);
}

// Per-operation sums (3 per operation on the device) to the caller: copied
// and waited for, or, while capturing, written at each replay's wait.
int Instance::deliverDerivativeSums(const double* sums, int count, double* outD1,
                                    double* outD1Sq, double* outD2) {
    if (capturing_) {
        const size_t off = captureResult(sums, 3 * (size_t)count);
        double* dst[3] = {outD1, outD1Sq, outD2};
        for (int k = 0; k < 3; ++k)
            if (dst[k]) cap_.outs.push_back({dst[k], off + k * sizeof(double), count, 3, false});
        return BEAGLE_SUCCESS;
    }
    const double* h = stageOut(sums, 3 * (size_t)count, 0);
    sync();
    for (int o = 0; o < count; ++o) {
        if (outD1) outD1[o] = h[3 * o];
        if (outD1Sq) outD1Sq[o] = h[3 * o + 1];
        if (outD2) outD2[o] = h[3 * o + 2];
    }
    return BEAGLE_SUCCESS;
}

// Pinned staging for results copied back: reserveStage(n) makes room for n
// doubles (waiting for the stream if it must grow), stageOut copies n
// doubles from the device to element `at` on the instance stream (the
// caller waits before reading).
void Instance::reserveStage(size_t n) {
    if (stageCap_ >= n) return;
    if (capturing_) captureBreak("results cannot be staged while capturing");
    sync();
    if (hStage_) HIPCHECK(hipHostFree(hStage_));
    stageCap_ = std::max<size_t>(2 * n, 1024);
    HIPCHECK(hipHostMalloc((void**)&hStage_, stageCap_ * sizeof(double), hipHostMallocDefault));
}

const double* Instance::stageOut(const double* dsrc, size_t n, size_t at) {
    reserveStage(at + n);
    HIPCHECK(hipMemcpyAsync(hStage_ + at, dsrc, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    return hStage_ + at;
}

// The fused 4-state derivative pass, enqueued; with `reduce` the per-
// operation sums (3 per operation) end in *dSums on the device.
//
// With qg (4 states, FP32, 1, 2 or 4 categories) the pass also forms the
// rate-matrix gradient sums H of eigen-system qg->eigen: qg->slots slots of
// qg->hacc ([slot][16], at most node jobs x tiles); the work buffer also
// holds their block sums, the root's workgroup rows (qg->rootRows, [1 + S]
// [P / 256]), H (qg->sums) and the call's outputs in one block (qg->out:
// logL, dlogL/dpi, dlogL/dQ, then the 3 sums per operation, which the
// reduction writes there).  qg->fused tells whether the node kernels ran
// (otherwise nothing of the gradient was formed).
int Instance::enqueuePreDerivatives(const int* ops, int count, const int* postIdx,
                                    const int* d1Idx, const int* d2Idx, int wIdx, bool reduce,
                                    const double** dSums, QGradPass* qg) {
    int rc = buildPreJobs(ops, count, false, true);
    if (rc != BEAGLE_SUCCESS || count == 0) return rc;
    rc = attachOwnOperands(count, postIdx, d1Idx, d2Idx);
    if (rc != BEAGLE_SUCCESS) return rc;
    const int ntiles = (dims_.P + 63) / 64;
    const size_t nPart = (size_t)count * 3 * ntiles;
    const int S = dims_.S, nb = (dims_.P + 255) / 256;
    const long cap = qg ? (long)(count / 2 + 1) * ntiles : 0;
    const size_t nq = qg ? 16 * ((size_t)cap + (size_t)cap / 128 + 1) + (size_t)(1 + S) * nb +
                               16 + (1 + (size_t)S + 16 + 3 * (size_t)count)
                         : 0;
    double* wk = reinterpret_cast<double*>(
        grow(dWork_, workCap_, (nPart + 3 * (size_t)count + nq) * sizeof(double)));
    double* sums = wk + nPart;
    preWeights_ = wIdx;
    bool nodes;
    if (qg) {
        qg->hacc = sums + 3 * (size_t)count;
        qg->slotPart = qg->hacc + 16 * (size_t)cap;
        qg->rootRows = qg->slotPart + 16 * ((size_t)cap / 128 + 1);
        qg->sums = qg->rootRows + (size_t)(1 + S) * nb;
        qg->out = qg->sums + 16;
        sums = qg->out + 1 + S + 16;   // the outputs' block: [logL | S | 16 | 3 count]
        NodeQGrad q{dF_, dVVi_ + (size_t)qg->eigen * 2 * S * S, qg->hacc, cap, 0};
        nodes = launchPreNodes(ops, count, postIdx, wk, &q);
        qg->fused = nodes;
        qg->slots = q.used;
    } else {
        nodes = launchPreNodes(ops, count, postIdx, wk);
    }
    if (!nodes) launchPre(ops, BEAGLE_OP_COUNT, count, wk);
    if (reduce) launch_reduce_rows(wk, ntiles, 3 * count, sums, stream_);
    HIPCHECK(hipGetLastError());
    *dSums = sums;
    return BEAGLE_SUCCESS;
}

// Extension: root log-likelihood, root pre-order partials and the fused
// derivative pass, with one wait (see beagle_amd_ext.h).
int Instance::logLikelihoodAndDerivatives(int rootIdx, int wIdx, int fIdx, int cumIdx,
                                          int rootPreIdx, const int* ops, int count,
                                          const int* postIdx, const int* d1Idx, const int* d2Idx,
                                          double* outLogL, double* outD1, double* outD1Sq,
                                          double* outD2) {
    const bool fused = dims_.S == 4 && dims_.C <= 8 && !preTransposed_ && d2Idx != nullptr;
    if (!fused) {
        int rc = calculateRootLogLikelihoods(&rootIdx, &wIdx, &fIdx, &cumIdx, 1, outLogL);
        if (rc != BEAGLE_SUCCESS) return rc;
        rc = setRootPrePartials(&rootPreIdx, &fIdx, 1);
        if (rc != BEAGLE_SUCCESS) return rc;
        return updatePrePartialsWithDerivatives(ops, count, postIdx, d1Idx, d2Idx, wIdx, outD1,
                                                outD1Sq, outD2);
    }
    if (!validEigen(wIdx) || hCatWeights_[wIdx].empty()) return BEAGLE_ERROR_OUT_OF_RANGE;
    int rc = enqueueRoot(rootIdx, wIdx, fIdx, cumIdx);
    if (rc != BEAGLE_SUCCESS) return rc;
    rc = setRootPrePartials(&rootPreIdx, &fIdx, 1);
    if (rc != BEAGLE_SUCCESS) return rc;
    const double* sums = nullptr;
    rc = enqueuePreDerivatives(ops, count, postIdx, d1Idx, d2Idx, wIdx, true, &sums);
    if (rc != BEAGLE_SUCCESS) return rc;
    if (capturing_) {
        cap_.outs.push_back({outLogL, captureResult(dResult_, 1), 1, 1, true});
        const size_t off = captureResult(sums, 3 * (size_t)count);
        double* dst[3] = {outD1, outD1Sq, outD2};
        for (int k = 0; k < 3; ++k)
            if (dst[k]) cap_.outs.push_back({dst[k], off + k * sizeof(double), count, 3, false});
        return BEAGLE_SUCCESS;
    }
    reserveStage(1 + 3 * (size_t)count);
    const double* hl = stageOut(dResult_, 1, 0);
    const double* h = stageOut(sums, 3 * (size_t)count, 1);
    sync();
    lastLogL_ = hl[0];
    *outLogL = lastLogL_;
    for (int o = 0; o < count; ++o) {
        if (outD1) outD1[o] = h[3 * o];
        if (outD1Sq) outD1Sq[o] = h[3 * o + 1];
        if (outD2) outD2[o] = h[3 * o + 2];
    }
    return lastLogL_ != lastLogL_ ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
}

// The operands of a rate-matrix gradient of eigen-system e: every
// operation's branch matrix computed from e (its Frechet factors were then
// written beside it), category weights wIdx set.
int Instance::checkQGradOperands(int e, const int* ops, int count, int wIdx) {
    if (!validEigen(e) || !eigenSet_[e] || hEigen_[e].lam.empty() || !validEigen(wIdx) ||
        hCatWeights_[wIdx].empty())
        return BEAGLE_ERROR_OUT_OF_RANGE;
    for (int o = 0; o < count; ++o) {
        const int m = ops[(size_t)o * BEAGLE_OP_COUNT + 4];
        if (!validMatrix(m)) return BEAGLE_ERROR_OUT_OF_RANGE;
        const MatrixOrigin& org = matOrigin_[m];
        if (org.eigen != e || hCatRates_[org.rates].empty()) return BEAGLE_ERROR_OUT_OF_RANGE;
    }
    return BEAGLE_SUCCESS;
}

// Extension: logLikelihoodAndDerivatives and rateMatrixGradient with one
// wait (see beagle_amd_ext.h).  On the FP32 4-state node path (1, 2 or 4
// categories) the rate-matrix gradient is formed inside the derivative
// pass from the operands it already holds; otherwise the two calls run in
// turn.
int Instance::logLikelihoodAndGradient(int rootIdx, int wIdx, int fIdx, int cumIdx,
                                       int rootPreIdx, const int* ops, int count,
                                       const int* postIdx, const int* d1Idx, const int* d2Idx,
                                       int e, double* outLogL, double* outD1, double* outD1Sq,
                                       double* outD2, double* outQ, double* outFreqs) {
    const int S = dims_.S;
    const bool fused = S == 4 && !dp_ && node_f32_supported(dims_.C) && !preTransposed_ &&
                       !complex_ && d2Idx != nullptr && dF_ != nullptr && count > 0 &&
                       partialsBytes_ < (size_t(1) << 32);
    const bool none = outLogL == nullptr;   // every output NULL (the front end checks)
    if (!fused) {
        if (none) return BEAGLE_ERROR_OUT_OF_RANGE;
        int rc = logLikelihoodAndDerivatives(rootIdx, wIdx, fIdx, cumIdx, rootPreIdx, ops, count,
                                             postIdx, d1Idx, d2Idx, outLogL, outD1, outD1Sq,
                                             outD2);
        if (rc != BEAGLE_SUCCESS && rc != BEAGLE_ERROR_FLOATING_POINT) return rc;
        const int rq = rateMatrixGradient(e, ops, count, postIdx, wIdx, rootIdx, fIdx, outQ,
                                          outFreqs);
        return rq != BEAGLE_SUCCESS ? rq : rc;
    }
    int rc = checkQGradOperands(e, ops, count, wIdx);
    if (rc != BEAGLE_SUCCESS) return rc;
    const bool freqs = outFreqs != nullptr;
    if (freqs && (!validBuffer(rootIdx) || !dPartials_[rootIdx] || !validEigen(fIdx) ||
                  hFreqs_[fIdx].empty()))
        return BEAGLE_ERROR_OUT_OF_RANGE;
    // the root's validity first: its kernel runs after the pass
    if (!validBuffer(rootIdx) || !dPartials_[rootIdx] || !validEigen(fIdx) ||
        hFreqs_[fIdx].empty() || dims_.sStride != 1)
        return BEAGLE_ERROR_OUT_OF_RANGE;
    const int cs = always_ ? rootIdx - tipCount_ : (auto_ ? BEAGLE_OP_NONE : cumIdx);
    if (cs != BEAGLE_OP_NONE && !validScale(cs)) return BEAGLE_ERROR_OUT_OF_RANGE;
    rc = setRootPrePartials(&rootPreIdx, &fIdx, 1);
    if (rc != BEAGLE_SUCCESS) return rc;
    const double* sums = nullptr;
    QGradPass qg{e, false, nullptr, nullptr, nullptr, nullptr, nullptr, 0};
    rc = enqueuePreDerivatives(ops, count, postIdx, d1Idx, d2Idx, wIdx, true, &sums, &qg);
    if (rc != BEAGLE_SUCCESS) return rc;
    if (!qg.fused) {
        // the operations do not pair into node jobs: logL, the derivative
        // results as logLikelihoodAndDerivatives delivers them, then the
        // separate gradient
        if (capturing_) captureBreak("the rate-matrix gradient needs node-paired operations");
        if (none) return BEAGLE_ERROR_OUT_OF_RANGE;
        rc = enqueueRoot(rootIdx, wIdx, fIdx, cumIdx);
        if (rc != BEAGLE_SUCCESS) return rc;
        reserveStage(1 + 3 * (size_t)count);
        const double* hl = stageOut(dResult_, 1, 0);
        const double* h = stageOut(sums, 3 * (size_t)count, 1);
        sync();
        lastLogL_ = hl[0];
        *outLogL = lastLogL_;
        for (int o = 0; o < count; ++o) {
            if (outD1) outD1[o] = h[3 * o];
            if (outD1Sq) outD1Sq[o] = h[3 * o + 1];
            if (outD2) outD2[o] = h[3 * o + 2];
        }
        const int rq = rateMatrixGradient(e, ops, count, postIdx, wIdx, rootIdx, fIdx, outQ,
                                          outFreqs);
        if (rq != BEAGLE_SUCCESS) return rq;
        return lastLogL_ != lastLogL_ ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
    }
    // the root after the pass (nothing in it reads logL): its workgroup
    // rows of w_p log L_p and of the root-frequency gradient, reduced in one
    // launch into the outputs' block [logL | dlogL/dpi | dlogL/dQ | sums]
    const double* cw = dCatWeights_ + (size_t)wIdx * dims_.C;
    const double* eig = dVVi_ + (size_t)e * 2 * S * S;
    const short* ex = auto_ ? dExp_[rootIdx] : nullptr;
    const int nb = launch_root_rows<float>(dPartials_[rootIdx], ex, cw,
                                           dFreqs_ + (size_t)fIdx * S,
                                           cs >= 0 ? scale(cs) : nullptr, dPatternWeights_,
                                           dSiteL_, qg.rootRows, dims_, stream_);
    launch_reduce_rows(qg.rootRows, nb, 1 + S, qg.out, stream_);
    launch_qgrad_slot_sums(qg.hacc, qg.slots, qg.slotPart, qg.sums, stream_);
    launch_qgrad_finish(qg.sums, eig, 0, qg.out + 1 + S, stream_);
    HIPCHECK(hipGetLastError());
    if (none) {
        // nothing to return: no copy; an asynchronous instance returns
        // without waiting
        finishCall();
        return BEAGLE_SUCCESS;
    }
    const size_t nOut = 1 + (size_t)S + 16 + 3 * (size_t)count;
    const size_t dOff = (1 + (size_t)S + 16) * sizeof(double);   // the sums' first byte
    double* dst[3] = {outD1, outD1Sq, outD2};
    if (capturing_) {
        const size_t off = captureResult(qg.out, nOut);
        cap_.outs.push_back({outLogL, off, 1, 1, true});
        if (freqs) cap_.outs.push_back({outFreqs, off + sizeof(double), S, 1, false});
        if (outQ) cap_.outs.push_back({outQ, off + (1 + (size_t)S) * sizeof(double), 16, 1, false});
        for (int k = 0; k < 3; ++k)
            if (dst[k]) cap_.outs.push_back({dst[k], off + dOff + k * sizeof(double), count, 3, false});
        return BEAGLE_SUCCESS;
    }
    const double* h = stageOut(qg.out, nOut, 0);
    sync();
    lastLogL_ = h[0];
    *outLogL = lastLogL_;
    if (freqs)
        for (int i = 0; i < S; ++i) outFreqs[i] = h[1 + i];
    if (outQ)
        for (int q = 0; q < 16; ++q) outQ[q] = h[1 + S + q];
    const double* hd = h + 1 + S + 16;
    for (int o = 0; o < count; ++o) {
        if (outD1) outD1[o] = hd[3 * o];
        if (outD1Sq) outD1Sq[o] = hd[3 * o + 1];
        if (outD2) outD2[o] = hd[3 * o + 2];
    }
    return lastLogL_ != lastLogL_ ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
}

// Pairs the two operations of each node's children (same parent, each the
// other's sibling, matrices crossed) into node jobs, in the order the second
// operation of each pair appears.  False when some operation has no such
// partner or writes scale factors.  Uses hPre_ (buildPreJobs) and fills
// acc with each job's buffer accesses.
bool Instance::makeNodeJobs(const int* ops, int count, const int* postIdx,
                            std::vector<NodeJob>& nodes, std::vector<Access>& acc) {
    std::unordered_map<int, int> open;   // parent buffer -> operation awaiting its sibling
    std::vector<std::array<int, 2>> pairs;
    for (int o = 0; o < count; ++o) {
        if (hPre_[o].mode != SCALE_NONE) return false;
        const int* q = ops + (size_t)o * BEAGLE_OP_COUNT;
        auto it = open.find(q[3]);
        if (it == open.end()) {
            open.emplace(q[3], o);
            continue;
        }
        const int a = it->second;
        const int* r = ops + (size_t)a * BEAGLE_OP_COUNT;
        if (r[5] != postIdx[o] || q[5] != postIdx[a] || r[4] != q[6] || r[6] != q[4]) return false;
        pairs.push_back({a, o});
        open.erase(it);
    }
    if (!open.empty() || pairs.empty()) return false;
    const int n = (int)pairs.size();
    nodes.assign(n, NodeJob{});
    acc.assign(n, Access{});
    for (int j = 0; j < n; ++j) {
        const PreJob& A = hPre_[pairs[j][0]];
        const PreJob& B = hPre_[pairs[j][1]];
        NodeJob& N = nodes[j];
        std::memset(&N, 0, sizeof(NodeJob));
        N.parent = A.parent;
        N.x[0] = A.xself ? A.xself : (const void*)A.sself;
        N.k[0] = A.kSelf;
        N.x[1] = B.xself ? B.xself : (const void*)B.sself;
        N.k[1] = B.kSelf;
        N.dest[0] = A.dest;
        N.dest[1] = B.dest;
        N.eParent = A.e1;
        N.ex[0] = B.e2;   // a is b's sibling
        N.ex[1] = A.e2;
        N.ed[0] = A.ed;
        N.ed[1] = B.ed;
        N.moP[0] = A.moSelf;
        N.moP[1] = B.moSelf;
        N.moD1[0] = A.moD1;
        N.moD2[0] = A.moD2;
        N.moD1[1] = B.moD1;
        N.moD2[1] = B.moD2;
        N.oidx[0] = pairs[j][0];
        N.oidx[1] = pairs[j][1];
        const int* qa = ops + (size_t)pairs[j][0] * BEAGLE_OP_COUNT;
        const int* qb = ops + (size_t)pairs[j][1] * BEAGLE_OP_COUNT;
        acc[j] = Access{{qa[3], -1, -1}, {qa[0], qb[0]}};
        // the FP32 kernels' buffer ranges (a forwarded parent: launchNodePlan)
        N.rU = -1;
        N.rEu = N.eParent ? -1 : 0;
        for (int ch = 0; ch < 2; ++ch) {
            N.rX[ch] = N.k[ch] == 0 ? -1 : 0;
            N.rS[ch] = N.k[ch] != 0 ? -1 : 0;
            N.rEx[ch] = N.ex[ch] ? -1 : 0;
            N.rD[ch] = N.dest[ch] ? dims_.P * 16 : 0;
            N.rEd[ch] = N.dest[ch] && N.ed[ch] ? dims_.P * 2 : 0;
        }
    }
    return true;
}

// Node-fused path of the derivative extension: the two operations of each
// node's children become one job that reads pre(v), x_a and x_b once.
// Returns false, launching nothing, when the operations do not pair up
// (makeNodeJobs).  Jobs are planned like operations: wide dependency
// levels batched, narrow runs fused with the parent forwarded from
// registers.
//
// With q the jobs also form the rate-matrix gradient (4 states, FP32; the
// operations' matrices checked by checkQGradOperands): each job carries its
// branches' factor offsets, and every run (each timed candidate too) writes
// its H slots from the first, so the last run's slots are what remain.
bool Instance::launchPreNodes(const int* ops, int count, const int* postIdx, double* partial,
                              NodeQGrad* q) {
    if (partialsBytes_ >= (size_t(1) << 32)) return false;   // 32-bit offsets in the kernel
    std::vector<NodeJob> nodes;
    std::vector<Access> acc;
    if (!makeNodeJobs(ops, count, postIdx, nodes, acc)) return false;
    const int n = (int)nodes.size();
    if (q) {
        for (NodeJob& J : nodes)
            for (int ch = 0; ch < 2; ++ch)
                J.mF[ch] = ops[(size_t)J.oidx[ch] * BEAGLE_OP_COUNT + 4] * dims_.C * 16;
    }
    std::vector<int> level;
    const int nl = dependencyLevels(acc, bufferCount_, level);
    auto run = [&](int k) {
        if (q) q->used = 0;
        std::vector<Segment> plan;
        if (kPlans[k].fused) {
            Segment sg{false, {}};
            for (int j = 0; j < n; ++j) sg.ops.push_back(j);
            plan.push_back(std::move(sg));
        } else {
            plan = planSegments(level, nl, kPlans[k].narrow);
        }
        launchNodePlan(nodes, plan, partial, q);
    };
    const int k = tunedChoice(shapeKey(q ? 6 : 2, n, nl), kNumPlans, forcedPlan(), run);
    if (k >= 0) run(k);
    return true;
}

// Launch node jobs as planned: fused segments forward each job's parent from
// the previous job's registers where it is that job's output.
void Instance::launchNodePlan(const std::vector<NodeJob>& nodes, const std::vector<Segment>& plan,
                              double* partial, NodeQGrad* q) {
    std::vector<NodeJob> all;
    struct Launch { int first, n; bool batched; };
    std::vector<Launch> launches;
    for (const Segment& s : plan) {
        const int first = (int)all.size();
        for (int j : s.ops) all.push_back(nodes[j]);
        if (!s.batched) {
            for (size_t j = first + 1; j < all.size(); ++j) {
                NodeJob& J = all[j];
                const NodeJob& Pv = all[j - 1];
                if (Pv.dest[0] && J.parent == Pv.dest[0]) J.kPar = 2;
                else if (Pv.dest[1] && J.parent == Pv.dest[1]) J.kPar = 3;
                if (J.kPar != 0) J.rU = J.rEu = 0;   // forwarded: nothing loaded
            }
        }
        launches.push_back({first, (int)s.ops.size(), s.batched});
    }
    const double* cw = dCatWeights_ + (size_t)preWeights_ * dims_.C;
    const bool f32 = !dp_ && node_f32_supported(dims_.C);
    const size_t perJob = (size_t)dims_.C * (f32 ? kNodeJobF32 : kNodeJobMats);
    for (size_t j = 0; j < all.size(); ++j) all[j].mbase = (int)(j * perJob);
    void* jm = grow(dJobMats_, jobMatsCap_, all.size() * perJob * realSize());
    auto d = reinterpret_cast<const NodeJob*>(uploadJobs(all.data(), all.size() * sizeof(NodeJob)));
    if (dp_) launch_gather_node_matrices<double>(d, (int)all.size(), dims_, dMatrices_, jm, stream_);
    else launch_gather_node_matrices<float>(d, (int)all.size(), dims_, dMatrices_, jm, stream_,
                                            f32, q ? q->F : nullptr);
    for (const Launch& L : launches) {
        if (dp_) launch_pre_nodes<double>(d + L.first, L.n, L.batched, levelChunk_, dims_, jm, dDummy_, cw, dPatternWeights_, partial, stream_);
        else launch_pre_nodes<float>(d + L.first, L.n, L.batched, levelChunk_, dims_, jm, dDummy_, cw, dPatternWeights_, partial, stream_, q);
    }
}

// Each pre-order job's own post-order operand (postIdx) and derivative
// matrices (d1Idx, d2Idx; without them the job's own matrix stands in).
int Instance::attachOwnOperands(int count, const int* postIdx, const int* d1Idx,
                                const int* d2Idx) {
    const size_t mstride = matrixBytes_ / realSize();
    for (int o = 0; o < count; ++o) {
        const int po = postIdx[o];
        if (!validBuffer(po) || (!dStates_[po] && !dPartials_[po]) ||
            (d1Idx && !validMatrix(d1Idx[o])) || (d2Idx && !validMatrix(d2Idx[o])))
            return BEAGLE_ERROR_OUT_OF_RANGE;
        PreJob& J = hPre_[o];
        J.xself = dStates_[po] ? nullptr : dPartials_[po];
        J.sself = dStates_[po];
        J.kSelf = dStates_[po] ? 1 : 0;
        J.moD1 = d1Idx ? (int)(d1Idx[o] * mstride) : J.moSelf;
        J.moD2 = d2Idx ? (int)(d2Idx[o] * mstride) : J.moSelf;
    }
    return BEAGLE_SUCCESS;
}

// Extension: dlogL/dQ for one eigen-system and dlogL/dpi through the root
// distribution (see beagle_amd_ext.h and k_qgrad4).  The pre-order
// partials of every parent in `ops` must be current, and every branch
// matrix must come from beagleUpdateTransitionMatrices(WithMultipleModels)
// with eigen-system `e`.
int Instance::rateMatrixGradient(int e, const int* ops, int count, const int* postIdx, int wIdx,
                                 int rootIdx, int fIdx, double* outQ, double* outFreqs) {
    const int S = dims_.S, C = dims_.C;
    if (preTransposed_ || complex_) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    if (!validEigen(e) || !eigenSet_[e] || hEigen_[e].lam.empty() || !validEigen(wIdx) ||
        hCatWeights_[wIdx].empty())
        return BEAGLE_ERROR_OUT_OF_RANGE;
    if (S != 4 || C > 8)
        return rateMatrixGradientGeneric(e, ops, count, postIdx, wIdx, rootIdx, fIdx, outQ,
                                         outFreqs);
    if (partialsBytes_ >= (size_t(1) << 32) || !dF_) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    int rc = buildPreJobs(ops, count, false, true);
    if (rc != BEAGLE_SUCCESS) return rc;
    rc = attachOwnOperands(count, postIdx, nullptr, nullptr);
    if (rc != BEAGLE_SUCCESS) return rc;
    std::vector<NodeJob> nodes;
    std::vector<Access> acc;
    if (!makeNodeJobs(ops, count, postIdx, nodes, acc)) return BEAGLE_ERROR_OUT_OF_RANGE;
    const int n = (int)nodes.size();
    // every branch's matrix must come from eigen-system e: the matrix kernel
    // then wrote its Frechet factors F (r_c t) next to it, on the device, so
    // a graph replay with new branch lengths uses the new factors
    const size_t perJob = (size_t)C * 6 * 20;
    for (int j = 0; j < n; ++j) {
        for (int ch = 0; ch < 2; ++ch) {
            const int m = ops[(size_t)nodes[j].oidx[ch] * BEAGLE_OP_COUNT + 4];
            const MatrixOrigin& org = matOrigin_[m];
            if (org.eigen != e || hCatRates_[org.rates].empty()) return BEAGLE_ERROR_OUT_OF_RANGE;
            nodes[j].mF[ch] = m * C * 16;
        }
        nodes[j].mbase = (int)(j * perJob);
    }
    const NodeJob* dj =
        reinterpret_cast<const NodeJob*>(upload(nodes.data(), (size_t)n * sizeof(NodeJob)));
    const double* eig = dVVi_ + (size_t)e * 2 * S * S;
    void* jm = grow(dJobMats_, jobMatsCap_, (size_t)n * perJob * realSize());
    const int ntiles = (dims_.P + 63) / 64;
    const int nb = (dims_.P + 255) / 256;
    const int cols = qgrad_partial_cols(dims_, n);
    const size_t nq = (size_t)16 * cols, nf = (size_t)S * nb;
    double* wk = reinterpret_cast<double*>(
        grow(dWork_, workCap_, (nq + nf + 2 * (16 + S)) * sizeof(double)));
    double* sums = wk + nq + nf;
    double* out = sums + 16 + S;
    const double* cw = dCatWeights_ + (size_t)wIdx * C;
    if (C != 4) {   // the category-per-wave kernel reads gathered job matrices
        if (dp_) launch_gather_node_matrices<double>(dj, n, dims_, dMatrices_, jm, stream_);
        else launch_gather_node_matrices<float>(dj, n, dims_, dMatrices_, jm, stream_);
    }
    if (dp_)
        launch_qgrad<double>(dj, n, dims_, jm, dMatrices_, dF_, eig, dDummy_, cw,
                             dPatternWeights_, wk, stream_);
    else
        launch_qgrad<float>(dj, n, dims_, jm, dMatrices_, dF_, eig, dDummy_, cw,
                            dPatternWeights_, wk, stream_);
    launch_reduce_rows(wk, cols, 16, sums, stream_);
    const bool freqs = outFreqs != nullptr;
    if (freqs) {
        if (!validBuffer(rootIdx) || !dPartials_[rootIdx] || !validEigen(fIdx) || hFreqs_[fIdx].empty())
            return BEAGLE_ERROR_OUT_OF_RANGE;
        if (dp_) launch_root_freq_grad<double>(dPartials_[rootIdx], cw, dFreqs_ + (size_t)fIdx * S, dPatternWeights_, wk + nq, dims_, stream_);
        else launch_root_freq_grad<float>(dPartials_[rootIdx], cw, dFreqs_ + (size_t)fIdx * S, dPatternWeights_, wk + nq, dims_, stream_);
        launch_reduce_rows(wk + nq, nb, S, sums + 16, stream_);
    }
    // dlogL/dQ = V^-T H V^T on the device, then the frequency sums
    launch_qgrad_finish(sums, eig, freqs ? S : 0, out, stream_);
    HIPCHECK(hipGetLastError());
    if (capturing_) {
        if (outQ) cap_.outs.push_back({outQ, captureResult(out, 16), 16, 1, false});
        if (freqs) cap_.outs.push_back({outFreqs, captureResult(out + 16, S), S, 1, false});
        return BEAGLE_SUCCESS;
    }
    const double* h = stageOut(out, 16 + (freqs ? S : 0), 0);
    sync();
    if (outQ)
        for (int q = 0; q < 16; ++q) outQ[q] = h[q];
    if (freqs)
        for (int s = 0; s < S; ++s) outFreqs[s] = h[16 + s];
    return BEAGLE_SUCCESS;
}

// Any S up to 64 (and 4 states with more than 8 categories): one branch
// per pre-order operation, see launch_qgrad_generic.  Same contract as the
// 4-state path.
int Instance::rateMatrixGradientGeneric(int e, const int* ops, int count, const int* postIdx,
                                        int wIdx, int rootIdx, int fIdx, double* outQ,
                                        double* outFreqs) {
    const int S = dims_.S, C = dims_.C;
    if (S > 64) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    int rc = buildPreJobs(ops, count, false, true);
    if (rc != BEAGLE_SUCCESS) return rc;
    rc = attachOwnOperands(count, postIdx, nullptr, nullptr);
    if (rc != BEAGLE_SUCCESS) return rc;
    const HostEigen& E = hEigen_[e];
    std::vector<double> tau((size_t)count * C);
    for (int o = 0; o < count; ++o) {
        const MatrixOrigin& org = matOrigin_[ops[(size_t)o * BEAGLE_OP_COUNT + 4]];
        if (org.eigen != e || hCatRates_[org.rates].empty()) return BEAGLE_ERROR_OUT_OF_RANGE;
        for (int c = 0; c < C; ++c) tau[(size_t)o * C + c] = hCatRates_[org.rates][c] * org.t;
    }
    const bool freqs = outFreqs != nullptr;
    if (freqs && (!validBuffer(rootIdx) || !dPartials_[rootIdx] || !validEigen(fIdx) ||
                  hFreqs_[fIdx].empty()))
        return BEAGLE_ERROR_OUT_OF_RANGE;
    // FP32 partials with 5-64 states and at most 4 categories: the matrix
    // cores, with each branch's P_sib and P packed (as a partials operand's
    // matrix, or by state for a tip)
    const bool mfmaQ = !dp_ && mfmaAdjoint() && C <= 4 &&
                       qgrad_mfma_lds(dims_) <= (size_t)maxLdsPerBlock_;
    const float* pack = nullptr;
    if (mfmaQ) {
        const size_t per = packFloats();
        std::vector<PackJob> pj(2 * (size_t)count);
        for (int o = 0; o < count; ++o) {
            PreJob& J = hPre_[o];
            J.pkSib = (int)(2 * o * per);
            J.pkSelf = (int)((2 * o + 1) * per);
            pj[2 * o] = PackJob{J.moSib, J.pkSib, J.ss ? PACK_ROWS : PACK_OPERAND};
            pj[2 * o + 1] = PackJob{J.moSelf, J.pkSelf, J.sself ? PACK_ROWS : PACK_OPERAND};
        }
        float* pk = static_cast<float*>(
            grow(reinterpret_cast<void*&>(dPack_), packCap_, pj.size() * per * sizeof(float)));
        auto dpj = reinterpret_cast<const PackJob*>(upload(pj.data(), pj.size() * sizeof(PackJob)));
        launch_pack_matrices<float>(dpj, (int)pj.size(), S, C, dMatrices_, pk, stream_);
        pack = pk;
    }
    // one upload: jobs | V, V^-1, eigenvalues | tau
    const size_t S2 = (size_t)S * S;
    const size_t jobBytes = (size_t)count * sizeof(PreJob);
    const size_t eOff = (jobBytes + 255) / 256 * 256;
    const size_t tOff = eOff + (2 * S2 + S) * sizeof(double);
    std::vector<char> blob(tOff + tau.size() * sizeof(double));
    std::memcpy(blob.data(), hPre_.data(), jobBytes);
    double* he = reinterpret_cast<double*>(blob.data() + eOff);
    std::copy(E.V.begin(), E.V.end(), he);
    std::copy(E.Vi.begin(), E.Vi.end(), he + S2);
    std::copy(E.lam.begin(), E.lam.end(), he + 2 * S2);
    std::memcpy(blob.data() + tOff, tau.data(), tau.size() * sizeof(double));
    char* dblob = reinterpret_cast<char*>(upload(blob.data(), blob.size()));
    // work: coef [count][Ppad] | X chunks [count][C][chunks][S^2] |
    // z [S^2][count C] | root-frequency block sums [S][nb] | sums [S^2 + S]
    const int ntiles = dims_.Ppad / 32;
    // chunks of tiles per branch (and category): enough workgroups to fill
    // the device, X partial sums of at most a few thousand patterns each
    const int chunks = mfmaQ ? std::max(1, std::min(ntiles, std::max((2048 + count - 1) / count,
                                                                     (ntiles + 39) / 40)))
                             : std::max(1, std::min(ntiles, (1024 + count * C - 1) / (count * C)));
    const int nb = root_freq_wide_blocks(dims_);
    const size_t nCoef = (size_t)count * dims_.Ppad, nX = (size_t)count * C * chunks * S2,
                 nZ = S2 * count * C, nF = (size_t)S * nb;
    double* wk = reinterpret_cast<double*>(
        grow(dWork_, workCap_, (nCoef + nX + nZ + nF + S2 + S) * sizeof(double)));
    double* coef = wk;
    double* xs = coef + nCoef;
    double* z = xs + nX;
    double* fb = z + nZ;
    double* sums = fb + nF;
    const double* cw = dCatWeights_ + (size_t)wIdx * C;
    const PreJob* dj = reinterpret_cast<const PreJob*>(dblob);
    const double* deig = reinterpret_cast<const double*>(dblob + eOff);
    const double* dtau = reinterpret_cast<const double*>(dblob + tOff);
    if (mfmaQ)
        launch_qgrad_mfma(dj, count, dims_, pack, cw, dPatternWeights_, deig, dtau, xs, chunks, z,
                          stream_);
    else if (dp_)
        launch_qgrad_generic<double>(dj, count, dims_, cw, dPatternWeights_, deig, dtau, coef, xs,
                                     chunks, z, stream_);
    else
        launch_qgrad_generic<float>(dj, count, dims_, cw, dPatternWeights_, deig, dtau, coef, xs,
                                    chunks, z, stream_);
    launch_reduce_rows(z, count * C, (int)S2, sums, stream_);
    if (freqs) {
        if (dp_) launch_root_freq_grad_wide<double>(dPartials_[rootIdx], cw, dFreqs_ + (size_t)fIdx * S, dPatternWeights_, fb, dims_, stream_);
        else launch_root_freq_grad_wide<float>(dPartials_[rootIdx], cw, dFreqs_ + (size_t)fIdx * S, dPatternWeights_, fb, dims_, stream_);
        launch_reduce_rows(fb, nb, S, sums + S2, stream_);
    }
    HIPCHECK(hipGetLastError());
    const double* h = stageOut(sums, S2 + (freqs ? S : 0), 0);
    sync();
    // dlogL/dQ = V^-T H V^T
    if (outQ) {
        std::vector<double> t(S2, 0.0);   // t = H V^T
        for (int k = 0; k < S; ++k)
            for (int b = 0; b < S; ++b) {
                double v = 0.0;
                for (int l = 0; l < S; ++l) v += h[k * S + l] * E.V[b * S + l];
                t[k * S + b] = v;
            }
        for (int a = 0; a < S; ++a)
            for (int b = 0; b < S; ++b) {
                double g = 0.0;
                for (int k = 0; k < S; ++k) g += E.Vi[k * S + a] * t[k * S + b];
                outQ[a * S + b] = g;
            }
    }
    if (freqs)
        for (int s = 0; s < S; ++s) outFreqs[s] = h[S2 + s];
    return BEAGLE_SUCCESS;
}

int Instance::waitForPartials() {
    sync();
    return BEAGLE_SUCCESS;
}

// ---- scale buffers -------------------------------------------------------------------

// cumulative[p] += sign * (LOG ? s : log s) over the listed buffers, for all
// patterns or one partition (BeagleCPUImpl.hpp:2763-2860).
int Instance::accumulateScaleFactors(const int* idx, int count, int cumIdx, double sign, int part) {
    if (auto_) return BEAGLE_SUCCESS;  // exponents already propagate with the partials
    if (!validScale(cumIdx)) return BEAGLE_ERROR_OUT_OF_RANGE;
    int p0 = 0, p1 = dims_.P;
    if (part >= 0) {
        if (part >= partitionCount_) return BEAGLE_ERROR_OUT_OF_RANGE;
        p0 = partStart_[part];
        p1 = partStart_[part + 1];
    }
    std::vector<const double*> bufs(count);
    for (int i = 0; i < count; ++i) {
        if (!validScale(idx[i])) return BEAGLE_ERROR_OUT_OF_RANGE;
        bufs[i] = scale(idx[i]);
    }
    if (count == 0) return BEAGLE_SUCCESS;
    auto d = reinterpret_cast<const double* const*>(upload(bufs.data(), count * sizeof(double*)));
    launch_accumulate(scale(cumIdx), d, count, sign, logScalers_ ? 1 : 0, p0, p1, stream_);
    HIPCHECK(hipGetLastError());
    return BEAGLE_SUCCESS;
}

// Cleared with a fill kernel, not a memset node, so the call stays
// capturable into a HIP graph.
int Instance::resetScaleFactors(int cumIdx, int part) {
    if (auto_) return BEAGLE_SUCCESS;
    if (!validScale(cumIdx)) return BEAGLE_ERROR_OUT_OF_RANGE;
    int p0 = 0, p1 = dims_.Ppad;
    if (part >= 0) {
        if (part >= partitionCount_) return BEAGLE_ERROR_OUT_OF_RANGE;
        p0 = partStart_[part];
        p1 = partStart_[part + 1];
    }
    launch_fill(scale(cumIdx) + p0, 0.0, p1 - p0, stream_);
    HIPCHECK(hipGetLastError());
    return BEAGLE_SUCCESS;
}

int Instance::copyScaleFactors(int dst, int src) {
    if (!validScale(dst) || !validScale(src)) return BEAGLE_ERROR_OUT_OF_RANGE;
    HIPCHECK(hipMemcpyAsync(scale(dst), scale(src), dims_.P * sizeof(double),
                            hipMemcpyDeviceToDevice, stream_));
    return BEAGLE_SUCCESS;
}

int Instance::getScaleFactors(int src, double* out) {
    if (!validScale(src)) return BEAGLE_ERROR_OUT_OF_RANGE;
    copyOut(out, scale(src), dims_.P * sizeof(double));
    return BEAGLE_SUCCESS;
}

// ---- likelihoods --------------------------------------------------------------------------

int Instance::calculateRootLogLikelihoods(const int* bufIdx, const int* wIdx, const int* fIdx,
                                          const int* cumIdx, int count, double* out) {
    if (count < 1) return BEAGLE_ERROR_OUT_OF_RANGE;
    if (count > 1) return rootLogLikelihoodsMulti(bufIdx, wIdx, fIdx, cumIdx, count, out);
    if (wIdx[0] < 0)
        return rootLogLikelihoodsPerCategory(bufIdx[0], fIdx[0],
                                             cumIdx ? cumIdx[0] : BEAGLE_OP_NONE, out);
    const int rc = enqueueRoot(bufIdx[0], wIdx[0], fIdx[0], cumIdx[0]);
    if (rc != BEAGLE_SUCCESS) return rc;
    if (capturing_) {
        cap_.outs.push_back({out, captureResult(dResult_, 1), 1, 1, true});
        return BEAGLE_SUCCESS;
    }
    HIPCHECK(hipMemcpyAsync(hResult_, dResult_, sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    lastLogL_ = hResult_[0];
    *out = lastLogL_;
    return lastLogL_ != lastLogL_ ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
}

// The single-root case of calculateRootLogLikelihoods with the root
// kernels' workgroup sums and the result slot in this GPU's fixed scratch
// array instead of dWork_ (which grows on demand): no device allocation and
// no host staging beyond the pinned result mirror made at creation.  The
// kernels write every workgroup sum before the reduction reads it and
// overwrite the result slot, so nothing needs clearing between calls; calls
// on one GPU that overlap in time share the array and race.  Not capturable
// into this library's graphs.
int Instance::calculateRootLogLikelihoodsNoAlloc(const int* bufIdx, const int* wIdx,
                                                 const int* fIdx, const int* cumIdx, int count,
                                                 double* out) {
    if (capturing_) captureBreak("the allocation-free root log-likelihood cannot be captured");
    if (count != 1 || wIdx[0] < 0) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    if (!noAllocScratch_) return BEAGLE_ERROR_GENERAL;
    const int b = bufIdx[0], w = wIdx[0], f = fIdx[0];
    const int cum = cumIdx ? cumIdx[0] : BEAGLE_OP_NONE;
    if (!validBuffer(b) || !dPartials_[b] || !validEigen(w) || !validEigen(f) ||
        hCatWeights_[w].empty() || hFreqs_[f].empty())
        return BEAGLE_ERROR_OUT_OF_RANGE;
    const int cs = always_ ? b - tipCount_ : (auto_ ? BEAGLE_OP_NONE : cum);
    if (cs != BEAGLE_OP_NONE && !validScale(cs)) return BEAGLE_ERROR_OUT_OF_RANGE;
    if (root_block_count(dims_, 0, dims_.P) > kNoAllocScratchDoubles - 1)
        return BEAGLE_ERROR_OUT_OF_RANGE;
    double* sums = noAllocScratch_;
    double* result = noAllocScratch_ + (kNoAllocScratchDoubles - 1);
    const short* ex = auto_ ? dExp_[b] : nullptr;
    if (dp_)
        launch_root<double>(dPartials_[b], ex, dCatWeights_ + (size_t)w * dims_.C,
                            dFreqs_ + (size_t)f * dims_.S, cs >= 0 ? scale(cs) : nullptr,
                            dPatternWeights_, dSiteL_, sums, result, dims_, stream_);
    else
        launch_root<float>(dPartials_[b], ex, dCatWeights_ + (size_t)w * dims_.C,
                           dFreqs_ + (size_t)f * dims_.S, cs >= 0 ? scale(cs) : nullptr,
                           dPatternWeights_, dSiteL_, sums, result, dims_, stream_);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(hResult_, result, sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    lastLogL_ = hResult_[0];
    *out = lastLogL_;
    return lastLogL_ != lastLogL_ ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
}

// Several roots, one likelihood (BeagleCPUImpl.hpp:2474-2586): scale
// factors apply when the first subset names a cumulative buffer (or with
// SCALING_ALWAYS, each root's own), and then every subset must name one;
// automatic scaling uses each root's exponents.  Site values are kept for
// beagleGetSiteLogLikelihoods.
int Instance::rootLogLikelihoodsMulti(const int* bufIdx, const int* wIdx, const int* fIdx,
                                      const int* cumIdx, int count, double* out) {
    const bool manual = !auto_ && (always_ || (cumIdx && cumIdx[0] != BEAGLE_OP_NONE));
    std::vector<RootSubset> subs(count);
    for (int s = 0; s < count; ++s) {
        const int b = bufIdx[s], w = wIdx[s], f = fIdx[s];
        if (!validBuffer(b) || !dPartials_[b] || !validEigen(w) || !validEigen(f) ||
            hCatWeights_[w].empty() || hFreqs_[f].empty())
            return BEAGLE_ERROR_OUT_OF_RANGE;
        const int cs = manual ? (always_ ? b - tipCount_ : cumIdx[s]) : BEAGLE_OP_NONE;
        if (manual && !validScale(cs)) return BEAGLE_ERROR_OUT_OF_RANGE;
        subs[s] = RootSubset{dPartials_[b], dCatWeights_ + (size_t)w * dims_.C,
                             dFreqs_ + (size_t)f * dims_.S, manual ? scale(cs) : nullptr,
                             auto_ ? dExp_[b] : nullptr};
    }
    const int nb = (dims_.P + 255) / 256;   // the kernel's workgroup sums
    double* work = reinterpret_cast<double*>(grow(dWork_, workCap_, (size_t)nb * sizeof(double)));
    auto d = reinterpret_cast<const RootSubset*>(upload(subs.data(), count * sizeof(RootSubset)));
    if (dp_) launch_root_multi<double>(d, count, dPatternWeights_, dSiteL_, work, dims_, stream_);
    else launch_root_multi<float>(d, count, dPatternWeights_, dSiteL_, work, dims_, stream_);
    launch_reduce_rows(work, nb, 1, dResult_, stream_);
    HIPCHECK(hipGetLastError());
    copyOut(hResult_, dResult_, sizeof(double));
    lastLogL_ = hResult_[0];
    *out = lastLogL_;
    return lastLogL_ != lastLogL_ ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
}

// categoryWeightsIndices[0] < 0: out[c P + p] = per-category site
// log-likelihoods, C x P values, in the instance's pattern order, as
// calcRootLogLikelihoodsPerCategory (BeagleCPUImpl.hpp:2589-2625) writes.
int Instance::rootLogLikelihoodsPerCategory(int b, int f, int cum, double* out) {
    if (!validBuffer(b) || !dPartials_[b] || !validEigen(f) || hFreqs_[f].empty())
        return BEAGLE_ERROR_OUT_OF_RANGE;
    const int cs = always_ ? b - tipCount_ : (auto_ ? BEAGLE_OP_NONE : cum);
    if (cs != BEAGLE_OP_NONE && !validScale(cs)) return BEAGLE_ERROR_OUT_OF_RANGE;
    const size_t n = (size_t)dims_.C * dims_.P;
    double* dst = reinterpret_cast<double*>(grow(dWork_, workCap_, n * sizeof(double)));
    if (dp_)
        launch_root_per_category<double>(dPartials_[b], dFreqs_ + (size_t)f * dims_.S,
                                         cs >= 0 ? scale(cs) : nullptr,
                                         auto_ ? dExp_[b] : nullptr, dst, dims_, stream_);
    else
        launch_root_per_category<float>(dPartials_[b], dFreqs_ + (size_t)f * dims_.S,
                                        cs >= 0 ? scale(cs) : nullptr,
                                        auto_ ? dExp_[b] : nullptr, dst, dims_, stream_);
    HIPCHECK(hipGetLastError());
    copyOut(out, dst, n * sizeof(double));
    return BEAGLE_SUCCESS;
}

// Root log-likelihood of buffer b into dResult_[0] (device), not waited for.
int Instance::enqueueRoot(int b, int w, int f, int cum) {
    if (w < 0) return BEAGLE_ERROR_NO_IMPLEMENTATION;  // per-category output: not here
    if (!validBuffer(b) || !dPartials_[b] || !validEigen(w) || !validEigen(f) ||
        hCatWeights_[w].empty() || hFreqs_[f].empty())
        return BEAGLE_ERROR_OUT_OF_RANGE;
    int cs = always_ ? b - tipCount_ : (auto_ ? BEAGLE_OP_NONE : cum);
    if (cs != BEAGLE_OP_NONE && !validScale(cs)) return BEAGLE_ERROR_OUT_OF_RANGE;
    const short* ex = auto_ ? dExp_[b] : nullptr;
    const int nb = (dims_.P + 31) / 32;   // workgroup sums of the root kernels (at most P/32)
    double* work = reinterpret_cast<double*>(grow(dWork_, workCap_, (size_t)nb * sizeof(double)));
    if (dp_)
        launch_root<double>(dPartials_[b], ex, dCatWeights_ + (size_t)w * dims_.C,
                            dFreqs_ + (size_t)f * dims_.S, cs >= 0 ? scale(cs) : nullptr,
                            dPatternWeights_, dSiteL_, work, dResult_, dims_, stream_);
    else
        launch_root<float>(dPartials_[b], ex, dCatWeights_ + (size_t)w * dims_.C,
                           dFreqs_ + (size_t)f * dims_.S, cs >= 0 ? scale(cs) : nullptr,
                           dPatternWeights_, dSiteL_, work, dResult_, dims_, stream_);
    HIPCHECK(hipGetLastError());
    return BEAGLE_SUCCESS;
}

int Instance::calculateRootLogLikelihoodsByPartition(const int* bufIdx, const int* wIdx,
                                                     const int* fIdx, const int* cumIdx,
                                                     const int* partIdx, int partitionCount,
                                                     int count, double* outByPart,
                                                     double* outSum) {
    if (count != 1 || always_) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    if (partitionCount < 1) return BEAGLE_ERROR_OUT_OF_RANGE;
    for (int q = 0; q < partitionCount; ++q) {
        const int part = partIdx[q], b = bufIdx[q], w = wIdx[q], f = fIdx[q];
        const int cs = auto_ ? BEAGLE_OP_NONE : cumIdx[q];
        if (part < 0 || part >= partitionCount_ || !validBuffer(b) || !dPartials_[b] ||
            !validEigen(w) || !validEigen(f) || hCatWeights_[w].empty() || hFreqs_[f].empty() ||
            (cs != BEAGLE_OP_NONE && !validScale(cs)))
            return BEAGLE_ERROR_OUT_OF_RANGE;
    }
    // work: one result per entry, then the workgroup sums (reused per entry)
    const size_t nbMax = (size_t)root_block_count(dims_, 0, dims_.P);
    double* res = reinterpret_cast<double*>(
        grow(dWork_, workCap_, ((size_t)partitionCount + nbMax) * sizeof(double)));
    double* bs = res + partitionCount;
    for (int q = 0; q < partitionCount; ++q) {
        const int part = partIdx[q], b = bufIdx[q], w = wIdx[q], f = fIdx[q];
        const int cs = auto_ ? BEAGLE_OP_NONE : cumIdx[q];
        const int p0 = partStart_[part], p1 = partStart_[part + 1];
        if (dp_)
            launch_root_range<double>(dPartials_[b], auto_ ? dExp_[b] : nullptr,
                                      dCatWeights_ + (size_t)w * dims_.C,
                                      dFreqs_ + (size_t)f * dims_.S, cs >= 0 ? scale(cs) : nullptr,
                                      dPatternWeights_, dSiteL_, bs, res + q, p0, p1, dims_,
                                      stream_);
        else
            launch_root_range<float>(dPartials_[b], auto_ ? dExp_[b] : nullptr,
                                     dCatWeights_ + (size_t)w * dims_.C,
                                     dFreqs_ + (size_t)f * dims_.S, cs >= 0 ? scale(cs) : nullptr,
                                     dPatternWeights_, dSiteL_, bs, res + q, p0, p1, dims_,
                                     stream_);
        HIPCHECK(hipGetLastError());
    }
    const double* h = stageOut(res, partitionCount, 0);
    sync();
    double total = 0.0;
    for (int q = 0; q < partitionCount; ++q) {
        outByPart[q] = h[q];
        total += h[q];
    }
    *outSum = total;
    lastLogL_ = total;
    return total != total ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
}

// Derivatives with client matrices D per node (beagle.cpp:1794-1824 passes
// no rates and no scale buffers): see k_edge_deriv.
int Instance::calculateEdgeDerivatives(const int* postIdx, const int* preIdx, const int* dIdx,
                                       const int* wIdx, int count, double* outD, double* outSum,
                                       double* outSumSq) {
    const int w = wIdx[0];
    if (!validEigen(w) || hCatWeights_[w].empty()) return BEAGLE_ERROR_OUT_OF_RANGE;
    std::vector<EdgeJob> jobs(count);
    for (int n = 0; n < count; ++n) {
        const int po = postIdx[n], pr = preIdx[n], dm = dIdx[n];
        if (!validBuffer(po) || !validBuffer(pr) || !validMatrix(dm) || !dPartials_[pr] ||
            (!dStates_[po] && !dPartials_[po]))
            return BEAGLE_ERROR_OUT_OF_RANGE;
        jobs[n].pre = dPartials_[pr];
        jobs[n].states = dStates_[po];
        jobs[n].post = dStates_[po] ? nullptr : dPartials_[po];
        jobs[n].D = matrix(dm);
        jobs[n].scale = 0.0;
    }
    if (count == 0) return BEAGLE_SUCCESS;
    if (mfmaAdjoint()) {
        // sites go to a buffer of their own after the work area
        double* sites = nullptr;
        if (outD) {
            sites = reinterpret_cast<double*>(
                grow(dSites_, sitesCap_, (size_t)count * dims_.P * sizeof(double)));
        }
        double* dSums = nullptr;
        const int rc = enqueueDerivMfma(postIdx, preIdx, dIdx, nullptr, w, count, sites, &dSums);
        if (rc != BEAGLE_SUCCESS) return rc;
        std::vector<double> sums(2 * (size_t)count);
        copyOut(sums.data(), dSums, sums.size() * sizeof(double));
        for (int n = 0; n < count; ++n) {
            if (outSum) outSum[n] = sums[2 * n];
            if (outSumSq) outSumSq[n] = sums[2 * n + 1];
        }
        if (outD) copyOut(outD, sites, (size_t)count * dims_.P * sizeof(double));
        return BEAGLE_SUCCESS;
    }
    const int nb = (dims_.P + 255) / 256;
    const size_t work = (size_t)2 * count * nb + 2 * (size_t)count;
    size_t extra = outD ? (size_t)count * dims_.P : 0;
    double* wk = reinterpret_cast<double*>(grow(dWork_, workCap_, (work + 2 * count + extra) * sizeof(double)));
    double* dSums = wk + work;
    double* dSites = outD ? dSums + 2 * count : nullptr;
    auto dj = reinterpret_cast<const EdgeJob*>(uploadJobs(jobs.data(), count * sizeof(EdgeJob)));
    if (dp_)
        launch_edge_derivatives<double>(dj, count, dCatWeights_ + (size_t)w * dims_.C,
                                        dPatternWeights_, dSites, wk, dSums, dSums + count,
                                        preTransposed_ ? 1 : 0, dims_, stream_);
    else
        launch_edge_derivatives<float>(dj, count, dCatWeights_ + (size_t)w * dims_.C,
                                       dPatternWeights_, dSites, wk, dSums, dSums + count,
                                       preTransposed_ ? 1 : 0, dims_, stream_);
    HIPCHECK(hipGetLastError());
    std::vector<double> sums(2 * count);
    copyOut(sums.data(), dSums, 2 * count * sizeof(double));
    if (outSum) std::copy(sums.begin(), sums.begin() + count, outSum);
    if (outSumSq) std::copy(sums.begin() + count, sums.end(), outSumSq);
    if (outD) copyOut(outD, dSites, (size_t)count * dims_.P * sizeof(double));
    return BEAGLE_SUCCESS;
}

int Instance::calculateCrossProducts(const int* postIdx, const int* preIdx, const int* rIdx,
                                     const int* wIdx, const double* lengths, int count,
                                     double* out) {
    const int r = rIdx[0], w = wIdx[0];
    if (!validEigen(r) || !validEigen(w) || hCatRates_[r].empty() || hCatWeights_[w].empty())
        return BEAGLE_ERROR_OUT_OF_RANGE;
    // generic kernel: S*S sums over <= 16 per lane, and its LDS staging of
    // 16 patterns x C x S (pre and post) must fit in 64 KiB
    if (dims_.S > 64 || (dims_.S != 4 && (2 * 16 * dims_.C * dims_.S + 16 * dims_.C) * 8 > 65536))
        return BEAGLE_ERROR_NO_IMPLEMENTATION;
    std::vector<EdgeJob> jobs(count);
    for (int n = 0; n < count; ++n) {
        const int po = postIdx[n], pr = preIdx[n];
        if (!validBuffer(po) || !validBuffer(pr) || !dPartials_[pr] ||
            (!dStates_[po] && !dPartials_[po]))
            return BEAGLE_ERROR_OUT_OF_RANGE;
        jobs[n].pre = dPartials_[pr];
        jobs[n].states = dStates_[po];
        jobs[n].post = dStates_[po] ? nullptr : dPartials_[po];
        jobs[n].D = nullptr;
        jobs[n].scale = lengths[n];
    }
    if (count == 0) return BEAGLE_SUCCESS;
    const int S2 = dims_.S * dims_.S;
    const size_t rows = cross_products_rows(dims_);
    double* wk = reinterpret_cast<double*>(
        grow(dWork_, workCap_, (rows * S2 + S2) * sizeof(double)));
    double* dOut = wk + rows * S2;
    copyIn(dOut, out, S2 * sizeof(double));
    auto dj = reinterpret_cast<const EdgeJob*>(uploadJobs(jobs.data(), count * sizeof(EdgeJob)));
    if (dp_)
        launch_cross_products<double>(dj, count, dCatWeights_ + (size_t)w * dims_.C,
                                      dCatRates_ + (size_t)r * dims_.C, dPatternWeights_, wk, dOut,
                                      dims_, stream_);
    else
        launch_cross_products<float>(dj, count, dCatWeights_ + (size_t)w * dims_.C,
                                     dCatRates_ + (size_t)r * dims_.C, dPatternWeights_, wk, dOut,
                                     dims_, stream_);
    HIPCHECK(hipGetLastError());
    copyOut(out, dOut, S2 * sizeof(double));
    return BEAGLE_SUCCESS;
}

// Edge log-likelihood (+ derivatives along the edge); scaling-always sums
// the two node scale buffers (BeagleCPUImpl.hpp:2918-3005).
int Instance::calculateEdgeLogLikelihoods(const int* parIdx, const int* childIdx, const int* pIdx,
                                          const int* d1Idx, const int* d2Idx, const int* wIdx,
                                          const int* fIdx, const int* cumIdx, int count,
                                          double* outL, double* outD1, double* outD2) {
    if (count != 1) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    const int par = parIdx[0], ch = childIdx[0], pm = pIdx[0], w = wIdx[0], f = fIdx[0];
    if (!validBuffer(par) || !dPartials_[par] || !validBuffer(ch) ||
        (!dStates_[ch] && !dPartials_[ch]) || !validMatrix(pm) || !validEigen(w) ||
        !validEigen(f))
        return BEAGLE_ERROR_OUT_OF_RANGE;
    const void* D1 = nullptr;
    const void* D2 = nullptr;
    if (d1Idx) {
        if (!validMatrix(d1Idx[0])) return BEAGLE_ERROR_OUT_OF_RANGE;
        D1 = matrix(d1Idx[0]);
        if (d2Idx) {
            if (!validMatrix(d2Idx[0])) return BEAGLE_ERROR_OUT_OF_RANGE;
            D2 = matrix(d2Idx[0]);
        }
    }
    const double* cumA = nullptr;
    const double* cumB = nullptr;
    if (always_) {
        if (par >= tipCount_) cumA = scale(par - tipCount_);
        if (ch >= tipCount_) cumB = scale(ch - tipCount_);
    } else if (!auto_ && cumIdx[0] != BEAGLE_OP_NONE) {
        if (!validScale(cumIdx[0])) return BEAGLE_ERROR_OUT_OF_RANGE;
        cumA = scale(cumIdx[0]);
    }
    const short* exA = auto_ ? dExp_[par] : nullptr;
    const short* exB = (auto_ && !dStates_[ch]) ? dExp_[ch] : nullptr;
    const int nb = (dims_.P + 255) / 256;
    double* wk = reinterpret_cast<double*>(grow(dWork_, workCap_, (size_t)3 * nb * sizeof(double)));
    if (dp_)
        launch_edge_loglik<double>(dPartials_[par], dPartials_[ch], dStates_[ch], matrix(pm), D1,
                                   D2, dCatWeights_ + (size_t)w * dims_.C,
                                   dFreqs_ + (size_t)f * dims_.S, cumA, cumB, exA, exB,
                                   dPatternWeights_, dSiteL_, dSiteD1_, dSiteD2_, wk, dResult_,
                                   0, dims_.P, dims_, stream_);
    else
        launch_edge_loglik<float>(dPartials_[par], dPartials_[ch], dStates_[ch], matrix(pm), D1,
                                  D2, dCatWeights_ + (size_t)w * dims_.C,
                                  dFreqs_ + (size_t)f * dims_.S, cumA, cumB, exA, exB,
                                  dPatternWeights_, dSiteL_, dSiteD1_, dSiteD2_, wk, dResult_,
                                  0, dims_.P, dims_, stream_);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(hResult_, dResult_, 3 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    lastLogL_ = hResult_[0];
    *outL = lastLogL_;
    haveDerivs_ = D1 != nullptr;
    lastD1_ = D1 ? hResult_[1] : 0.0;
    lastD2_ = D2 ? hResult_[2] : 0.0;
    if (outD1 && D1) *outD1 = lastD1_;
    if (outD2 && D2) *outD2 = lastD2_;
    return lastLogL_ != lastLogL_ ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
}

// Edge log-likelihoods (and derivatives) per pattern partition
// (BeagleCPUImpl.hpp:3008-3110, 3452-3590): entry q is the edge
// parIdx[q] - childIdx[q] over partition partIdx[q] with its own matrices,
// weights, frequencies and cumulative buffer.  As the reference: one set
// (count = 1), no automatic or always-on scaling, derivatives both or none.
int Instance::calculateEdgeLogLikelihoodsByPartition(
    const int* parIdx, const int* childIdx, const int* pIdx, const int* d1Idx, const int* d2Idx,
    const int* wIdx, const int* fIdx, const int* cumIdx, const int* partIdx, int partitionCount,
    int count, double* outLByPart, double* outL, double* outD1ByPart, double* outD1,
    double* outD2ByPart, double* outD2) {
    if (count != 1 || auto_ || always_) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    if ((d1Idx != nullptr) != (d2Idx != nullptr)) return BEAGLE_ERROR_NO_IMPLEMENTATION;
    if (partitionCount < 1) return BEAGLE_ERROR_OUT_OF_RANGE;
    const bool derivs = d1Idx != nullptr;
    for (int q = 0; q < partitionCount; ++q) {
        const int par = parIdx[q], ch = childIdx[q], part = partIdx[q];
        if (part < 0 || part >= partitionCount_ || !validBuffer(par) || !dPartials_[par] ||
            !validBuffer(ch) || (!dStates_[ch] && !dPartials_[ch]) || !validMatrix(pIdx[q]) ||
            !validEigen(wIdx[q]) || !validEigen(fIdx[q]) || hCatWeights_[wIdx[q]].empty() ||
            hFreqs_[fIdx[q]].empty() ||
            (derivs && (!validMatrix(d1Idx[q]) || !validMatrix(d2Idx[q]))) ||
            (cumIdx[q] != BEAGLE_OP_NONE && !validScale(cumIdx[q])))
            return BEAGLE_ERROR_OUT_OF_RANGE;
    }
    // work: three results per entry, then the workgroup sums (reused)
    const size_t nbMax = (size_t)(dims_.P + 255) / 256;
    double* res = reinterpret_cast<double*>(
        grow(dWork_, workCap_, (3 * (size_t)partitionCount + 3 * nbMax) * sizeof(double)));
    double* bs = res + 3 * (size_t)partitionCount;
    for (int q = 0; q < partitionCount; ++q) {
        const int par = parIdx[q], ch = childIdx[q], part = partIdx[q];
        const int p0 = partStart_[part], p1 = partStart_[part + 1];
        const void* D1 = derivs ? matrix(d1Idx[q]) : nullptr;
        const void* D2 = derivs ? matrix(d2Idx[q]) : nullptr;
        const double* cum = cumIdx[q] != BEAGLE_OP_NONE ? scale(cumIdx[q]) : nullptr;
        const double* cw = dCatWeights_ + (size_t)wIdx[q] * dims_.C;
        const double* fr = dFreqs_ + (size_t)fIdx[q] * dims_.S;
        if (dp_)
            launch_edge_loglik<double>(dPartials_[par], dPartials_[ch], dStates_[ch],
                                       matrix(pIdx[q]), D1, D2, cw, fr, cum, nullptr, nullptr,
                                       nullptr, dPatternWeights_, dSiteL_, dSiteD1_, dSiteD2_, bs,
                                       res + 3 * q, p0, p1, dims_, stream_);
        else
            launch_edge_loglik<float>(dPartials_[par], dPartials_[ch], dStates_[ch],
                                      matrix(pIdx[q]), D1, D2, cw, fr, cum, nullptr, nullptr,
                                      nullptr, dPatternWeights_, dSiteL_, dSiteD1_, dSiteD2_, bs,
                                      res + 3 * q, p0, p1, dims_, stream_);
        HIPCHECK(hipGetLastError());
    }
    const double* h = stageOut(res, 3 * (size_t)partitionCount, 0);
    sync();
    double tL = 0.0, t1 = 0.0, t2 = 0.0;
    for (int q = 0; q < partitionCount; ++q) {
        outLByPart[q] = h[3 * q];
        tL += h[3 * q];
        if (derivs) {
            outD1ByPart[q] = h[3 * q + 1];
            outD2ByPart[q] = h[3 * q + 2];
            t1 += h[3 * q + 1];
            t2 += h[3 * q + 2];
        }
    }
    int rc = BEAGLE_SUCCESS;
    if (derivs) {
        *outD1 = t1;
        *outD2 = t2;
        if (t1 != t1 || t2 != t2) rc = BEAGLE_ERROR_FLOATING_POINT;
    }
    *outL = tL;
    lastLogL_ = tL;
    haveDerivs_ = derivs;
    lastD1_ = t1;
    lastD2_ = t2;
    if (tL != tL) rc = BEAGLE_ERROR_FLOATING_POINT;
    return rc;
}

int Instance::getLogLikelihood(double* out) {
    sync();
    *out = lastLogL_;
    return lastLogL_ != lastLogL_ ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
}

int Instance::getDerivatives(double* d1, double* d2) {
    sync();
    if (d1) *d1 = lastD1_;
    if (d2) *d2 = lastD2_;
    return BEAGLE_SUCCESS;
}

int Instance::getSiteLogLikelihoods(double* out) {
    if (newOrder_.empty()) {
        copyOut(out, dSiteL_, dims_.P * sizeof(double));
    } else {
        // reordered patterns: back to the client's order (BeagleCPUImpl.hpp:1009-1021)
        std::vector<double> tmp(dims_.P);
        copyOut(tmp.data(), dSiteL_, dims_.P * sizeof(double));
        for (int k = 0; k < dims_.P; ++k) out[k] = tmp[newOrder_[k]];
    }
    return BEAGLE_SUCCESS;
}

int Instance::getSiteDerivatives(double* d1, double* d2) {
    copyOut(d1, dSiteD1_, dims_.P * sizeof(double));
    if (d2) copyOut(d2, dSiteD2_, dims_.P * sizeof(double));
    return BEAGLE_SUCCESS;
}

// ---- HIP-graph capture ---------------------------------------------------------

void Instance::captureBreak(const char* why) {
    if (captureError_.empty()) captureError_ = why;
    throw std::runtime_error(why);
}

size_t Instance::captureAlloc(size_t bytes) {
    const size_t off = (cap_.used + 255) / 256 * 256;
    if (off + bytes > kCaptureArena) captureBreak("capture arena full");
    cap_.used = off + bytes;
    return off;
}

// An output of the captured calls, written by a copy kernel into the
// graph's host memory at its place in the stream (no copy node, whose
// engine would add its own latency to every replay).
size_t Instance::captureResult(const double* dsrc, size_t n) {
    const size_t off = captureAlloc(n * sizeof(double));
    launch_copy_to_host(reinterpret_cast<double*>(cap_.hostDev + off), dsrc, (long)n, stream_);
    return off;
}

// Calls between beginCapture and endCapture are recorded, not run (their
// outputs are written at each replay's wait); calls that must wait for the
// device (getters, copies, the generic-state derivative path) fail and
// spoil the capture.  Work buffers must already have their size, i.e. the
// same calls run once before capturing.
int Instance::beginCapture() {
    if (capturing_) return BEAGLE_ERROR_GENERAL;
    sync();
    cap_ = Graph{};
    HIPCHECK(hipHostMalloc((void**)&cap_.host, kCaptureArena,
                           hipHostMallocCoherent | hipHostMallocMapped));
    HIPCHECK(hipHostGetDevicePointer((void**)&cap_.hostDev, cap_.host, 0));
    HIPCHECK(hipMalloc((void**)&cap_.dev, kCaptureArena));
    captureError_.clear();
    HIPCHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
    capturing_ = true;
    return BEAGLE_SUCCESS;
}

int Instance::endCapture(int* outGraph) {
    if (!capturing_) return BEAGLE_ERROR_GENERAL;
    capturing_ = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream_, &g);
    if (e != hipSuccess || !captureError_.empty() || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        (void)hipHostFree(cap_.host);
        (void)hipFree(cap_.dev);
        cap_ = Graph{};
        return BEAGLE_ERROR_GENERAL;
    }
    const hipError_t ie = hipGraphInstantiate(&cap_.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    // the job lists the captured calls uploaded, to the device once
    const hipError_t ce = ie == hipSuccess ? hipMemcpy(cap_.dev, cap_.host, cap_.used,
                                                       hipMemcpyHostToDevice)
                                           : hipSuccess;
    if (ie != hipSuccess || ce != hipSuccess) {
        if (ie == hipSuccess) (void)hipGraphExecDestroy(cap_.exec);
        (void)hipGetLastError();
        (void)hipHostFree(cap_.host);
        (void)hipFree(cap_.dev);
        cap_ = Graph{};
        return BEAGLE_ERROR_GENERAL;
    }
    cap_.live = true;
    graphs_.push_back(cap_);
    cap_ = Graph{};
    *outGraph = (int)graphs_.size() - 1;
    return BEAGLE_SUCCESS;
}

int Instance::launchGraph(int graph) {
    if (capturing_) return BEAGLE_ERROR_GENERAL;
    if (graph < 0 || graph >= (int)graphs_.size() || !graphs_[graph].live)
        return BEAGLE_ERROR_OUT_OF_RANGE;
    HIPCHECK(hipGraphLaunch(graphs_[graph].exec, stream_));
    if (!(flags_ & BEAGLE_FLAG_COMPUTATION_ASYNCH)) return waitGraph(graph);
    return BEAGLE_SUCCESS;
}

int Instance::waitGraph(int graph) {
    if (graph < 0 || graph >= (int)graphs_.size() || !graphs_[graph].live)
        return BEAGLE_ERROR_OUT_OF_RANGE;
    sync();
    const Graph& G = graphs_[graph];
    bool nan = false;
    for (const CapturedOut& o : G.outs) {
        const double* v = reinterpret_cast<const double*>(G.host + o.src);
        for (int k = 0; k < o.n; ++k) o.dst[k] = v[(size_t)k * o.stride];
        if (o.logL) {
            lastLogL_ = v[0];
            nan |= v[0] != v[0];
        }
    }
    return nan ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
}

// New branch lengths for the captured transition-matrix updates, in the
// order their matrix jobs were captured (one per matrix index and call).
int Instance::setGraphBranchLengths(int graph, const double* lengths, int count) {
    if (graph < 0 || graph >= (int)graphs_.size() || !graphs_[graph].live)
        return BEAGLE_ERROR_OUT_OF_RANGE;
    Graph& G = graphs_[graph];
    int total = 0;
    for (const auto& m : G.matJobs) total += m.second;
    if (count != total) return BEAGLE_ERROR_OUT_OF_RANGE;
    sync();   // a replay in flight still reads the staged jobs
    int k = 0;
    for (const auto& m : G.matJobs) {
        MatrixJob* J = reinterpret_cast<MatrixJob*>(G.host + m.first);
        for (int j = 0; j < m.second; ++j, ++k) {
            J[j].t = lengths[k];
            const long idx = (static_cast<char*>(J[j].P) - static_cast<char*>(dMatrices_)) /
                             (long)matrixBytes_;
            if (idx >= 0 && idx < matrixCount_ && J[j].cat < 0) matOrigin_[idx].t = lengths[k];
        }
    }
    return BEAGLE_SUCCESS;
}

int Instance::destroyGraph(int graph) {
    if (graph < 0 || graph >= (int)graphs_.size() || !graphs_[graph].live)
        return BEAGLE_ERROR_OUT_OF_RANGE;
    sync();
    Graph& G = graphs_[graph];
    (void)hipGraphExecDestroy(G.exec);
    (void)hipHostFree(G.host);
    (void)hipFree(G.dev);
    G = Graph{};
    return BEAGLE_SUCCESS;
}

}  // namespace hipbeagle
