/*
 * beagle_oracle.c -- FP64 CPU restatement of the reference beagle-lib
 * likelihood path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, never the product: only tests/, the smoke test
 * in __graft_entry__.py and the cpu_baseline leg of bench.py may load it.
 * The shipped library (beagle-lib_amd/) neither links nor calls it, and has
 * no CPU fallback.
 *
 * It exports the reference's API shape under the prefix "orc" instead of
 * "beagle" (orcCreateInstance, orcUpdatePartials, ...), with the same
 * arguments, so a single test driver can run the reference build
 * (oracle/_ref), this restatement and the HIP plugin on identical inputs.
 *
 * Layout follows the reference CPU implementation, which this restates:
 *   partials  [category][pattern][state], stride stateCount, no padding
 *             (libhmsbeagle/CPU/BeagleCPUImpl.hpp:4485-4490)
 *   matrices  [category][row i][col j], row length stateCount+1; column
 *             stateCount is the pad read by a missing tip state: 1.0 for
 *             P(t), 0.0 for derivative matrices
 *             (libhmsbeagle/CPU/EigenDecompositionCube.hpp:222-225,268-271)
 *   scale     one double per pattern
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   - every function is compared with the reference build oracle/_ref on
 *     randomized problems (4, 20 and 61 states, tips as states and as
 *     partials, missing states, manual raw/log scaling, scaling-always);
 *   - committed golden vectors under tests/golden/ produced by that build;
 *   - closed forms: JC69 P(t), rows of P summing to 1, P(0)=I, dP/dt(0)=Q;
 *   - brute-force enumeration of ancestral states on 3-5 taxon trees;
 *   - central finite differences for the edge derivatives;
 *   - the reference's own known answers (hmctest5 logL and per-site
 *     gradients, tinytest, hmcGapTest) restated as test vectors.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* Values of the reference's public enums (libhmsbeagle/beagle.h:110-205). */
enum {
    ORC_SUCCESS = 0, ORC_ERROR_GENERAL = -1, ORC_ERROR_OUT_OF_MEMORY = -2,
    ORC_ERROR_UNINITIALIZED_INSTANCE = -4, ORC_ERROR_OUT_OF_RANGE = -5,
    ORC_ERROR_NO_IMPLEMENTATION = -7, ORC_ERROR_FLOATING_POINT = -8
};
#define F_SCALING_MANUAL   (1L << 6)
#define F_SCALING_AUTO     (1L << 7)
#define F_SCALING_ALWAYS   (1L << 8)
#define F_SCALING_DYNAMIC  (1L << 25)
#define F_SCALERS_RAW      (1L << 9)
#define F_SCALERS_LOG      (1L << 10)
#define F_INVEVEC_STANDARD (1L << 20)
#define F_INVEVEC_TRANSPOSED (1L << 21)
#define F_PRECISION_DOUBLE (1L << 1)
#define F_EIGEN_REAL       (1L << 4)
#define OP_NONE (-1)

typedef struct {
    int tipCount, bufferCount, S, P, eigenCount, matrixCount, C, scaleCount;
    long flags;
    double** partials;   /* bufferCount, NULL until allocated */
    int** tipStates;     /* bufferCount, NULL unless set */
    double** evals;      /* eigenCount x S */
    double** cube;       /* eigenCount x S^3, C_ijk = V_ik Vinv_kj */
    double** freqs;      /* eigenCount x S */
    double** catWeights; /* eigenCount x C */
    double** catRates;   /* eigenCount x C */
    double* patternWeights;
    double** matrices;   /* matrixCount x C*S*(S+1) */
    double** scale;      /* scaleCount x P */
    double* siteLogL;    /* last root evaluation, per pattern */
    double* siteD1;      /* last edge-derivative evaluation */
    double* siteD2;
    double* numTmp;      /* P */
    double* denTmp;      /* P */
    int* partStart;      /* partition boundaries, partitionCount+1 */
    int partitionCount;
    int* newOrder;       /* after a pattern reorder: internal index of pattern i */
} Orc;

#define MAX_INSTANCES 4096
static Orc* g_inst[MAX_INSTANCES];

static Orc* get(int i) {
    if (i < 0 || i >= MAX_INSTANCES) return NULL;
    return g_inst[i];
}

static void* xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

/* Scaling-mode precedence AUTO > ALWAYS > DYNAMIC > LOG > RAW
 * (libhmsbeagle/CPU/BeagleCPUImpl.hpp:294-311). */
int orcCreateInstance(int tipCount, int partialsBufferCount, int compactBufferCount,
                      int stateCount, int patternCount, int eigenBufferCount,
                      int matrixBufferCount, int categoryCount, int scaleBufferCount,
                      int* resourceList, int resourceCount, long preferenceFlags,
                      long requirementFlags, void* returnInfo) {
    (void)resourceList; (void)resourceCount; (void)returnInfo;
    long both = preferenceFlags | requirementFlags;
    int slot = -1;
    for (int i = 0; i < MAX_INSTANCES; i++) if (!g_inst[i]) { slot = i; break; }
    if (slot < 0) return ORC_ERROR_GENERAL;
    if (stateCount < 2 || patternCount < 1 || categoryCount < 1) return ORC_ERROR_OUT_OF_RANGE;
    Orc* o = (Orc*)xcalloc(1, sizeof(Orc));
    o->tipCount = tipCount;
    o->bufferCount = partialsBufferCount + compactBufferCount;
    o->S = stateCount; o->P = patternCount; o->C = categoryCount;
    o->eigenCount = eigenBufferCount; o->matrixCount = matrixBufferCount;
    o->scaleCount = scaleBufferCount;
    int internal = o->bufferCount - tipCount;
    if (both & F_SCALING_AUTO) {
        free(o);
        return ORC_ERROR_NO_IMPLEMENTATION;
    } else if (both & F_SCALING_ALWAYS) {
        o->flags |= F_SCALING_ALWAYS | F_SCALERS_LOG;
        o->scaleCount = internal + 1;
    } else if (both & F_SCALING_DYNAMIC) {
        free(o);
        return ORC_ERROR_NO_IMPLEMENTATION;
    } else if (both & F_SCALERS_LOG) {
        o->flags |= F_SCALING_MANUAL | F_SCALERS_LOG;
    } else {
        o->flags |= F_SCALING_MANUAL | F_SCALERS_RAW;
    }
    o->flags |= F_EIGEN_REAL | F_PRECISION_DOUBLE;
    o->flags |= (both & F_INVEVEC_TRANSPOSED) ? F_INVEVEC_TRANSPOSED : F_INVEVEC_STANDARD;
    int S = o->S, P = o->P, C = o->C;
    o->partials = (double**)xcalloc(o->bufferCount, sizeof(double*));
    o->tipStates = (int**)xcalloc(o->bufferCount, sizeof(int*));
    for (int b = tipCount; b < o->bufferCount; b++)
        o->partials[b] = (double*)xcalloc((size_t)C * P * S, sizeof(double));
    o->evals = (double**)xcalloc(o->eigenCount, sizeof(double*));
    o->cube = (double**)xcalloc(o->eigenCount, sizeof(double*));
    o->freqs = (double**)xcalloc(o->eigenCount, sizeof(double*));
    o->catWeights = (double**)xcalloc(o->eigenCount, sizeof(double*));
    o->catRates = (double**)xcalloc(o->eigenCount > 0 ? o->eigenCount : 1, sizeof(double*));
    for (int e = 0; e < o->eigenCount; e++) {
        o->evals[e] = (double*)xcalloc(S, sizeof(double));
        o->cube[e] = (double*)xcalloc((size_t)S * S * S, sizeof(double));
    }
    o->patternWeights = (double*)xcalloc(P, sizeof(double));
    o->matrices = (double**)xcalloc(o->matrixCount, sizeof(double*));
    for (int m = 0; m < o->matrixCount; m++)
        o->matrices[m] = (double*)xcalloc((size_t)C * S * (S + 1), sizeof(double));
    o->scale = (double**)xcalloc(o->scaleCount, sizeof(double*));
    for (int s = 0; s < o->scaleCount; s++)
        o->scale[s] = (double*)xcalloc(P, sizeof(double));
    o->siteLogL = (double*)xcalloc(P, sizeof(double));
    o->siteD1 = (double*)xcalloc(P, sizeof(double));
    o->siteD2 = (double*)xcalloc(P, sizeof(double));
    o->numTmp = (double*)xcalloc(P, sizeof(double));
    o->denTmp = (double*)xcalloc(P, sizeof(double));
    o->partitionCount = 1;
    o->partStart = (int*)xcalloc(2, sizeof(int));
    o->partStart[0] = 0; o->partStart[1] = P;
    g_inst[slot] = o;
    return slot;
}

int orcFinalizeInstance(int instance) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    for (int b = 0; b < o->bufferCount; b++) { free(o->partials[b]); free(o->tipStates[b]); }
    for (int e = 0; e < o->eigenCount; e++) {
        free(o->evals[e]); free(o->cube[e]); free(o->freqs[e]); free(o->catWeights[e]);
        free(o->catRates[e]);
    }
    for (int m = 0; m < o->matrixCount; m++) free(o->matrices[m]);
    for (int s = 0; s < o->scaleCount; s++) free(o->scale[s]);
    free(o->partials); free(o->tipStates); free(o->evals); free(o->cube); free(o->freqs);
    free(o->catWeights); free(o->catRates); free(o->patternWeights); free(o->matrices);
    free(o->scale); free(o->siteLogL); free(o->siteD1); free(o->siteD2); free(o->numTmp);
    free(o->denTmp); free(o->partStart); free(o->newOrder);
    free(o);
    g_inst[instance] = NULL;
    return ORC_SUCCESS;
}

/* States >= S become S, meaning missing (BeagleCPUImpl.hpp:570-584). */
int orcSetTipStates(int instance, int tipIndex, const int* inStates) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (tipIndex < 0 || tipIndex >= o->tipCount) return ORC_ERROR_OUT_OF_RANGE;
    if (!o->tipStates[tipIndex]) o->tipStates[tipIndex] = (int*)xcalloc(o->P, sizeof(int));
    for (int k = 0; k < o->P; k++)
        o->tipStates[tipIndex][k] = inStates[k] < o->S ? inStates[k] : o->S;
    return ORC_SUCCESS;
}

/* Replicated over categories (BeagleCPUImpl.hpp:587-620). */
int orcSetTipPartials(int instance, int tipIndex, const double* in) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (tipIndex < 0 || tipIndex >= o->tipCount) return ORC_ERROR_OUT_OF_RANGE;
    size_t n = (size_t)o->P * o->S;
    if (!o->partials[tipIndex]) o->partials[tipIndex] = (double*)xcalloc(n * o->C, sizeof(double));
    for (int l = 0; l < o->C; l++) memcpy(o->partials[tipIndex] + l * n, in, n * sizeof(double));
    return ORC_SUCCESS;
}

/* BeagleCPUImpl.hpp:680-709 */
int orcSetPartials(int instance, int bufferIndex, const double* in) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (bufferIndex < 0 || bufferIndex >= o->bufferCount) return ORC_ERROR_OUT_OF_RANGE;
    size_t n = (size_t)o->C * o->P * o->S;
    if (!o->partials[bufferIndex]) o->partials[bufferIndex] = (double*)xcalloc(n, sizeof(double));
    memcpy(o->partials[bufferIndex], in, n * sizeof(double));
    return ORC_SUCCESS;
}

/* Root pre-order partials = state frequencies replicated over C x P
 * (BeagleCPUImpl.hpp:623-677). */
int orcSetRootPrePartials(int instance, const int* bufferIndices, const int* freqIndices, int count) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (count != 1) return ORC_ERROR_NO_IMPLEMENTATION;
    int b = bufferIndices[0];
    if (b < 0 || b >= o->bufferCount) return ORC_ERROR_OUT_OF_RANGE;
    size_t n = (size_t)o->C * o->P * o->S;
    if (!o->partials[b]) o->partials[b] = (double*)xcalloc(n, sizeof(double));
    const double* f = o->freqs[freqIndices[0]];
    for (int l = 0; l < o->C; l++)
        for (int k = 0; k < o->P; k++)
            memcpy(o->partials[b] + ((size_t)l * o->P + k) * o->S, f, o->S * sizeof(double));
    return ORC_SUCCESS;
}

/* With a cumulative scale index, the reference multiplies only the first
 * patternCount*stateCount outputs (category 0) by exp(scale[pattern])
 * (BeagleCPUImpl.hpp:743-754); restated as is. */
int orcGetPartials(int instance, int bufferIndex, int scaleIndex, double* out) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (bufferIndex < 0 || bufferIndex >= o->bufferCount || !o->partials[bufferIndex])
        return ORC_ERROR_OUT_OF_RANGE;
    size_t n = (size_t)o->C * o->P * o->S;
    memcpy(out, o->partials[bufferIndex], n * sizeof(double));
    if (scaleIndex != OP_NONE) {
        const double* cs = o->scale[scaleIndex];
        int idx = 0;
        for (int k = 0; k < o->P; k++) {
            double f = exp(cs[k]);
            for (int i = 0; i < o->S; i++) out[idx++] *= f;
        }
    }
    return ORC_SUCCESS;
}

/* C_ijk = V_ik * Vinv_kj, stored [i][j][k] (EigenDecompositionCube.hpp:69-100). */
int orcSetEigenDecomposition(int instance, int e, const double* V, const double* Vi, const double* lam) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (e < 0 || e >= o->eigenCount) return ORC_ERROR_OUT_OF_RANGE;
    int S = o->S, l = 0;
    int trans = (o->flags & F_INVEVEC_TRANSPOSED) != 0;
    for (int i = 0; i < S; i++) {
        o->evals[e][i] = lam[i];
        for (int j = 0; j < S; j++)
            for (int k = 0; k < S; k++)
                o->cube[e][l++] = V[i * S + k] * (trans ? Vi[k + j * S] : Vi[k * S + j]);
    }
    return ORC_SUCCESS;
}

static int set_vec(Orc* o, double** slots, int idx, int n, const double* in) {
    if (idx < 0 || idx >= o->eigenCount) return ORC_ERROR_OUT_OF_RANGE;
    if (!slots[idx]) slots[idx] = (double*)xcalloc(n, sizeof(double));
    memcpy(slots[idx], in, n * sizeof(double));
    return ORC_SUCCESS;
}

int orcSetStateFrequencies(int instance, int idx, const double* in) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    return set_vec(o, o->freqs, idx, o->S, in);
}
int orcSetCategoryWeights(int instance, int idx, const double* in) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    return set_vec(o, o->catWeights, idx, o->C, in);
}
/* Default rates live in slot 0 (BeagleCPUImpl.hpp:769-781). */
int orcSetCategoryRates(int instance, const double* in) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (!o->catRates[0]) o->catRates[0] = (double*)xcalloc(o->C, sizeof(double));
    memcpy(o->catRates[0], in, o->C * sizeof(double));
    return ORC_SUCCESS;
}
int orcSetCategoryRatesWithIndex(int instance, int idx, const double* in) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    return set_vec(o, o->catRates, idx, o->C, in);
}
int orcSetPatternWeights(int instance, const double* in) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    memcpy(o->patternWeights, in, o->P * sizeof(double));
    return ORC_SUCCESS;
}

/* Partition boundaries (BeagleCPUImpl.hpp:803-920).  When the partition
 * indices are not already 0, 1, 2, ... in runs, the patterns are reordered
 * so that each partition is contiguous, keeping their order within a
 * partition (BeagleCPUImpl.hpp:4140-4219): pattern i moves to
 * start[part[i]] + (its rank in its partition); the tips' states or
 * partials and the pattern weights are permuted, other buffers are not, and
 * site log-likelihoods are returned in the original order
 * (BeagleCPUImpl.hpp:1009-1021).  A second reorder is refused. */
int orcSetPatternPartitions(int instance, int partitionCount, const int* parts) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (partitionCount < 1) return ORC_ERROR_OUT_OF_RANGE;
    for (int k = 0; k < o->P; k++)
        if (parts[k] < 0 || parts[k] >= partitionCount) return ORC_ERROR_OUT_OF_RANGE;
    int reorder = 0, run = 0;
    for (int k = 0; k < o->P; k++) {
        if (k > 0 && parts[k] != parts[k - 1]) run++;
        if (parts[k] != run) { reorder = 1; break; }
    }
    free(o->partStart);
    o->partitionCount = partitionCount;
    o->partStart = (int*)xcalloc(partitionCount + 1, sizeof(int));
    if (!reorder) {
        for (int k = 1; k < o->P; k++)
            if (parts[k] != parts[k - 1]) o->partStart[parts[k]] = k;
        for (int q = parts[o->P - 1] + 1; q <= partitionCount; q++) o->partStart[q] = o->P;
        return ORC_SUCCESS;
    }
    if (o->newOrder) return ORC_ERROR_NO_IMPLEMENTATION;
    int* size = (int*)xcalloc(partitionCount, sizeof(int));
    o->newOrder = (int*)xcalloc(o->P, sizeof(int));
    for (int k = 0; k < o->P; k++) o->newOrder[k] = size[parts[k]]++;
    for (int q = 0; q < partitionCount; q++) o->partStart[q + 1] = o->partStart[q] + size[q];
    for (int k = 0; k < o->P; k++) o->newOrder[k] += o->partStart[parts[k]];
    free(size);
    double* w = (double*)xcalloc(o->P, sizeof(double));
    for (int k = 0; k < o->P; k++) w[o->newOrder[k]] = o->patternWeights[k];
    memcpy(o->patternWeights, w, o->P * sizeof(double));
    free(w);
    int S = o->S;
    for (int t = 0; t < o->tipCount; t++) {
        if (o->tipStates[t]) {
            int* st = (int*)xcalloc(o->P, sizeof(int));
            for (int k = 0; k < o->P; k++) st[o->newOrder[k]] = o->tipStates[t][k];
            memcpy(o->tipStates[t], st, o->P * sizeof(int));
            free(st);
        } else if (o->partials[t]) {
            size_t n = (size_t)o->C * o->P * S;
            double* x = (double*)xcalloc(n, sizeof(double));
            for (int l = 0; l < o->C; l++)
                for (int k = 0; k < o->P; k++)
                    for (int j = 0; j < S; j++)
                        x[((size_t)l * o->P + o->newOrder[k]) * S + j] =
                            o->partials[t][((size_t)l * o->P + k) * S + j];
            memcpy(o->partials[t], x, n * sizeof(double));
            free(x);
        }
    }
    return ORC_SUCCESS;
}

/* P, dP/dt and d2P/dt2 from the eigensystem, per category rate r_l:
 *   P_ij   = max(0, sum_k C_ijk exp(lambda_k t r_l)),  pad column 1
 *   dP_ij  = sum_k C_ijk (lambda_k r_l) exp(.),        pad column 0
 *   d2P_ij = sum_k C_ijk (lambda_k r_l)^2 exp(.),      pad column 0
 * (EigenDecompositionCube.hpp:105-322). */
static void update_matrices(Orc* o, int e, const double* rates, int pIdx, int d1Idx, int d2Idx,
                            double t) {
    int S = o->S;
    double* ex = (double*)xcalloc(S, sizeof(double));
    double* e1 = (double*)xcalloc(S, sizeof(double));
    double* e2 = (double*)xcalloc(S, sizeof(double));
    for (int l = 0; l < o->C; l++) {
        for (int k = 0; k < S; k++) {
            double sl = o->evals[e][k] * rates[l];
            if (d1Idx < 0) ex[k] = exp(o->evals[e][k] * (t * rates[l]));
            else ex[k] = exp(sl * t);
            e1[k] = sl * ex[k];
            e2[k] = sl * e1[k];
        }
        const double* cm = o->cube[e];
        double* Pm = o->matrices[pIdx] + (size_t)l * S * (S + 1);
        double* D1 = d1Idx >= 0 ? o->matrices[d1Idx] + (size_t)l * S * (S + 1) : NULL;
        double* D2 = d2Idx >= 0 ? o->matrices[d2Idx] + (size_t)l * S * (S + 1) : NULL;
        int m = 0;
        for (int i = 0; i < S; i++) {
            for (int j = 0; j < S; j++) {
                double s0 = 0, s1 = 0, s2 = 0;
                for (int k = 0; k < S; k++, m++) {
                    s0 += cm[m] * ex[k];
                    s1 += cm[m] * e1[k];
                    s2 += cm[m] * e2[k];
                }
                Pm[i * (S + 1) + j] = s0 > 0 ? s0 : 0;
                if (D1) D1[i * (S + 1) + j] = s1;
                if (D2) D2[i * (S + 1) + j] = s2;
            }
            Pm[i * (S + 1) + S] = 1.0;
            if (D1) D1[i * (S + 1) + S] = 0.0;
            if (D2) D2[i * (S + 1) + S] = 0.0;
        }
    }
    free(ex); free(e1); free(e2);
}

/* BeagleCPUImpl.hpp:1225-1238 (category rates from slot 0). */
int orcUpdateTransitionMatrices(int instance, int eigenIndex, const int* probIdx, const int* d1Idx,
                                const int* d2Idx, const double* edgeLengths, int count) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (eigenIndex < 0 || eigenIndex >= o->eigenCount || !o->catRates[0]) return ORC_ERROR_OUT_OF_RANGE;
    for (int u = 0; u < count; u++) {
        int d1 = d1Idx ? d1Idx[u] : -1;
        int d2 = (d1Idx && d2Idx) ? d2Idx[u] : -1;
        if (probIdx[u] < 0 || probIdx[u] >= o->matrixCount) return ORC_ERROR_OUT_OF_RANGE;
        update_matrices(o, eigenIndex, o->catRates[0], probIdx[u], d1, d2, edgeLengths[u]);
    }
    return ORC_SUCCESS;
}

/* BeagleCPUImpl.hpp:1256-1290 */
int orcUpdateTransitionMatricesWithMultipleModels(int instance, const int* eigenIdx, const int* rateIdx,
                                                  const int* probIdx, const int* d1Idx, const int* d2Idx,
                                                  const double* edgeLengths, int count) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    for (int u = 0; u < count; u++) {
        int d1 = d1Idx ? d1Idx[u] : -1;
        int d2 = (d1Idx && d2Idx) ? d2Idx[u] : -1;
        update_matrices(o, eigenIdx[u], o->catRates[rateIdx[u]], probIdx[u], d1, d2, edgeLengths[u]);
    }
    return ORC_SUCCESS;
}

/* Category l uses eigen-system eigenIdx[l] with rate 1
 * (EigenDecompositionCube.hpp:326-470, via BeagleCPUImpl.hpp:1242-1253):
 * the per-category form of update_matrices. */
int orcUpdateTransitionMatricesWithModelCategories(int instance, const int* eigenIdx,
                                                   const int* probIdx, const int* d1Idx,
                                                   const int* d2Idx, const double* edgeLengths,
                                                   int count) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    int S = o->S;
    for (int l = 0; l < o->C; l++)
        if (eigenIdx[l] < 0 || eigenIdx[l] >= o->eigenCount) return ORC_ERROR_OUT_OF_RANGE;
    double* ex = (double*)xcalloc(S, sizeof(double));
    double* e1 = (double*)xcalloc(S, sizeof(double));
    double* e2 = (double*)xcalloc(S, sizeof(double));
    for (int u = 0; u < count; u++) {
        int d1 = d1Idx ? d1Idx[u] : -1;
        int d2 = (d1Idx && d2Idx) ? d2Idx[u] : -1;
        for (int l = 0; l < o->C; l++) {
            int e = eigenIdx[l];
            for (int k = 0; k < S; k++) {
                double lam = o->evals[e][k];
                ex[k] = exp(lam * edgeLengths[u]);
                e1[k] = lam * ex[k];
                e2[k] = lam * e1[k];
            }
            const double* cm = o->cube[e];
            double* Pm = o->matrices[probIdx[u]] + (size_t)l * S * (S + 1);
            double* D1 = d1 >= 0 ? o->matrices[d1] + (size_t)l * S * (S + 1) : NULL;
            double* D2 = d2 >= 0 ? o->matrices[d2] + (size_t)l * S * (S + 1) : NULL;
            int m = 0;
            for (int i = 0; i < S; i++) {
                for (int j = 0; j < S; j++) {
                    double s0 = 0, s1 = 0, s2 = 0;
                    for (int k = 0; k < S; k++, m++) {
                        s0 += cm[m] * ex[k];
                        s1 += cm[m] * e1[k];
                        s2 += cm[m] * e2[k];
                    }
                    Pm[i * (S + 1) + j] = s0 > 0 ? s0 : 0;
                    if (D1) D1[i * (S + 1) + j] = s1;
                    if (D2) D2[i * (S + 1) + j] = s2;
                }
                Pm[i * (S + 1) + S] = 1.0;
                if (D1) D1[i * (S + 1) + S] = 0.0;
                if (D2) D2[i * (S + 1) + S] = 0.0;
            }
        }
    }
    free(ex); free(e1); free(e2);
    return ORC_SUCCESS;
}

/* result[l] = first[l] * second[l] (matrix product per category), pad
 * column 1; in-place results are refused (BeagleCPUImpl.hpp:1096-1162). */
int orcConvolveTransitionMatrices(int instance, const int* first, const int* second,
                                  const int* result, int count) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    int S = o->S, W = S + 1;
    for (int u = 0; u < count; u++) {
        if (first[u] == result[u] || second[u] == result[u]) return ORC_ERROR_OUT_OF_RANGE;
        for (int l = 0; l < o->C; l++) {
            const double* A = o->matrices[first[u]] + (size_t)l * S * W;
            const double* B = o->matrices[second[u]] + (size_t)l * S * W;
            double* Cm = o->matrices[result[u]] + (size_t)l * S * W;
            for (int i = 0; i < S; i++) {
                for (int j = 0; j < S; j++) {
                    double sum = 0.0;
                    for (int k = 0; k < S; k++) sum += A[i * W + k] * B[k * W + j];
                    Cm[i * W + j] = sum;
                }
                Cm[i * W + S] = 1.0;
            }
        }
    }
    return ORC_SUCCESS;
}

/* Client matrices [C][S][S] plus the given pad value (BeagleCPUImpl.hpp:1035-1091). */
int orcSetTransitionMatrix(int instance, int m, const double* in, double padValue) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (m < 0 || m >= o->matrixCount) return ORC_ERROR_OUT_OF_RANGE;
    int S = o->S;
    for (int l = 0; l < o->C; l++)
        for (int i = 0; i < S; i++) {
            double* row = o->matrices[m] + ((size_t)l * S + i) * (S + 1);
            memcpy(row, in + ((size_t)l * S + i) * S, S * sizeof(double));
            row[S] = padValue;
        }
    return ORC_SUCCESS;
}
int orcSetDifferentialMatrix(int instance, int m, const double* in) {
    return orcSetTransitionMatrix(instance, m, in, 0.0);
}
int orcSetTransitionMatrices(int instance, const int* idx, const double* in, const double* pads, int count) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    for (int k = 0; k < count; k++) {
        int rc = orcSetTransitionMatrix(instance, idx[k], in + (size_t)k * o->S * o->S * o->C, pads[k]);
        if (rc) return rc;
    }
    return ORC_SUCCESS;
}
int orcGetTransitionMatrix(int instance, int m, double* out) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (m < 0 || m >= o->matrixCount) return ORC_ERROR_OUT_OF_RANGE;
    int S = o->S;
    for (int l = 0; l < o->C; l++)
        for (int i = 0; i < S; i++)
            memcpy(out + ((size_t)l * S + i) * S, o->matrices[m] + ((size_t)l * S + i) * (S + 1),
                   S * sizeof(double));
    return ORC_SUCCESS;
}
/* Per category transpose of the S x S block; pad column kept
 * (BeagleCPUImpl.hpp:1175-1222). */
int orcTransposeTransitionMatrices(int instance, const int* in, const int* outIdx, int count) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    int S = o->S;
    double* tmp = (double*)xcalloc((size_t)S * (S + 1), sizeof(double));
    for (int u = 0; u < count; u++) {
        for (int l = 0; l < o->C; l++) {
            const double* a = o->matrices[in[u]] + (size_t)l * S * (S + 1);
            double* b = o->matrices[outIdx[u]] + (size_t)l * S * (S + 1);
            for (int i = 0; i < S; i++) {
                for (int j = 0; j < S; j++) tmp[j * (S + 1) + i] = a[i * (S + 1) + j];
                tmp[i * (S + 1) + S] = a[i * (S + 1) + S];
            }
            memcpy(b, tmp, (size_t)S * (S + 1) * sizeof(double));
        }
    }
    free(tmp);
    return ORC_SUCCESS;
}

/* Per pattern: max over all categories and states; 0 -> 1; multiply by
 * 1/max; store log(max) or max; cumulative += log(max)
 * (BeagleCPUImpl.hpp:4008-4056, BeagleCPU4StateImpl.hpp:665-715). */
static void rescale(Orc* o, double* d, double* sf, double* cum, int k0, int k1) {
    int S = o->S, P = o->P;
    for (int k = k0; k < k1; k++) {
        double mx = 0;
        for (int l = 0; l < o->C; l++)
            for (int i = 0; i < S; i++) {
                double v = d[((size_t)l * P + k) * S + i];
                if (v > mx) mx = v;
            }
        if (mx == 0) mx = 1.0;
        double inv = 1.0 / mx;
        for (int l = 0; l < o->C; l++)
            for (int i = 0; i < S; i++) d[((size_t)l * P + k) * S + i] *= inv;
        if (o->flags & F_SCALERS_LOG) {
            double lm = log(mx);
            sf[k] = lm;
            if (cum) cum[k] += lm;
        } else {
            sf[k] = mx;
            if (cum) cum[k] += log(mx);
        }
    }
}

/* One child's contribution to a parent entry: sum_j M[i][j] x[j] for a
 * partials child, M[i][state] for a state child (state S hits the pad). */
static inline double child_term(const Orc* o, const double* M, const double* x, const int* st,
                                int l, int k, int i) {
    int S = o->S;
    const double* row = M + ((size_t)l * S + i) * (S + 1);
    if (st) return row[st[k]];
    const double* xv = x + ((size_t)l * o->P + k) * S;
    double s = 0;
    for (int j = 0; j < S; j++) s += row[j] * xv[j];
    return s;
}

static int accumulate(Orc* o, const int* idx, int count, int cumIdx, int sign, int k0, int k1);

/* Post-order operations (BeagleCPUImpl.hpp:1497-1686; kernels
 * BeagleCPU4StateImpl.hpp:186-377, 519-560 and BeagleCPUImpl.hpp:4226-4524):
 *   dest[c,p,i] = term(child1) * term(child2)  [/ readScale[p] if fixed]
 * writeScale >= 0 -> rescale after computing; else readScale >= 0 ->
 * divide by that (raw) factor; SCALING_ALWAYS rescales every op into
 * buffer dest-tipCount and accumulates the children's buffers into it. */
static int up_partials(Orc* o, const int* ops, int count, int cumIdx, int stride) {
    int S = o->S;
    for (int op = 0; op < count; op++) {
        const int* q = ops + (size_t)op * stride;
        int dest = q[0], wS = q[1], rS = q[2], c1 = q[3], m1 = q[4], c2 = q[5], m2 = q[6];
        int k0 = 0, k1 = o->P;
        int cum = cumIdx;
        if (stride == 9) {
            int part = q[7];
            if (part < 0 || part >= o->partitionCount) return ORC_ERROR_OUT_OF_RANGE;
            k0 = o->partStart[part]; k1 = o->partStart[part + 1];
            cum = q[8];
        }
        if (dest < o->tipCount || dest >= o->bufferCount || c1 < 0 || c1 >= o->bufferCount ||
            c2 < 0 || c2 >= o->bufferCount || m1 < 0 || m1 >= o->matrixCount || m2 < 0 ||
            m2 >= o->matrixCount)
            return ORC_ERROR_OUT_OF_RANGE;
        const int* s1 = o->tipStates[c1];
        const int* s2 = o->tipStates[c2];
        const double* x1 = o->partials[c1];
        const double* x2 = o->partials[c2];
        if ((!s1 && !x1) || (!s2 && !x2)) return ORC_ERROR_OUT_OF_RANGE;
        /* scale indices in range (the reference does not check them) */
        if (wS >= o->scaleCount || rS >= o->scaleCount || cum >= o->scaleCount ||
            ((o->flags & F_SCALING_ALWAYS) && dest - o->tipCount >= o->scaleCount))
            return ORC_ERROR_OUT_OF_RANGE;
        double* d = o->partials[dest];
        int mode = OP_NONE;
        double* sf = NULL;
        if (o->flags & F_SCALING_ALWAYS) {
            mode = 1;
            sf = o->scale[dest - o->tipCount];
        } else if (wS >= 0) {
            mode = 1;
            sf = o->scale[wS];
        } else if (rS >= 0) {
            mode = 0;
            sf = o->scale[rS];
        }
        for (int l = 0; l < o->C; l++)
            for (int k = k0; k < k1; k++)
                for (int i = 0; i < S; i++) {
                    double v = child_term(o, o->matrices[m1], x1, s1, l, k, i) *
                               child_term(o, o->matrices[m2], x2, s2, l, k, i);
                    if (mode == 0) v = v / sf[k];
                    d[((size_t)l * o->P + k) * S + i] = v;
                }
        if (mode == 1) rescale(o, d, sf, cum >= 0 ? o->scale[cum] : NULL, k0, k1);
        if (o->flags & F_SCALING_ALWAYS) {
            int idx[2], n = 0;
            if (c1 - o->tipCount >= 0) idx[n++] = c1 - o->tipCount;
            if (c2 - o->tipCount >= 0) idx[n++] = c2 - o->tipCount;
            accumulate(o, idx, n, dest - o->tipCount, +1, k0, k1);
        }
    }
    return ORC_SUCCESS;
}

int orcUpdatePartials(int instance, const int* ops, int count, int cumIdx) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    return up_partials(o, ops, count, cumIdx, 7);
}
int orcUpdatePartialsByPartition(int instance, const int* ops, int count) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    return up_partials(o, ops, count, OP_NONE, 9);
}

/* Pre-order operations [destPre, wScale, rScale, parentPre, P_self,
 * sibling, P_sib] (BeagleCPUImpl.hpp:1689-1862; kernels
 * BeagleCPU4StateImpl.hpp:380-461, BeagleCPUImpl.hpp:4530-4662):
 *   pre_self[c,p,j] = sum_i P_self[c][i][j] pre_parent[c,p,i] term(sibling)_i
 * Rescaling as in post-order but never into a cumulative buffer; no
 * read-scale path (the reference leaves it unimplemented). */
static int up_pre_partials(Orc* o, const int* ops, int count, int stride) {
    int S = o->S;
    double* tmp = (double*)xcalloc(S, sizeof(double));
    for (int op = 0; op < count; op++) {
        const int* q = ops + (size_t)op * stride;
        int dest = q[0], wS = q[1], par = q[3], mSelf = q[4], sib = q[5], mSib = q[6];
        int k0 = 0, k1 = o->P;
        if (stride == 9) {
            int part = q[7];
            if (part < 0 || part >= o->partitionCount) { free(tmp); return ORC_ERROR_OUT_OF_RANGE; }
            k0 = o->partStart[part]; k1 = o->partStart[part + 1];
        }
        if (dest < 0 || dest >= o->bufferCount || par < 0 || par >= o->bufferCount || sib < 0 ||
            sib >= o->bufferCount || mSelf < 0 || mSelf >= o->matrixCount || mSib < 0 ||
            mSib >= o->matrixCount || !o->partials[par]) {
            free(tmp);
            return ORC_ERROR_OUT_OF_RANGE;
        }
        size_t n = (size_t)o->C * o->P * S;
        if (!o->partials[dest]) o->partials[dest] = (double*)xcalloc(n, sizeof(double));
        const int* ss = o->tipStates[sib];
        const double* xs = o->partials[sib];
        if (!ss && !xs) { free(tmp); return ORC_ERROR_OUT_OF_RANGE; }
        const double* pp = o->partials[par];
        double* d = o->partials[dest];
        double* sf = NULL;
        int mode = OP_NONE;
        if (o->flags & F_SCALING_ALWAYS) {
            mode = 1;
            sf = o->scale[dest - o->tipCount];
        } else if (wS >= 0) {
            if (wS >= o->scaleCount) { free(tmp); return ORC_ERROR_OUT_OF_RANGE; }
            mode = 1;
            sf = o->scale[wS];
        }
        if ((o->flags & F_SCALING_ALWAYS) && dest - o->tipCount >= o->scaleCount) {
            free(tmp);
            return ORC_ERROR_OUT_OF_RANGE;
        }
        for (int l = 0; l < o->C; l++) {
            const double* Ms = o->matrices[mSelf] + (size_t)l * S * (S + 1);
            for (int k = k0; k < k1; k++) {
                const double* pv = pp + ((size_t)l * o->P + k) * S;
                for (int i = 0; i < S; i++)
                    tmp[i] = pv[i] * child_term(o, o->matrices[mSib], xs, ss, l, k, i);
                for (int j = 0; j < S; j++) {
                    double s = 0;
                    for (int i = 0; i < S; i++) s += Ms[i * (S + 1) + j] * tmp[i];
                    d[((size_t)l * o->P + k) * S + j] = s;
                }
            }
        }
        if (mode == 1) rescale(o, d, sf, NULL, k0, k1);
    }
    free(tmp);
    return ORC_SUCCESS;
}

int orcUpdatePrePartials(int instance, const int* ops, int count, int cumIdx) {
    (void)cumIdx;
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    return up_pre_partials(o, ops, count, 7);
}
int orcUpdatePrePartialsByPartition(int instance, const int* ops, int count) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    return up_pre_partials(o, ops, count, 9);
}

int orcWaitForPartials(int instance, const int* d, int n) {
    (void)d; (void)n;
    return get(instance) ? ORC_SUCCESS : ORC_ERROR_UNINITIALIZED_INSTANCE;
}

/* cumulative[p] += (LOG ? s : log s), or -= for removal
 * (BeagleCPUImpl.hpp:2763-2860). */
static int accumulate(Orc* o, const int* idx, int count, int cumIdx, int sign, int k0, int k1) {
    if (cumIdx < 0 || cumIdx >= o->scaleCount) return ORC_ERROR_OUT_OF_RANGE;
    for (int i = 0; i < count; i++)
        if (idx[i] < 0 || idx[i] >= o->scaleCount) return ORC_ERROR_OUT_OF_RANGE;
    double* cum = o->scale[cumIdx];
    for (int i = 0; i < count; i++) {
        if (idx[i] < 0 || idx[i] >= o->scaleCount) return ORC_ERROR_OUT_OF_RANGE;
        const double* s = o->scale[idx[i]];
        for (int k = k0; k < k1; k++) {
            double v = (o->flags & F_SCALERS_LOG) ? s[k] : log(s[k]);
            cum[k] += sign * v;
        }
    }
    return ORC_SUCCESS;
}

int orcAccumulateScaleFactors(int instance, const int* idx, int count, int cumIdx) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    return accumulate(o, idx, count, cumIdx, +1, 0, o->P);
}
int orcRemoveScaleFactors(int instance, const int* idx, int count, int cumIdx) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    return accumulate(o, idx, count, cumIdx, -1, 0, o->P);
}
int orcAccumulateScaleFactorsByPartition(int instance, const int* idx, int count, int cumIdx, int part) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (part < 0 || part >= o->partitionCount) return ORC_ERROR_OUT_OF_RANGE;
    return accumulate(o, idx, count, cumIdx, +1, o->partStart[part], o->partStart[part + 1]);
}
int orcRemoveScaleFactorsByPartition(int instance, const int* idx, int count, int cumIdx, int part) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (part < 0 || part >= o->partitionCount) return ORC_ERROR_OUT_OF_RANGE;
    return accumulate(o, idx, count, cumIdx, -1, o->partStart[part], o->partStart[part + 1]);
}
int orcResetScaleFactors(int instance, int cumIdx) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (cumIdx < 0 || cumIdx >= o->scaleCount) return ORC_ERROR_OUT_OF_RANGE;
    memset(o->scale[cumIdx], 0, o->P * sizeof(double));
    return ORC_SUCCESS;
}
int orcResetScaleFactorsByPartition(int instance, int cumIdx, int part) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (cumIdx < 0 || cumIdx >= o->scaleCount || part < 0 || part >= o->partitionCount)
        return ORC_ERROR_OUT_OF_RANGE;
    int k0 = o->partStart[part], k1 = o->partStart[part + 1];
    memset(o->scale[cumIdx] + k0, 0, (size_t)(k1 - k0) * sizeof(double));
    return ORC_SUCCESS;
}
int orcCopyScaleFactors(int instance, int dst, int src) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (dst < 0 || dst >= o->scaleCount || src < 0 || src >= o->scaleCount) return ORC_ERROR_OUT_OF_RANGE;
    memcpy(o->scale[dst], o->scale[src], o->P * sizeof(double));
    return ORC_SUCCESS;
}
/* The oracle returns the buffer (the reference CPU path is a no-op here,
 * BeagleCPUImpl.hpp:2911-2915, so tests compare it with the GPU only). */
int orcGetScaleFactors(int instance, int src, double* out) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (src < 0 || src >= o->scaleCount) return ORC_ERROR_OUT_OF_RANGE;
    memcpy(out, o->scale[src], o->P * sizeof(double));
    return ORC_SUCCESS;
}

/* site[p] = log(sum_i pi_i sum_c w_c root[c,p,i]) + cumulative[p];
 * logL = sum_p weight_p site[p], sequential in double
 * (BeagleCPUImpl.hpp:2294-2346, 2630-2691; BeagleCPU4StateImpl.hpp:563-605). */
static double root_site(const Orc* o, const double* root, const double* w, const double* f, int k) {
    int S = o->S;
    double sum = 0;
    for (int i = 0; i < S; i++) {
        double t = 0;
        for (int l = 0; l < o->C; l++) t += root[((size_t)l * o->P + k) * S + i] * w[l];
        sum += f[i] * t;
    }
    return log(sum);
}

int orcCalculateRootLogLikelihoods(int instance, const int* bufIdx, const int* wIdx, const int* fIdx,
                                   const int* cumIdx, int count, double* out) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (count != 1) return ORC_ERROR_NO_IMPLEMENTATION;
    int b = bufIdx[0];
    if (b < 0 || b >= o->bufferCount || !o->partials[b]) return ORC_ERROR_OUT_OF_RANGE;
    if (wIdx[0] < 0 || wIdx[0] >= o->eigenCount || fIdx[0] < 0 || fIdx[0] >= o->eigenCount)
        return ORC_ERROR_OUT_OF_RANGE;
    int cs = (o->flags & F_SCALING_ALWAYS) ? b - o->tipCount : cumIdx[0];
    const double* w = o->catWeights[wIdx[0]];
    const double* f = o->freqs[fIdx[0]];
    for (int k = 0; k < o->P; k++) {
        o->siteLogL[k] = root_site(o, o->partials[b], w, f, k);
        if (cs >= 0) o->siteLogL[k] += o->scale[cs][k];
    }
    double s = 0;
    for (int k = 0; k < o->P; k++) s += o->siteLogL[k] * o->patternWeights[k];
    *out = s;
    return s != s ? ORC_ERROR_FLOATING_POINT : ORC_SUCCESS;
}

/* Per-partition root sums (BeagleCPUImpl.hpp:2349-2388 and the
 * integrate-by-partition loops it calls). */
int orcCalculateRootLogLikelihoodsByPartition(int instance, const int* bufIdx, const int* wIdx,
                                              const int* fIdx, const int* cumIdx, const int* partIdx,
                                              int partitionCount, int count, double* outByPart,
                                              double* outSum) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (count != 1) return ORC_ERROR_NO_IMPLEMENTATION;
    double total = 0;
    for (int q = 0; q < partitionCount; q++) {
        int part = partIdx[q];
        if (part < 0 || part >= o->partitionCount) return ORC_ERROR_OUT_OF_RANGE;
        const double* root = o->partials[bufIdx[q]];
        if (!root) return ORC_ERROR_OUT_OF_RANGE;
        const double* w = o->catWeights[wIdx[q]];
        const double* f = o->freqs[fIdx[q]];
        double s = 0;
        for (int k = o->partStart[part]; k < o->partStart[part + 1]; k++) {
            o->siteLogL[k] = root_site(o, root, w, f, k);
            if (cumIdx[q] >= 0) o->siteLogL[k] += o->scale[cumIdx[q]][k];
            s += o->siteLogL[k] * o->patternWeights[k];
        }
        outByPart[q] = s;
        total += s;
    }
    *outSum = total;
    return total != total ? ORC_ERROR_FLOATING_POINT : ORC_SUCCESS;
}

int orcGetSiteLogLikelihoods(int instance, double* out) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (o->newOrder)
        for (int k = 0; k < o->P; k++) out[k] = o->siteLogL[o->newOrder[k]];
    else
        memcpy(out, o->siteLogL, o->P * sizeof(double));
    return ORC_SUCCESS;
}
int orcGetLogLikelihood(int instance, double* out) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    double s = 0;
    for (int k = 0; k < o->P; k++) s += o->siteLogL[k] * o->patternWeights[k];
    *out = s;
    return s != s ? ORC_ERROR_FLOATING_POINT : ORC_SUCCESS;
}

/* Edge derivatives (beagle.cpp:1794-1824 passes no rates/scales;
 * BeagleCPUImpl.hpp:1940-2083; 4-state BeagleCPU4StateImpl.hpp:1003-1083;
 * generic BeagleCPUImpl.hpp:2086-2119, 2240-2285):
 *   num_p = sum_c w_c sum_i pre[c,p,i] sum_j D[c][i][j] post[c,p,j]
 *   den_p = sum_c w_c sum_i pre[c,p,i] post[c,p,i]
 * A state child uses column `state` of D for num (pad 0 when missing) and
 * pre[state % S] for den (the reference's own indexing for a missing
 * state).  ratio = num/den, outputs per site, sum_p w ratio and
 * sum_p w ratio^2. */
int orcCalculateEdgeDerivatives(int instance, const int* postIdx, const int* preIdx, const int* dIdx,
                                const int* wIdx, int count, double* outD, double* outSum,
                                double* outSumSq) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    int S = o->S, P = o->P;
    const double* w = o->catWeights[wIdx[0]];
    for (int n = 0; n < count; n++) {
        const double* pre = o->partials[preIdx[n]];
        const int* st = o->tipStates[postIdx[n]];
        const double* post = o->partials[postIdx[n]];
        if (!pre || (!st && !post) || dIdx[n] < 0 || dIdx[n] >= o->matrixCount)
            return ORC_ERROR_OUT_OF_RANGE;
        const double* D = o->matrices[dIdx[n]];
        memset(o->numTmp, 0, P * sizeof(double));
        memset(o->denTmp, 0, P * sizeof(double));
        for (int l = 0; l < o->C; l++) {
            const double* Dl = D + (size_t)l * S * (S + 1);
            for (int k = 0; k < P; k++) {
                const double* pv = pre + ((size_t)l * P + k) * S;
                double num = 0, den = 0;
                if (st) {
                    int s = st[k];
                    for (int i = 0; i < S; i++) num += Dl[i * (S + 1) + s] * pv[i];
                    den = pv[s % S];
                } else {
                    const double* qv = post + ((size_t)l * P + k) * S;
                    for (int i = 0; i < S; i++) {
                        double t = 0;
                        for (int j = 0; j < S; j++) t += Dl[i * (S + 1) + j] * qv[j];
                        num += t * pv[i];
                        den += qv[i] * pv[i];
                    }
                }
                o->numTmp[k] += w[l] * num;
                o->denTmp[k] += w[l] * den;
            }
        }
        double s1 = 0, s2 = 0;
        for (int k = 0; k < P; k++) {
            double r = o->numTmp[k] / o->denTmp[k];
            if (outD) outD[(size_t)n * P + k] = r;
            s1 += r * o->patternWeights[k];
            s2 += r * r * o->patternWeights[k];
        }
        if (outSum) outSum[n] = s1;
        if (outSumSq) outSumSq[n] = s2;
    }
    return ORC_SUCCESS;
}

/* Cross products added into out[k*S+j] (BeagleCPUImpl.hpp:1865-1937,
 * 2122-2237; BeagleCPU4StateImpl.hpp:1088-1254):
 *   sum_n sum_p (w_p / den_p) sum_c w_c r_c t_n pre[c,p,k] post[c,p,j]
 * with den_p = sum_c w_c pre.post; a missing-state tip contributes the
 * full outer product with post = 1 and den = sum_c w_c sum_k pre. */
int orcCalculateCrossProductDerivative(int instance, const int* postIdx, const int* preIdx,
                                       const int* rIdx, const int* wIdx, const double* t, int count,
                                       double* out, double* outSq) {
    (void)outSq;
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    int S = o->S, P = o->P;
    const double* r = o->catRates[rIdx[0]];
    const double* w = o->catWeights[wIdx[0]];
    double* tmp = (double*)xcalloc((size_t)S * S, sizeof(double));
    for (int n = 0; n < count; n++) {
        const double* pre = o->partials[preIdx[n]];
        const int* st = o->tipStates[postIdx[n]];
        const double* post = o->partials[postIdx[n]];
        if (!pre || (!st && !post)) { free(tmp); return ORC_ERROR_OUT_OF_RANGE; }
        for (int k = 0; k < P; k++) {
            memset(tmp, 0, (size_t)S * S * sizeof(double));
            double den = 0;
            int s = st ? st[k] : -1;
            for (int l = 0; l < o->C; l++) {
                double ws = w[l] * (r[l] * t[n]);
                const double* pv = pre + ((size_t)l * P + k) * S;
                if (st && s < S) {
                    den += pv[s] * w[l];
                    for (int a = 0; a < S; a++) tmp[a * S + s] += pv[a] * ws;
                } else if (st) {
                    double d = 0;
                    for (int a = 0; a < S; a++) d += pv[a];
                    den += d * w[l];
                    for (int a = 0; a < S; a++)
                        for (int b = 0; b < S; b++) tmp[a * S + b] += pv[a] * ws;
                } else {
                    const double* qv = post + ((size_t)l * P + k) * S;
                    double d = 0;
                    for (int a = 0; a < S; a++) d += pv[a] * qv[a];
                    den += d * w[l];
                    for (int a = 0; a < S; a++)
                        for (int b = 0; b < S; b++) tmp[a * S + b] += pv[a] * qv[b] * ws;
                }
            }
            double pw = o->patternWeights[k] / den;
            for (int a = 0; a < S * S; a++) out[a] += tmp[a] * pw;
        }
    }
    free(tmp);
    return ORC_SUCCESS;
}

/* Site values and pattern-weighted sums [logL, d1, d2] of the edge
 * parent - child over patterns [k0, k1) (the per-pattern body of
 * BeagleCPUImpl.hpp:3209-3322 and :3452-3590); Ms = {P, D1 or NULL, D2 or
 * NULL}, cum a cumulative scale buffer or NULL, alw the two node scale
 * buffers under SCALING_ALWAYS. */
static void edge_range(Orc* o, const double* par, const double* ch, const int* st,
                       const double* const Ms[3], const double* w, const double* f,
                       const double* cum, const double* const alw[2], int k0, int k1,
                       double sums[3]) {
    const int S = o->S, P = o->P;
    double sL = 0, s1 = 0, s2 = 0;
    for (int k = k0; k < k1; k++) {
        double acc[3] = {0, 0, 0};
        for (int m = 0; m < 3; m++) {
            if (!Ms[m]) continue;
            for (int l = 0; l < o->C; l++) {
                const double* pv = par + ((size_t)l * P + k) * S;
                double s = 0;
                for (int i = 0; i < S; i++) {
                    double t = child_term(o, Ms[m], ch, st, l, k, i);
                    s += f[i] * pv[i] * t;
                }
                acc[m] += w[l] * s;
            }
        }
        o->siteLogL[k] = log(acc[0]) + (cum ? cum[k] : 0.0);
        if (alw[0]) o->siteLogL[k] += alw[0][k];
        if (alw[1]) o->siteLogL[k] += alw[1][k];
        sL += o->siteLogL[k] * o->patternWeights[k];
        if (Ms[1]) {
            double d1 = acc[1] / acc[0];
            o->siteD1[k] = d1;
            s1 += d1 * o->patternWeights[k];
            if (Ms[2]) {
                double d2 = acc[2] / acc[0] - d1 * d1;
                o->siteD2[k] = d2;
                s2 += d2 * o->patternWeights[k];
            }
        }
    }
    sums[0] = sL;
    sums[1] = s1;
    sums[2] = s2;
}

/* Edge log-likelihood between a parent (post-order above the edge) and a
 * child, with optional first and second derivatives along the edge
 * (BeagleCPUImpl.hpp:2918-3005, 3209-3322):
 *   L_p   = sum_c w_c sum_i pi_i parent[c,p,i] sum_j P[c][i][j] child[c,p,j]
 *   L'_p, L''_p with dP, d2P in place of P
 *   site = log L_p + cumulative;  d1 = L'/L;  d2 = L''/L - (L'/L)^2. */
int orcCalculateEdgeLogLikelihoods(int instance, const int* parIdx, const int* childIdx,
                                   const int* pIdx, const int* d1Idx, const int* d2Idx,
                                   const int* wIdx, const int* fIdx, const int* cumIdx, int count,
                                   double* outL, double* outD1, double* outD2) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (count != 1) return ORC_ERROR_NO_IMPLEMENTATION;
    int P = o->P;
    const double* par = o->partials[parIdx[0]];
    const int* st = o->tipStates[childIdx[0]];
    const double* ch = o->partials[childIdx[0]];
    if (!par || (!st && !ch)) return ORC_ERROR_OUT_OF_RANGE;
    const double* w = o->catWeights[wIdx[0]];
    const double* f = o->freqs[fIdx[0]];
    const double* Ms[3] = {o->matrices[pIdx[0]], d1Idx ? o->matrices[d1Idx[0]] : NULL,
                           (d1Idx && d2Idx) ? o->matrices[d2Idx[0]] : NULL};
    /* SCALING_ALWAYS: the cumulative factor is the sum of the parent's and
     * child's node scale buffers (BeagleCPUImpl.hpp:2934-2949). */
    int cs = (o->flags & F_SCALING_ALWAYS) ? -1 : cumIdx[0];
    const double* alw[2] = {NULL, NULL};
    if (o->flags & F_SCALING_ALWAYS) {
        if (parIdx[0] - o->tipCount >= 0) alw[0] = o->scale[parIdx[0] - o->tipCount];
        if (childIdx[0] - o->tipCount >= 0) alw[1] = o->scale[childIdx[0] - o->tipCount];
    }
    double sums[3];
    edge_range(o, par, ch, st, Ms, w, f, cs >= 0 ? o->scale[cs] : NULL, alw, 0, P, sums);
    *outL = sums[0];
    if (outD1 && Ms[1]) *outD1 = sums[1];
    if (outD2 && Ms[2]) *outD2 = sums[2];
    return sums[0] != sums[0] ? ORC_ERROR_FLOATING_POINT : ORC_SUCCESS;
}

/* Edge log-likelihoods per pattern partition
 * (BeagleCPUImpl.hpp:3008-3110, calcEdgeLogLikelihoodsSecondDerivByPartition
 * :3452-3590): for partition entry q the edge parentBufferIndices[q] -
 * childBufferIndices[q] over the patterns of partitionIndices[q], with its
 * own matrices, weights, frequencies and cumulative scale buffer; sums per
 * entry and over all entries.  As the reference: count must be 1, automatic
 * and always-on scaling are not implemented, and derivatives come both or
 * not at all (a first derivative alone is not implemented). */
int orcCalculateEdgeLogLikelihoodsByPartition(
    int instance, const int* parIdx, const int* childIdx, const int* pIdx, const int* d1Idx,
    const int* d2Idx, const int* wIdx, const int* fIdx, const int* cumIdx, const int* partIdx,
    int partitionCount, int count, double* outLByPart, double* outL, double* outD1ByPart,
    double* outD1, double* outD2ByPart, double* outD2) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    if (count != 1 || (o->flags & (F_SCALING_AUTO | F_SCALING_ALWAYS)))
        return ORC_ERROR_NO_IMPLEMENTATION;
    const int derivs = d1Idx != NULL;
    if ((d1Idx != NULL) != (d2Idx != NULL)) return ORC_ERROR_NO_IMPLEMENTATION;
    double tL = 0, t1 = 0, t2 = 0;
    for (int q = 0; q < partitionCount; q++) {
        const int part = partIdx[q];
        if (part < 0 || part >= o->partitionCount) return ORC_ERROR_OUT_OF_RANGE;
        const double* par = o->partials[parIdx[q]];
        const int* st = o->tipStates[childIdx[q]];
        const double* ch = o->partials[childIdx[q]];
        if (!par || (!st && !ch)) return ORC_ERROR_OUT_OF_RANGE;
        const double* Ms[3] = {o->matrices[pIdx[q]], derivs ? o->matrices[d1Idx[q]] : NULL,
                               derivs ? o->matrices[d2Idx[q]] : NULL};
        const double* alw[2] = {NULL, NULL};
        double sums[3];
        edge_range(o, par, ch, st, Ms, o->catWeights[wIdx[q]], o->freqs[fIdx[q]],
                   cumIdx[q] >= 0 ? o->scale[cumIdx[q]] : NULL, alw, o->partStart[part],
                   o->partStart[part + 1], sums);
        outLByPart[q] = sums[0];
        tL += sums[0];
        if (derivs) {
            outD1ByPart[q] = sums[1];
            outD2ByPart[q] = sums[2];
            t1 += sums[1];
            t2 += sums[2];
        }
    }
    int rc = ORC_SUCCESS;
    if (derivs) {
        *outD1 = t1;
        *outD2 = t2;
        if (t1 != t1 || t2 != t2) rc = ORC_ERROR_FLOATING_POINT;
    }
    *outL = tL;
    if (tL != tL) rc = ORC_ERROR_FLOATING_POINT;
    return rc;
}

int orcGetSiteDerivatives(int instance, double* d1, double* d2) {
    Orc* o = get(instance);
    if (!o) return ORC_ERROR_UNINITIALIZED_INSTANCE;
    memcpy(d1, o->siteD1, o->P * sizeof(double));
    if (d2) memcpy(d2, o->siteD2, o->P * sizeof(double));
    return ORC_SUCCESS;
}
