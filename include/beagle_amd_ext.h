/*
 * beagle_amd_ext.h -- extension entry points of libhmsbeagle_amd.so that
 * have no counterpart in the reference API (libhmsbeagle/beagle.h).  They
 * add to beagle.h without changing any reference signature.
 */
#ifndef BEAGLE_AMD_EXT_H
#define BEAGLE_AMD_EXT_H

#include "beagle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Returns the HIP stream (hipStream_t) on which the instance enqueues all
 * of its work, so a caller can order its own work after it or bracket it
 * with events.  No counterpart in the reference. */
int beagleExtGetStream(int instance, void** outStream);

/* Matrix-core engine of the partials products of an FP32 instance with 5
 * to 64 states: post-order, pre-order and edge-derivative operations (no
 * counterpart in the reference; the rate-matrix gradient stays on the FP32
 * matrix cores).
 * terms = 0: FP32 matrix cores (the default).  terms = 3: BF16 matrix cores
 * with every FP32 operand split into a BF16 high and low part, x = hi + lo,
 * and three products per pair (lo*hi + hi*lo + hi*hi, FP32 accumulation);
 * terms = 4 adds lo*lo.  Returns BEAGLE_ERROR_NO_IMPLEMENTATION for a
 * double-precision instance or another state count, BEAGLE_ERROR_OUT_OF_RANGE
 * for other values of terms.  Applies to operations enqueued afterwards. */
int beagleExtSetSplitBF16(int instance, int terms);

/* Autotuning of launch plans (no counterpart in the reference; on by
 * default, BEAGLE_AMD_AUTOTUNE=0 in the environment turns it off).  Calls
 * are grouped into shape classes: the call kind (post-order, pre-order,
 * fused pre-order + derivatives, matrix-core post-order), the number of
 * operations and the mean width of their dependency levels (below 2, below
 * 16, 16 or more); the instance's sizes are fixed.  The first call of a
 * class times each candidate plan (4 states: one fused launch over the
 * list, or dependency levels batched above a width of 4, 8, 16 or 1
 * operations; S > 4 in FP32: partial loads pipelined or a whole tile issued
 * at once) with events on the instance's stream, without waiting; a later
 * call of the class reads the timings once they are complete and keeps the
 * fastest plan, and runs the default plan until then.  New topologies of
 * one tree size therefore tune once.  Plans give identical results.  Lists
 * that add to a cumulative scale buffer, calls being captured into a graph
 * and classes beyond the 64th use the default plan.  on = 0 turns it off
 * and forgets the choices. */
int beagleExtSetAutotune(int instance, int on);

/* The number of shape classes the instance has tuned so far and of those
 * whose timings have not completed yet (no counterpart in the reference;
 * either pointer may be NULL). */
int beagleExtGetAutotuneStats(int instance, int* outTuned, int* outPending);

/* Log-likelihood and per-branch derivative sums of one tree under a
 * partials budget (no counterpart in the reference; the reference keeps
 * every post-order and pre-order partials buffer).  `operations` is the
 * tree's post-order traversal as for beagleUpdatePartials, with node ids in
 * place of destination buffers: a tip is its tip buffer, every internal
 * node an id >= the tip count that need not be a buffer; scale indices must
 * be BEAGLE_OP_NONE and the last operation's destination is the root.
 * firstDerivativeIndices / secondDerivativeIndices give, per node id, the
 * derivative matrices of the branch above the node (r_c Q and r_c^2 Q^2).
 * The call uses only the bufferCount partials buffers in bufferIndices
 * (distinct, not tip buffers):
 * it keeps the post-order partials of internal nodes whose height (edges to
 * the deepest tip below) is a multiple of `interval`, and of the root,
 * recomputes the others from those below them when the pre-order pass
 * reaches them, and holds each node's pre-order partials only between the
 * jobs that write and read them.  interval = 0 picks the smallest interval
 * whose schedule fits the buffers.  Outputs, per node id (the root's
 * entries untouched): sum_p w_p L'/L, sum_p w_p (L'/L)^2 and sum_p w_p L''/L
 * of the branch above the node; the log-likelihood; the interval used and
 * the most buffers in use at once.  With all four result pointers null the
 * call only plans.  BEAGLE_ERROR_GENERAL when no schedule fits the buffers;
 * BEAGLE_ERROR_NO_IMPLEMENTATION with partitions, SCALING_ALWAYS or the
 * transposed pre-order convention (S > 4 without PREORDER_TRANSPOSE_AUTO).
 * The results equal those of a full evaluation of the same tree. */
int beagleExtCheckpointedGradient(int instance, const BeagleOperation* operations, int count,
                                  const int* firstDerivativeIndices,
                                  const int* secondDerivativeIndices,
                                  int categoryWeightsIndex, int stateFrequenciesIndex,
                                  const int* bufferIndices, int bufferCount, int interval,
                                  int* outInterval, int* outPeakBuffers, double* outLogL,
                                  double* outSumFirstDerivatives,
                                  double* outSumSquaredFirstDerivatives,
                                  double* outSumSecondDerivatives);

/* The schedule beagleExtCheckpointedGradient would run for a tree with
 * bufferCount buffers, without an instance or a device (no counterpart in
 * the reference): the interval (interval = 0: the smallest that fits), the
 * most buffers in use at once, the internal nodes kept as checkpoints and
 * the post-order operations run again.  tipPreorderBuffers = 1 when the
 * instance stores tips' pre-order partials too (every state count other
 * than 4, and more than 8 categories).  Return codes as
 * beagleExtCheckpointedGradient. */
int beagleExtPlanCheckpoints(int tipCount, const BeagleOperation* operations, int count,
                             int bufferCount, int interval, int tipPreorderBuffers,
                             int* outInterval, int* outPeakBuffers, int* outCheckpoints,
                             int* outRecomputed);

/* HIP-graph capture of an instance's calls (no counterpart in the
 * reference).  Between beagleExtBeginCapture and beagleExtEndCapture the
 * instance's calls are recorded into a graph instead of run:
 * transition-matrix updates, post- and pre-order traversals, root
 * log-likelihoods, beagleExtUpdatePrePartialsWithDerivatives,
 * beagleExtCalculateLogLikelihoodAndDerivatives and, for 4 states with at
 * most 8 categories, beagleExtCalculateRateMatrixGradient (its Frechet
 * factors come from the captured transition-matrix updates, so replays at
 * new branch lengths give the new gradient).  A call that must wait for the
 * device (getters, set* copies, the rate-matrix gradient for other state
 * counts) fails and spoils the capture: EndCapture then returns
 * BEAGLE_ERROR_GENERAL and records nothing.  Work buffers must already have
 * their size, so run the same calls once before capturing.  The output
 * arrays passed to captured calls receive each replay's results when
 * beagleExtWaitGraph returns; they must stay valid while the graph lives.
 * Uploads are staged in pinned memory owned by the graph; no memset node is
 * recorded. */
int beagleExtBeginCapture(int instance);
int beagleExtEndCapture(int instance, int* outGraph);
/* Enqueues one replay on the instance's stream; an instance without
 * BEAGLE_FLAG_COMPUTATION_ASYNCH also waits, as beagleExtWaitGraph. */
int beagleExtLaunchGraph(int instance, int graph);
/* Waits for the instance's stream and writes the last replay's results to
 * the captured output arrays; BEAGLE_ERROR_FLOATING_POINT on a NaN
 * log-likelihood. */
int beagleExtWaitGraph(int instance, int graph);
/* New branch lengths for the graph's transition-matrix updates, in the
 * order the matrix indices were captured (count = their total number);
 * waits for a replay in flight first. */
int beagleExtSetGraphBranchLengths(int instance, int graph, const double* lengths, int count);
int beagleExtDestroyGraph(int instance, int graph);

/* Pre-order operations fused with the derivative sums of each operation's
 * branch.  Equivalent to
 *   beagleUpdatePrePartials(instance, operations, operationCount, OP_NONE);
 *   beagleCalculateEdgeDerivatives(instance, postBufferIndices, pre, D1, w,
 *       operationCount, NULL, outSumFirstDerivatives,
 *       outSumSquaredFirstDerivatives);
 *   beagleCalculateEdgeDerivatives(instance, postBufferIndices, pre, D2, w,
 *       operationCount, NULL, outSumSecondDerivatives, NULL);
 * with pre[o] = operations[o].destinationPartials, D1 =
 * firstDerivativeIndices, D2 = secondDerivativeIndices and w =
 * {categoryWeightsIndex}: with D1 = r_c Q and D2 = r_c^2 Q^2 (the caller of
 * examples/hmctest/hmctest.cpp:398-427 sets them so) the outputs give per
 * branch dlogL/dt = outSumFirstDerivatives[o] and d2logL/dt2 =
 * outSumSecondDerivatives[o] - outSumSquaredFirstDerivatives[o].  For
 * 4 states the derivative sums are computed inside the pre-order traversal
 * (post- and pre-order partials read once).  Output pointers may be NULL;
 * with all three NULL only the pre-order partials are produced, and an
 * instance created with BEAGLE_FLAG_COMPUTATION_ASYNCH returns without
 * waiting for the device.
 * An operation whose destinationPartials is BEAGLE_OP_NONE yields only its
 * derivative sums and keeps no pre-order partials (for branches to tips,
 * whose pre-order partials no later operation reads); this needs the
 * fused 4-state path (instances with 4 states and at most 8 categories)
 * or, for 5 to 64 states, a single-precision instance and a tip set with
 * beagleSetTipStates as the operation's post-order buffer (its derivative
 * sums are then formed from the tip's state without the tip's pre-order
 * partials), and returns BEAGLE_ERROR_OUT_OF_RANGE elsewhere.
 * No counterpart in the reference. */
int beagleExtUpdatePrePartialsWithDerivatives(int instance, const BeagleOperation* operations,
                                              int operationCount, const int* postBufferIndices,
                                              const int* firstDerivativeIndices,
                                              const int* secondDerivativeIndices,
                                              int categoryWeightsIndex,
                                              double* outSumFirstDerivatives,
                                              double* outSumSquaredFirstDerivatives,
                                              double* outSumSecondDerivatives);

/* One log-likelihood + branch-gradient evaluation with a single wait for
 * the device.  Equivalent to
 *   beagleCalculateRootLogLikelihoods(instance, &rootBufferIndex,
 *       &categoryWeightsIndex, &stateFrequenciesIndex,
 *       &cumulativeScaleIndex, 1, outLogLikelihood);
 *   beagleSetRootPrePartials(instance, &rootPreBufferIndex,
 *       &stateFrequenciesIndex, 1);
 *   beagleExtUpdatePrePartialsWithDerivatives(instance, operations,
 *       operationCount, postBufferIndices, firstDerivativeIndices,
 *       secondDerivativeIndices, categoryWeightsIndex,
 *       outSumFirstDerivatives, outSumSquaredFirstDerivatives,
 *       outSumSecondDerivatives);
 * with the three steps enqueued back to back and their results copied
 * back together (for 4 states; otherwise the three calls in turn).  Made
 * for gradient-based samplers that need logL and the gradient at each
 * step.  outLogLikelihood must not be NULL.  No counterpart in the
 * reference. */
int beagleExtCalculateLogLikelihoodAndDerivatives(
    int instance, int rootBufferIndex, int categoryWeightsIndex, int stateFrequenciesIndex,
    int cumulativeScaleIndex, int rootPreBufferIndex, const BeagleOperation* operations,
    int operationCount, const int* postBufferIndices, const int* firstDerivativeIndices,
    const int* secondDerivativeIndices, double* outLogLikelihood, double* outSumFirstDerivatives,
    double* outSumSquaredFirstDerivatives, double* outSumSecondDerivatives);

/* Gradient of the log-likelihood with respect to the rate matrix of one
 * eigen-system and to the root frequencies.  `operations` are the tree's
 * pre-order operations (as given to beagleUpdatePrePartials) with
 * postBufferIndices[o] the post-order buffer of operation o's node; the
 * pre-order partials of every parent must be current (after
 * beagleUpdatePrePartials or one of the derivative extensions; the
 * operations' own destinations are not read), the post-order partials
 * too, and every P_self must have been computed by
 * beagleUpdateTransitionMatrices(WithMultipleModels) from eigen-system
 * eigenIndex.  With Q = V diag(lambda) V^-1 as given to
 * beagleSetEigenDecomposition and P = exp(r_c t Q):
 *   outGradientRateMatrix[i*S + j] = dlogL / dQ_ij   (S*S, may be NULL)
 *   outGradientFrequencies[i] = dlogL / dpi_i through the root
 *       distribution pi = stateFrequenciesIndex of root buffer
 *       rootBufferIndex (S, may be NULL).
 * Exact (Frechet derivative of the matrix exponential through the
 * eigen-system), unlike beagleCalculateCrossProductDerivative.  Chain rules
 * to exchangeabilities and frequencies inside Q belong to the caller.
 * Any state count up to 64: 4 states with at most 8 categories go through
 * the node-fused kernel, every other case one branch per operation.  For
 * S > 4 the instance needs BEAGLE_FLAG_PREORDER_TRANSPOSE_AUTO (the
 * operations name untransposed matrices); otherwise, and above 64 states,
 * BEAGLE_ERROR_NO_IMPLEMENTATION.  No counterpart in the reference. */
int beagleExtCalculateRateMatrixGradient(int instance, int eigenIndex,
                                          const BeagleOperation* operations, int operationCount,
                                          const int* postBufferIndices, int categoryWeightsIndex,
                                          int rootBufferIndex, int stateFrequenciesIndex,
                                          double* outGradientRateMatrix,
                                          double* outGradientFrequencies);

/* beagleExtCalculateLogLikelihoodAndDerivatives and
 * beagleExtCalculateRateMatrixGradient in one call with one wait:
 *   beagleExtCalculateLogLikelihoodAndDerivatives(instance, rootBufferIndex,
 *       categoryWeightsIndex, stateFrequenciesIndex, cumulativeScaleIndex,
 *       rootPreBufferIndex, operations, operationCount, postBufferIndices,
 *       firstDerivativeIndices, secondDerivativeIndices, outLogLikelihood,
 *       outSumFirstDerivatives, outSumSquaredFirstDerivatives,
 *       outSumSecondDerivatives);
 *   beagleExtCalculateRateMatrixGradient(instance, eigenIndex, operations,
 *       operationCount, postBufferIndices, categoryWeightsIndex,
 *       rootBufferIndex, stateFrequenciesIndex, outGradientRateMatrix,
 *       outGradientFrequencies);
 * the full gradient a sampler of branch lengths and substitution-model
 * parameters needs at each step.  For 4 states in FP32 with 1, 2 or 4
 * categories and node-paired operations the rate-matrix gradient is formed
 * inside the derivative pass, from the partials and matrices that pass
 * already reads (no second traversal); otherwise the two calls run in turn.
 * Same arguments, outputs and errors as the two calls; recorded by
 * beagleExtBeginCapture on the fused path.  With every output pointer NULL
 * (fused path only; otherwise BEAGLE_ERROR_OUT_OF_RANGE) the work is
 * enqueued and nothing is copied back: an instance created with
 * BEAGLE_FLAG_COMPUTATION_ASYNCH returns without waiting.  The summation order of the
 * rate-matrix sums follows the traversal plan the instance picks, so with
 * autotuning on the first (tuning) call may differ from later ones in the
 * last bits.  No counterpart in the reference. */
int beagleExtCalculateLogLikelihoodAndGradient(
    int instance, int rootBufferIndex, int categoryWeightsIndex, int stateFrequenciesIndex,
    int cumulativeScaleIndex, int rootPreBufferIndex, const BeagleOperation* operations,
    int operationCount, const int* postBufferIndices, const int* firstDerivativeIndices,
    const int* secondDerivativeIndices, int eigenIndex, double* outLogLikelihood,
    double* outSumFirstDerivatives, double* outSumSquaredFirstDerivatives,
    double* outSumSecondDerivatives, double* outGradientRateMatrix,
    double* outGradientFrequencies);

/* beagleCalculateRootLogLikelihoods for one root (count = 1, a category
 * weights index >= 0) that allocates no device memory.  No counterpart in
 * the reference: for callers that cannot call hipMalloc (serving loops,
 * harnesses that forbid allocation, code running while a stream is being
 * captured).  Arguments and results as beagleCalculateRootLogLikelihoods;
 * the site log-likelihoods are kept for beagleGetSiteLogLikelihoods.  The
 * root kernels' workgroup sums and the result go to a fixed 256 KiB device
 * array per GPU (held until the process ends) instead of instance buffers
 * grown on demand; the result is returned through pinned memory made when
 * the instance was created.  Limits: at most 1,048,544 patterns with more
 * than 4 states, 8,388,352 with 4 or fewer (BEAGLE_ERROR_OUT_OF_RANGE
 * beyond); count != 1 or a negative weights index gives
 * BEAGLE_ERROR_NO_IMPLEMENTATION.  Every instance on one GPU shares that
 * GPU's array: calls that overlap in time on one GPU (instances driven from
 * different threads) race.  Not recorded by beagleExtBeginCapture. */
int beagleExtCalculateRootLogLikelihoodsNoAlloc(int instance, const int* bufferIndices,
                                                const int* categoryWeightsIndices,
                                                const int* stateFrequenciesIndices,
                                                const int* cumulativeScaleIndices, int count,
                                                double* outSumLogLikelihood);

#ifdef __cplusplus
}
#endif

#endif /* BEAGLE_AMD_EXT_H */
