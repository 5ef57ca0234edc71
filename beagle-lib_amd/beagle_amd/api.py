"""ctypes binding of the BEAGLE C API.

One class wraps any shared library that exports the API: this package's
MI355X library (``libhmsbeagle_amd.so``, prefix ``beagle``), the reference
build, or the CPU restatement used by the tests (prefix ``orc``).  Function
names, argument order and meaning follow the reference header
(libhmsbeagle/beagle.h:280-1399); arrays are numpy arrays converted to the
C types the header declares.
"""
import ctypes
import os

import numpy as np

c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)

# Return codes and flags (libhmsbeagle/beagle.h:110-177).
SUCCESS = 0
ERROR_GENERAL = -1
ERROR_OUT_OF_MEMORY = -2
ERROR_UNIDENTIFIED_EXCEPTION = -3
ERROR_UNINITIALIZED_INSTANCE = -4
ERROR_OUT_OF_RANGE = -5
ERROR_NO_RESOURCE = -6
ERROR_NO_IMPLEMENTATION = -7
ERROR_FLOATING_POINT = -8

FLAG_PRECISION_SINGLE = 1 << 0
FLAG_PRECISION_DOUBLE = 1 << 1
FLAG_COMPUTATION_SYNCH = 1 << 2
FLAG_COMPUTATION_ASYNCH = 1 << 3
FLAG_EIGEN_REAL = 1 << 4
FLAG_EIGEN_COMPLEX = 1 << 5
FLAG_SCALING_MANUAL = 1 << 6
FLAG_SCALING_AUTO = 1 << 7
FLAG_SCALING_ALWAYS = 1 << 8
FLAG_SCALING_DYNAMIC = 1 << 25
FLAG_SCALERS_RAW = 1 << 9
FLAG_SCALERS_LOG = 1 << 10
FLAG_INVEVEC_STANDARD = 1 << 20
FLAG_INVEVEC_TRANSPOSED = 1 << 21
FLAG_VECTOR_SSE = 1 << 11
FLAG_VECTOR_AVX = 1 << 24
FLAG_VECTOR_NONE = 1 << 12
FLAG_THREADING_CPP = 1 << 30
FLAG_THREADING_OPENMP = 1 << 13
FLAG_THREADING_NONE = 1 << 14
FLAG_PROCESSOR_CPU = 1 << 15
FLAG_PROCESSOR_GPU = 1 << 16
FLAG_FRAMEWORK_CUDA = 1 << 22
FLAG_FRAMEWORK_OPENCL = 1 << 23
FLAG_FRAMEWORK_CPU = 1 << 27
FLAG_PARALLELOPS_STREAMS = 1 << 28
FLAG_PARALLELOPS_GRID = 1 << 29
FLAG_PREORDER_TRANSPOSE_MANUAL = 1 << 30
FLAG_PREORDER_TRANSPOSE_AUTO = -(1 << 31)  # 1 << 31 as a C int enumerator
FLAG_FRAMEWORK_HIP = 1 << 32
OP_NONE = -1


class BeagleError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} returned {code}")
        self.fn = fn
        self.code = code


class InstanceDetails(ctypes.Structure):
    _fields_ = [("resourceNumber", ctypes.c_int), ("resourceName", ctypes.c_char_p),
                ("implName", ctypes.c_char_p), ("implDescription", ctypes.c_char_p),
                ("flags", ctypes.c_long)]


class Resource(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("description", ctypes.c_char_p),
                ("supportFlags", ctypes.c_long), ("requiredFlags", ctypes.c_long)]


class ResourceList(ctypes.Structure):
    _fields_ = [("list", ctypes.POINTER(Resource)), ("length", ctypes.c_int)]


# name -> (restype, argtypes) for every exported function of the header.
_I, _L, _D = ctypes.c_int, ctypes.c_long, ctypes.c_double
_IP, _DP, _VP = c_int_p, c_double_p, ctypes.c_void_p
SIGNATURES = {
    "GetVersion": (ctypes.c_char_p, []),
    "GetCitation": (ctypes.c_char_p, []),
    "GetResourceList": (ctypes.POINTER(ResourceList), []),
    "GetBenchmarkedResourceList": (_VP, [_I, _I, _I, _I, _I, _IP, _I, _L, _L, _I, _I, _I, _L]),
    "CreateInstance": (_I, [_I, _I, _I, _I, _I, _I, _I, _I, _I, _IP, _I, _L, _L, _VP]),
    "FinalizeInstance": (_I, [_I]),
    "Finalize": (_I, []),
    "SetCPUThreadCount": (_I, [_I, _I]),
    "SetTipStates": (_I, [_I, _I, _IP]),
    "SetTipPartials": (_I, [_I, _I, _DP]),
    "SetRootPrePartials": (_I, [_I, _IP, _IP, _I]),
    "SetPartials": (_I, [_I, _I, _DP]),
    "GetPartials": (_I, [_I, _I, _I, _DP]),
    "SetEigenDecomposition": (_I, [_I, _I, _DP, _DP, _DP]),
    "SetStateFrequencies": (_I, [_I, _I, _DP]),
    "SetCategoryWeights": (_I, [_I, _I, _DP]),
    "SetCategoryRates": (_I, [_I, _DP]),
    "SetCategoryRatesWithIndex": (_I, [_I, _I, _DP]),
    "SetPatternWeights": (_I, [_I, _DP]),
    "SetPatternPartitions": (_I, [_I, _I, _IP]),
    "ConvolveTransitionMatrices": (_I, [_I, _IP, _IP, _IP, _I]),
    "AddTransitionMatrices": (_I, [_I, _IP, _IP, _IP, _I]),
    "TransposeTransitionMatrices": (_I, [_I, _IP, _IP, _I]),
    "UpdateTransitionMatrices": (_I, [_I, _I, _IP, _IP, _IP, _DP, _I]),
    "UpdateTransitionMatricesWithModelCategories": (_I, [_I, _IP, _IP, _IP, _IP, _DP, _I]),
    "UpdateTransitionMatricesWithMultipleModels": (_I, [_I, _IP, _IP, _IP, _IP, _IP, _DP, _I]),
    "SetTransitionMatrix": (_I, [_I, _I, _DP, _D]),
    "SetDifferentialMatrix": (_I, [_I, _I, _DP]),
    "GetTransitionMatrix": (_I, [_I, _I, _DP]),
    "SetTransitionMatrices": (_I, [_I, _IP, _DP, _DP, _I]),
    "UpdatePartials": (_I, [_I, _IP, _I, _I]),
    "UpdatePrePartials": (_I, [_I, _IP, _I, _I]),
    "UpdatePartialsByPartition": (_I, [_I, _IP, _I]),
    "UpdatePrePartialsByPartition": (_I, [_I, _IP, _I]),
    "WaitForPartials": (_I, [_I, _IP, _I]),
    "AccumulateScaleFactors": (_I, [_I, _IP, _I, _I]),
    "AccumulateScaleFactorsByPartition": (_I, [_I, _IP, _I, _I, _I]),
    "RemoveScaleFactors": (_I, [_I, _IP, _I, _I]),
    "RemoveScaleFactorsByPartition": (_I, [_I, _IP, _I, _I, _I]),
    "ResetScaleFactors": (_I, [_I, _I]),
    "ResetScaleFactorsByPartition": (_I, [_I, _I, _I]),
    "CopyScaleFactors": (_I, [_I, _I, _I]),
    "GetScaleFactors": (_I, [_I, _I, _DP]),
    "CalculateRootLogLikelihoods": (_I, [_I, _IP, _IP, _IP, _IP, _I, _DP]),
    "CalculateRootLogLikelihoodsByPartition": (_I, [_I, _IP, _IP, _IP, _IP, _IP, _I, _I, _DP, _DP]),
    "CalculateEdgeDerivative": (_I, [_I, _IP, _IP, _I, _IP, _IP, _I, _I, _I, _IP, _I, _DP, _DP]),
    "CalculateEdgeDerivatives": (_I, [_I, _IP, _IP, _IP, _IP, _I, _DP, _DP, _DP]),
    "CalculateCrossProductDerivative": (_I, [_I, _IP, _IP, _IP, _IP, _DP, _I, _DP, _DP]),
    "CalculateEdgeLogLikelihoods": (_I, [_I, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _I, _DP, _DP, _DP]),
    "CalculateEdgeLogLikelihoodsByPartition": (
        _I, [_I, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _I, _I, _DP, _DP, _DP, _DP, _DP, _DP]),
    "GetLogLikelihood": (_I, [_I, _DP]),
    "GetDerivatives": (_I, [_I, _DP, _DP]),
    "GetSiteLogLikelihoods": (_I, [_I, _DP]),
    "GetSiteDerivatives": (_I, [_I, _DP, _DP]),
    # extensions (include/beagle_amd_ext.h)
    "ExtGetStream": (_I, [_I, ctypes.POINTER(ctypes.c_void_p)]),
    "ExtSetSplitBF16": (_I, [_I, _I]),
    "ExtSetAutotune": (_I, [_I, _I]),
    "ExtGetAutotuneStats": (_I, [_I, _IP, _IP]),
    "ExtCheckpointedGradient": (_I, [_I, _IP, _I, _IP, _IP, _I, _I, _IP, _I, _I, _IP, _IP, _DP,
                                     _DP, _DP, _DP]),
    "ExtPlanCheckpoints": (_I, [_I, _IP, _I, _I, _I, _I, _IP, _IP, _IP, _IP]),
    "ExtBeginCapture": (_I, [_I]),
    "ExtEndCapture": (_I, [_I, ctypes.POINTER(ctypes.c_int)]),
    "ExtLaunchGraph": (_I, [_I, _I]),
    "ExtWaitGraph": (_I, [_I, _I]),
    "ExtSetGraphBranchLengths": (_I, [_I, _I, _DP, _I]),
    "ExtDestroyGraph": (_I, [_I, _I]),
    "ExtUpdatePrePartialsWithDerivatives": (_I, [_I, _IP, _I, _IP, _IP, _IP, _I, _DP, _DP, _DP]),
    "ExtCalculateLogLikelihoodAndDerivatives": (_I, [_I, _I, _I, _I, _I, _I, _IP, _I, _IP, _IP,
                                                     _IP, _DP, _DP, _DP, _DP]),
    "ExtCalculateRateMatrixGradient": (_I, [_I, _I, _IP, _I, _IP, _I, _I, _I, _DP, _DP]),
    "ExtCalculateLogLikelihoodAndGradient": (_I, [_I, _I, _I, _I, _I, _I, _IP, _I, _IP, _IP,
                                                  _IP, _I, _DP, _DP, _DP, _DP, _DP, _DP]),
    "ExtCalculateRootLogLikelihoodsNoAlloc": (_I, [_I, _IP, _IP, _IP, _IP, _I, _DP]),
}


def _ip(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(c_int_p)


def _dp(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(c_double_p)


class Library:
    """A loaded library exporting the BEAGLE API under ``prefix``."""

    def __init__(self, path, prefix="beagle", mode=ctypes.RTLD_LOCAL):
        self.path = os.path.abspath(path)
        self.prefix = prefix
        self.lib = ctypes.CDLL(self.path, mode=mode)
        self.fn = {}
        for name, (res, args) in SIGNATURES.items():
            sym = getattr(self.lib, prefix + name, None)
            if sym is None:
                continue
            sym.restype = res
            sym.argtypes = args
            self.fn[name] = sym

    def has(self, name):
        return name in self.fn

    def plan_checkpoints(self, tip_count, ops, buffers, interval=0, tip_preorder=False):
        """Extension (no instance): the checkpointed gradient's schedule for
        a tree with `buffers` partials buffers.  Returns (interval, peak
        buffers, checkpoints, recomputed operations)."""
        a = _ip(np.asarray(ops).reshape(-1))
        out = [ctypes.c_int(0) for _ in range(4)]
        rc = self.fn["ExtPlanCheckpoints"](tip_count, a[1], len(a[0]) // 7, buffers, interval,
                                           int(bool(tip_preorder)),
                                           *[ctypes.byref(x) for x in out])
        _check(self.prefix + "ExtPlanCheckpoints", rc)
        return tuple(x.value for x in out)

    def call(self, name, *args):
        return self.fn[name](*args)

    def version(self):
        return self.fn["GetVersion"]().decode()

    def resources(self):
        rl = self.fn["GetResourceList"]().contents
        out = []
        for i in range(rl.length):
            r = rl.list[i]
            out.append(dict(number=i, name=(r.name or b"").decode(),
                            description=(r.description or b"").decode(),
                            supportFlags=r.supportFlags, requiredFlags=r.requiredFlags))
        return out

    def create(self, tips, partials_buffers, compact_buffers, states, patterns, eigen_buffers,
               matrix_buffers, categories, scale_buffers, resources=None, preference=0,
               requirement=0):
        det = InstanceDetails()
        rl = _ip(resources) if resources is not None else None
        rc = self.fn["CreateInstance"](tips, partials_buffers, compact_buffers, states, patterns,
                                       eigen_buffers, matrix_buffers, categories, scale_buffers,
                                       rl[1] if rl else None, len(resources) if resources else 0,
                                       preference, requirement, ctypes.byref(det))
        if rc < 0:
            raise BeagleError(self.prefix + "CreateInstance", rc)
        return Instance(self, rc, det, states, patterns, categories)


def _check(name, rc, allow=()):
    if rc < 0 and rc not in allow:
        raise BeagleError(name, rc)
    return rc


class Instance:
    """Thin object wrapper; every method is one API call."""

    def __init__(self, lib, handle, details, states, patterns, categories):
        self.lib = lib
        self.h = handle
        self.S = states
        self.P = patterns
        self.C = categories
        self.flags = details.flags
        self.impl_name = (details.implName or b"").decode()
        self.resource_name = (details.resourceName or b"").decode()
        self.resource_number = details.resourceNumber

    def _c(self, name, *args, allow=()):
        return _check(self.lib.prefix + name, self.lib.call(name, self.h, *args), allow)

    def finalize(self):
        if self.h is not None:
            self.lib.call("FinalizeInstance", self.h)
            self.h = None

    def set_tip_states(self, tip, states):
        a = _ip(states)
        return self._c("SetTipStates", tip, a[1])

    def set_tip_partials(self, tip, partials):
        a = _dp(partials)
        return self._c("SetTipPartials", tip, a[1])

    def set_partials(self, buf, partials):
        a = _dp(partials)
        return self._c("SetPartials", buf, a[1])

    def get_partials(self, buf, scale_index=OP_NONE):
        out = np.zeros(self.C * self.P * self.S)
        self._c("GetPartials", buf, scale_index, out.ctypes.data_as(c_double_p))
        return out

    def set_root_pre_partials(self, bufs, freqs):
        a, b = _ip(bufs), _ip(freqs)
        return self._c("SetRootPrePartials", a[1], b[1], len(a[0]))

    def set_eigen(self, index, evec, ivec, evals):
        a, b, c = _dp(evec), _dp(ivec), _dp(evals)
        return self._c("SetEigenDecomposition", index, a[1], b[1], c[1])

    def set_state_frequencies(self, index, f):
        a = _dp(f)
        return self._c("SetStateFrequencies", index, a[1])

    def set_category_weights(self, index, w):
        a = _dp(w)
        return self._c("SetCategoryWeights", index, a[1])

    def set_category_rates(self, r):
        a = _dp(r)
        return self._c("SetCategoryRates", a[1])

    def set_category_rates_with_index(self, index, r):
        a = _dp(r)
        return self._c("SetCategoryRatesWithIndex", index, a[1])

    def set_pattern_weights(self, w):
        a = _dp(w)
        return self._c("SetPatternWeights", a[1])

    def set_pattern_partitions(self, count, parts):
        a = _ip(parts)
        return self._c("SetPatternPartitions", count, a[1])

    def update_transition_matrices(self, eigen, prob, d1, d2, lengths):
        p, a, b, t = _ip(prob), _ip(d1), _ip(d2), _dp(lengths)
        return self._c("UpdateTransitionMatrices", eigen, p[1], a[1] if a else None,
                       b[1] if b else None, t[1], len(p[0]))

    def update_transition_matrices_multi(self, eigens, rates, prob, d1, d2, lengths):
        e, r, p, a, b, t = _ip(eigens), _ip(rates), _ip(prob), _ip(d1), _ip(d2), _dp(lengths)
        return self._c("UpdateTransitionMatricesWithMultipleModels", e[1], r[1], p[1],
                       a[1] if a else None, b[1] if b else None, t[1], len(p[0]))

    def update_transition_matrices_model_categories(self, eigens, prob, d1, d2, lengths):
        e, p, a, b, t = _ip(eigens), _ip(prob), _ip(d1), _ip(d2), _dp(lengths)
        return self._c("UpdateTransitionMatricesWithModelCategories", e[1], p[1],
                       a[1] if a else None, b[1] if b else None, t[1], len(p[0]))

    def convolve_transition_matrices(self, first, second, result):
        a, b, c = _ip(first), _ip(second), _ip(result)
        return self._c("ConvolveTransitionMatrices", a[1], b[1], c[1], len(c[0]))

    def set_transition_matrix(self, index, m, pad):
        a = _dp(m)
        return self._c("SetTransitionMatrix", index, a[1], float(pad))

    def set_differential_matrix(self, index, m):
        a = _dp(m)
        return self._c("SetDifferentialMatrix", index, a[1])

    def set_transition_matrices(self, idx, ms, pads):
        a, b, c = _ip(idx), _dp(ms), _dp(pads)
        return self._c("SetTransitionMatrices", a[1], b[1], c[1], len(a[0]))

    def get_transition_matrix(self, index):
        out = np.zeros(self.C * self.S * self.S)
        self._c("GetTransitionMatrix", index, out.ctypes.data_as(c_double_p))
        return out

    def transpose_transition_matrices(self, src, dst):
        a, b = _ip(src), _ip(dst)
        return self._c("TransposeTransitionMatrices", a[1], b[1], len(a[0]))

    def update_partials(self, ops, cumulative=OP_NONE):
        a = _ip(np.asarray(ops).reshape(-1))
        return self._c("UpdatePartials", a[1], len(a[0]) // 7, cumulative)

    def update_pre_partials(self, ops, cumulative=OP_NONE):
        a = _ip(np.asarray(ops).reshape(-1))
        return self._c("UpdatePrePartials", a[1], len(a[0]) // 7, cumulative)

    def update_partials_by_partition(self, ops):
        a = _ip(np.asarray(ops).reshape(-1))
        return self._c("UpdatePartialsByPartition", a[1], len(a[0]) // 9)

    def update_pre_partials_by_partition(self, ops):
        a = _ip(np.asarray(ops).reshape(-1))
        return self._c("UpdatePrePartialsByPartition", a[1], len(a[0]) // 9)

    def wait_for_partials(self, bufs):
        a = _ip(bufs)
        return self._c("WaitForPartials", a[1], len(a[0]))

    def accumulate_scale_factors(self, idx, cumulative):
        a = _ip(idx)
        return self._c("AccumulateScaleFactors", a[1], len(a[0]), cumulative)

    def accumulate_scale_factors_by_partition(self, idx, cumulative, part):
        a = _ip(idx)
        return self._c("AccumulateScaleFactorsByPartition", a[1], len(a[0]), cumulative, part)

    def remove_scale_factors(self, idx, cumulative):
        a = _ip(idx)
        return self._c("RemoveScaleFactors", a[1], len(a[0]), cumulative)

    def remove_scale_factors_by_partition(self, idx, cumulative, part):
        a = _ip(idx)
        return self._c("RemoveScaleFactorsByPartition", a[1], len(a[0]), cumulative, part)

    def reset_scale_factors(self, cumulative):
        return self._c("ResetScaleFactors", cumulative)

    def reset_scale_factors_by_partition(self, cumulative, part):
        return self._c("ResetScaleFactorsByPartition", cumulative, part)

    def copy_scale_factors(self, dst, src):
        return self._c("CopyScaleFactors", dst, src)

    def get_scale_factors(self, index):
        out = np.zeros(self.P)
        self._c("GetScaleFactors", index, out.ctypes.data_as(c_double_p))
        return out

    def root_log_likelihood(self, buf, weights_index=0, freqs_index=0, cumulative=OP_NONE,
                            allow=(), out=None):
        """out: a caller-owned float64 array of one element (needed while
        capturing a graph, whose replays write to it)."""
        res = np.zeros(1) if out is None else out
        a, b, c, d = _ip([buf]), _ip([weights_index]), _ip([freqs_index]), _ip([cumulative])
        rc = self._c("CalculateRootLogLikelihoods", a[1], b[1], c[1], d[1], 1,
                     res.ctypes.data_as(c_double_p), allow=allow)
        return float(res[0]) if rc >= 0 else rc

    def root_log_likelihood_no_alloc(self, buf, weights_index=0, freqs_index=0,
                                     cumulative=OP_NONE, allow=()):
        """Extension: the single-root log-likelihood through the entry point
        that allocates no device memory."""
        res = np.zeros(1)
        a, b, c, d = _ip([buf]), _ip([weights_index]), _ip([freqs_index]), _ip([cumulative])
        rc = self._c("ExtCalculateRootLogLikelihoodsNoAlloc", a[1], b[1], c[1], d[1], 1,
                     res.ctypes.data_as(c_double_p), allow=allow)
        return float(res[0]) if rc >= 0 else rc

    def root_log_likelihoods(self, bufs, weights, freqs, cumulative):
        """Several roots combined into one likelihood per pattern
        (count = len(bufs) > 1); returns the summed log-likelihood."""
        res = np.zeros(1)
        a, b, c, d = _ip(bufs), _ip(weights), _ip(freqs), _ip(cumulative)
        self._c("CalculateRootLogLikelihoods", a[1], b[1], c[1], d[1], len(a[0]),
                res.ctypes.data_as(c_double_p))
        return float(res[0])

    def root_log_likelihood_per_category(self, buf, freqs_index=0, cumulative=OP_NONE):
        """Category weights index -1: per-category site log-likelihoods as a
        (categories, patterns) array."""
        res = np.zeros(self.C * self.P)
        a, b, c, d = _ip([buf]), _ip([-1]), _ip([freqs_index]), _ip([cumulative])
        self._c("CalculateRootLogLikelihoods", a[1], b[1], c[1], d[1], 1,
                res.ctypes.data_as(c_double_p))
        return res.reshape(self.C, self.P)

    def root_log_likelihood_by_partition(self, bufs, weights, freqs, cumulative, parts):
        n = len(parts)
        by = np.zeros(n)
        tot = ctypes.c_double(0.0)
        a, b, c, d, e = _ip(bufs), _ip(weights), _ip(freqs), _ip(cumulative), _ip(parts)
        self._c("CalculateRootLogLikelihoodsByPartition", a[1], b[1], c[1], d[1], e[1], n, 1,
                by.ctypes.data_as(c_double_p), ctypes.byref(tot))
        return by, tot.value

    def site_log_likelihoods(self):
        out = np.zeros(self.P)
        self._c("GetSiteLogLikelihoods", out.ctypes.data_as(c_double_p))
        return out

    def log_likelihood(self):
        out = ctypes.c_double(0.0)
        self._c("GetLogLikelihood", ctypes.byref(out))
        return out.value

    def edge_derivatives(self, post, pre, dmat, weights=(0,), per_site=True):
        n = len(post)
        a, b, c, d = _ip(post), _ip(pre), _ip(dmat), _ip(list(weights))
        sites = np.zeros(n * self.P) if per_site else None
        s1 = np.zeros(n)
        s2 = np.zeros(n)
        self._c("CalculateEdgeDerivatives", a[1], b[1], c[1], d[1], n,
                sites.ctypes.data_as(c_double_p) if per_site else None,
                s1.ctypes.data_as(c_double_p), s2.ctypes.data_as(c_double_p))
        return (sites.reshape(n, self.P) if per_site else None), s1, s2

    def cross_products(self, post, pre, rates_index, weights_index, lengths, out=None):
        n = len(post)
        a, b, c, d, t = _ip(post), _ip(pre), _ip([rates_index]), _ip([weights_index]), _dp(lengths)
        if out is None:
            out = np.zeros(self.S * self.S)
        self._c("CalculateCrossProductDerivative", a[1], b[1], c[1], d[1], t[1], n,
                out.ctypes.data_as(c_double_p), None)
        return out

    def edge_log_likelihood(self, parent, child, pmat, d1=None, d2=None, weights_index=0,
                            freqs_index=0, cumulative=OP_NONE):
        out = ctypes.c_double(0.0)
        o1 = ctypes.c_double(0.0)
        o2 = ctypes.c_double(0.0)
        a, b, c = _ip([parent]), _ip([child]), _ip([pmat])
        e1 = _ip([d1]) if d1 is not None else None
        e2 = _ip([d2]) if d2 is not None else None
        w, f, s = _ip([weights_index]), _ip([freqs_index]), _ip([cumulative])
        self._c("CalculateEdgeLogLikelihoods", a[1], b[1], c[1], e1[1] if e1 else None,
                e2[1] if e2 else None, w[1], f[1], s[1], 1, ctypes.byref(out), ctypes.byref(o1),
                ctypes.byref(o2))
        return out.value, o1.value, o2.value

    def edge_log_likelihood_by_partition(self, parents, children, pmats, d1, d2, weights,
                                         freqs, cumulative, parts):
        """Per-partition edge log-likelihoods (and, with d1 and d2 lists,
        first and second derivatives): returns (logL by partition, total,
        d1 by partition, total, d2 by partition, total)."""
        n = len(parts)
        by = [np.zeros(n) for _ in range(3)]
        tot = [ctypes.c_double(0.0) for _ in range(3)]
        a, b, c = _ip(parents), _ip(children), _ip(pmats)
        e1 = _ip(d1) if d1 is not None else None
        e2 = _ip(d2) if d2 is not None else None
        w, f, s, q = _ip(weights), _ip(freqs), _ip(cumulative), _ip(parts)
        self._c("CalculateEdgeLogLikelihoodsByPartition", a[1], b[1], c[1],
                e1[1] if e1 else None, e2[1] if e2 else None, w[1], f[1], s[1], q[1], n, 1,
                by[0].ctypes.data_as(c_double_p), ctypes.byref(tot[0]),
                by[1].ctypes.data_as(c_double_p), ctypes.byref(tot[1]),
                by[2].ctypes.data_as(c_double_p), ctypes.byref(tot[2]))
        return by[0], tot[0].value, by[1], tot[1].value, by[2], tot[2].value

    def pre_partials_with_derivatives(self, ops, post, d1, d2, weights_index=0, outputs=True):
        """Extension: pre-order pass fused with per-branch derivative sums.
        Returns (sum w L'/L, sum w (L'/L)^2, sum w L''/L) per operation;
        with outputs=False all output pointers are NULL (only the pre-order
        partials are produced, and an asynchronous instance returns at
        once)."""
        a = _ip(np.asarray(ops).reshape(-1))
        n = len(a[0]) // 7
        b, c, d = _ip(post), _ip(d1), _ip(d2)
        if not outputs:
            self._c("ExtUpdatePrePartialsWithDerivatives", a[1], n, b[1], c[1], d[1],
                    weights_index, None, None, None)
            return None
        s1, sq, s2 = np.zeros(n), np.zeros(n), np.zeros(n)
        self._c("ExtUpdatePrePartialsWithDerivatives", a[1], n, b[1], c[1], d[1], weights_index,
                s1.ctypes.data_as(c_double_p), sq.ctypes.data_as(c_double_p),
                s2.ctypes.data_as(c_double_p))
        return s1, sq, s2

    def log_likelihood_and_derivatives(self, root, root_pre, ops, post, d1, d2,
                                       weights_index=0, freqs_index=0, cumulative=OP_NONE,
                                       out=None):
        """Extension: root logL, root pre-order partials = frequencies, and
        the fused pre-order + derivative pass with one wait.  Returns
        (logL, sum w L'/L, sum w (L'/L)^2, sum w L''/L).  out: caller-owned
        float64 arrays (logL[1], s1[n], sq[n], s2[n]), needed while
        capturing a graph."""
        a = _ip(np.asarray(ops).reshape(-1))
        n = len(a[0]) // 7
        b, c, d = _ip(post), _ip(d1), _ip(d2)
        lnl, s1, sq, s2 = out if out is not None else (np.zeros(1), np.zeros(n), np.zeros(n),
                                                       np.zeros(n))
        self._c("ExtCalculateLogLikelihoodAndDerivatives", root, weights_index, freqs_index,
                cumulative, root_pre, a[1], n, b[1], c[1], d[1], lnl.ctypes.data_as(c_double_p),
                s1.ctypes.data_as(c_double_p), sq.ctypes.data_as(c_double_p),
                s2.ctypes.data_as(c_double_p))
        return float(lnl[0]), s1, sq, s2

    def log_likelihood_and_gradient(self, root, root_pre, ops, post, d1, d2, eigen=0,
                                    weights_index=0, freqs_index=0, cumulative=OP_NONE,
                                    out=None):
        """Extension: log_likelihood_and_derivatives() and
        rate_matrix_gradient(eigen, ops, post, root=root) with one wait.
        Returns (logL, s1, sq, s2, dlogL/dQ [S][S], dlogL/dpi_root [S]).
        out: caller-owned float64 arrays (logL[1], s1[n], sq[n], s2[n],
        q[S*S], pi[S]), needed while capturing a graph; False: enqueue only,
        nothing returned (no copy-back or wait on an asynchronous
        instance)."""
        a = _ip(np.asarray(ops).reshape(-1))
        n = len(a[0]) // 7
        b, c, d = _ip(post), _ip(d1), _ip(d2)
        S = self.S
        if out is None:
            out = (np.zeros(1), np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(S * S),
                   np.zeros(S))
        if out is False:
            # enqueue only (no outputs, no copy-back)
            self._c("ExtCalculateLogLikelihoodAndGradient", root, weights_index, freqs_index,
                    cumulative, root_pre, a[1], n, b[1], c[1], d[1], eigen,
                    None, None, None, None, None, None)
            return None
        lnl, s1, sq, s2, gq, gf = out
        self._c("ExtCalculateLogLikelihoodAndGradient", root, weights_index, freqs_index,
                cumulative, root_pre, a[1], n, b[1], c[1], d[1], eigen,
                lnl.ctypes.data_as(c_double_p), s1.ctypes.data_as(c_double_p),
                sq.ctypes.data_as(c_double_p), s2.ctypes.data_as(c_double_p),
                gq.ctypes.data_as(c_double_p), gf.ctypes.data_as(c_double_p))
        return float(lnl[0]), s1, sq, s2, gq.reshape(S, S), gf

    def checkpointed_gradient(self, ops, d1, d2, buffers, interval=0, weights_index=0,
                              freqs_index=0, plan_only=False):
        """Extension: logL and per-branch derivative sums of one tree using
        only the partials buffers in `buffers` (post-order partials kept at
        checkpoint heights, the rest recomputed).  ops: post-order operations
        over node ids; d1, d2: derivative matrix index per node id.  Returns
        (logL, s1, sq, s2, interval, peak buffers), the sums indexed by node
        id; plan_only: (None, None, None, None, interval, peak)."""
        a = _ip(np.asarray(ops).reshape(-1))
        n = len(a[0]) // 7
        size = int(np.asarray(ops).reshape(-1, 7)[:, [0, 3, 5]].max()) + 1
        b, c, p = _ip(d1), _ip(d2), _ip(buffers)
        k, peak = ctypes.c_int(0), ctypes.c_int(0)
        if plan_only:
            self._c("ExtCheckpointedGradient", a[1], n, b[1], c[1], weights_index, freqs_index,
                    p[1], len(p[0]), int(interval), ctypes.byref(k), ctypes.byref(peak),
                    None, None, None, None)
            return None, None, None, None, k.value, peak.value
        lnl, s1, sq, s2 = np.zeros(1), np.zeros(size), np.zeros(size), np.zeros(size)
        self._c("ExtCheckpointedGradient", a[1], n, b[1], c[1], weights_index, freqs_index,
                p[1], len(p[0]), int(interval), ctypes.byref(k), ctypes.byref(peak),
                lnl.ctypes.data_as(c_double_p), s1.ctypes.data_as(c_double_p),
                sq.ctypes.data_as(c_double_p), s2.ctypes.data_as(c_double_p))
        return float(lnl[0]), s1, sq, s2, k.value, peak.value

    def rate_matrix_gradient(self, eigen, ops, post, weights_index=0, root=None,
                             freqs_index=0, out=None):
        """Extension: (dlogL/dQ as an S x S array, dlogL/dpi_root or None).
        out: caller-owned float64 arrays (S*S, S or None) that stay alive
        while a graph that captured the call replays."""
        a = _ip(np.asarray(ops).reshape(-1))
        n = len(a[0]) // 7
        b = _ip(post)
        if out is not None:
            gq, gf = out
        else:
            gq = np.zeros(self.S * self.S)
            gf = np.zeros(self.S) if root is not None else None
        self._c("ExtCalculateRateMatrixGradient", eigen, a[1], n, b[1], weights_index,
                -1 if root is None else root, freqs_index, gq.ctypes.data_as(c_double_p),
                gf.ctypes.data_as(c_double_p) if gf is not None else None)
        return gq.reshape(self.S, self.S), gf

    # HIP-graph capture (extension entry points): calls between
    # begin_capture() and end_capture() are recorded; their output arrays
    # (which must stay alive) receive each replay's results at wait_graph().
    def autotune_stats(self):
        """Extension: (shape classes tuned so far, of which pending)."""
        tuned, pending = ctypes.c_int(0), ctypes.c_int(0)
        self._c("ExtGetAutotuneStats", ctypes.byref(tuned), ctypes.byref(pending))
        return tuned.value, pending.value

    def set_autotune(self, on=True):
        """Extension: per-shape-class choice of launch plans by timing."""
        self._c("ExtSetAutotune", int(bool(on)))

    def begin_capture(self):
        self._c("ExtBeginCapture")

    def end_capture(self):
        g = ctypes.c_int(-1)
        self._c("ExtEndCapture", ctypes.byref(g))
        return g.value

    def launch_graph(self, graph):
        self._c("ExtLaunchGraph", int(graph))

    def wait_graph(self, graph):
        self._c("ExtWaitGraph", int(graph))

    def set_graph_branch_lengths(self, graph, lengths):
        a = _dp(lengths)
        self._c("ExtSetGraphBranchLengths", int(graph), a[1], len(a[0]))

    def destroy_graph(self, graph):
        self._c("ExtDestroyGraph", int(graph))

    def set_split_bf16(self, terms=3):
        """S > 4, FP32: post-order products on split-BF16 matrix cores
        (terms 3 or 4), or back on the FP32 matrix cores (terms 0)."""
        self._c("ExtSetSplitBF16", int(terms))

    def stream(self):
        """hipStream_t of the instance (extension entry point)."""
        s = ctypes.c_void_p(0)
        self._c("ExtGetStream", ctypes.byref(s))
        return s.value

    def site_derivatives(self):
        d1 = np.zeros(self.P)
        d2 = np.zeros(self.P)
        self._c("GetSiteDerivatives", d1.ctypes.data_as(c_double_p), d2.ctypes.data_as(c_double_p))
        return d1, d2
