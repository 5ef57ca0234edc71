"""Benchmark: BEAGLE log-likelihood + gradient evaluations on one MI355X.

Workload: 4-state GTR + Gamma(4 categories, shape 0.5), 128 taxa,
100,000 unique site patterns of uniform random tip states, pattern
weights 1 + Poisson(2), branch lengths ~ Exp(mean 0.1), two tree shapes (a
127-deep caterpillar and a balanced tree).  Synthetic data, seed 1.

One evaluation = the log-likelihood and, for all 254 branches, dlogL/dt
and d2logL/dt2: transition matrices from the eigen-system, the post-order
traversal, the root log-likelihood, the root pre-order partials, the
pre-order traversal and the two derivative sums per branch (derivative
matrices r_c Q and r_c^2 Q^2).  One step = one evaluation of each tree.

value: evaluations per second of the FP32 path (partials in FP32 with
exact power-of-two rescaling; derivative ratios summed in FP32 within
64-pattern tiles and in FP64 across tiles; scale exponents and the
log-likelihood in FP64; pre-order fused with the derivative sums), each
tree's evaluation replayed as a HIP graph with its branch lengths set per
evaluation, the two trees run as independent chains (a tree's next
evaluation is launched once its previous result is collected, while the
other tree's runs), with its logL and gradient error against the FP64 path
in the same line (and the eager calls' rate beside it).
Also reported: the plain FP64 path through the beagle.h calls
(beagleUpdatePrePartials + two beagleCalculateEdgeDerivatives), the FP64
fused path, the FP64 post-order traversal's HBM bandwidth, and the
reference's threaded CPU plugin on this host's cores.

codon: the 61-state GY94 codon workload (64 taxa, 20,000 patterns, Gamma(4);
transition matrices + post-order + root logL per evaluation) on the FP64,
FP32 and split-BF16 matrix-core post-order kernels: evaluations/s, error
against FP64, partials GB/s and algorithmic TFLOP/s against each MFMA
peak; logL + full branch-gradient evaluations/s (matrix-core pre-order and
derivative passes); and the reference CPU plugin on a pattern sample.

partitions: 48 partitions with 4-, 20- and 61-state models and their own
trees, FP32 logL + gradient per partition, eager calls against HIP-graph
replays launched across the instances' streams.

by_partition: the same partitions on one shared tree, one instance per
state count holding its partitions as pattern ranges (the calls by
partition, one launch per dependency level over a work list of (operation,
pattern tile) items), against the full tiles x operations grid and against
one instance per partition.

roofline: the dominant kernel of the value path, the FP32 node kernel of
the one-call gradient (pre-order partials + branch derivatives + rate-
matrix sums of the 512-taxon tree's 511 node jobs, one launch per
dependency level), timed per pass with HIP events on the library's stream
after the timed loop: the one call captured alone as a HIP graph and its
replays bracketed back to back, so no host work falls inside the bracket;
algorithmic bytes per pass from preorder_bytes() plus the root's read.  Inputs are resident in
HBM before timing.
"""
import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "beagle-lib_amd"))

HBM_PEAK_GBS = 8000.0
N_TAXA, N_PATTERNS, SEED = 128, 100_000, 1
SHAPES = ("caterpillar", "balanced")
# the adjoint workload the metric is quoted on: 4-state GTR + Gamma(4),
# a 512-taxon Yule tree, 50,000 patterns
ADJ = dict(taxa=512, patterns=50_000, shape="yule")
# the plain FP64 pruning workload: 16 taxa, 10,000 patterns, root logL only
PLAIN = dict(taxa=16, patterns=10_000, shape="yule")
METRIC = "logL+gradient evals/s & partials GB/s vs plain FP64, at ≤1e-6 rel err"
SHORT = 2   # bytes of a power-of-two scale exponent


def postorder_bytes(prob, w, auto=False):
    """Algorithmic HBM bytes of one post-order traversal: per operation the
    destination write C*P*S*w, each child (4*P for tip states, C*P*S*w for
    partials), the two C*S*(S+1)*w matrices, and with automatic rescaling
    the children's and the destination's exponents (2 bytes per pattern)."""
    C, P, S = prob.C, prob.P, prob.S
    n = prob.tree.n
    part = C * P * S * w
    total = 0
    for v in prob.tree.postorder():
        a, b = prob.tree.children[v]
        total += part + 2 * C * S * (S + 1) * w
        total += sum(4 * P if x < n else part for x in (a, b))
        if auto:
            total += SHORT * P * (1 + sum(1 for x in (a, b) if x >= n))
    return total


def preorder_bytes(prob, ops, post_of, w, auto):
    """Algorithmic HBM bytes of one node-fused pre-order + derivative
    launch over `ops` (the library pairs the two operations of each node's
    children into one job, in the order the second one appears).  Per job:
    the node's pre-order partials unless the previous job produced them
    (then they come from registers), each child's post-order partials
    (4*P for tip states), each kept destination write, six C*S*(S+1)*w
    matrices and 48 bytes of derivative sums per 64-pattern tile; with
    automatic rescaling 2 bytes per pattern for each exponent read or
    written.  Plus the pattern weights (8*P), read once per launch."""
    C, P, S = prob.C, prob.P, prob.S
    n = prob.tree.n
    part = C * P * S * w
    tiles = (P + 63) // 64
    waiting, jobs = {}, []
    for o in ops:
        par = int(o[3])
        if par in waiting:
            jobs.append((waiting.pop(par), o))
        else:
            waiting[par] = o
    total, prev = 8 * P, None
    for a, b in jobs:
        par = int(a[3])
        if not (prev is not None and par in (int(prev[0][0]), int(prev[1][0]))):
            total += part + (SHORT * P if auto else 0)
        for x in (a, b):
            own = post_of(x)
            total += 4 * P if own < n else part + (SHORT * P if auto else 0)
            if int(x[0]) >= 0:
                total += part + (SHORT * P if auto else 0)
        total += 6 * C * S * (S + 1) * w + 6 * 8 * tiles
        prev = (a, b)
    return total


def problems(patterns=N_PATTERNS):
    from beagle_amd import synth
    return [synth.make_problem(SEED, N_TAXA, patterns, S=4, shape=s) for s in SHAPES]


def adjoint_problem(patterns=None):
    from beagle_amd import synth
    return synth.make_problem(SEED, ADJ["taxa"], patterns or ADJ["patterns"], S=4,
                              shape=ADJ["shape"])


def full_gradient_bench(lib, steps, warmup):
    """The metric's own workload: one evaluation = logL, dlogL/dt and
    d2logL/dt2 of all 1,022 branches, dlogL/dQ (4 x 4) and dlogL/dpi_root
    of the 512-taxon tree (matrices + post-order + root + fused pre-order
    and derivative pass + rate-matrix and root-frequency gradients).

    value path: FP32 partials, power-of-two rescaling, the evaluation
    captured once as a HIP graph and replayed with its branch lengths set
    before every replay, one evaluation after the other (a gradient-based
    sampler's chain: each evaluation's result is collected before the next
    is launched).  Plain path: the FP64 beagle.h calls as a client makes
    them (manual scale buffers; post-order, root, beagleUpdatePrePartials,
    two beagleCalculateEdgeDerivatives) plus the FP64 rate-matrix gradient,
    every call waiting for its result.  Errors of the value path are taken
    against the plain path."""
    import numpy as np
    import torch
    from beagle_amd import api, driver
    prob = adjoint_problem()
    asyn = api.FLAG_COMPUTATION_ASYNCH
    ev = driver.Evaluator(lib, prob, precision="single", scaling="auto", extra_flags=asyn)
    lengths = prob.tree.lengths
    t0, w = time.perf_counter(), 0
    while w < warmup or time.perf_counter() - t0 < 1.0:
        ev.fused_gradient(rate_matrix=True)
        w += 1
    torch.cuda.synchronize()
    ne = max(steps // 2, 5)
    t0 = time.perf_counter()
    for _ in range(ne):
        eager = ev.fused_gradient(rate_matrix=True)
    torch.cuda.synchronize()
    eager_rate = ne / (time.perf_counter() - t0)
    ev.capture_fused_gradient(rate_matrix=True)
    for _ in range(max(warmup, 3)):
        ev.replay_fused_gradient(lengths)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res32 = ev.replay_fused_gradient(lengths)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    identical = all(np.array_equal(np.asarray(a), np.asarray(b)) for a, b in zip(res32, eager))
    phases = phase_timing(ev, steps)
    ops, post, _, _ = ev.fused_operations()
    post_idx = dict(zip(map(tuple, ops.tolist()), post.tolist()))
    pre_bytes = preorder_bytes(prob, ops, lambda o: post_idx[tuple(o.tolist())], 4, True)
    post_bytes = postorder_bytes(prob, 4, True)
    ev.close()

    ev = driver.Evaluator(lib, prob, scaling="manual")

    def plain():
        lnl, g1, g2 = ev.gradient()
        q, pi = ev.rate_matrix_gradient()
        return lnl, g1, g2, q, pi

    for _ in range(2):
        plain()
    n64 = max(steps // 4, 3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n64):
        res64 = plain()
    plain_rate = n64 / (time.perf_counter() - t0)
    ev.close()
    value = steps / dt
    err = {"logL": rel(res32[0], res64[0]), "dlogL_dt": grad_err(res32[1], res64[1]),
           "d2logL_dt2": grad_err(res32[2], res64[2]), "dlogL_dQ": grad_err(res32[3], res64[3]),
           "dlogL_dpi": grad_err(res32[4], res64[4])}
    return {"value": value, "ms_per_eval": dt * 1e3 / steps, "eager_evals_per_s": eager_rate,
            "graph_results_identical_to_eager": bool(identical), "error_vs_fp64": err,
            "phases_ms": phases, "preorder_bytes": pre_bytes, "postorder_bytes": post_bytes,
            "plain_fp64": {"evals_per_s": plain_rate,
                           "what": "FP64 beagle.h calls with manual scale buffers: "
                                   "beagleUpdateTransitionMatrices, beagleResetScaleFactors, "
                                   "beagleUpdatePartials, beagleCalculateRootLogLikelihoods, "
                                   "beagleSetRootPrePartials, beagleUpdatePrePartials, two "
                                   "beagleCalculateEdgeDerivatives, then "
                                   "beagleExtCalculateRateMatrixGradient in FP64; each call "
                                   "waits for its result",
                           "speedup_of_value": value / plain_rate},
            "logL": res64[0]}


def phase_timing(ev, steps):
    """Device time of each phase of one evaluation (HIP events on the
    instance's stream, ASYNCH instance, medians over `steps` evaluations):
    matrices, post-order, then the one call that forms the root logL, the
    pre-order partials, the branch derivatives and the rate-matrix and
    root-frequency gradients (beagleExtCalculateLogLikelihoodAndGradient),
    and for comparison the same work as two calls (the derivative pass,
    then beagleExtCalculateRateMatrixGradient's own pass)."""
    import numpy as np
    import torch
    st = torch.cuda.ExternalStream(ev.inst.stream())
    ops, post, d1, d2 = ev.fused_operations()
    names = ("matrices", "postorder", "root_preorder_derivatives_rate_matrix",
             "two_calls_root_preorder_derivatives", "two_calls_rate_matrix_gradient")
    acc = {k: [] for k in names}
    root = ev.prob.tree.root
    for _ in range(max(steps // 2, 5)):
        ev_ = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        ev_[0].record(st)
        ev.update_matrices()
        ev_[1].record(st)
        ev.post_order()
        ev_[2].record(st)
        ev.inst.log_likelihood_and_gradient(root, ev.pre_base + root, ops, post, d1, d2)
        ev_[3].record(st)
        ev.inst.log_likelihood_and_derivatives(root, ev.pre_base + root, ops, post, d1, d2)
        ev_[4].record(st)
        ev.rate_matrix_gradient()
        ev_[5].record(st)
        torch.cuda.synchronize()
        for k, name in enumerate(names):
            acc[name].append(ev_[k].elapsed_time(ev_[k + 1]))
    return {k: float(np.median(v)) for k, v in acc.items()}


def fused_eval(ev):
    """One logL + gradient evaluation through the library's extensions:
    matrices, post-order, then one call for the root logL, the root
    pre-order partials and the fused pre-order + derivative pass."""
    return ev.fused_gradient()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """Cores this process may use: the affinity mask, capped by the job's
    thread allowance (OMP_NUM_THREADS) where one is set."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else n


def cpu_baseline(budget_s):
    """The reference CPU plugin (CPU-4State-SSE-Double, threaded) on this
    host's cores, on the adjoint workload's generators (512-taxon Yule tree,
    4-state GTR + Gamma(4)) at a pattern sample sized to the budget: one
    evaluation = the post-order + root phase (beagleUpdateTransitionMatrices,
    beagleUpdatePartials with manual scale buffers,
    beagleCalculateRootLogLikelihoods) and the adjoint phase
    (beagleUpdatePrePartials, beagleCalculateEdgeDerivatives with r_c Q and
    r_c^2 Q^2, beagleCalculateCrossProductDerivative, the reference's
    substitution-parameter gradient).  1 thread and all cores
    (beagleSetCPUThreadCount; the reference runs the adjoint phase on one
    thread either way), 2 warm-ups, then at least 10 evaluations per thread
    count: median and best.  Rates are scaled to 50,000 patterns (the work is
    linear in the pattern count).  Falls back to the oracle (one thread)
    without the reference build.  Runs in its own process (both libraries
    export beagle*)."""
    import numpy as np
    from beagle_amd import api, driver
    ref_path = os.path.join(REPO, "oracle", "_ref", "libhmsbeagle.so")
    have_ref = os.path.exists(ref_path)
    cores = host_cores() if have_ref else 1

    def make(patterns, threads):
        prob = adjoint_problem(patterns)
        if have_ref:
            lib = api.Library(ref_path, "beagle")
            flags = api.FLAG_PROCESSOR_CPU | api.FLAG_VECTOR_SSE | api.FLAG_THREADING_CPP
            ev = driver.Evaluator(lib, prob, extra_flags=flags, transpose="none",
                                  scaling="manual")
            lib.call("SetCPUThreadCount", ev.inst.h, threads)
            return ev
        lib = api.Library(os.path.join(REPO, "oracle", "liboracle.so"), "orc")
        return driver.Evaluator(lib, prob, transpose="none", scaling="manual")

    def phases(ev):
        t0 = time.perf_counter()
        ev.log_likelihood()
        t1 = time.perf_counter()
        ev.pre_order()
        ev.edge_derivatives(False, per_site=False)
        ev.edge_derivatives(True, per_site=False)
        ev.cross_products()
        return t1 - t0, time.perf_counter() - t1

    # size the sample from one small single-thread evaluation: the
    # single-thread run takes about half the budget
    ev = make(500, 1)
    phases(ev)
    per_pattern = sum(phases(ev)) / 500
    ev.close()
    sample = int(max(500, min(ADJ["patterns"], budget_s / 2 / 12 / per_pattern)) // 100 * 100)
    scale = sample / ADJ["patterns"]
    runs = {}
    for threads in sorted({1, cores}):
        ev = make(sample, threads)
        for _ in range(2):
            phases(ev)
        t = [phases(ev) for _ in range(10)]
        ev.close()
        tot = [a + b for a, b in t]
        runs[threads] = {
            "threads": threads,
            "evals_per_s_median": scale / float(np.median(tot)),
            "evals_per_s_best": scale / float(np.min(tot)),
            "postorder_root_ms_median": float(np.median([a for a, _ in t])) / scale * 1e3,
            "adjoint_ms_median": float(np.median([b for _, b in t])) / scale * 1e3}
    best = runs[cores]
    return {"value": best["evals_per_s_median"], "unit": "evals/s", "cores": cores,
            "kind": "reference" if have_ref else "port",
            "sample": f"the adjoint workload's generators at {sample} patterns (512-taxon Yule "
                      f"tree), {'CPU-4State-SSE-Double' if have_ref else 'the oracle'}, FP64 with "
                      f"manual scale buffers; 2 warm-ups + 10 timed evaluations per thread "
                      f"count; rates scaled by {sample}/{ADJ['patterns']}",
            "cpu_model": cpu_model(), "host_cpus_in_affinity": len(os.sched_getaffinity(0)),
            "one_thread": runs[1], "all_cores": best}


def cpu_baseline_plain(budget_s):
    """The plain FP64 pruning workload (16 taxa, 10,000 patterns, root logL
    only) on the reference CPU plugin, 1 thread and all cores: median
    evaluations/s over about budget_s seconds each."""
    import numpy as np
    from beagle_amd import api, driver, synth
    ref_path = os.path.join(REPO, "oracle", "_ref", "libhmsbeagle.so")
    if not os.path.exists(ref_path):
        return None
    prob = synth.make_problem(SEED, PLAIN["taxa"], PLAIN["patterns"], S=4, shape=PLAIN["shape"])
    out = {}
    for threads in sorted({1, host_cores()}):
        lib = api.Library(ref_path, "beagle")
        flags = api.FLAG_PROCESSOR_CPU | api.FLAG_VECTOR_SSE | api.FLAG_THREADING_CPP
        ev = driver.Evaluator(lib, prob, extra_flags=flags, transpose="none", preorder=False)
        lib.call("SetCPUThreadCount", ev.inst.h, threads)
        for _ in range(3):
            ev.log_likelihood()
        t, t_end = [], time.perf_counter() + budget_s
        while time.perf_counter() < t_end or len(t) < 10:
            t0 = time.perf_counter()
            ev.log_likelihood()
            t.append(time.perf_counter() - t0)
        ev.close()
        out[str(threads)] = {"evals_per_s_median": 1.0 / float(np.median(t)),
                             "evals_per_s_best": 1.0 / float(np.min(t))}
    return out


CODON = dict(S=61, taxa=64, patterns=20_000, shape="yule")
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}


def codon_problem(patterns=None):
    from beagle_amd import synth
    return synth.make_problem(SEED, CODON["taxa"], patterns or CODON["patterns"], S=CODON["S"],
                              shape=CODON["shape"])


def postorder_flops(prob):
    """Algorithmic flops of one post-order traversal (synthetictest's
    convention, examples/synthetictest/synthetictest.cpp:2567-2580): per
    operation C*P*S*f with f = 4S-1 (two partials children), 2S (one tip),
    1 (two tips)."""
    C, P, S = prob.C, prob.P, prob.S
    n = prob.tree.n
    total = 0
    for v in prob.tree.postorder():
        tips = sum(1 for x in prob.tree.children[v] if x < n)
        total += C * P * S * (4 * S - 1 if tips == 0 else 2 * S if tips == 1 else 1)
    return total


def codon_bench(lib, steps):
    """61-state GY94 codon model, 64 taxa, 20,000 patterns, Gamma(4): one
    evaluation = transition matrices + post-order + root log-likelihood.
    FP64 (v_mfma_f64_16x16x4f64), FP32 (v_mfma_f32_32x32x2f32) and split BF16
    (v_mfma_f32_32x32x16_bf16, three products per hi/lo pair) post-order
    kernels, all with exact power-of-two rescaling; errors against FP64."""
    import numpy as np
    import torch
    from beagle_amd import api, driver
    prob = codon_problem()
    flops = postorder_flops(prob)
    out, ref = {}, None
    for name, prec, terms, peak in (("fp64", "double", 0, "f64"), ("fp32_mfma", "single", 0, "f32"),
                                    ("split_bf16x3", "single", 3, "bf16")):
        ev = driver.Evaluator(lib, prob, precision=prec, preorder=False, scaling="auto",
                              extra_flags=api.FLAG_COMPUTATION_ASYNCH)
        if terms:
            ev.inst.set_split_bf16(terms)
        st = torch.cuda.ExternalStream(ev.inst.stream())
        for _ in range(3):
            ev.log_likelihood()
        n = max(steps, 10)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            logl = ev.log_likelihood()
        torch.cuda.synchronize()
        rate = n / (time.perf_counter() - t0)
        pairs = []
        for _ in range(n):
            ev.update_matrices()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            ev.post_order()
            e1.record(st)
            ev.root()
            pairs.append((e0, e1))
        torch.cuda.synchronize()
        ms = float(np.mean([a.elapsed_time(b) for a, b in pairs]))
        ev.close()
        # logL + dlogL/dt and d2logL/dt2 of every branch (matrix-core
        # pre-order and derivative passes; the post-order on the path's engine)
        ev = driver.Evaluator(lib, prob, precision=prec, scaling="auto",
                              extra_flags=api.FLAG_COMPUTATION_ASYNCH)
        if terms:
            ev.inst.set_split_bf16(terms)
        grad = ev.fused_gradient()
        ng = max(steps // 3, 5)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(ng):
            grad = ev.fused_gradient()
        torch.cuda.synchronize()
        grad_rate = ng / (time.perf_counter() - t0)
        # exact dlogL/dQ and dlogL/dpi_root from the same partials (one
        # branch per pre-order operation for S > 4; synchronous call)
        G, _ = ev.rate_matrix_gradient()
        t0 = time.perf_counter()
        for _ in range(3):
            G, _ = ev.rate_matrix_gradient()
        qgrad_ms = (time.perf_counter() - t0) / 3 * 1e3
        ev.close()
        if name == "fp64":
            grad_ref = grad
            G_ref = G
        w = 8 if prec == "double" else 4
        gbs = postorder_bytes(prob, w, True) / (ms * 1e-3) / 1e9
        tf = flops / (ms * 1e-3) / 1e12
        if ref is None:
            ref = logl
        out[name] = {"evals_per_s": rate, "logL": logl, "logL_rel_err_vs_fp64": rel(logl, ref),
                     "gradient_evals_per_s": grad_rate,
                     "gradient_err_vs_fp64": max(grad_err(grad[1], grad_ref[1]),
                                                 grad_err(grad[2], grad_ref[2])),
                     "rate_matrix_gradient_ms": qgrad_ms,
                     "rate_matrix_gradient_err_vs_fp64":
                         float(np.abs(G - G_ref).max() / np.abs(G_ref).max()),
                     "postorder_ms": ms, "partials_GBps": gbs,
                     "partials_frac_of_hbm_peak": gbs / HBM_PEAK_GBS,
                     "algorithmic_TFLOPs": tf, "mfma_peak_TFLOPs": PEAK_TFLOPS[peak],
                     "frac_of_mfma_peak": tf / PEAK_TFLOPS[peak]}
    for name in ("fp32_mfma", "split_bf16x3"):
        out[name]["speedup_vs_fp64"] = out[name]["evals_per_s"] / out["fp64"]["evals_per_s"]
        out[name]["gradient_speedup_vs_fp64"] = (out[name]["gradient_evals_per_s"] /
                                                 out["fp64"]["gradient_evals_per_s"])
    out["workload"] = ("61-state GY94 codon, 64 taxa (Yule), 20,000 patterns, Gamma(4); "
                       "matrices + post-order + root logL per evaluation; algorithmic flops "
                       f"{flops:.3e} per traversal")
    return out


PARTS = dict(count=48, pmin=200, pmax=50_000, taxa=(32, 256), states=(4, 20, 61))


def partition_problems():
    """48 partitions: state counts cycling 4 (GTR), 20 (exchangeabilities
    ~Exp(1), pi ~Dirichlet(1)) and 61 (GY94 codon); patterns ~LogUniform(200,
    50,000); Yule trees of 32-256 taxa; Gamma(4)."""
    import numpy as np
    from beagle_amd import synth
    rng = np.random.default_rng(SEED + 100)
    out = []
    for k in range(PARTS["count"]):
        S = PARTS["states"][k % 3]
        P = int(round(np.exp(rng.uniform(np.log(PARTS["pmin"]), np.log(PARTS["pmax"])))))
        taxa = int(rng.integers(PARTS["taxa"][0], PARTS["taxa"][1] + 1))
        out.append(synth.make_problem(SEED + 1000 + k, taxa, P, S=S, shape="yule"))
    return out


def partitions_bench(lib, rounds):
    """One round = FP32 logL + every branch's dlogL/dt and d2logL/dt2 for all
    48 partitions (one instance each, matrix-core kernels for 20 and 61
    states).  Eager: the extension calls instance after instance (each
    waits for its results).  Graphs: each instance's evaluation captured
    once as a HIP graph; a round launches all 48 replays (one stream per
    instance, so partitions overlap) and then collects them."""
    import numpy as np
    import torch
    from beagle_amd import api, driver
    probs = partition_problems()
    evs = [driver.Evaluator(lib, p, precision="single", scaling="auto",
                            extra_flags=api.FLAG_COMPUTATION_ASYNCH) for p in probs]
    for ev in evs:
        ev.fused_gradient()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(rounds):
        eager = [ev.fused_gradient() for ev in evs]
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / rounds
    for ev in evs:
        ev.capture_fused_gradient()
    for ev in evs:
        ev.replay_fused_gradient(wait=False)
    for ev in evs:
        ev.collect_fused_gradient()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(rounds):
        for ev in evs:
            ev.replay_fused_gradient(wait=False)
        graph = [ev.collect_fused_gradient() for ev in evs]
    t_graph = (time.perf_counter() - t0) / rounds
    same = all(a[0] == b[0] and np.array_equal(a[1], b[1]) for a, b in zip(eager, graph))
    for ev in evs:
        ev.close()
    n = len(probs)
    return {"workload": "48 partitions: 4-, 20- and 61-state models in turn, patterns "
                        "~LogUniform(200, 50000), Yule trees of 32-256 taxa, Gamma(4); FP32 "
                        "logL + per-branch first and second derivatives per partition",
            "total_patterns": int(sum(p.P for p in probs)),
            "patterns_by_states": {str(S): int(sum(p.P for p in probs if p.S == S))
                                   for S in PARTS["states"]},
            "eager_ms_per_round": t_eager * 1e3, "graph_ms_per_round": t_graph * 1e3,
            "eager_partition_evals_per_s": n / t_eager,
            "graph_partition_evals_per_s": n / t_graph,
            "graph_speedup": t_eager / t_graph,
            "graph_results_identical_to_eager": bool(same), "rounds": rounds}


def by_partition_bench(lib, rounds):
    """The 48 partitions' state and pattern counts on one shared 128-taxon
    Yule tree, each partition with its own model (GTR, exchangeabilities
    ~Exp(1) and pi ~Dirichlet(1) for 20 states, GY94), Gamma(4), FP32.
    By partition: one instance per state count holding its 16 partitions as
    pattern ranges; a round = the partitions' transition matrices, one
    beagleUpdatePartialsByPartition over every partition's operations (one
    launch per dependency level, a work list of (operation, pattern tile)
    items), one beagleUpdatePrePartialsByPartition, then root
    log-likelihoods by partition.  Against one instance per partition
    making the same calls, and the by-partition instances with the work
    lists switched off (grid of every tile x every operation)."""
    import dataclasses
    import os
    import numpy as np
    import torch
    from beagle_amd import api, driver, synth
    rng = np.random.default_rng(SEED + 200)
    tree = synth.make_tree(rng, 128, "yule", 0.1)
    probs = []
    for k in range(PARTS["count"]):
        S = PARTS["states"][k % 3]
        P = int(round(np.exp(rng.uniform(np.log(PARTS["pmin"]), np.log(PARTS["pmax"])))))
        probs.append(dataclasses.replace(synth.make_problem(SEED + 3000 + k, 128, P, S=S), tree=tree))
    n = tree.n
    nodes, nb = 2 * n - 1, 2 * n - 2
    post = np.asarray(synth.postorder_ops(tree), np.int32)
    flags = api.FLAG_EIGEN_REAL | api.FLAG_SCALING_AUTO | api.FLAG_PREORDER_TRANSPOSE_AUTO | \
        api.FLAG_COMPUTATION_ASYNCH

    def build(group):
        K = len(group)
        S = group[0].S
        sizes = [g.P for g in group]
        inst = lib.create(n, (n - 1) + nodes, n, S, sum(sizes), K, K * nb, group[0].C, n,
                          preference=flags, requirement=api.FLAG_PRECISION_SINGLE)
        for t in range(n):
            inst.set_tip_states(t, np.concatenate([g.tips[t] for g in group]))
        for k, g in enumerate(group):
            m = g.model
            inst.set_eigen(k, m.evec.reshape(-1), m.ivec.reshape(-1), m.evals)
            inst.set_state_frequencies(k, m.freqs)
        inst.set_category_weights(0, group[0].weights)
        inst.set_category_rates(group[0].rates)
        inst.set_pattern_weights(np.concatenate([g.pattern_weights for g in group]))
        inst.set_pattern_partitions(K, np.repeat(np.arange(K), sizes).astype(np.int32))
        pops, qops = [], []
        pre = np.asarray(synth.preorder_ops(tree, nodes), np.int32)
        for k in range(K):
            a = post.copy()
            a[:, [4, 6]] += k * nb
            pops.append(np.hstack([a, np.full((len(a), 1), k), np.full((len(a), 1), -1)]))
            b = pre.copy()
            b[:, [4, 6]] += k * nb
            qops.append(np.hstack([b, np.full((len(b), 1), k), np.full((len(b), 1), -1)]))
        # root pre-order partials: each partition's frequencies over its range
        root_pre = np.concatenate([np.tile(g.model.freqs, g.P) for g in group])
        inst.set_partials(nodes + tree.root, np.tile(root_pre, group[0].C))
        return inst, np.vstack(pops).astype(np.int32), np.vstack(qops).astype(np.int32)

    def round_by_partition(insts):
        for inst, group, pops, qops in insts:
            for k, g in enumerate(group):
                inst.update_transition_matrices(k, np.arange(nb, dtype=np.int32) + k * nb, None, None,
                                                tree.lengths[:nb])
            inst.update_partials_by_partition(pops)
            inst.update_pre_partials_by_partition(qops)
        out = []
        for inst, group, pops, qops in insts:
            K = len(group)
            by, _ = inst.root_log_likelihood_by_partition([tree.root] * K, [0] * K, list(range(K)),
                                                          [-1] * K, list(range(K)))
            out.append(by)
        return np.concatenate(out)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, r

    res = {}
    for lists in ("1", "0"):
        os.environ["BEAGLE_AMD_TILE_LISTS"] = lists
        insts = []
        for S in PARTS["states"]:
            group = [g for g in probs if g.S == S]
            inst, pops, qops = build(group)
            insts.append((inst, group, pops, qops))
        t, by = timed(lambda: round_by_partition(insts), rounds)
        res[lists] = (t, by)
        for inst, *_ in insts:
            inst.finalize()
    os.environ.pop("BEAGLE_AMD_TILE_LISTS", None)
    evs = [driver.Evaluator(lib, g, precision="single", scaling="auto",
                            extra_flags=api.FLAG_COMPUTATION_ASYNCH) for g in probs]
    order = [i for S in PARTS["states"] for i, g in enumerate(probs) if g.S == S]

    def round_per_instance():
        for i in order:
            evs[i].update_matrices()
            evs[i].post_order()
            evs[i].pre_order()
        return np.array([evs[i].root() for i in order])

    t_inst, by_inst = timed(round_per_instance, rounds)
    for ev in evs:
        ev.close()
    t_list, by_list = res["1"]
    t_grid, by_grid = res["0"]
    return {"workload": "48 partitions (state and pattern counts as the partitions workload) on one "
                        "shared 128-taxon Yule tree, own model each, Gamma(4), FP32: matrices, "
                        "post-order, pre-order and root logL by partition",
            "total_patterns": int(sum(g.P for g in probs)),
            "by_partition_ms_per_round": t_list * 1e3,
            "by_partition_full_grid_ms_per_round": t_grid * 1e3,
            "instance_per_partition_ms_per_round": t_inst * 1e3,
            "work_list_speedup_vs_full_grid": t_grid / t_list,
            "by_partition_speedup_vs_instance_per_partition": t_inst / t_list,
            "logl_identical_with_and_without_lists": bool(np.array_equal(by_list, by_grid)),
            "max_rel_diff_vs_instance_per_partition": float(np.max(np.abs(by_list - by_inst) /
                                                                   np.abs(by_inst))),
            "rounds": rounds}


def cpu_baseline_codon(budget_s):
    """The reference CPU plugin (threaded SSE, double) on the codon workload:
    log-likelihood evaluations (matrices, post-order, root) of a pattern
    sample, scaled to 20,000 patterns."""
    from beagle_amd import api, driver
    ref_path = os.path.join(REPO, "oracle", "_ref", "libhmsbeagle.so")
    if not os.path.exists(ref_path):
        return None
    cores = host_cores()
    sample = 1000
    lib = api.Library(ref_path, "beagle")
    flags = api.FLAG_PROCESSOR_CPU | api.FLAG_VECTOR_SSE | api.FLAG_THREADING_CPP
    ev = driver.Evaluator(lib, codon_problem(sample), preorder=False, scaling="manual",
                          extra_flags=flags, transpose="none")
    lib.call("SetCPUThreadCount", ev.inst.h, cores)
    ev.log_likelihood()
    t0, n = time.perf_counter(), 0
    while n < 2 or time.perf_counter() - t0 < budget_s:
        ev.log_likelihood()
        n += 1
    dt = time.perf_counter() - t0
    name = ev.inst.impl_name
    ev.close()
    return {"value": n / dt * sample / CODON["patterns"], "unit": "evals/s", "cores": cores,
            "kind": "reference",
            "sample": f"{n} FP64 logL evaluations of the 64-taxon codon tree on a {sample}-pattern "
                      f"alignment from the same generators, {name} ({dt:.1f} s), rate scaled by "
                      f"{sample}/20000"}


def rel(a, b):
    return abs(a - b) / abs(b)


def grad_err(a, b):
    import numpy as np
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def checkpoint_bench(lib, probs, steps):
    """FP32 logL + gradient of each benchmark tree through
    beagleExtCheckpointedGradient with the smallest pool of partials
    buffers any checkpoint interval fits (found with its planning call),
    against the full fused evaluation (every post- and pre-order partials
    buffer of the internal nodes held): buffers, bytes, evaluations/s and
    whether the results are identical."""
    import numpy as np
    import torch
    from beagle_amd import api, driver
    out = {}
    for prob, shape in zip(probs, SHAPES):
        ev = driver.Evaluator(lib, prob, precision="single", scaling="auto")
        n = prob.tree.n
        pool = list(range(ev.pre_base, ev.pre_base + 2 * n - 1))
        m = None
        for size in range(4, len(pool) + 1):
            try:
                ev.checkpointed_gradient(pool[:size], 0, plan_only=True)
            except api.BeagleError:
                continue
            m = size
            break
        full = ev.fused_gradient()
        ck = ev.checkpointed_gradient(pool[:m], 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ck = ev.checkpointed_gradient(pool[:m], 0)
        rate = steps / (time.perf_counter() - t0)
        t0 = time.perf_counter()
        for _ in range(steps):
            full = ev.fused_gradient()
        full_rate = steps / (time.perf_counter() - t0)
        buf = prob.C * prob.P * prob.S * 4 + 2 * prob.P
        full_bufs = 2 * (n - 1)    # internal post-order + internal pre-order partials
        out[shape] = {
            "pool_buffers": m, "interval": ck[3], "peak_buffers": ck[4],
            "peak_partials_MB": ck[4] * buf / 1e6,
            "full_evaluation_buffers": full_bufs, "full_partials_MB": full_bufs * buf / 1e6,
            "evals_per_s": rate, "full_evals_per_s": full_rate,
            "identical_to_full": bool(ck[0] == full[0] and np.array_equal(ck[1], full[1])
                                      and np.array_equal(ck[2], full[2]))}
        ev.close()
    return out


def plain_post_timing(lib, prob, steps):
    """The plain path's FP64 post-order traversal of the caterpillar (one
    fused launch of 127 operations, no scaling) bracketed by HIP events on
    the library's stream.  Returns (mean ms, algorithmic GB/s)."""
    import numpy as np
    import torch
    from beagle_amd import api, driver
    ev = driver.Evaluator(lib, prob, preorder=False, extra_flags=api.FLAG_COMPUTATION_ASYNCH)
    st = torch.cuda.ExternalStream(ev.inst.stream())
    for _ in range(3):
        ev.log_likelihood()
    pairs = []
    for _ in range(max(steps // 2, 5)):
        ev.update_matrices()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        ev.post_order()
        e1.record(st)
        ev.root()
        pairs.append((e0, e1))
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in pairs]))
    ev.close()
    return ms, postorder_bytes(prob, 8) / (ms * 1e-3) / 1e9


def calibrate_copy():
    """A plain 1 GiB -> 1 GiB elementwise copy kernel, the reference size
    for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings."""
    import torch
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    del src, dst


def roofline_timing(ev, stream, prob, steps):
    """The dominant kernel of the value path: the one-call gradient's
    node pass (pre-order partials, branch derivatives and rate-matrix sums of
    all 511 node jobs: for this tree one launch of the chained node kernel
    k_pren4f for the narrow top levels, 13 launches of k_pren4f_level, one
    per dependency level, and one k_pren4f launch for the narrow bottom
    levels, as the tuner picks them).  The call (beagleExtCalculateLogLikelihoodAndGradient)
    is captured alone as a HIP graph and each replay is bracketed by HIP
    events on the library's stream, back to back, so no host work falls
    inside the bracket; the root, the reductions and the result copy that
    follow the pass (about 4% of the bracket) are inside it.  Returns (mean
    ms per pass, algorithmic GB/s, algorithmic bytes per pass:
    preorder_bytes() plus one read of the root partials)."""
    import numpy as np
    import torch
    ops, post, d1, d2 = ev.fused_operations()
    post_idx = dict(zip(map(tuple, ops.tolist()), post.tolist()))
    pre_bytes = preorder_bytes(prob, ops, lambda o: post_idx[tuple(o.tolist())], 4, True)
    pre_bytes += prob.C * prob.P * prob.S * 4 + SHORT * prob.P   # the root's read
    root = prob.tree.root
    n, S = len(post), prob.S
    outs = (np.zeros(1), np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(S * S), np.zeros(S))
    ev.fused_gradient(rate_matrix=True)
    ev.inst.begin_capture()
    try:
        ev.inst.log_likelihood_and_gradient(root, ev.pre_base + root, ops, post, d1, d2,
                                            out=outs)
    finally:
        graph = ev.inst.end_capture()
    for _ in range(3):
        ev.inst.launch_graph(graph)
    ev.inst.wait_graph(graph)
    pairs = []
    for _ in range(max(steps, 10)):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ev.inst.launch_graph(graph)
        e1.record(stream)
        pairs.append((e0, e1))
    ev.inst.wait_graph(graph)
    torch.cuda.synchronize()
    ev.inst.destroy_graph(graph)
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in pairs]))
    return kernel_ms, pre_bytes / (kernel_ms * 1e-3) / 1e9, pre_bytes


def tree128_bench(lib, steps, warmup):
    """The 128-taxon, 100,000-pattern workload (caterpillar and balanced
    trees): logL + dlogL/dt and d2logL/dt2 of all 254 branches per
    evaluation.  FP32: each tree's evaluation replayed as a HIP graph, the
    two trees as independent chains (a tree's next evaluation launched once
    its result is collected, while the other tree's runs); the same
    evaluations as eager calls; the plain FP64 beagle.h calls and the FP64
    fused extension; errors against the plain path."""
    import numpy as np
    import torch
    from beagle_amd import api, driver
    probs = problems()
    asyn = api.FLAG_COMPUTATION_ASYNCH
    evs = [driver.Evaluator(lib, p, precision="single", scaling="auto", extra_flags=asyn)
           for p in probs]
    t0, w = time.perf_counter(), 0
    while w < warmup or time.perf_counter() - t0 < 1.0:
        for ev in evs:
            fused_eval(ev)
        w += 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eager32 = [fused_eval(evs[0]), fused_eval(evs[1])]
    torch.cuda.synchronize()
    eager_rate = 2 * steps / (time.perf_counter() - t0)
    lengths = [ev.prob.tree.lengths for ev in evs]
    for ev in evs:
        ev.capture_fused_gradient()

    def chains(n):
        for ev, x in zip(evs, lengths):
            ev.replay_fused_gradient(x, wait=False)
        out = None
        for step in range(n):
            out = []
            for ev, x in zip(evs, lengths):
                out.append(ev.collect_fused_gradient())
                if step + 1 < n:
                    ev.replay_fused_gradient(x, wait=False)
        return out

    chains(max(warmup, 3))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res32 = chains(steps)
    torch.cuda.synchronize()
    value = 2 * steps / (time.perf_counter() - t0)
    identical = all(a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
                    for a, b in zip(res32, eager32))
    for ev in evs:
        ev.close()
    n64 = max(steps // 3, 3)
    rates, res64 = {}, {}
    for name in ("api", "fused"):
        evs = [driver.Evaluator(lib, p, extra_flags=asyn) for p in probs]
        fn = (lambda ev: ev.gradient()) if name == "api" else fused_eval
        for _ in range(2):
            for ev in evs:
                fn(ev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n64):
            res64[name] = [fn(ev) for ev in evs]
        torch.cuda.synchronize()
        rates[name] = 2 * n64 / (time.perf_counter() - t0)
        for ev in evs:
            ev.close()
    api64 = res64["api"]
    err = {"logL": max(rel(a[0], b[0]) for a, b in zip(res32, api64)),
           "dlogL_dt": max(grad_err(a[1], b[1]) for a, b in zip(res32, api64)),
           "d2logL_dt2": max(grad_err(a[2], b[2]) for a, b in zip(res32, api64))}
    post_ms, post_gbs = plain_post_timing(lib, probs[0], steps)
    return {"workload": "4-state GTR+G4, 128 taxa, 100,000 patterns, caterpillar + balanced "
                        "trees; logL + dlogL/dt and d2logL/dt2 of all 254 branches",
            "evals_per_s": value, "eager_evals_per_s": eager_rate,
            "graph_results_identical_to_eager": bool(identical), "error_vs_fp64": err,
            "plain_fp64": {"api_evals_per_s": rates["api"],
                           "what": "beagle.h calls: update matrices, post-order, root, "
                                   "beagleUpdatePrePartials, two beagleCalculateEdgeDerivatives",
                           "speedup_of_fp32": value / rates["api"],
                           "fused_evals_per_s": rates["fused"],
                           "postorder_GBps": post_gbs,
                           "postorder_frac_of_hbm_peak": post_gbs / HBM_PEAK_GBS,
                           "postorder_ms": post_ms,
                           "postorder_kernel": "k_post4<double,false>: caterpillar, 127 "
                                               "operations, one launch"},
            "logL": [r[0] for r in api64]}


def plain_workload_bench(lib, steps):
    """The plain FP64 pruning workload (16 taxa, 10,000 patterns, root logL
    only): the beagle.h calls (matrices, post-order, root) one evaluation
    after the other, eagerly and replayed as a HIP graph."""
    import numpy as np
    import torch
    from beagle_amd import api, driver, synth
    prob = synth.make_problem(SEED, PLAIN["taxa"], PLAIN["patterns"], S=4, shape=PLAIN["shape"])
    ev = driver.Evaluator(lib, prob, preorder=False)
    for _ in range(5):
        lnl = ev.log_likelihood()
    n = max(steps * 5, 50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        lnl = ev.log_likelihood()
    eager = n / (time.perf_counter() - t0)
    res = np.zeros(1)
    ev.inst.begin_capture()
    ev.update_matrices()
    ev.post_order()
    ev.inst.root_log_likelihood(prob.tree.root, 0, 0, api.OP_NONE, out=res)
    g = ev.inst.end_capture()
    for _ in range(5):
        ev.inst.launch_graph(g)
        ev.inst.wait_graph(g)
    t0 = time.perf_counter()
    for _ in range(n):
        ev.inst.set_graph_branch_lengths(g, prob.tree.lengths[:2 * prob.tree.n - 2])
        ev.inst.launch_graph(g)
        ev.inst.wait_graph(g)
    graph = n / (time.perf_counter() - t0)
    same = float(res[0]) == lnl
    ev.close()
    return {"workload": "4-state GTR+G4, 16 taxa (Yule), 10,000 patterns, FP64, root logL only",
            "eager_evals_per_s": eager, "graph_evals_per_s": graph, "logL": lnl,
            "graph_identical_to_eager": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-only", action="store_true")
    ap.add_argument("--only", choices=("headline", "tree128", "codon", "partitions", "by_partition", "roofline",
                                       "plain-post"),
                    help="run one workload alone (for profiling): the metric's adjoint "
                         "workload, the 128-taxon workload, the codon or partition workload, "
                         "the roofline timing of the dominant kernel, or the plain path's "
                         "FP64 post-order traversal")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB -> 1 GiB elementwise copy kernel, the "
                         "reference size for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()

    if args.cpu_baseline_only:
        res = cpu_baseline(args.cpu_budget)
        res["plain_workload"] = cpu_baseline_plain(1.0)
        res["codon"] = cpu_baseline_codon(min(4.0, args.cpu_budget / 3))
        print(json.dumps(res))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(patterns shard across GPUs with one all-reduce of the per-pattern sums "
              "per evaluation)", file=sys.stderr)
        sys.exit(2)

    import torch
    import beagle_amd
    from beagle_amd import api, driver

    assert torch.cuda.is_available(), "needs an MI355X"
    lib = beagle_amd.load()
    if args.only == "headline":
        print(json.dumps(full_gradient_bench(lib, args.steps, args.warmup)))
        return
    if args.only == "tree128":
        print(json.dumps(tree128_bench(lib, args.steps, args.warmup)))
        return
    if args.only == "codon":
        print(json.dumps({"codon": codon_bench(lib, args.steps)}))
        return
    if args.only == "partitions":
        print(json.dumps({"partitions": partitions_bench(lib, max(args.steps // 10, 3))}))
        return
    if args.only == "by_partition":
        print(json.dumps({"by_partition": by_partition_bench(lib, max(args.steps // 10, 3))}))
        return
    if args.only == "plain-post":
        probs = problems()
        ms, gbs = plain_post_timing(lib, probs[0], args.steps)
        if args.calibrate_copy:
            calibrate_copy()
        print(json.dumps({"plain_post": {"ms": ms, "GBps": gbs,
                                         "algorithmic_bytes": postorder_bytes(probs[0], 8)}}))
        return
    asyn = api.FLAG_COMPUTATION_ASYNCH
    if args.only == "roofline":
        prob = adjoint_problem()
        ev = driver.Evaluator(lib, prob, precision="single", scaling="auto", extra_flags=asyn)
        for _ in range(3):
            fused_eval(ev)
        kernel_ms, achieved, pre_bytes = roofline_timing(
            ev, torch.cuda.ExternalStream(ev.inst.stream()), prob, args.steps)
        ev.close()
        if args.calibrate_copy:
            calibrate_copy()
        print(json.dumps({"roofline": {"kernel_ms": kernel_ms, "achieved": achieved,
                                       "algorithmic_bytes": pre_bytes}}))
        return

    head = full_gradient_bench(lib, args.steps, args.warmup)
    prob = adjoint_problem()
    ev = driver.Evaluator(lib, prob, precision="single", scaling="auto", extra_flags=asyn)
    for _ in range(3):
        fused_eval(ev)
    kernel_ms, achieved, pre_bytes = roofline_timing(
        ev, torch.cuda.ExternalStream(ev.inst.stream()), prob, args.steps)
    ev.close()
    tree128 = tree128_bench(lib, args.steps, args.warmup)
    plain16 = plain_workload_bench(lib, args.steps)
    checkpointed = checkpoint_bench(lib, problems(), max(args.steps // 3, 3))
    codon = codon_bench(lib, args.steps)
    partitions = partitions_bench(lib, max(args.steps // 10, 3))
    by_partition = by_partition_bench(lib, max(args.steps // 10, 3))
    if args.calibrate_copy:
        calibrate_copy()

    cpu = None
    if not args.no_cpu_baseline:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-only",
                            "--cpu-budget", str(args.cpu_budget)], capture_output=True, text=True,
                           timeout=900)
        if r.returncode == 0:
            cpu = json.loads(r.stdout.strip().splitlines()[-1])
        else:
            print(r.stderr[-2000:], file=sys.stderr)
    if cpu:
        if cpu.get("codon"):
            codon["cpu_reference"] = cpu.pop("codon")
        if cpu.get("plain_workload"):
            plain16["cpu_reference"] = cpu.pop("plain_workload")

    traffic = None
    tpath = os.path.join(REPO, "profiles", "pren4f_fp32_adjoint_512taxa_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")

    value = head.pop("value")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_eval"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": "adjoint: 4-state GTR+G4, 512-taxon Yule tree, 50,000 patterns; "
                               "one evaluation = logL + dlogL/dt and d2logL/dt2 of all 1,022 "
                               "branches + dlogL/dQ + dlogL/dpi_root",
                   "taxa": ADJ["taxa"], "patterns": ADJ["patterns"], "categories": 4,
                   "steps_are": "evaluations, one after the other",
                   "path": "FP32 partials, power-of-two rescaling, fused pre-order + "
                           "derivative extension and rate-matrix gradient; the evaluation "
                           "replayed as one HIP graph with its branch lengths set per replay"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_pren4f_level<4,true,true>: the one-call gradient's node "
                               "pass of the 512-taxon tree (511 node jobs = 1,022 branches; "
                               "pre-order partials, branch derivatives, rate-matrix sums): "
                               "13 level launches plus two chained k_pren4f launches for "
                               "the narrow top and bottom levels per pass; kernel_ms is "
                               "per pass, bracketing one graph replay of the call "
                               "(root, reductions and result copy included)",
                     "kernel_ms": kernel_ms, "algorithmic_bytes": pre_bytes},
        "cpu_baseline": cpu,
        "adjoint": head,
        "tree128": tree128,
        "plain_workload": plain16,
        "checkpointed": checkpointed,
        "codon": codon,
        "partitions": partitions,
        "by_partition": by_partition,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
