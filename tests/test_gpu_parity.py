"""HIP plugin vs the FP64 oracle, through the C ABI, on cuda:0.

Each case builds the same synthetic problem, runs the same call sequence
(driver.Evaluator) on both libraries and compares: total and per-site
log-likelihoods, partials buffers, scale buffers, per-branch first and
second derivatives (sums and per site) and cross products.

Two matrix sources:
- "explicit": identical P(t) set on both sides with
  beagleSetTransitionMatrices, so the traversal kernels are compared on
  identical inputs: partials, scale factors, derivatives and cross
  products must agree to ~1e-12 (FMA contraction and the device's tree
  reductions are the only differences);
- "eigen": P(t) computed by each library from the eigen-system.  The
  device exp() and the host libm may differ in the last ulp, which is
  ~1e-16 absolute in P but up to ~1e-10 relative in P entries near zero
  (and more for 61-state codon entries of order 1e-15), so here only the
  quantities the accuracy targets bound are compared: log-likelihood within
  1e-10 relative, site log-likelihoods, derivative sums.
"""
import numpy as np
import pytest

from beagle_amd import api, driver, synth

import sequences

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def close(a, b, rtol=RTOL, atol=0.0):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    assert a.shape == b.shape
    err = np.abs(a - b)
    lim = atol + rtol * np.abs(b)
    bad = err > lim
    assert not bad.any(), f"max err {err.max():.3e}, worst rel {rel(a, b):.3e}"


def close_rows(a, b, S, rtol=1e-11):
    """Partials: each (category, pattern) vector compared relative to its own
    largest entry (P(t) entries near zero carry ~1e-16 absolute rounding
    from the eigen-sum, so element-wise relative error is not meaningful
    for entries far below the vector's maximum)."""
    a = np.asarray(a, float).reshape(-1, S)
    b = np.asarray(b, float).reshape(-1, S)
    scale = np.max(np.abs(b), axis=1, keepdims=True)
    err = np.abs(a - b)
    bad = err > rtol * scale
    assert not bad.any(), f"worst row-relative err {np.max(err / np.maximum(scale, 1e-300)):.3e}"


def run_both(amd, oracle, prob, **kw):
    res = []
    for lib in (amd, oracle):
        ev = driver.Evaluator(lib, prob, **kw)
        out = {"explicit": kw.get("matrices") == "explicit"}
        out["logl"] = ev.log_likelihood()
        out["sites"] = ev.inst.site_log_likelihoods()
        out["root"] = ev.inst.get_partials(prob.tree.root)
        if kw.get("scaling", "none") in ("manual", "manual_log"):
            out["cum"] = ev.inst.get_scale_factors(ev.n - 1)
            out["scale0"] = ev.inst.get_scale_factors(0)
        ev.pre_order()
        d1, s1, q1 = ev.edge_derivatives(False)
        d2, s2, q2 = ev.edge_derivatives(True)
        out.update(d1=d1, s1=s1, q1=q1, d2=d2, s2=s2)
        out["pre_tip0"] = ev.inst.get_partials(ev.pre_base + 0)
        out["cross"] = ev.cross_products()
        out["impl"] = ev.inst.impl_name
        ev.close()
        res.append(out)
    return res


CASES = [
    # S, taxa, patterns, shape, tips_as, scaling, missing
    (4, 16, 10000, "yule", "states", "none", 0.0),
    (4, 16, 1001, "yule", "states", "manual", 0.05),
    (4, 24, 777, "caterpillar", "partials", "manual_log", 0.1),
    (4, 12, 513, "balanced", "states", "always", 0.05),
    (20, 10, 300, "yule", "states", "manual", 0.05),
    (61, 6, 64, "yule", "states", "none", 0.05),
]


@pytest.mark.parametrize("S,n,P,shape,tips_as,scaling,missing", CASES)
def test_kernels_match_oracle_on_identical_matrices(amd, oracle, S, n, P, shape, tips_as,
                                                    scaling, missing):
    prob = synth.make_problem(11 + S + n, n, P, S=S, shape=shape, missing=missing)
    g, o = run_both(amd, oracle, prob, tips_as=tips_as, scaling=scaling, matrices="explicit")
    assert g["impl"] == "MI355X-HIP-Double"
    assert rel(g["logl"], o["logl"]) < 1e-12
    close(g["sites"], o["sites"], 1e-12, 1e-12)
    close_rows(g["root"], o["root"], S, 1e-12)
    if "cum" in o:
        close(g["cum"], o["cum"], 1e-12, 1e-12)
        close(g["scale0"], o["scale0"], 1e-12, 1e-13)
    close(g["s1"], o["s1"], 1e-10, 1e-10)
    close(g["q1"], o["q1"], 1e-10, 1e-10)
    close(g["s2"], o["s2"], 1e-10, 1e-10)
    close(g["d1"], o["d1"], 1e-10, 1e-10)
    close_rows(g["pre_tip0"], o["pre_tip0"], S, 1e-12)
    close(g["cross"], o["cross"], 1e-10, 1e-10)


@pytest.mark.parametrize("S,n,P,shape,tips_as,scaling,missing", CASES)
def test_eigen_path_matches_oracle(amd, oracle, S, n, P, shape, tips_as, scaling, missing):
    prob = synth.make_problem(11 + S + n, n, P, S=S, shape=shape, missing=missing)
    g, o = run_both(amd, oracle, prob, tips_as=tips_as, scaling=scaling)
    assert rel(g["logl"], o["logl"]) < RTOL
    close(g["sites"], o["sites"], 1e-10, 1e-10)
    close(g["s1"], o["s1"], 1e-8, 1e-8)
    close(g["s2"], o["s2"], 1e-8, 1e-8)
    close(g["cross"], o["cross"], 1e-8, 1e-8)


@pytest.mark.parametrize("precision,shape,n", [("double", "caterpillar", 64),
                                              ("single", "caterpillar", 64),
                                              ("single", "balanced", 128),
                                              ("double", "yule", 40)])
def test_auto_scaling_matches_oracle(amd, oracle, precision, shape, n):
    """SCALING_AUTO: exact power-of-two rescaling inside every operation.
    FP64 must equal the oracle's (manually rescaled) FP64 result to 1e-10;
    FP32 must stay within 1e-6 although unscaled FP32 partials underflow."""
    prob = synth.make_problem(41 + n, n, 3000, S=4, shape=shape, missing=0.03)
    ev = driver.Evaluator(amd, prob, precision=precision, scaling="auto")
    assert ev.inst.flags & api.FLAG_SCALING_AUTO
    lg = ev.log_likelihood()
    ev.pre_order()
    _, s1, _ = ev.edge_derivatives(False, per_site=False)
    ev.close()
    ev = driver.Evaluator(oracle, prob, scaling="manual")
    lo = ev.log_likelihood()
    ev.pre_order()
    _, o1, _ = ev.edge_derivatives(False, per_site=False)
    ev.close()
    tol = 1e-10 if precision == "double" else 1e-6
    assert rel(lg, lo) < tol, (lg, lo)
    close(s1, o1, 1e-8 if precision == "double" else 1e-4, 1e-6 if precision == "double" else 1e-2)


def test_single_precision_close_to_oracle(amd, oracle):
    prob = synth.make_problem(5, 16, 2000, S=4, missing=0.02)
    ev = driver.Evaluator(amd, prob, precision="single", scaling="manual")
    assert ev.inst.impl_name == "MI355X-HIP-Single"
    lg = ev.log_likelihood()
    ev.close()
    ev = driver.Evaluator(oracle, prob, scaling="manual")
    lo = ev.log_likelihood()
    ev.close()
    assert rel(lg, lo) < 1e-6


MFMA_CASES = [
    # S, taxa, patterns, shape, tips_as, scaling, missing
    (20, 12, 700, "yule", "states", "none", 0.05),
    (20, 10, 513, "caterpillar", "partials", "manual", 0.05),
    (61, 8, 300, "yule", "states", "always", 0.05),
    (61, 10, 257, "balanced", "states", "auto", 0.05),
    (61, 7, 130, "yule", "partials", "manual_log", 0.1),
    (5, 9, 100, "yule", "partials", "manual", 0.1),
    (33, 7, 97, "caterpillar", "states", "auto", 0.0),
]


# engine: FP32 matrix cores, or split BF16 (hi/lo operands, three products);
# tolerances (logL, sites, root rows) against the FP64 oracle
MFMA_ENGINES = {"fp32": (0, 1e-6, 2e-6, 1e-5),
                "bf16x3": (3, 1e-6, 1e-5, 1e-4)}


@pytest.mark.parametrize("engine", sorted(MFMA_ENGINES))
@pytest.mark.parametrize("S,n,P,shape,tips_as,scaling,missing", MFMA_CASES)
def test_mfma_post_order_single_matches_oracle(amd, oracle, S, n, P, shape, tips_as, scaling,
                                               missing, engine):
    """S > 4 in FP32: the post-order runs on the FP32 matrix cores
    (v_mfma_f32_32x32x2f32) or, after beagleExtSetSplitBF16(3), on the BF16
    matrix cores with hi/lo split operands.  Identical P(t) on both
    sides; FP32 must stay within 1e-6 relative log-likelihood of the FP64
    oracle, site log-likelihoods within 2e-6 relative and the root partials
    (where both sides share a scale) within 1e-5 of each row's maximum;
    split BF16 within 1e-6, 1e-5 and 1e-4."""
    terms, tl, ts, tr = MFMA_ENGINES[engine]
    prob = synth.make_problem(17 + S + n, n, P, S=S, shape=shape, missing=missing)
    ev = driver.Evaluator(amd, prob, precision="single", tips_as=tips_as, scaling=scaling,
                          matrices="explicit")
    assert ev.inst.impl_name == "MI355X-HIP-Single"
    if terms:
        ev.inst.set_split_bf16(terms)
    lg = ev.log_likelihood()
    sg = ev.inst.site_log_likelihoods()
    rg = ev.inst.get_partials(prob.tree.root)
    ev.close()
    ev = driver.Evaluator(oracle, prob, tips_as=tips_as,
                          scaling="manual" if scaling == "auto" else scaling, matrices="explicit")
    lo = ev.log_likelihood()
    so = ev.inst.site_log_likelihoods()
    ro = ev.inst.get_partials(prob.tree.root)
    ev.close()
    print(f"{engine} S={S}: logL rel err {rel(lg, lo):.2e}, worst site rel err {rel(sg, so):.2e}")
    assert rel(lg, lo) < tl, (lg, lo)
    close(sg, so, ts, 0)
    if scaling != "auto":
        # unscaled rows whose FP64 maximum lies below FP32's normal range
        # (1.2e-38) underflow in FP32 and are left out
        rg = np.asarray(rg).reshape(-1, S)
        ro = np.asarray(ro).reshape(-1, S)
        keep = ro.max(axis=1) > 1e-30
        assert keep.mean() > 0.5
        close_rows(rg[keep], ro[keep], S, tr)


@pytest.mark.parametrize("S,tips_as,precision", [(4, "states", "double"),
                                                  (4, "partials", "double"),
                                                  (20, "states", "double"),
                                                  (61, "partials", "double"),
                                                  (4, "states", "single"),
                                                  (20, "partials", "single")])
def test_partition_reorder_matches_oracle(amd, oracle, S, tips_as, precision):
    """Non-contiguous partitions reorder the patterns (tips and weights
    permuted on the device, site log-likelihoods returned in the client's
    order); the oracle's reorder is pinned to the reference CPU library in
    test_oracle.py."""
    from test_oracle import reordered_partition_run
    prob = synth.make_problem(61 + S, 7, 150 if S > 4 else 700, S=S, missing=0.05)
    parts = np.random.default_rng(3).integers(0, 3, prob.P).astype(np.int32)
    g = reordered_partition_run(amd, prob, parts, tips_as, precision=precision,
                                matrices="explicit")
    o = reordered_partition_run(oracle, prob, parts, tips_as, matrices="explicit")
    assert g["second_reorder_rc"][0] == api.ERROR_NO_IMPLEMENTATION
    tol = 1e-12 if precision == "double" else 1e-6
    close(g["by_partition"], o["by_partition"], tol, 0)
    close(g["total"], o["total"], tol, 0)
    close(g["sites"], o["sites"], tol, 0)
    if precision == "double":
        close_rows(g["root"], o["root"], S, 1e-12)


@pytest.mark.parametrize("precision,scaling,shape,asynch", [("single", "auto", "caterpillar", True),
                                                            ("double", "none", "yule", False),
                                                            ("double", "manual", "balanced", True)])
def test_graph_replays_match_eager(amd, precision, scaling, shape, asynch):
    """HIP-graph capture of a fused logL + gradient evaluation: replays with
    new branch lengths, with eager evaluations and waits between them,
    reproduce the eager results exactly, and a second replay of the same
    lengths gives the same values again."""
    prob = synth.make_problem(91, 24, 3000, S=4, shape=shape, missing=0.03)
    rng = np.random.default_rng(2)
    lens = [prob.tree.lengths * rng.uniform(0.5, 1.5, prob.tree.lengths.shape) for _ in range(4)]
    ev = driver.Evaluator(amd, prob, precision=precision, scaling=scaling,
                          extra_flags=api.FLAG_COMPUTATION_ASYNCH if asynch else 0)
    ref = [ev.fused_gradient(x) for x in lens]
    ev.capture_fused_gradient()
    for k, x in enumerate(lens):
        got = ev.replay_fused_gradient(x, wait=not asynch)
        if asynch:
            got = ev.collect_fused_gradient()
        eager = ev.fused_gradient(lens[(k + 1) % len(lens)])     # eager work in between
        for a, b in zip(got, ref[k]):
            np.testing.assert_array_equal(np.asarray(a), np.asarray(b))
        for a, b in zip(eager, ref[(k + 1) % len(lens)]):
            np.testing.assert_array_equal(np.asarray(a), np.asarray(b))
    again = ev.replay_fused_gradient(lens[-1])
    for a, b in zip(again, ref[-1]):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(b))
    ev.close()


def test_graph_codon_root_and_spoiled_capture(amd):
    """A 61-state FP32 (MFMA) matrices + post-order + root logL graph
    replays to the eager value; a capture that contains a waiting call
    (a getter) is refused and leaves the instance usable."""
    prob = synth.make_problem(93, 12, 500, S=61)
    ev = driver.Evaluator(amd, prob, precision="single", scaling="auto", preorder=False)
    want = ev.log_likelihood()
    out = np.zeros(1)
    ev.inst.begin_capture()
    ev.update_matrices()
    ev.post_order()
    ev.inst.root_log_likelihood(prob.tree.root, 0, 0, api.OP_NONE, out=out)
    g = ev.inst.end_capture()
    for _ in range(3):
        out[0] = 0.0
        ev.inst.launch_graph(g)
        ev.inst.wait_graph(g)
        assert out[0] == want
        assert ev.log_likelihood() == want
    ev.inst.begin_capture()
    with pytest.raises(api.BeagleError):
        ev.inst.site_log_likelihoods()
    with pytest.raises(api.BeagleError):
        ev.inst.end_capture()
    assert ev.log_likelihood() == want
    ev.inst.destroy_graph(g)
    ev.close()


ADJ_CASES = [
    # S, taxa, patterns, shape, tips_as, scaling, transpose
    (20, 10, 400, "yule", "states", "none", "auto"),
    (20, 9, 333, "caterpillar", "partials", "manual", "auto"),
    (61, 8, 200, "balanced", "states", "auto", "auto"),
    (61, 7, 150, "yule", "partials", "always", "auto"),
    (33, 6, 97, "yule", "states", "auto", "auto"),
    (20, 8, 300, "yule", "states", "none", "manual"),
    (61, 6, 130, "balanced", "partials", "auto", "manual"),
]


def _adjoint_run(lib, prob, precision, tips_as, scaling, transpose, terms=0):
    ev = driver.Evaluator(lib, prob, precision=precision, tips_as=tips_as, scaling=scaling,
                          matrices="explicit", transpose=transpose)
    if terms:
        ev.inst.set_split_bf16(terms)
    out = {"logl": ev.log_likelihood()}
    ev.pre_order()
    out["d1"], out["s1"], out["q1"] = ev.edge_derivatives(False)
    _, out["s2"], _ = ev.edge_derivatives(True, per_site=False)
    out["pre"] = ev.inst.get_partials(ev.pre_base + ev.branch_nodes()[1])
    ev.close()
    return out


@pytest.mark.parametrize("S,n,P,shape,tips_as,scaling,transpose", ADJ_CASES)
def test_mfma_adjoint_single_matches_oracle(amd, oracle, S, n, P, shape, tips_as, scaling,
                                            transpose):
    """S > 4 in FP32: pre-order partials (P_self^T (pre_parent * P_sib x_sib)
    with the second product fed from accumulator registers) and edge
    derivatives (D x with D = r_c Q, r_c^2 Q^2) on the FP32 matrix cores,
    against the FP64 oracle: logL within 1e-6 relative, per-site
    derivative ratios and per-branch sums within 1e-4 of their largest
    magnitude, pre-order partials (unscaled runs) within 1e-5 of each row's
    maximum.  transpose="manual": the client passes the pre-order self and
    derivative matrices transposed (the reference GPU convention)."""
    prob = synth.make_problem(29 + S + n, n, P, S=S, shape=shape, missing=0.05, mean_length=0.2)
    g = _adjoint_run(amd, prob, "single", tips_as, scaling, transpose)
    o = _adjoint_run(oracle, prob, "double", tips_as,
                     "manual" if scaling == "auto" else scaling, "auto")
    assert rel(g["logl"], o["logl"]) < 1e-6
    for k in ("d1", "s1", "q1", "s2"):
        a, b = np.asarray(g[k]), np.asarray(o[k])
        assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max(), (k, np.abs(a - b).max())
    if scaling == "none":
        pg = np.asarray(g["pre"]).reshape(-1, S)
        po = np.asarray(o["pre"]).reshape(-1, S)
        keep = po.max(axis=1) > 1e-30
        close_rows(pg[keep], po[keep], S, 1e-5)


BF16_ADJ_CASES = [
    # S, taxa, patterns, shape, tips_as, scaling, transpose, products per hi/lo pair
    (12, 7, 180, "yule", "states", "auto", "auto", 3),
    (20, 10, 400, "yule", "states", "none", "auto", 3),
    (33, 6, 97, "yule", "partials", "manual", "auto", 4),
    (61, 8, 200, "balanced", "states", "auto", "manual", 3),
    (61, 7, 150, "yule", "partials", "none", "auto", 4),
]


@pytest.mark.parametrize("S,n,P,shape,tips_as,scaling,transpose,terms", BF16_ADJ_CASES)
def test_split_bf16_adjoint_matches_oracle(amd, oracle, S, n, P, shape, tips_as, scaling,
                                           transpose, terms):
    """beagleExtSetSplitBF16 on the adjoint: post-order, pre-order (P_sib
    x_sib, then P_self^T fed from the accumulator registers) and edge
    derivatives on v_mfma_f32_32x32x16_bf16 with every operand split into
    BF16 hi and lo parts (3 or 4 products per pair), against the FP64
    oracle: logL within 1e-6 relative, per-site derivative ratios and
    per-branch sums within 1e-4 of their largest magnitude, pre-order
    partials (unscaled runs) within 1e-4 of each row's maximum.  The
    pre-order partials differ from the FP32 engine's, so the BF16 kernels
    ran."""
    prob = synth.make_problem(29 + S + n, n, P, S=S, shape=shape, missing=0.05, mean_length=0.2)
    g = _adjoint_run(amd, prob, "single", tips_as, scaling, transpose, terms)
    f = _adjoint_run(amd, prob, "single", tips_as, scaling, transpose)
    o = _adjoint_run(oracle, prob, "double", tips_as,
                     "manual" if scaling == "auto" else scaling, "auto")
    assert rel(g["logl"], o["logl"]) < 1e-6
    for k in ("d1", "s1", "q1", "s2"):
        a, b = np.asarray(g[k]), np.asarray(o[k])
        assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max(), (k, np.abs(a - b).max())
    assert not np.array_equal(np.asarray(g["pre"]), np.asarray(f["pre"]))
    if scaling == "none":
        pg = np.asarray(g["pre"]).reshape(-1, S)
        po = np.asarray(o["pre"]).reshape(-1, S)
        keep = po.max(axis=1) > 1e-30
        close_rows(pg[keep], po[keep], S, 1e-4)


@pytest.mark.parametrize("S,tips_as,scaling,transpose", [(20, "states", "manual", "auto"),
                                                        (61, "partials", "none", "manual"),
                                                        (33, "states", "auto", "auto")])
def test_mfma_adjoint_double_matches_oracle(amd, oracle, S, tips_as, scaling, transpose):
    """S > 4 in FP64 (v_mfma_f64_16x16x4f64 pre-order and edge
    derivatives) against the oracle on identical matrices: logL within
    1e-12, derivative ratios and sums within 1e-10 of their largest
    magnitude, pre-order partials (unscaled) within 1e-12 of each row's
    maximum."""
    prob = synth.make_problem(31 + S, 8, 200, S=S, missing=0.05, mean_length=0.2)
    g = _adjoint_run(amd, prob, "double", tips_as, scaling, transpose)
    o = _adjoint_run(oracle, prob, "double", tips_as, "manual" if scaling == "auto" else scaling,
                     "auto")
    assert rel(g["logl"], o["logl"]) < 1e-12
    for k in ("d1", "s1", "q1", "s2"):
        a, b = np.asarray(g[k]), np.asarray(o[k])
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max(), (k, np.abs(a - b).max())
    if scaling == "none":
        close_rows(g["pre"], o["pre"], S, 1e-12)


@pytest.mark.parametrize("S,transpose,terms", [(20, "auto", 0), (61, "auto", 0), (61, "manual", 0),
                                              (20, "manual", 3), (61, "auto", 3)])
def test_mfma_fused_gradient_single(amd, oracle, S, transpose, terms):
    """beagleExtCalculateLogLikelihoodAndDerivatives for S > 4 in FP32
    (matrix-core pre-order, then both derivative matrices in one pass; FP32
    or, terms = 3, split-BF16 matrix cores; the tips' operations without a
    destination, their branch sums formed in the pre-order kernel) against
    the oracle's beagleUpdatePrePartials + two beagleCalculateEdgeDerivatives."""
    prob = synth.make_problem(37 + S, 9, 300, S=S, missing=0.05, mean_length=0.2)
    ev = driver.Evaluator(amd, prob, precision="single", scaling="auto", matrices="explicit",
                          transpose=transpose)
    if terms:
        ev.inst.set_split_bf16(terms)
    lg, g1, g2 = ev.fused_gradient()
    # tips' pre-order partials kept: the same sums (a tip branch's come from
    # the pre-order kernel either way)
    lk, k1, k2 = ev.fused_gradient(keep_tips=True)
    ev.close()
    assert lk == lg and np.array_equal(k1, g1) and np.array_equal(k2, g2)
    ev = driver.Evaluator(oracle, prob, scaling="manual", matrices="explicit")
    lo, o1, o2 = ev.gradient()
    ev.close()
    assert rel(lg, lo) < 1e-6
    assert np.abs(g1 - o1).max() <= 1e-4 * np.abs(o1).max()
    assert np.abs(g2 - o2).max() <= 1e-4 * np.abs(o2).max()


def test_split_bf16_engine_selection(amd):
    """beagleExtSetSplitBF16: FP32 instances with 5..64 states only, terms 0,
    3 or 4."""
    for S, prec, ok in ((20, "single", True), (20, "double", False), (4, "single", False)):
        prob = synth.make_problem(3, 6, 70, S=S)
        ev = driver.Evaluator(amd, prob, precision=prec)
        rc = amd.call("ExtSetSplitBF16", ev.inst.h, 3)
        assert rc == (api.SUCCESS if ok else api.ERROR_NO_IMPLEMENTATION), (S, prec, rc)
        if ok:
            assert amd.call("ExtSetSplitBF16", ev.inst.h, 2) == api.ERROR_OUT_OF_RANGE
        ev.close()


def test_mfma_post_order_partitions(amd, oracle):
    """Partitioned post-order (pattern ranges that start and end inside a
    32-pattern tile) on the MFMA path, 20 states, FP32."""
    prob = synth.make_problem(23, 10, 333, S=20)
    parts = np.repeat(np.arange(3), [70, 101, 162]).astype(np.int32)
    out = []
    for lib, prec in ((amd, "single"), (oracle, "double")):
        ev = driver.Evaluator(lib, prob, precision=prec, matrices="explicit")
        ev.update_matrices()
        ev.inst.set_pattern_partitions(3, parts)
        ops = synth.postorder_ops(prob.tree)
        pops = []
        for q in range(3):
            for o in ops:
                pops.append(list(o) + [q, -1])
        ev.inst.update_partials_by_partition(np.array(pops, np.int32))
        r = prob.tree.root
        by, tot = ev.inst.root_log_likelihood_by_partition([r] * 3, [0] * 3, [0] * 3, [-1] * 3,
                                                           [0, 1, 2])
        out.append((by, tot))
        ev.close()
    close(out[0][0], out[1][0], 1e-6, 0)
    assert rel(out[0][1], out[1][1]) < 1e-6


@pytest.mark.parametrize("S,traversal,precision", [(4, "1", "single"), (4, None, "single"),
                                                  (20, None, "single"), (61, None, "single"),
                                                  (4, "1", "double"), (20, None, "double"),
                                                  (61, None, "bf16x3")])
def test_partition_work_lists_match_oracle(amd, oracle, monkeypatch, S, traversal, precision):
    """Operations by partition over ten uneven partitions (one without
    patterns, some smaller than a tile, ranges that start and end inside
    tiles): every level launch over pattern ranges runs one workgroup per
    (operation, tile) item of its work list.  Post-order, root
    log-likelihoods by partition and pre-order partials by partition against
    the oracle: logL 1e-6 relative (FP64 1e-10) per partition and total,
    pre-order partials 1e-5 (FP64 1e-10) of each row's maximum.  traversal
    "1" forces the 4-state level plan (batched levels); bf16x3: the split-BF16
    engine (pre-order partials 1e-4 of each row's maximum)."""
    if traversal:
        monkeypatch.setenv("BEAGLE_AMD_TRAVERSAL", traversal)
    sizes = [0, 3, 40, 70, 1, 150, 33, 64, 5, 200]
    K = len(sizes)
    prob = synth.make_problem(41 + S, 9, sum(sizes), S=S, missing=0.05, mean_length=0.2)
    parts = np.repeat(np.arange(K), sizes).astype(np.int32)
    out = []
    for lib, prec in ((amd, precision), (oracle, "double")):
        ev = driver.Evaluator(lib, prob, precision="single" if prec == "bf16x3" else prec,
                              matrices="explicit")
        if prec == "bf16x3":
            ev.inst.set_split_bf16(3)
        ev.update_matrices()
        ev.inst.set_pattern_partitions(K, parts)
        pops = [list(o) + [q, -1] for q in range(K) for o in synth.postorder_ops(prob.tree)]
        ev.inst.update_partials_by_partition(np.array(pops, np.int32))
        r = prob.tree.root
        by, tot = ev.inst.root_log_likelihood_by_partition([r] * K, [0] * K, [0] * K, [-1] * K,
                                                           list(range(K)))
        ev.set_root_pre()
        pre = ev.preorder_operations()
        qops = [list(o) + [q, -1] for q in range(K) for o in pre]
        ev.inst.update_pre_partials_by_partition(np.array(qops, np.int32))
        pp = [np.asarray(ev.inst.get_partials(ev.pre_base + v)).reshape(-1, S)
              for v in ev.branch_nodes()[:4]]
        out.append((np.asarray(by), tot, pp))
        ev.close()
    (bg, tg, pg), (bo, to, po) = out
    tl, tp = {"single": (1e-6, 1e-5), "bf16x3": (1e-6, 1e-4), "double": (1e-10, 1e-10)}[precision]
    assert rel(tg, to) < tl
    close(bg, bo, tl, 0)
    for a, b in zip(pg, po):
        keep = b.max(axis=1) > 1e-30
        close_rows(a[keep], b[keep], S, tp)


def test_matrices_match_oracle(amd, oracle):
    prob = synth.make_problem(3, 8, 100, S=20)
    mats = []
    for lib in (amd, oracle):
        ev = driver.Evaluator(lib, prob)
        n = 2 * ev.n - 2
        ev.inst.update_transition_matrices(0, np.arange(3), [n + 1, n + 2, n + 3],
                                           None, prob.tree.lengths[:3])
        mats.append([ev.inst.get_transition_matrix(m) for m in range(3)])
        ev.close()
    for a, b in zip(*mats):
        close(a, b, 1e-12, 1e-15)


def test_edge_log_likelihood_matches_oracle(amd, oracle):
    prob = synth.make_problem(9, 10, 500, S=4, missing=0.05)
    res = []
    for lib in (amd, oracle):
        ev = driver.Evaluator(lib, prob)
        ev.log_likelihood()
        ev.pre_order()
        # edge above tip 0: parent side = pre-order partials of tip 0 times freqs is
        # the reference's convention of a rooted edge; use root's children here
        a, b = prob.tree.children[prob.tree.root]
        m = ev.n_matrices - 1
        res.append(ev.inst.edge_log_likelihood(prob.tree.root, a, a, None, None))
        ev.close()
    assert rel(res[0][0], res[1][0]) < RTOL


def test_partitions_match_oracle(amd, oracle):
    prob = synth.make_problem(21, 10, 900, S=4)
    parts = np.repeat(np.arange(3), [200, 300, 400]).astype(np.int32)
    out = []
    for lib in (amd, oracle):
        ev = driver.Evaluator(lib, prob)
        ev.update_matrices()
        ev.inst.set_pattern_partitions(3, parts)
        ops = synth.postorder_ops(prob.tree)
        pops = []
        for q in range(3):
            for o in ops:
                pops.append(list(o) + [q, -1])
        ev.inst.update_partials_by_partition(np.array(pops, np.int32))
        r = prob.tree.root
        by, tot = ev.inst.root_log_likelihood_by_partition([r] * 3, [0] * 3, [0] * 3, [-1] * 3,
                                                           [0, 1, 2])
        out.append((by, tot))
        ev.close()
    close(out[0][0], out[1][0], 1e-10, 0)
    assert rel(out[0][1], out[1][1]) < RTOL


@pytest.mark.parametrize("S,shape,scaling,precision,keep", [(4, "yule", "none", "double", True),
                                                            (4, "balanced", "manual", "double", False),
                                                            (4, "caterpillar", "auto", "double", True),
                                                            (4, "yule", "auto", "single", False),
                                                            (4, "balanced", "auto", "single", True),
                                                            (20, "yule", "none", "double", True)])
def test_fused_pre_order_derivatives_match_oracle(amd, oracle, S, shape, scaling, precision, keep):
    """Extension beagleExtUpdatePrePartialsWithDerivatives against the
    oracle's beagleUpdatePrePartials + two beagleCalculateEdgeDerivatives
    calls, per branch: dlogL/dt and d2logL/dt2; with keep, also the tips'
    pre-order partials (otherwise the tips' operations have no
    destination)."""
    prob = synth.make_problem(77 + S, 20, 1500 if S == 4 else 200, S=S, shape=shape,
                              missing=0.05)
    ev = driver.Evaluator(amd, prob, scaling=scaling, precision=precision, matrices="explicit")
    lg, g1, g2 = ev.fused_gradient(keep_tips=keep)
    pre_tip = ev.inst.get_partials(ev.pre_base + 3) if keep else None
    ev.close()
    ev = driver.Evaluator(oracle, prob, scaling="manual" if scaling == "auto" else scaling,
                          matrices="explicit")
    lo, o1, o2 = ev.gradient()
    pre_tip_o = ev.inst.get_partials(ev.pre_base + 3)
    ev.close()
    if precision == "double":
        assert rel(lg, lo) < 1e-12
        close(g1, o1, 1e-9, 1e-8)
        close(g2, o2, 1e-9, 1e-8)
    else:
        assert rel(lg, lo) < 1e-6
        close(g1, o1, 1e-3, 1e-2)
        close(g2, o2, 1e-3, 1e-1)
    if keep and scaling == "none":
        close(pre_tip, pre_tip_o, 1e-10 if precision == "double" else 1e-4, 0)


@pytest.mark.parametrize("C", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("shape", ["caterpillar", "yule"])
def test_fused_gradient_category_counts(amd, oracle, monkeypatch, C, shape):
    """The FP32 node-fused derivative pass for each category count (the
    kernels specialised for 1, 2 and 4 categories, the general one for the
    others), with the whole list fused and with every level batched: logL
    within 1e-6 and each branch's derivatives within 1e-5 of the largest
    against the oracle (FP64), and the two plans bit for bit."""
    prob = synth.make_problem(90 + C, 24, 1800, S=4, shape=shape, C=C, missing=0.04)
    res = []
    for mode in ("2", "1"):
        monkeypatch.setenv("BEAGLE_AMD_TRAVERSAL", mode)
        ev = driver.Evaluator(amd, prob, precision="single", scaling="auto", matrices="explicit")
        res.append(ev.fused_gradient())
        ev.close()
    ev = driver.Evaluator(oracle, prob, scaling="manual", matrices="explicit")
    lo, o1, o2 = ev.gradient()
    ev.close()
    lg, g1, g2 = res[0]
    assert rel(lg, lo) < 1e-6
    assert np.max(np.abs(g1 - o1)) <= 1e-5 * np.max(np.abs(o1))
    assert np.max(np.abs(g2 - o2)) <= 1e-5 * np.max(np.abs(o2))
    for a, b in zip(res[0], res[1]):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("P,C", [(1, 4), (37, 1), (64, 2), (65, 4), (130, 4)])
def test_fused_gradient_small_pattern_counts(amd, oracle, P, C):
    """The FP32 node-fused derivative pass with fewer patterns than one
    64-pattern tile, exactly one tile and one pattern past it (lanes past
    the last pattern neither store nor add to the sums): logL and each
    branch's derivatives against the oracle."""
    prob = synth.make_problem(500 + P, 12, P, S=4, shape="yule", C=C, missing=0.1)
    ev = driver.Evaluator(amd, prob, precision="single", scaling="auto", matrices="explicit")
    lg, g1, g2 = ev.fused_gradient(keep_tips=True)
    ev.close()
    ev = driver.Evaluator(oracle, prob, scaling="manual", matrices="explicit")
    lo, o1, o2 = ev.gradient()
    ev.close()
    assert rel(lg, lo) < 1e-6
    assert np.max(np.abs(g1 - o1)) <= 1e-5 * max(np.max(np.abs(o1)), 1e-30)
    assert np.max(np.abs(g2 - o2)) <= 1e-5 * max(np.max(np.abs(o2)), 1e-30)


@pytest.mark.parametrize("mode,chunk", [("1", "1"), ("2", "1"), ("0", "3"), ("1", "2")])
@pytest.mark.parametrize("precision,scaling", [("double", "none"), ("double", "always"),
                                               ("single", "auto")])
def test_traversal_plans_agree(amd, monkeypatch, mode, chunk, precision, scaling):
    """Every dependency level batched (mode 1), the whole list fused (mode
    2), and batched levels walked in chunks of independent operations by
    the pipelined kernels (chunk > 1) give the automatic plan's logL,
    pre-order partials and derivative sums bit for bit on a balanced tree
    (wide levels) and a caterpillar (one op per level)."""
    for shape in ("balanced", "caterpillar"):
        prob = synth.make_problem(5, 32, 3000, S=4, shape=shape)
        res = []
        for m, ch in (("0", "1"), (mode, chunk)):
            monkeypatch.setenv("BEAGLE_AMD_TRAVERSAL", m)
            monkeypatch.setenv("BEAGLE_AMD_LEVEL_CHUNK", ch)
            ev = driver.Evaluator(amd, prob, scaling=scaling, precision=precision)
            lg, g1, g2 = ev.fused_gradient(keep_tips=True)
            ev.pre_order()
            res.append((lg, g1, g2, ev.inst.get_partials(ev.pre_base + 1),
                        ev.inst.get_partials(ev.pre_base + prob.tree.n + 2)))
            ev.close()
        for a, b in zip(res[0], res[1]):
            np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


# ---- the reference's golden vectors ---------------------------------------------

def _golden():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "reference_outputs.npz")
    return np.load(path, allow_pickle=False)


def _cases():
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import cases
    return cases


@pytest.mark.parametrize("name", ["s4_yule_none_states", "s4_balanced_manual_partials",
                                  "s4_caterpillar_manual_log", "s4_yule_always",
                                  "s20_yule_none", "s61_yule_manual", "s4_partitions"])
def test_hip_matches_reference_fixtures(amd, name):
    """The reference CPU library's outputs (tests/golden/reference_outputs.npz)
    reproduced by the HIP plugin from the stored inputs.  P(t) comes from
    each library's own eigen-sum, so arrays are compared relative to their
    largest entry (near-zero P entries carry ~1e-16 absolute rounding);
    the log-likelihood within 1e-10 relative."""
    cases, g = _cases(), _golden()
    prob = synth.problem_from_arrays(g, name + "/in/")
    _, opts, parts = cases.CASES[name]
    out = cases.run(amd, prob, opts, parts, transpose="auto")
    for k in out:
        want = g[name + "/out/" + k]
        got = np.asarray(out[k]).reshape(want.shape)
        if k in ("logL", "partition_logL", "site_logL"):
            close(got, want, 1e-10)
        else:
            scale = np.abs(want).max()
            tol = 1e-6 if prob.S == 61 else 1e-9
            assert np.abs(got - want).max() <= tol * scale, (k, np.abs(got - want).max() / scale)


# ---- full size (128 taxa, 100,000 patterns): size-independent properties ------

@pytest.mark.parametrize("shape", ["caterpillar", "balanced"])
def test_full_size_consistency(amd, shape):
    """At the benchmark size, where the oracle would take minutes: the
    log-likelihood equals the weighted sum of site log-likelihoods; no
    scaling, manual scaling and FP64 power-of-two scaling agree; the fused
    derivative extension equals the beagle.h call sequence; the FP32 path
    is within 1e-6 of FP64 for logL and the gradient (relative to its
    largest component)."""
    prob = synth.make_problem(1, 128, 100_000, S=4, shape=shape)
    res = {}
    for key, kw in (("none", {}), ("manual", {"scaling": "manual"}),
                    ("auto64", {"scaling": "auto"}),
                    ("auto32", {"scaling": "auto", "precision": "single"})):
        ev = driver.Evaluator(amd, prob, **kw)
        lnl = ev.log_likelihood()
        sites = ev.inst.site_log_likelihoods()
        if key != "auto32":
            assert rel(lnl, float(np.dot(sites, prob.pattern_weights))) < 1e-12
        res[key] = (lnl, ev.fused_gradient(keep_tips=False))
        if key == "none":
            res["api"] = (lnl, ev.gradient())
        ev.close()
    base = res["api"][1]
    for key in ("none", "manual", "auto64"):
        assert rel(res[key][0], res["none"][0]) < 1e-12
        for a, b in zip(res[key][1][1:], base[1:]):
            assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    lnl32, g32 = res["auto32"]
    assert rel(lnl32, res["none"][0]) < 1e-6
    for a, b in zip(g32[1:], base[1:]):
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


@pytest.mark.parametrize("S,precision", [(4, "double"), (20, "double"), (4, "single")])
def test_matrix_utilities_match_oracle(amd, oracle, S, precision):
    """Model-category matrices and convolution against the oracle."""
    rng = np.random.default_rng(11 + S)
    C = 3
    models = [synth.gtr(rng, S) for _ in range(C)]
    out = []
    for lib, kw in ((amd, {"requirement": api.FLAG_PRECISION_SINGLE if precision == "single"
                           else api.FLAG_PRECISION_DOUBLE}), (oracle, {})):
        inst = lib.create(2, 3, 2, S, 8, C, 12, C, 1, **kw)
        for e, m in enumerate(models):
            inst.set_eigen(e, m.evec.reshape(-1), m.ivec.reshape(-1), m.evals)
        idx = np.arange(3)
        inst.update_transition_matrices_model_categories(list(range(C)), idx, idx + 3, idx + 6,
                                                         np.array([0.03, 0.4, 1.1]))
        inst.convolve_transition_matrices([0, 1], [1, 2], [9, 10])
        out.append(np.stack([inst.get_transition_matrix(i) for i in list(range(9)) + [9, 10]]))
        inst.finalize()
    tol = 1e-12 if precision == "double" else 2e-6
    np.testing.assert_allclose(out[0], out[1], rtol=0, atol=tol * np.abs(out[1]).max())


def _oracle_loglik(oracle, prob):
    ev = driver.Evaluator(oracle, prob, preorder=False, transpose="none", matrices="explicit")
    lnl = ev.log_likelihood()
    ev.close()
    return lnl


def _with_model(prob, Q=None, freqs=None):
    import dataclasses
    m = dataclasses.replace(prob.model, Q=prob.model.Q if Q is None else Q,
                            freqs=prob.model.freqs if freqs is None else freqs)
    return dataclasses.replace(prob, model=m)


@pytest.mark.parametrize("scaling,precision", [("none", "double"), ("auto", "double"),
                                               ("auto", "single")])
def test_rate_matrix_gradient_against_oracle_finite_differences(amd, oracle, scaling, precision):
    """dlogL/dQ and dlogL/dpi_root from the extension against central
    differences of the oracle's log-likelihood with P = expm(r t (Q + h E))
    set explicitly, for random directions E (not generators), and pi_root
    moved one component at a time."""
    prob = synth.make_problem(61, 9, 400, S=4, shape="yule", missing=0.05, mean_length=0.2)
    ev = driver.Evaluator(amd, prob, scaling=scaling, precision=precision)
    ev.fused_gradient(keep_tips=False)
    G, gpi = ev.rate_matrix_gradient()
    ev.close()
    rng = np.random.default_rng(3)
    # curvature is large (|dlogL/dQ| ~ 1e5): h = 2.5e-5 with a fourth-order
    # stencil converges to ~1e-10 relative
    h = 2.5e-5
    tol = 1e-8 if precision == "double" else 2e-5

    def fd4(f):
        # fourth-order central difference
        return (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)

    for _ in range(3):
        E = rng.normal(size=(4, 4))
        fd = fd4(lambda s: _oracle_loglik(oracle, _with_model(prob, Q=prob.model.Q + s * E)))
        got = float(np.sum(G * E))
        assert abs(got - fd) <= tol * max(abs(fd), np.abs(G).max()), (got, fd)
    for i in range(4):
        e = np.eye(4)[i]
        fd = fd4(lambda s: _oracle_loglik(oracle, _with_model(prob, freqs=prob.model.freqs + s * e)))
        assert abs(gpi[i] - fd) <= tol * max(abs(fd), np.abs(gpi).max()), (i, gpi[i], fd)


@pytest.mark.parametrize("C,scaling,keep_tips,precision,tips_as",
                         [(4, "auto", False, "single", "states"),
                          (4, "none", True, "single", "partials"),
                          (2, "auto", False, "single", "states"),
                          (1, "auto", True, "single", "states"),
                          (3, "auto", False, "single", "states"),
                          (4, "auto", False, "double", "states")])
def test_log_likelihood_and_gradient_one_call(amd, C, scaling, keep_tips, precision, tips_as):
    """beagleExtCalculateLogLikelihoodAndGradient against the two calls it
    stands for (beagleExtCalculateLogLikelihoodAndDerivatives, then
    beagleExtCalculateRateMatrixGradient) on the same instance: logL,
    branch derivatives and dlogL/dpi_root bit for bit; dlogL/dQ (formed
    inside the FP32 derivative pass for 1, 2 and 4 categories, by the two
    calls otherwise) within 1e-6 of the largest component in FP32 and bit
    for bit on the two-call fallbacks (3 categories, FP64)."""
    prob = synth.make_problem(71 + C, 23, 1500, S=4, C=C, shape="yule", missing=0.05,
                              mean_length=0.2)
    ev = driver.Evaluator(amd, prob, scaling=scaling, precision=precision, tips_as=tips_as)
    sep = ev.fused_gradient(keep_tips=keep_tips, rate_matrix="separate")
    ev.fused_gradient(keep_tips=keep_tips, rate_matrix=True)   # plan tuned once
    one = ev.fused_gradient(keep_tips=keep_tips, rate_matrix=True)
    again = ev.fused_gradient(keep_tips=keep_tips, rate_matrix=True)
    ev.close()
    assert one[0] == sep[0]
    for k in (1, 2, 4):
        np.testing.assert_array_equal(one[k], sep[k])
    for a, b in zip(one, again):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(b))
    fused = precision == "single" and C in (1, 2, 4)
    if fused:
        assert np.abs(one[3] - sep[3]).max() <= 1e-6 * np.abs(sep[3]).max()
    else:
        np.testing.assert_array_equal(one[3], sep[3])


@pytest.mark.parametrize("precision", ["single", "double"])
def test_log_likelihood_and_gradient_enqueue_only(amd, precision):
    """beagleExtCalculateLogLikelihoodAndGradient with every output NULL:
    on the fused FP32 path the work is enqueued (success) and the next call
    with outputs returns what a call without the enqueue-only one returns,
    bit for bit; on the FP64 two-call path the call is refused with
    BEAGLE_ERROR_OUT_OF_RANGE."""
    prob = synth.make_problem(83, 23, 1500, S=4, shape="yule", missing=0.05, mean_length=0.2)
    ev = driver.Evaluator(amd, prob, scaling="auto", precision=precision,
                          extra_flags=api.FLAG_COMPUTATION_ASYNCH)
    ref = ev.fused_gradient(rate_matrix=True)
    ref = ev.fused_gradient(rate_matrix=True)   # after the tuning call
    ops, post, d1, d2 = ev.fused_operations()
    root = prob.tree.root
    if precision == "double":
        with pytest.raises(api.BeagleError) as e:
            ev.inst.log_likelihood_and_gradient(root, ev.pre_base + root, ops, post, d1, d2,
                                                out=False)
        assert e.value.code == api.ERROR_OUT_OF_RANGE
        ev.close()
        return
    for _ in range(3):
        assert ev.inst.log_likelihood_and_gradient(root, ev.pre_base + root, ops, post, d1,
                                                   d2, out=False) is None
    got = ev.fused_gradient(rate_matrix=True)
    ev.close()
    assert got[0] == ref[0]
    for a, b in zip(got[1:], ref[1:]):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


def test_log_likelihood_and_gradient_against_oracle_finite_differences(amd, oracle):
    """The one-call FP32 gradient (dlogL/dQ formed inside the derivative
    pass) against fourth-order differences of the oracle's log-likelihood
    along random directions, as the separate rate-matrix gradient is
    checked; and dlogL/dt against the oracle's edge derivatives."""
    prob = synth.make_problem(61, 9, 400, S=4, shape="yule", missing=0.05, mean_length=0.2)
    ev = driver.Evaluator(amd, prob, scaling="auto", precision="single")
    lnl, d1, d2, G, gpi = ev.fused_gradient(rate_matrix=True)
    ev.close()
    ref = driver.Evaluator(oracle, prob)
    want = ref.gradient()
    ref.close()
    assert rel(lnl, want[0]) < 1e-6
    assert np.abs(d1 - want[1]).max() <= 1e-5 * np.abs(want[1]).max()
    rng = np.random.default_rng(5)
    h = 2.5e-5

    def fd4(f):
        return (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)

    for _ in range(3):
        E = rng.normal(size=(4, 4))
        fd = fd4(lambda s: _oracle_loglik(oracle, _with_model(prob, Q=prob.model.Q + s * E)))
        assert abs(float(np.sum(G * E)) - fd) <= 2e-5 * max(abs(fd), np.abs(G).max())
    for i in range(4):
        e = np.eye(4)[i]
        fd = fd4(lambda s: _oracle_loglik(oracle, _with_model(prob, freqs=prob.model.freqs + s * e)))
        assert abs(gpi[i] - fd) <= 2e-5 * max(abs(fd), np.abs(gpi).max())


@pytest.mark.parametrize("S,C,scaling,precision", [(20, 4, "none", "double"),
                                                   (20, 4, "auto", "single"),
                                                   (20, 3, "auto", "double"),
                                                   (61, 4, "none", "single"),
                                                   (4, 9, "none", "double")])
def test_rate_matrix_gradient_generic_against_oracle_finite_differences(amd, oracle, S, C,
                                                                        scaling, precision):
    """The rate-matrix gradient for state counts other than 4 (and 4 states
    with more than 8 categories), one branch per pre-order operation:
    dlogL/dQ along random generator directions E (E_ij = Q_ij z_ij, rows
    summing to 0, so Q + h E stays a rate matrix and P stays >= 0: with 20
    Dirichlet(1) frequencies P has entries near 1e-9 that an arbitrary
    direction drives below the clamp at 0) and dlogL/dpi_root for a few
    components, against fourth-order central differences of the oracle's
    log-likelihood with P = expm(r t (Q + h E)) set explicitly."""
    prob = synth.make_problem(83 + S + C, 9, 240, S=S, C=C, shape="yule", missing=0.05,
                              mean_length=0.2)
    ev = driver.Evaluator(amd, prob, scaling=scaling, precision=precision)
    ev.fused_gradient(keep_tips=False)
    G, gpi = ev.rate_matrix_gradient()
    ev.close()
    assert G.shape == (S, S) and gpi.shape == (S,)
    rng = np.random.default_rng(5)
    h = 2.5e-5
    tol = 1e-8 if precision == "double" else 3e-5

    def fd4(f):
        return (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)

    for _ in range(3):
        E = prob.model.Q * rng.normal(size=(S, S))
        np.fill_diagonal(E, 0.0)
        np.fill_diagonal(E, -E.sum(axis=1))
        fd = fd4(lambda s: _oracle_loglik(oracle, _with_model(prob, Q=prob.model.Q + s * E)))
        got = float(np.sum(G * E))
        assert abs(got - fd) <= tol * max(abs(fd), np.abs(G * E).sum()), (got, fd)
    for i in rng.choice(S, size=4, replace=False):
        e = np.eye(S)[i]
        fd = fd4(lambda s: _oracle_loglik(oracle, _with_model(prob, freqs=prob.model.freqs + s * e)))
        assert abs(gpi[i] - fd) <= tol * max(abs(fd), np.abs(gpi).max()), (i, gpi[i], fd)


@pytest.mark.parametrize("S,C,tips_as", [(20, 4, "states"), (61, 2, "partials"),
                                         (33, 3, "states"), (5, 1, "states"),
                                         (64, 4, "partials")])
def test_rate_matrix_gradient_matrix_cores_against_fp64(amd, S, C, tips_as):
    """The rate-matrix gradient of an FP32 instance with 5-64 states and at
    most 4 categories (the matrix-core kernel k_qgm) against the FP64
    instance's (vector units) on the same problem: dlogL/dQ within 1e-6 of
    its largest entry, dlogL/dpi_root likewise; state counts on both sides
    of a 32-state block, tips as states (with missing ones) and as
    partials, pattern counts that leave a partial last tile."""
    prob = synth.make_problem(400 + S + C, 24, 1900, S=S, C=C, shape="yule", missing=0.05,
                              mean_length=0.2)
    res = []
    for precision, scaling in (("single", "auto"), ("double", "none")):
        ev = driver.Evaluator(amd, prob, scaling=scaling, precision=precision, tips_as=tips_as)
        ev.fused_gradient(keep_tips=False)
        res.append(ev.rate_matrix_gradient())
        ev.close()
    (g32, p32), (g64, p64) = res
    assert np.abs(g32 - g64).max() <= 1e-6 * np.abs(g64).max()
    assert np.abs(p32 - p64).max() <= 1e-6 * np.abs(p64).max()


@pytest.mark.parametrize("S,terms", [(61, 3), (20, 4)])
def test_rate_matrix_gradient_split_bf16_error_bound(amd, S, terms):
    """The rate-matrix gradient of a split-BF16 instance (post-order,
    pre-order and derivatives on the BF16 matrix cores) against the FP64
    instance's: within 2e-5 of the largest dQ entry (the BF16 engine's
    pre-order partials carry up to 2^-17 relative rounding per split
    operand; the FP32 engine meets 1e-6), dlogL/dpi_root likewise, and the
    branch gradient within 1e-6."""
    prob = synth.make_problem(410 + S, 24, 1900, S=S, C=4, shape="yule", missing=0.05,
                              mean_length=0.2)
    res = []
    for precision, scaling, t in (("single", "auto", terms), ("double", "none", 0)):
        ev = driver.Evaluator(amd, prob, scaling=scaling, precision=precision)
        if t:
            ev.inst.set_split_bf16(t)
        g = ev.fused_gradient(keep_tips=False)
        res.append((g, ev.rate_matrix_gradient()))
        ev.close()
    (gb, (qb, pb)), (g64, (q64, p64)) = res
    assert np.abs(qb - q64).max() <= 2e-5 * np.abs(q64).max()
    assert np.abs(pb - p64).max() <= 2e-5 * np.abs(p64).max()
    for a, b in zip(gb[1:], g64[1:]):
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


@pytest.mark.parametrize("precision,S", [("double", 4), ("single", 4), ("double", 20)])
def test_torch_autograd_binding(amd, oracle, precision, S):
    """beagle_amd.torch_fn: logL of (branch lengths, GTR(exch, freqs)) and
    its autograd gradients against the oracle's log-likelihood and
    fourth-order central differences of it (20 states: 190
    exchangeabilities, frequencies ~ Dirichlet(5), none near the stencil's
    reach)."""
    import torch
    from beagle_amd import torch_fn
    prob = synth.make_problem(71, 8, 300, S=S, shape="yule", missing=0.05, mean_length=0.2)
    lik = torch_fn.PhyloLikelihood(amd, prob.tree, prob.tips, prob.pattern_weights, prob.rates,
                                   prob.weights, precision=precision, states=S)
    rng = np.random.default_rng(9)
    exch0 = rng.exponential(1.0, S * (S - 1) // 2)
    freqs0 = rng.dirichlet(np.full(S, 1.0 if S == 4 else 5.0))
    lengths0 = prob.tree.lengths[:-1].copy()
    exch = torch.tensor(exch0, requires_grad=True)
    freqs = torch.tensor(freqs0, requires_grad=True)
    lengths = torch.tensor(lengths0, requires_grad=True)
    Q = torch_fn.gtr_rate_matrix(exch, freqs)
    logl = lik(lengths, Q, freqs)
    logl.backward()
    lik.close()

    def oracle_logl(e, f, t):
        import dataclasses
        Qn = torch_fn.gtr_rate_matrix(torch.tensor(e), torch.tensor(f)).numpy()
        m = dataclasses.replace(prob.model, Q=Qn, freqs=f)
        tt = np.append(t, 0.0)
        p = dataclasses.replace(prob, model=m,
                                tree=dataclasses.replace(prob.tree, lengths=tt))
        ev = driver.Evaluator(oracle, p, preorder=False, transpose="none", matrices="explicit")
        v = ev.log_likelihood()
        ev.close()
        return v

    ref = oracle_logl(exch0, freqs0, lengths0)
    assert rel(float(logl), ref) < (1e-10 if precision == "double" else 1e-6)
    h = 1e-4

    def fd4(f):
        return (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)

    tol = 1e-6 if precision == "double" else 2e-4
    for name, x0, grad, call in (
            ("exch", exch0, exch.grad, lambda x: oracle_logl(x, freqs0, lengths0)),
            ("freqs", freqs0, freqs.grad, lambda x: oracle_logl(exch0, x, lengths0)),
            ("lengths", lengths0, lengths.grad, lambda x: oracle_logl(exch0, freqs0, x))):
        g = grad.numpy()
        for i in range(0, len(x0), max(1, len(x0) // 5)):
            e = np.zeros(len(x0))
            e[i] = 1.0
            fd = fd4(lambda s: call(x0 + s * e))
            assert abs(g[i] - fd) <= tol * max(abs(fd), np.abs(g).max()), (name, i, g[i], fd)


# ---- gradient under a partials budget ----------------------------------------

@pytest.mark.parametrize("precision,scaling,shape,tips_as",
                         [("single", "auto", "caterpillar", "states"),
                          ("single", "auto", "yule", "states"),
                          ("double", "none", "balanced", "partials"),
                          ("double", "none", "yule", "states")])
def test_checkpointed_gradient_matches_full(amd, precision, scaling, shape, tips_as):
    """beagleExtCheckpointedGradient keeps post-order partials every k
    heights and recomputes the rest in the pre-order pass: for several k,
    and for the k it picks for a pool far smaller than a full evaluation
    uses, logL and every branch's dlogL/dt and d2logL/dt2 equal the fused
    full evaluation's bit for bit; the peak buffer count it reports stays
    within the pool, and a pool no schedule fits is refused."""
    prob = synth.make_problem(71, 40, 2000, S=4, shape=shape, missing=0.03)
    ev = driver.Evaluator(amd, prob, precision=precision, scaling=scaling, tips_as=tips_as)
    full = ev.fused_gradient()
    n = prob.tree.n
    pool = list(range(ev.pre_base, ev.pre_base + 2 * n - 1))
    for k in (1, 2, 5):
        lnl, g1, g2, used, peak = ev.checkpointed_gradient(pool, k)
        assert used == k and peak <= len(pool)
        assert lnl == full[0]
        np.testing.assert_array_equal(g1, full[1])
        np.testing.assert_array_equal(g2, full[2])
    # the smallest pool any interval fits, found by planning
    best = None
    for m in range(4, len(pool) + 1):
        try:
            _, _, _, k, peak = ev.checkpointed_gradient(pool[:m], 0, plan_only=True)
        except api.BeagleError as e:
            assert e.code == api.ERROR_GENERAL
            continue
        best = (m, k, peak)
        break
    assert best is not None and best[2] <= best[0] and best[0] < n - 1
    m, k, _ = best
    lnl, g1, g2, used, peak = ev.checkpointed_gradient(pool[:m], 0)
    assert used == k and peak <= m
    assert lnl == full[0]
    np.testing.assert_array_equal(g1, full[1])
    np.testing.assert_array_equal(g2, full[2])
    with pytest.raises(api.BeagleError) as e:
        ev.checkpointed_gradient(pool[:m - 1], 0)
    assert e.value.code == api.ERROR_GENERAL
    with pytest.raises(api.BeagleError) as e:
        ev.checkpointed_gradient(pool[:m] + pool[:1], 0)     # a buffer twice
    assert e.value.code == api.ERROR_OUT_OF_RANGE
    ev.close()


@pytest.mark.parametrize("S,precision,scaling", [(20, "single", "auto"), (61, "single", "none"),
                                                 (20, "double", "none")])
def test_checkpointed_gradient_states(amd, S, precision, scaling):
    """The checkpointed gradient for S > 4 (matrix-core post-order,
    pre-order and derivative passes; tips' pre-order partials taken from
    the pool): at interval 3 and at the interval picked for a small pool,
    logL and every branch's derivatives equal the full evaluation's."""
    prob = synth.make_problem(40 + S, 18, 300, S=S, shape="yule", missing=0.03)
    ev = driver.Evaluator(amd, prob, precision=precision, scaling=scaling)
    full = ev.fused_gradient()
    n = prob.tree.n
    pool = list(range(ev.pre_base, ev.pre_base + 2 * n - 1))
    _, _, _, k, peak = ev.checkpointed_gradient(pool, 3, plan_only=True)
    assert k == 3 and peak <= len(pool)
    runs = [ev.checkpointed_gradient(pool, 3)]
    for m in range(4, len(pool) + 1):
        try:
            ev.checkpointed_gradient(pool[:m], 0, plan_only=True)
        except api.BeagleError:
            continue
        runs.append(ev.checkpointed_gradient(pool[:m], 0))
        break
    assert len(runs) == 2
    for lnl, g1, g2, used, peak in runs:
        assert lnl == full[0]
        np.testing.assert_array_equal(g1, full[1])
        np.testing.assert_array_equal(g2, full[2])
    ev.close()


@pytest.mark.parametrize("S,precision", [(4, "double"), (5, "double"), (20, "double"),
                                         (61, "double"), (20, "single")])
def test_complex_eigen_matrices(amd, S, precision):
    """BEAGLE_FLAG_EIGEN_COMPLEX: conjugate pairs as 2 x 2 rotation blocks
    (EigenDecompositionSquare.hpp:108-185); P(r t) against scipy expm,
    clamped at 0, for the category-rate and the model-category updates; the
    derivative matrices are left alone, as in the reference."""
    from scipy.linalg import expm
    rng = np.random.default_rng(40 + S)
    m = synth.nonreversible(rng, S)
    assert np.any(m.evals[S:] != 0)
    rates = np.array([0.4, 1.0, 1.6])
    lengths = np.array([0.02, 0.3, 1.1])
    prefs = api.FLAG_PRECISION_DOUBLE if precision == "double" else api.FLAG_PRECISION_SINGLE
    inst = amd.create(2, 3, 2, S, 8, 3, 12, 3, 1, preference=prefs,
                      requirement=api.FLAG_EIGEN_COMPLEX)
    tol = 1e-12 if precision == "double" else 1e-6
    try:
        assert inst.flags & api.FLAG_EIGEN_COMPLEX
        for e in range(3):
            inst.set_eigen(e, m.evec.reshape(-1), m.ivec.reshape(-1), m.evals)
        inst.set_category_rates(rates)
        marker = np.full(3 * S * S, 7.0)
        inst.set_transition_matrix(3, marker, 0.0)
        inst.update_transition_matrices(0, [0, 1, 2], [3, 4, 5], None, lengths)
        for u, t in enumerate(lengths):
            got = inst.get_transition_matrix(u).reshape(3, S, S)
            for c, r in enumerate(rates):
                np.testing.assert_allclose(got[c], np.maximum(expm(m.Q * (r * t)), 0.0),
                                           rtol=0, atol=tol)
        np.testing.assert_array_equal(inst.get_transition_matrix(3), marker)
        inst.update_transition_matrices_model_categories([0, 1, 2], [6, 7, 8], None, None,
                                                         lengths)
        for u, t in enumerate(lengths):
            got = inst.get_transition_matrix(6 + u).reshape(3, S, S)
            for c in range(3):
                np.testing.assert_allclose(got[c], np.maximum(expm(m.Q * t), 0.0), rtol=0,
                                           atol=tol)
    finally:
        inst.finalize()


@pytest.mark.parametrize("S,scaling", [(4, "none"), (20, "auto")])
def test_complex_eigen_likelihood_matches_oracle(amd, oracle, S, scaling):
    """A non-reversible model through the complex eigen path: logL, site
    logL and branch derivative sums against the oracle fed expm(Q r t)
    explicitly (derivative matrices r Q, r^2 Q^2 on both sides); the
    rate-matrix gradient is refused for complex systems, as is a complex
    instance of more than 62 states."""
    rng = np.random.default_rng(60 + S)
    model = synth.nonreversible(rng, S)
    prob = synth.make_problem(61 + S, 12, 400, S=S, shape="yule", missing=0.05, model=model)
    ev = driver.Evaluator(amd, prob, scaling=scaling, extra_flags=api.FLAG_EIGEN_COMPLEX)
    assert ev.inst.flags & api.FLAG_EIGEN_COMPLEX
    lg = ev.log_likelihood()
    sites = ev.inst.site_log_likelihoods()
    ev.pre_order()
    _, s1, _ = ev.edge_derivatives(False, per_site=False)
    with pytest.raises(api.BeagleError) as e:
        ev.rate_matrix_gradient()
    assert e.value.code == api.ERROR_NO_IMPLEMENTATION
    ev.close()
    ev = driver.Evaluator(oracle, prob, scaling="manual" if scaling == "auto" else scaling,
                          matrices="explicit", transpose="none")
    lo = ev.log_likelihood()
    so = ev.inst.site_log_likelihoods()
    ev.pre_order()
    _, o1, _ = ev.edge_derivatives(False, per_site=False)
    ev.close()
    assert rel(lg, lo) < 1e-10, (lg, lo)
    close(sites, so, 1e-10, 1e-10)
    close(s1, o1, 1e-8, 1e-8)
    with pytest.raises(api.BeagleError) as e:
        amd.create(2, 3, 2, 64, 8, 1, 4, 1, 1, requirement=api.FLAG_EIGEN_COMPLEX)
    assert e.value.code == api.ERROR_NO_IMPLEMENTATION


def _two_model_roots(lib, prob, model2, scaling, precision):
    """One tree under two models in two buffer sets, then the
    several-root likelihood (count 2, weights/frequencies 0 and 1, each
    subset's cumulative scale buffer), its site values and the first
    subset's per-category output."""
    S, n, C, P = prob.S, prob.tree.n, prob.C, prob.tips.shape[1]
    nodes = 2 * n - 1
    manual = scaling == "manual"
    pref = api.FLAG_PRECISION_DOUBLE if precision == "double" else api.FLAG_PRECISION_SINGLE
    pref |= {"none": 0, "manual": api.FLAG_SCALING_MANUAL, "auto": api.FLAG_SCALING_AUTO}[scaling]
    inst = lib.create(n, 2 * nodes, n, S, P, 2, 2 * nodes, C, 2 * n, preference=pref)
    try:
        for t in range(n):
            inst.set_tip_states(t, prob.tips[t])
        for e, m in enumerate((prob.model, model2)):
            inst.set_eigen(e, m.evec.reshape(-1), m.ivec.reshape(-1), m.evals)
            inst.set_state_frequencies(e, m.freqs)
        inst.set_category_weights(0, prob.weights)
        inst.set_category_weights(1, prob.weights[::-1].copy())
        inst.set_category_rates(prob.rates)
        inst.set_pattern_weights(prob.pattern_weights)
        idx = np.arange(nodes - 1)
        lengths = prob.tree.lengths[:nodes - 1]
        inst.update_transition_matrices(0, idx, None, None, lengths)
        inst.update_transition_matrices(1, idx + nodes, None, None, lengths)
        ops1 = synth.postorder_ops(prob.tree, (lambda v: v - n) if manual else None)
        ops2 = ops1.copy()
        ops2[:, 0] += nodes
        for col in (3, 5):
            ops2[:, col] = np.where(ops2[:, col] >= n, ops2[:, col] + nodes, ops2[:, col])
        ops2[:, 4] += nodes
        ops2[:, 6] += nodes
        if manual:
            ops2[:, 1] += n - 1
        cum = [2 * n - 2, 2 * n - 1] if manual else [api.OP_NONE, api.OP_NONE]
        inst.update_partials(ops1, cum[0] if manual else api.OP_NONE)
        inst.update_partials(ops2, cum[1] if manual else api.OP_NONE)
        if manual:
            for k, ops in enumerate((ops1, ops2)):
                inst.reset_scale_factors(cum[k])
                inst.accumulate_scale_factors(ops[:, 1], cum[k])
        root = prob.tree.root
        lnl = inst.root_log_likelihoods([root, root + nodes], [0, 1], [0, 1], cum)
        sites = inst.site_log_likelihoods()
        per_cat = inst.root_log_likelihood_per_category(root, 0, cum[0])
        single = inst.root_log_likelihood(root, 0, 0, cum[0])
        single_sites = inst.site_log_likelihoods()
        return lnl, sites, per_cat, single, single_sites
    finally:
        inst.finalize()


@pytest.mark.parametrize("S,scaling,precision", [(4, "none", "double"), (4, "manual", "double"),
                                                 (20, "manual", "double"),
                                                 (4, "auto", "single")])
def test_several_roots_and_per_category_root(amd, ref, S, scaling, precision):
    """calculateRootLogLikelihoods with count > 1 (one likelihood from
    several roots, per-pattern maximum of the subsets' scale factors,
    BeagleCPUImpl.hpp:2474-2586) and with category weights index -1
    (per-category site log-likelihoods, :2589-2625) against the reference
    CPU library on the same calls; automatic scaling against the reference
    with manual scaling.  Closed form: with category weights w, the
    per-category values give the single-root site values,
    log sum_c w_c exp(per_cat[c]) (no scale factors)."""
    rng = np.random.default_rng(90 + S)
    prob = synth.make_problem(91 + S, 14 if S == 4 else 9, 700 if S == 4 else 150, S=S,
                              shape="yule", missing=0.05)
    model2 = synth.gtr(rng, S)
    g = _two_model_roots(amd, prob, model2, scaling, precision)
    r = _two_model_roots(ref, prob, model2, "manual" if scaling == "auto" else scaling, "double")
    tol = 1e-10 if precision == "double" else 1e-6
    assert rel(g[0], r[0]) < tol, (g[0], r[0])
    close(g[1], r[1], tol, tol)
    close(g[2], r[2], 1e-10 if precision == "double" else 1e-5, tol)
    assert rel(g[3], r[3]) < tol
    if scaling == "none":
        mix = np.log(np.exp(g[2]).T @ prob.weights)
        close(mix, g[4], 1e-12, 1e-12)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_job_list_copies_evicted_and_reused(amd, oracle, precision):
    """More distinct operation lists than an instance keeps on the device
    (eight): twelve Yule topologies over the same tips, evaluated in turn,
    three passes, on one instance (each list is reused from its device copy
    or replaces the least recently used one); every logL equals the
    oracle's for that topology, and repeats of a topology are
    bit-identical."""
    rng = np.random.default_rng(123)
    base = synth.make_problem(124, 10, 500, S=4, missing=0.05)
    n, S, C, P = 10, 4, base.C, base.tips.shape[1]
    trees = [synth.make_tree(rng, n, "yule", 0.1) for _ in range(12)]

    def run(lib, pref, passes):
        inst = lib.create(n, 2 * n, n, S, P, 1, 2 * n, C, 0, preference=pref)
        try:
            for t in range(n):
                inst.set_tip_states(t, base.tips[t])
            m = base.model
            inst.set_eigen(0, m.evec.reshape(-1), m.ivec.reshape(-1), m.evals)
            inst.set_state_frequencies(0, m.freqs)
            inst.set_category_weights(0, base.weights)
            inst.set_category_rates(base.rates)
            inst.set_pattern_weights(base.pattern_weights)
            out = []
            for _ in range(passes):
                for tr in trees:
                    inst.update_transition_matrices(0, np.arange(2 * n - 2), None, None,
                                                    tr.lengths[:2 * n - 2])
                    inst.update_partials(synth.postorder_ops(tr))
                    out.append(inst.root_log_likelihood(tr.root))
            return np.array(out)
        finally:
            inst.finalize()

    pref = api.FLAG_PRECISION_DOUBLE if precision == "double" else api.FLAG_PRECISION_SINGLE
    g = run(amd, pref, 3).reshape(3, -1)
    o = run(oracle, 0, 1)
    np.testing.assert_array_equal(g[0], g[1])
    np.testing.assert_array_equal(g[0], g[2])
    close(g[0], o, 1e-10 if precision == "double" else 1e-6)


def test_complex_eigen_transposed_inverse(amd):
    """BEAGLE_FLAG_INVEVEC_TRANSPOSED with a complex eigen-system: V^-1
    passed transposed gives the same P(r t) = expm(Q r t) (clamped)."""
    from scipy.linalg import expm
    S = 20
    m = synth.nonreversible(np.random.default_rng(77), S)
    rates = np.array([0.5, 1.5])
    lengths = np.array([0.05, 0.4])
    inst = amd.create(2, 3, 2, S, 8, 1, 4, 2, 1, preference=api.FLAG_PRECISION_DOUBLE,
                      requirement=api.FLAG_EIGEN_COMPLEX | api.FLAG_INVEVEC_TRANSPOSED)
    try:
        inst.set_eigen(0, m.evec.reshape(-1), m.ivec.T.copy().reshape(-1), m.evals)
        inst.set_category_rates(rates)
        inst.update_transition_matrices(0, [0, 1], None, None, lengths)
        for u, t in enumerate(lengths):
            got = inst.get_transition_matrix(u).reshape(2, S, S)
            for c, r in enumerate(rates):
                np.testing.assert_allclose(got[c], np.maximum(expm(m.Q * (r * t)), 0.0),
                                           rtol=0, atol=1e-12)
    finally:
        inst.finalize()


# ---- calls no fixture holds: scale-factor bookkeeping, pre-order write
# scales, partitions, matrix derivatives (the same sequences the CPU tests
# pin between the oracle and the reference CPU library) -------------------


@pytest.mark.parametrize("S,precision", [(4, "double"), (4, "single"), (20, "double"),
                                         (20, "single"), (61, "double")])
def test_api_sequence_matches_oracle(amd, oracle, S, precision):
    """sequences.api_sequence on the HIP library against the oracle:
    getPartials with a cumulative buffer, remove / accumulate / copy of
    scale factors and the root log-likelihood over them, pre-order
    operations with write-scale buffers (rescaled partials, their factors,
    the derivatives after them), reset / accumulate / remove of scale
    factors by partition with root log-likelihoods by partition, and the
    partitioned pre-order pass with write scales.  Identical explicit
    matrices on both sides.  FP64: 1e-11 of each array's largest entry;
    FP32: 1e-5 (log-likelihoods 1e-6 relative)."""
    P = {4: 300, 20: 120, 61: 64}[S]
    prob = synth.make_problem(151 + S, 9, P, S=S, missing=0.05, mean_length=0.3)
    got = sequences.api_sequence(amd, prob, precision, transpose="auto", matrices="explicit")
    want = sequences.api_sequence(oracle, prob, "double", transpose="auto", matrices="explicit")
    assert set(got) == set(want)
    for k in want:
        a, b = np.asarray(got[k]), np.asarray(want[k])
        assert a.shape == b.shape, k
        if k.endswith("logL"):
            assert rel(a, b) < (1e-11 if precision == "double" else 1e-6), k
            continue
        scale = max(np.abs(b).max(), 1e-300)
        tol = 1e-11 if precision == "double" else 1e-5
        err = np.abs(a - b).max() / scale
        assert err < tol, f"{k}: {err:.2e}"


@pytest.mark.parametrize("S,precision", [(4, "double"), (4, "single"), (20, "double"),
                                         (20, "single"), (61, "double"), (61, "single")])
def test_transition_derivatives_match_oracle(amd, oracle, S, precision):
    """beagleUpdateTransitionMatrices with first- and second-derivative
    index lists: P, dP/dt and d2P/dt2 on the device against the oracle
    (which the CPU tests pin against the reference and the closed forms
    dP/dt = rQP, d2P/dt2 = r^2 Q^2 P), t = 0 included.  Entries compared to
    the largest magnitude of each matrix kind: FP64 1e-13, FP32 2e-6."""
    prob = synth.make_problem(160 + S, 6, 32, S=S)
    lengths = [0.0, 0.003, 0.05, 0.4, 1.7]
    got = sequences.transition_derivatives(amd, prob, lengths, precision)
    want = sequences.transition_derivatives(oracle, prob, lengths)
    tol = 1e-13 if precision == "double" else 2e-6
    for kind, a, b in zip(("P", "dP", "d2P"), got, want):
        scale = np.abs(b).max(axis=(2, 3), keepdims=True)
        err = (np.abs(a - b) / np.maximum(scale, 1e-300)).max()
        assert err < tol, f"{kind}: {err:.2e}"


def _oracle_gradient(oracle, prob, lengths, scaling, matrices="eigen"):
    ev = driver.Evaluator(oracle, prob, scaling="manual" if scaling == "auto" else scaling,
                          matrices=matrices)
    res = ev.gradient(lengths)
    ev.close()
    return res


def _assert_gradient_close(got, want, precision):
    lg, g1, g2 = got
    lo, o1, o2 = want
    if precision == "double":
        assert rel(lg, lo) < 1e-10
        assert np.max(np.abs(g1 - o1)) <= 1e-9 * np.max(np.abs(o1))
        assert np.max(np.abs(g2 - o2)) <= 1e-9 * np.max(np.abs(o2))
    else:
        assert rel(lg, lo) < 1e-6
        assert np.max(np.abs(g1 - o1)) <= 1e-5 * np.max(np.abs(o1))
        assert np.max(np.abs(g2 - o2)) <= 1e-5 * np.max(np.abs(o2))


@pytest.mark.parametrize("precision,scaling,shape", [("single", "auto", "caterpillar"),
                                                     ("double", "none", "yule"),
                                                     ("double", "manual", "balanced")])
def test_graph_replays_match_oracle(amd, oracle, precision, scaling, shape):
    """Each replay of a captured logL + gradient evaluation, at branch
    lengths set after the capture, against the oracle's beagle.h call
    sequence at those lengths: logL 1e-10 (FP64) / 1e-6 (FP32) relative,
    every branch's dlogL/dt and d2logL/dt2 within 1e-9 / 1e-5 of the
    largest."""
    prob = synth.make_problem(191, 24, 3000, S=4, shape=shape, missing=0.03)
    rng = np.random.default_rng(12)
    lens = [prob.tree.lengths * rng.uniform(0.5, 1.5, prob.tree.lengths.shape) for _ in range(3)]
    ev = driver.Evaluator(amd, prob, precision=precision, scaling=scaling)
    ev.capture_fused_gradient()
    for x in lens:
        got = ev.replay_fused_gradient(x)
        _assert_gradient_close(got, _oracle_gradient(oracle, prob, x, scaling), precision)
    ev.close()


@pytest.mark.parametrize("S,precision,scaling,shape", [(4, "single", "auto", "caterpillar"),
                                                       (4, "double", "none", "yule"),
                                                       (20, "single", "auto", "yule"),
                                                       (61, "double", "none", "balanced")])
def test_checkpointed_gradient_matches_oracle(amd, oracle, S, precision, scaling, shape):
    """beagleExtCheckpointedGradient at interval 2 and with the smallest pool
    that fits, against the oracle's full beagle.h evaluation on identical
    explicit matrices (tolerances as the graph replays)."""
    P = 1500 if S == 4 else 200
    prob = synth.make_problem(230 + S, 30 if S == 4 else 14, P, S=S, shape=shape, missing=0.03)
    want = _oracle_gradient(oracle, prob, None, scaling, matrices="explicit")
    ev = driver.Evaluator(amd, prob, precision=precision, scaling=scaling, matrices="explicit")
    n = prob.tree.n
    pool = list(range(ev.pre_base, ev.pre_base + 2 * n - 1))
    lnl, g1, g2, used, _ = ev.checkpointed_gradient(pool, 2)
    assert used == 2
    _assert_gradient_close((lnl, g1, g2), want, precision)
    for m in range(4, len(pool) + 1):
        try:
            ev.checkpointed_gradient(pool[:m], 0, plan_only=True)
        except api.BeagleError:
            continue
        lnl, g1, g2, _, peak = ev.checkpointed_gradient(pool[:m], 0)
        assert peak <= m
        _assert_gradient_close((lnl, g1, g2), want, precision)
        break
    else:
        raise AssertionError("no pool fits")
    ev.close()


def _small_partitions(seed, count):
    """The 48-partition workload's generators at test size: state counts
    cycling 4, 20, 61; patterns ~LogUniform over a per-state-count range;
    Yule trees of 32-96 taxa (32-48 for 61 states); Gamma(4)."""
    rng = np.random.default_rng(seed)
    ranges = {4: (100, 3000), 20: (50, 400), 61: (40, 160)}
    out = []
    for k in range(count):
        S = (4, 20, 61)[k % 3]
        lo, hi = ranges[S]
        P = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        taxa = int(rng.integers(32, 49 if S == 61 else 97))
        out.append(synth.make_problem(seed + k, taxa, P, S=S, shape="yule", missing=0.02))
    return out


def test_partitions_workload_matches_oracle(amd, oracle):
    """The 48-partition workload scaled down to 9 partitions (4-, 20- and
    61-state models in turn, each with its own tree and pattern count): one
    ASYNCH FP32 instance per partition, each logL + gradient evaluation
    captured as a HIP graph; all 9 replays, at new branch lengths, are
    launched across the instances' streams before any is collected, twice.
    Every partition's logL (1e-6 relative) and every branch's dlogL/dt and
    d2logL/dt2 (1e-4 of the largest; 1e-5 for 4 states) against the
    oracle at the same lengths."""
    probs = _small_partitions(4100, 9)
    rng = np.random.default_rng(3)
    evs = [driver.Evaluator(amd, p, precision="single", scaling="auto",
                            extra_flags=api.FLAG_COMPUTATION_ASYNCH) for p in probs]
    try:
        for ev in evs:
            ev.capture_fused_gradient()
        for _ in range(2):
            lens = [p.tree.lengths * rng.uniform(0.5, 1.5, p.tree.lengths.shape) for p in probs]
            for ev, x in zip(evs, lens):
                ev.replay_fused_gradient(x, wait=False)
            got = [ev.collect_fused_gradient() for ev in evs]
            for p, x, (lg, g1, g2) in zip(probs, lens, got):
                lo, o1, o2 = _oracle_gradient(oracle, p, x, "auto")
                tol = 1e-5 if p.S == 4 else 1e-4
                assert rel(lg, lo) < 1e-6, (p.S, p.P)
                assert np.abs(g1 - o1).max() <= tol * np.abs(o1).max(), (p.S, p.P)
                assert np.abs(g2 - o2).max() <= tol * np.abs(o2).max(), (p.S, p.P)
    finally:
        for ev in evs:
            ev.close()


def test_codon_full_size_engines(amd):
    """The 61-state GY94 codon workload at full size (64 taxa, 20,000
    patterns, Gamma(4)), where the oracle would take minutes: the FP32 and
    split-BF16 (3 and 4 products) matrix-core post-order give logL within
    1e-6 relative of the FP64 matrix-core path, and the FP32 and split-BF16
    (3 products: post-order, pre-order and derivatives) logL + branch
    gradient are within 1e-6 of FP64 (relative to the largest component)."""
    prob = synth.make_problem(3, 64, 20_000, S=61, shape="yule")
    ev = driver.Evaluator(amd, prob)
    l64 = ev.log_likelihood()
    assert rel(l64, float(np.dot(ev.inst.site_log_likelihoods(), prob.pattern_weights))) < 1e-12
    g64 = ev.fused_gradient()
    ev.close()
    ev = driver.Evaluator(amd, prob, precision="single", scaling="auto")
    for terms in (0, 3, 4):
        ev.inst.set_split_bf16(terms)
        assert rel(ev.log_likelihood(), l64) < 1e-6, terms
    ev.inst.set_split_bf16(3)
    gbf = ev.fused_gradient()
    ev.inst.set_split_bf16(0)
    g32 = ev.fused_gradient()
    ev.close()
    for g in (g32, gbf):
        assert rel(g[0], g64[0]) < 1e-6
        for a, b in zip(g[1:], g64[1:]):
            assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


def test_adjoint_full_size_512_taxa(amd):
    """The adjoint workload at full size (512-taxon Yule tree, 50,000
    patterns, 4-state GTR + Gamma(4); unscaled FP64 partials underflow at
    this depth): the fused FP64 gradient extension with power-of-two
    scaling equals the beagle.h call sequence with manual scale buffers
    (1e-10); FP32 logL, branch gradient and rate-matrix / root-frequency
    gradient within 1e-6 of FP64 (relative to the largest component)."""
    prob = synth.make_problem(4, 512, 50_000, S=4, shape="yule")
    ev = driver.Evaluator(amd, prob, scaling="manual")
    api64 = ev.gradient()
    ev.close()
    ev = driver.Evaluator(amd, prob, scaling="auto")
    f64 = ev.fused_gradient()
    q64, pi64 = ev.rate_matrix_gradient()
    ev.close()
    assert np.isfinite(api64[0])
    assert rel(f64[0], api64[0]) < 1e-12
    for a, b in zip(f64[1:], api64[1:]):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    ev = driver.Evaluator(amd, prob, precision="single", scaling="auto")
    f32 = ev.fused_gradient()
    q32, pi32 = ev.rate_matrix_gradient()
    ev.close()
    assert rel(f32[0], f64[0]) < 1e-6
    for a, b in zip(f32[1:], f64[1:]):
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()
    assert np.abs(q32 - q64).max() <= 1e-6 * np.abs(q64).max()
    assert np.abs(pi32 - pi64).max() <= 1e-6 * np.abs(pi64).max()
    ev = driver.Evaluator(amd, prob, precision="single", scaling="auto")
    one = ev.fused_gradient(rate_matrix=True)
    ev.close()
    assert one[0] == f32[0]
    assert np.abs(one[3] - q64).max() <= 1e-6 * np.abs(q64).max()
    np.testing.assert_array_equal(one[4], pi32)


@pytest.mark.parametrize("S,precision", [(4, "single"), (20, "single")])
def test_autotuner_tunes_once_per_tree_size(amd, S, precision):
    """A sampler proposing topologies: 20 random Yule trees of one size over
    the same tips, each a new operation list, evaluated (logL + branch
    gradient) on one instance with autotuning and one without.  Each call
    kind is tuned at most once (shape classes, not operation lists), no
    timing is left pending after the calls, and every result equals the
    untuned instance's bit for bit."""
    rng = np.random.default_rng(8 + S)
    n = 64
    prob = synth.make_problem(88 + S, n, 3000 if S == 4 else 300, S=S, shape="yule")
    trees = [synth.make_tree(rng, n, "yule", 0.1) for _ in range(20)]
    on = driver.Evaluator(amd, prob, precision=precision, scaling="auto")
    off = driver.Evaluator(amd, prob, precision=precision, scaling="auto")
    off.inst.set_autotune(False)
    try:
        for t in trees:
            on.set_tree(t)
            off.set_tree(t)
            a = on.fused_gradient()
            b = off.fused_gradient()
            assert a[0] == b[0]
            np.testing.assert_array_equal(a[1], b[1])
            np.testing.assert_array_equal(a[2], b[2])
        tuned, pending = on.inst.autotune_stats()
        # 4 states: post-order and node-fused pre-order; S > 4: the FP32
        # matrix-core post-order (its pre-order has a single plan)
        assert 1 <= tuned <= (2 if S == 4 else 1), tuned
        assert pending == 0
        assert off.inst.autotune_stats() == (0, 0)
    finally:
        on.close()
        off.close()


@pytest.mark.parametrize("S,precision", [(4, "double"), (4, "single"), (20, "double"),
                                         (20, "single")])
def test_edge_log_likelihood_by_partition_matches_oracle(amd, oracle, S, precision):
    """beagleCalculateEdgeLogLikelihoodsByPartition against the oracle
    (which the CPU tests pin against the reference CPU library) on identical
    explicit matrices: per-partition and total log-likelihoods, first and
    second derivatives, site log-likelihoods, and the whole-pattern edge
    call; FP64 1e-11, FP32 1e-6 of each array's largest magnitude."""
    prob = synth.make_problem(181 + S, 9, 700 if S == 4 else 150, S=S, missing=0.05,
                              mean_length=0.3)
    got = sequences.edge_by_partition(amd, prob, precision, matrices="explicit")
    want = sequences.edge_by_partition(oracle, prob, "double", matrices="explicit")
    tol = 1e-11 if precision == "double" else 1e-6
    for k in want:
        a, b = np.asarray(got[k]), np.asarray(want[k])
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err < tol, f"{k}: {err:.2e}"


@pytest.mark.parametrize("precision", ["double", "single"])
def test_root_by_partition_beyond_64_partitions(amd, oracle, precision):
    """Root log-likelihoods by partition for 100 partitions of uneven sizes
    (one with no pattern): each against the oracle (FP64 1e-11, FP32 1e-6
    relative), and their sum against the whole-pattern root call."""
    P = 3000
    prob = synth.make_problem(321, 12, P, S=4, missing=0.05)
    rng = np.random.default_rng(5)
    cuts = np.sort(rng.choice(np.arange(1, P), 98, replace=False))
    parts = np.zeros(P, dtype=np.int32)
    for c in cuts:
        parts[c:] += 1
    parts[parts >= 50] += 1          # partition 50 holds no pattern
    res = []
    for lib, prec in ((amd, precision), (oracle, "double")):
        ev = driver.Evaluator(lib, prob, scaling="manual", precision=prec, matrices="explicit",
                              preorder=False)
        whole = ev.log_likelihood()
        ev.inst.set_pattern_partitions(100, parts)
        cum = ev.n - 1
        by, tot = ev.inst.root_log_likelihood_by_partition([prob.tree.root] * 100, [0] * 100,
                                                           [0] * 100, [cum] * 100,
                                                           list(range(100)))
        res.append((whole, by, tot))
        ev.close()
    (wg, bg, tg), (wo, bo, to) = res
    tol = 1e-11 if precision == "double" else 1e-6
    assert bg[50] == 0.0
    assert np.abs(bg - bo).max() <= tol * np.abs(bo).max()
    assert rel(tg, to) < tol and rel(tg, wg) < tol


@pytest.mark.parametrize("precision,scaling", [("single", "auto"), ("double", "none")])
def test_graph_replays_rate_matrix_gradient(amd, oracle, precision, scaling):
    """logL, branch derivatives, dlogL/dQ and dlogL/dpi_root captured in one
    HIP graph (4 states): each replay at new branch lengths equals the eager
    calls at those lengths bit for bit (the Frechet factors come from the
    replayed matrix updates), and dlogL/dQ matches fourth-order differences
    of the oracle's log-likelihood along a random direction at the last
    lengths."""
    prob = synth.make_problem(261, 16, 800, S=4, shape="yule", missing=0.05, mean_length=0.2)
    rng = np.random.default_rng(21)
    lens = [prob.tree.lengths * rng.uniform(0.5, 1.5, prob.tree.lengths.shape) for _ in range(3)]
    ev = driver.Evaluator(amd, prob, precision=precision, scaling=scaling)
    ev.fused_gradient(rate_matrix=True)   # the traversal plan tuned before the references
    want = [ev.fused_gradient(x, rate_matrix=True) for x in lens]
    ev.capture_fused_gradient(rate_matrix=True)
    for x, w in zip(lens, want):
        got = ev.replay_fused_gradient(x)
        assert len(got) == 5
        for a, b in zip(got, w):
            np.testing.assert_array_equal(np.asarray(a), np.asarray(b))
    ev.close()
    G = want[-1][3]
    last = _with_lengths(prob, lens[-1])
    h = 2.5e-5
    E = rng.normal(size=(4, 4))
    f = [_oracle_loglik(oracle, _with_model(last, Q=last.model.Q + s * E))
         for s in (-2 * h, -h, h, 2 * h)]
    fd = (8 * (f[2] - f[1]) - (f[3] - f[0])) / (12 * h)
    tol = 1e-8 if precision == "double" else 2e-5
    assert abs(float(np.sum(G * E)) - fd) <= tol * max(abs(fd), np.abs(G).max())


def _with_lengths(prob, lengths):
    import dataclasses
    return dataclasses.replace(prob, tree=dataclasses.replace(prob.tree, lengths=lengths))
