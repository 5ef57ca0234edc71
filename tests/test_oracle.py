"""Pins of the FP64 CPU oracle (oracle/beagle_oracle.c), which the GPU
parity tests use as their checker.  Each test would fail on a plausible
mistake in the oracle (a dropped term, a wrong sign, index or transpose):

  * known answers of the reference's example programs (PAUP and R values
    printed by examples/tinytest/tinytest.cpp:332,
    examples/hmctest/hmctest5.cpp:589,675-679 and
    examples/hmctest/hmcGapTest.cpp:564);
  * closed forms: JC69 transition probabilities, P(0) = I, rows of P
    summing to 1, dP/dt = r Q P(t) and d2P/dt2 = r^2 Q^2 P(t) against
    scipy's matrix exponential;
  * brute force: the likelihood of small trees by enumerating every
    assignment of internal-node states;
  * central finite differences of the log-likelihood for the branch
    derivatives (reference CPU pre-order + calculateEdgeDerivatives
    semantics, BeagleCPUImpl.hpp:1940-2083);
  * golden vectors of the reference's own CPU library (tests/golden/
    reference_outputs.npz, written by tests/golden/make_golden.py), and,
    where the reference build is present, live comparisons for cases the
    fixtures do not hold (a non-generator derivative matrix, whose tip
    denominator follows BeagleCPU4StateImpl.hpp:1032).
"""
import ctypes
import itertools
import json
import os
import sys

import numpy as np
import pytest

from beagle_amd import api, driver, synth

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases  # noqa: E402
import sequences  # noqa: E402

GOLDEN = os.path.join(HERE, "golden", "reference_outputs.npz")
KNOWN = os.path.join(HERE, "golden", "known_answers.json")


def rel(a, b):
    return abs(a - b) / abs(b)


seq_states = sequences.seq_states
small_problem = sequences.small_problem


# ---- known answers -------------------------------------------------------------

def test_tinytest_paup_loglikelihood(oracle):
    spec = json.load(open(KNOWN))["tinytest"]
    prob = small_problem(spec)
    ev = driver.Evaluator(oracle, prob, tips_as="partials", preorder=False, transpose="none")
    lnl = ev.log_likelihood()
    ev.close()
    assert abs(lnl - spec["logL"]) < 0.6 * 10 ** -spec["logL_digits"]


@pytest.mark.parametrize("name", ["hmctest5", "hmcGapTest"])
def test_hmc_examples_r_loglikelihood(oracle, name):
    spec = json.load(open(KNOWN))[name]
    prob = small_problem(spec)
    ev = driver.Evaluator(oracle, prob, preorder=False, transpose="none")
    lnl = ev.log_likelihood()
    ev.close()
    assert rel(lnl, spec["logL"]) < 1e-12


def test_hmctest5_r_site_gradient(oracle):
    """Per-site dlogL_p/dt (D = r_c Q) against the R table; rows are the
    branches above post-order buffers 1, 0, 2, 3, columns the patterns."""
    spec = json.load(open(KNOWN))["hmctest5"]
    prob = small_problem(spec)
    ev = driver.Evaluator(oracle, prob, transpose="none")
    ev.log_likelihood()
    ev.pre_order()
    d1, _, _ = ev.edge_derivatives(False, per_site=True)
    nodes = ev.branch_nodes()
    ev.close()
    d1 = np.asarray(d1).reshape(len(nodes), -1)
    got = np.stack([d1[nodes.index(v)] for v in spec["gradient_nodes"]])
    want = np.array(spec["site_dlogL_dt"])
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-7)


# ---- closed forms --------------------------------------------------------------

def matrices(oracle, model, rates, lengths, derivs=True):
    C, S = len(rates), model.S
    inst = oracle.create(2, 2, 0, S, 4, 1, 3 * len(lengths), C, 0)
    inst.set_eigen(0, model.evec.reshape(-1), model.ivec.reshape(-1), model.evals)
    inst.set_category_rates(rates)
    n = len(lengths)
    idx = np.arange(n)
    inst.update_transition_matrices(0, idx, idx + n if derivs else None,
                                    idx + 2 * n if derivs else None, lengths)
    out = [np.stack([inst.get_transition_matrix(i + k * n).reshape(C, S, S) for i in range(n)])
           for k in range(3 if derivs else 1)]
    inst.finalize()
    return out


def test_jc69_closed_form(oracle):
    rates = np.array([0.3, 1.0, 2.5])
    lengths = np.array([0.0, 0.05, 0.4, 2.0])
    (P,) = matrices(oracle, synth.jc69(), rates, lengths, derivs=False)
    for u, t in enumerate(lengths):
        for c, r in enumerate(rates):
            e = np.exp(-4.0 * r * t / 3.0)
            want = np.full((4, 4), 0.25 - 0.25 * e)
            np.fill_diagonal(want, 0.25 + 0.75 * e)
            np.testing.assert_allclose(P[u, c], want, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("S,seed", [(4, 1), (20, 2), (61, 3)])
def test_transition_matrices_against_expm(oracle, S, seed):
    from scipy.linalg import expm
    rng = np.random.default_rng(seed)
    model = synth.gy94(rng) if S == 61 else synth.gtr(rng, S)
    rates = np.array([0.2, 1.7])
    lengths = np.array([0.0, 0.013, 0.31])
    P, D1, D2 = matrices(oracle, model, rates, lengths)
    Q = model.Q
    for u, t in enumerate(lengths):
        for c, r in enumerate(rates):
            E = expm(Q * r * t)
            scale = np.abs(E).max()
            np.testing.assert_allclose(P[u, c], np.maximum(E, 0.0), atol=1e-12 * scale)
            np.testing.assert_allclose(D1[u, c], r * Q @ E, atol=1e-11 * np.abs(r * Q).max())
            np.testing.assert_allclose(D2[u, c], r * r * Q @ Q @ E,
                                       atol=1e-11 * np.abs(r * r * Q @ Q).max())
    # invariants: P(0) = I, rows of P sum to 1
    np.testing.assert_allclose(P[0], np.broadcast_to(np.eye(S), P[0].shape), atol=1e-12)
    np.testing.assert_allclose(P.sum(axis=-1), 1.0, atol=1e-11)


# ---- brute force ---------------------------------------------------------------

def brute_force_loglikelihood(prob):
    from scipy.linalg import expm
    tree, m = prob.tree, prob.model
    n, S = tree.n, prob.S
    internal = list(range(n, 2 * n - 1))
    Ps = {(v, c): expm(m.Q * r * tree.lengths[v]) for v in range(2 * n - 2)
          for c, r in enumerate(prob.rates)}
    total = 0.0
    for p in range(prob.P):
        lp = 0.0
        for c in range(prob.C):
            acc = 0.0
            for assign in itertools.product(range(S), repeat=len(internal)):
                st = dict(zip(internal, assign))
                prod = m.freqs[st[tree.root]]
                for v in internal:
                    for ch in tree.children[v]:
                        row = Ps[(ch, c)][st[v]]
                        if ch < n:
                            s = prob.tips[ch, p]
                            prod *= row.sum() if s >= S else row[s]
                        else:
                            prod *= row[st[ch]]
                acc += prod
            lp += prob.weights[c] * acc
        total += prob.pattern_weights[p] * np.log(lp)
    return total


@pytest.mark.parametrize("shape", ["yule", "caterpillar", "balanced"])
def test_brute_force_enumeration(oracle, shape):
    prob = synth.make_problem(21, 4, 6, S=4, shape=shape, C=2, missing=0.15, mean_length=0.3)
    ev = driver.Evaluator(oracle, prob, preorder=False, transpose="none", matrices="explicit")
    lnl = ev.log_likelihood()
    ev.close()
    assert rel(lnl, brute_force_loglikelihood(prob)) < 1e-13


# ---- finite differences --------------------------------------------------------

@pytest.mark.parametrize("S,scaling", [(4, "none"), (4, "manual"), (20, "none")])
def test_branch_derivatives_against_finite_differences(oracle, S, scaling):
    prob = synth.make_problem(31 + S, 7, 40 if S == 4 else 12, S=S, missing=0.05)
    ev = driver.Evaluator(oracle, prob, scaling=scaling, transpose="none")
    lnl, g1, g2 = ev.gradient()
    nodes = ev.branch_nodes()
    h = 1e-6
    base = prob.tree.lengths.copy()
    fd1, fd2 = [], []
    for v in nodes:
        vals = []
        for sgn in (1, -1):
            t = base.copy()
            t[v] += sgn * h
            vals.append(ev.gradient(t)[1][nodes.index(v)])
            if sgn == 1:
                lp = ev.log_likelihood(t)
            else:
                lm = ev.log_likelihood(t)
        fd1.append((lp - lm) / (2 * h))
        fd2.append((vals[0] - vals[1]) / (2 * h))
    ev.close()
    np.testing.assert_allclose(g1, fd1, rtol=1e-5, atol=1e-6 * abs(lnl) / 1e3)
    np.testing.assert_allclose(g2, fd2, rtol=1e-5, atol=1e-5 * np.abs(g2).max())


# ---- reference golden vectors ---------------------------------------------------

def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.mark.parametrize("name", list(cases.CASES))
def test_oracle_matches_reference_fixtures(oracle, name):
    g = golden()
    prob = synth.problem_from_arrays(g, name + "/in/")
    _, opts, parts = cases.CASES[name]
    out = cases.run(oracle, prob, opts, parts)
    keys = [k[len(name) + 5:] for k in g.files if k.startswith(name + "/out/")]
    assert set(keys) == set(out)
    for k in keys:
        want = g[name + "/out/" + k]
        got = np.asarray(out[k]).reshape(want.shape)
        scale = max(np.abs(want).max(), 1e-300)
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12 * scale, err_msg=k)


def test_fixture_inputs_match_generators():
    """The stored inputs are the seeded generators' output, so a change of
    generator cannot silently change what the fixtures mean."""
    g = golden()
    for name in cases.CASES:
        prob = cases.problem(name)
        for k, v in synth.problem_arrays(prob, name + "/in/").items():
            np.testing.assert_array_equal(g[k], v, err_msg=k)


def test_non_generator_derivative_matrix_against_reference(oracle, ref):
    """A derivative matrix whose rows do not sum to 0, with missing tip
    states: the tip denominator uses pre[state & 3]
    (BeagleCPU4StateImpl.hpp:1032), which only a reference run pins."""
    prob = synth.make_problem(41, 6, 50, S=4, missing=0.3)
    rng = np.random.default_rng(5)
    D = rng.normal(size=(prob.C, 4, 4)).reshape(-1)
    res = []
    for lib in (oracle, ref):
        ev = driver.Evaluator(lib, prob, transpose="none")
        ev.log_likelihood()
        ev.pre_order()
        ev.inst.set_differential_matrix(ev.d1_index, D)
        res.append(ev.edge_derivatives(False, per_site=True))
        ev.close()
    for a, b in zip(res[0], res[1]):
        np.testing.assert_allclose(np.asarray(a), np.asarray(b), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("S", [4, 20])
def test_remaining_calls_against_reference(oracle, ref, S):
    """sequences.api_sequence on the oracle and on the reference CPU
    library: scaled getPartials, scale-factor bookkeeping (also by
    partition), root log-likelihoods by partition, pre-order operations
    with write-scale buffers (whole and by partition); 4 states go through
    the reference's 4-state kernels, 20 through its generic ones."""
    prob = synth.make_problem(51 + S, 9, 120 if S == 4 else 60, S=S, missing=0.05,
                              mean_length=0.3)
    a = sequences.api_sequence(oracle, prob)
    b = sequences.api_sequence(ref, prob)
    assert set(a) == set(b)
    if S != 4:
        # The reference's generic pre-order kernel clears the whole
        # destination buffer before it computes one partition
        # (BeagleCPUImpl.hpp:4545), so after the partition-1 operations only
        # partition 1 is left; its 4-state kernel and the oracle keep the
        # other partitions.  Compare partition 1 here; partition 0 is pinned
        # against the whole-pattern pass below.
        part1 = (np.arange(prob.C * prob.P * S) // S) % prob.P >= prob.P // 3
        for k in ("pre_by_partition", "pre_by_partition_unscaled"):
            a[k], b[k] = a[k][..., part1], b[k][..., part1]
    for k in a:
        scale = np.abs(b[k]).max()
        np.testing.assert_allclose(a[k], b[k], rtol=1e-12, atol=1e-13 * scale, err_msg=k)


@pytest.mark.parametrize("S", [4, 20])
def test_pre_order_by_partition_equals_whole_pass(oracle, S):
    """The partitioned pre-order pass (every operation once per partition)
    gives the whole-pattern pass's pre-order partials in every partition."""
    prob = synth.make_problem(71 + S, 9, 90, S=S, missing=0.05)
    ev = driver.Evaluator(oracle, prob, transpose="none")
    ev.log_likelihood()
    ops = ev.pre_order()
    nodes = ev.branch_nodes()
    whole = np.stack([ev.inst.get_partials(ev.pre_base + v) for v in nodes])
    parts = np.zeros(prob.P, dtype=np.int32)
    parts[prob.P // 3:] = 1
    ev.inst.set_pattern_partitions(2, parts)
    ev.set_root_pre()
    ev.inst.update_pre_partials_by_partition(
        np.array([list(o) + [q, -1] for q in range(2) for o in ops], np.int32))
    split = np.stack([ev.inst.get_partials(ev.pre_base + v) for v in nodes])
    ev.close()
    np.testing.assert_array_equal(split, whole)


def test_prescaled_derivatives_equal_unscaled(oracle):
    """Per-pattern factors cancel from the derivative ratios: the edge
    derivatives after a pre-order pass that rescales every operation equal
    those of the unscaled pass, and the rescaled pre-order partials have a
    per-pattern maximum of 1 (BeagleCPUImpl.hpp:4008-4056)."""
    prob = synth.make_problem(58, 9, 120, S=4, missing=0.05, mean_length=0.3)
    seq = sequences.api_sequence(oracle, prob)
    ev = driver.Evaluator(oracle, prob, scaling="manual", transpose="none")
    ev.log_likelihood()
    ev.pre_order()
    d1 = ev.edge_derivatives(False, per_site=True)
    d2 = ev.edge_derivatives(True, per_site=False)
    ev.close()
    np.testing.assert_allclose(seq["pre_scaled_d1_sites"], np.asarray(d1[0]).reshape(-1),
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(seq["pre_scaled_d1_sums"],
                               np.concatenate([np.asarray(d1[1]), np.asarray(d1[2])]),
                               rtol=1e-11)
    np.testing.assert_allclose(seq["pre_scaled_d2_sums"], np.asarray(d2[1]), rtol=1e-11)
    x = seq["pre_rescaled"].reshape(4, prob.C, prob.P, prob.S)
    np.testing.assert_allclose(x.max(axis=(1, 3)), 1.0, rtol=0, atol=1e-15)


@pytest.mark.parametrize("S", [4, 20, 61])
def test_transition_derivatives_against_reference(oracle, ref, S):
    """beagleUpdateTransitionMatrices with first- and second-derivative
    index lists (EigenDecompositionCube.hpp:105-322): P, dP/dt and d2P/dt2
    of the oracle equal the reference CPU library's, and dP/dt = r Q P,
    d2P/dt2 = r^2 Q^2 P (closed forms of the exponential)."""
    prob = synth.make_problem(60 + S, 6, 20, S=S)
    lengths = [0.0, 0.01, 0.1, 0.7]
    a = sequences.transition_derivatives(oracle, prob, lengths)
    b = sequences.transition_derivatives(ref, prob, lengths)
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=0, atol=1e-14 * max(1.0, np.abs(y).max()))
    Q = prob.model.Q
    for u, t in enumerate(lengths):
        for c, r in enumerate(prob.rates):
            P = a[0][u, c]
            np.testing.assert_allclose(a[1][u, c], r * Q @ P, rtol=0,
                                       atol=1e-11 * max(1.0, r * np.abs(Q).max()))
            np.testing.assert_allclose(a[2][u, c], r * r * (Q @ Q) @ P, rtol=0,
                                       atol=1e-10 * max(1.0, (r * np.abs(Q).max()) ** 2))


def reordered_partition_run(lib, prob, parts, tips_as, **kw):
    """Non-contiguous pattern partitions (BeagleCPUImpl.hpp:803-920,
    4140-4219): set after the tip data, so the library reorders the
    patterns; then a partitioned post-order and per-partition root
    log-likelihoods, the site log-likelihoods (original order,
    :1009-1021), and a second reorder, which is refused."""
    ev = driver.Evaluator(lib, prob, tips_as=tips_as, transpose="none", preorder=False, **kw)
    inst, tree = ev.inst, prob.tree
    out = {}
    try:
        nparts = int(parts.max()) + 1
        ev.update_matrices()
        inst.set_pattern_partitions(nparts, parts)
        ops = synth.postorder_ops(tree)
        pops = [list(o) + [q, -1] for q in range(nparts) for o in ops]
        inst.update_partials_by_partition(np.array(pops, np.int32))
        r = tree.root
        by, tot = inst.root_log_likelihood_by_partition([r] * nparts, [0] * nparts, [0] * nparts,
                                                        [-1] * nparts, list(range(nparts)))
        out["by_partition"] = np.asarray(by)
        out["total"] = np.array([tot])
        out["sites"] = np.asarray(inst.site_log_likelihoods())
        out["root"] = np.asarray(inst.get_partials(r))
        out["second_reorder_rc"] = np.array([lib.call("SetPatternPartitions", inst.h, nparts,
                                                      np.ascontiguousarray(parts[::-1]).ctypes
                                                      .data_as(ctypes.POINTER(ctypes.c_int)))])
    finally:
        ev.close()
    return out


@pytest.mark.parametrize("S,tips_as", [(4, "states"), (4, "partials"), (20, "states")])
def test_partition_reorder_against_reference(oracle, ref, S, tips_as):
    prob = synth.make_problem(61 + S, 7, 150, S=S, missing=0.05)
    parts = np.random.default_rng(3).integers(0, 3, prob.P).astype(np.int32)
    a = reordered_partition_run(oracle, prob, parts, tips_as)
    b = reordered_partition_run(ref, prob, parts, tips_as)
    assert b["second_reorder_rc"][0] == api.ERROR_NO_IMPLEMENTATION
    for k in a:
        scale = max(np.abs(b[k]).max(), 1e-300)
        np.testing.assert_allclose(a[k], b[k], rtol=1e-12, atol=1e-13 * scale, err_msg=k)


def test_multiple_model_matrices_against_reference(oracle, ref):
    """beagleUpdateTransitionMatricesWithMultipleModels: per-branch eigen
    buffer and category-rate buffer (BeagleCPUImpl.hpp:1240-1290)."""
    rng = np.random.default_rng(7)
    models = [synth.gtr(rng, 4), synth.gtr(rng, 4)]
    rates = [np.array([0.3, 0.9, 1.8]), np.array([0.5, 1.0, 1.5])]
    lengths = np.array([0.05, 0.2, 0.7, 0.01])
    out = []
    for lib in (oracle, ref):
        inst = lib.create(2, 3, 2, 4, 8, 2, 12, 3, 1)
        for e, m in enumerate(models):
            inst.set_eigen(e, m.evec.reshape(-1), m.ivec.reshape(-1), m.evals)
            inst.set_category_rates_with_index(e, rates[e])
        idx = np.arange(4)
        inst.update_transition_matrices_multi([0, 1, 1, 0], [1, 0, 1, 0], idx, idx + 4, idx + 8,
                                              lengths)
        out.append(np.stack([inst.get_transition_matrix(i) for i in range(12)]))
        inst.finalize()
    np.testing.assert_allclose(out[0], out[1], rtol=1e-12, atol=1e-14)


def _matrix_utilities(lib, S):
    """updateTransitionMatricesWithModelCategories (category l uses eigen
    buffer l, rate 1) and convolveTransitionMatrices."""
    rng = np.random.default_rng(11 + S)
    C = 3
    models = [synth.gtr(rng, S) for _ in range(C)]
    inst = lib.create(2, 3, 2, S, 8, C, 12, C, 1)
    try:
        for e, m in enumerate(models):
            inst.set_eigen(e, m.evec.reshape(-1), m.ivec.reshape(-1), m.evals)
        idx = np.arange(3)
        inst.update_transition_matrices_model_categories(list(range(C)), idx, idx + 3, idx + 6,
                                                         np.array([0.03, 0.4, 1.1]))
        inst.convolve_transition_matrices([0, 1], [1, 2], [9, 10])
        return np.stack([inst.get_transition_matrix(i) for i in list(range(9)) + [9, 10]])
    finally:
        inst.finalize()


@pytest.mark.parametrize("S", [4, 20])
def test_matrix_utilities_against_reference(oracle, ref, S):
    a = _matrix_utilities(oracle, S)
    b = _matrix_utilities(ref, S)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)


def test_model_categories_closed_form(oracle):
    """Model categories: category l's P(t) = expm(Q_l t), rate 1; the
    convolution of P(s) and P(t) from one model is P(s + t)."""
    from scipy.linalg import expm
    got = _matrix_utilities(oracle, 4)
    rng = np.random.default_rng(15)
    models = [synth.gtr(rng, 4) for _ in range(3)]
    for u, t in enumerate([0.03, 0.4, 1.1]):
        P = got[u].reshape(3, 4, 4)
        for l, m in enumerate(models):
            np.testing.assert_allclose(P[l], expm(m.Q * t), atol=1e-12)
    conv = got[9].reshape(3, 4, 4)
    for l, m in enumerate(models):
        np.testing.assert_allclose(conv[l], expm(m.Q * 0.43), atol=1e-12)


@pytest.mark.parametrize("S", [4, 20])
def test_complex_eigen_reference_matches_expm(ref, S):
    """BEAGLE_FLAG_EIGEN_COMPLEX in the reference CPU library
    (EigenDecompositionSquare.hpp:108-185) with the real block form that
    synth.nonreversible builds: P(r t) = expm(Q r t), clamped at 0 — pins
    the convention the HIP library's complex path follows."""
    from scipy.linalg import expm
    rng = np.random.default_rng(20 + S)
    m = synth.nonreversible(rng, S)
    assert np.any(m.evals[S:] != 0)
    rates = np.array([0.4, 1.0, 1.6])
    lengths = np.array([0.02, 0.3, 1.1])
    inst = ref.create(2, 3, 2, S, 8, 1, 4, 3, 1, requirement=api.FLAG_EIGEN_COMPLEX)
    try:
        inst.set_eigen(0, m.evec.reshape(-1), m.ivec.reshape(-1), m.evals)
        inst.set_category_rates(rates)
        inst.update_transition_matrices(0, [0, 1, 2], None, None, lengths)
        for u, t in enumerate(lengths):
            got = inst.get_transition_matrix(u).reshape(3, S, S)
            for c, r in enumerate(rates):
                np.testing.assert_allclose(got[c], np.maximum(expm(m.Q * (r * t)), 0.0),
                                           rtol=0, atol=1e-12)
    finally:
        inst.finalize()


@pytest.mark.parametrize("S", [4, 7, 20, 61])
def test_nonreversible_generator(S):
    """synth.nonreversible: Q is a rate matrix (rows sum to 0, off-diagonal
    >= 0) normalised to one substitution at its stationary pi (pi Q = 0),
    has complex eigenvalues, and V B V^-1 = Q with B block-diagonal:
    a on the diagonal, [[a, b], [-b, a]] for each pair (b on the first
    row); expm through the blocks equals scipy's expm."""
    from scipy.linalg import expm
    m = synth.nonreversible(np.random.default_rng(S), S)
    Q = m.Q
    assert np.allclose(Q.sum(axis=1), 0.0, atol=1e-13)
    assert (Q - np.diag(np.diag(Q)) >= 0).all()
    assert abs(m.freqs.sum() - 1.0) < 1e-13 and np.abs(m.freqs @ Q).max() < 1e-12
    assert abs(-(m.freqs * np.diag(Q)).sum() - 1.0) < 1e-12
    re, im = m.evals[:S], m.evals[S:]
    assert np.any(im != 0)
    B = np.diag(re)
    k = 0
    while k < S:
        if im[k] != 0:
            assert im[k + 1] == -im[k] and re[k + 1] == re[k]
            B[k, k + 1], B[k + 1, k] = im[k], -im[k]
            k += 2
        else:
            k += 1
    np.testing.assert_allclose(m.evec @ B @ m.ivec, Q, rtol=0, atol=1e-12)
    np.testing.assert_allclose(m.evec @ expm(0.7 * B) @ m.ivec, expm(0.7 * Q), rtol=0,
                               atol=1e-12)


@pytest.mark.parametrize("S,tips_state_child", [(4, True), (20, False)])
def test_edge_log_likelihood_by_partition_against_reference(oracle, ref, S, tips_state_child):
    """Edge log-likelihoods by partition on the oracle and the reference CPU
    library (values per partition and totals, with and without
    derivatives, site log-likelihoods), and the partitions' sums of one edge
    equal the whole-pattern edge call."""
    n = 3 if tips_state_child else 9      # 3 taxa: a root child is a tip
    prob = synth.make_problem(81 + S, n, 150 if S == 4 else 70, S=S, missing=0.05,
                              mean_length=0.3)
    a = sequences.edge_by_partition(oracle, prob)
    b = sequences.edge_by_partition(ref, prob)
    for k in a:
        np.testing.assert_allclose(a[k], b[k], rtol=1e-12, atol=1e-13 * np.abs(b[k]).max(),
                                   err_msg=k)
    np.testing.assert_allclose(a["without_derivatives"][-1], a["whole_edge"][0], rtol=1e-12)
