"""Runs a whole likelihood (and adjoint) evaluation of a synthetic problem
through any library that exports the BEAGLE API, issuing the calls a
client such as BEAST or the reference's examples/hmctest makes:

  set tips / eigen-system / frequencies / category weights and rates /
  pattern weights -> update transition matrices -> post-order partials ->
  (accumulate scale factors) -> root log-likelihood -> root pre-order
  partials = frequencies -> pre-order partials -> edge derivatives with
  D = r_c Q and D = r_c^2 Q^2 -> cross products.

Buffer numbering: tips 0..n-1, post-order partials of internal node v at
v (n..2n-2), pre-order partials of node v at 2n-1+v.  Matrix of the
branch above v is v; the two derivative matrices follow at 2n-1 and 2n
(their transposes, for the S>4 manual-transpose convention, at 2n+1, 2n+2).
The scale buffer of internal node v is v-n; the cumulative buffer is n-1.
With pre_scaling, pre-order operation k rescales its output into scale
buffer n+k (the pattern of examples/hmctest/hmctest.cpp:503-506).
"""
import dataclasses

import numpy as np

from . import api, synth


def explicit_matrices(prob, lengths):
    from scipy.linalg import expm
    out = np.empty((len(lengths), prob.C, prob.S, prob.S))
    for u, t in enumerate(lengths):
        for c, r in enumerate(prob.rates):
            out[u, c] = np.maximum(expm(prob.model.Q * (r * t)), 0.0)
    return out


class Evaluator:
    def __init__(self, lib, prob, *, tips_as="states", scaling="none", preorder=True,
                 precision="double", extra_flags=0, resources=None, transpose="auto",
                 matrices="eigen", pre_scaling=False):
        self.lib = lib
        self.prob = prob
        self.precision = precision
        self.scaling = scaling
        self.preorder = preorder
        self.transpose = transpose
        self.matrices = matrices
        self.pre_scaling = pre_scaling
        n = prob.tree.n
        self.n = n
        nodes = 2 * n - 1
        self.pre_base = nodes
        pre = nodes if preorder else 0
        if tips_as == "states":
            compact = n
            partials = (n - 1) + pre
        else:
            compact = 0
            partials = nodes + pre
        self.tips_as = tips_as
        # transpose="manual" (S > 4): the reference GPU convention, pre-order
        # self and derivative matrices passed transposed; the branch
        # matrices' transposes live at t_offset + index
        self.manual_transpose = transpose == "manual" and prob.S > 4
        self.t_offset = nodes + 4
        self.n_matrices = nodes + 4 + (nodes if self.manual_transpose else 0)
        self.d1_index = nodes
        self.d2_index = nodes + 1
        pref = api.FLAG_EIGEN_REAL
        req = 0
        if precision == "double":
            req |= api.FLAG_PRECISION_DOUBLE
        else:
            req |= api.FLAG_PRECISION_SINGLE
        if scaling == "manual_log":
            pref |= api.FLAG_SCALERS_LOG
        elif scaling == "always":
            pref |= api.FLAG_SCALING_ALWAYS
        elif scaling == "auto":
            pref |= api.FLAG_SCALING_AUTO
        if transpose == "auto":
            pref |= api.FLAG_PREORDER_TRANSPOSE_AUTO
        pref |= extra_flags
        scale_buffers = n + (2 * n - 2 if pre_scaling else 0)
        self.inst = lib.create(n, partials, compact, prob.S, prob.P, 1, self.n_matrices, prob.C,
                               scale_buffers, resources=resources, preference=pref,
                               requirement=req)
        self._setup()
        # operation lists depend on the topology only: built once, as a
        # client builds them once per tree
        self._branch_idx = np.arange(2 * n - 2, dtype=np.int32)
        self._post_ops = {}
        self._pre = None

    def set_tree(self, tree):
        """Another topology over the same tips (a sampler's proposal): the
        operation lists are rebuilt on next use."""
        if tree.n != self.n:
            raise ValueError("the new tree must have the same tips")
        self.prob = dataclasses.replace(self.prob, tree=tree)
        self._post_ops = {}
        self._pre = None
        self.__dict__.pop("_nodes", None)

    # -- data ---------------------------------------------------------------
    def _setup(self):
        p, inst = self.prob, self.inst
        for t in range(self.n):
            if self.tips_as == "states":
                inst.set_tip_states(t, p.tips[t])
            else:
                inst.set_tip_partials(t, synth.states_to_partials(p.tips[t], p.S).reshape(-1))
        m = p.model
        inst.set_eigen(0, m.evec.reshape(-1), m.ivec.reshape(-1), m.evals)
        inst.set_state_frequencies(0, m.freqs)
        inst.set_category_weights(0, p.weights)
        inst.set_category_rates(p.rates)
        inst.set_pattern_weights(p.pattern_weights)
        # derivative matrices r_c Q and r_c^2 Q^2 with pad 0
        # (examples/hmctest/hmctest.cpp:398-427)
        Q = m.Q
        d1 = np.concatenate([(r * Q).reshape(-1) for r in p.rates])
        d2 = np.concatenate([(r * r * (Q @ Q)).reshape(-1) for r in p.rates])
        inst.set_differential_matrix(self.d1_index, d1)
        inst.set_differential_matrix(self.d2_index, d2)
        if self.manual_transpose:
            inst.transpose_transition_matrices([self.d1_index, self.d2_index],
                                               [self.d1_index + 2, self.d2_index + 2])

    def _scale_write(self, v):
        return v - self.n

    # -- evaluation ---------------------------------------------------------
    def update_matrices(self, lengths=None):
        tree = self.prob.tree
        if lengths is None:
            lengths = tree.lengths
        idx = self._branch_idx
        if self.matrices == "explicit":
            # identical P(t) on every library: exp(Q r t) from scipy, clamped
            # at 0, set with pad 1 (beagleSetTransitionMatrices)
            ms = explicit_matrices(self.prob, lengths[:2 * self.n - 2])
            self.inst.set_transition_matrices(idx, ms.reshape(-1), np.ones(len(idx)))
        else:
            self.inst.update_transition_matrices(0, idx, None, None, lengths[:2 * self.n - 2])
        if self.manual_transpose:
            # pre-order "self" matrices are passed transposed (the reference
            # GPU convention for S > 4, examples/hmctest/hmctest.cpp:495)
            self.inst.transpose_transition_matrices(idx, np.asarray(idx) + self.t_offset)

    def post_order(self):
        tree = self.prob.tree
        manual = self.scaling in ("manual", "manual_log")
        if manual not in self._post_ops:
            self._post_ops[manual] = synth.postorder_ops(tree, self._scale_write if manual else None)
        ops = self._post_ops[manual]
        cum = self.n - 1
        if manual:
            self.inst.reset_scale_factors(cum)
            self.inst.update_partials(ops, cum)
        else:
            self.inst.update_partials(ops, api.OP_NONE)
        return ops

    def root(self):
        tree = self.prob.tree
        manual = self.scaling in ("manual", "manual_log")
        cum = self.n - 1 if manual else api.OP_NONE
        return self.inst.root_log_likelihood(tree.root, 0, 0, cum)

    def log_likelihood(self, lengths=None):
        self.update_matrices(lengths)
        self.post_order()
        return self.root()

    def set_root_pre(self):
        """Root pre-order partials = state frequencies (the reference's CPU
        path has beagleSetRootPrePartials; its GPU path does not, and
        examples/hmctest/hmctest.cpp:550-559 uses beagleSetPartials)."""
        root = self.pre_base + self.prob.tree.root
        if self.lib.prefix == "beagle" and self.lib.path.endswith("libhmsbeagle_amd.so"):
            self.inst.set_root_pre_partials([root], [0])
        else:
            self.inst.set_partials(root, np.tile(self.prob.model.freqs, self.prob.C * self.prob.P))

    def preorder_operations(self):
        ops = self._transposed_self(synth.preorder_ops(self.prob.tree, self.pre_base))
        if self.pre_scaling:
            ops = np.array(ops, copy=True)
            ops[:, 1] = self.n + np.arange(len(ops))
        return ops

    def pre_order(self):
        self.set_root_pre()
        ops = self.preorder_operations()
        self.inst.update_pre_partials(ops, api.OP_NONE)
        return ops

    def _transposed_self(self, ops):
        if self.manual_transpose:
            ops = np.array(ops, copy=True)
            ops[:, 4] += self.t_offset
        return ops

    def _deriv_index(self, second):
        d = self.d2_index if second else self.d1_index
        return d + 2 if self.manual_transpose else d

    def fused_operations(self, keep_tips=False):
        """Pre-order operations and matching post-order buffers for the
        derivative extension; for 4 states, and 5-64 states in single
        precision with tips as states, the tips' operations have no
        destination unless keep_tips (no later operation reads them)."""
        key = bool(keep_tips)
        if self._pre is None or self._pre[0] != key:
            ops = self._transposed_self(synth.preorder_ops(self.prob.tree, self.pre_base))
            post = (ops[:, 0] - self.pre_base).astype(np.int32)
            S, C = self.prob.S, self.prob.C
            if not keep_tips and C <= 8 and (S == 4 or (4 < S <= 64 and self.precision == "single"
                                                        and self.tips_as == "states")):
                ops[post < self.n, 0] = api.OP_NONE
            n = len(post)
            self._pre = (key, ops, post, np.full(n, self._deriv_index(False), np.int32),
                         np.full(n, self._deriv_index(True), np.int32))
        return self._pre[1:]

    def fused_gradient(self, lengths=None, keep_tips=False, rate_matrix=False):
        """logL, dlogL/dt and d2logL/dt2 per branch through the fused
        pre-order + derivative extensions (this package's library only):
        matrices and the post-order traversal, then one call for the root
        log-likelihood, root pre-order partials and the derivative pass;
        with rate_matrix also dlogL/dQ and dlogL/dpi_root (True: in the same
        call, beagleExtCalculateLogLikelihoodAndGradient; "separate": a
        second call, beagleExtCalculateRateMatrixGradient)."""
        ops, post, d1, d2 = self.fused_operations(keep_tips)
        self.update_matrices(lengths)
        self.post_order()
        manual = self.scaling in ("manual", "manual_log")
        cum = self.n - 1 if manual else api.OP_NONE
        root, root_pre = self.prob.tree.root, self.pre_base + self.prob.tree.root
        if rate_matrix is True:
            logl, s1, sq, s2, q, pi = self.inst.log_likelihood_and_gradient(
                root, root_pre, ops, post, d1, d2, cumulative=cum)
        else:
            logl, s1, sq, s2 = self.inst.log_likelihood_and_derivatives(
                root, root_pre, ops, post, d1, d2, cumulative=cum)
        g1 = np.zeros(2 * self.n - 1)
        g2 = np.zeros(2 * self.n - 1)
        g1[post] = s1
        g2[post] = s2 - sq
        nodes = self.branch_nodes()
        if rate_matrix == "separate":
            q, pi = self.rate_matrix_gradient()
        if rate_matrix:
            return logl, g1[nodes], g2[nodes], q, pi
        return logl, g1[nodes], g2[nodes]

    def checkpointed_gradient(self, buffers, interval=0, lengths=None, plan_only=False):
        """logL, dlogL/dt and d2logL/dt2 per branch (as fused_gradient())
        through beagleExtCheckpointedGradient, using only the partials
        buffers in `buffers`: post-order partials kept every `interval`
        heights (0: the smallest interval that fits) and the rest recomputed.
        Returns (logL, d1, d2, interval, peak buffers in use); plan_only:
        (None, None, None, interval, peak)."""
        if self.scaling in ("manual", "manual_log", "always"):
            raise ValueError("the checkpointed gradient takes automatic scaling or none")
        tree = self.prob.tree
        ops = synth.postorder_ops(tree, None)
        size = 2 * self.n - 1
        d1 = np.full(size, self.d1_index, dtype=np.int32)
        d2 = np.full(size, self.d2_index, dtype=np.int32)
        if not plan_only:
            self.update_matrices(lengths)
        lnl, s1, sq, s2, k, peak = self.inst.checkpointed_gradient(
            ops, d1, d2, buffers, interval, plan_only=plan_only)
        if plan_only:
            return None, None, None, k, peak
        nodes = self.branch_nodes()
        return lnl, s1[nodes], (s2 - sq)[nodes], k, peak

    def capture_fused_gradient(self, keep_tips=False, rate_matrix=False):
        """Record one fused_gradient() evaluation (matrices from the
        eigen-system, post-order, root logL, fused derivative pass and, with
        rate_matrix, dlogL/dQ and dlogL/dpi_root) as a HIP graph of this
        package's library; replay_fused_gradient(lengths) reruns it with new
        branch lengths.  The calls run once first so that the library's work
        buffers have their size."""
        if self.matrices != "eigen":
            raise ValueError("graphs record matrices computed from the eigen-system")
        self.fused_gradient(keep_tips=keep_tips, rate_matrix=rate_matrix)
        ops, post, d1, d2 = self.fused_operations(keep_tips)
        S = self.prob.S
        n = len(post)
        self._gout = (np.zeros(1), np.zeros(n), np.zeros(n), np.zeros(n))
        self._qout = (np.zeros(S * S), np.zeros(S)) if rate_matrix else None
        self._gpost = post
        self._gperm = None
        manual = self.scaling in ("manual", "manual_log")
        cum = self.n - 1 if manual else api.OP_NONE
        self.inst.begin_capture()
        try:
            self.update_matrices()
            self.post_order()
            root, root_pre = self.prob.tree.root, self.pre_base + self.prob.tree.root
            if rate_matrix is True:
                self.inst.log_likelihood_and_gradient(root, root_pre, ops, post, d1, d2,
                                                      cumulative=cum,
                                                      out=self._gout + self._qout)
            else:
                self.inst.log_likelihood_and_derivatives(root, root_pre, ops, post, d1, d2,
                                                         cumulative=cum, out=self._gout)
            if rate_matrix == "separate":
                self.inst.rate_matrix_gradient(0, ops, post, root=root, out=self._qout)
        finally:
            self._graph = self.inst.end_capture()
        return self._graph

    def replay_fused_gradient(self, lengths=None, wait=True):
        """One replay of the captured evaluation; returns what
        fused_gradient() returns (wait=False: launch only, collect with
        collect_fused_gradient())."""
        if lengths is not None:
            self.inst.set_graph_branch_lengths(self._graph, np.asarray(lengths)[:2 * self.n - 2])
        self.inst.launch_graph(self._graph)
        return self.collect_fused_gradient() if wait else None

    def collect_fused_gradient(self):
        """(logL, dlogL/dt, d2logL/dt2) of the last replay, and with a
        captured rate-matrix gradient also (dlogL/dQ, dlogL/dpi_root)."""
        self.inst.wait_graph(self._graph)
        lnl, s1, sq, s2 = self._gout
        if getattr(self, "_gperm", None) is None:
            # position in the operation list of each branch's node
            at = np.empty(2 * self.n - 1, dtype=np.int64)
            at[self._gpost] = np.arange(len(self._gpost))
            self._gperm = at[np.asarray(self.branch_nodes())]
        perm = self._gperm
        res = (float(lnl[0]), s1[perm], (s2 - sq)[perm])
        if self._qout is not None:
            S = self.prob.S
            res += (self._qout[0].reshape(S, S).copy(), self._qout[1].copy())
        return res

    def rate_matrix_gradient(self, root_frequencies=True):
        """dlogL/dQ (S x S) for eigen-system 0 and, optionally, dlogL/dpi
        through the root distribution, from the current post- and
        pre-order partials (after fused_gradient() or pre_order()); this
        package's library only."""
        ops, post, _, _ = self.fused_operations()
        root = self.prob.tree.root if root_frequencies else None
        return self.inst.rate_matrix_gradient(0, ops, post, root=root)

    def branch_nodes(self):
        if not hasattr(self, "_nodes"):
            self._nodes = [v for v in range(2 * self.n - 1) if v != self.prob.tree.root]
        return self._nodes

    def edge_derivatives(self, second=False, per_site=True):
        nodes = self.branch_nodes()
        d = self._deriv_index(second)
        return self.inst.edge_derivatives(nodes, [self.pre_base + v for v in nodes],
                                          [d] * len(nodes), per_site=per_site)

    def cross_products(self):
        nodes = self.branch_nodes()
        return self.inst.cross_products(nodes, [self.pre_base + v for v in nodes], 0, 0,
                                        self.prob.tree.lengths[nodes])

    def gradient(self, lengths=None):
        """logL, dlogL/dt per branch, d2logL/dt2 per branch."""
        logl = self.log_likelihood(lengths)
        self.pre_order()
        _, g1, sq1 = self.edge_derivatives(False, per_site=False)
        _, g2, _ = self.edge_derivatives(True, per_site=False)
        return logl, g1, g2 - sq1

    def close(self):
        self.inst.finalize()
