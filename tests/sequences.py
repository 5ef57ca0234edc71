"""Call sequences shared by the oracle pins (oracle against the reference CPU
library, tests/test_oracle.py) and the GPU parity tests (HIP library
against the oracle, tests/test_gpu_parity.py): one function drives any
library that exports the BEAGLE API and returns the arrays to compare.
"""
import numpy as np

from beagle_amd import driver, synth


def seq_states(seq):
    return np.array(["ACGT".index(ch) if ch in "ACGT" else 4 for ch in seq], dtype=np.int32)


def small_problem(spec):
    """A Problem from a known-answer entry: three tips, ((0,1)3,2)4."""
    seqs = spec["sequences"]
    n = len(seqs)
    children = {n + i: tuple(c) for i, c in enumerate(spec["tree"]["children"])}
    parent = np.full(2 * n - 1, -1)
    for v, (a, b) in children.items():
        parent[a] = v
        parent[b] = v
    lengths = np.zeros(2 * n - 1)
    lengths[:2 * n - 2] = spec["tree"]["lengths"]
    tree = synth.Tree(n=n, children=children, parent=parent, lengths=lengths)
    if spec["model"] == "JC69":
        model = synth.jc69()
    else:  # an explicit eigen-system, Q = evec diag(evals) ivec
        e = spec["eigen"]
        V, Vi, lam = np.array(e["evec"]), np.array(e["ivec"]), np.array(e["evals"])
        model = synth.Model(Q=V @ np.diag(lam) @ Vi, freqs=np.array(spec["freqs"]), evec=V,
                            ivec=Vi, evals=lam)
    rates = np.array(spec["rates"])
    return synth.Problem(tree=tree, model=model, rates=rates,
                         weights=np.full(len(rates), 1.0 / len(rates)),
                         tips=np.stack([seq_states(s) for s in seqs]),
                         pattern_weights=np.ones(len(seqs[0])))


def api_sequence(lib, prob, precision="double", transpose="none", matrices="eigen"):
    """Calls the golden fixtures do not hold.

    - getPartials with a cumulative scale buffer (BeagleCPUImpl.hpp:743-754);
    - remove / accumulate / copy of scale factors (:2763-2908) followed by a
      root log-likelihood over the rebuilt cumulative buffer;
    - pre-order operations that rescale into write-scale buffers
      (upPrePartials, :1771-1823, the pattern of
      examples/hmctest/hmctest.cpp:503-506): the rescaled pre-order
      partials, each operation's scale factors exposed through getPartials
      with a cumulative buffer that holds only them, and the edge
      derivatives, which the per-pattern factors cancel out of;
    - the ByPartition scale-factor calls (:2804-2900): reset, accumulate
      and remove restricted to one partition each, then root
      log-likelihoods by partition (:2349-2388) over the result;
    - the partitioned pre-order pass (:1415-1433) with write-scale buffers
      (rescalePartialsByPartition, :4059-4102).

    Buffer numbering as driver.Evaluator with pre_scaling: the scale
    buffer of internal node v is v - n, the cumulative buffer n - 1,
    pre-order operation k writes scale buffer n + k.  transpose="auto"
    sets BEAGLE_FLAG_PREORDER_TRANSPOSE_AUTO, under which a GPU library takes
    untransposed pre-order matrices for S > 4 as the CPU libraries do;
    matrices="explicit" sets identical P(t) on every library."""
    ev = driver.Evaluator(lib, prob, scaling="manual", transpose=transpose, precision=precision,
                          pre_scaling=True, matrices=matrices)
    inst, tree, n = ev.inst, prob.tree, prob.tree.n
    out = {}
    try:
        ev.log_likelihood()
        cum = n - 1
        out["scaled_root"] = inst.get_partials(tree.root, cum)
        inst.remove_scale_factors([1, 2], cum)
        inst.accumulate_scale_factors([1, 2], cum)
        inst.copy_scale_factors(1, 0)              # buffer 1 := buffer 0
        inst.accumulate_scale_factors([1], cum)
        inst.remove_scale_factors([0], cum)
        out["cumulative_logL"] = np.array([inst.root_log_likelihood(tree.root, 0, 0, cum)])

        # pre-order with write-scale buffers
        ops = ev.pre_order()
        picks = [0, 1, len(ops) // 2, len(ops) - 1]
        out["pre_rescaled"] = np.stack([inst.get_partials(int(ops[k, 0])) for k in picks])
        unscaled = []
        for k in picks:
            inst.reset_scale_factors(cum)
            inst.accumulate_scale_factors([int(ops[k, 1])], cum)
            unscaled.append(inst.get_partials(int(ops[k, 0]), cum))
        out["pre_unscaled"] = np.stack(unscaled)
        d1 = ev.edge_derivatives(False, per_site=True)
        d2 = ev.edge_derivatives(True, per_site=False)
        out["pre_scaled_d1_sites"] = np.asarray(d1[0]).reshape(-1)
        out["pre_scaled_d1_sums"] = np.concatenate([np.asarray(d1[1]), np.asarray(d1[2])])
        out["pre_scaled_d2_sums"] = np.asarray(d2[1])

        # scale factors by partition, then root log-likelihood by partition
        ev.post_order()                            # cumulative buffer rebuilt
        parts = np.zeros(prob.P, dtype=np.int32)
        parts[prob.P // 3:] = 1
        inst.set_pattern_partitions(2, parts)
        internal = [v - n for v in range(n, 2 * n - 1)]
        inst.reset_scale_factors_by_partition(cum, 0)
        inst.accumulate_scale_factors_by_partition(internal, cum, 0)   # part 0 rebuilt
        inst.remove_scale_factors_by_partition(internal[:2], cum, 1)   # part 1 changed
        by, tot = inst.root_log_likelihood_by_partition([tree.root] * 2, [0, 0], [0, 0],
                                                        [cum, cum], [0, 1])
        out["by_partition_logL"] = np.concatenate([by, [tot]])
        out["by_partition_cumulative"] = inst.get_partials(tree.root, cum)
        inst.reset_scale_factors_by_partition(cum, 1)
        by, tot = inst.root_log_likelihood_by_partition([tree.root] * 2, [0, 0], [0, 0],
                                                        [cum, cum], [0, 1])
        out["by_partition_reset_logL"] = np.concatenate([by, [tot]])

        # partitioned pre-order, rescaled into write-scale buffers
        ev.set_root_pre()
        pops = [list(o) + [q, -1] for q in range(2) for o in ops]
        inst.update_pre_partials_by_partition(np.array(pops, np.int32))
        out["pre_by_partition"] = np.stack([inst.get_partials(ev.pre_base + v)
                                            for v in ev.branch_nodes()[:3]])
        inst.reset_scale_factors(cum)
        inst.accumulate_scale_factors([int(ops[2, 1])], cum)
        out["pre_by_partition_unscaled"] = inst.get_partials(int(ops[2, 0]), cum)
    finally:
        ev.close()
    return out


def transition_derivatives(lib, prob, lengths, precision="double"):
    """P(t), dP/dt and d2P/dt2 from beagleUpdateTransitionMatrices with all
    three index lists (EigenDecompositionCube.hpp:105-322), for the given
    lengths: returns three arrays (len(lengths), C, S, S)."""
    ev = driver.Evaluator(lib, prob, preorder=False, precision=precision)
    try:
        m = len(lengths)
        if ev.n_matrices < 3 * m:
            raise ValueError("too many lengths for the instance's matrices")
        p, d1, d2 = np.arange(m), np.arange(m, 2 * m), np.arange(2 * m, 3 * m)
        ev.inst.update_transition_matrices(0, p, d1, d2, np.asarray(lengths, float))
        get = ev.inst.get_transition_matrix
        shape = (prob.C, prob.S, prob.S)
        res = [np.stack([get(int(i)).reshape(shape) for i in idx]) for idx in (p, d1, d2)]
    finally:
        ev.close()
    return res


def edge_by_partition(lib, prob, precision="double", matrices="eigen"):
    """beagleCalculateEdgeLogLikelihoodsByPartition
    (BeagleCPUImpl.hpp:3008-3110, 3452-3590) over three uneven contiguous
    partitions, after a post-order pass with manual scaling: the root's
    post-order partials as the parent and, per partition, one of the
    root's children (its own branch matrix, the derivative matrices r_c Q
    and r_c^2 Q^2) as the child, with the cumulative scale buffer; with and
    without derivatives; the site log-likelihoods after it; and the same
    edge over all patterns through beagleCalculateEdgeLogLikelihoods."""
    ev = driver.Evaluator(lib, prob, scaling="manual", transpose="none", precision=precision,
                          matrices=matrices, preorder=False)
    inst, tree, n = ev.inst, prob.tree, prob.tree.n
    out = {}
    try:
        ev.log_likelihood()
        cum = n - 1
        root = tree.root
        a, b = tree.children[root]
        parts = np.zeros(prob.P, dtype=np.int32)
        parts[prob.P // 5:] = 1
        parts[prob.P // 2:] = 2
        inst.set_pattern_partitions(3, parts)
        ch = [a, b, a]
        d1 = [ev.d1_index] * 3
        d2 = [ev.d2_index] * 3
        res = inst.edge_log_likelihood_by_partition([root] * 3, ch, ch, d1, d2, [0] * 3, [0] * 3,
                                                    [cum] * 3, [0, 1, 2])
        out["with_derivatives"] = np.concatenate([res[0], [res[1]], res[2], [res[3]], res[4],
                                                  [res[5]]])
        out["sites_after"] = inst.site_log_likelihoods()
        res = inst.edge_log_likelihood_by_partition([root] * 3, [a] * 3, [a] * 3, None, None,
                                                    [0] * 3, [0] * 3, [cum] * 3, [2, 0, 1])
        out["without_derivatives"] = np.concatenate([res[0], [res[1]]])
        whole = inst.edge_log_likelihood(root, a, a, ev.d1_index, ev.d2_index, cumulative=cum)
        out["whole_edge"] = np.array(whole)
    finally:
        ev.close()
    return out
