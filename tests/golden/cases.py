"""Golden-vector cases: the BEAGLE call sequence run on a small problem and
the outputs recorded.  make_golden.py runs these cases through the
reference's CPU library (built from its sources into oracle/_ref) and
stores inputs and outputs in reference_outputs.npz; the oracle tests
replay them through the oracle and compare.

Each case names a problem (seeded generator arguments) and the options of
the run: tip representation, scaling mode, state count, partitions.
"""
import numpy as np

from beagle_amd import api, driver, synth

CASES = {
    # name: (make_problem kwargs, Evaluator kwargs, partitions)
    "s4_yule_none_states": (dict(seed=11, n_taxa=10, P=300, S=4, shape="yule", missing=0.05),
                            dict(tips_as="states", scaling="none"), 0),
    "s4_balanced_manual_partials": (dict(seed=12, n_taxa=12, P=250, S=4, shape="balanced",
                                         missing=0.05),
                                    dict(tips_as="partials", scaling="manual"), 0),
    "s4_caterpillar_manual_log": (dict(seed=13, n_taxa=16, P=200, S=4, shape="caterpillar",
                                       missing=0.05, mean_length=0.4),
                                  dict(tips_as="states", scaling="manual_log"), 0),
    "s4_yule_always": (dict(seed=14, n_taxa=9, P=180, S=4, shape="yule", missing=0.05),
                       dict(tips_as="states", scaling="always"), 0),
    "s20_yule_none": (dict(seed=15, n_taxa=8, P=60, S=20, shape="yule", missing=0.05),
                      dict(tips_as="states", scaling="none"), 0),
    "s61_yule_manual": (dict(seed=16, n_taxa=5, P=24, S=61, shape="yule", missing=0.05),
                        dict(tips_as="states", scaling="manual"), 0),
    "s4_partitions": (dict(seed=17, n_taxa=8, P=240, S=4, shape="yule", missing=0.05),
                      dict(tips_as="states", scaling="none"), 3),
}

DERIV_NODES = 3   # per-site derivative rows kept for the first branches


def problem(name):
    kw = dict(CASES[name][0])
    seed, n, P = kw.pop("seed"), kw.pop("n_taxa"), kw.pop("P")
    return synth.make_problem(seed, n, P, **kw)


def run(lib, prob, opts, partitions, extra_flags=0, transpose="none"):
    """The call sequence of one case; returns a dict of output arrays.
    A GPU library takes transpose="auto" (PREORDER_TRANSPOSE_AUTO) to get
    the CPU's untransposed pre-order convention for S > 4."""
    ev = driver.Evaluator(lib, prob, transpose=transpose, extra_flags=extra_flags, **opts)
    inst, tree, n = ev.inst, prob.tree, prob.tree.n
    out = {}
    try:
        out["logL"] = np.array([ev.log_likelihood()])
        out["site_logL"] = inst.site_log_likelihoods()
        # (scale buffers are pinned through the site log-likelihoods: the
        # reference CPU beagleGetScaleFactors copies nothing,
        # BeagleCPUImpl.hpp:2911-2915)
        out["root_partials"] = inst.get_partials(tree.root)
        ev.pre_order()
        nodes = ev.branch_nodes()
        out["pre_partials"] = np.stack([inst.get_partials(ev.pre_base + v)
                                        for v in nodes[:2]])
        d1, s1, q1 = ev.edge_derivatives(False, per_site=True)
        _, s2, _ = ev.edge_derivatives(True, per_site=False)
        out["site_d1"] = np.asarray(d1).reshape(len(nodes), -1)[:DERIV_NODES]
        out["sum_d1"], out["sumsq_d1"], out["sum_d2"] = s1, q1, s2
        out["cross_products"] = ev.cross_products()
        v = nodes[0]
        parent = ev.pre_base + v
        lnl, e1, e2 = inst.edge_log_likelihood(parent, v, v, ev.d1_index, ev.d2_index)
        out["edge_logL"] = np.array([lnl, e1, e2])
        if partitions:
            bounds = np.linspace(0, prob.P, partitions + 1).astype(int)
            parts = np.zeros(prob.P, dtype=np.int32)
            for q in range(partitions):
                parts[bounds[q]:bounds[q + 1]] = q
            inst.set_pattern_partitions(partitions, parts)
            ops = synth.postorder_ops(tree)
            pops = [list(o) + [q, -1] for q in range(partitions) for o in ops]
            inst.update_partials_by_partition(np.array(pops, np.int32))
            by, tot = inst.root_log_likelihood_by_partition(
                [tree.root] * partitions, [0] * partitions, [0] * partitions,
                [-1] * partitions, list(range(partitions)))
            out["partition_logL"] = np.append(np.asarray(by), tot)
    finally:
        ev.close()
    return out
