"""Error paths of the HIP library through the C ABI on cuda:0, each forced
from the host (no kernel faults or hangs): NaN inputs give
BEAGLE_ERROR_FLOATING_POINT with the NaN still written (beagle.h:110-122,
BeagleCPUImpl.hpp:2294-2346), out-of-range buffer, matrix and scale indices
give BEAGLE_ERROR_OUT_OF_RANGE before anything is launched, a partials
allocation larger than the GPU's memory gives BEAGLE_ERROR_OUT_OF_MEMORY and
leaves the library usable, and the oracle returns the same codes."""
import numpy as np
import pytest

from beagle_amd import api, driver, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["double", "single"])
def test_nan_input_gives_floating_point_error(amd, oracle, precision):
    prob = synth.make_problem(901, 6, 200, S=4)
    codes = []
    for lib, prec in ((amd, precision), (oracle, "double")):
        ev = driver.Evaluator(lib, prob, tips_as="partials", precision=prec, preorder=False)
        x = synth.states_to_partials(prob.tips[0], 4).reshape(-1).copy()
        x[17] = np.nan
        ev.inst.set_tip_partials(0, x)
        ev.update_matrices()
        ev.post_order()
        out = np.zeros(1)
        rc = ev.inst.root_log_likelihood(prob.tree.root, out=out,
                                         allow=(api.ERROR_FLOATING_POINT,))
        codes.append(rc)
        assert np.isnan(out[0])
        sites = ev.inst.site_log_likelihoods()
        assert np.isnan(sites[17 // 4]) and np.isfinite(np.delete(sites, 17 // 4)).all()
        ev.close()
    assert codes == [api.ERROR_FLOATING_POINT] * 2


def test_nan_pattern_weight_in_edge_and_partitions(amd):
    """A NaN pattern weight: the edge log-likelihood and the root
    log-likelihood by partition (the partition holding it) report
    BEAGLE_ERROR_FLOATING_POINT."""
    prob = synth.make_problem(902, 6, 300, S=4)
    w = prob.pattern_weights.copy()
    w[250] = np.nan
    ev = driver.Evaluator(amd, prob, preorder=False)
    ev.inst.set_pattern_weights(w)
    ev.update_matrices()
    ev.post_order()
    a, _ = prob.tree.children[prob.tree.root]
    with pytest.raises(api.BeagleError) as e:
        ev.inst.edge_log_likelihood(prob.tree.root, a, a)
    assert e.value.code == api.ERROR_FLOATING_POINT
    parts = (np.arange(prob.P) >= 200).astype(np.int32)
    ev.inst.set_pattern_partitions(2, parts)
    with pytest.raises(api.BeagleError) as e:
        ev.inst.root_log_likelihood_by_partition([prob.tree.root] * 2, [0, 0], [0, 0],
                                                 [api.OP_NONE] * 2, [0, 1])
    assert e.value.code == api.ERROR_FLOATING_POINT
    ev.close()


def test_out_of_range_indices(amd, oracle):
    """Bad buffer, matrix and scale indices in operations and likelihood
    calls are refused with BEAGLE_ERROR_OUT_OF_RANGE by the HIP library and
    the oracle alike, and the instance still evaluates afterwards."""
    prob = synth.make_problem(903, 8, 300, S=4)
    for lib in (amd, oracle):
        ev = driver.Evaluator(lib, prob, scaling="manual")
        want = ev.log_likelihood()
        inst, n = ev.inst, ev.n
        nb = 2 * n - 1 + 2 * n - 1          # partials buffers incl. pre-order
        good = np.array(synth.postorder_ops(prob.tree, ev._scale_write), np.int32)
        for col, bad in ((0, nb + 3), (3, -2), (4, 10_000), (5, nb), (1, 10 * n)):
            ops = good.copy()
            ops[0, col] = bad
            with pytest.raises(api.BeagleError) as e:
                inst.update_partials(ops, n - 1)
            assert e.value.code == api.ERROR_OUT_OF_RANGE, (lib.prefix, col)
        pre = ev.preorder_operations().copy()
        pre[1, 3] = nb + 7
        with pytest.raises(api.BeagleError) as e:
            inst.update_pre_partials(pre)
        assert e.value.code == api.ERROR_OUT_OF_RANGE
        with pytest.raises(api.BeagleError) as e:
            inst.root_log_likelihood(nb + 1, 0, 0, n - 1)
        assert e.value.code == api.ERROR_OUT_OF_RANGE
        with pytest.raises(api.BeagleError) as e:
            inst.accumulate_scale_factors([0, 10 * n], n - 1)
        assert e.value.code == api.ERROR_OUT_OF_RANGE
        assert ev.log_likelihood() == want
        ev.close()


def test_allocation_failure_gives_out_of_memory(amd):
    """An instance whose first partials buffer (8 categories x 2e9 patterns
    x 4 states in FP64, 512 GB) exceeds the GPU's memory: creation returns
    BEAGLE_ERROR_OUT_OF_MEMORY, frees what it had made, and a following
    instance evaluates normally."""
    with pytest.raises(api.BeagleError) as e:
        amd.create(3, 2, 3, 4, 2_000_000_000, 1, 4, 8, 0,
                   requirement=api.FLAG_PRECISION_DOUBLE)
    assert e.value.code == api.ERROR_OUT_OF_MEMORY
    prob = synth.make_problem(904, 6, 500, S=4)
    ev = driver.Evaluator(amd, prob)
    assert np.isfinite(ev.log_likelihood())
    ev.close()


@pytest.mark.parametrize("precision,tips_as", [("double", "states"), ("single", "partials")])
def test_no_destination_pre_order_rules(amd, precision, tips_as):
    """A pre-order operation without a destination in the derivative
    extension, 20 states: refused (BEAGLE_ERROR_OUT_OF_RANGE) in double
    precision and for a tip set as partials (its derivative sums need the
    tip's state); a single-precision instance with the tip set as states
    accepts it (test_gpu_parity checks those sums)."""
    prob = synth.make_problem(5, 7, 90, S=20, missing=0.05)
    ev = driver.Evaluator(amd, prob, precision=precision, tips_as=tips_as)
    ops, post, d1, d2 = ev.fused_operations(keep_tips=True)
    ops = np.array(ops, copy=True)
    ops[post < ev.n, 0] = api.OP_NONE
    ev.update_matrices()
    ev.post_order()
    root = prob.tree.root
    with pytest.raises(api.BeagleError) as e:
        ev.inst.log_likelihood_and_derivatives(root, ev.pre_base + root, ops, post, d1, d2,
                                               cumulative=api.OP_NONE)
    assert e.value.code == api.ERROR_OUT_OF_RANGE
    ev.close()
