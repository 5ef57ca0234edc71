"""The checkpoint schedule of beagleExtCheckpointedGradient, planned on the
host without a device (beagleExtPlanCheckpoints): which nodes are kept,
how many operations are recomputed, the buffers in use at once, and the
trees and pools it refuses.  The schedule's results are compared with the
full evaluation on the GPU (test_gpu_parity.py)."""
import numpy as np
import pytest

from beagle_amd import api, synth


def heights(ops, n_tips):
    h = {}
    for d, _, _, a, _, b, _ in np.asarray(ops).reshape(-1, 7):
        h[d] = 1 + max(h.get(a, 0) if a >= n_tips else 0, h.get(b, 0) if b >= n_tips else 0)
    return h


def tree_ops(n, shape, seed=1):
    tree = synth.make_tree(np.random.default_rng(seed), n, shape, 0.1)
    return synth.postorder_ops(tree, None), tree


@pytest.mark.parametrize("n,shape", [(16, "caterpillar"), (64, "balanced"), (128, "caterpillar"),
                                     (128, "balanced"), (100, "yule"), (257, "yule")])
@pytest.mark.parametrize("tip_pre", [False, True])
def test_checkpoints_and_recomputation(amd, n, shape, tip_pre):
    """For intervals 1..7 the kept nodes are exactly the internal nodes
    whose height is a multiple of the interval, plus the root; every other
    internal node is computed again once in the pre-order pass; the peak
    fits the pool it was planned for."""
    ops, tree = tree_ops(n, shape)
    h = heights(ops, n)
    root = int(np.asarray(ops).reshape(-1, 7)[-1, 0])
    big = 4 * n
    for k in range(1, 8):
        used, peak, ckpt, again = amd.plan_checkpoints(n, ops, big, k, tip_pre)
        want = sum(1 for v, hv in h.items() if hv % k == 0 or v == root)
        assert used == k and ckpt == want
        assert again == (n - 1) - ckpt
        assert 0 < peak <= big


@pytest.mark.parametrize("n,shape", [(128, "caterpillar"), (128, "balanced"), (200, "yule")])
def test_smallest_pool(amd, n, shape):
    """With interval 0 the planner takes the smallest interval whose
    schedule fits: at the smallest pool any interval fits, every smaller
    interval is refused; one buffer less is refused for every interval.
    The pool is a fraction of the 2 (n - 1) buffers a full evaluation
    holds."""
    ops, _ = tree_ops(n, shape)

    def fits(m, k):
        try:
            amd.plan_checkpoints(n, ops, m, k)
            return True
        except api.BeagleError as e:
            assert e.code == api.ERROR_GENERAL
            return False
    m = next(m for m in range(2, 2 * n) if fits(m, 0))
    used, peak, _, _ = amd.plan_checkpoints(n, ops, m, 0)
    assert peak <= m
    assert all(not fits(m, k) for k in range(1, used))
    assert not fits(m - 1, 0)
    assert m < (n - 1) / 2


def test_refuses_malformed_trees(amd):
    ops, _ = tree_ops(16, "yule")
    ops = np.asarray(ops).reshape(-1, 7)
    bad = ops.copy()
    bad[3, 0] = 2                               # a destination that is a tip
    bad2 = ops.copy()
    bad2[-1, 3] = bad2[-2, 3]                   # a child with two parents
    bad3 = ops[::-1].copy()                     # children before their operation
    for b in (bad, bad2, bad3):
        with pytest.raises(api.BeagleError) as e:
            amd.plan_checkpoints(16, b, 64, 0)
        assert e.value.code == api.ERROR_OUT_OF_RANGE
    scaled = ops.copy()
    scaled[0, 1] = 0                            # a scale buffer
    with pytest.raises(api.BeagleError) as e:
        amd.plan_checkpoints(16, scaled, 64, 0)
    assert e.value.code == api.ERROR_NO_IMPLEMENTATION
