"""beagleExtCalculateRootLogLikelihoodsNoAlloc, the allocation-free variant
of beagleCalculateRootLogLikelihoods, through the C ABI on cuda:0: its
log-likelihood and site log-likelihoods equal the existing entry point's
(same kernels, same summation order: bit for bit) on two calls in a row on
instances of one GPU, the second with fewer patterns (the shared scratch
array then still holds the first call's workgroup sums), and pattern counts
beyond the scratch array are refused."""
import numpy as np
import pytest

from beagle_amd import api, driver, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("S,precision,scaling", [(4, "double", "manual"), (4, "single", "auto"),
                                                 (20, "single", "none"), (61, "double", "auto")])
def test_no_alloc_root_matches_existing_entry_point(amd, S, precision, scaling):
    sizes = {4: (100_000, 1_000), 20: (20_000, 333), 61: (6_000, 77)}[S]
    for k, P in enumerate(sizes):
        prob = synth.make_problem(700 + 10 * S + k, 16, P, S=S, shape="yule", missing=0.03)
        ev = driver.Evaluator(amd, prob, precision=precision, scaling=scaling, preorder=False)
        ev.update_matrices()
        ev.post_order()
        cum = ev.n - 1 if scaling == "manual" else api.OP_NONE
        root = prob.tree.root
        got = ev.inst.root_log_likelihood_no_alloc(root, 0, 0, cum)
        got_sites = ev.inst.site_log_likelihoods()
        want = ev.inst.root_log_likelihood(root, 0, 0, cum)
        want_sites = ev.inst.site_log_likelihoods()
        again = ev.inst.root_log_likelihood_no_alloc(root, 0, 0, cum)
        ev.close()
        assert np.isfinite(want)
        assert got == want and again == want, (P, got, want)
        np.testing.assert_array_equal(got_sites, want_sites)


def test_no_alloc_root_limits(amd):
    """More patterns than the scratch array holds workgroup sums for:
    BEAGLE_ERROR_OUT_OF_RANGE; several roots or per-category output:
    BEAGLE_ERROR_NO_IMPLEMENTATION."""
    P = 32767 * 32 + 1          # one 32-pattern workgroup sum too many (S > 4)
    inst = amd.create(2, 1, 2, 20, P, 1, 2, 1, 1,
                      requirement=api.FLAG_PRECISION_SINGLE)
    try:
        inst.set_category_weights(0, np.ones(1))
        inst.set_state_frequencies(0, np.full(20, 0.05))
        with pytest.raises(api.BeagleError) as e:
            inst.root_log_likelihood_no_alloc(2, 0, 0)
        assert e.value.code == api.ERROR_OUT_OF_RANGE
        with pytest.raises(api.BeagleError) as e:
            inst.root_log_likelihood_no_alloc(2, -1, 0)
        assert e.value.code == api.ERROR_NO_IMPLEMENTATION
    finally:
        inst.finalize()
