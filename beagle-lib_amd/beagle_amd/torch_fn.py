"""torch.autograd binding over the C ABI.

PhyloLikelihood holds one instance of the HIP library for a fixed
topology and alignment; calling it with branch lengths, a rate matrix Q
and root frequencies pi returns the log-likelihood as a differentiable
torch scalar:

    lik = PhyloLikelihood(lib, tree, tips, pattern_weights, rates, weights)
    Q = gtr_rate_matrix(exch, freqs)          # differentiable in torch
    logl = lik(lengths, Q, freqs)
    logl.backward()                           # d/d lengths, exch, freqs

Forward: the eigen-system of Q (FP64, numpy), transition matrices, the
post-order traversal, then one call for the root log-likelihood and the
fused pre-order + branch-derivative pass.  Backward: dlogL/dt from those
sums, dlogL/dQ and dlogL/dpi (root distribution) from the rate-matrix
gradient extension (exact, through the eigen-system); torch's autograd
then carries dlogL/dQ to whatever parametrizes Q.  The device work runs
on the library's own HIP stream; tensors cross as host FP64 arrays (the
alignment stays resident on the device between calls).
"""
import numpy as np
import torch

from . import driver, synth


def gtr_rate_matrix(exch, freqs):
    """Reversible rate matrix from exchangeabilities (upper triangle,
    row-major) and frequencies, normalised to one expected substitution per
    unit time; differentiable torch code."""
    S = freqs.shape[0]
    iu = torch.triu_indices(S, S, 1)
    R = torch.zeros(S, S, dtype=exch.dtype, device=exch.device)
    R = R.index_put((iu[0], iu[1]), exch)
    R = R + R.T
    Q = R * freqs[None, :]
    Q = Q - torch.diag(Q.sum(dim=1))
    mu = -(freqs * torch.diagonal(Q)).sum()
    return Q / mu


def _eigen(Q):
    lam, V = np.linalg.eig(Q)
    if np.max(np.abs(lam.imag)) > 1e-9 * max(1.0, np.max(np.abs(lam.real))):
        raise ValueError("rate matrix with complex eigenvalues")
    V = V.real
    return V, np.linalg.inv(V), lam.real


class PhyloLikelihood:
    def __init__(self, lib, tree, tips, pattern_weights, rates, weights, *, precision="double",
                 scaling="auto", states=4):
        """tree: synth.Tree; tips: (taxa, patterns) int states 0..states-1
        (>= states missing); rates, weights: the rate categories.  Any state
        count up to 64 (the rate-matrix gradient covers them); the model is
        a placeholder until the first evaluation sets Q and pi."""
        S = int(states)
        model = synth.model_from(np.ones(S * (S - 1) // 2), np.full(S, 1.0 / S))
        self.prob = synth.Problem(tree=tree, model=model, rates=np.asarray(rates, float),
                                  weights=np.asarray(weights, float), tips=np.asarray(tips),
                                  pattern_weights=np.asarray(pattern_weights, float))
        self.ev = driver.Evaluator(lib, self.prob, precision=precision, scaling=scaling)

    def close(self):
        self.ev.close()

    def _set_model(self, Q, pi):
        ev, p = self.ev, self.prob
        V, Vi, lam = _eigen(Q)
        ev.inst.set_eigen(0, V.reshape(-1), Vi.reshape(-1), lam)
        ev.inst.set_state_frequencies(0, pi)
        d1 = np.concatenate([(r * Q).reshape(-1) for r in p.rates])
        d2 = np.concatenate([(r * r * (Q @ Q)).reshape(-1) for r in p.rates])
        ev.inst.set_differential_matrix(ev.d1_index, d1)
        ev.inst.set_differential_matrix(ev.d2_index, d2)
        p.model = synth.Model(Q=Q, freqs=pi, evec=V, ivec=Vi, evals=lam)

    def evaluate(self, lengths, Q, pi):
        """(logL, dlogL/dt per branch, dlogL/dQ, dlogL/dpi) as numpy."""
        self._set_model(Q, pi)
        n = self.prob.tree.n
        t = np.zeros(2 * n - 1)
        t[:2 * n - 2] = lengths
        logl, g1, _ = self.ev.fused_gradient(t)
        G, gpi = self.ev.rate_matrix_gradient()
        # branch order of the derivative sums: nodes 0..2n-2 without the root
        nodes = self.ev.branch_nodes()
        gt = np.zeros(2 * n - 2)
        gt[np.asarray(nodes)] = g1
        return logl, gt, G, gpi

    def __call__(self, lengths, Q, pi):
        return _LogLikelihood.apply(lengths, Q, pi, self)


class _LogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lengths, Q, pi, lik):
        logl, gt, G, gpi = lik.evaluate(lengths.detach().cpu().double().numpy(),
                                        Q.detach().cpu().double().numpy(),
                                        pi.detach().cpu().double().numpy())
        ctx.save_for_backward(torch.as_tensor(gt, dtype=lengths.dtype, device=lengths.device),
                              torch.as_tensor(G, dtype=Q.dtype, device=Q.device),
                              torch.as_tensor(gpi, dtype=pi.dtype, device=pi.device))
        return torch.tensor(logl, dtype=torch.float64, device=lengths.device)

    @staticmethod
    def backward(ctx, grad):
        gt, G, gpi = ctx.saved_tensors
        g = grad.to(gt.dtype)
        return g * gt, g * G, g * gpi, None
