"""Synthetic inputs for tests and benchmarks: substitution models, trees,
site patterns and the BEAGLE operation lists that traverse them.

Everything is seeded (numpy PCG64), so a seed names a problem.  The
generators follow the workload definitions used for the measurements:
GTR with exchangeabilities ~ Exp(1) and frequencies ~ Dirichlet(1),
discrete-Gamma rates (4 categories, shape 0.5, mean-of-category rates,
equal weights), Yule (pure-birth) topologies or caterpillar/balanced
shapes, branch lengths ~ Exp(mean 0.1), uniform random tip states and
pattern weights 1 + Poisson(2).  The codon model is GY94-style over the 61
sense codons of the universal code (kappa ~ LogNormal(1, 0.3),
omega ~ Gamma(2, 0.25), frequencies ~ Dirichlet(1)).

The eigen-system is computed through the symmetric form
Pi^(1/2) Q Pi^(-1/2) (reversible models), giving real eigenvalues and
Vinv = V^T Pi^(1/2)-scaled, as a client such as BEAST supplies it.
"""
import dataclasses
import itertools

import numpy as np
from scipy import special, stats


@dataclasses.dataclass
class Model:
    Q: np.ndarray          # S x S generator, rows sum to 0, normalised to 1 subst/unit
    freqs: np.ndarray      # S
    evec: np.ndarray       # S x S, Q = evec diag(evals) ivec
    ivec: np.ndarray
    evals: np.ndarray
    exch: np.ndarray = None  # exchangeabilities (upper triangle, row-major) if reversible

    @property
    def S(self):
        return len(self.freqs)


def eigen_reversible(Q, freqs):
    s = np.sqrt(freqs)
    A = (Q * s[:, None]) / s[None, :]
    A = 0.5 * (A + A.T)
    lam, U = np.linalg.eigh(A)
    V = U / s[:, None]
    Vi = U.T * s[None, :]
    return V, Vi, lam


def reversible_Q(exch, freqs):
    S = len(freqs)
    R = np.zeros((S, S))
    iu = np.triu_indices(S, 1)
    R[iu] = exch
    R = R + R.T
    Q = R * freqs[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    mu = -np.dot(freqs, np.diag(Q))
    return Q / mu


def model_from(exch, freqs):
    Q = reversible_Q(exch, freqs)
    V, Vi, lam = eigen_reversible(Q, freqs)
    return Model(Q=Q, freqs=np.asarray(freqs, float), evec=V, ivec=Vi, evals=lam,
                 exch=np.asarray(exch, float))


def gtr(rng, S=4):
    """GTR-style reversible model on S states: exchangeabilities ~ Exp(1),
    frequencies ~ Dirichlet(1)."""
    exch = rng.exponential(1.0, S * (S - 1) // 2)
    freqs = rng.dirichlet(np.ones(S))
    return model_from(exch, freqs)


def nonreversible(rng, S=4):
    """Non-reversible model with complex eigenvalues: random rates R_ij ~
    Exp(1) (i != j), Q = R - diag(R 1), normalised to one expected
    substitution at its stationary distribution.  The eigen-system in the
    real form BEAGLE_FLAG_EIGEN_COMPLEX takes (EigenDecompositionSquare.hpp:
    108-185): a conjugate pair a +- ib with eigenvector x + iy becomes the
    columns x, y of V with the block [[a, b], [-b, a]], evals = S real parts
    then S imaginary parts (b on the pair's first row, -b on its second)."""
    for _ in range(100):
        R = rng.exponential(1.0, (S, S))
        np.fill_diagonal(R, 0.0)
        Q = R - np.diag(R.sum(axis=1))
        w, vl = np.linalg.eig(Q.T)
        pi = np.real(vl[:, np.argmin(np.abs(w))])
        pi = pi / pi.sum()
        Q = Q / -(pi * np.diag(Q)).sum()
        lam, vec = np.linalg.eig(Q)
        if np.any(np.abs(lam.imag) > 1e-6):
            break
    V = np.zeros((S, S))
    re = np.zeros(S)
    im = np.zeros(S)
    order = np.argsort(-lam.real, kind="stable")
    done = np.zeros(S, bool)
    k = 0
    for i in order:
        if done[i]:
            continue
        if abs(lam[i].imag) < 1e-12:
            V[:, k], re[k] = np.real(vec[:, i]), lam[i].real
            done[i] = True
            k += 1
            continue
        j = min((m for m in range(S) if not done[m] and m != i),
                key=lambda m: abs(lam[m] - np.conj(lam[i])))
        a, b = (i, j) if lam[i].imag > 0 else (j, i)   # a: positive imaginary part
        V[:, k], V[:, k + 1] = np.real(vec[:, a]), np.imag(vec[:, a])
        re[k] = re[k + 1] = lam[a].real
        im[k], im[k + 1] = lam[a].imag, -lam[a].imag
        done[a] = done[b] = True
        k += 2
    Vi = np.linalg.inv(V)
    return Model(Q=Q, freqs=pi, evec=V, ivec=Vi, evals=np.concatenate([re, im]))


def jc69():
    return model_from(np.ones(6), np.full(4, 0.25))


def hky(kappa, freqs):
    # order A,C,G,T: transitions A<->G, C<->T
    exch = np.array([1.0, kappa, 1.0, 1.0, kappa, 1.0])
    return model_from(exch, np.asarray(freqs, float))


BASES = "TCAG"
_AA = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG"


def sense_codons():
    cods = ["".join(c) for c in itertools.product(BASES, repeat=3)]
    return [(c, _AA[i]) for i, c in enumerate(cods) if _AA[i] != "*"]


def gy94(rng, kappa=None, omega=None, freqs=None):
    """GY94-style codon model over the 61 sense codons."""
    cods = sense_codons()
    S = len(cods)
    if kappa is None:
        kappa = float(np.exp(rng.normal(1.0, 0.3)))
    if omega is None:
        omega = float(rng.gamma(2.0, 0.25))
    if freqs is None:
        freqs = rng.dirichlet(np.ones(S))
    purines = set("AG")
    exch = []
    for a in range(S):
        for b in range(a + 1, S):
            ca, aa = cods[a]
            cb, ab = cods[b]
            diff = [(x, y) for x, y in zip(ca, cb) if x != y]
            if len(diff) != 1:
                exch.append(0.0)
                continue
            x, y = diff[0]
            r = 1.0
            if (x in purines) == (y in purines):
                r *= kappa
            if aa != ab:
                r *= omega
            exch.append(r)
    m = model_from(np.array(exch), freqs)
    m.kappa, m.omega = kappa, omega
    return m


def gamma_rates(alpha=0.5, K=4):
    """Mean-of-category discrete-Gamma rates (mean 1) and equal weights."""
    if K == 1:
        return np.ones(1), np.ones(1)
    cuts = stats.gamma.ppf(np.arange(1, K) / K, a=alpha, scale=1.0 / alpha)
    lo = np.concatenate([[0.0], cuts])
    hi = np.concatenate([cuts, [np.inf]])
    inc = special.gammainc(alpha + 1, alpha * hi) - special.gammainc(alpha + 1, alpha * lo)
    rates = inc * K
    return rates, np.full(K, 1.0 / K)


@dataclasses.dataclass
class Tree:
    """Rooted binary tree.  Tips are 0..n-1, internal nodes n..2n-2 with the
    root last; children[i] = (a, b) for internal i; lengths[v] is the branch
    above node v (root excluded)."""
    n: int
    children: dict
    parent: np.ndarray
    lengths: np.ndarray

    @property
    def root(self):
        return 2 * self.n - 2

    def postorder(self):
        out = []

        def rec(v):
            if v < self.n:
                return
            a, b = self.children[v]
            rec(a)
            rec(b)
            out.append(v)
        # iterative to allow deep caterpillars
        stack = [(self.root, False)]
        while stack:
            v, done = stack.pop()
            if v < self.n:
                continue
            if done:
                out.append(v)
            else:
                stack.append((v, True))
                a, b = self.children[v]
                stack.append((b, False))
                stack.append((a, False))
        return out

    def preorder(self):
        return list(reversed(self.postorder()))

    def depth(self):
        d = np.zeros(2 * self.n - 1, int)
        for v in self.preorder():
            for c in self.children[v]:
                d[c] = d[v] + 1
        return d.max()


def _relabel(n, edges_children, root_id, lengths_by_id, rng_order=None):
    """Map arbitrary internal ids to n..2n-2 in post-order, root last."""
    order = []
    stack = [(root_id, False)]
    while stack:
        v, done = stack.pop()
        if v < n:
            continue
        if done:
            order.append(v)
        else:
            stack.append((v, True))
            a, b = edges_children[v]
            stack.append((b, False))
            stack.append((a, False))
    mapping = {v: v for v in range(n)}
    for i, v in enumerate(order):
        mapping[v] = n + i
    children = {mapping[v]: (mapping[a], mapping[b]) for v, (a, b) in edges_children.items()}
    parent = np.full(2 * n - 1, -1)
    for v, (a, b) in children.items():
        parent[a] = v
        parent[b] = v
    lengths = np.zeros(2 * n - 1)
    for v, t in lengths_by_id.items():
        lengths[mapping[v]] = t
    return Tree(n=n, children=children, parent=parent, lengths=lengths)


def yule_tree(rng, n, mean_length=0.1):
    """Pure-birth topology: repeatedly split a uniformly random extant lineage."""
    next_id = n
    root = 10 ** 9
    children = {}
    lineages = [root]
    while len(lineages) < n:
        i = rng.integers(len(lineages))
        v = lineages.pop(i)
        a, b = next_id, next_id + 1
        next_id += 2
        children[v] = (a, b)
        lineages += [a, b]
    # assign tips: the n extant lineages become tips 0..n-1 in random order
    perm = rng.permutation(n)
    tipmap = {lin: int(perm[i]) for i, lin in enumerate(lineages)}

    def m(v):
        return tipmap.get(v, v)
    ch = {m(v): (m(a), m(b)) for v, (a, b) in children.items()}
    ids = set(ch) | {x for ab in ch.values() for x in ab}
    lengths = {v: rng.exponential(mean_length) for v in ids if v != m(root)}
    return _relabel(n, ch, m(root), lengths)


def caterpillar_tree(rng, n, mean_length=0.1):
    ch = {}
    cur = 0
    nid = n + 1000
    for t in range(1, n):
        ch[nid] = (cur, t)
        cur = nid
        nid += 1
    lengths = {v: rng.exponential(mean_length) for v in list(ch) + list(range(n)) if v != cur}
    return _relabel(n, ch, cur, lengths)


def balanced_tree(rng, n, mean_length=0.1):
    level = list(range(n))
    ch = {}
    nid = n + 1000
    while len(level) > 1:
        nxt = []
        for i in range(0, len(level) - 1, 2):
            ch[nid] = (level[i], level[i + 1])
            nxt.append(nid)
            nid += 1
        if len(level) % 2:
            nxt.append(level[-1])
        level = nxt
    root = level[0]
    lengths = {v: rng.exponential(mean_length) for v in list(ch) + list(range(n)) if v != root}
    return _relabel(n, ch, root, lengths)


def make_tree(rng, n, shape="yule", mean_length=0.1):
    return {"yule": yule_tree, "caterpillar": caterpillar_tree,
            "balanced": balanced_tree}[shape](rng, n, mean_length)


def random_patterns(rng, n_tips, P, S=4, missing=0.0):
    X = rng.integers(0, S, size=(n_tips, P)).astype(np.int32)
    if missing > 0:
        X[rng.random((n_tips, P)) < missing] = S
    return X


def pattern_weights(rng, P):
    return 1.0 + rng.poisson(2.0, P).astype(float)


def states_to_partials(states, S):
    """One-hot partials for tip states; a missing state (>= S) is all ones."""
    P = len(states)
    out = np.zeros((P, S))
    ok = states < S
    out[np.arange(P)[ok], states[ok]] = 1.0
    out[~ok] = 1.0
    return out


def postorder_ops(tree, scale_write=None, cumulative=False):
    """BEAGLE operations for a full post-order traversal.  Partials buffer of
    node v is v; the matrix of the branch above v is v.  scale_write maps a
    node to its scale-buffer index (or is None for no rescaling)."""
    ops = []
    for v in tree.postorder():
        a, b = tree.children[v]
        w = scale_write(v) if scale_write else -1
        ops.append([v, w, -1, a, a, b, b])
    return np.array(ops, dtype=np.int32)


def preorder_ops(tree, pre_base):
    """Pre-order operations.  The pre-order buffer of node v is pre_base + v;
    the root's pre-order buffer holds the state frequencies.  Each op:
    [pre(v), -1, -1, pre(parent), P(v), post(sibling), P(sibling)]."""
    ops = []
    for v in tree.preorder():
        a, b = tree.children[v]
        ops.append([pre_base + a, -1, -1, pre_base + v, a, b, b])
        ops.append([pre_base + b, -1, -1, pre_base + v, b, a, a])
    return np.array(ops, dtype=np.int32)


@dataclasses.dataclass
class Problem:
    tree: Tree
    model: Model
    rates: np.ndarray
    weights: np.ndarray      # category weights
    tips: np.ndarray         # n x P int states (S means missing)
    pattern_weights: np.ndarray

    @property
    def S(self):
        return self.model.S

    @property
    def P(self):
        return self.tips.shape[1]

    @property
    def C(self):
        return len(self.rates)


def make_problem(seed, n_taxa, P, S=4, shape="yule", C=4, alpha=0.5, missing=0.0,
                 mean_length=0.1, model=None):
    rng = np.random.default_rng(seed)
    tree = make_tree(rng, n_taxa, shape, mean_length)
    if model is None:
        model = gy94(rng) if S == 61 else gtr(rng, S)
    rates, cw = gamma_rates(alpha, C)
    tips = random_patterns(rng, n_taxa, P, model.S, missing)
    w = pattern_weights(rng, P)
    return Problem(tree=tree, model=model, rates=rates, weights=cw, tips=tips, pattern_weights=w)


def problem_arrays(prob, prefix=""):
    """A Problem as a flat dict of arrays (for fixtures)."""
    t, m = prob.tree, prob.model
    n = t.n
    ch = np.array([t.children[v] for v in range(n, 2 * n - 1)], dtype=np.int64)
    return {prefix + "children": ch, prefix + "lengths": t.lengths, prefix + "Q": m.Q,
            prefix + "freqs": m.freqs, prefix + "evec": m.evec, prefix + "ivec": m.ivec,
            prefix + "evals": m.evals, prefix + "rates": prob.rates,
            prefix + "weights": prob.weights, prefix + "tips": prob.tips,
            prefix + "pattern_weights": prob.pattern_weights}


def problem_from_arrays(d, prefix=""):
    """Inverse of problem_arrays."""
    ch = np.asarray(d[prefix + "children"])
    n = ch.shape[0] + 1
    children = {n + i: (int(a), int(b)) for i, (a, b) in enumerate(ch)}
    parent = np.full(2 * n - 1, -1)
    for v, (a, b) in children.items():
        parent[a] = v
        parent[b] = v
    tree = Tree(n=n, children=children, parent=parent, lengths=np.asarray(d[prefix + "lengths"]))
    model = Model(Q=np.asarray(d[prefix + "Q"]), freqs=np.asarray(d[prefix + "freqs"]),
                  evec=np.asarray(d[prefix + "evec"]), ivec=np.asarray(d[prefix + "ivec"]),
                  evals=np.asarray(d[prefix + "evals"]))
    return Problem(tree=tree, model=model, rates=np.asarray(d[prefix + "rates"]),
                   weights=np.asarray(d[prefix + "weights"]),
                   tips=np.asarray(d[prefix + "tips"]),
                   pattern_weights=np.asarray(d[prefix + "pattern_weights"]))
