"""The PhyloLik compatibility entries on the GPU (pcsf_lik_columns,
pcsf_lik_posteriors) against the oracle's restatement of
lib/CamlPaml/PhyloLik.ml:9-174 and PhyloEM.collect_sufficient_statistics
(lib/CamlPaml/PhyloEM.ml:14-20); paths relative to the mlin/PhyloCSF tree.

* every leaf kind (`Certain, `Distribution, `Marginalize, and a state >= k,
  which the reference makes a vector of zeros), on a small alphabet and on 64
  codons: likelihoods, node posteriors of every node, and expected
  substitution counts summed over columns with weights, added to what the
  array held;
* trees whose workspace exceeds 64 KiB (100vertebrates, 120mammals at
  k = 64), which the first version of this entry refused;
* what the mathematics fixes: each column's branch posterior sums to 1, and
  its row sums are the parent's node posterior.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import phylocsf_amd as P  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    return P.Context(0)


def _trees(params_dir, name):
    text = open(os.path.join(params_dir, name + ".nh")).read()
    return P.newick_tree(text), O.newick_tree(text)


def _random_P(rng, nbr, k):
    P_ = rng.random((nbr, k, k)) + 0.05
    return P_ / P_.sum(axis=2, keepdims=True)


def _leaves(rng, ncols, nl, k):
    leaves = rng.integers(0, k, (ncols, nl)).astype(np.int32)
    r = rng.random((ncols, nl))
    leaves[r < 0.15] = -1
    leaves[(r >= 0.15) & (r < 0.3)] = -2
    leaves[(r >= 0.3) & (r < 0.32)] = k + 3  # `Certain beyond the alphabet
    dist = rng.random((ncols, nl, k))
    return leaves, dist


@pytest.mark.parametrize("k", [2, 5, 64])
def test_posteriors_and_counts_match_oracle(ctx, params_dir, k):
    _say("PhyloLik posteriors, k = %d" % k)
    tree, otree = _trees(params_dir, "12flies")
    rng = np.random.default_rng(k)
    Pm = _random_P(rng, tree.n - 1, k)
    prior = rng.random(k)
    prior /= prior.sum()
    leaves, dist = _leaves(rng, 700, tree.nl, k)
    w = rng.uniform(0.5, 2.0, 700)
    it = P.PhyloLik(ctx).prepare(tree, Pm, prior, leaves, dist)
    z0, post0, ec0 = O.lik_posteriors(otree, Pm, prior, leaves, dist, w)
    assert np.allclose(it.likelihood(), z0, rtol=1e-12, atol=0)
    for i in (0, tree.nl - 1, tree.nl, tree.n - 2, tree.n - 1):
        assert np.allclose(it.node_posterior(i), post0[:, i, :], rtol=1e-11, atol=1e-300), i
    ec = np.full((tree.n - 1, k, k), 0.25)  # added to, as add_branch_posteriors does
    it.add_branch_posteriors(ec, w)
    assert np.allclose(ec, ec0 + 0.25, rtol=1e-11, atol=1e-14)


def test_branch_posterior_identities(ctx, params_dir):
    _say("branch posterior identities")
    tree, _ = _trees(params_dir, "29mammals")
    rng = np.random.default_rng(3)
    k = 4
    Pm = _random_P(rng, tree.n - 1, k)
    prior = np.full(k, 0.25)
    leaves, dist = _leaves(rng, 1, tree.nl, k)
    leaves[leaves > k] = 1
    it = P.PhyloLik(ctx).prepare(tree, Pm, prior, leaves, dist)
    ec = np.zeros((tree.n - 1, k, k))
    it.add_branch_posteriors(ec)
    parent = tree.parent()
    for br in range(tree.n - 1):
        assert abs(ec[br].sum() - 1.0) < 1e-12, br
        p = parent[br]
        assert np.allclose(ec[br].sum(axis=1), it.node_posterior(p)[0], rtol=1e-10, atol=1e-15)
        if br >= tree.nl:  # column sums: the node's own posterior
            assert np.allclose(ec[br].sum(axis=0), it.node_posterior(br)[0], rtol=1e-10,
                               atol=1e-15)


@pytest.mark.parametrize("name", ["100vertebrates", "120mammals"])
def test_large_trees_at_64_codons(ctx, params_dir, name):
    _say("PhyloLik on %s at k = 64" % name)
    tree, otree = _trees(params_dir, name)
    S, pi = O.ecm_read(os.path.join(params_dir, name + "_coding.ECM"))
    Q = O.build_q(S, pi)
    Pm = np.stack([O.pt(Q, pi, t) for t in otree.branch[:-1]])
    m = P.PhyloCSFModel(None, tree, [(S, pi)])
    codes, off = m.simulate([60], [0], [1.0], gap_p=0.2, seed=4)
    leaves = np.where(codes == 255, -1, codes.astype(np.int32))
    z = P.PhyloLik(ctx).prepare(tree, Pm, pi, leaves).likelihood()
    z0, post0, ec0 = O.lik_posteriors(otree, Pm, pi, leaves)
    assert np.allclose(z, z0, rtol=1e-11, atol=0)
    it = P.PhyloLik(ctx).prepare(tree, Pm, pi, leaves)
    assert np.allclose(it.node_posterior(tree.n - 1), post0[:, -1, :], rtol=1e-10, atol=1e-300)
    assert np.allclose(it.node_posterior(tree.nl + 3), post0[:, tree.nl + 3, :], rtol=1e-10,
                       atol=1e-300)
    ss = np.zeros((tree.n - 1, 64, 64))
    zz = P.collect_sufficient_statistics(ctx, tree, Pm, pi, leaves, ss)
    assert np.allclose(zz, z0, rtol=1e-11, atol=0)
    assert np.allclose(ss, ec0, rtol=1e-10, atol=1e-13)


def test_estep_fast_path_matches_oracle(ctx, params_dir):
    """The E-step at 64 states without `Distribution leaves runs on the FP64
    matrix cores (kernels/estep.hip): counts and likelihoods against the
    oracle with weights, `Marginalize leaves and states beyond the alphabet,
    over a column count that leaves a partial 64-column block."""
    _say("E-step fast path, 58mammals")
    tree, otree = _trees(params_dir, "58mammals")
    S, pi = O.ecm_read(os.path.join(params_dir, "58mammals_coding.ECM"))
    Q = O.build_q(S, pi)
    Pm = np.stack([O.pt(Q, pi, 1.3 * t) for t in otree.branch[:-1]])
    m = P.PhyloCSFModel(None, tree, [(S, pi)])
    codes, off = m.simulate([150], [0], [1.3], gap_p=0.2, seed=8)
    leaves = np.where(codes == 255, -1, codes.astype(np.int32))
    rng = np.random.default_rng(8)
    leaves[rng.random(leaves.shape) < 0.01] = 64 + 5  # a vector of zeros
    w = rng.uniform(0.5, 2.0, leaves.shape[0])
    z0, _, ec0 = O.lik_posteriors(otree, Pm, pi, leaves, None, w, posteriors=False)
    ec = np.full((tree.n - 1, 64, 64), 0.5)
    z = P.collect_sufficient_statistics(ctx, tree, Pm, pi, leaves, ec, weights=w)
    assert np.allclose(z, z0, rtol=1e-11, atol=0)
    assert np.allclose(ec, ec0 + 0.5, rtol=1e-10, atol=1e-13)


def test_estep_chunks_add_up(ctx, params_dir):
    """Over more columns than one device chunk (32,768) the counts are the sum
    of the counts of the parts (the E-step is additive over columns)."""
    _say("E-step over 40,000 columns = its two parts")
    tree, otree = _trees(params_dir, "29mammals")
    S, pi = O.ecm_read(os.path.join(params_dir, "29mammals_coding.ECM"))
    Q = O.build_q(S, pi)
    Pm = np.stack([O.pt(Q, pi, t) for t in otree.branch[:-1]])
    m = P.PhyloCSFModel(None, tree, [(S, pi)])
    codes, off = m.simulate([40000], [0], [1.0], gap_p=0.1, seed=9)
    leaves = np.where(codes == 255, -1, codes.astype(np.int32))
    whole = np.zeros((tree.n - 1, 64, 64))
    zw = P.collect_sufficient_statistics(ctx, tree, Pm, pi, leaves, whole)
    parts = np.zeros((tree.n - 1, 64, 64))
    za = P.collect_sufficient_statistics(ctx, tree, Pm, pi, leaves[:17000], parts)
    zb = P.collect_sufficient_statistics(ctx, tree, Pm, pi, leaves[17000:], parts)
    assert np.array_equal(zw, np.concatenate([za, zb]))
    assert np.allclose(whole, parts, rtol=1e-11, atol=1e-9)
    # each branch's counts sum to the number of columns (a posterior per column)
    assert np.allclose(whole.sum(axis=(1, 2)), leaves.shape[0], rtol=1e-9)


def test_bad_leaves_refused(ctx, params_dir):
    tree, _ = _trees(params_dir, "12flies")
    Pm = np.tile(np.eye(3), (tree.n - 1, 1, 1))
    lv = np.zeros((2, tree.nl), np.int32)
    lv[1, 2] = -2
    with pytest.raises(P.InvalidArgument):
        P.PhyloLik(ctx).prepare(tree, Pm, np.ones(3) / 3, lv)  # `Distribution without leaf_dist
    lv[1, 2] = -3
    with pytest.raises(P.InvalidArgument):
        P.PhyloLik(ctx).prepare(tree, Pm, np.ones(3) / 3, lv)
