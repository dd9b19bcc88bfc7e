"""The synthetic-alignment generator (pcsf_simulate, pcsf_simulate_nucleotides)
against what PhyloModel.simulate's semantics fix (lib/CamlPaml/PhyloModel.ml:
38-54, lib/CamlPaml/Tools.ml:22-41; paths relative to the mlin/PhyloCSF
tree), on a host-only model (no device):

* every leaf's codon is distributed as the root prior pi (stationarity of a
  reversible chain), checked on the reference row, which is never masked;
* two sibling leaves a, b are jointly distributed as pi_x P(t_a + t_b)_xy
  (reversibility: the path through their parent is one branch of length
  t_a + t_b), with P from the oracle's P(t) -- so the generator's own
  transition matrices, row sampling and tree walk are all on trial;
* masks: i.i.d. gaps at rate gap_p with no correlation between neighbouring
  columns; gap runs covering gap_p with geometric lengths of mean gap_run;
  N codons at rate n_p with the N in each of the three positions equally
  often; the nucleotide output decodes to the code output of the same seed.

All tests are chi-squared or binomial checks at fixed seeds, held to
p >= 1e-4 (z <= 4).
"""
import os

import numpy as np
import pytest
from scipy import stats

import phylocsf_amd as P
from oracle import oracle as O


@pytest.fixture(scope="module")
def flies(params_dir):
    return P.PhyloCSFModel.from_paramset(None, "12flies")


def _chi2_p(counts, expected, min_expected=5.0):
    """chi-squared p-value with cells of small expectation pooled."""
    counts = np.asarray(counts, float).ravel()
    expected = np.asarray(expected, float).ravel()
    order = np.argsort(expected)
    c, e = counts[order], expected[order]
    small = np.cumsum(e) < min_expected * 4
    pooled_c = [c[small].sum()] if small.any() else []
    pooled_e = [e[small].sum()] if small.any() else []
    cc = np.concatenate([pooled_c, c[~small]])
    ee = np.concatenate([pooled_e, e[~small]])
    chi2 = float(((cc - ee) ** 2 / ee).sum())
    return stats.chi2.sf(chi2, len(cc) - 1)


def test_leaf_marginal_is_the_prior(flies):
    R, L = 200, 1000
    codes, off = flies.simulate(np.full(R, L), np.zeros(R, np.int32), np.ones(R), seed=11,
                                n_threads=8)
    counts = np.bincount(codes[:, 0], minlength=64)
    assert counts.sum() == R * L and counts[64:].sum() == 0
    p = _chi2_p(counts, flies.pi[0] * R * L)
    assert p >= 1e-4, p
    # and under the other model at another tree scale
    codes, _ = flies.simulate(np.full(R, L), np.ones(R, np.int32), np.full(R, 2.5), seed=12,
                              n_threads=8)
    assert _chi2_p(np.bincount(codes[:, 3], minlength=64), flies.pi[1] * R * L) >= 1e-4


def test_sibling_pairs_follow_the_pulley_principle(flies, params_dir):
    tree = flies.tree
    nl = tree.nl
    # a cherry: an internal node whose children are both leaves
    cherry = next(i for i in range(nl, tree.n)
                  if tree.children[i - nl][0] < nl and tree.children[i - nl][1] < nl)
    a, b = tree.children[cherry - nl]
    rho = 1.7
    R, L = 400, 2500
    codes, _ = flies.simulate(np.full(R, L), np.zeros(R, np.int32), np.full(R, rho), seed=21,
                              n_threads=8)
    counts = np.zeros((64, 64))
    np.add.at(counts, (codes[:, a], codes[:, b]), 1)
    prefix = os.path.join(params_dir, "12flies")
    S, pi = O.ecm_read(prefix + "_coding.ECM")
    Pab = O.pt(O.build_q(S, pi), pi, rho * (tree.branch[a] + tree.branch[b]))
    expected = pi[:, None] * Pab * (R * L)
    p = _chi2_p(counts, expected)
    assert p >= 1e-4, p
    # the same counts are far from independence (the test has power)
    indep = np.outer(counts.sum(1), counts.sum(0)) / counts.sum()
    assert _chi2_p(counts, indep) < 1e-12


def _z(k, n, p):
    return abs(k - n * p) / np.sqrt(n * p * (1 - p))


def test_iid_gaps_and_n_codons(flies):
    R, L = 100, 2000
    nl = flies.tree.nl
    codes, _ = flies.simulate(np.full(R, L), np.zeros(R, np.int32), np.ones(R), gap_mode=0,
                              gap_p=0.1, n_p=0.0, seed=31, n_threads=8)
    masked = codes[:, 1:] == P.PCSF_MARGINALIZE
    assert not (codes[:, 0] == P.PCSF_MARGINALIZE).any()  # the reference row is never masked
    assert _z(masked.sum(), masked.size, 0.1) <= 4
    # no correlation between neighbouring columns of one species (within regions)
    m = masked.reshape(R, L, nl - 1)
    both = (m[:, 1:, :] & m[:, :-1, :]).sum()
    assert _z(both, m[:, 1:, :].size, 0.01) <= 4
    # N codons alone
    codes, _ = flies.simulate(np.full(R, L), np.zeros(R, np.int32), np.ones(R), gap_p=0.0,
                              n_p=0.01, seed=32, n_threads=8)
    masked = codes[:, 1:] == P.PCSF_MARGINALIZE
    assert _z(masked.sum(), masked.size, 0.01) <= 4


def test_gap_runs(flies):
    R, L = 60, 5000
    nl = flies.tree.nl
    codes, _ = flies.simulate(np.full(R, L), np.zeros(R, np.int32), np.ones(R), gap_mode=1,
                              gap_p=0.25, gap_run=5.0, seed=41, n_threads=8)
    m = (codes[:, 1:] == P.PCSF_MARGINALIZE).reshape(R, L, nl - 1)
    cover = m.mean()
    # runs are correlated, so the coverage's spread is wider than a binomial's:
    # for a two-state chain the variance of the mean is inflated by (1 + r)/(1 - r),
    # r = 1 - 1/5 - p_start the lag-one correlation
    p_start = 0.25 / (5.0 * 0.75 + 0.25)
    r = 1 - 0.2 - p_start
    sd = np.sqrt(0.25 * 0.75 / m.size * (1 + r) / (1 - r))
    assert abs(cover - 0.25) <= 4 * sd, cover
    # run lengths of runs that neither start at a region's first column nor
    # reach its last: geometric with mean 5
    lengths = []
    for reg in range(R):
        for j in range(nl - 1):
            x = m[reg, :, j].astype(np.int8)
            d = np.diff(np.concatenate([[0], x, [0]]))
            starts, ends = np.flatnonzero(d == 1), np.flatnonzero(d == -1)
            for s0, e0 in zip(starts, ends):
                if s0 > 0 and e0 < L:
                    lengths.append(e0 - s0)
    lengths = np.array(lengths)
    mean_sd = np.sqrt(20.0 / len(lengths))  # geometric(1/5): variance (1 - p) / p^2 = 20
    assert abs(lengths.mean() - 5.0) <= 4 * mean_sd, lengths.mean()
    # geometric shape: P(len = 1) = 1/5
    assert _z((lengths == 1).sum(), len(lengths), 0.2) <= 4


def test_nucleotides_decode_to_codes(flies):
    lengths = np.array([50, 1, 0, 333])
    kw = dict(gap_mode=1, gap_p=0.2, gap_run=3.0, n_p=0.05, seed=51, n_threads=3)
    rm, rr = np.array([0, 1, 0, 1], np.int32), np.array([0.5, 1.0, 2.0, 3.0])
    codes, coff = flies.simulate(lengths, rm, rr, **kw)
    nuc, noff = flies.simulate_nucleotides(lengths, rm, rr, **kw)
    assert np.array_equal(noff, 3 * coff)
    idx = np.full(256, -1)
    for i, ch in enumerate(b"ACGT"):
        idx[ch] = i
    trip = idx[nuc.reshape(flies.tree.nl, -1, 3)]  # [nl, ncols, 3]
    ok = (trip >= 0).all(axis=2)
    dec = np.where(ok, 16 * trip[..., 0] + 4 * trip[..., 1] + trip[..., 2], 255).T
    assert np.array_equal(dec.astype(np.uint8), codes)
    # a masked codon is either a gap "---" or carries exactly one N
    rows = nuc.reshape(flies.tree.nl, -1, 3)
    n_count = (rows == ord("N")).sum(axis=2)
    gap = (rows == ord("-")).all(axis=2)
    assert np.array_equal(~ok.T, (gap | (n_count == 1)).T)
    # the thread count does not change the alignment
    codes1, _ = flies.simulate(lengths, rm, rr, **dict(kw, n_threads=1))
    assert np.array_equal(codes1, codes)


def test_n_position_is_uniform(flies):
    R, L = 50, 4000
    nuc, _ = flies.simulate_nucleotides(np.full(R, L), np.zeros(R, np.int32), np.ones(R),
                                        gap_p=0.0, n_p=0.02, seed=61, n_threads=8)
    rows = nuc.reshape(flies.tree.nl, -1, 3)
    pos = (rows == ord("N")).sum(axis=(0, 1))
    assert pos.sum() > 10000
    assert stats.chisquare(pos).pvalue >= 1e-4, pos
