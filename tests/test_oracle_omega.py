"""Pins the oracle's restatement of the omega test (src/OmegaModel.ml; paths
relative to the mlin/PhyloCSF tree) by what the mathematics fixes:

* with kappa = omega = sigma = 1 and uniform codon-position frequencies the
  codon process is three independent Jukes-Cantor nucleotide processes, so
  P(t) is the Kronecker cube of JC69's P (rate 1/9 per alternative after the
  mean-rate normalisation);
* for any settings Q is reversible with respect to the codon frequencies
  (detailed balance) and has them as its stationary distribution;
* the exchangeabilities, codon by codon: 9 neighbours per codon, kappa on
  transitions only, omega on nonsynonymous changes between sense codons
  only (checked on hand-picked pairs and by counts from the genetic code);
* F3x4 against counts made with numpy; the two priors against scipy's
  Cauchy and gamma densities;
* the MAP fit: H1 with H0's settings (omega_H1 = sigma_H1 = 1) scores 0 up to
  the maximiser's tolerance, and alignments simulated under purifying
  selection / neutrally score positive / negative.
"""
import os

import numpy as np
import pytest
import scipy.linalg
from scipy import stats

import phylocsf_amd as P
from oracle import oracle as O

AA = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"
IDX = {c: i for i, c in enumerate("ACGT")}


def codon(s):
    return 16 * IDX[s[0]] + 4 * IDX[s[1]] + IDX[s[2]]


def _q(qs):
    pi = O.omega_pi(qs)
    return O.build_q(O.omega_s(qs), pi), pi


def test_neutral_uniform_model_is_three_jukes_cantor_processes():
    qs = np.ones(12)
    Q, pi = _q(qs)
    assert np.allclose(pi, 1 / 64, rtol=0, atol=1e-17)
    for t in (0.05, 0.4, 2.0):
        e = np.exp(-4 * t / 9)
        J = np.full((4, 4), 0.25 - 0.25 * e)
        np.fill_diagonal(J, 0.25 + 0.75 * e)
        K = np.kron(np.kron(J, J), J)
        assert np.abs(O.pt(Q, pi / pi.sum(), t) - K).max() < 1e-13
        assert np.abs(scipy.linalg.expm(Q * t) - K).max() < 1e-13


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_reversible_and_stationary(seed):
    rng = np.random.default_rng(seed)
    qs = np.concatenate([[rng.uniform(1, 5), rng.uniform(0.05, 2), rng.uniform(0.01, 1)],
                         rng.uniform(0.3, 3, 9)])
    Q, pi = _q(qs)
    p = pi / pi.sum()
    F = p[:, None] * Q
    assert np.abs(F - F.T).max() < 1e-15
    assert np.abs(p @ Q).max() < 1e-15
    assert np.abs(Q.sum(axis=1)).max() < 1e-14
    # the mean-rate normalisation uses the settings' (unnormalised) frequencies
    assert abs(-(pi * np.diag(Q)).sum() - 1.0) < 1e-12


def test_exchangeabilities_codon_by_codon():
    kappa, omega = 3.0, 0.25
    S = O.omega_s(np.array([kappa, omega, 1.0] + [1.0] * 9))
    assert ((S > 0).sum(axis=1) == 9).all() and np.allclose(S, S.T)
    cases = [("TTT", "TTC", kappa),          # Phe -> Phe, T<->C transition
             ("GCT", "GAT", omega),          # Ala -> Asp, C->A transversion
             ("GCT", "GTT", kappa * omega),  # Ala -> Val, C->T transition
             ("TAA", "CAA", kappa),          # stop -> Gln: no omega
             ("TGG", "TAG", kappa),          # Trp -> stop, G->A transition
             ("TGG", "TGT", omega),          # Trp -> Cys, G->T transversion
             ("CTA", "TTA", kappa),          # Leu -> Leu at the first position
             ("AAA", "CCC", 0.0)]            # three changes apart
    for a, b, want in cases:
        assert S[codon(a), codon(b)] == pytest.approx(want, rel=1e-15), (a, b)
    # counts from the genetic code: nonsynonymous single changes between sense codons
    nonsyn = 0
    for i in range(64):
        for j in range(64):
            d = [((i >> s) & 3) != ((j >> s) & 3) for s in (4, 2, 0)]
            if sum(d) == 1 and AA[i] != "*" and AA[j] != "*" and AA[i] != AA[j]:
                nonsyn += 1
    Sw = O.omega_s(np.array([1.0, 7.0, 1.0] + [1.0] * 9))
    assert int((Sw == 7.0).sum()) == nonsyn


def test_codon_frequencies_and_stop_factor():
    rng = np.random.default_rng(5)
    r = rng.uniform(0.3, 3, 9)
    for sigma in (1.0, 0.01):
        qs = np.concatenate([[2.0, 0.5, sigma], r])
        pi = O.omega_pi(qs)
        f = [np.append(r[3 * p:3 * p + 3], 1.0) for p in range(3)]
        f = [x / x.sum() for x in f]
        sc = np.einsum("a,b,c->abc", *f).ravel()
        denom = 1 - (1 - sigma) * (sc[codon("TAA")] + sc[codon("TAG")] + sc[codon("TGA")])
        assert np.allclose(pi, sc / denom, rtol=1e-15, atol=0)


def test_f3x4_counts():
    rng = np.random.default_rng(6)
    codes = rng.integers(0, 64, (300, 12)).astype(np.uint8)
    codes[rng.random(codes.shape) < 0.2] = 255
    got = O.omega_f3x4(codes)
    v = codes[codes != 255].astype(int)
    cnt = np.ones((3, 4))
    for p, sh in enumerate((4, 2, 0)):
        cnt[p] += np.bincount((v >> sh) & 3, minlength=4)
    assert np.allclose(got, (cnt[:, :3] / cnt[:, 3:4]).ravel(), rtol=1e-15)


def test_priors_against_scipy():
    c = stats.cauchy(loc=1.0, scale=0.5)
    for x in (0.001, 0.3, 1.0, 2.5, 10.0):
        want = c.logpdf(x) - np.log(c.sf(0.0))
        assert O.half_cauchy_lpdf(1.0, 0.5, x) == pytest.approx(want, rel=1e-13)
    g = stats.gamma(a=7.0, scale=0.25)
    for k in (1.0, 1.5, 2.5, 4.0, 10.0):
        want = g.logpdf(k - 1 + np.finfo(float).eps)
        assert O.lpr_kappa(k) == pytest.approx(want, rel=1e-12, abs=1e-12)


@pytest.fixture(scope="module")
def flies(params_dir):
    return O.newick_tree(open(os.path.join(params_dir, "12flies.nh")).read())


def _simulate(tree_text, qs, lengths, seed, rho=1.0):
    """Alignments simulated under the omega model's Q (rate scale aside)."""
    S = O.omega_s(qs)
    pi = O.omega_pi(qs)
    pi = pi / pi.sum()
    m = P.PhyloCSFModel(None, P.newick_tree(tree_text), [(S, pi)])
    return m.simulate(lengths, np.zeros(len(lengths), np.int32), np.full(len(lengths), rho),
                      seed=seed)


def test_map_fit_identical_hypotheses_score_zero(flies, params_dir):
    text = open(os.path.join(params_dir, "12flies.nh")).read()
    codes, off = _simulate(text, np.array([2.0, 0.3, 1.0] + [1.0] * 9), [150], 3)
    r = O.omega_score(flies, codes, omega_H1=1.0, sigma_H1=1.0)
    # H1 restarts kr_map at H0's estimates with H0's settings: it can only
    # climb, by at most what Brent's 1% stopping interval left
    assert -1e-9 <= r["score"] < 0.05, r
    assert r["lpr_H1"] >= r["lpr_H0"] - 1e-9


def test_selection_is_detected(flies, params_dir):
    text = open(os.path.join(params_dir, "12flies.nh")).read()
    pur, off = _simulate(text, np.array([2.0, 0.05, 0.01] + [1.0] * 9), [300], 11, rho=2.0)
    neu, _ = _simulate(text, np.array([2.0, 1.0, 1.0] + [1.0] * 9), [300], 12, rho=2.0)
    s_pur = O.omega_score(flies, pur)["score"]
    s_neu = O.omega_score(flies, neu)["score"]
    assert s_pur > 10 and s_neu < 0, (s_pur, s_neu)
