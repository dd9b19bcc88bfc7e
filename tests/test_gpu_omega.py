"""The omega test (pcsf_score_omega) on the GPU against the oracle's
restatement of OmegaModel.score (src/OmegaModel.ml:160-219; paths relative to
the mlin/PhyloCSF tree).

With FP64 likelihoods the library's lock-step kr_map follows the oracle's
maximiser (find_init, GSL Brent, three rounds of coordinate ascent): each
Brent run stops once its bracket is within 1% of its estimate, and on the
flat optimum of kappa (likelihood and gamma prior) a comparison in the last
steps can fall either way on the 1e-13-level differences between the
library's GPU eigensolver (parallel-order Jacobi) and the oracle's; six
chained maximisations carry such a difference along.  The estimates then
agree to 1e-4 relative (measured: 1.3e-5 at most), each hypothesis's MAP lpr
-- flat to second order there -- to 1e-8 relative (measured: 1.4e-9), and the
scores to what that allows, 1e-5 dB + DB 1e-8 (|lpr_H0| + |lpr_H1|).
Checked on simulated regions and
on the reference's example alignments, for the default H1 (omega 0.2, sigma
0.01) and another one.  The MFMA precisions change the likelihoods by their
rounding only; the maximiser may then stop elsewhere inside its 1% intervals,
so they are held to 0.05 dB.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import phylocsf_amd as P  # noqa: E402
from oracle import cli_oracle as CO  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    return P.Context(0)


def _sim_regions(paramset, lengths, seed):
    """Regions simulated under the coding and noncoding ECMs (any codon
    alignments will do for the omega test)."""
    m = P.PhyloCSFModel.from_paramset(None, paramset)
    rng = np.random.default_rng(seed)
    R = len(lengths)
    return m.simulate(lengths, rng.integers(0, 2, R), rng.uniform(0.3, 3.0, R), gap_p=0.1,
                      seed=seed)


def _check(res, codes, off, otree, h1, tol_rho=1e-4, tol_db=1e-5):
    worst = [0.0, 0.0]
    for r in range(len(off) - 1):
        ref = O.omega_score(otree, codes[off[r]:off[r + 1]], *h1)
        worst[0] = max(worst[0], abs(res["score"][r] - ref["score"]))
        worst[1] = max(worst[1], max(abs(res["kappa"][r, h] / ref["kappa_" + k] - 1)
                                     for h, k in ((0, "H0"), (1, "H1"))))
        assert res["status"][r] == 0, r
        tol = tol_db + P.DB * 1e-8 * (abs(ref["lpr_H0"]) + abs(ref["lpr_H1"]))
        assert abs(res["score"][r] - ref["score"]) <= tol, (r, res["score"][r], ref["score"])
        for h, key in ((0, "H0"), (1, "H1")):
            assert abs(res["rho"][r, h] - ref["rho_" + key]) <= tol_rho * ref["rho_" + key], (r, h)
            assert abs(res["kappa"][r, h] - ref["kappa_" + key]) <= tol_rho * ref["kappa_" + key]
            assert abs(res["lpr"][r, h] - ref["lpr_" + key]) <= 1e-8 * abs(ref["lpr_" + key])
    _say("  max |d score| %.3g dB, max relative d kappa %.3g" % tuple(worst))


@pytest.mark.parametrize("paramset", ["12flies", "29mammals"])
def test_omega_matches_oracle(ctx, params_dir, paramset):
    _say("omega test vs oracle, " + paramset)
    tree = P.newick_tree(open(os.path.join(params_dir, paramset + ".nh")).read())
    otree = O.newick_tree(open(os.path.join(params_dir, paramset + ".nh")).read())
    codes, off = _sim_regions(paramset, np.array([40, 1, 75, 12]), 3)
    m = P.PhyloCSFModel.tree_only(ctx, tree)
    b = P.Batch(m, codes, off)
    for h1 in ((0.2, 0.01), (0.5, 0.3)):
        res = P.score_omega(b, *h1, precision=P.PCSF_FP64)
        _check(res, codes, off, otree, h1)


def test_omega_example_alignments(ctx, params_dir, examples_dir):
    """The reference's examples through the omega strategy's per-region path
    (frames and ORFs as the CLI makes them)."""
    for paramset, fa, kw in (("12flies", "tal-AA.fa", {}),
                             ("29mammals", "ALDH2.exon5.fa", {"frames": 3})):
        _say("omega test, " + fa)
        tree = P.newick_tree(open(os.path.join(params_dir, paramset + ".nh")).read())
        otree = O.newick_tree(open(os.path.join(params_dir, paramset + ".nh")).read())
        sp, seqs = CO.input_mfa(open(os.path.join(examples_dir, fa)).read())
        which = {s: i for i, s in enumerate(sp)}
        leaf_ord = [which.get(tree.labels[i]) for i in range(tree.nl)]
        rgns = CO.candidate_regions(seqs[0], kw.get("frames", 1), "AsIs", 25)
        rc_seqs = [CO.revcomp(s) for s in seqs]
        blocks = [CO.pleaves(rc_seqs if rc else seqs, leaf_ord, lo, hi) for rc, lo, hi in rgns]
        m = P.PhyloCSFModel.tree_only(ctx, tree)
        b = m.batch(blocks)
        res = P.score_omega(b, precision=P.PCSF_FP64)
        codes = np.concatenate(blocks)
        off = np.cumsum([0] + [x.shape[0] for x in blocks])
        _check(res, codes, off, otree, (0.2, 0.01))


def test_omega_mfma_precisions(ctx, params_dir):
    _say("omega test at FP32 and split FP16 vs FP64")
    codes, off = _sim_regions("29mammals", np.array([60, 120, 33]), 8)
    tree = P.newick_tree(open(os.path.join(params_dir, "29mammals.nh")).read())
    b = P.Batch(P.PhyloCSFModel.tree_only(ctx, tree), codes, off)
    r64 = P.score_omega(b, precision=P.PCSF_FP64)
    for prec in (P.PCSF_FP32, P.PCSF_FP16S):
        r = P.score_omega(b, precision=prec)
        assert (r["status"] == 0).all()
        d = np.abs(r["score"] - r64["score"])
        _say("  precision %d: max |d score| %.3g dB" % (prec, d.max()))
        assert d.max() <= 0.05


def test_omega_warm_started_eigensystems(ctx, params_dir, monkeypatch):
    """With every (region, hypothesis) set in one chunk the rate models'
    eigensystems start from each set's first one (and unchanged sets keep
    theirs); with a 1 MiB set budget every region is a chunk of its own and
    every eigensystem is built cold.  Both follow the same maximiser, so they
    agree as the oracle comparison does: estimates to 1e-4 relative, MAP lpr
    to 1e-8 relative, scores to 1e-5 dB + DB 1e-8 (|lpr_H0| + |lpr_H1|)."""
    _say("omega test, warm-started against cold eigensystems")
    codes, off = _sim_regions("29mammals", np.array([90, 14, 150, 60, 33, 120]), 11)
    tree = P.newick_tree(open(os.path.join(params_dir, "29mammals.nh")).read())
    m = P.PhyloCSFModel.tree_only(ctx, tree)
    monkeypatch.delenv("PHYLOCSF_SET_BUDGET_MB", raising=False)
    warm = P.score_omega(P.Batch(m, codes, off), precision=P.PCSF_FP64M)
    monkeypatch.setenv("PHYLOCSF_SET_BUDGET_MB", "1")
    cold = P.score_omega(P.Batch(m, codes, off), precision=P.PCSF_FP64M)
    assert (warm["status"] == 0).all() and (cold["status"] == 0).all()
    for r in range(len(off) - 1):
        tol = 1e-5 + P.DB * 1e-8 * (abs(cold["lpr"][r, 0]) + abs(cold["lpr"][r, 1]))
        assert abs(warm["score"][r] - cold["score"][r]) <= tol, r
        for h in range(2):
            assert abs(warm["lpr"][r, h] - cold["lpr"][r, h]) <= 1e-8 * abs(cold["lpr"][r, h])
            for key in ("rho", "kappa"):
                assert abs(warm[key][r, h] - cold[key][r, h]) <= 1e-4 * cold[key][r, h], (r, h)


def test_tree_only_model_refuses_ecm_entry_points(ctx, params_dir):
    tree = P.newick_tree(open(os.path.join(params_dir, "12flies.nh")).read())
    m = P.PhyloCSFModel.tree_only(ctx, tree)
    b = P.Batch(m, np.zeros((5, tree.nl), np.uint8), np.array([0, 5]))
    with pytest.raises(P.InvalidArgument):
        b.forward(1.0, P.PCSF_FP32)
    with pytest.raises(P.InvalidArgument):
        P.score(b, "mle")
    with pytest.raises(P.InvalidArgument):
        m.transition(0, 1.0)
