"""GPU parity of per-region tree scales and the mle strategy.

* forward with one tree scale per (region, model) against the oracle's lpr at
  the same scales (FP64 kernel: relative 1e-10; FP32 MFMA: relative 1e-6);
* the mle driver (Fit.find_init + GSL Brent for every region in lock step)
  against the oracle's sequential restatement of the same maximiser, and the
  reference's published CLI results (src/test.ml:27-59), which use mle.
  With the FP64 kernel the maximiser follows the oracle's iterate sequence
  (a near-tie at Brent's last sqrt(eps)-sized step can flip on 1e-12-level
  differences), so rho agrees to 1e-6 and the score to 1e-6 dB.  With FP32 likelihoods
  the Brent comparisons can take another path near the flat optimum; those
  scores are held to the reference's published 0.01-wide windows and to the
  bound its 1% stopping rule allows around the FP64 run (tests/mle_bounds.py)
  plus 1e-3 dB.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import phylocsf_amd as P  # noqa: E402
from mle_bounds import brent_stop_bound_db  # noqa: E402
from oracle import cli_oracle as CO  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    return P.Context(0)


def _oracle_model(params_dir, name):
    prefix = os.path.join(params_dir, name)
    t = O.newick_tree(open(prefix + ".nh").read())
    Sc, pc = O.ecm_read(prefix + "_coding.ECM")
    Sn, pn = O.ecm_read(prefix + "_noncoding.ECM")
    return O.Model(t, [Sc, Sn], [pc, pn])


def test_forward_per_region_rho(ctx, params_dir):
    _say("per-region rho forward")
    m = P.PhyloCSFModel.from_paramset(ctx, "58mammals")
    rng = np.random.default_rng(21)
    lengths = np.array([5, 0, 64, 65, 1, 130, 63, 200, 17, 128])
    codes, off = m.simulate(lengths, rng.integers(0, 2, len(lengths)),
                            rng.uniform(0.3, 3, len(lengths)), gap_p=0.1, n_p=0.01, seed=9)
    rho = rng.uniform(0.05, 6.0, (len(lengths), 2))
    om = _oracle_model(params_dir, "58mammals")
    b = P.Batch(m, codes, off)
    for prec, tol in ((P.PCSF_FP64, 1e-10), (P.PCSF_FP32, 1e-6)):
        ll, el = b.forward(rho, prec, per_region=True)
        for r in range(len(lengths)):
            for q in range(2):
                ref_l, ref_e = om.lpr(q, rho[r, q], codes[off[r]:off[r + 1]])
                assert abs(ll[r, q] - ref_l) <= tol * max(1.0, abs(ref_l)), (prec, r, q)
                assert abs(el[r, q] - ref_e) <= 10 * tol * max(1.0, abs(ref_e)), (prec, r, q)
    # a shared-rho forward after per-region ones is unaffected
    ll1, _ = b.forward(1.0, P.PCSF_FP32)
    assert abs(ll1[2, 0] - om.lpr(0, 1.0, codes[off[2]:off[3]])[0]) < 1e-6 * abs(ll1[2, 0])


def test_mle_matches_oracle_maximizer(ctx, params_dir):
    _say("mle vs oracle")
    m = P.PhyloCSFModel.from_paramset(ctx, "12flies")
    rng = np.random.default_rng(5)
    R = 12
    lengths = rng.integers(20, 90, R)
    codes, off = m.simulate(lengths, rng.integers(0, 2, R), rng.uniform(0.3, 3.0, R),
                            gap_p=0.05, seed=3)
    om = _oracle_model(params_dir, "12flies")
    b = P.Batch(m, codes, off)
    g64 = P.score(b, "mle", P.PCSF_FP64)
    g32 = P.score(b, "mle", P.PCSF_FP32)
    for r in range(R):
        ref = om.score_mle(codes[off[r]:off[r + 1]])
        assert g64["status"][r] == 0 and g32["status"][r] == 0
        assert abs(g64["rho"][r, 0] - ref["rho_c"]) <= 1e-6 * ref["rho_c"], r
        assert abs(g64["rho"][r, 1] - ref["rho_n"]) <= 1e-6 * ref["rho_n"], r
        assert abs(g64["score"][r] - ref["score"]) <= 1e-6, r
        assert abs(g64["anc_comp"][r] - ref["anc_comp"]) <= 1e-5, r
    # FP32 likelihoods may steer Brent along another path: held to the 1%
    # stopping rule's own bound around the FP64 run, plus the 1e-3 dB rounding
    # contract of one evaluation
    bound = brent_stop_bound_db(b, g64["rho"], g32["rho"]) + 1e-3
    d = np.abs(g32["score"] - g64["score"])
    _say("  FP32 mle vs FP64 mle: max |d score| %.3g dB, max d/bound %.3g"
         % (d.max(), np.max(d / bound)))
    assert np.all(d <= bound)


def _regions_from_mfa(m, text, frames=1, orf="AsIs", remove_ref_gaps=False, min_codons=25):
    sp, seqs = CO.input_mfa(text)
    if remove_ref_gaps:
        seqs = CO.remove_ref_gaps(seqs)
    rc_seqs = [CO.revcomp(s) for s in seqs]
    which = {s: i for i, s in enumerate(sp)}
    leaf_ord = [which.get(m.tree.labels[i]) for i in range(m.tree.nl)]
    rgns = CO.candidate_regions(seqs[0], frames, orf, min_codons)
    blocks = [CO.pleaves(rc_seqs if rc else seqs, leaf_ord, lo, hi) for rc, lo, hi in rgns]
    return rgns, blocks


def _best(res, rgns):
    keys = [(res["score"][i], res["anc_comp"][i], rgns[i][0], rgns[i][1], rgns[i][2])
            for i in range(len(rgns))]
    bi = 0
    for i in range(1, len(keys)):
        if keys[i] > keys[bi]:
            bi = i
    return bi


@pytest.mark.parametrize("prec", [P.PCSF_FP64, P.PCSF_FP32])
def test_published_examples_mle(ctx, examples_dir, prec):
    """src/test.ml:27-59 through the GPU mle driver."""
    _say("published examples, precision %d" % prec)
    cases = [
        ("12flies", "tal-AA.fa", {}, (297.62, 297.63), (48.25, 48.26), None),
        ("29mammals", "ALDH2.exon5.fa", {}, (-178.93, -178.92), (-38.29, -38.28), None),
        ("29mammals", "ALDH2.exon5.fa", {"frames": 6}, (218.26, 218.27), None, (False, 1, 111)),
        ("29mammals", "Aldh2.mRNA.fa", {"frames": 3, "orf": "ATGStop", "remove_ref_gaps": True},
         (2013.92, 2013.93), None, (False, 343, 1899)),
    ]
    for pset, fa, kw, win, awin, where in cases:
        m = P.PhyloCSFModel.from_paramset(ctx, pset)
        text = open(os.path.join(examples_dir, fa)).read()
        rgns, blocks = _regions_from_mfa(m, text, **kw)
        b = m.batch(blocks)
        res = P.score(b, "mle", prec)
        bi = _best(res, rgns)
        assert win[0] < res["score"][bi] < win[1], (fa, kw, res["score"][bi])
        if awin:
            assert awin[0] < res["anc_comp"][bi] < awin[1], (fa, res["anc_comp"][bi])
        if where:
            assert tuple(rgns[bi]) == where, (fa, rgns[bi])


def test_forward_per_region_many_sets(ctx, params_dir):
    """Enough (region, model) sets that the P builder makes several branches per
    workgroup (100vertebrates: 240 sets x 198 branches), with a branch count
    that is not a multiple of that group; sampled regions against the oracle."""
    _say("per-region rho forward, 100vertebrates, 120 regions")
    m = P.PhyloCSFModel.from_paramset(ctx, "100vertebrates")
    rng = np.random.default_rng(23)
    n = 120
    lengths = rng.integers(1, 9, n)
    codes, off = m.simulate(lengths, rng.integers(0, 2, n), rng.uniform(0.3, 3, n),
                            gap_p=0.1, n_p=0.01, seed=17)
    rho = rng.uniform(0.05, 6.0, (n, 2))
    om = _oracle_model(params_dir, "100vertebrates")
    b = P.Batch(m, codes, off)
    for prec, tol in ((P.PCSF_FP64, 1e-10), (P.PCSF_FP32, 1e-6)):
        ll, el = b.forward(rho, prec, per_region=True)
        for r in range(0, n, 7):
            for q in range(2):
                ref_l, ref_e = om.lpr(q, rho[r, q], codes[off[r]:off[r + 1]])
                assert abs(ll[r, q] - ref_l) <= tol * max(1.0, abs(ref_l)), (prec, r, q)
                assert abs(el[r, q] - ref_e) <= 10 * tol * max(1.0, abs(ref_e)), (prec, r, q)


def test_mle_find_init_retry_path_matches_oracle(ctx, params_dir):
    """Regions whose likelihood at rho = 1 loses to an end of [0.01, 10] send
    Fit.find_init into its retry loop (lib/CamlPaml/Fit.ml:35-47), drawing
    from OCaml's Random.init 0 stream: the library's restatement of that
    stream (host/maximize.cpp) and the oracle's must give the same retries,
    hence the same starting point for Brent and the same estimates (a
    different start ends elsewhere inside Brent's 1% interval)."""
    _say("mle retry path vs oracle")
    m = P.PhyloCSFModel.from_paramset(ctx, "12flies")
    lengths = np.array([60, 80, 40, 70, 50, 90])
    true_rho = np.array([0.012, 0.02, 0.03, 0.05, 9.0, 12.0])
    codes, off = m.simulate(lengths, np.array([0, 1, 0, 1, 0, 1]), true_rho, seed=77)
    om = _oracle_model(params_dir, "12flies")
    b = P.Batch(m, codes, off)
    g = P.score(b, "mle", P.PCSF_FP64)
    retried = 0
    for r in range(len(lengths)):
        ref = om.score_mle(codes[off[r]:off[r + 1]])
        assert g["status"][r] == 0
        assert abs(g["rho"][r, 0] - ref["rho_c"]) <= 1e-6 * ref["rho_c"], r
        assert abs(g["rho"][r, 1] - ref["rho_n"]) <= 1e-6 * ref["rho_n"], r
        assert abs(g["score"][r] - ref["score"]) <= 1e-6, r
        for q in range(2):
            l1, _ = om.lpr(q, 1.0, codes[off[r]:off[r + 1]])
            llo, _ = om.lpr(q, 0.01, codes[off[r]:off[r + 1]])
            lhi, _ = om.lpr(q, 10.0, codes[off[r]:off[r + 1]])
            retried += int(l1 <= llo or l1 <= lhi)
    assert retried >= 6  # the retry loop ran for most (region, model) pairs


@pytest.mark.parametrize("paramset", ["12flies", "100vertebrates"])
def test_region_chunks_agree_with_one_chunk(ctx, paramset, monkeypatch):
    """Per-region passes go through the regions in chunks of as many
    transition sets as fit the device budget (PHYLOCSF_SET_BUDGET_MB): with
    a 1 MiB budget every region is a chunk of its own, and the per-region
    forwards (FP32, FP64, FP64 on matrix cores) and both maximisers give what
    one chunk gives.  Columns and regions are computed independently of the
    chunking (one workgroup per region in the reduction), so the tolerances
    only allow for nothing."""
    _say(f"region chunks, {paramset}")
    m = P.PhyloCSFModel.from_paramset(ctx, paramset)
    rng = np.random.default_rng(8)
    R = 23
    lengths = rng.integers(1, 150, R)
    lengths[3] = 0
    codes, off = m.simulate(lengths, rng.integers(0, 2, R), rng.uniform(0.3, 3.0, R),
                            gap_p=0.1, seed=12)
    rho = rng.uniform(0.05, 6.0, (R, 2))

    def run():
        b = P.Batch(m, codes, off)
        out = {}
        for prec in (P.PCSF_FP32, P.PCSF_FP64, P.PCSF_FP64M):
            out[prec] = b.forward(rho, prec, per_region=True)
        out["grad"] = P.score(b, "mle_grad", P.PCSF_FP16S)
        out["brent"] = P.score(b, "mle", P.PCSF_FP64M)
        b.close()
        return out

    monkeypatch.delenv("PHYLOCSF_SET_BUDGET_MB", raising=False)
    one = run()
    monkeypatch.setenv("PHYLOCSF_SET_BUDGET_MB", "1")
    many = run()
    for prec in (P.PCSF_FP32, P.PCSF_FP64, P.PCSF_FP64M):
        for a, c in zip(one[prec], many[prec]):
            np.testing.assert_allclose(c, a, rtol=1e-13, atol=0, err_msg=str(prec))
    for k in ("grad", "brent"):
        np.testing.assert_array_equal(many[k]["status"], one[k]["status"])
        np.testing.assert_allclose(many[k]["score"], one[k]["score"], rtol=0, atol=1e-9)
        np.testing.assert_allclose(many[k]["rho"], one[k]["rho"], rtol=1e-9, atol=0)
