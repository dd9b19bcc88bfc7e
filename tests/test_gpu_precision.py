"""GPU parity of the reduced-precision GEMM inputs of the fused kernel.

* split FP16 (PCSF_FP16S: P and the partial as hi + lo FP16 pairs, three MFMA
  products, FP32 accumulation) is held to the FP32 contract: |d score|
  <= 1e-3 dB and relative log-likelihood <= 1e-6 against the FP64 path, on
  the fixed and the per-region (mle) paths;
* plain FP16 and BF16 inputs are reported per precision.  Rounding P to 11
  or 8 significant bits changes the model itself, so they do not meet the
  1e-3 dB contract; they are bounded loosely here (a wrong fragment layout
  gives unrelated numbers and still fails).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import phylocsf_amd as P  # noqa: E402
from phylocsf_amd import workloads as W  # noqa: E402
from mle_bounds import brent_stop_bound_db  # noqa: E402
from oracle import oracle as O  # noqa: E402

TOL_DB = 1e-3
TOL_REL = 1e-6
# loose sanity bounds of the single-rounding precisions (relative lpr)
LOOSE_REL = {P.PCSF_FP16: 2e-3, P.PCSF_BF16: 2e-2}
NAMES = {P.PCSF_FP32: "FP32", P.PCSF_FP16S: "FP16 split", P.PCSF_FP16: "FP16",
         P.PCSF_BF16: "BF16"}


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


def _errors(res, ref):
    d = np.abs(res["score"] - ref["score"])
    rel = max(float(np.max(np.abs(res[k] - ref[k]) / np.maximum(1.0, np.abs(ref[k]))))
              for k in ("lpr_c", "lpr_n"))
    return float(d.max()), float(d.mean()), rel


def test_precisions_c2_sample(ctx):
    _say("C2 sample, every precision vs FP64")
    m = W.model_for("C2", ctx)
    wl = W.make("C2", m, n_regions=400)
    b = P.Batch(m, wl.codes, wl.offsets)
    ref = P.score_fixed(b, P.PCSF_FP64)
    for prec in (P.PCSF_FP32, P.PCSF_FP16S, P.PCSF_FP16, P.PCSF_BF16):
        dmax, dmean, rel = _errors(P.score_fixed(b, prec), ref)
        _say("  %-10s max |d score| %.3g dB (mean %.3g), max rel lpr %.3g"
             % (NAMES[prec], dmax, dmean, rel))
        if prec in (P.PCSF_FP32, P.PCSF_FP16S):
            assert dmax <= TOL_DB and rel <= TOL_REL, NAMES[prec]
        else:
            assert rel <= LOOSE_REL[prec], NAMES[prec]


def test_fp16s_against_oracle(ctx, params_dir):
    _say("FP16 split vs oracle: C1 and 100vertebrates")
    m = W.model_for("C1", ctx)
    wl = W.make("C1", m)
    om = _oracle_model(params_dir, "12flies")
    res = P.score_fixed(P.Batch(m, wl.codes, wl.offsets), P.PCSF_FP16S)
    ref = om.score_fixed(wl.codes)
    assert abs(res["score"][0] - ref["score"]) <= TOL_DB
    assert abs(res["lpr_c"][0] - ref["lpr_c"]) <= TOL_REL * abs(ref["lpr_c"])
    m = P.PhyloCSFModel.from_paramset(ctx, "100vertebrates")
    rng = np.random.default_rng(8)
    lengths = np.array([30, 1, 77, 0, 130])
    codes, off = m.simulate(lengths, rng.integers(0, 2, len(lengths)),
                            rng.uniform(0.3, 3, len(lengths)), gap_p=0.15, n_p=0.01, seed=4)
    om = _oracle_model(params_dir, "100vertebrates")
    res = P.score_fixed(P.Batch(m, codes, off), P.PCSF_FP16S)
    for r in range(len(lengths)):
        s = om.score_fixed(codes[off[r]:off[r + 1]])
        assert abs(res["score"][r] - s["score"]) <= TOL_DB, r
        for k in ("lpr_c", "lpr_n"):
            assert abs(res[k][r] - s[k]) <= TOL_REL * max(1.0, abs(s[k])), (r, k)


def test_fp16s_per_region_scales(ctx, params_dir):
    _say("FP16 split, per-region tree scales vs oracle")
    m = P.PhyloCSFModel.from_paramset(ctx, "58mammals")
    rng = np.random.default_rng(12)
    lengths = np.array([5, 0, 64, 65, 1, 130, 63, 200])
    codes, off = m.simulate(lengths, rng.integers(0, 2, len(lengths)),
                            rng.uniform(0.3, 3, len(lengths)), gap_p=0.1, n_p=0.01, seed=6)
    rho = rng.uniform(0.05, 6.0, (len(lengths), 2))
    om = _oracle_model(params_dir, "58mammals")
    b = P.Batch(m, codes, off)
    for prec in (P.PCSF_FP16S, P.PCSF_FP16, P.PCSF_BF16):
        ll, _ = b.forward(rho, prec, per_region=True)
        worst = 0.0
        for r in range(len(lengths)):
            for q in range(2):
                ref_l, _ = om.lpr(q, rho[r, q], codes[off[r]:off[r + 1]])
                worst = max(worst, abs(ll[r, q] - ref_l) / max(1.0, abs(ref_l)))
        _say("  %-10s max rel lpr %.3g" % (NAMES[prec], worst))
        assert worst <= (TOL_REL if prec == P.PCSF_FP16S else LOOSE_REL[prec]), NAMES[prec]


def test_fp16s_mle(ctx):
    _say("FP16 split mle vs FP64 mle")
    m = P.PhyloCSFModel.from_paramset(ctx, "12flies")
    rng = np.random.default_rng(5)
    R = 10
    lengths = rng.integers(20, 90, R)
    codes, off = m.simulate(lengths, rng.integers(0, 2, R), rng.uniform(0.3, 3.0, R),
                            gap_p=0.05, seed=3)
    b = P.Batch(m, codes, off)
    g64 = P.score(b, "mle", P.PCSF_FP64)
    g16 = P.score(b, "mle", P.PCSF_FP16S)
    assert (g16["status"] == 0).all()
    bound = brent_stop_bound_db(b, g64["rho"], g16["rho"]) + TOL_DB
    d = np.abs(g16["score"] - g64["score"])
    _say("  max |d score| %.3g dB, max d/bound %.3g" % (d.max(), np.max(d / bound)))
    assert np.all(d <= bound)
    # the gradient maximiser: 1e-3 dB of its own FP64 run
    q64 = P.score(b, "mle_grad", P.PCSF_FP64)
    q16 = P.score(b, "mle_grad", P.PCSF_FP16S)
    assert (q16["status"] == 0).all()
    assert np.all(np.abs(q16["score"] - q64["score"]) <= TOL_DB)


def test_edge_columns_fp16s(ctx):
    _say("FP16 split edge columns")
    m = P.PhyloCSFModel.from_paramset(ctx, "29mammals")
    nl = m.tree.nl
    allm = np.full((7, nl), P.PCSF_MARGINALIZE, dtype=np.uint8)
    one = np.full((3, nl), P.PCSF_MARGINALIZE, dtype=np.uint8)
    one[:, 5] = [0, 17, 63]
    codes = np.concatenate([allm, one])
    off = np.array([0, 7, 10])
    b = P.Batch(m, codes, off)
    r16 = P.score_fixed(b, P.PCSF_FP16S)
    r64 = P.score_fixed(b, P.PCSF_FP64)
    assert abs(r16["lpr_c"][0]) <= 1e-9 and abs(r16["lpr_n"][0]) <= 1e-9
    for k in ("lpr_c", "lpr_n"):
        assert abs(r16[k][1] - r64[k][1]) <= TOL_REL * abs(r64[k][1])


def test_autotune_each_precision(ctx):
    _say("autotune per precision")
    m = W.model_for("C2", ctx)
    wl = W.make("C2", m, n_regions=60)
    b = P.Batch(m, wl.codes, wl.offsets)
    ref = P.score_fixed(b, P.PCSF_FP64)
    for prec in (P.PCSF_FP32, P.PCSF_FP16S):
        chosen, ms = b.autotune(reps=1, precision=prec)
        assert 0 <= chosen < P.PCSF_N_VARIANTS and np.all(ms > 0)
        dmax, _, _ = _errors(P.score_fixed(b, prec), ref)
        assert dmax <= TOL_DB
    with pytest.raises(P.InvalidArgument):
        b.autotune(reps=1, precision=P.PCSF_FP64)


@pytest.mark.parametrize("specialized", [True, False])
def test_every_launch_shape(ctx, specialized):
    """Every launch shape of the fused kernel (waves per workgroup, register
    cap, gather prefetch, 16-column tiles per wave), tree-specialised and
    interpreted, fixed and per-region scales, against FP64."""
    _say("every launch shape, specialized=%s" % specialized)
    m = W.model_for("C2", ctx)
    m.specialized(specialized)
    wl = W.make("C2", m, n_regions=40)
    b = P.Batch(m, wl.codes, wl.offsets)
    ref = P.score_fixed(b, P.PCSF_FP64)
    rng = np.random.default_rng(4)
    rho = rng.uniform(0.3, 3.0, (b.n_regions, 2))
    ll64, _ = b.forward(rho, P.PCSF_FP64, per_region=True)
    try:
        for prec in (P.PCSF_FP32, P.PCSF_FP16S):
            for v in range(P.PCSF_N_VARIANTS):
                m.set_variant(v, prec)
                dmax, _, _ = _errors(P.score_fixed(b, prec), ref)
                assert dmax <= TOL_DB, (prec, v, dmax)
                ll, _ = b.forward(rho, prec, per_region=True)
                rel = np.max(np.abs(ll - ll64) / np.abs(ll64))
                assert rel <= 1e-6, (prec, v, rel)
    finally:
        m.specialized(True)
