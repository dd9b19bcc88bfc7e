"""FP64 on matrix cores (PCSF_FP64M): the inside pass of the FP64
v_mfma_f64_16x16x4f64 kernel as a forward.

It computes what the plain FP64 kernel computes (lpr_leaves and the root's
expected ln pi per region, src/PhyloCSFModel.ml:48-62 over
lib/CamlPaml/PhyloLik.ml:73-93; paths relative to the mlin/PhyloCSF tree)
with FP64 products and sums in another order, so:

* per-region forwards at random tree scales agree with the plain FP64 kernel
  to 1e-12 relative (both are within ~1e-14 of exact on these sizes);
* the maximisers and the omega test at PCSF_FP64M follow the same searches
  as at PCSF_FP64 (scores within 1e-5 dB: Brent's sqrt(eps)-sized last steps
  can fall either way on 1e-15-level differences);
* a 200-leaf tree, whose FP64 workspace (199 internal nodes x 64 x 8 bytes)
  exceeds the plain kernel's 64 KiB of LDS, is scored against the oracle;
  the plain kernel refuses it with Invalid_argument;
* the no-allocation forward refuses PCSF_FP64M.
The fixed-scale parity against the oracle on every parameter set and the
edge cases (empty, one-column and all-`Marginalize regions) are in
test_gpu_parity.py.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import phylocsf_amd as P  # noqa: E402
from phylocsf_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    return P.Context(0)


def test_per_region_forward_matches_fp64(ctx):
    _say("FP64M per-region forward vs FP64")
    m = W.model_for("C2", ctx)
    rng = np.random.default_rng(21)
    lengths = np.array([0, 1, 63, 64, 65, 300, 17, 1000, 2, 129])
    codes, off = m.simulate(lengths, rng.integers(0, 2, len(lengths)),
                            rng.uniform(0.3, 3, len(lengths)), gap_p=0.2, n_p=0.02, seed=3)
    b = P.Batch(m, codes, off)
    rho = rng.uniform(0.05, 5.0, (len(lengths), 2))
    l64, e64 = b.forward(rho, P.PCSF_FP64, per_region=True)
    lm, em = b.forward(rho, P.PCSF_FP64M, per_region=True)
    assert np.all(np.abs(lm - l64) <= 1e-12 * np.maximum(1.0, np.abs(l64)))
    assert np.all(np.abs(em - e64) <= 1e-12 * np.maximum(1.0, np.abs(e64)))
    assert lm[0, 0] == 0.0 and lm[0, 1] == 0.0  # the empty region
    # shared scale on the same batch (the model's transition sets)
    l64, e64 = b.forward(0.7, P.PCSF_FP64)
    lm, em = b.forward(0.7, P.PCSF_FP64M)
    assert np.all(np.abs(lm - l64) <= 1e-12 * np.maximum(1.0, np.abs(l64)))
    assert np.all(np.abs(em - e64) <= 1e-12 * np.maximum(1.0, np.abs(e64)))


@pytest.mark.parametrize("strategy", ["mle", "mle_grad"])
def test_maximisers_match_fp64(ctx, strategy):
    _say("FP64M %s vs FP64" % strategy)
    m = W.model_for("C2", ctx)
    rng = np.random.default_rng(5)
    lengths = rng.integers(5, 200, 12)
    codes, off = m.simulate(lengths, rng.integers(0, 2, 12), rng.uniform(0.3, 3, 12),
                            gap_p=0.1, seed=9)
    b = P.Batch(m, codes, off)
    r64 = P.score(b, strategy, P.PCSF_FP64)
    rm = P.score(b, strategy, P.PCSF_FP64M)
    assert (rm["status"] == 0).all()
    assert np.abs(rm["score"] - r64["score"]).max() <= 1e-5
    assert np.all(np.abs(rm["rho"] - r64["rho"]) <= 1e-5 * r64["rho"])


def test_omega_matches_fp64(ctx, params_dir):
    _say("FP64M omega test vs FP64")
    tree = P.newick_tree(open(os.path.join(params_dir, "12flies.nh")).read())
    m = P.PhyloCSFModel.from_paramset(None, "12flies")
    codes, off = m.simulate(np.array([40, 90, 7]), np.array([0, 1, 0]), np.ones(3), gap_p=0.1,
                            seed=2)
    b = P.Batch(P.PhyloCSFModel.tree_only(ctx, tree), codes, off)
    r64 = P.score_omega(b, precision=P.PCSF_FP64)
    rm = P.score_omega(b, precision=P.PCSF_FP64M)
    assert (rm["status"] == 0).all()
    assert np.abs(rm["score"] - r64["score"]).max() <= 1e-5
    assert np.all(np.abs(rm["lpr"] - r64["lpr"]) <= 1e-9 * np.abs(r64["lpr"]) + 1e-4)


def _random_tree_text(n_leaves, rng):
    """A random rooted binary tree in Newick (leaves s0, s1, ...)."""
    nodes = ["s%d:%.4f" % (i, rng.uniform(0.01, 0.3)) for i in range(n_leaves)]
    while len(nodes) > 1:
        i, j = sorted(rng.choice(len(nodes), 2, replace=False))
        b = nodes.pop(j)
        a = nodes.pop(i)
        nodes.append("(%s,%s):%.4f" % (a, b, rng.uniform(0.01, 0.3)))
    return nodes[0].rsplit(":", 1)[0] + ";"


def test_tree_beyond_lds_workspace(ctx, params_dir):
    _say("FP64M on a 200-leaf tree vs oracle")
    rng = np.random.default_rng(200)
    text = _random_tree_text(200, rng)
    prefix = os.path.join(params_dir, "58mammals")
    Sc, pc = O.ecm_read(prefix + "_coding.ECM")
    Sn, pn = O.ecm_read(prefix + "_noncoding.ECM")
    tree = P.newick_tree(text)
    assert tree.nl == 200
    m = P.PhyloCSFModel(ctx, tree, [(Sc, pc), (Sn, pn)])
    lengths = np.array([1, 40, 70])
    codes, off = m.simulate(lengths, np.array([0, 1, 0]), np.ones(3), gap_p=0.3, n_p=0.02,
                            seed=6)
    b = P.Batch(m, codes, off)
    with pytest.raises(P.InvalidArgument):
        P.score_fixed(b, P.PCSF_FP64)
    res = P.score_fixed(b, P.PCSF_FP64M)
    om = O.Model(O.newick_tree(text), [Sc, Sn], [pc, pn])
    for r in range(len(lengths)):
        ref = om.score_fixed(codes[off[r]:off[r + 1]])
        assert abs(res["score"][r] - ref["score"]) <= 1e-6, r
        for key in ("lpr_c", "lpr_n"):
            assert abs(res[key][r] - ref[key]) <= 1e-10 * max(1.0, abs(ref[key])), (r, key)
        assert abs(res["anc_comp"][r] - ref["anc_comp"]) <= 1e-5, r


def test_noalloc_refuses_fp64m(ctx):
    """Refused before any argument is read (the per-node tiles are allocated)."""
    from phylocsf_amd import _lib
    m = W.model_for("C1", ctx)
    rho = np.ones(2)
    rc = _lib.lib().pcsf_forward_noalloc(m.h, None, 0, None, 0,
                                         rho.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)),
                                         P.PCSF_FP64M, None, None)
    assert rc == _lib.PCSF_EINVAL
    assert b"PCSF_FP64M" in _lib.lib().pcsf_last_error(ctx.h)
