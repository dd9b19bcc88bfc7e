"""Error paths of the C ABI on the GPU.

* a leaf code outside 0..63 and 255 is refused by pcsf_batch_create itself
  (checked on the device while the codes are transposed);
* a failed device allocation leaves nothing half-built: the call returns
  PCSF_ENOMEM and the next call, with memory available, allocates afresh
  and gives the same results as an undisturbed batch;
* a non-symmetric exchangeability matrix (a non-reversible model) is refused;
* Q.Diag.real_to_Pt's clamp and failure (lib/CamlPaml/Q.ml:226-247) on
  crafted models whose Q has negative off-diagonal entries: entries of
  exp(Qt) in (-1e-6, 0) become 0 and the P(t) equals the oracle's; an entry
  below -1e-6 fails with the reference's message for the first failing
  branch, row and column in the reference's order, whatever the launch order.
"""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu

import phylocsf_amd as P  # noqa: E402
from phylocsf_amd import _lib  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    return P.Context(0)


def _ecms(params_dir, name):
    prefix = os.path.join(params_dir, name)
    return [P.import_parameters(prefix + "_coding.ECM"),
            P.import_parameters(prefix + "_noncoding.ECM")]


def _tree(params_dir, name):
    return P.newick_tree(open(os.path.join(params_dir, name + ".nh")).read())


def test_batch_create_rejects_bad_codes_on_device(ctx):
    _say("bad codes through the C ABI")
    m = P.PhyloCSFModel.from_paramset(ctx, "12flies")
    nl = m.tree.nl
    off = np.array([0, 130], np.int64)
    for bad in (64, 100, 254):
        codes = np.zeros((130, nl), np.uint8)
        codes[129, nl - 1] = bad
        h = C.c_void_p()
        rc = _lib.lib().pcsf_batch_create(m.h, P.api._u8(codes), 130, P.api._i64(off), 1,
                                          C.byref(h))
        assert rc == _lib.PCSF_EINVAL and not h.value, bad
        assert b"neither a codon index" in _lib.lib().pcsf_last_error(ctx.h)
    codes = np.full((130, nl), 255, np.uint8)
    codes[:, 0] = 63
    b = P.Batch(m, codes, off)  # 63 and 255 are accepted
    got, _ = b.codes()
    assert np.array_equal(got, codes)


def test_failed_allocation_recovers(ctx):
    m = P.PhyloCSFModel.from_paramset(ctx, "58mammals")
    rng = np.random.default_rng(4)
    lengths = np.array([40, 70, 3])
    codes, off = m.simulate(lengths, [0, 1, 0], [1.0, 0.7, 2.0], gap_p=0.1, seed=5)
    rho = rng.uniform(0.3, 3.0, (3, 2))
    ref = P.Batch(m, codes, off).forward(rho, P.PCSF_FP32, per_region=True)
    nl = m.tree.nl
    # the region-aligned code copy is nl * 256 bytes; the transition sets are MBs
    for limit, what in ((nl * 256 - 1, "aligned codes"), (1 << 20, "transition sets")):
        _say("allocation limit %d (%s)" % (limit, what))
        b = P.Batch(m, codes, off)
        os.environ["PCSF_DEVICE_ALLOC_LIMIT"] = str(limit)
        try:
            with pytest.raises(P.PcsfError) as ei:
                b.forward(rho, P.PCSF_FP32, per_region=True)
            assert ei.value.code == _lib.PCSF_ENOMEM
            with pytest.raises(P.PcsfError):  # and again: still refused, no kernel ran
                b.forward(rho, P.PCSF_FP32, per_region=True)
        finally:
            del os.environ["PCSF_DEVICE_ALLOC_LIMIT"]
        got = b.forward(rho, P.PCSF_FP32, per_region=True)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_nonsymmetric_exchangeabilities_refused(ctx, params_dir):
    _say("non-reversible model")
    (Sc, pc), (Sn, pn) = _ecms(params_dir, "12flies")
    S = Sc.copy()
    S[5, 9] *= 1.5
    with pytest.raises(P.InvalidArgument) as ei:
        P.PhyloCSFModel(ctx, _tree(params_dir, "12flies"), [(S, pc), (Sn, pn)])
    assert "not reversible" in str(ei.value)


def _crafted(S, pairs, eps):
    S = S.copy()
    for (a, b), f in pairs:
        S[a, b] = S[b, a] = -eps * f
    return S


def _expm_min(S, pi, branches):
    Q = O.build_q(S, pi)
    return min(scipy.linalg.expm(Q * t).min() for t in branches)


# two single-nucleotide neighbours each for codons 10 and 33 (index 16 n1 + 4 n2 + n3):
# rows 10 and 33 both go negative; row 10 comes first in the reference's order
# although the pair through row 33 is made the more negative
PAIRS = [((10, 14), 1.0), ((33, 37), 3.0)]


def test_pt_clamp_matches_oracle(ctx, params_dir):
    _say("P(t) clamp")
    (Sc, pc), (Sn, pn) = _ecms(params_dir, "12flies")
    tree = _tree(params_dir, "12flies")
    br = tree.branch[:-1]
    # the largest eps whose most negative exp(Qt) entry over the branches stays above -1e-6
    lo, hi = 0.0, 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if _expm_min(_crafted(Sc, PAIRS, mid), pc, br) > -0.5e-6:
            lo = mid
        else:
            hi = mid
    S = _crafted(Sc, PAIRS, lo)
    assert -1e-6 < _expm_min(S, pc, br) < -1e-8  # the clamp branch is reached
    m = P.PhyloCSFModel(ctx, tree, [(S, pc), (Sn, pn)])
    Pg = m.transition(0, 1.0)
    Q = O.build_q(S, pc)
    clamped = 0
    for i, t in enumerate(br):
        Po = O.pt(Q, pc, t)
        raw = scipy.linalg.expm(Q * t)
        assert (Pg[i] >= 0).all()
        assert np.abs(Pg[i] - Po).max() <= 1e-12, i
        neg = (raw < 0) & (np.abs(raw) > 1e-12)
        assert (Pg[i][neg] == 0).all()
        clamped += int(neg.sum())
    assert clamped > 0


def test_pt_failure_message_is_the_references_first(ctx, params_dir):
    _say("P(t) failure")
    (Sc, pc), (Sn, pn) = _ecms(params_dir, "12flies")
    tree = _tree(params_dir, "12flies")
    br = tree.branch[:-1]
    S = _crafted(Sc, PAIRS, 0.5)
    Q = O.build_q(S, pc)
    first = None
    for t in br:  # PhyloModel.make builds branch 0, 1, ... and stops at the first Failure
        try:
            O.pt(Q, pc, t)
        except O.OracleError as e:
            first = str(e)
            break
    assert first is not None and "[10," in first
    for rho in (1.0, 2.5):  # fixed-scale and per-region builds report the same entry
        m = P.PhyloCSFModel(ctx, tree, [(S, pc), (Sn, pn)])
        with pytest.raises(P.Failure) as ei:
            m.transition(0, rho)
        msg = str(ei.value)
        t0 = next(t for t in br * rho if (scipy.linalg.expm(Q * t) < -1e-6).any())
        assert ("expm(%.2e*Q)[10," % t0) in msg, msg
        if rho == 1.0:
            assert msg.split(" (status")[0].split(" = ")[0] == first.split(" = ")[0], (msg, first)
    m = P.PhyloCSFModel(ctx, tree, [(S, pc), (Sn, pn)])
    b = P.Batch(m, np.zeros((10, tree.nl), np.uint8), np.array([0, 5, 10]))
    with pytest.raises(P.Failure) as ei:
        b.forward(np.array([[1.0, 1.0], [3.0, 1.0]]), P.PCSF_FP32, per_region=True)
    assert "[10," in str(ei.value)
