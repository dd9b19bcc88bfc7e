"""GPU parity of the adjoint (outside-pass) kernel and the gradient mle.

* d lpr / d rho from pcsf_backward against central finite differences of
  the FP64 forward (step 1e-5 in rho) and against the oracle's gradient
  through branch posteriors (PhyloLik.add_branch_posteriors + PhyloEM's
  d ELL/dt, lib/CamlPaml/PhyloEM.ml:140-182), per branch and per region;
* the torch.autograd.Function contract (torch.autograd.gradcheck);
* the gradient maximiser: at its estimate the derivative vanishes (or the
  estimate is a bound of [0.01, 10] with the derivative pointing out), its
  lpr is at least the reference maximiser's, and its rho agrees with the
  reference maximiser's within that maximiser's 1% stopping tolerance.
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


def _oracle_model(params_dir, name):
    prefix = os.path.join(params_dir, name)
    t = O.newick_tree(open(prefix + ".nh").read())
    Sc, pc = O.ecm_read(prefix + "_coding.ECM")
    Sn, pn = O.ecm_read(prefix + "_noncoding.ECM")
    return O.Model(t, [Sc, Sn], [pc, pn])


@pytest.fixture(scope="module")
def mammals(ctx):
    m = P.PhyloCSFModel.from_paramset(ctx, "29mammals")
    rng = np.random.default_rng(31)
    lengths = np.array([40, 1, 70, 0, 23, 66])
    codes, off = m.simulate(lengths, rng.integers(0, 2, len(lengths)),
                            rng.uniform(0.3, 3, len(lengths)), gap_p=0.15, n_p=0.01, seed=12)
    rho = rng.uniform(0.2, 4.0, (len(lengths), 2))
    return m, codes, off, rho


def test_backward_matches_finite_differences(ctx, mammals):
    _say("adjoint vs finite differences")
    m, codes, off, rho = mammals
    b = P.Batch(m, codes, off)
    ll, _ = b.forward(rho, P.PCSF_FP64, per_region=True)
    grho, _ = b.backward(np.ones_like(rho))
    h = 1e-5
    lp, _ = b.forward(rho + h, P.PCSF_FP64, per_region=True)
    lm, _ = b.forward(rho - h, P.PCSF_FP64, per_region=True)
    fd = (lp - lm) / (2 * h)
    assert np.all(np.abs(grho - fd) <= 1e-6 * np.maximum(1.0, np.abs(fd))), (grho, fd)
    # weights scale the gradient linearly
    b.forward(rho, P.PCSF_FP64, per_region=True)
    w = np.arange(1.0, 1.0 + rho.size).reshape(rho.shape)
    g2, _ = b.backward(w)
    assert np.allclose(g2, w * grho, rtol=1e-12, atol=1e-12)


def test_backward_matches_oracle_branch_gradients(ctx, mammals, params_dir):
    _say("adjoint vs oracle branch posteriors")
    m, codes, off, rho = mammals
    om = _oracle_model(params_dir, "29mammals")
    b = P.Batch(m, codes, off)
    b.forward(rho, P.PCSF_FP64, per_region=True)
    grho, gbr = b.backward(np.ones_like(rho), branches=True)
    for r in range(len(off) - 1):
        for q in range(2):
            _, d, g = om.lpr_grad(q, rho[r, q], codes[off[r]:off[r + 1]])
            assert abs(grho[r, q] - d) <= 1e-9 * max(1.0, abs(d)), (r, q)
            assert np.allclose(gbr[r, q], g, rtol=1e-8, atol=1e-9), (r, q)


@pytest.mark.parametrize("precision", [P.PCSF_FP32, P.PCSF_FP16S])
def test_tangent_pass_matches_fp64_adjoint(ctx, mammals, precision):
    """Backward of an MFMA-precision forward: the FP32 tangent pass of the
    fused kernel against the FP64 outside pass (and the FP64 central
    differences it is checked against above), empty and one-column regions,
    gaps and N included."""
    _say(f"tangent pass, precision {precision}")
    m, codes, off, rho = mammals
    b = P.Batch(m, codes, off)
    b.forward(rho, P.PCSF_FP64, per_region=True)
    g64, _ = b.backward(np.ones_like(rho))
    b.forward(rho, precision, per_region=True)
    g32, _ = b.backward(np.ones_like(rho))
    assert np.all(np.abs(g32 - g64) <= 2e-4 * np.maximum(1.0, np.abs(g64))), (g32, g64)
    assert np.all(g32[3] == 0.0)  # the empty region
    # linear in the incoming gradient, as the FP64 pass
    w = np.arange(1.0, 1.0 + rho.size).reshape(rho.shape)
    g2, _ = b.backward(w)
    assert np.allclose(g2, w * g32, rtol=1e-12, atol=1e-12)
    # per-branch gradients always come from the FP64 outside pass
    gb, gbr = b.backward(np.ones_like(rho), branches=True)
    assert np.allclose(gb, g64, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("precision", [P.PCSF_FP32, P.PCSF_FP16S])
def test_tangent_pass_58mammals(ctx, precision):
    """The FP32 tangent pass after an FP32 or split-FP16 forward on the
    58-leaf tree at several tree scales, against the FP64 outside pass."""
    _say(f"tangent pass, 58mammals, precision {precision}")
    m = P.PhyloCSFModel.from_paramset(ctx, "58mammals")
    rng = np.random.default_rng(77)
    R = 40
    lengths = rng.integers(1, 300, R)
    codes, off = m.simulate(lengths, rng.integers(0, 2, R), rng.uniform(0.3, 3.0, R),
                            gap_p=0.2, n_p=0.01, seed=8)
    rho = np.exp(rng.uniform(np.log(0.01), np.log(10.0), (R, 2)))
    b = P.Batch(m, codes, off)
    b.forward(rho, P.PCSF_FP64, per_region=True)
    g64, _ = b.backward(np.ones_like(rho))
    b.forward(rho, precision, per_region=True)
    g32, _ = b.backward(np.ones_like(rho))
    assert np.all(np.abs(g32 - g64) <= 2e-4 * np.maximum(1.0, np.abs(g64))), \
        np.max(np.abs(g32 - g64) / np.maximum(1.0, np.abs(g64)))
    # the tree-specialised tangent kernel against the interpreted one
    assert m.specialized()["plain"]
    m.specialized(False)
    try:
        b.forward(rho, precision, per_region=True)
        gi, _ = b.backward(np.ones_like(rho))
    finally:
        m.specialized(True)
    assert np.allclose(gi, g32, rtol=1e-9, atol=1e-9), np.max(np.abs(gi - g32))


def test_autograd_function(ctx, mammals):
    _say("torch.autograd.Function")
    import torch

    from phylocsf_amd.autograd import lpr
    m, codes, off, rho = mammals
    b = P.Batch(m, codes[:off[3]], off[:4])
    r = torch.tensor(rho[:3], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: lpr(x, b), (r,), eps=1e-6, atol=1e-5, rtol=1e-6)
    out = lpr(r, b).sum()
    out.backward()
    b.forward(rho[:3], P.PCSF_FP64, per_region=True)
    g, _ = b.backward(np.ones((3, 2)))
    assert np.allclose(r.grad.numpy(), g, rtol=1e-12)


def test_mle_grad_optimality(ctx, params_dir):
    _say("gradient mle")
    m = P.PhyloCSFModel.from_paramset(ctx, "12flies")
    rng = np.random.default_rng(5)
    R = 12
    lengths = rng.integers(20, 90, R)
    codes, off = m.simulate(lengths, rng.integers(0, 2, R), rng.uniform(0.3, 3.0, R),
                            gap_p=0.05, seed=3)
    b = P.Batch(m, codes, off)
    gg = P.score(b, "mle_grad", P.PCSF_FP64)
    gb = P.score(b, "mle", P.PCSF_FP64)
    assert (gg["status"] == 0).all()
    # derivative at the estimate
    b.forward(gg["rho"], P.PCSF_FP64, per_region=True)
    d, _ = b.backward(np.ones((R, 2)))
    for r in range(R):
        for q in range(2):
            x = gg["rho"][r, q]
            if 0.0100001 < x < 9.99999:
                assert abs(d[r, q] * x) <= 1e-5 * max(1.0, abs(gg["lpr"][r, q])), (r, q)
            else:
                assert (x >= 9.99999 and d[r, q] >= 0) or (x <= 0.0100001 and d[r, q] <= 0)
            assert gg["lpr"][r, q] >= gb["lpr"][r, q] - 1e-9 * abs(gb["lpr"][r, q])
            assert abs(gg["rho"][r, q] - gb["rho"][r, q]) <= 0.01 * gb["rho"][r, q] + 1e-9
    g32 = P.score(b, "mle_grad", P.PCSF_FP32)
    assert np.all(np.abs(g32["score"] - gg["score"]) <= 1e-3)


def test_device_results_survive_backward(ctx, mammals):
    """pcsf_batch_device_results keeps the last forward's values through a
    backward (the gradient passes reduce into buffers of their own)."""
    import ctypes as C
    from phylocsf_amd import _lib
    _say("device results after backward")
    m, codes, off, rho = mammals
    b = P.Batch(m, codes, off)
    from hip_runtime import runtime
    hip = runtime()
    for prec in (P.PCSF_FP64, P.PCSF_FP32):
        ll, el = b.forward(rho, prec, per_region=True)
        b.backward(np.ones_like(rho))
        dl, de = C.c_void_p(), C.c_void_p()
        assert _lib.lib().pcsf_batch_device_results(b.h, C.byref(dl), C.byref(de)) == 0
        got_l, got_e = np.zeros_like(ll), np.zeros_like(el)
        assert hip.hipMemcpy(got_l.ctypes.data, dl, got_l.nbytes, 2) == 0
        assert hip.hipMemcpy(got_e.ctypes.data, de, got_e.nbytes, 2) == 0
        assert np.array_equal(got_l, ll) and np.array_equal(got_e, el), prec
