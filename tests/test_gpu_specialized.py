"""GPU parity of the tree-specialised kernels (straight-line op programs
generated at build time for the shipped trees) against the op-program
interpreter and the FP64 path: the MFMA kernels at FP32 and split FP16, and
the FP64 forward on matrix cores.

Both run the same op code (kernels/prune_mfma.hpp); the specialised kernel
only replaces the dispatch.  They agree to 1e-12 relative (the compiler may
contract a multiply-add differently in the two instances), and both stay
within the FP32 contract of the FP64 path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import phylocsf_amd as P  # noqa: E402

SHIPPED = ["12flies", "29mammals", "58mammals", "100vertebrates", "120mammals"]


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    return P.Context(0)


def _sample(m, seed, R=24):
    rng = np.random.default_rng(seed)
    lengths = rng.integers(0, 140, R)
    lengths[0] = 1
    codes, off = m.simulate(lengths, rng.integers(0, 2, R), rng.uniform(0.3, 3.0, R),
                            gap_p=0.12, n_p=0.01, seed=seed)
    rho = rng.uniform(0.05, 5.0, (R, 2))
    return codes, off, rho


@pytest.mark.parametrize("name", SHIPPED)
def test_specialised_matches_interpreter(ctx, name):
    m = P.PhyloCSFModel.from_paramset(ctx, name)
    av = m.specialized()
    _say("%s: specialised programs %s" % (name, av))
    assert av == {"plain": True, "tables": True}
    codes, off, rho = _sample(m, 40 + len(name))
    b = P.Batch(m, codes, off)
    ref, _ = b.forward(1.0, P.PCSF_FP64)
    ref_r, _ = b.forward(rho, P.PCSF_FP64, per_region=True)
    for prec in (P.PCSF_FP32, P.PCSF_FP16S):
        for shape in range(6):
            m.set_variant(shape, prec)
            out = {}
            for spec in (True, False):
                m.specialized(spec)
                fixed, _ = b.forward(1.0, prec)
                per, _ = b.forward(rho, prec, per_region=True)
                out[spec] = (fixed, per)
            m.specialized(True)
            for k in range(2):
                a, i = out[True][k], out[False][k]
                assert np.allclose(a, i, rtol=1e-12, atol=1e-12), (name, prec, shape, k)
            assert np.all(np.abs(out[True][0] - ref) <= 1e-6 * np.maximum(1.0, np.abs(ref)))
            assert np.all(np.abs(out[True][1] - ref_r) <= 1e-6 * np.maximum(1.0, np.abs(ref_r)))
    # the FP64 forward on matrix cores: the generated program against the
    # interpreter and the plain FP64 kernel, to FP64 rounding
    out = {}
    for spec in (True, False):
        m.specialized(spec)
        fixed, _ = b.forward(1.0, P.PCSF_FP64M)
        per, _ = b.forward(rho, P.PCSF_FP64M, per_region=True)
        out[spec] = (fixed, per)
    m.specialized(True)
    for k, r in ((0, ref), (1, ref_r)):
        assert np.allclose(out[True][k], out[False][k], rtol=1e-12, atol=1e-12), (name, k)
        assert np.all(np.abs(out[True][k] - r) <= 1e-12 * np.maximum(1.0, np.abs(r))), (name, k)


def test_species_subtree_uses_interpreter(ctx):
    sp = "Human,Chimp,Rhesus,Mouse,Rat,Dog,Cow"
    m = P.PhyloCSFModel.from_paramset(ctx, "29mammals", species=sp.split(","))
    assert m.specialized() == {"plain": False, "tables": False}
    codes, off, _ = _sample(m, 7, R=6)
    b = P.Batch(m, codes, off)
    r64, _ = b.forward(1.0, P.PCSF_FP64)
    r32, _ = b.forward(1.0, P.PCSF_FP32)
    assert np.all(np.abs(r32 - r64) <= 1e-6 * np.maximum(1.0, np.abs(r64)))
