"""The C3, C4 and C5 scoring workloads on the HIP path.

C3  100vertebrates, lognormal(300, 0.8) region lengths, coding model at a
    per-region tree scale U(0.3, 3.0), mle:
    * 10 regions clipped to 150 codons: the reference maximiser at FP64
      against the oracle's (rho to 1e-6 relative, score to 1e-6 dB); the
      gradient maximiser at FP64 reaches a vanishing gradient with lpr at
      least the reference maximiser's and rho within its 1% tolerance;
    * the FP64 adjoint against central finite differences (step 1e-5) on
      200 full-length regions, to a tolerance derived from the differences'
      rounding noise and truncation (below);
    * all 2,000 regions: the gradient maximiser at FP32 and split FP16 within
      1e-3 dB of the FP64 run of the same maximiser, and the reference
      maximiser at split FP16 within its own stopping bound of FP64's.
C4  120mammals, lengths U{20..600}, gap runs (mean 5 codons) covering 0.25,
    six reading frames built on the GPU:
    * frame codes byte for byte against the oracle's pleaves on 30 regions;
    * 2,000 regions: FP32 within 1e-3 dB of the FP64 kernel on every frame;
    * 3 short regions: the FP64 kernel against the oracle.
C5  100vertebrates, 20,000 regions U{100..3000}, gaps 0.15:
    * FP32 and split FP16 within 1e-3 dB of the FP64 kernel on every region;
      plain FP16 and BF16 errors measured and bounded;
    * per-branch gradients against the oracle's branch posteriors on 5
      short regions.
Paths are relative to the mlin/PhyloCSF tree: src/PhyloCSFModel.ml:84-99,
src/PhyloCSF.ml:198-246, lib/CamlPaml/PhyloEM.ml:140-182.

Finite-difference tolerance.  With lpr known to relative rounding u (the
FP64 kernel's lpr agrees with the oracle's to 1e-12 relative or better,
measured here on 5 regions; u is taken as ten times that), a central
difference with step h carries rounding noise at most 2 u |lpr| / (2 h) and
truncation h^2 |lpr'''| / 6; |lpr'''| is
estimated from a five-point stencil at step 1e-2.  Near rho = 0.01 with a
step relative to rho (1e-7) the noise term alone reaches 1e-2 of the
derivative on long regions, which is what a relative-step check there
shows; the bound is checked there too.

Reference-maximiser bound: tests/mle_bounds.py.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import phylocsf_amd as P  # noqa: E402
from phylocsf_amd import workloads as W  # noqa: E402
from mle_bounds import brent_stop_bound_db  # noqa: E402
from oracle import cli_oracle as CO  # noqa: E402
from oracle import oracle as O  # noqa: E402

TOL_DB = 1e-3


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
def vert(ctx):
    return W.model_for("C3", ctx)


@pytest.fixture(scope="module")
def c3_full(vert):
    wl = W.make("C3", vert, n_threads=16)
    return wl, P.Batch(vert, wl.codes, wl.offsets)




# ---- C3 ---------------------------------------------------------------------

def test_c3_short_mle_vs_oracle(ctx, vert, params_dir):
    _say("C3: 10 regions <= 150 codons, mle and mle_grad at FP64 vs the oracle")
    lengths, rmodel, rrho, kw = W.region_params("C3", 10)
    lengths = np.minimum(lengths, 150)
    codes, off = vert.simulate(lengths, rmodel, rrho, seed=W.SPECS["C3"]["seed"], **kw)
    om = _oracle_model(params_dir, "100vertebrates")
    b = P.Batch(vert, codes, off)
    g = P.score(b, "mle", P.PCSF_FP64)
    gg = P.score(b, "mle_grad", P.PCSF_FP64)
    assert (g["status"] == 0).all() and (gg["status"] == 0).all()
    for r in range(10):
        ref = om.score_mle(codes[off[r]:off[r + 1]])
        assert abs(g["rho"][r, 0] - ref["rho_c"]) <= 1e-6 * ref["rho_c"], r
        assert abs(g["rho"][r, 1] - ref["rho_n"]) <= 1e-6 * ref["rho_n"], r
        assert abs(g["score"][r] - ref["score"]) <= 1e-6, r
        for q, key in ((0, "lpr_c"), (1, "lpr_n")):
            # the gradient maximiser's lpr is at least the reference's ...
            assert gg["lpr"][r, q] >= ref[key] - 1e-9 * abs(ref[key]), (r, q)
            # ... its rho within the reference's 1% stopping tolerance ...
            rr = ref["rho_c"] if q == 0 else ref["rho_n"]
            assert abs(gg["rho"][r, q] - rr) <= 0.0101 * rr or rr in (0.01, 10.0), (r, q)
            # ... and the oracle's gradient vanishes there (or points out of [0.01, 10])
            l, dl, _ = om.lpr_grad(q, gg["rho"][r, q], codes[off[r]:off[r + 1]])
            x = gg["rho"][r, q]
            hl = x * dl
            assert (abs(hl) <= 1e-4 * max(1.0, abs(l)) or (x <= 0.0100001 and hl < 0)
                    or (x >= 9.99999 and hl > 0)), (r, q, x, hl)


def test_c3_fp64_adjoint_vs_finite_differences(ctx, vert, c3_full, params_dir):
    wl, _ = c3_full
    n = 200
    sub = P.Batch(vert, wl.codes[:wl.offsets[n]], wl.offsets[:n + 1])
    # u: the FP64 kernel's relative lpr rounding, measured against the oracle
    om = _oracle_model(params_dir, "100vertebrates")
    rho0 = np.repeat(wl.region_rho[:n, None], 2, axis=1)
    ll, _ = sub.forward(rho0, P.PCSF_FP64, per_region=True)
    u = 0.0
    for r in range(5):
        for q in range(2):
            ref, _ = om.lpr(q, rho0[r, q], wl.codes[wl.offsets[r]:wl.offsets[r + 1]])
            u = max(u, abs(ll[r, q] - ref) / abs(ref))
    _say("C3: FP64 kernel vs oracle, max relative lpr difference %.2e" % u)
    assert u <= 1e-12
    u = max(10 * u, 1e-14)
    for what, rho, h in (("true scales, step 1e-5", rho0, 1e-5 * np.ones_like(rho0)),
                         ("mle estimates, step 1e-5 rho", None, None)):
        if rho is None:
            rho = P.score(sub, "mle_grad", P.PCSF_FP64)["rho"]
            h = 1e-5 * rho
        ll, _ = sub.forward(rho, P.PCSF_FP64, per_region=True)
        d, _ = sub.backward(np.ones_like(rho))
        f = {}
        for k in (-1, 1):
            f[k], _ = sub.forward(rho + k * h, P.PCSF_FP64, per_region=True)
        fd = (f[1] - f[-1]) / (2 * h)
        H = 1e-2 * np.minimum(rho, 1.0) * 0.5
        g = {}
        for k in (-2, -1, 1, 2):
            g[k], _ = sub.forward(rho + k * H, P.PCSF_FP64, per_region=True)
        d3 = np.abs(g[2] - 2 * g[1] + 2 * g[-1] - g[-2]) / (2 * H ** 3)
        bound = 2 * u * np.abs(ll) / (2 * h) + 2 * h ** 2 * d3 / 6 + 1e-9 * np.maximum(1, np.abs(d))
        err = np.abs(d - fd)
        _say("C3 FD (%s): max |adjoint - FD| / max(1,|FD|) %.3g, max err/bound %.3g"
             % (what, float(np.max(err / np.maximum(1, np.abs(fd)))), float(np.max(err / bound))))
        assert np.all(err <= bound), what


def test_c3_full_gradient_maximiser_precisions(ctx, c3_full):
    wl, b = c3_full
    _say("C3: all %d regions, mle_grad at FP64 / FP32 / split FP16" % wl.n_regions)
    g64 = P.score(b, "mle_grad", P.PCSF_FP64)
    assert (g64["status"] == 0).all()
    lengths = np.diff(wl.offsets)
    worst = {}
    for prec, name in ((P.PCSF_FP32, "FP32"), (P.PCSF_FP16S, "split FP16")):
        g = P.score(b, "mle_grad", prec)
        assert (g["status"] == 0).all()
        dd = np.abs(g["score"] - g64["score"])
        # the forward's own rounding at the FP64 estimates
        ll, _ = b.forward(g64["rho"], prec, per_region=True)
        fw = np.abs(P.DB * ((ll[:, 0] - ll[:, 1]) - (g64["lpr"][:, 0] - g64["lpr"][:, 1])))
        rel = np.max(np.abs(ll - g64["lpr"]) / np.abs(g64["lpr"]))
        i = int(np.argmax(dd))
        _say("  %s mle_grad vs FP64 mle_grad: max |d score| %.3g dB (mean %.3g; region %d, "
             "%d codons), max |d ln rho| %.3g; forward at the FP64 estimates: max %.3g dB, "
             "max rel lpr %.3g" % (name, dd.max(), dd.mean(), i, lengths[i],
                                   np.max(np.abs(np.log(g["rho"] / g64["rho"]))), fw.max(), rel))
        worst[name] = dd.max()
    assert max(worst.values()) <= TOL_DB, worst


def test_c3_full_reference_maximiser_within_stopping_bound(ctx, c3_full):
    wl, b = c3_full
    _say("C3: all regions, mle (GSL Brent) at split FP16 vs FP64")
    m64 = P.score(b, "mle", P.PCSF_FP64)
    m16 = P.score(b, "mle", P.PCSF_FP16S)
    assert (m64["status"] == 0).all() and (m16["status"] == 0).all()
    bound = brent_stop_bound_db(b, m64["rho"], m16["rho"]) + TOL_DB
    dd = np.abs(m16["score"] - m64["score"])
    _say("  max |d score| %.3g dB; bound median %.3g dB, max %.3g dB; max d/bound %.3g"
         % (dd.max(), np.median(bound), bound.max(), np.max(dd / bound)))
    assert np.all(dd <= bound)


# ---- C4 ---------------------------------------------------------------------

@pytest.fixture(scope="module")
def mammals120(ctx):
    return W.model_for("C4", ctx)


def test_c4_frame_codes_byte_equal(ctx, mammals120):
    _say("C4: 30 regions, 6-frame codes vs oracle pleaves")
    m = mammals120
    lengths, rmodel, rrho, kw = W.region_params("C4", 30)
    nuc, noff = m.simulate_nucleotides(lengths, rmodel, rrho, seed=W.SPECS["C4"]["seed"],
                                       n_p=0.01, **kw)
    # trim some regions to lengths that are not multiples of 3, lower-case one
    pieces = []
    for r in range(30):
        lo, hi = noff[r], noff[r + 1] - (r % 3)
        pieces.append(nuc[:, lo:hi])
    low = pieces[7].copy()
    low[(low >= ord("A")) & (low <= ord("Z"))] += 32
    pieces[7] = low
    off = np.zeros(31, np.int64)
    np.cumsum([p.shape[1] for p in pieces], out=off[1:])
    packed = np.ascontiguousarray(np.concatenate(pieces, axis=1))
    bg = P.Batch.frames(m, packed, off, 6)
    codes, coff = bg.codes()
    ident = list(range(m.tree.nl))
    expect = []
    for p in pieces:
        rows = ["".join(map(chr, p[l])) for l in range(m.tree.nl)]
        rc_rows = [CO.revcomp(s) for s in rows]
        L = p.shape[1]
        for f in range(6):
            expect.append(CO.pleaves(rows if f < 3 else rc_rows, ident, f % 3, L - 1))
    eoff = np.zeros(len(expect) + 1, np.int64)
    np.cumsum([e.shape[0] for e in expect], out=eoff[1:])
    assert np.array_equal(coff, eoff)
    assert np.array_equal(codes, np.concatenate(expect, axis=0))
    assert (codes == P.PCSF_MARGINALIZE).mean() > 0.2  # the gap runs reached the codes


def test_c4_fp32_vs_fp64_every_frame(ctx, mammals120, params_dir):
    m = mammals120
    lengths, rmodel, rrho, kw = W.region_params("C4", 2000)
    nuc, noff = m.simulate_nucleotides(lengths, rmodel, rrho, seed=W.SPECS["C4"]["seed"],
                                       n_threads=16, **kw)
    b = P.Batch.frames(m, nuc, noff, 6)
    _say("C4: 2,000 regions x 6 frames (%d codon columns), FP32 vs FP64" % b.n_cols)
    s64 = P.score_fixed(b, P.PCSF_FP64)
    s32 = P.score_fixed(b, P.PCSF_FP32)
    dd = np.abs(s32["score"] - s64["score"])
    _say("  max |d score| %.3g dB (mean %.3g)" % (dd.max(), dd.mean()))
    assert dd.max() <= TOL_DB
    # the FP64 kernel against the oracle on 3 short regions (all six frames)
    om = _oracle_model(params_dir, "120mammals")
    codes, coff = b.codes()
    for k in range(18):
        ref = om.score_fixed(codes[coff[k]:coff[k + 1]])
        assert abs(s64["score"][k] - ref["score"]) <= 1e-9 + 1e-11 * abs(ref["score"]), k
        assert abs(s64["lpr_c"][k] - ref["lpr_c"]) <= 1e-10 * max(1, abs(ref["lpr_c"])), k


# ---- C5 ---------------------------------------------------------------------

def test_c5_full_precisions(ctx, vert):
    wl = W.make("C5", vert, n_threads=16)
    b = P.Batch(vert, wl.codes, wl.offsets)
    _say("C5: %d regions, %d columns, every precision vs the FP64 kernel"
         % (wl.n_regions, wl.n_cols))
    s64 = P.score_fixed(b, P.PCSF_FP64)
    stated = {P.PCSF_FP16: 1.0, P.PCSF_BF16: 8.0}
    for prec, name in ((P.PCSF_FP32, "FP32"), (P.PCSF_FP16S, "split FP16"),
                       (P.PCSF_FP16, "FP16"), (P.PCSF_BF16, "BF16")):
        s = P.score_fixed(b, prec)
        dd = np.abs(s["score"] - s64["score"])
        rel = max(float(np.max(np.abs(s[k] - s64[k]) / np.abs(s64[k]))) for k in ("lpr_c", "lpr_n"))
        _say("  %-10s max |d score| %.3g dB (mean %.3g), max rel lpr %.3g"
             % (name, dd.max(), dd.mean(), rel))
        if prec in (P.PCSF_FP32, P.PCSF_FP16S):
            assert dd.max() <= TOL_DB, name
        else:
            # rounding P to 11 (FP16) or 8 (BF16) bits changes the model itself:
            # measured 0.56 dB and 4.7 dB at most on this workload
            assert dd.max() <= stated[prec], name


def test_c5_branch_gradients_vs_oracle(ctx, vert, params_dir):
    _say("C5: per-branch gradients, 5 short regions, vs oracle")
    lengths, rmodel, rrho, kw = W.region_params("C5", 5)
    lengths = np.minimum(lengths, 40)
    codes, off = vert.simulate(lengths, rmodel, rrho, seed=W.SPECS["C5"]["seed"], **kw)
    b = P.Batch(vert, codes, off)
    g = P.score(b, "mle_grad", P.PCSF_FP64)
    om = _oracle_model(params_dir, "100vertebrates")
    b.forward(g["rho"], P.PCSF_FP64, per_region=True)
    grho, gbr = b.backward(np.ones_like(g["rho"]), branches=True)
    for r in range(5):
        for q in range(2):
            _, d, gb = om.lpr_grad(q, g["rho"][r, q], codes[off[r]:off[r + 1]])
            assert abs(grho[r, q] - d) <= 1e-8 * max(1.0, abs(d)), (r, q)
            assert np.all(np.abs(gbr[r, q] - gb) <= 1e-8 * np.maximum(1.0, np.abs(gb))), (r, q)
