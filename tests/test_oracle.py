"""Pins the FP64 oracle before anything is compared against it.

Sources of truth, in order:
  * the reference's own known-answer tests (lib/CamlPaml/test.ml:21-99);
  * the reference's published CLI results (src/test.ml:27-59);
  * closed forms and invariants of the mathematics (JC69 P(t), stationarity,
    reversibility, Chapman-Kolmogorov, brute-force enumeration, finite
    differences), chosen so a dropped term, wrong index or transposed
    operand fails at least one of them.
"""
import itertools
import os
import re

import numpy as np
import pytest
import scipy.linalg

from oracle import cli_oracle as CO
from oracle import oracle as O

REF_TEST_ML = os.path.join(os.path.dirname(__file__), "golden", "tinytest_sequences.txt")


def _params(params_dir, name):
    return os.path.join(params_dir, name)


# ---------------------------------------------------------------- helpers

def test_md5_rfc1321_vectors():
    assert O.md5(b"").hex() == "d41d8cd98f00b204e9800998ecf8427e"
    assert O.md5(b"abc").hex() == "900150983cd24fb0d6963f7d28e17f72"
    assert O.md5(b"message digest").hex() == "f96b697d7cb7938d525a2f31aaf161d0"
    assert O.md5(b"a" * 1000).hex() == __import__("hashlib").md5(b"a" * 1000).hexdigest()


def test_md5_matches_hashlib_at_every_padding_length():
    import hashlib
    rng = np.random.default_rng(1)
    for n in list(range(0, 130)) + [1000, 4097]:
        msg = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert O.md5(msg) == hashlib.md5(msg).digest(), n


def test_ocaml_random_range_and_determinism():
    a = O.ocaml_random_floats(0, 6.9, 1000)
    b = O.ocaml_random_floats(0, 6.9, 1000)
    assert np.array_equal(a, b)
    assert (a >= 0).all() and (a < 6.9).all()
    assert not np.array_equal(a, O.ocaml_random_floats(1, 6.9, 1000))




def test_ocaml_random_recurrence():
    """The stream obeys the generator's own recurrence, read from its outputs:
    the 30-bit output n is (out[n-31] + (out[n-55] xor ((out[n-55] lsr 25)
    land 31))) land (2^30 - 1) for n >= 55 (55 words of state, the new word
    built from the ones written 31 and 55 outputs earlier).  Random.float
    takes two outputs per draw, so output 2k is draw k's low part r1, whose
    lags 31 and 55 are high parts r2 (known exactly): r1 is predicted and
    the draw rebuilt from it in the same double arithmetic.  A wrong lag,
    shift, mask or order of r1 and r2 in the float fails."""
    n = 20000
    f = O.ocaml_random_floats(0, 1.0, n)
    x = f * 2.0 ** 30                 # r1 / 2^30 + r2, as rounded to a double
    r2 = np.floor(x).astype(np.int64)  # output 2k + 1, exact
    mask = (1 << 30) - 1
    for k in range(28, n):
        a = int(r2[k - 16])  # output 2k - 31
        b = int(r2[k - 28])  # output 2k - 55
        r1 = (a + (b ^ ((b >> 25) & 31))) & mask
        assert (float(r1) / 2.0 ** 30 + float(r2[k])) / 2.0 ** 30 == f[k], k


def test_ocaml_random_seeding_against_hashlib():
    """Random.init seed = full_init [| seed |]: state word i starts as i, then
    for i = 0 .. 109 the accumulator becomes MD5(accu ^ string_of_int seed)
    (from accu = "x") and word i mod 55 is xor-ed with the digest's first four
    bytes, little-endian, masked to 30 bits.  Restated here on hashlib's MD5
    and the first outputs compared with the oracle's (C) restatement."""
    import hashlib
    for seed in (0, 1, 42, -7):
        st = list(range(55))
        accu = b"x"
        for i in range(110):
            accu = hashlib.md5(accu + str(seed).encode()).digest()
            ex = accu[0] | (accu[1] << 8) | (accu[2] << 16) | (accu[3] << 24)
            st[i % 55] = (st[i % 55] ^ ex) & 0x3FFFFFFF
        idx = 0
        outs = []
        for _ in range(40):
            idx = (idx + 1) % 55
            cur = st[idx]
            new = (st[(idx + 24) % 55] + (cur ^ ((cur >> 25) & 0x1F))) & 0x3FFFFFFF
            st[idx] = new
            outs.append(new)
        f = O.ocaml_random_floats(seed, 1.0, 20)
        exp = [(outs[2 * k] / 2.0 ** 30 + outs[2 * k + 1]) / 2.0 ** 30 for k in range(20)]
        assert np.array_equal(f, np.array(exp)), seed


def test_ocaml_random_uniformity():
    """Random.float over 1e5 draws: chi-squared over 100 bins, Kolmogorov-
    Smirnov, and no lag-one correlation."""
    from scipy import stats
    a = O.ocaml_random_floats(0, 1.0, 100000)
    assert stats.chisquare(np.bincount((a * 100).astype(int), minlength=100)).pvalue > 1e-4
    assert stats.kstest(a, "uniform").pvalue > 1e-4
    r = np.corrcoef(a[1:], a[:-1])[0, 1]
    assert abs(r) < 4 / np.sqrt(len(a))


def test_find_init_retries_until_inside_beats_both_ends():
    """Fit.find_init (lib/CamlPaml/Fit.ml:27-48) on a function whose ends beat
    f(init): it retries at exp(ln lo + Random.float (ln hi - ln lo)) until a
    point beats both ends; that point lies in [lo, hi], beats both ends, and
    is the first retry draw that does (the draws come from Random.init 0)."""
    import ctypes as C
    L = O.lib()
    FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
    L.o_find_init.argtypes = [FN, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                              C.c_double, C.POINTER(C.c_int)]
    L.o_find_init.restype = C.c_double
    lo, hi = 0.01, 10.0

    def bump(x):  # maximum at 0.05; f(0.01) and f(10) both above f(1)
        return -(np.log(x) - np.log(0.05)) ** 2 + (3.0 if x in (lo, hi) else 0.0) * 0

    seen = []
    cb = FN(lambda x, ctx: (seen.append(x), bump(x))[1])
    ev = C.c_int(0)
    x = L.o_find_init(cb, None, 250, 1, 1.0, lo, hi, C.byref(ev))
    assert bump(1.0) <= bump(lo)  # init loses to an end: the retry loop runs
    assert lo <= x <= hi and bump(x) > bump(lo) and bump(x) > bump(hi)
    assert ev.value > 3 and ev.value == len(seen)
    # the retries are the Random.init 0 stream mapped into [ln lo, ln hi]
    draws = O.ocaml_random_floats(0, np.log(hi) - np.log(lo), ev.value - 3)
    assert np.allclose(seen[3:], np.exp(np.log(lo) + draws), rtol=1e-15, atol=0)
    first_ok = next(i for i, d in enumerate(np.exp(np.log(lo) + draws))
                    if bump(d) > bump(lo) and bump(d) > bump(hi))
    assert first_ok == ev.value - 4
    # maxtries reached: the better end is taken
    seen.clear()
    x = L.o_find_init(FN(lambda x, ctx: 1.0 if x == hi else 0.0), None, 250, 1, 1.0, lo, hi,
                      C.byref(ev))
    assert x == hi and ev.value == 253


def test_brent_finds_parabola_max():
    f = lambda x, ctx: -(x - 2.5) ** 2  # noqa: E731
    import ctypes as C
    cb = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)(f)
    x = C.c_double(1.0)
    ev = C.c_int(0)
    err = C.create_string_buffer(256)
    L = O.lib()
    L.o_brent_max.argtypes = [C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p), C.c_void_p,
                              C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                              C.POINTER(C.c_int), C.c_char_p, C.c_int]
    rc = L.o_brent_max(cb, None, 0.01, 10.0, 1e-6, C.byref(x), C.byref(ev), err, 256)
    assert rc == 0
    assert abs(x.value - 2.5) < 1e-5


# ---------------------------------------------------------------- parsing

def test_newick_numbering_matches_T(params_dir):
    t = O.newick_tree("(A,(B,C))")
    # leaves left to right, internal nodes in post order, root last
    assert t.labels[:3] == ["A", "B", "C"]
    assert list(t.parent[:4]) == [4, 3, 3, 4]
    assert (t.lc[3], t.rc[3]) == (1, 2) and (t.lc[4], t.rc[4]) == (0, 3)
    nh = open(_params(params_dir, "58mammals.nh")).read()
    t = O.newick_tree(nh)
    assert t.n == 115 and t.nl == 58
    for i in range(t.nl, t.n):
        assert t.lc[i] < i and t.rc[i] < i


def test_newick_subtree_merges_unary(params_dir):
    t = O.newick_tree("((a:1,b:2):3,(c:4,d:5):6);", keep=["a", "c", "d"])
    assert t.n == 5 and t.labels[:3] == ["a", "c", "d"]
    # a's branch absorbs its dropped parent: 1 + 3
    assert t.branch[0] == 4.0
    assert O.newick_total_length("((a:1,b:2):3,(c:4,d:5):6);") == 21.0
    assert O.newick_total_length("((a:1,b:2):3,(c:4,d:5):6);", keep=["a", "c"]) == 14.0


def test_newick_rejects_nonbinary_and_numeric_leaf():
    with pytest.raises(O.OracleError):
        O.newick_tree("(a,b,c);")
    with pytest.raises(O.OracleError):
        O.newick_tree("(123,b);")  # BRANCHLEN token where a LABEL is needed


def test_ecm_reader(params_dir):
    for name in ["12flies", "58mammals"]:
        for kind in ["coding", "noncoding"]:
            S, pi = O.ecm_read(_params(params_dir, "%s_%s.ECM" % (name, kind)))
            assert np.allclose(S, S.T) and np.all(np.diag(S) == 0)
            assert abs(pi.sum() - 1) < 1e-6
            Q = O.build_q(S, pi)
            assert np.allclose(Q.sum(1), 0, atol=1e-12)
            # normalised to unit mean rate
            assert abs(-(pi * np.diag(Q)).sum() - 1) < 1e-12


# ---------------------------------------------------------------- P(t)

def test_pt_jc69_closed_form():
    Q = (np.ones((4, 4)) - 4 * np.eye(4)) / 3.0
    pi = np.full(4, 0.25)
    for t in [0.0, 0.01, 0.3, 2.0]:
        P = O.pt(Q, pi, t)
        same = 0.25 + 0.75 * np.exp(-4 * t / 3)
        diff = 0.25 - 0.25 * np.exp(-4 * t / 3)
        expect = np.where(np.eye(4) == 1, same, diff)
        assert np.allclose(P, expect, atol=1e-14, rtol=0)


def test_pt_properties_codon(params_dir):
    S, pi = O.ecm_read(_params(params_dir, "58mammals_coding.ECM"))
    Q = O.build_q(S, pi)
    for t in [0.002, 0.1, 1.37]:
        P = O.pt(Q, pi, t)
        assert np.allclose(P.sum(1), 1, atol=1e-13)
        assert (P >= 0).all()
        assert np.allclose(pi[:, None] * P, (pi[:, None] * P).T, atol=1e-15)  # detailed balance
        assert np.allclose(P, scipy.linalg.expm(Q * t), atol=1e-12)
    assert np.allclose(O.pt(Q, pi, 0.0), np.eye(64), atol=1e-13)
    Pa, Pb, Pab = O.pt(Q, pi, 0.2), O.pt(Q, pi, 0.5), O.pt(Q, pi, 0.7)
    assert np.allclose(Pa @ Pb, Pab, atol=1e-13)  # Chapman-Kolmogorov
    assert np.allclose(O.pt(Q, pi, 200.0), np.tile(pi, (64, 1)), atol=1e-10)
    dP = O.dpt(Q, Pa)
    h = 1e-6
    fd = (O.pt(Q, pi, 0.2 + h) - O.pt(Q, pi, 0.2 - h)) / (2 * h)
    assert np.allclose(dP, fd, atol=1e-8)


# ---------------------------------------------------------------- pruning

def test_testinfer_kat():
    """lib/CamlPaml/test.ml:8-54: all 8 leaf patterns, closed-form alpha/beta."""
    t = O.newick_tree("(A,(B,C))")
    sm0 = np.array([[0.8, 0.2], [0.25, 0.75]])
    sm1 = np.array([[0.9, 0.1], [0.85, 0.15]])
    P = np.stack([sm0, sm1, sm1, sm1])
    prior = np.array([0.6, 0.4])
    for lvs in itertools.product([0, 1], repeat=3):
        a0, a1, a2 = [np.eye(2)[v] for v in lvs]
        a3 = (sm1 @ a1) * (sm1 @ a2)
        a4 = (sm0 @ a0) * (sm1 @ a3)
        z = prior @ a4
        b4 = prior
        b3 = np.array([(b4[0] * sm0[0] @ a0) * sm1[0, 0] + (b4[1] * sm0[1] @ a0) * sm1[1, 0],
                       (b4[0] * sm0[0] @ a0) * sm1[0, 1] + (b4[1] * sm0[1] @ a0) * sm1[1, 1]])
        zz, post, _ = O.prune(t, P, prior, np.array([lvs]), posteriors=True)
        assert abs(zz[0] - z) < 1e-12
        assert np.allclose(post[0, 4], b4 * a4 / z, atol=1e-12)
        assert np.allclose(post[0, 3], b3 * a3 / z, atol=1e-12)


def test_tinytest_jc69_kat(examples_dir):
    """lib/CamlPaml/test.ml:56-99: sum log L = -1574.63623 +- 0.001."""
    seqs = {}
    for line in open(REF_TEST_ML):
        k, v = line.split()
        seqs[k] = v
    t = O.newick_tree("((human:0.1,chimp:0.1):0.1,gorilla:0.2);")
    Q = (np.ones((4, 4)) - 4 * np.eye(4)) / 3.0
    pi = np.full(4, 0.25)
    P = np.stack([O.pt(Q, pi, b) for b in t.branch[:-1]])
    idx = {"A": 0, "C": 1, "G": 2, "T": 3}
    L = np.array([[idx.get(seqs[s][i], -1) for s in ("human", "chimp", "gorilla")]
                  for i in range(len(seqs["human"]))])
    z, _, _ = O.prune(t, P, pi, L)
    assert abs(np.log(z).sum() - (-1574.63623)) < 1e-3


@pytest.fixture(scope="module")
def codon_small(params_dir):
    S, pi = O.ecm_read(_params(params_dir, "12flies_coding.ECM"))
    Q = O.build_q(S, pi)
    return Q, pi


def test_pruning_invariants(codon_small):
    Q, pi = codon_small
    t = O.newick_tree("((a:0.1,b:0.3):0.2,(c:0.05,d:0.4):0.15);")
    P = np.stack([O.pt(Q, pi, b) for b in t.branch[:-1]])
    rng = np.random.default_rng(0)
    # all marginalised -> 1 ; one present -> pi_c ; two present -> pi_a P(ta+tb)_ab
    z, _, _ = O.prune(t, P, pi, np.array([[-1, -1, -1, -1]]))
    assert abs(z[0] - 1) < 1e-13
    for c in rng.integers(0, 64, 5):
        z, _, _ = O.prune(t, P, pi, np.array([[-1, -1, int(c), -1]]))
        assert abs(z[0] - pi[c]) < 1e-15
    Pab = O.pt(Q, pi, 0.1 + 0.3)
    Pac = O.pt(Q, pi, 0.1 + 0.2 + 0.15 + 0.05)
    for a, b in rng.integers(0, 64, (5, 2)):
        z, _, _ = O.prune(t, P, pi, np.array([[int(a), int(b), -1, -1]]))
        assert abs(z[0] - pi[a] * Pab[a, b]) < 1e-15
        z, _, _ = O.prune(t, P, pi, np.array([[int(a), -1, int(b), -1]]))
        assert abs(z[0] - pi[a] * Pac[a, b]) < 1e-15


def test_pruning_brute_force(codon_small):
    """Enumerate both internal states of ((x,y),z) plus the root: 64^2 terms."""
    Q, pi = codon_small
    t = O.newick_tree("((x:0.3,y:0.2):0.1,z:0.5);")
    P = np.stack([O.pt(Q, pi, b) for b in t.branch[:-1]])
    rng = np.random.default_rng(1)
    for x, y, zc in rng.integers(0, 64, (4, 3)):
        brute = 0.0
        for r in range(64):
            for u in range(64):
                brute += pi[r] * P[3][r, u] * P[0][u, x] * P[1][u, y] * P[2][r, zc]
        z, post, _ = O.prune(t, P, pi, np.array([[x, y, zc]]), posteriors=True)
        assert abs(z[0] - brute) / brute < 1e-12
        assert abs(post[0, 4].sum() - 1) < 1e-12 and abs(post[0, 3].sum() - 1) < 1e-12


def test_reroot_invariance(codon_small):
    """Reversible model: moving the root along a branch keeps z (pulley principle)."""
    Q, pi = codon_small
    t1 = O.newick_tree("((a:0.1,b:0.3):0.2,c:0.4);")
    t2 = O.newick_tree("(a:0.1,(b:0.3,c:0.6):0.0);")
    P1 = np.stack([O.pt(Q, pi, b) for b in t1.branch[:-1]])
    P2 = np.stack([O.pt(Q, pi, b) for b in t2.branch[:-1]])
    rng = np.random.default_rng(2)
    L = rng.integers(-1, 64, (20, 3))
    z1, _, _ = O.prune(t1, P1, pi, L)
    z2, _, _ = O.prune(t2, P2, pi, L)
    assert np.allclose(z1, z2, rtol=1e-12)


# ---------------------------------------------------------------- model level

@pytest.fixture(scope="module")
def flies(params_dir):
    t = O.newick_tree(open(_params(params_dir, "12flies.nh")).read())
    Sc, pc = O.ecm_read(_params(params_dir, "12flies_coding.ECM"))
    Sn, pn = O.ecm_read(_params(params_dir, "12flies_noncoding.ECM"))
    return t, O.Model(t, [Sc, Sn], [pc, pn]), O.Model(t, [Sn, Sc], [pn, pc]), (Sc, pc)


def test_fixed_score_invariants(flies):
    t, m, m_swapped, (Sc, pc) = flies
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 64, (30, t.nl)).astype(np.uint8)
    codes[rng.random(codes.shape) < 0.2] = 255
    a = m.score_fixed(codes)
    b = m_swapped.score_fixed(codes)
    assert abs(a["score"] + b["score"]) < 1e-9 and abs(a["anc_comp"] + b["anc_comp"]) < 1e-9
    same = O.Model(t, [Sc, Sc], [pc, pc])
    s = same.score_fixed(codes)
    assert s["score"] == 0.0 and s["anc_comp"] == 0.0
    assert abs(a["score"] - 10 * (a["lpr_c"] - a["lpr_n"]) / np.log(10)) < 1e-9


def test_gradient_matches_fd_and_sum(flies):
    t, m, _, _ = flies
    rng = np.random.default_rng(4)
    codes = rng.integers(0, 64, (12, t.nl)).astype(np.uint8)
    codes[rng.random(codes.shape) < 0.15] = 255
    for model in (0, 1):
        for rho in (0.4, 1.0, 2.3):
            l, d, g = m.lpr_grad(model, rho, codes)
            h = 1e-5
            fd = (m.lpr(model, rho + h, codes)[0] - m.lpr(model, rho - h, codes)[0]) / (2 * h)
            assert abs(d - fd) < 1e-6 * max(1.0, abs(fd))
            assert abs(d - (g * t.branch[:-1]).sum()) < 1e-9 * max(1.0, abs(d))
            assert abs(l - m.lpr(model, rho, codes)[0]) < 1e-12 * abs(l)


# ---------------------------------------------------------------- published CLI results

def _run(params_dir, examples_dir, pset, fa, **kw):
    return CO.run(_params(params_dir, pset), open(os.path.join(examples_dir, fa)).read(), **kw)


def test_cli_talAA(params_dir, examples_dir):
    """src/test.ml:27-31."""
    _, (s, rc, lo, hi) = _run(params_dir, examples_dir, "12flies", "tal-AA.fa")
    assert 297.62 < s["score"] < 297.63
    assert 48.25 < s["anc_comp"] < 48.26


def test_cli_aldh2_exon5_out(params_dir, examples_dir):
    """src/test.ml:33-37."""
    _, (s, rc, lo, hi) = _run(params_dir, examples_dir, "29mammals", "ALDH2.exon5.fa")
    assert -178.93 < s["score"] < -178.92
    assert -38.29 < s["anc_comp"] < -38.28


def test_cli_aldh2_exon5_in_6frames(params_dir, examples_dir):
    """src/test.ml:39-47."""
    _, (s, rc, lo, hi) = _run(params_dir, examples_dir, "29mammals", "ALDH2.exon5.fa", frames=6)
    assert 218.26 < s["score"] < 218.27
    assert (lo, hi, rc) == (1, 111, False)


@pytest.mark.slow
def test_cli_aldh2_mrna_orf(params_dir, examples_dir):
    """src/test.ml:49-57."""
    _, (s, rc, lo, hi) = _run(params_dir, examples_dir, "29mammals", "Aldh2.mRNA.fa", frames=3,
                              orf="ATGStop", remove_ref_gaps_=True)
    assert 2013.92 < s["score"] < 2013.93
    assert (lo, hi) == (343, 1899)
    sp, seqs = CO.input_mfa(open(os.path.join(examples_dir, "Aldh2.mRNA.fa")).read())
    seqs = CO.remove_ref_gaps(seqs)
    assert re.match("^MLRAALTTVRRGPRLSRLLSAAA", CO.translate(seqs[0][lo:hi + 1]))
