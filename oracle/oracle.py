"""ctypes front end of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry and the
bench's cpu_baseline leg, as the checker.  The product never imports it.

Each wrapper names the reference function it restates (paths relative to
the mlin/PhyloCSF source tree).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    """Compile liboracle.so from the C restatement (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        u8 = C.POINTER(C.c_uint8)
        L.o_ecm_read.argtypes = [C.c_char_p, dp, dp, C.c_char_p, C.c_int]
        L.o_newick_tree.argtypes = [C.c_char_p, C.c_char_p, C.c_int, ip, ip, ip, ip, dp,
                                    C.c_char_p, C.c_int, C.c_char_p, C.c_int]
        L.o_newick_total_length.argtypes = [C.c_char_p, C.c_char_p, dp, C.c_char_p, C.c_int]
        L.o_build_q.argtypes = [C.c_int, dp, dp, dp, C.c_char_p, C.c_int]
        L.o_pt.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, dp, C.c_char_p, C.c_int]
        L.o_dpt.argtypes = [C.c_int, dp, dp, dp]
        L.o_prune.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, ip, C.c_int64, dp, dp, dp, dp]
        L.o_model_new.restype = C.c_void_p
        L.o_model_new.argtypes = [C.c_int, ip, ip, dp, C.c_int, dp, dp, C.c_char_p, C.c_int]
        L.o_model_free.argtypes = [C.c_void_p]
        L.o_lpr.argtypes = [C.c_void_p, C.c_int, C.c_double, u8, C.c_int64, dp, dp,
                            C.c_char_p, C.c_int]
        L.o_lpr_grad.argtypes = [C.c_void_p, C.c_int, C.c_double, u8, C.c_int64, dp, dp, dp,
                                 C.c_char_p, C.c_int]
        L.o_score_fixed.argtypes = [C.c_void_p, u8, C.c_int64, dp, C.c_char_p, C.c_int]
        L.o_score_mle.argtypes = [C.c_void_p, u8, C.c_int64, dp, C.c_char_p, C.c_int]
        L.o_ocaml_random_floats.argtypes = [C.c_int, C.c_double, C.c_int, dp]
        L.o_md5.argtypes = [C.c_char_p, C.c_uint64, C.c_char_p]
        L.o_lik_posteriors.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, ip, dp, C.c_int64, dp,
                                       dp, dp, dp]
        L.o_omega_pi.argtypes = [dp, dp]
        L.o_omega_s.argtypes = [dp, dp]
        L.o_omega_f3x4.argtypes = [u8, C.c_int64, C.c_int, dp]
        L.o_half_cauchy_lpdf.argtypes = [C.c_double, C.c_double, C.c_double]
        L.o_half_cauchy_lpdf.restype = C.c_double
        L.o_lpr_kappa.argtypes = [C.c_double]
        L.o_lpr_kappa.restype = C.c_double
        L.o_omega_lpr.argtypes = [C.c_int, ip, ip, dp, dp, C.c_double, u8, C.c_int64, dp,
                                  C.c_char_p, C.c_int]
        L.o_omega_score.argtypes = [C.c_int, ip, ip, dp, u8, C.c_int64, C.c_double, C.c_double,
                                    dp, C.c_char_p, C.c_int]
        _LIB = L
    return _LIB


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _check(rc, err):
    if rc < 0:
        raise OracleError(rc, err.value.decode(errors="replace"))


def ecm_read(path):
    """ECM.import_parameters (src/ECM.ml:18-72) -> (S[64,64], pi[64])."""
    S = np.zeros((64, 64))
    pi = np.zeros(64)
    err = C.create_string_buffer(512)
    _check(lib().o_ecm_read(path.encode(), _dp(S), _dp(pi), err, 512), err)
    return S, pi


class Tree:
    """T.t (lib/CamlPaml/T.ml): leaves 0..nl-1, post-order internals, root n-1."""

    def __init__(self, n, parent, lc, rc, branch, labels):
        self.n, self.nl = n, (n + 1) // 2
        self.parent, self.lc, self.rc = parent, lc, rc
        self.branch, self.labels = branch, labels


def newick_tree(text, keep=None, max_n=4096):
    """NewickParser + [Newick.subtree] + T.of_newick."""
    n = C.c_int(0)
    par = np.zeros(max_n, dtype=np.int32)
    lc = np.zeros(max_n, dtype=np.int32)
    rc = np.zeros(max_n, dtype=np.int32)
    br = np.zeros(max_n)
    lab = C.create_string_buffer(1 << 20)
    err = C.create_string_buffer(512)
    keep_b = None if keep is None else ",".join(keep).encode()
    rc_ = lib().o_newick_tree(text.encode(), keep_b, max_n, C.byref(n), _ip(par), _ip(lc),
                              _ip(rc), _dp(br), lab, 1 << 20, err, 512)
    _check(rc_, err)
    m = n.value
    labels = lab.value.decode().split("\n")
    return Tree(m, par[:m].copy(), lc[:m].copy(), rc[:m].copy(), br[:m].copy(), labels)


def newick_total_length(text, keep=None):
    out = C.c_double(0)
    err = C.create_string_buffer(512)
    keep_b = None if keep is None else ",".join(keep).encode()
    rc_ = lib().o_newick_total_length(text.encode(), keep_b, C.byref(out), err, 512)
    _check(rc_, err)
    return out.value if rc_ == 0 else None


def build_q(S, pi):
    k = len(pi)
    S = np.ascontiguousarray(S, dtype=np.float64)
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    Q = np.zeros((k, k))
    err = C.create_string_buffer(512)
    _check(lib().o_build_q(k, _dp(S), _dp(pi), _dp(Q), err, 512), err)
    return Q


def pt(Q, pi, t, tol=1e-6):
    """Q.Diag.real_to_Pt (lib/CamlPaml/Q.ml:211-249)."""
    k = len(pi)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    P = np.zeros((k, k))
    err = C.create_string_buffer(512)
    _check(lib().o_pt(k, _dp(Q), _dp(pi), float(t), float(tol), _dp(P), err, 512), err)
    return P


def dpt(Q, P):
    k = Q.shape[0]
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    P = np.ascontiguousarray(P, dtype=np.float64)
    out = np.zeros((k, k))
    lib().o_dpt(k, _dp(Q), _dp(P), _dp(out))
    return out


def prune(tree, P, prior, leaves, posteriors=False, dP=None):
    """PhyloLik.prepare/likelihood/node_posterior (+ branch posteriors -> d ln z/dt).

    leaves: int array [ncols, nl], -1 = `Marginalize.
    Returns z [ncols], post [ncols, n, k] or None, dlogz_dt [ncols, n-1] or None.
    """
    prior = np.ascontiguousarray(prior, dtype=np.float64)
    k = len(prior)
    leaves = np.ascontiguousarray(leaves, dtype=np.int32)
    ncols = leaves.shape[0]
    P = np.ascontiguousarray(P, dtype=np.float64)
    z = np.zeros(ncols)
    post = np.zeros((ncols, tree.n, k)) if posteriors else None
    g = np.zeros((ncols, tree.n - 1)) if dP is not None else None
    dPa = np.ascontiguousarray(dP, dtype=np.float64) if dP is not None else None
    lib().o_prune(k, tree.n, _ip(tree.lc.astype(np.int32)), _ip(tree.rc.astype(np.int32)),
                  _dp(P), _dp(prior), _ip(leaves), ncols, _dp(z),
                  _dp(post) if post is not None else None,
                  _dp(dPa) if dPa is not None else None,
                  _dp(g) if g is not None else None)
    return z, post, g


def lik_posteriors(tree, P, prior, leaves, leaf_dist=None, weights=None, posteriors=True,
                   ecounts=True):
    """PhyloLik.prepare with `Certain / `Distribution / `Marginalize leaves
    (leaves: [ncols, nl] int, -1 marginalise, -2 distribution from
    leaf_dist [ncols, nl, k]), likelihood, node_posterior of every node
    and add_branch_posteriors of every branch summed with weights
    (PhyloEM.collect_sufficient_statistics).
    -> z [ncols], post [ncols, n, k] or None, ecounts [n-1, k, k] or None."""
    prior = np.ascontiguousarray(prior, dtype=np.float64)
    k = len(prior)
    leaves = np.ascontiguousarray(leaves, dtype=np.int32)
    ncols = leaves.shape[0]
    P = np.ascontiguousarray(P, dtype=np.float64)
    ld = None if leaf_dist is None else np.ascontiguousarray(leaf_dist, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    z = np.zeros(ncols)
    post = np.zeros((ncols, tree.n, k)) if posteriors else None
    ec = np.zeros((tree.n - 1, k, k)) if ecounts else None
    lib().o_lik_posteriors(k, tree.n, _ip(tree.lc.astype(np.int32)), _ip(tree.rc.astype(np.int32)),
                           _dp(P), _dp(prior), _ip(leaves), None if ld is None else _dp(ld), ncols,
                           None if w is None else _dp(w), _dp(z),
                           None if post is None else _dp(post), None if ec is None else _dp(ec))
    return z, post, ec


class Model:
    """PhyloCSFModel.make over one tree and nm ECMs (0 = coding, 1 = noncoding)."""

    def __init__(self, tree, S_list, pi_list):
        self.tree = tree
        S = np.ascontiguousarray(np.stack(S_list), dtype=np.float64)
        pi = np.ascontiguousarray(np.stack(pi_list), dtype=np.float64)
        err = C.create_string_buffer(512)
        self._lc = tree.lc.astype(np.int32)
        self._rc = tree.rc.astype(np.int32)
        self._br = tree.branch.astype(np.float64)
        h = lib().o_model_new(tree.n, _ip(self._lc), _ip(self._rc), _dp(self._br), len(S_list),
                              _dp(S), _dp(pi), err, 512)
        if not h:
            raise OracleError(-1, err.value.decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().o_model_free(self.h)
            self.h = None

    @staticmethod
    def _codes(codes):
        return np.ascontiguousarray(codes, dtype=np.uint8)

    def lpr(self, model, rho, codes):
        """lpr_leaves (src/PhyloCSFModel.ml:67-82) -> (lpr, elpr_anc)."""
        c = self._codes(codes)
        a, b = C.c_double(0), C.c_double(0)
        err = C.create_string_buffer(512)
        _check(lib().o_lpr(self.h, model, float(rho), _u8(c), c.shape[0], C.byref(a),
                           C.byref(b), err, 512), err)
        return a.value, b.value

    def lpr_grad(self, model, rho, codes):
        """lpr with d lpr/d rho and d lpr/d t_br (PhyloEM.ml:140-182 via branch posteriors)."""
        c = self._codes(codes)
        a, d = C.c_double(0), C.c_double(0)
        g = np.zeros(self.tree.n - 1)
        err = C.create_string_buffer(512)
        _check(lib().o_lpr_grad(self.h, model, float(rho), _u8(c), c.shape[0], C.byref(a),
                                C.byref(d), _dp(g), err, 512), err)
        return a.value, d.value, g

    def score_fixed(self, codes):
        """llr_FixedLik 1.0 -> dict(score, anc_comp, lpr_c, lpr_n)."""
        c = self._codes(codes)
        out = np.zeros(4)
        err = C.create_string_buffer(512)
        _check(lib().o_score_fixed(self.h, _u8(c), c.shape[0], _dp(out), err, 512), err)
        return dict(score=out[0], anc_comp=out[1], lpr_c=out[2], lpr_n=out[3])

    def score_mle(self, codes):
        """llr_MaxLik ~init:1 with GSL Brent + Fit.find_init."""
        c = self._codes(codes)
        out = np.zeros(7)
        err = C.create_string_buffer(512)
        _check(lib().o_score_mle(self.h, _u8(c), c.shape[0], _dp(out), err, 512), err)
        return dict(score=out[0], anc_comp=out[1], lpr_c=out[2], lpr_n=out[3], rho_c=out[4],
                    rho_n=out[5], evals=int(out[6]))


def ocaml_random_floats(seed, bound, n):
    out = np.zeros(n)
    lib().o_ocaml_random_floats(int(seed), float(bound), int(n), _dp(out))
    return out


def md5(data: bytes) -> bytes:
    out = C.create_string_buffer(16)
    lib().o_md5(data, len(data), out)
    return out.raw


# ---- OmegaModel (src/OmegaModel.ml) -----------------------------------------

def omega_pi(qs):
    """pi_exprs (OmegaModel.ml:24-40) for the 12 settings, not normalised."""
    qs = np.ascontiguousarray(qs, dtype=np.float64)
    out = np.zeros(64)
    lib().o_omega_pi(_dp(qs), _dp(out))
    return out


def omega_s(qs):
    """exchangeabilities s (q_ij = pi_j s_ij) of q_p14n (OmegaModel.ml:42-74)."""
    qs = np.ascontiguousarray(qs, dtype=np.float64)
    out = np.zeros((64, 64))
    lib().o_omega_s(_dp(qs), _dp(out))
    return out


def omega_f3x4(codes):
    """update_f3x4 (OmegaModel.ml:102-134): settings 3..11 from codes [ncols, nl]."""
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.zeros(9)
    lib().o_omega_f3x4(_u8(c), c.shape[0], c.shape[1], _dp(out))
    return out


def half_cauchy_lpdf(mode, scale, x):
    return lib().o_half_cauchy_lpdf(float(mode), float(scale), float(x))


def lpr_kappa(k):
    return lib().o_lpr_kappa(float(k))


def omega_lpr(tree, qs, rho, codes):
    """OmegaModel.lpr_leaves (OmegaModel.ml:137-144)."""
    qs = np.ascontiguousarray(qs, dtype=np.float64)
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    out = C.c_double(0)
    err = C.create_string_buffer(512)
    _check(lib().o_omega_lpr(tree.n, _ip(tree.lc.astype(np.int32)), _ip(tree.rc.astype(np.int32)),
                             _dp(tree.branch.astype(np.float64)), _dp(qs), float(rho), _u8(c),
                             c.shape[0], C.byref(out), err, 512), err)
    return out.value


def omega_score(tree, codes, omega_H1=0.2, sigma_H1=0.01):
    """OmegaModel.score (OmegaModel.ml:195-219) -> dict."""
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.zeros(8)
    err = C.create_string_buffer(512)
    _check(lib().o_omega_score(tree.n, _ip(tree.lc.astype(np.int32)),
                               _ip(tree.rc.astype(np.int32)), _dp(tree.branch.astype(np.float64)),
                               _u8(c), c.shape[0], float(omega_H1), float(sigma_H1), _dp(out),
                               err, 512), err)
    return dict(score=out[0], lpr_H0=out[1], lpr_H1=out[2], rho_H0=out[3], kappa_H0=out[4],
                rho_H1=out[5], kappa_H1=out[6], evals=int(out[7]))
