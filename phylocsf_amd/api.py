"""Host-side mirror of the reference's scoring interface over the HIP C ABI.

Names follow the reference so a caller of CamlPaml / PhyloCSF finds the
same entry points (paths relative to the mlin/PhyloCSF tree):

  import_parameters ........ src/ECM.ml:18-72
  newick_tree .............. NewickParser + Newick.subtree + T.of_newick
  PhyloCSFModel ............ src/PhyloCSFModel.ml:107-146 (make / lpr_leaves / score)
  PhyloLik ................. lib/CamlPaml/PhyloLik.mli:6-35 (prepare / likelihood /
                             node_posterior at the root), batched over columns
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from ._lib import (PCSF_BF16, PCSF_FP16, PCSF_FP16S, PCSF_FP32, PCSF_FP64,  # noqa: F401
                   PCSF_FP64M, PCSF_MARGINALIZE, PCSF_N_VARIANTS, PCSF_STRATEGY_FIXED, PCSF_STRATEGY_MLE,
                   PCSF_STRATEGY_MLE_GRAD, InvalidArgument, lib, raise_for)

DB = 10.0 / np.log(10.0)  # PhyloCSFModel.db (src/PhyloCSFModel.ml:113)
_HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS_DIR = os.path.join(_HERE, "data", "PhyloCSF_Parameters")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def import_parameters(path: str):
    """ECM.import_parameters: -> (S [64,64], pi [64])."""
    S = np.zeros((64, 64))
    pi = np.zeros(64)
    err = C.create_string_buffer(512)
    rc = lib().pcsf_ecm_read(path.encode(), _dp(S), _dp(pi), err, 512)
    if rc:
        raise_for(rc, err.value.decode())
    return S, pi


@dataclass
class Tree:
    """T.t: leaves 0..nl-1 left to right, post-order internals, root n-1."""
    n: int
    children: np.ndarray  # (n-nl, 2) int32
    branch: np.ndarray    # (n,) float64, NaN where unspecified
    labels: List[str]

    @property
    def nl(self):
        return (self.n + 1) // 2

    def parent(self):
        p = np.full(self.n, -1, dtype=np.int32)
        for i in range(self.nl, self.n):
            l, r = self.children[i - self.nl]
            p[l] = p[r] = i
        return p


def newick_tree(text: str, species: Optional[Sequence[str]] = None, max_nodes=1 << 15) -> Tree:
    """NewickParser.parse + optional Newick.subtree + T.of_newick."""
    n = C.c_int(0)
    ch = np.zeros(max_nodes, dtype=np.int32)
    br = np.zeros(max_nodes)
    lab = C.create_string_buffer(1 << 20)
    err = C.create_string_buffer(512)
    keep = None if species is None else ",".join(species).encode()
    rc = lib().pcsf_newick_tree(text.encode(), keep, max_nodes, C.byref(n), _i32(ch), _dp(br),
                                lab, 1 << 20, err, 512)
    if rc:
        raise_for(rc, err.value.decode())
    m = n.value
    nl = (m + 1) // 2
    return Tree(m, ch[:2 * (m - nl)].reshape(m - nl, 2).copy(), br[:m].copy(),
                lab.value.decode().split("\n"))


def paramset_prefix(paramset: str) -> str:
    """Parameter-set lookup (src/PhyloCSF.ml:407-421): a bare name resolves under
    $PHYLOCSF_BASE/PhyloCSF_Parameters (default: the copy shipped in this
    package); a name containing '/' is a path prefix."""
    if "/" in paramset:
        return paramset
    base = os.environ.get("PHYLOCSF_BASE")
    if base:
        if not os.path.isdir(base):
            raise RuntimeError("PHYLOCSF_BASE environment variable must be set to the root "
                               "directory of the source or executable distribution")
        return os.path.join(base, "PhyloCSF_Parameters", paramset)
    return os.path.join(PARAMS_DIR, paramset)


def program(tree: Tree, with_tables=False):
    """The fused kernel's op program for a tree: (ops [n_ops, 4] int32, slots)."""
    ch = np.ascontiguousarray(tree.children, dtype=np.int32)
    n = C.c_int(0)
    sl = C.c_int(0)
    rc = lib().pcsf_program(tree.n, _i32(ch), 1 if with_tables else 0, None, 0, C.byref(n),
                            C.byref(sl))
    if rc:
        raise_for(rc, "pcsf_program: invalid tree")
    ops = np.zeros(n.value * 4, dtype=np.int32)
    rc = lib().pcsf_program(tree.n, _i32(ch), 1 if with_tables else 0, _i32(ops), ops.size,
                            C.byref(n), C.byref(sl))
    if rc:
        raise_for(rc, "pcsf_program failed")
    return ops.reshape(-1, 4), sl.value


class Context:
    """One HIP stream + the device memory of its models and batches."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib().pcsf_ctx_create(device, C.byref(h))
        if rc:
            raise_for(rc, "pcsf_ctx_create(device=%d) failed" % device)
        self.h = h

    def check(self, rc):
        if rc:
            raise_for(rc, lib().pcsf_last_error(self.h).decode(errors="replace"))

    def set_timing(self, on=True):
        self.check(lib().pcsf_ctx_set_timing(self.h, 1 if on else 0))

    def synchronize(self):
        self.check(lib().pcsf_ctx_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().pcsf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _HostOnly:
    """Stands in for a Context when a model is host-only (no device)."""
    h = None

    def check(self, rc):
        if rc:
            raise_for(rc, "host-only model: the call needs a device context")


class PhyloCSFModel:
    """PhyloCSFModel.make over one tree and n ECMs (0 = coding, 1 = noncoding)."""

    def __init__(self, ctx: Optional[Context], tree: Tree, ecms: Sequence):
        """ctx None: a host-only model (simulation and model facts only)."""
        self.ctx = ctx if ctx is not None else _HostOnly()
        self.tree = tree
        S = np.ascontiguousarray(np.stack([e[0] for e in ecms]), dtype=np.float64)
        pi = np.ascontiguousarray(np.stack([e[1] for e in ecms]), dtype=np.float64)
        self.pi = pi
        self.nm = len(ecms)
        ch = np.ascontiguousarray(tree.children, dtype=np.int32)
        br = np.ascontiguousarray(tree.branch, dtype=np.float64).copy()
        br[-1] = 0.0  # the root branch is unused
        h = C.c_void_p()
        self.ctx.check(lib().pcsf_model_create(self.ctx.h, tree.n, _i32(ch), _dp(br), self.nm,
                                               _dp(S), _dp(pi), C.byref(h)))
        self.h = h

    @classmethod
    def tree_only(cls, ctx: Context, tree: Tree):
        """A tree without rate models, for the omega test (pcsf_model_create
        with n_models = 0)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.tree = tree
        self.nm = 2
        self.pi = None
        ch = np.ascontiguousarray(tree.children, dtype=np.int32)
        br = np.ascontiguousarray(tree.branch, dtype=np.float64).copy()
        br[-1] = 0.0
        h = C.c_void_p()
        ctx.check(lib().pcsf_model_create(ctx.h, tree.n, _i32(ch), _dp(br), 0, None, None,
                                          C.byref(h)))
        self.h = h
        return self

    @classmethod
    def from_paramset(cls, ctx: Optional[Context], paramset: str, species=None):
        prefix = paramset_prefix(paramset)
        tree = newick_tree(open(prefix + ".nh").read(), species)
        return cls(ctx, tree, [import_parameters(prefix + "_coding.ECM"),
                               import_parameters(prefix + "_noncoding.ECM")])

    def specialized(self, enable=None):
        """Allow (True) or forbid (False) the build-time tree-specialised
        kernels for this model; None leaves the setting.  Returns which of the
        model's programs have one: {"plain": bool, "tables": bool}."""
        av = C.c_int(0)
        self.ctx.check(lib().pcsf_model_specialized(self.h, -1 if enable is None else int(bool(enable)),
                                                    C.byref(av)))
        return {"plain": bool(av.value & 1), "tables": bool(av.value & 2)}

    def set_variant(self, v, precision=PCSF_FP32):
        self.ctx.check(lib().pcsf_set_variant(self.h, int(precision), int(v)))

    def info(self):
        n, nl, nm, sl = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.ctx.check(lib().pcsf_model_info(self.h, C.byref(n), C.byref(nl), C.byref(nm),
                                             C.byref(sl)))
        return dict(n=n.value, nl=nl.value, nm=nm.value, slots=sl.value)

    def transition(self, model: int, rho: float = 1.0):
        """P(rho*b) for every branch, built on the GPU (Q.Diag.to_Pt)."""
        P = np.zeros((self.tree.n - 1, 64, 64))
        self.ctx.check(lib().pcsf_model_transition(self.h, model, float(rho), _dp(P)))
        return P

    def batch(self, regions: Sequence[np.ndarray]) -> "Batch":
        return Batch.from_regions(self, regions)

    def simulate(self, lengths, region_model, region_rho, gap_mode=0, gap_p=0.0, gap_run=5.0,
                 n_p=0.0, seed=1, n_threads=0):
        """Synthetic packed codes [n_cols, nl] and region offsets."""
        lengths = np.asarray(lengths, dtype=np.int64)
        off = np.zeros(len(lengths) + 1, dtype=np.int64)
        np.cumsum(lengths, out=off[1:])
        codes = np.zeros((int(off[-1]), self.tree.nl), dtype=np.uint8)
        rm = np.ascontiguousarray(region_model, dtype=np.int32)
        rr = np.ascontiguousarray(region_rho, dtype=np.float64)
        rc = lib().pcsf_simulate(self.h, len(lengths), _i64(off), _i32(rm), _dp(rr), gap_mode,
                                 gap_p, gap_run, n_p, seed, n_threads, _u8(codes))
        if rc:
            raise_for(rc, "pcsf_simulate failed")
        return codes, off

    def simulate_nucleotides(self, lengths, region_model, region_rho, gap_mode=0, gap_p=0.0,
                             gap_run=5.0, n_p=0.0, seed=1, n_threads=0):
        """The alignments of simulate() as nucleotide characters:
        -> (nuc [n_leaves, 3 * n_cols] uint8 in tree-leaf order, nucleotide
        region offsets [R+1])."""
        lengths = np.asarray(lengths, dtype=np.int64)
        off = np.zeros(len(lengths) + 1, dtype=np.int64)
        np.cumsum(lengths, out=off[1:])
        nuc = np.zeros((self.tree.nl, 3 * int(off[-1])), dtype=np.uint8)
        rm = np.ascontiguousarray(region_model, dtype=np.int32)
        rr = np.ascontiguousarray(region_rho, dtype=np.float64)
        rc = lib().pcsf_simulate_nucleotides(self.h, len(lengths), _i64(off), _i32(rm), _dp(rr),
                                             gap_mode, gap_p, gap_run, n_p, seed, n_threads,
                                             _u8(nuc))
        if rc:
            raise_for(rc, "pcsf_simulate_nucleotides failed")
        return nuc, 3 * off

    def close(self):
        if getattr(self, "h", None):
            lib().pcsf_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Packed codon columns of many regions, resident in HBM."""

    def __init__(self, model: PhyloCSFModel, codes: np.ndarray, offsets: np.ndarray):
        self.model = model
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        if codes.ndim != 2 or codes.shape[1] != model.tree.nl:
            raise InvalidArgument(-1, "codes must be [n_cols, n_leaves]")
        # a code is a codon index 0..63 or PCSF_MARGINALIZE (255); anything
        # else would index past the kernels' 65-row leaf tables
        if codes.size and (codes.max() > 63) and ((codes > 63) & (codes != PCSF_MARGINALIZE)).any():
            raise InvalidArgument(-1, "pcsf_batch_create: a leaf code is neither a codon index "
                                      "(0..63) nor PCSF_MARGINALIZE (255)")
        self.n_regions = len(offsets) - 1
        self.n_cols = codes.shape[0]
        h = C.c_void_p()
        model.ctx.check(lib().pcsf_batch_create(model.h, _u8(codes), self.n_cols, _i64(offsets),
                                                self.n_regions, C.byref(h)))
        self.h = h

    @classmethod
    def frames(cls, model: PhyloCSFModel, nuc: np.ndarray, offsets: np.ndarray, frames: int):
        """Every reading frame of packed nucleotide regions, built on the GPU
        (pcsf_batch_create_frames): nuc is [n_leaves, n_nuc] characters in
        tree-leaf order; batch region r * frames + f is frame f of region r."""
        nuc = np.ascontiguousarray(nuc, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        if nuc.ndim != 2 or nuc.shape[0] != model.tree.nl:
            raise InvalidArgument(-1, "nuc must be [n_leaves, n_nuc]")
        b = cls.__new__(cls)
        b.model = model
        h = C.c_void_p()
        model.ctx.check(lib().pcsf_batch_create_frames(model.h, _u8(nuc), nuc.shape[1],
                                                       _i64(offsets), len(offsets) - 1,
                                                       int(frames), C.byref(h)))
        b.h = h
        b.n_regions = (len(offsets) - 1) * int(frames)
        b.n_cols = int(lib().pcsf_batch_columns(h))
        return b

    @classmethod
    def from_regions(cls, model, regions):
        lens = [r.shape[0] for r in regions]
        off = np.zeros(len(regions) + 1, dtype=np.int64)
        np.cumsum(lens, out=off[1:])
        codes = (np.concatenate(regions, axis=0) if regions
                 else np.zeros((0, model.tree.nl), np.uint8))
        return cls(model, codes, off)

    def codes(self):
        """The leaf codes the kernels read, back from the device:
        -> (codes [n_cols, n_leaves] uint8, region offsets [R+1])."""
        codes = np.zeros((self.n_cols, self.model.tree.nl), np.uint8)
        off = np.zeros(self.n_regions + 1, np.int64)
        self.model.ctx.check(lib().pcsf_batch_codes(self.h, _u8(codes), _i64(off)))
        return codes, off

    def forward(self, rho=1.0, precision=PCSF_FP32, per_region=False):
        """-> (loglik [R, nm], elpr [R, nm]): lpr_leaves for every region/model."""
        nm = self.model.nm
        if per_region:
            r = np.ascontiguousarray(rho, dtype=np.float64).reshape(self.n_regions * nm)
        else:
            r = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, dtype=np.float64), (nm,)))
        ll = np.zeros((self.n_regions, nm))
        el = np.zeros((self.n_regions, nm))
        self.model.ctx.check(lib().pcsf_forward(self.h, _dp(r), 1 if per_region else 0,
                                                precision, _dp(ll), _dp(el), 0))
        self.forward_count = getattr(self, "forward_count", 0) + 1
        return ll, el

    def backward(self, grad_loglik, branches=False):
        """Gradients of the last forward: -> (dL/drho [R, nm], dL/dt_br [R, nm, n-1] or None)
        for dL/dloglik = grad_loglik [R, nm]: FP64 outside pass, or the FP32
        tangent pass when the last forward ran at an MFMA precision and
        branches is False."""
        nm = self.model.nm
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(grad_loglik, dtype=np.float64),
                                                 (self.n_regions, nm)))
        grho = np.zeros((self.n_regions, nm))
        gbr = np.zeros((self.n_regions, nm, self.model.tree.n - 1)) if branches else None
        self.model.ctx.check(lib().pcsf_backward(self.h, _dp(g), _dp(grho),
                                                 _dp(gbr) if branches else None))
        return grho, gbr

    def autotune(self, reps=3, precision=PCSF_FP32):
        """Time every launch shape of the fused kernel at one MFMA precision on
        this batch and keep the fastest for this model and precision:
        -> (chosen index, [ms per shape])."""
        ms = np.zeros(PCSF_N_VARIANTS)
        ch = C.c_int(0)
        self.model.ctx.check(lib().pcsf_autotune(self.h, int(precision), reps, _dp(ms),
                                                 C.byref(ch)))
        return ch.value, ms

    def timing(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self.model.ctx.check(lib().pcsf_batch_last_timing(self.h, C.byref(a), C.byref(b),
                                                          C.byref(c)))
        return dict(prune_ms=a.value, reduce_ms=b.value, pbuild_ms=c.value)

    def close(self):
        if getattr(self, "h", None):
            lib().pcsf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def score(batch: Batch, strategy="fixed", precision=PCSF_FP32):
    """PhyloCSFModel.score for every region of a batch (src/PhyloCSFModel.ml:138-146).

    strategy "fixed" (llr_FixedLik 1.0), "mle" (llr_MaxLik: Fit.find_init + GSL
    Brent per region and model) or "mle_grad" (root of d lpr / d ln rho through the
    adjoint kernel).  -> dict of arrays: score, anc_comp (decibans),
    rho, lpr, elpr ([R, 2], coding first), status, n_evals."""
    st = {"fixed": PCSF_STRATEGY_FIXED, "mle": PCSF_STRATEGY_MLE,
          "mle_grad": PCSF_STRATEGY_MLE_GRAD}[strategy]
    R = batch.n_regions
    out = dict(score=np.zeros(R), anc_comp=np.zeros(R), rho=np.zeros((R, 2)),
               lpr=np.zeros((R, 2)), elpr=np.zeros((R, 2)),
               status=np.zeros(R, np.int32), n_evals=np.zeros(R, np.int32))
    # the maximisers run forwards of their own on this batch
    batch.forward_count = getattr(batch, "forward_count", 0) + 1
    batch.model.ctx.check(lib().pcsf_score(batch.h, st, precision, _dp(out["score"]),
                                           _dp(out["anc_comp"]), _dp(out["rho"]),
                                           _dp(out["lpr"]), _dp(out["elpr"]),
                                           _i32(out["status"]), _i32(out["n_evals"])))
    return out


def score_omega(batch: Batch, omega_H1=0.2, sigma_H1=0.01, precision=PCSF_FP64):
    """OmegaModel.score for every region of a batch (src/OmegaModel.ml:195-219):
    -> dict of arrays score (decibans), lpr, rho, kappa ([R, 2], H0 first),
    status, n_evals."""
    R = batch.n_regions
    out = dict(score=np.zeros(R), lpr=np.zeros((R, 2)), rho=np.zeros((R, 2)),
               kappa=np.zeros((R, 2)), status=np.zeros(R, np.int32),
               n_evals=np.zeros(R, np.int32))
    batch.model.ctx.check(lib().pcsf_score_omega(batch.h, float(omega_H1), float(sigma_H1),
                                                 int(precision), _dp(out["score"]),
                                                 _dp(out["lpr"]), _dp(out["rho"]),
                                                 _dp(out["kappa"]), _i32(out["status"]),
                                                 _i32(out["n_evals"])))
    return out


def score_fixed(batch: Batch, precision=PCSF_FP32):
    """llr_FixedLik 1.0 per region (src/PhyloCSFModel.ml:122-128):
    -> dict of arrays score, anc_comp, lpr_c, lpr_n (decibans for the first two)."""
    ll, el = batch.forward(1.0, precision)
    return dict(score=DB * (ll[:, 0] - ll[:, 1]), anc_comp=DB * (el[:, 0] - el[:, 1]),
                lpr_c=ll[:, 0], lpr_n=ll[:, 1])


class PhyloLik:
    """CamlPaml.PhyloLik over many columns at once (lib/CamlPaml/PhyloLik.mli):
    prepare(tree, pms, prior, leaves) then likelihood / node_posterior /
    branch_posteriors / add_branch_posteriors of the intermediate.

    leaves: [ncols, nl] int -- a state index (`Certain), -1 (`Marginalize) or
    -2 (`Distribution, its probabilities in leaf_dist[col, leaf, :])."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    @staticmethod
    def _inputs(tree, pms, prior, leaves, leaf_dist):
        """The shape checks of PhyloLik.prepare (the library checks the leaf
        states, `Distribution leaves without leaf_dist included, in one pass);
        -> (P, prior, leaves, leaf_dist, children)."""
        k = len(prior)
        leaves = np.ascontiguousarray(leaves, dtype=np.int32)
        if leaves.ndim == 1:
            leaves = leaves[None, :]
        if leaves.shape[1] != tree.nl:
            raise InvalidArgument(-1, "CamlPaml.Infer.prepare: length(leaves) != leaves(t)")
        if pms.shape[0] < tree.n - 1:
            raise InvalidArgument(-1, "CamlPaml.Infer.prepare: not enough P matrices")
        P = np.ascontiguousarray(pms[:tree.n - 1], dtype=np.float64)
        pr = np.ascontiguousarray(prior, dtype=np.float64)
        ld = None if leaf_dist is None else np.ascontiguousarray(leaf_dist, dtype=np.float64)
        if ld is not None and ld.shape != (leaves.shape[0], tree.nl, k):
            raise InvalidArgument(-1, "CamlPaml.Infer.prepare: leaf_dist must be [ncols, nl, k]")
        return P, pr, leaves, ld, np.ascontiguousarray(tree.children, dtype=np.int32)

    def prepare(self, tree: Tree, pms: np.ndarray, prior: np.ndarray, leaves: np.ndarray,
                leaf_dist: Optional[np.ndarray] = None):
        k = len(prior)
        P, pr, leaves, ld, ch = self._inputs(tree, pms, prior, leaves, leaf_dist)
        ncols = leaves.shape[0]
        z = np.zeros(ncols)
        post = np.zeros((ncols, k))
        if ld is None:
            self.ctx.check(lib().pcsf_lik_columns(self.ctx.h, k, tree.n, _i32(ch), _dp(P),
                                                  _dp(pr), _i32(leaves), ncols, _dp(z),
                                                  _dp(post)))
        else:
            allpost = np.zeros((ncols, tree.n, k))
            self.ctx.check(lib().pcsf_lik_posteriors(self.ctx.h, k, tree.n, _i32(ch), _dp(P),
                                                     _dp(pr), _i32(leaves), _dp(ld), ncols,
                                                     None, _dp(z), _dp(allpost), None))
            post = allpost[:, -1, :].copy()
        it = _Intermediate(self.ctx, tree, P, pr, leaves, ld, z, post)
        if ld is not None:
            it._all = allpost
        return it


class _Intermediate:
    """The result of PhyloLik.prepare for many columns (one reference
    intermediate per column)."""

    def __init__(self, ctx, tree, P, prior, leaves, leaf_dist, z, root_post):
        self.ctx, self.tree, self.P, self.prior = ctx, tree, P, prior
        self.leaves, self.leaf_dist, self.z, self.root_post = leaves, leaf_dist, z, root_post
        self._all = None

    def _call(self, node_post, ecounts, weights):
        k = len(self.prior)
        ch = np.ascontiguousarray(self.tree.children, dtype=np.int32)
        z = np.zeros(self.leaves.shape[0])
        w = None if weights is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(weights, np.float64), z.shape))
        self.ctx.check(lib().pcsf_lik_posteriors(
            self.ctx.h, k, self.tree.n, _i32(ch), _dp(self.P), _dp(self.prior),
            _i32(self.leaves), None if self.leaf_dist is None else _dp(self.leaf_dist),
            len(z), None if w is None else _dp(w), _dp(z),
            None if node_post is None else _dp(node_post),
            None if ecounts is None else _dp(ecounts)))

    def likelihood(self):
        """PhyloLik.likelihood: z per column."""
        return self.z

    def node_posterior(self, i):
        """PhyloLik.node_posterior (PhyloLik.ml:127-138): [ncols, k]."""
        if not 0 <= i < self.tree.n:
            raise InvalidArgument(-1, "node_posterior: no such node")
        if i == self.tree.n - 1:
            return self.root_post
        if self._all is None:
            self._all = np.zeros((self.leaves.shape[0], self.tree.n, len(self.prior)))
            self._call(self._all, None, None)
        return self._all[:, i, :]

    def add_branch_posteriors(self, ecounts, weight=1.0):
        """PhyloLik.add_branch_posteriors (PhyloLik.ml:140-174) of every branch,
        summed over the columns: ecounts [n-1, k, k] += weight * P(parent = a,
        node = b | leaves); weight a scalar or one per column."""
        if ecounts.shape != (self.tree.n - 1, len(self.prior), len(self.prior)) or \
                ecounts.dtype != np.float64 or not ecounts.flags.c_contiguous:
            raise InvalidArgument(-1, "add_branch_posteriors: ecounts must be a C-contiguous "
                                      "float64 [n-1, k, k] array")
        self._call(None, ecounts, weight)
        return ecounts

    def branch_posteriors(self, br):
        """PhyloLik.branch_posteriors (PhyloLik.mli:31-32), summed over the
        columns: [k, k]."""
        if not 0 <= br < self.tree.n - 1:
            raise InvalidArgument(-1, "CamlPaml.Infer.add_branch_posteriors: invalid branch")
        k = len(self.prior)
        ec = np.zeros((self.tree.n - 1, k, k))
        self.add_branch_posteriors(ec)
        return ec[br]


def collect_sufficient_statistics(ctx: Context, tree: Tree, pms, prior, leaves, ss,
                                  leaf_dist=None, weights=None):
    """PhyloEM.collect_sufficient_statistics (lib/CamlPaml/PhyloEM.ml:14-20) over
    many columns: adds every column's branch posteriors (times its weight) to
    ss [n-1, k, k] and returns the likelihood of each column.  One call of
    pcsf_lik_posteriors (with 64 states and no `Distribution leaves: the FP64
    matrix-core E-step)."""
    P, pr, lv, ld, ch = PhyloLik._inputs(tree, pms, prior, leaves, leaf_dist)
    k = len(pr)
    if ss.shape != (tree.n - 1, k, k) or ss.dtype != np.float64 or not ss.flags.c_contiguous:
        raise InvalidArgument(-1, "collect_sufficient_statistics: ss must be a C-contiguous "
                                  "float64 [n-1, k, k] array")
    z = np.zeros(lv.shape[0])
    w = None if weights is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(weights, np.float64), z.shape))
    ctx.check(lib().pcsf_lik_posteriors(ctx.h, k, tree.n, _i32(ch), _dp(P), _dp(pr), _i32(lv),
                                        None if ld is None else _dp(ld), len(z),
                                        None if w is None else _dp(w), _dp(z), None, _dp(ss)))
    return z
