"""The fused kernel's tree schedule, checked on the CPU.

pcsf_program compiles a tree into the op program the MFMA kernel runs
(post-order, fewest live partial tiles, cherry tables on the fixed-scale
path).  Interpreting that program in FP64 numpy with the oracle's P(t) must
give the oracle's likelihood for every shipped tree, so a wrong order, a
clobbered slot, a missing multiply or a wrong table index fails here
without a GPU.
"""
import os

import numpy as np
import pytest

import phylocsf_amd as P
from phylocsf_amd.api import program
from oracle import oracle as O

OP_CHERRY, OP_GEMM, OP_LEAFMUL, OP_MULPOP, OP_ROOT, OP_END, OP_CHERRYT = 1, 2, 3, 4, 5, 6, 7


def interpret(ops, tree, Pm, prior, codes, cherries):
    nl = tree.nl

    def leaf(j, c):
        return np.ones(64) if codes[j] == 255 else Pm[j][:, codes[j]].copy()

    slots = {}
    z = None
    for w0, w1, w2, _ in ops:
        op, d = w0 & 0xFF, (w0 >> 8) & 0xFF
        a0, a1 = w1 & 0xFFFF, (w1 >> 16) & 0xFFFF
        if op == OP_CHERRY:
            slots[d] = leaf(a0, codes) * leaf(a1, codes)
        elif op == OP_GEMM:
            slots[d] = Pm[a0] @ slots[d]
        elif op == OP_LEAFMUL:
            slots[d] = slots[d] * leaf(a0, codes)
        elif op == OP_MULPOP:
            slots[d] = slots[d] * slots.pop(d + 1)
        elif op == OP_CHERRYT:
            c = cherries[w2]
            assert (tree.children[c - nl] == [a0, a1]).all()
            slots[d] = Pm[c] @ (leaf(a0, codes) * leaf(a1, codes))
        elif op == OP_ROOT:
            z = prior @ slots[0]
        elif op == OP_END:
            break
    return z


@pytest.mark.parametrize("name", ["7yeast", "12flies", "29mammals", "58mammals",
                                  "100vertebrates", "120mammals", "53birds"])
@pytest.mark.parametrize("tables", [False, True])
def test_program_matches_oracle(name, tables, params_dir):
    text = open(os.path.join(params_dir, name + ".nh")).read()
    tree = P.newick_tree(text)
    ot = O.newick_tree(text)
    S, pi = O.ecm_read(os.path.join(params_dir, name + "_coding.ECM"))
    Q = O.build_q(S, pi)
    Pm = np.stack([O.pt(Q, pi, b) for b in ot.branch[:-1]])
    ops, slots = program(tree, tables)
    assert slots <= 6
    depth = max(((w0 >> 8) & 0xFF) for w0 in ops[:, 0]) + 1
    assert depth == slots
    nl = tree.nl
    cherries = [i for i in range(nl, tree.n - 1)
                if tree.children[i - nl][0] < nl and tree.children[i - nl][1] < nl]
    n_gemm = int(((ops[:, 0] & 0xFF) == OP_GEMM).sum())
    assert n_gemm == (tree.n - 1 - nl) - (len(cherries) if tables else 0)
    rng = np.random.default_rng(len(name))
    codes = rng.integers(0, 64, (4, nl)).astype(np.int64)
    codes[rng.random(codes.shape) < 0.3] = 255
    z_ref, _, _ = O.prune(ot, Pm, pi, np.where(codes == 255, -1, codes))
    for c in range(codes.shape[0]):
        z = interpret(ops, tree, Pm, pi, codes[c], cherries)
        assert abs(z - z_ref[c]) <= 1e-12 * z_ref[c], (name, tables, c)


PF_TABLE = 0x80000000


@pytest.mark.parametrize("name", ["12flies", "58mammals", "100vertebrates", "120mammals"])
@pytest.mark.parametrize("tables", [False, True])
def test_prefetch_words_name_the_next_gather(name, tables, params_dir):
    """A GEMM's word 2 (leaf) or word 3 (table row) names exactly the first
    gather op after it, which carries the use-prefetch bit; no other op
    carries it, and no GEMM at the deepest GEMM level issues one."""
    tree = P.newick_tree(open(os.path.join(params_dir, name + ".nh")).read())
    ops, _ = program(tree, tables)
    ops = ops.astype(np.int64) & 0xFFFFFFFF
    opc = ops[:, 0] & 0xFF
    depth = (ops[:, 0] >> 8) & 0xFF
    deepest = depth[opc == OP_GEMM].max()
    claimed = set()
    n_pf = 0
    for k in np.nonzero(opc == OP_GEMM)[0]:
        leaf, row = ops[k, 2] & 0xFFFF, ops[k, 3]
        if leaf == 0xFFFF and not row & PF_TABLE:
            continue
        assert depth[k] < deepest, (k, depth[k])
        assert leaf == 0xFFFF or not row & PF_TABLE
        j = k + 1
        while opc[j] not in (OP_CHERRY, OP_LEAFMUL, OP_CHERRYT):
            assert opc[j] != OP_GEMM and opc[j] != OP_END, (k, j)
            j += 1
        assert (ops[j, 0] >> 16) & 1, (k, j)
        a0, a1 = ops[j, 1] & 0xFFFF, (ops[j, 1] >> 16) & 0xFFFF
        if leaf != 0xFFFF:
            assert opc[j] in (OP_CHERRY, OP_LEAFMUL) and a0 == leaf
        else:
            assert opc[j] == OP_CHERRYT
            assert (row & 0xFFF, (row >> 12) & 0xFFF, (row >> 24) & 0x7F) == (a0, a1, ops[j, 2])
        claimed.add(j)
        n_pf += 1
    flagged = set(np.nonzero((ops[:, 0] >> 16) & 1)[0])
    assert flagged == claimed
    assert n_pf > 0
    if tables:
        assert any(opc[j] == OP_CHERRYT for j in claimed)
