"""CPU-only checks of the product's host layer: the C ABI library loads and
exports every symbol include/phylocsf_hip.h declares, and the host parsers
agree with the oracle's restatement (no GPU calls)."""
import os
import re

import numpy as np
import pytest

import phylocsf_amd as P
from phylocsf_amd import _lib
from oracle import oracle as O

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "phylocsf_hip.h")


def test_library_exports_every_declared_symbol():
    decl = set(re.findall(r"\b(pcsf_[a-z_0-9]+)\s*\(", open(HEADER).read()))
    L = _lib.lib()
    missing = [s for s in sorted(decl) if not hasattr(L, s)]
    assert not missing, missing
    assert decl == set(_lib.EXPORTED)


@pytest.mark.parametrize("name", ["12flies", "29mammals", "58mammals", "100vertebrates",
                                  "120mammals", "7yeast", "20flies"])
def test_newick_matches_oracle(name, params_dir):
    text = open(os.path.join(params_dir, name + ".nh")).read()
    a = P.newick_tree(text)
    b = O.newick_tree(text)
    assert a.n == b.n and a.labels == b.labels
    assert np.array_equal(a.children[:, 0], b.lc[b.nl:]) and np.array_equal(a.children[:, 1],
                                                                             b.rc[b.nl:])
    assert np.array_equal(a.branch[:-1], b.branch[:-1])


def test_newick_subtree_matches_oracle(params_dir):
    text = open(os.path.join(params_dir, "58mammals.nh")).read()
    full = O.newick_tree(text)
    keep = full.labels[0:58:9]  # six species spread over the tree
    a = P.newick_tree(text, keep)
    b = O.newick_tree(text, keep)
    assert a.n == b.n == 2 * len(keep) - 1 and a.labels == b.labels
    assert np.allclose(a.branch[:-1], b.branch[:-1], rtol=0, atol=1e-15)
    with pytest.raises(P.Failure):
        P.newick_tree(text, ["Human"])


def test_newick_errors():
    with pytest.raises(P.InvalidArgument):
        P.newick_tree("(a,b,c);")
    with pytest.raises(P.Failure):
        P.newick_tree("(a,(b,c);")
    with pytest.raises(P.Failure):
        P.newick_tree("(123,b);")


@pytest.mark.parametrize("name", ["12flies", "58mammals", "7yeast", "26worms"])
def test_ecm_matches_oracle(name, params_dir):
    for kind in ("coding", "noncoding"):
        path = os.path.join(params_dir, "%s_%s.ECM" % (name, kind))
        S1, p1 = P.import_parameters(path)
        S2, p2 = O.ecm_read(path)
        assert np.array_equal(S1, S2) and np.array_equal(p1, p2)


def test_ecm_errors(tmp_path, params_dir):
    src = open(os.path.join(params_dir, "12flies_coding.ECM")).read().split("\n")
    bad = list(src)
    bad[68] = bad[68].replace("CCA", "CCC", 1)
    p = tmp_path / "bad.ECM"
    p.write_text("\n".join(bad))
    with pytest.raises(P.Failure):
        P.import_parameters(str(p))
    with pytest.raises(P.Failure):
        P.import_parameters(str(tmp_path / "missing.ECM"))


def test_batch_rejects_codes_outside_the_alphabet():
    """A leaf code is a codon index 0..63 or PCSF_MARGINALIZE (255); the Python
    Batch refuses anything else before the library is called (the library
    checks the same on the device, tests/test_gpu_errors.py)."""
    from types import SimpleNamespace
    model = SimpleNamespace(tree=SimpleNamespace(nl=4))
    off = np.array([0, 3], np.int64)
    for bad in (64, 100, 254):
        codes = np.zeros((3, 4), np.uint8)
        codes[1, 2] = bad
        with pytest.raises(P.InvalidArgument):
            P.Batch(model, codes, off)
