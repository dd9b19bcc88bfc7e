"""The PhyloCSF command-line program (phylocsf_amd/bin/PhyloCSF) on the CPU.

With --strategy=nop (src/PhyloCSF.ml:466: every region scores 0) the program
runs without a GPU, so everything around the scoring is checked here:
candidate regions for every frame / ORF mode against the oracle's
restatement of src/PhyloCSF.ml:133-217, the DNA and translation columns,
BLS against closed forms on synthetic alignments, the abort / failure lines
and exit codes, --files and --species.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import cli_oracle as CO
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "phylocsf_amd", "bin", "PhyloCSF")
MODES = ["AsIs", "ATGStop", "StopStop", "StopStop3", "ToFirstStop", "FromLastStop", "ToOrFromStop"]


def run(*args, stdin=None, env=None):
    if not os.path.exists(EXE):
        pytest.fail("PhyloCSF executable not built (run __graft_entry__.build())")
    e = dict(os.environ)
    e.pop("PHYLOCSF_BASE", None)
    if env:
        e.update(env)
    p = subprocess.run([EXE] + list(args), input=stdin, capture_output=True, text=True, env=e,
                       timeout=120)
    return p.returncode, p.stdout, p.stderr


def regions_of(out):
    rows = []
    for line in out.strip().split("\n"):
        f = line.split("\t")
        if f[1] != "orf_score(decibans)":
            continue
        rc = len(f) > 5 and f[5] == "-"
        rows.append((rc, int(f[3]), int(f[4])))
    return rows


@pytest.mark.parametrize("fa,extra", [("ALDH2.exon5.fa", []), ("tal-AA.fa", []),
                                      ("Aldh2.mRNA.fa", ["--removeRefGaps"])])
def test_candidate_regions_every_mode(examples_dir, fa, extra):
    path = os.path.join(examples_dir, fa)
    _, seqs = CO.input_mfa(open(path).read())
    if extra:
        seqs = CO.remove_ref_gaps(seqs)
    pset = "12flies" if fa == "tal-AA.fa" else "29mammals"
    for frames in (1, 3, 6):
        for mode in MODES:
            if mode == "AsIs" and frames == 1:
                continue  # single region, no lo/hi columns
            rc, out, err = run(pset, path, "--strategy=nop", "--frames=%d" % frames,
                               "--orf=" + mode, "--allScores", "--minCodons=20", *extra)
            want = CO.candidate_regions(seqs[0], frames, mode, 20)
            if not want:
                assert rc == 0 and "\tfailure\tFailure(\"no sufficiently long ORFs found\")" in out
                continue
            assert rc == 0, err
            got = regions_of(out)
            assert got == [(r, lo, hi) for (r, lo, hi) in want], (frames, mode)


def test_dna_and_translation(examples_dir):
    path = os.path.join(examples_dir, "Aldh2.mRNA.fa")
    rc, out, _ = run("29mammals", path, "--strategy=nop", "--orf=ATGStop", "--frames=3",
                     "--removeRefGaps", "--allScores", "--dna", "--aa")
    assert rc == 0
    _, seqs = CO.input_mfa(open(path).read())
    seqs = CO.remove_ref_gaps(seqs)
    lines = [l.split("\t") for l in out.strip().split("\n")]
    assert lines[-1][1] == "max_score(decibans)"
    for f in lines:
        lo, hi = int(f[3]), int(f[4])
        assert f[5] == seqs[0][lo:hi + 1]
        assert f[6] == CO.translate(f[5])
    # the published ORF translates to the published protein prefix
    assert any(f[3:5] == ["343", "1899"] and f[6].startswith("MLRAALTTVRRGPRLSRLLSAAA")
               for f in lines)


def _tree(params_dir, name):
    return O.newick_tree(open(os.path.join(params_dir, name + ".nh")).read())


def _path_length(t, a, b):
    """Patristic distance between leaves a and b of an oracle tree."""
    def anc(x):
        out = {}
        d = 0.0
        while x is not None:
            out[x] = d
            p = t.parent[x]
            if p is not None and p >= 0:
                d += t.branch[x]
                x = p
            else:
                x = None
        return out
    A, B = anc(a), anc(b)
    common = [x for x in A if x in B]
    return min(A[x] + B[x] for x in common)


def test_bls_closed_forms(tmp_path, params_dir):
    t = _tree(params_dir, "12flies")
    labels = t.labels[:t.nl]
    total = float(sum(t.branch[i] for i in range(t.n - 1)))
    n = 12
    seq = "ATG" * (n // 3)
    rng = np.random.default_rng(1)
    # column blocks: all present | reference alone | reference + one other
    rows = {sp: list(seq) for sp in labels}
    other = labels[5]
    for sp in labels:
        for j in range(3, 6):
            if sp != labels[0]:
                rows[sp][j] = "-"
        for j in range(6, 9):
            if sp not in (labels[0], other):
                rows[sp][j] = rng.choice(["-", "N"])
    text = "".join(">%s\n%s\n" % (sp, "".join(rows[sp])) for sp in labels)
    f = tmp_path / "bls.fa"
    f.write_text(text)
    d = _path_length(t, 0, labels.index(other))
    for lo, hi, want in ((0, 2, 1.0), (3, 5, 0.0), (6, 8, d / total),
                         (0, 8, (3 * 1.0 + 3 * 0.0 + 3 * d / total) / 9)):
        # single-frame AsIs regions span the whole alignment: cut it instead
        sub = "".join(">%s\n%s\n" % (sp, "".join(rows[sp][lo:hi + 1])) for sp in labels)
        g = tmp_path / ("bls_%d_%d.fa" % (lo, hi))
        g.write_text(sub)
        rc, out, err = run("12flies", str(g), "--strategy=nop", "--bls")
        assert rc == 0, err
        got = float(out.strip().split("\t")[3])
        assert abs(got - want) <= 5e-5, (lo, hi, got, want)


def test_abort_and_failure_lines(tmp_path, examples_dir):
    good = os.path.join(examples_dir, "ALDH2.exon5.fa")
    gapped = tmp_path / "gapped.fa"
    gapped.write_text(">Human\nAC-GTA\n>Mouse\nACGGTA\n")
    rc, out, _ = run("29mammals", good, str(gapped), "--strategy=nop")
    lines = out.strip().split("\n")
    assert rc == 255
    assert lines[0].startswith(good + "\tscore(decibans)\t0.0000")  # earlier output first
    assert lines[1] == (str(gapped) + "\tabort\tFailure(\"the reference sequence (first "
                        "alignment row) must be ungapped\")")
    unknown = tmp_path / "unknown.fa"
    unknown.write_text(">Human\nACGTAC\n>Zebra\nACGTAC\n>Aardwolf\nACGTAC\n")
    rc, out, _ = run("29mammals", str(unknown), "--strategy=nop")
    assert rc == 255 and out.strip().endswith(
        "abort\tFailure(\"parameters not available for species: Aardwolf Zebra\")")
    badnuc = tmp_path / "badnuc.fa"
    badnuc.write_text(">Human\nACGTAC\n>Mouse\nACXTAC\n")
    rc, out, _ = run("29mammals", str(badnuc), "--strategy=nop")
    assert rc == 255 and out.strip().endswith(
        "abort\tInvalid_argument(\"unrecognized nucleotide X\")")
    badlen = tmp_path / "badlen.fa"
    badlen.write_text(">Human\nACGTAC\n>Mouse\nACGTA\n")
    rc, out, _ = run("29mammals", str(badlen), "--strategy=nop")
    assert rc == 255 and out.strip().endswith(
        "abort\tFailure(\"invalid MFA alignment: empty species name or sequence, or sequence "
        "length mismatch\")")
    rc, out, _ = run("29mammals", str(tmp_path / "missing.fa"), "--strategy=nop")
    assert rc == 255 and "abort\tSys_error(" in out
    short = tmp_path / "short.fa"
    short.write_text(">Human\nATGAAATAA\n>Mouse\nATGAAATAA\n")
    rc, out, _ = run("29mammals", str(short), str(short), "--strategy=nop", "--orf=ATGStop",
                     "--frames=3")
    assert rc == 0
    assert out.strip().split("\n") == [
        str(short) + "\tfailure\tFailure(\"no sufficiently long ORFs found\")"] * 2


def test_files_lists_stdin_and_species(tmp_path, examples_dir):
    a = os.path.join(examples_dir, "ALDH2.exon5.fa")
    lst = tmp_path / "list.txt"
    lst.write_text(a + "\n" + a + "\n")
    rc, out, _ = run("29mammals", str(lst), "--files", "--strategy=nop", "--ancComp")
    assert rc == 0 and out.strip().split("\n") == [a + "\tscore(decibans)\t0.0000\t0.0000"] * 2
    rc, out, _ = run("29mammals", "--strategy=nop", stdin=open(a).read())
    assert rc == 0 and out.strip() == "(STDIN)\tscore(decibans)\t0.0000"
    rc, out, err = run("29mammals", a, "--strategy=nop", "--species=Human")
    assert rc == 2 and "specify at least two available --species" in err
    rc, out, _ = run("29mammals", a, "--strategy=nop", "--species=Human,Mouse,Rat")
    assert rc == 255 and "parameters not available for species" in out
    rc, _, err = run("29mammals", a, "--frames=4")
    assert rc == 2 and "invalid reading frame" in err
    rc, _, err = run("no_such_set", a, "--strategy=nop")
    assert rc == 2 and "could not find required parameter file" in err
