"""The PhyloCSF command-line program scoring on the GPU.

* the reference's four published CLI results (src/test.ml:27-59: tal-AA,
  ALDH2 exon 5 out of frame and in 6 frames, Aldh2 mRNA ATGStop with the
  translation), with the same command lines, at FP32 (default) and FP64;
* the fixed strategy against the oracle's restatement of the CLI
  (oracle/cli_oracle.py) on every frame of the examples;
* several alignments in one run (--files) give the same lines as one run each.
"""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

from oracle import cli_oracle as CO  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "phylocsf_amd", "bin", "PhyloCSF")


def _say(msg):
    print(msg, flush=True)


def run(*args):
    e = dict(os.environ)
    e.pop("PHYLOCSF_BASE", None)
    p = subprocess.run([EXE] + list(args), capture_output=True, text=True, env=e, timeout=300)
    assert p.returncode == 0, p.stderr
    return [l.split("\t") for l in p.stdout.strip().split("\n")]


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_published_cli_results(examples_dir, precision):
    ex = lambda f: os.path.join(examples_dir, f)  # noqa: E731
    pr = "--precision=" + precision
    _say("tal-AA " + precision)
    a = run("12flies", ex("tal-AA.fa"), "--ancComp", pr)[0]
    assert a[1] == "score(decibans)"
    assert 297.62 <= float(a[2]) <= 297.63 and 48.25 <= float(a[3]) <= 48.26
    _say("ALDH2 exon 5 out of frame " + precision)
    a = run("29mammals", ex("ALDH2.exon5.fa"), "--ancComp", pr)[0]
    assert a[1] == "score(decibans)"
    assert -178.93 <= float(a[2]) <= -178.92 and -38.29 <= float(a[3]) <= -38.28
    _say("ALDH2 exon 5, 6 frames " + precision)
    a = run("29mammals", ex("ALDH2.exon5.fa"), "--frames=6", pr)[0]
    assert a[1] == "max_score(decibans)"
    assert 218.26 <= float(a[2]) <= 218.27 and a[3:6] == ["1", "111", "+"]
    _say("Aldh2 mRNA " + precision)
    a = run("29mammals", ex("Aldh2.mRNA.fa"), "--orf=ATGStop", "--frames=3", "--removeRefGaps",
            "--aa", pr)[0]
    assert a[1] == "max_score(decibans)"
    assert 2013.92 <= float(a[2]) <= 2013.93 and a[3:5] == ["343", "1899"]
    assert re.match("^MLRAALTTVRRGPRLSRLLSAAA", a[5])


def test_fixed_strategy_against_oracle(examples_dir, params_dir):
    for pset, fa in (("12flies", "tal-AA.fa"), ("29mammals", "ALDH2.exon5.fa")):
        _say("fixed, 6 frames: " + fa)
        path = os.path.join(examples_dir, fa)
        lines = run(pset, path, "--strategy=fixed", "--frames=6", "--allScores", "--ancComp")
        results, best = CO.run(os.path.join(params_dir, pset), open(path).read(),
                               strategy="fixed", frames=6)
        assert len(lines) == len(results) + 1
        for f, (s, rc, lo, hi) in zip(lines, results):
            assert f[1] == "orf_score(decibans)"
            assert (f[3], f[4], f[5]) == (str(lo), str(hi), "-" if rc else "+")
            assert abs(float(f[2]) - s["score"]) <= 1e-3 + 5e-5
            assert abs(float(f[6]) - s["anc_comp"]) <= 1e-2 + 5e-5
        s, rc, lo, hi = best
        assert lines[-1][1] == "max_score(decibans)" and lines[-1][3:5] == [str(lo), str(hi)]


def test_batched_files_match_single_runs(tmp_path, examples_dir):
    fas = [os.path.join(examples_dir, f) for f in ("ALDH2.exon5.fa", "Aldh2.mRNA.fa")]
    lst = tmp_path / "list.txt"
    lst.write_text("\n".join(fas + fas[:1]) + "\n")
    args = ["--strategy=fixed", "--frames=3", "--orf=StopStop", "--allScores", "--removeRefGaps"]
    both = run("29mammals", str(lst), "--files", *args)
    single = []
    for f in fas + fas[:1]:
        single += run("29mammals", f, *args)
    assert both == single
