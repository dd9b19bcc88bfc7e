"""The PhyloCSF command-line program scoring on the GPU.

* the reference's four published CLI results (src/test.ml:27-59: tal-AA,
  ALDH2 exon 5 out of frame and in 6 frames, Aldh2 mRNA ATGStop with the
  translation), with the same command lines, at FP32, FP64 (the plain
  kernel) and FP64M (FP64 on matrix cores, the default);
* the fixed strategy (default precision) against the oracle's restatement of
  the CLI (oracle/cli_oracle.py) on every frame of the examples;
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


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp64m"])
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


def test_omega_strategy_against_oracle(examples_dir, params_dir, tmp_path):
    """--strategy=omega (src/PhyloCSF.ml:452-463, src/OmegaModel.ml:195-219)
    against the oracle's restatement: scores as printed, the max region, and
    the diagnostics; H1 from a parameter set's _omega file."""
    import shutil
    cases = [("12flies", "tal-AA.fa", []), ("29mammals", "ALDH2.exon5.fa", ["--frames=3"])]
    for pset, fa, extra in cases:
        _say("omega strategy, " + fa)
        out = run(pset, os.path.join(examples_dir, fa), "--strategy=omega", "--debug",
                  "--precision=fp64", "--ancComp", *extra)[0]
        frames = 3 if extra else 1
        res, best = CO.run(os.path.join(params_dir, pset), open(os.path.join(examples_dir, fa)).read(),
                           strategy="omega", frames=frames)
        assert abs(float(out[2]) - best[0]["score"]) <= 6e-5, (out, best)
        if frames > 1:
            assert out[3:5] == [str(best[2]), str(best[3])]
        assert out[-2] == "nan"  # anc_comp_score is nan for this test
        diag = dict(kv.split("=") for kv in out[-1].lstrip("# ").split(" "))
        assert diag["omega_H1"] == "0.20" and diag["sigma_H1"] == "0.01"
        assert diag["rho_H0"] == "%.2f" % best[0]["rho_H0"]
        assert diag["kappa_H1"] == "%.2f" % best[0]["kappa_H1"]
    # a parameter set with an _omega file
    for ext in (".nh",):
        shutil.copy(os.path.join(params_dir, "12flies" + ext), tmp_path / ("x" + ext))
    (tmp_path / "x_omega").write_text("# H1\nomega_H1\t0.5\nsigma_H1\t0.3\n")
    prefix = str(tmp_path / "x")
    out = run(prefix, os.path.join(examples_dir, "tal-AA.fa"), "--strategy=omega",
              "--precision=fp64", "--debug")[0]
    _, best = CO.run(prefix, open(os.path.join(examples_dir, "tal-AA.fa")).read(),
                     strategy="omega")
    assert abs(float(out[2]) - best[0]["score"]) <= 6e-5
    assert "omega_H1=0.50" in out[-1] and "sigma_H1=0.30" in out[-1]
    (tmp_path / "x_omega").write_text("omega_H1\t0.5\n")
    e = dict(os.environ)
    p = subprocess.run([EXE, prefix, os.path.join(examples_dir, "tal-AA.fa"), "--strategy=omega"],
                       capture_output=True, text=True, env=e, timeout=120)
    assert p.returncode != 0 and "does not specify valid omega_H1 and sigma_H1" in p.stderr
