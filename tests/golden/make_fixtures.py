"""Writes the small data fixtures the tests use from the reference tree.

Run in the build container (where the reference is mounted):
    python tests/golden/make_fixtures.py /path/to/PhyloCSF
It copies data only: the three JC69 TinyTest sequences of
lib/CamlPaml/test.ml:59-61 (inputs of the reference's known-answer test).
The parameter sets and example alignments are copied verbatim as data files
(phylocsf_amd/data/PhyloCSF_Parameters, tests/golden/PhyloCSF_Examples).
"""
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
here = os.path.dirname(os.path.abspath(__file__))
txt = open(os.path.join(ref, "lib", "CamlPaml", "test.ml")).read()
seqs = dict(re.findall(r'let (human|chimp|gorilla) = "([ACGT\-]+)"', txt))
with open(os.path.join(here, "tinytest_sequences.txt"), "w") as f:
    for k in ("human", "chimp", "gorilla"):
        f.write("%s %s\n" % (k, seqs[k]))
