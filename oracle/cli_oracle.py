"""Oracle restatement of the PhyloCSF CLI's alignment handling (pure Python).

TEST INFRASTRUCTURE ONLY (checker for the product's C++ CLI/host code).
Restates src/PhyloCSF.ml (paths relative to the mlin/PhyloCSF tree):

  input_mfa ............... src/PhyloCSF.ml:83-118
  maybe_remove_ref_gaps ... src/PhyloCSF.ml:122-132
  find_orfs ............... src/PhyloCSF.ml:134-196
  candidate_regions ....... src/PhyloCSF.ml:198-217
  pleaves ................. src/PhyloCSF.ml:219-246
  bls ..................... src/PhyloCSF.ml:252-262
  process_alignment ....... src/PhyloCSF.ml:280-389 (region loop, max, output)
  DNA.comp / revcomp ...... lib/CamlPaml/Code.ml:39-58
  translate ............... src/PhyloCSF.ml:268-278, Code.ml:159-181
"""
from __future__ import annotations

import os

import numpy as np

from . import oracle as O

_COMP = {'A': 'T', 'G': 'C', 'C': 'G', 'T': 'A', 'a': 't', 'g': 'c', 'c': 'g', 't': 'a',
         'N': 'N', 'n': 'n', '-': '-'}
_IDX = {'A': 0, 'C': 1, 'G': 2, 'T': 3, 'a': 0, 'c': 1, 'g': 2, 't': 3}
TRANSLATION = ("KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF")


class CliError(Exception):
    """abort / failure lines of process_alignment."""


def revcomp(s):
    out = []
    for ch in reversed(s):
        if ch not in _COMP:
            raise CliError("Invalid_argument(\"unrecognized nucleotide %s\")" % ch)
        out.append(_COMP[ch])
    return "".join(out)


def input_mfa(text):
    """-> (species list, sequences list)."""
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines = lines[:-1]
    i = 0
    recs = []
    # hdr: skip empty lines, then require '>'
    while i < len(lines) and len(lines[i]) == 0:
        i += 1
    if i >= len(lines):
        raise CliError("invalid MFA alignment: No_more_elements")
    while i < len(lines):
        line = lines[i]
        if len(line) == 0:
            i += 1
            continue
        if line[0] != '>':
            raise CliError("invalid MFA alignment: bad header")
        hdr = line[1:]
        sp = (hdr[:hdr.index('|')] if '|' in hdr else hdr).strip()
        i += 1
        buf = []
        while i < len(lines):
            ln = lines[i].strip()
            if len(ln) == 0:
                i += 1
                continue
            if ln[0] != '>':
                buf.append(ln)
                i += 1
                continue
            break
        recs.append((sp, "".join(buf)))
    species = [r[0] for r in recs]
    seqs = [r[1] for r in recs]
    L = len(seqs[0])
    if L == 0 or any(s == "" for s in species) or any(len(s) != L for s in seqs):
        raise CliError("invalid MFA alignment: empty species name or sequence, or sequence "
                       "length mismatch")
    return species, seqs


def remove_ref_gaps(seqs):
    ref = seqs[0]
    keep = [j for j in range(len(ref)) if ref[j] != '-']
    return ["".join(s[j] for j in keep) for s in seqs]


def find_orfs(dna, orf_mode, min_codons, ofs=0):
    atg = orf_mode == "ATGStop"

    def codon(pos):
        return dna[pos].upper() + dna[pos + 1].upper() + dna[pos + 2].upper()

    def is_start(pos):
        return codon(pos) == "ATG"

    def is_stop(pos):
        return codon(pos) in ("TAA", "TAG", "TGA")

    n = len(dna)
    orfs = []      # OCaml list: newest first
    starts = []
    for codon_lo in range(ofs, n - 2):
        if (codon_lo - ofs) % 3 == 0:
            codon_hi = codon_lo + 2
            if ((not atg) and starts == [] and not is_stop(codon_lo)) or (atg and is_start(codon_lo)):
                starts.insert(0, codon_lo)
            if codon_hi + 3 < n and is_stop(codon_hi + 1):
                for start in starts:
                    if codon_hi > start + 2:
                        orfs.insert(0, (start, codon_hi))
                starts = []
    if not atg:
        for start in starts:
            rem = n - start
            orfs.insert(0, (start, start + (rem // 3) * 3 - 1))
    if orf_mode == "StopStop3":
        allsub = []
        for (lo, hi) in orfs:
            sub = [(lo, hi)]
            codons = (hi - lo + 1) // 3
            lo2 = lo + (codons // 3) * 3
            if lo2 > lo:
                sub.insert(0, (lo2, hi))
            lo3 = lo + (2 * codons // 3) * 3
            if lo3 > lo2 and lo3 > lo:
                sub.insert(0, (lo3, hi))
            allsub.append(sub)
        orfs = [x for sub in allsub for x in sub]
    if orf_mode == "ToFirstStop" and orfs:
        first = orfs[-1]
        orfs = [first] if first[0] == ofs else []
    if orf_mode == "FromLastStop" and orfs:
        last = orfs[0]
        orfs = [last] if n - last[1] <= 3 else []
    if orf_mode == "ToOrFromStop" and orfs:
        first = orfs[-1]
        last = orfs[0]
        orfs = [first] if first[0] == ofs else []
        if n - last[1] <= 3 and first != last:
            orfs = [last] + orfs
    return [(lo, hi) for (lo, hi) in reversed(orfs) if (hi - lo + 1) // 3 >= min_codons]


def candidate_regions(dna, frames, orf_mode, min_codons):
    if orf_mode == "AsIs":
        hi = len(dna) - 1
        r = [(False, 0, hi)]
        if frames != 1:
            r += [(False, 1, hi), (False, 2, hi)]
        if frames == 6:
            r += [(True, 0, hi), (True, 1, hi), (True, 2, hi)]
        return r
    groups = [[(False, lo, hi) for lo, hi in find_orfs(dna, orf_mode, min_codons, 0)]]
    if frames != 1:
        groups.insert(0, [(False, lo, hi) for lo, hi in find_orfs(dna, orf_mode, min_codons, 1)])
        groups.insert(0, [(False, lo, hi) for lo, hi in find_orfs(dna, orf_mode, min_codons, 2)])
    if frames == 6:
        rc = revcomp(dna)
        for o in (0, 1, 2):
            groups.insert(0, [(True, lo, hi) for lo, hi in find_orfs(rc, orf_mode, min_codons, o)])
    return [x for g in reversed(groups) for x in g]


def pleaves(seqs, leaf_ord, lo, hi):
    """codes [ncols, nl] uint8 (255 = `Marginalize)."""
    cols = []
    pos = lo
    while pos + 2 <= hi:
        row = []
        for r in leaf_ord:
            if r is None:
                row.append(255)
                continue
            s = seqs[r]
            a, b, c = s[pos], s[pos + 1], s[pos + 2]
            if a in _IDX and b in _IDX and c in _IDX:
                row.append(16 * _IDX[a] + 4 * _IDX[b] + _IDX[c])
            else:
                row.append(255)
        cols.append(row)
        pos += 3
    return np.array(cols, dtype=np.uint8).reshape(len(cols), len(leaf_ord))


def translate(dna):
    out = []
    for i in range(len(dna) // 3):
        c = dna[3 * i:3 * i + 3]
        if all(ch in _IDX for ch in c):
            out.append(TRANSLATION[16 * _IDX[c[0]] + 4 * _IDX[c[1]] + _IDX[c[2]]])
        else:
            out.append('?')
    return "".join(out)


def omega_config(param_prefix):
    """initialize_strategy's OmegaTest (src/PhyloCSF.ml:452-463): (omega_H1,
    sigma_H1) from the optional key<TAB>value file <prefix>_omega."""
    fn = param_prefix + "_omega"
    if not os.path.exists(fn):
        return 0.2, 0.01
    kv = {}
    for line in open(fn).read().split("\n"):
        line = line.strip()
        if not line or line[0] == "#":
            continue
        k, _, v = line.partition("\t")
        kv[k] = v
    try:
        return float(kv["omega_H1"]), float(kv["sigma_H1"])
    except (KeyError, ValueError):
        raise CliError("configuration file %s exists but does not specify valid omega_H1 and "
                       "sigma_H1 values" % fn)


def run(param_prefix, mfa_text, strategy="mle", frames=1, orf="AsIs", min_codons=25,
        remove_ref_gaps_=False, species=None):
    """process_alignment for one MFA; returns list of (score dict, rc, lo, hi) and the max."""
    nh = open(param_prefix + ".nh").read()
    tree = O.newick_tree(nh, keep=species)
    if strategy == "omega":
        h1 = omega_config(param_prefix)
    else:
        Sc, pic = O.ecm_read(param_prefix + "_coding.ECM")
        Sn, pin = O.ecm_read(param_prefix + "_noncoding.ECM")
        model = O.Model(tree, [Sc, Sn], [pic, pin])
    sp, seqs = input_mfa(mfa_text)
    if remove_ref_gaps_:
        seqs = remove_ref_gaps(seqs)
    seqs = [s.replace('u', 't').replace('U', 'T') for s in seqs]
    if '-' in seqs[0]:
        raise CliError("the reference sequence (first alignment row) must be ungapped")
    rc_seqs = [revcomp(s) for s in seqs]
    tl = set(tree.labels[:tree.nl])
    extra = sorted(set(sp) - tl)
    if extra:
        raise CliError("parameters not available for species: " + " ".join(extra))
    which = {}
    for i, s in enumerate(sp):
        which[s] = i
    leaf_ord = [which.get(tree.labels[i]) for i in range(tree.nl)]
    rgns = candidate_regions(seqs[0], frames, orf, min_codons)
    if not rgns:
        raise CliError("Failure(\"no sufficiently long ORFs found\")")
    results = []
    for (rc, lo, hi) in rgns:
        codes = pleaves(rc_seqs if rc else seqs, leaf_ord, lo, hi)
        if strategy == "mle":
            s = model.score_mle(codes)
        elif strategy == "omega":
            s = O.omega_score(tree, codes, *h1)
            s["anc_comp"] = float("nan")
        else:
            s = model.score_fixed(codes)
        results.append((s, rc, lo, hi))
    # List.reduce max: polymorphic compare on (record, rc, lo, hi); the record
    # compares score first, then anc_comp_score, then diagnostics.
    best = results[0]
    for r in results[1:]:
        if _key(r) > _key(best):
            best = r
    return results, best


def _key(r):
    s, rc, lo, hi = r
    # OCaml's compare takes nan as equal to nan (the omega test's anc_comp_score)
    a = s["anc_comp"]
    return (s["score"], 0.0 if a != a else a, rc, lo, hi)
