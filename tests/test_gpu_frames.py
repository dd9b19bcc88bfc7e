"""GPU batch construction from nucleotide alignments (pcsf_batch_create_frames).

Every reading frame of packed nucleotide regions is read on the GPU; the
codes must equal the oracle's restatement of PhyloCSF's candidate regions
and pleaves (src/PhyloCSF.ml:198-246; oracle/cli_oracle.py) for frames 1, 3
and 6, including regions whose length is not a multiple of 3, shorter than
a codon, or empty.  The codes are read back from the device and compared
byte for byte, and the FP64 forwards of both batches agree bitwise.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import phylocsf_amd as P  # noqa: E402
from oracle import cli_oracle as CO  # noqa: E402


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    return P.Context(0)


def _host_frames(rows, frames):
    """codes of every frame of one region (rows: tree-leaf-ordered strings)."""
    L = len(rows[0])
    rc_rows = [CO.revcomp(r) for r in rows]
    ident = list(range(len(rows)))
    out = []
    for f in range(frames):
        src = rows if f < 3 else rc_rows
        out.append(CO.pleaves(src, ident, f % 3, L - 1))
    return out


def test_frames_match_oracle_pleaves(ctx):
    m = P.PhyloCSFModel.from_paramset(ctx, "29mammals")
    nl = m.tree.nl
    rng = np.random.default_rng(17)
    codons = np.array([30, 1, 0, 7, 22, 3])
    nuc, noff = m.simulate_nucleotides(codons, rng.integers(0, 2, len(codons)),
                                       rng.uniform(0.4, 2.5, len(codons)), gap_p=0.2, n_p=0.05,
                                       seed=23)
    # trim regions to lengths that are not multiples of 3 (and shorter than a codon)
    trims = [0, 1, 0, 2, 1, 2]
    rows_per_region, pieces = [], []
    for r, t in enumerate(trims):
        lo, hi = noff[r], noff[r + 1] - min(t, noff[r + 1] - noff[r])
        block = nuc[:, lo:hi]
        pieces.append(block)
        rows_per_region.append(["".join(map(chr, block[l])) for l in range(nl)])
    # a lowercase region and one with a character outside ACGTN-
    rows_per_region[0] = [s.lower() for s in rows_per_region[0]]
    pieces[0] = np.array([[ord(c) for c in s] for s in rows_per_region[0]], np.uint8)
    rows_per_region[4][3] = "N" * len(rows_per_region[4][3])
    pieces[4][3, :] = ord("N")
    lens = [p.shape[1] for p in pieces]
    off = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    packed = np.concatenate(pieces, axis=1)
    for frames in (1, 3, 6):
        _say("frames %d" % frames)
        bg = P.Batch.frames(m, packed, off, frames)
        regions = [c for rows in rows_per_region for c in _host_frames(rows, frames)]
        bh = P.Batch.from_regions(m, regions)
        assert bg.n_regions == bh.n_regions == len(regions)
        assert bg.n_cols == bh.n_cols
        # the GPU-built codes, byte for byte
        codes, coff = bg.codes()
        assert np.array_equal(codes, np.concatenate(regions, axis=0))
        assert np.array_equal(coff, np.cumsum([0] + [r.shape[0] for r in regions]))
        lg, eg = bg.forward(1.0, P.PCSF_FP64)
        lh, eh = bh.forward(1.0, P.PCSF_FP64)
        assert np.array_equal(lg, lh) and np.array_equal(eg, eh)


def test_frame_one_equals_simulated_codes(ctx):
    m = P.PhyloCSFModel.from_paramset(ctx, "58mammals")
    rng = np.random.default_rng(3)
    lengths = rng.integers(1, 90, 12)
    kw = dict(gap_mode=1, gap_p=0.25, gap_run=5.0, n_p=0.01, seed=99)
    rm, rr = rng.integers(0, 2, 12), rng.uniform(0.3, 3.0, 12)
    codes, coff = m.simulate(lengths, rm, rr, **kw)
    nuc, noff = m.simulate_nucleotides(lengths, rm, rr, **kw)
    assert nuc.shape == (m.tree.nl, 3 * codes.shape[0])
    bg = P.Batch.frames(m, nuc, noff, 1)
    bh = P.Batch(m, codes, coff)
    assert np.array_equal(bg.codes()[0], codes) and np.array_equal(bh.codes()[0], codes)
    assert np.array_equal(bg.forward(1.0, P.PCSF_FP64)[0], bh.forward(1.0, P.PCSF_FP64)[0])
    # the reference row is never masked
    assert set(np.unique(nuc[0])) <= set(b"ACGT")


def test_frames_bad_arguments(ctx):
    m = P.PhyloCSFModel.from_paramset(ctx, "12flies")
    nuc = np.full((m.tree.nl, 9), ord("A"), np.uint8)
    for frames, off in ((2, [0, 9]), (3, [0, 8]), (3, [1, 9])):
        with pytest.raises(P.InvalidArgument):
            P.Batch.frames(m, nuc, np.array(off), frames)
