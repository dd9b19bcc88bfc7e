"""pcsf_forward_noalloc: the forward over device-resident codes that allocates
nothing (its per-column work lives in a fixed per-GPU device array).

Called through the C ABI with device buffers of the test's own, twice in a row
-- the second call with a smaller input and other tree scales and precision,
over the first call's leftovers in the shared array -- and held bit for bit
to pcsf_forward on a batch of the same codes (the same kernels in the same
launch shapes); inputs beyond the array's size, bad codes and bad offsets are
refused before any pruning kernel runs.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import phylocsf_amd as P  # noqa: E402
from hip_runtime import DeviceBuffer, runtime  # noqa: E402
from phylocsf_amd import _lib  # noqa: E402


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    return P.Context(0)


def _noalloc(ctx, m, codes, off, rho, precision):
    """Inputs and outputs in device buffers of the test's own (hipMalloc
    through the library's runtime); one pcsf_forward_noalloc call."""
    hip = runtime()
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    off = np.ascontiguousarray(off, dtype=np.int64)
    R = len(off) - 1
    bufs = [DeviceBuffer(hip, codes.nbytes), DeviceBuffer(hip, off.nbytes),
            DeviceBuffer(hip, max(R, 1) * m.nm * 8), DeviceBuffer(hip, max(R, 1) * m.nm * 8)]
    try:
        d_codes, d_off, d_ll, d_el = bufs
        d_codes.upload(codes)
        d_off.upload(off)
        nan = np.full((max(R, 1), m.nm), np.nan)
        d_ll.upload(nan)
        d_el.upload(nan)
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, np.float64), (m.nm,)))
        rc = _lib.lib().pcsf_forward_noalloc(
            m.h, d_codes.ptr, codes.shape[0], d_off.ptr, R,
            r.ctypes.data_as(C.POINTER(C.c_double)), precision, d_ll.ptr, d_el.ptr)
        ll = d_ll.download(np.zeros((max(R, 1), m.nm)))[:R]
        el = d_el.download(np.zeros((max(R, 1), m.nm)))[:R]
        return rc, ll, el
    finally:
        for b in bufs:
            b.close()


def test_noalloc_forward_matches_pcsf_forward(ctx):
    m = P.PhyloCSFModel.from_paramset(ctx, "58mammals")
    rng = np.random.default_rng(7)
    big = rng.integers(50, 2000, 200)
    small = np.array([5, 0, 130, 1, 64, 77])
    for lengths, rho, prec in ((big, 1.0, P.PCSF_FP16S), (small, 1.7, P.PCSF_FP32),
                               (small, 0.4, P.PCSF_FP64)):
        _say("noalloc: %d regions, %d columns, rho %.2f, precision %d"
             % (len(lengths), lengths.sum(), rho, prec))
        codes, off = m.simulate(lengths, rng.integers(0, 2, len(lengths)), np.ones(len(lengths)),
                                gap_p=0.1, n_p=0.01, seed=int(lengths.sum()))
        rc, ll, el = _noalloc(ctx, m, codes, off, rho, prec)
        assert rc == 0, _lib.lib().pcsf_last_error(ctx.h)
        want_ll, want_el = P.Batch(m, codes, off).forward(rho, prec)
        assert np.array_equal(ll, want_ll) and np.array_equal(el, want_el), prec


def test_noalloc_refusals(ctx):
    m = P.PhyloCSFModel.from_paramset(ctx, "12flies")
    nl = m.tree.nl
    # more columns than the scratch holds
    n = _lib.PCSF_NOALLOC_SCRATCH_BYTES // (nl + 16 * m.nm) + 1
    codes = np.zeros((n, nl), np.uint8)
    _say("noalloc: %d columns (over the limit)" % n)
    rc, _, _ = _noalloc(ctx, m, codes, np.array([0, n]), 1.0, P.PCSF_FP32)
    assert rc == _lib.PCSF_EINVAL
    assert b"exceed" in _lib.lib().pcsf_last_error(ctx.h)
    # a bad code, bad offsets
    codes = np.zeros((40, nl), np.uint8)
    codes[17, 3] = 70
    rc, _, _ = _noalloc(ctx, m, codes, np.array([0, 40]), 1.0, P.PCSF_FP32)
    assert rc == _lib.PCSF_EINVAL and b"leaf code" in _lib.lib().pcsf_last_error(ctx.h)
    codes[17, 3] = 5
    for off in ([0, 41], [1, 40], [0, 30, 20, 40]):
        rc, _, _ = _noalloc(ctx, m, codes, np.array(off), 1.0, P.PCSF_FP32)
        assert rc == _lib.PCSF_EINVAL and b"offsets" in _lib.lib().pcsf_last_error(ctx.h), off
    # and a good call after the refusals
    rc, ll, _ = _noalloc(ctx, m, codes, np.array([0, 10, 40]), 1.0, P.PCSF_FP32)
    assert rc == 0
    assert np.array_equal(ll, P.Batch(m, codes, np.array([0, 10, 40])).forward(1.0, P.PCSF_FP32)[0])
