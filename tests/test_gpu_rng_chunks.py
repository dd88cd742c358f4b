"""The chunked RNG at the sizes the workloads draw, against the serial
stream.

The device splits a draw into chunks of kc ranlxs2 blocks (kc = 4, 8, 16,
32, 64 by draw size) and starts each chunk from a 576-bit modular jump;
draws of about 4.6e7 Gaussians and more (every level of 256^3 and above
once combined, the 1024^3 workload at 1.1e9) use kc = 64 with chunk
indices up to 2^20.  Windows far into a stream are checked against the
oracle's serial generator continued from a state found by an independent
big-integer jump (tests/rng_jump.py, pinned against the serial stream in
tests/test_cpu_oracle.py)."""
import numpy as np
import pytest

import pyoracle as po
import rng_jump as rj
from genetic_amd import gic

pytestmark = pytest.mark.gpu

# raw numbers per Gaussian of gsl_ran_gaussian_ziggurat on ranlxs2 (measured
# over 4e7 oracle draws: 2.05333, 2.05344); only a hint for where a window
# sits, the comparison itself locates it by value
RHO = 2.05338


def _kc_of_gaussians(n):
    """The chunk length the device plan picks for n Gaussians (blocks per
    chunk; the plan halves 64 until it has about 65536 chunks)."""
    need = 2.2 * n + 4096.0
    kc = 64
    while kc > 4 and need / (kc * 24.0) < 65536.0:
        kc //= 2
    return kc


def _check_raw_window(got, seed, start, count):
    k0 = start // 24
    words, carry = rj.state_at_block(seed, k0)
    want = po.ranlxs_raw_from_state(words, carry, count + 24)
    off = start - 24 * k0
    seg = got[start:start + count]
    assert np.array_equal(seg, want[off:off + len(seg)]), "raw window at %d" % start


def _check_gaussian_window(got, seed, g, count=20000, margin=60000):
    """The stream values around index g equal the serial stream's, located
    by value within [g - 3 margin, g + 2 margin)."""
    vals = rj.gaussians_near(po, seed, g, RHO, count, margin=margin)
    lo = max(0, g - 3 * margin)
    hi = min(len(got), g + 2 * margin)
    i = rj.locate(got[lo:hi], vals)
    assert i >= 0, "window near %d: serial values not found" % g
    s = lo + i
    n = min(count, len(got) - s)
    np.testing.assert_allclose(got[s:s + n], vals[:n], rtol=0, atol=1e-13, err_msg="window near %d" % g)
    return s, s + n


@pytest.mark.parametrize("n", [15_000_000, 30_000_000, 60_000_000, 120_000_000, 480_000_000])
def test_raw_stream_chunk_lengths(n):
    """k_rng_raw at every chunk length (the plan for n raw numbers is that of
    n/2 + 1 Gaussians: kc 8, 16, 32, 64, 64); the first 1e6 numbers against
    the serial stream, then windows across chunk boundaries up to the end
    (chunk indices to 2^18 at the largest n)."""
    seed = 20261203 + n % 997
    kc = _kc_of_gaussians(n // 2 + 1)
    print("raw stream n=%d kc=%d" % (n, kc), flush=True)
    ic = gic.ICGenerator(precision=gic.FLOAT64)
    got = ic.raw_stream(seed, n)
    ic.close()
    assert np.array_equal(got[:1_000_000], po.ranlxs_raw(seed, 1_000_000))
    ch = 24 * kc
    nch = n // ch
    rng = np.random.default_rng(n)
    starts = [ch * c - 37 for c in (1, 2, 3, nch // 2, nch - 1)] + [int(x) for x in rng.integers(0, n - 5000, 8)]
    starts.append(n - 3000)
    for s in starts:
        _check_raw_window(got, seed, max(0, s), 3000)


@pytest.mark.parametrize("n", [7_000_000, 15_000_000, 30_000_000, 50_000_000])
def test_gaussian_stream_chunk_lengths(n):
    """The production parse (summary, scan, emit) at kc 8, 16, 32, 64: the
    first 1e6 Gaussians exactly as the serial stream, then windows spread to
    the end."""
    seed = 977 + n % 1013
    kc = _kc_of_gaussians(n)
    assert kc == {7_000_000: 8, 15_000_000: 16, 30_000_000: 32, 50_000_000: 64}[n]
    print("gaussian stream n=%d kc=%d" % (n, kc), flush=True)
    ic = gic.ICGenerator(precision=gic.FLOAT64)
    got = ic.gaussian_stream(seed, n)
    ic.close()
    want, _ = po.gaussians(seed, 1_000_000)
    np.testing.assert_allclose(got[:1_000_000], want, rtol=0, atol=1e-13)
    for g in (n // 3, (2 * n) // 3, n - 20000):
        _check_gaussian_window(got, seed, g)


def test_gaussian_stream_workload_size():
    """The whole draw of the 1024^3 + 2 x 256^3 workload (1,107,296,256
    Gaussians, kc = 64, chunk indices to 2^20.6): the start, both level
    boundaries and the end equal the serial stream."""
    seed = 8896131
    n0, n1 = 1024 ** 3, 256 ** 3
    n = n0 + 2 * n1
    print("gaussian stream n=%d" % n, flush=True)
    ic = gic.ICGenerator(precision=gic.FLOAT64)
    got = ic.gaussian_stream(seed, n)
    ic.close()
    want, _ = po.gaussians(seed, 1_000_000)
    np.testing.assert_allclose(got[:1_000_000], want, rtol=0, atol=1e-13)
    for g in (n0, n0 + n1):
        lo, hi = _check_gaussian_window(got, seed, g - 10000, count=150000)
        assert lo <= g < hi, "the level boundary %d lies outside the checked window [%d, %d)" % (g, lo, hi)
    _check_gaussian_window(got, seed, n - 20000)
    # the whole stream is N(0, 1): mean and variance to 6 sigma
    m = float(np.mean(got))
    v = float(np.var(got))
    assert abs(m) < 6 / np.sqrt(n)
    assert abs(v - 1) < 6 * np.sqrt(2.0 / n)
