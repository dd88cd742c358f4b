"""gic_gaussian_stream_noalloc: the Gaussian stream drawn into a caller's
device array with the library's fixed per-GPU workspace, against
gic_gaussian_stream (the allocating entry point, itself checked against the
serial oracle) -- bit-identical, on two calls in a row with the second
smaller, with no device memory taken by either call."""
import numpy as np
import pytest

from genetic_amd import gic

pytestmark = pytest.mark.gpu


def test_noalloc_stream_matches_allocating_entry_point():
    import torch
    print("gaussian_stream_noalloc", flush=True)
    ic = gic.ICGenerator(precision=gic.FLOAT64)
    out = torch.full((1 << 20,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    # a first call (code objects, tables) and a first large copy back (the
    # runtime's staging for device-to-host copies); then the device memory
    # must not change across the entry point's calls
    ic.gaussian_stream_noalloc(8896131, out.data_ptr(), 1 << 20)
    out.cpu()
    torch.cuda.synchronize()
    held = ic.device_bytes()
    got = {}
    for seed, n in ((8896131, 1 << 20), (42, 300_007)):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        ic.gaussian_stream_noalloc(seed, out.data_ptr(), n)
        free1, _ = torch.cuda.mem_get_info()
        assert free1 == free0, "device memory changed by %d bytes" % (free0 - free1)
        assert ic.device_bytes() == held
        got[(seed, n)] = out[:n].cpu().numpy()
        assert np.isnan(out[n:].cpu().numpy()).all()   # nothing written past n
    # the allocating entry point (after the memory checks) on the same seeds
    for (seed, n), g in got.items():
        assert np.array_equal(g, ic.gaussian_stream(seed, n)), (seed, n)
    # beyond the fixed workspace: a clear error and the array left as it was
    before = out.cpu().numpy()
    with pytest.raises(gic.GICError, match="at most 1048576"):
        ic.gaussian_stream_noalloc(7, out.data_ptr(), (1 << 20) + 1)
    assert np.array_equal(out.cpu().numpy(), before, equal_nan=True)
    ic.close()


def test_noalloc_stream_from_a_second_context():
    """Two contexts on one GPU share the device's workspace; calls in turn
    give each its own stream."""
    import torch
    print("gaussian_stream_noalloc two contexts", flush=True)
    a = gic.ICGenerator(precision=gic.FLOAT32)
    b = gic.ICGenerator(precision=gic.FLOAT64)
    out = torch.empty((5000,), dtype=torch.float64, device="cuda")
    a.gaussian_stream_noalloc(5, out.data_ptr(), 5000)
    x = out.cpu().numpy()
    b.gaussian_stream_noalloc(6, out.data_ptr(), 4000)
    y = out[:4000].cpu().numpy()
    assert np.array_equal(x, b.gaussian_stream(5, 5000))
    assert np.array_equal(y, a.gaussian_stream(6, 4000))
    a.close()
    b.close()
