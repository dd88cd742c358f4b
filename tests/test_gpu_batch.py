"""Batched constrained realisations (gic_batch_constrained) against the
one-at-a-time path of the same context API, which the other GPU tests hold
to the oracle and the reference fixtures: reseed, draw, select_sphere +
modify overdensity (absolute) per constraint, apply_modifications,
apply_power_spectrum, zeldovich (ic.hpp:1458-1500,1613-1648).  Tolerances:
FP64 1e-9 and FP32 2e-4 of each field's largest value; repeated batches,
with or without a graph and on any number of chains, are bit-identical."""
import numpy as np
import pytest

from genetic_amd import gic
from helpers import cosmology_from_paramfile, golden, select_sphere_cells

pytestmark = pytest.mark.gpu

BOX = 50.0


def _ctx(precision, n):
    ic = gic.ICGenerator(precision=precision)
    ic.set_cosmology(**cosmology_from_paramfile(golden("test_11_import", "paramfile.txt")))
    ic.set_camb(golden("camb_transfer_kmax40_z0.dat"))
    ic.init_base_grid(BOX, n)
    return ic


def _constraints(nreal, ncons, seed):
    rng = np.random.default_rng(seed)
    c = np.empty((nreal, ncons, 5))
    c[..., :3] = rng.uniform(0, BOX, (nreal, ncons, 3))
    c[..., 3] = rng.uniform(1.5, 4.0, (nreal, ncons))
    c[..., 4] = rng.normal(0, 1.5, (nreal, ncons))
    return c


def _one_at_a_time(ic, n, seed, cons):
    ic.reseed(int(seed))
    ic.draw()
    ic.clear_modifications()
    for x, y, z, r, t in cons:
        ic.set_flags(0, select_sphere_cells(n, BOX, (x, y, z), r))
        ic.modify("overdensity", False, float(t))
    ic.apply_modifications()
    ic.apply_power_spectrum()
    ic.zeldovich()
    return [ic.field(f, 0) for f in (gic.FIELD_OVERDENSITY, gic.FIELD_PSI_X, gic.FIELD_PSI_Y, gic.FIELD_PSI_Z)]


def _batch(ic, n, seeds, cons, dtype, lanes, graph):
    import torch
    out = torch.full((len(seeds), 4, n, n, n), float("nan"), dtype=dtype, device="cuda")
    torch.cuda.synchronize()
    ms = ic.batch_constrained(seeds, cons, out.data_ptr(), lanes=lanes, graph=graph)
    assert ms > 0
    return out.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("precision,tol", [(gic.FLOAT64, 1e-9), (gic.FLOAT32, 2e-4)])
def test_batch_matches_one_at_a_time(precision, tol):
    import torch
    n, seeds = 64, [11, 12, 13]
    print("batch vs sequential", precision, n, seeds, flush=True)
    cons = _constraints(len(seeds), 5, 7)
    dtype = torch.float64 if precision == gic.FLOAT64 else torch.float32
    ic = _ctx(precision, n)
    got = _batch(ic, n, seeds, cons, dtype, lanes=2, graph=True)
    for r, s in enumerate(seeds):
        want = _one_at_a_time(ic, n, s, cons[r])
        for f in range(4):
            scale = np.abs(want[f]).max()
            np.testing.assert_allclose(got[r, f], want[f], atol=tol * scale, err_msg="realisation %d field %d" % (r, f))
        # the constraints hold in the batch output: mean delta in each sphere
        for i, (x, y, z, rad, t) in enumerate(cons[r]):
            cells = select_sphere_cells(n, BOX, (x, y, z), rad)
            assert abs(got[r, 0].ravel()[cells].mean() - t) < 50 * tol * max(1.0, abs(t)), (r, i)
    ic.close()


def test_batch_graph_and_chains_are_bit_identical():
    import torch
    n, seeds = 64, [3, 1, 4, 1, 5]
    print("batch determinism", n, seeds, flush=True)
    cons = _constraints(len(seeds), 8, 11)
    ic = _ctx(gic.FLOAT32, n)
    a = _batch(ic, n, seeds, cons, torch.float32, lanes=1, graph=False)
    b = _batch(ic, n, seeds, cons, torch.float32, lanes=3, graph=True)
    # eager work on the context between two recorded batches
    ic.reseed(9)
    ic.draw()
    ic.apply_power_spectrum()
    c = _batch(ic, n, seeds[:2], cons[:2], torch.float32, lanes=2, graph=True)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(c, a[:2])
    # the recorded batch replayed for other seeds and constraints of the same shape
    seeds2, cons2 = [7, 8, 9, 10, 11], _constraints(len(seeds), 8, 12)
    d = _batch(ic, n, seeds2, cons2, torch.float32, lanes=3, graph=True)
    e = _batch(ic, n, seeds2, cons2, torch.float32, lanes=3, graph=True)
    f = _batch(ic, n, seeds2, cons2, torch.float32, lanes=1, graph=False)
    np.testing.assert_array_equal(d, f)
    np.testing.assert_array_equal(e, f)
    assert not np.array_equal(d, a)
    ic.close()


def test_batch_rejects_bad_input():
    import torch
    n = 32
    print("batch errors", n, flush=True)
    ic = _ctx(gic.FLOAT32, n)
    out = torch.empty((1, 4, n, n, n), dtype=torch.float32, device="cuda")
    # a sphere between cell centroids: no cell selected
    cons = np.array([[[0.0, 0.0, 0.0, 0.05, 1.0]]])
    with pytest.raises(gic.GICError, match="selects no cells"):
        ic.batch_constrained([1], cons, out.data_ptr())
    with pytest.raises(gic.GICError, match="constraints"):
        ic.batch_constrained([1], np.zeros((1, 33, 5)) + [10, 10, 10, 2, 0], out.data_ptr())
    # the context still works afterwards
    ok = _constraints(1, 2, 3)
    ic.batch_constrained([1], ok, out.data_ptr())
    assert torch.isfinite(out).all()
    ic.init_zoom_grid_with_origin_at(8, 8, 8, 2, 32)
    with pytest.raises(gic.GICError, match="one grid level"):
        ic.batch_constrained([1], ok, out.data_ptr())
    ic.close()


def test_batch_large_spheres_match_one_at_a_time():
    """Spheres wider than a box of the grid take the whole-field covectors;
    small and large ones mixed in one batch."""
    import torch
    n, seeds = 32, [21, 22]
    print("batch large spheres", n, seeds, flush=True)
    cons = _constraints(len(seeds), 3, 5)
    cons[:, 0, 3] = 22.0   # wider than the 32-cell grid's box can hold
    ic = _ctx(gic.FLOAT64, n)
    got = _batch(ic, n, seeds, cons, torch.float64, lanes=2, graph=True)
    for r, s in enumerate(seeds):
        want = _one_at_a_time(ic, n, s, cons[r])
        for f in range(4):
            np.testing.assert_allclose(got[r, f], want[f], atol=1e-9 * np.abs(want[f]).max(),
                                       err_msg="realisation %d field %d" % (r, f))
    ic.close()
