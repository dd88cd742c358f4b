"""Second-order LPT on the GPU.  The reference has no 2LPT, so parity is
pinned by closed forms (shared with the oracle's test) and by the oracle's
restatement on random fields."""
import numpy as np
import pytest

import pyoracle as po
from genetic_amd import gic
from helpers import golden, cosmology_from_paramfile

pytestmark = pytest.mark.gpu


def _ic(precision=gic.FLOAT64, n=32, box=50.0, zoom=None, seed=1):
    c = cosmology_from_paramfile(golden("test_10a_gadget/paramfile.txt"))
    ic = gic.ICGenerator(precision=precision)
    ic.set_cosmology(**c)
    ic.set_camb(golden("camb_transfer_kmax40_z0.dat"))
    ic.init_base_grid(box, n)
    o = po.OracleIC()
    o.set_cosmology(Om=c["Om"], Ol=c["Ol"], s8=c["s8"], ns=c["ns"], z=c["z"], hubble=c["hubble"])
    tab, nc = po.read_camb(golden("camb_transfer_kmax40_z0.dat"))
    o.set_camb(tab, nc)
    o.init_base_grid(box, n)
    if zoom:
        ic.init_zoom_grid_with_origin_at(*zoom)
        o.init_zoom_grid_at(*zoom)
    ic.set_seed(seed)
    o.seed(seed)
    ic.set_lpt_order(2)
    o.set_lpt_order(2)
    return ic, o


@pytest.mark.parametrize("precision,tol", [(gic.FLOAT64, 1e-12), (gic.FLOAT32, 2e-6)])
def test_2lpt_two_waves_closed_form(precision, tol):
    print("2lpt closed form", precision, flush=True)
    n, box, A, B = 32, 50.0, 0.3, 0.2
    ic, _ = _ic(precision, n, box)
    ic.apply_power_spectrum()
    x = np.arange(n) * box / n
    k1, k2 = 2 * np.pi * 2 / box, 2 * np.pi * 3 / box
    X, Y = np.meshgrid(x, x, indexing="ij")
    delta = (A * np.cos(k1 * X) + B * np.cos(k2 * Y))[:, :, None] + np.zeros((n, n, n))
    ic.set_field(0, delta)
    ic.zeldovich()
    ex = -(3 / 7) * A * B * k1 * np.sin(k1 * X) * np.cos(k2 * Y) / (k1 ** 2 + k2 ** 2)
    ey = -(3 / 7) * A * B * k2 * np.cos(k1 * X) * np.sin(k2 * Y) / (k1 ** 2 + k2 ** 2)
    s = np.abs(ex).max()
    assert np.abs(ic.field(gic.FIELD_PSI2_X, 0) - ex[:, :, None]).max() < tol * s * 10
    assert np.abs(ic.field(gic.FIELD_PSI2_Y, 0) - ey[:, :, None]).max() < tol * s * 10
    assert np.abs(ic.field(gic.FIELD_PSI2_Z, 0)).max() < tol * s * 10


@pytest.mark.parametrize("precision,tol", [(gic.FLOAT64, 1e-12), (gic.FLOAT32, 2e-6)])
def test_2lpt_mixed_derivatives_closed_form(precision, tol):
    """delta = A cos(k (x + y)) + B cos(k2 z): the phi_xy^2 term cancels
    phi_xx phi_yy and S = A B cos(k (x + y)) cos(k2 z); a (1,1,1) plane wave
    gives S = 0 only through the mixed terms."""
    print("2lpt mixed derivatives", precision, flush=True)
    from test_cpu_oracle import _diagonal_waves
    n, box = 32, 50.0
    delta, ex, ey, ez, plane = _diagonal_waves(n, box)
    ic, _ = _ic(precision, n, box)
    ic.apply_power_spectrum()
    ic.set_field(0, delta)
    ic.zeldovich()
    s = np.abs(ex).max()
    for f, e in zip((gic.FIELD_PSI2_X, gic.FIELD_PSI2_Y, gic.FIELD_PSI2_Z), (ex, ey, ez)):
        assert np.abs(ic.field(f, 0) - e).max() < tol * s * 10
    ic2, _ = _ic(precision, n, box)
    ic2.apply_power_spectrum()
    ic2.set_field(0, plane)
    ic2.zeldovich()
    for f in (gic.FIELD_PSI2_X, gic.FIELD_PSI2_Y, gic.FIELD_PSI2_Z):
        assert np.abs(ic2.field(f, 0)).max() < (1e-14 if precision == gic.FLOAT64 else 1e-6)


def test_2lpt_plane_wave_has_no_second_order():
    n, box = 32, 50.0
    ic, _ = _ic(gic.FLOAT64, n, box)
    ic.apply_power_spectrum()
    x = np.arange(n) * box / n
    delta = 0.4 * np.cos(2 * np.pi * 5 * x / box)[None, :, None] + np.zeros((n, n, n))
    ic.set_field(0, delta)
    ic.zeldovich()
    for f in (gic.FIELD_PSI2_X, gic.FIELD_PSI2_Y, gic.FIELD_PSI2_Z):
        assert np.abs(ic.field(f, 0)).max() < 1e-15


@pytest.mark.parametrize("precision,rtol", [(gic.FLOAT64, 1e-10), (gic.FLOAT32, 5e-5)])
def test_2lpt_random_field_matches_oracle(precision, rtol):
    """64^3 single level and a x2 zoom: psi2 on every level and the particle
    positions and velocities match the oracle."""
    print("2lpt random", precision, flush=True)
    for zoom, n in ((None, 64), ((8, 8, 8, 2, 32), 32)):
        ic, o = _ic(precision, n, 50.0, zoom, seed=77)
        ic.zeldovich()
        o.zeldovich()
        for lev in range(2 if zoom else 1):
            for f in range(4, 7):
                a, b = ic.field(f, lev), o.field(f, lev)
                assert np.abs(a - b).max() <= rtol * np.abs(b).max() * 10, (zoom, lev, f)
            pa, va = ic.particles(lev)
            pb, vb, _, _ = o.particles(lev)
            assert np.abs(pa - pb).max() < 1e-4
            assert np.abs(va - vb).max() <= 10 * rtol * np.abs(vb).max() + 1e-4
