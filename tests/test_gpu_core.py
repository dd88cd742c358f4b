"""GPU parity of the core path (RNG, FFT, P(k), Zel'dovich, levels,
particles) against the CPU oracle and the reference's fixtures, through the
C-ABI of libgic_hip.so."""
import numpy as np
import pytest

import pyoracle as po
from genetic_amd import gic
from helpers import golden, read_gadget, assert_decimal, cosmology_from_paramfile

pytestmark = pytest.mark.gpu

SEED = 8896131


def _ic(precision=gic.FLOAT64):
    return gic.ICGenerator(precision=precision)


@pytest.mark.parametrize("seed,n", [(8896131, 100_000), (42, 77_777), (13842314, 3_000_000), (1, 50)])
def test_raw_stream_bit_exact(seed, n):
    """The 24-bit ranlxs2 integers are identical to the serial restatement."""
    ic = _ic()
    got = ic.raw_stream(seed, n)
    want = po.ranlxs_raw(seed, n)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("seed,n", [(8896131, 64 ** 3), (99, 1_000_003)])
def test_gaussian_stream(seed, n):
    """Ziggurat Gaussians follow the serial stream; only the tail values go
    through log/exp, so agreement is to a few ulp."""
    ic = _ic()
    got = ic.gaussian_stream(seed, n)
    want, _ = po.gaussians(seed, n)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64, 128])
@pytest.mark.parametrize("precision,tol", [(gic.FLOAT64, 1e-13), (gic.FLOAT32, 2e-6)])
def test_fft_matches_oracle(n, precision, tol):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n, n))
    ic = _ic(precision)
    ic.init_base_grid(50.0, n)
    ic.set_field(0, x)
    spec = ic.field(gic.FIELD_OVERDENSITY, 0, fourier=True)
    want = po.fft_r2c(x)
    scale = np.abs(want).max()
    assert np.abs(spec - want).max() <= tol * scale * 4
    back = ic.field(gic.FIELD_OVERDENSITY, 0, fourier=False)
    assert np.abs(back - x).max() <= tol * np.abs(x).max() * 8


@pytest.mark.parametrize("n", [256, 512])
@pytest.mark.parametrize("precision,tol", [(gic.FLOAT64, 1e-13), (gic.FLOAT32, 2e-6)])
def test_fft_matches_oracle_large(n, precision, tol):
    """The tile shapes of the large sizes (FP32: 512 threads x 16 columns at
    256, 1024 x 16 at 512; FP64: 512 x 8 at 256, 256 x 8 at 512) against the
    oracle's transform, both directions."""
    print("fft vs oracle n=%d precision=%d" % (n, precision), flush=True)
    rng = np.random.default_rng(n + precision)
    x = rng.standard_normal((n, n, n))
    ic = _ic(precision)
    ic.init_base_grid(50.0, n)
    ic.set_field(0, x if precision == gic.FLOAT64 else x.astype(np.float32))
    spec = ic.field(gic.FIELD_OVERDENSITY, 0, fourier=True)
    want = po.fft_r2c(x)
    assert np.abs(spec - want).max() <= tol * np.abs(want).max() * 4
    del spec, want
    back = ic.field(gic.FIELD_OVERDENSITY, 0, fourier=False)
    assert np.abs(back - x).max() <= tol * np.abs(x).max() * 8
    ic.close()


def _separable_spectrum_slab(A, Bk, C, kx, n):
    """Spectrum slab kx of sum_i a_i(x) b_i(y) c_i(z), unitary (N^-1.5),
    forward sign: sum_i A_i(kx) B_i(ky) C_i(kz), kz in [0, N/2]."""
    s = np.zeros((n, n // 2 + 1), dtype=np.complex128)
    for i in range(len(A)):
        s += A[i][kx] * np.outer(Bk[i], C[i])
    return s / n ** 1.5


@pytest.mark.parametrize("precision,tol", [(gic.FLOAT64, 1e-12), (gic.FLOAT32, 1e-5)])
def test_fft_1024_separable_field(precision, tol):
    """1024^3 (FP32 512 x 8 tiles, FP64 1024 x 8): a sum of three separable
    products, whose spectrum is the product of 1-D transforms, in every mode;
    Parseval; and the round trip back to the field."""
    n = 1024
    print("fft 1024 separable precision=%d" % precision, flush=True)
    rng = np.random.default_rng(1024 + precision)
    a, b, c = (rng.standard_normal((3, n)) for _ in range(3))
    dt = np.float64 if precision == gic.FLOAT64 else np.float32
    x = np.zeros((n, n, n), dtype=dt)
    for i in range(3):
        x += (a[i][:, None, None] * b[i][None, :, None]).astype(dt) * c[i][None, None, :].astype(dt)
    A = np.fft.fft(a, axis=1)
    Bk = np.fft.fft(b, axis=1)
    C = np.fft.rfft(c, axis=1)
    ic = _ic(precision)
    ic.init_base_grid(200.0, n)
    ic.set_field(0, x)
    spec = ic.field(gic.FIELD_OVERDENSITY, 0, fourier=True)
    scale = 0.0
    err = 0.0
    for kx in range(n):
        want = _separable_spectrum_slab(A, Bk, C, kx, n)
        err = max(err, float(np.abs(spec[kx] - want).max()))
        scale = max(scale, float(np.abs(want).max()))
    assert err <= tol * scale, (err, scale)
    # Parseval (unitary): the half-space modes count twice off the kz = 0, N/2 planes
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    e_k = sum(float(np.sum(np.abs(spec[kx]) ** 2 * w)) for kx in range(n))
    e_x = sum(float(np.sum(np.square(x[i], dtype=np.float64))) for i in range(n))
    assert abs(e_k - e_x) <= 10 * tol * e_x
    del spec
    back = ic.field(gic.FIELD_OVERDENSITY, 0, fourier=False)
    assert np.abs(back - x).max() <= 8 * tol * np.abs(x).max()
    ic.close()


def _setup_pair(paramfile, camb, levels, precision=gic.FLOAT64, seed=SEED, mode=gic.SEED_REAL_SPACE):
    c = cosmology_from_paramfile(golden(paramfile))
    ic = _ic(precision)
    ic.set_cosmology(**c)
    ic.set_camb(golden(camb))
    o = po.OracleIC()
    o.set_cosmology(Om=c["Om"], Ol=c["Ol"], s8=c["s8"], ns=c["ns"], z=c["z"], hubble=c["hubble"])
    tab, nc = po.read_camb(golden(camb))
    o.set_camb(tab, nc)
    box, n = levels[0]
    ic.init_base_grid(box, n)
    o.init_base_grid(box, n)
    for (x0, y0, z0, fac, nz) in levels[1:]:
        ic.init_zoom_grid_with_origin_at(x0, y0, z0, fac, nz)
        o.init_zoom_grid_at(x0, y0, z0, fac, nz)
    ic.set_seed(seed, mode)
    o.seed(seed, mode)
    return ic, o


@pytest.mark.parametrize("precision,rtol", [(gic.FLOAT64, 1e-10), (gic.FLOAT32, 2e-5)])
def test_single_level_delta_psi(precision, rtol):
    ic, o = _setup_pair("test_10a_gadget/paramfile.txt", "camb_transfer_kmax40_z0.dat", [(50.0, 64)], precision)
    for x in (ic, o):
        x.draw()
    w_gpu = ic.field(gic.FIELD_OVERDENSITY, 0, fourier=True)
    w_orc = o.field(0, 0, fourier=True)
    assert np.abs(w_gpu - w_orc).max() <= rtol * np.abs(w_orc).max() * 10
    ic.apply_power_spectrum()
    o.apply_power_spectrum()
    ic.zeldovich()
    o.zeldovich()
    for which in range(4):
        a = ic.field(which, 0)
        b = o.field(which, 0)
        assert np.abs(a - b).max() <= rtol * np.abs(b).max() * 10, which


@pytest.mark.parametrize("precision,rtol", [(gic.FLOAT64, 1e-9), (gic.FLOAT32, 5e-5)])
def test_two_level_zoom_matches_oracle_and_fixture(precision, rtol):
    """test_01c geometry: 64^3 base in 50 Mpc/h, x3 zoom of 64^3 at (21,21,21)."""
    ic, o = _setup_pair("test_01c_zoom_grid/paramfile.txt", "camb_transfer_kmax40_z0.dat",
                        [(50.0, 64), (21, 21, 21, 3, 64)], precision)
    for x in (ic, o):
        x.zeldovich()
    for level in range(2):
        for which in range(4):
            a = ic.field(which, level)
            b = o.field(which, level)
            assert np.abs(a - b).max() <= rtol * np.abs(b).max() * 10, (level, which)
        ref = np.load(golden("test_01c_zoom_grid/reference_grid/grid-%d.npy" % level))
        assert_decimal(ic.field(0, level), ref, 4, "grid-%d" % level)


@pytest.mark.parametrize("precision,rtol", [(gic.FLOAT64, 1e-10), (gic.FLOAT32, 5e-5)])
def test_three_level_zoom_matches_oracle_and_fixture(precision, rtol):
    """test_01d geometry: 64^3 base in 50 Mpc/h, x3 zoom of 64^3 at
    (21,21,21), x2 zoom of 64^3 at (16,16,16) inside it, seed 889613 (the
    stream continues over three levels; level 2 is combined from the
    already combined level 1)."""
    print("three-level zoom precision=%d" % precision, flush=True)
    ic, o = _setup_pair("test_01d_two_zooms/paramfile.txt", "camb_transfer_kmax40_z0.dat",
                        [(50.0, 64), (21, 21, 21, 3, 64), (16, 16, 16, 2, 64)], precision, seed=889613)
    for x in (ic, o):
        x.zeldovich()
    for level in range(3):
        for which in range(4):
            a = ic.field(which, level)
            b = o.field(which, level)
            assert np.abs(a - b).max() <= rtol * np.abs(b).max() * 10, (level, which)
    ref = np.load(golden("test_01d_two_zooms/reference_grid/grid-2.npy"))
    assert_decimal(ic.field(0, 2), ref, 4, "grid-2")
    ic.close()


def test_gadget_fixture_particles():
    """test_10a_gadget: 8^3 base grid, gadget3 positions/velocities/mass."""
    ref = read_gadget(golden("test_10a_gadget/reference_output"))
    c = cosmology_from_paramfile(golden("test_10a_gadget/paramfile.txt"))
    ic = _ic()
    ic.set_cosmology(**c)
    ic.set_camb(golden("camb_transfer_kmax40_z0.dat"))
    ic.init_base_grid(50.0, 8)
    ic.set_seed(SEED)
    ic.zeldovich()
    pos, vel = ic.particles(0)
    mass, _ = ic.particle_mass_eps(0)
    assert mass == ref["mass"][1]
    assert_decimal(pos, ref["pos"], 4, "pos")
    assert_decimal(vel, ref["vel"], 4, "vel")


FOURIER_MODES = [gic.SEED_FOURIER_SERIAL, gic.SEED_FOURIER_PARALLEL, gic.SEED_FOURIER_REVERSE]


@pytest.mark.parametrize("mode", FOURIER_MODES)
@pytest.mark.parametrize("n", [4, 8, 16, 64])
def test_fourier_seeded_noise_matches_oracle(mode, n):
    """Fourier-space white noise: every stored mode, including the slots
    written twice through the conjugate mirror, equals the serial oracle's
    (FP64: a few ulp from the ziggurat tails)."""
    print("fourier noise", mode, n, flush=True)
    ic, o = _setup_pair("test_10a_gadget/paramfile.txt", "camb_transfer_kmax40_z0.dat", [(50.0, n)],
                        seed=8896131, mode=mode)
    ic.draw()
    o.draw()
    a = ic.field(gic.FIELD_OVERDENSITY, 0, fourier=True)
    b = o.field(0, 0, fourier=True)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-13)
    assert np.count_nonzero(b) > 0.8 * b.size * (1 - 3.0 / n)


@pytest.mark.parametrize("mode", FOURIER_MODES)
@pytest.mark.parametrize("precision,rtol", [(gic.FLOAT64, 1e-10), (gic.FLOAT32, 2e-5)])
def test_fourier_seeded_delta_psi(mode, precision, rtol):
    ic, o = _setup_pair("test_10a_gadget/paramfile.txt", "camb_transfer_kmax40_z0.dat", [(50.0, 64)], precision,
                        seed=42, mode=mode)
    ic.zeldovich()
    o.zeldovich()
    for which in range(4):
        a = ic.field(which, 0)
        b = o.field(which, 0)
        assert np.abs(a - b).max() <= rtol * np.abs(b).max() * 10, which


@pytest.mark.parametrize("mode", [gic.SEED_FOURIER_SERIAL, gic.SEED_FOURIER_PARALLEL])
def test_fourier_seeded_two_levels(mode):
    """Serial mode continues one stream into the zoom level; parallel mode
    restarts the per-shell seeds on every level."""
    ic, o = _setup_pair("test_01c_zoom_grid/paramfile.txt", "camb_transfer_kmax40_z0.dat",
                        [(50.0, 32), (10, 10, 10, 2, 32)], seed=1234, mode=mode)
    ic.zeldovich()
    o.zeldovich()
    for level in range(2):
        for which in range(4):
            a = ic.field(which, level)
            b = o.field(which, level)
            assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), (level, which)
