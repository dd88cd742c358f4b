"""The paramfile front-end (genetIC-hip) run on the reference's own test
paramfiles, its outputs compared with the reference outputs at the
tolerances of genetIC's tools/compare.py (particles: mass/eps 6 decimals,
pos/vel 4 decimals, gadget IDs exact; grids 4 decimals; .ps rtol 1e-4)."""
import os

import numpy as np
import pytest

from helpers import (assert_decimal, compare_particles, golden, read_gadget, read_tipsy, read_tipsy_array, run_frontend,
                     stage_test)

pytestmark = pytest.mark.gpu


def _run(test, tmp_path, args=(), drop=()):
    print("frontend case", test, args, flush=True)
    d = stage_test(test, tmp_path)
    for f in drop:
        os.remove(os.path.join(d, f))
    rc, err = run_frontend(d, args)
    assert rc == 0, err
    return d, err


def particle_file(d):
    """The one particle output of a run, found as compare.py does."""
    import glob
    files = glob.glob(os.path.join(d, "*.tipsy")) + glob.glob(os.path.join(d, "*.gadget?"))
    assert len(files) == 1, files
    return files[0]


@pytest.mark.parametrize("test", [
    "test_07a_CAMB_new_interface",
    "test_10a_gadget",
    "test_01g_zoom_grid_vel_offset",
    "test_23_doublezoom_wrap_issue_90",
    "test_25_zoom_factor_1",
    "test_13_parallel_seed",
    "test_05a_seed_old",
    "test_10h_gadget_zoom_ptype_gas",
])
def test_particle_output_matches_reference(tmp_path, test):
    d, _ = _run(test, tmp_path)
    compare_particles(particle_file(d), golden(test, "reference_output"))


def test_zoom_factor_1_overdensity_array(tmp_path):
    """dump_tipsy writes delta at every particle in output order: one value
    per particle, each equal to the delta of its cell on its own level."""
    text = open(golden("test_25_zoom_factor_1", "paramfile.txt")).read() + "\ndump_grid 0\ndump_grid 1\n"
    d = stage_test("test_25_zoom_factor_1", tmp_path, text)
    rc, err = run_frontend(d)
    assert rc == 0, err
    arr = read_tipsy_array(os.path.join(d, "test_25.tipsy.overden"))
    t = read_tipsy(os.path.join(d, "test_25.tipsy"))
    assert arr.size == t["n"]
    g0 = np.load(os.path.join(d, "grid-0.npy")).ravel().astype(np.float32)
    g1 = np.load(os.path.join(d, "grid-1.npy")).ravel().astype(np.float32)
    n_zoom = t["n"] - (8 ** 3 - (t["n"] - 8 ** 3) // 7)
    assert np.isin(arr[:t["n"] - n_zoom], g0).all()
    assert np.isin(arr[t["n"] - n_zoom:], g1).all()


def test_id_file_zoom_grids_and_crosscheck(tmp_path):
    """test_01c: idfile flags, dump_idfile round trip, zoom grid placement
    and the combined delta of both levels."""
    d, err = _run("test_01c_zoom_grid", tmp_path, drop=["zoomtest_crosscheck.txt"])
    assert open(os.path.join(d, "zoomtest_crosscheck.txt")).read().split() == \
        open(golden("test_01c_zoom_grid", "zoomtest_crosscheck.txt")).read().split()
    for lev in (0, 1):
        ref = np.load(golden("test_01c_zoom_grid", "reference_grid", "grid-%d.npy" % lev))
        got = np.load(os.path.join(d, "grid-%d.npy" % lev))
        assert_decimal(got, ref, 4, "grid-%d" % lev)
        info = open(os.path.join(d, "grid-info-%d.txt" % lev)).read().split("\n")[0].split()
        rinfo = open(golden("test_01c_zoom_grid", "reference_grid", "grid-info-%d.txt" % lev)).read().split("\n")[0]
        np.testing.assert_allclose([float(x) for x in info], [float(x) for x in rinfo.split()], rtol=1e-5)


def test_two_zoom_grids_and_power_spectra(tmp_path):
    """test_01d: grids of the deepest level and dump_ps of all three levels
    (reference_ps, rtol 1e-4 as compare.py)."""
    d, _ = _run("test_01d_two_zooms", tmp_path)
    ref = np.load(golden("test_01d_two_zooms", "reference_grid", "grid-2.npy"))
    got = np.load(os.path.join(d, "grid-2.npy"))
    assert_decimal(got, ref, 4, "grid-2")
    for lev in (0, 1, 2):
        r = np.loadtxt(golden("test_01d_two_zooms", "reference_ps", "test_01d_%d.ps" % lev))
        g = np.loadtxt(os.path.join(d, "test_1_%d.ps" % lev))
        assert r.shape == g.shape
        np.testing.assert_allclose(g, r, rtol=1e-4)


def test_float32_fields_stay_close(tmp_path):
    """-f holds the fields in float32: the gadget output stays within the
    float precision of the FP64 run (positions 1e-4 Mpc/h, velocities 1e-3
    of their range)."""
    d64, _ = _run("test_10a_gadget", tmp_path / "f64")
    d32, _ = _run("test_10a_gadget", tmp_path / "f32", ["-f"])
    from helpers import read_gadget
    a = read_gadget(particle_file(d64))
    b = read_gadget(particle_file(d32))
    np.testing.assert_array_equal(a["ids"], b["ids"])
    assert np.abs(a["pos"] - b["pos"]).max() < 1e-4
    assert np.abs(a["vel"] - b["vel"]).max() < 1e-3 * np.abs(a["vel"]).max()


def _log_lines_present(test, err):
    """compare.py's compare_outputlogfile: every line of reference.txt occurs
    in the (time-stamp stripped) log."""
    lines = err.splitlines()
    for want in open(golden(test, "reference.txt")).read().splitlines():
        assert want in lines, "missing log line: %r" % want


@pytest.mark.parametrize("test", [
    "test_02k_chisquare_onelevel",
    "test_02l_chisquare_modification",
    "test_02m_chisquare_twolevels",
    "test_02f_onelevelvariance",
    "test_02g_multilevelvariance",
    "test_02j_invertible_modifs",
])
def test_modification_logs_match_reference(tmp_path, test):
    d, err = _run(test, tmp_path)
    _log_lines_present(test, err)


@pytest.mark.parametrize("test", [
    "test_02_simple_modif",
    "test_14_velocity_modif",
    "test_22_angular_momentum_modif",
    "test_20_velmodif_zoom",
    "test_02e_two_separate_modifs",
])
def test_modified_particle_output_matches_reference(tmp_path, test):
    d, _ = _run(test, tmp_path)
    compare_particles(particle_file(d), golden(test, "reference_output"))


@pytest.mark.parametrize("test", [
    "test_02a_onemodif_zeroed_zoom",
    "test_02b_onemodif_zoom",
    "test_02d_modif_double_zoom",
    "test_02i_multilevel_algorithm",
    "test_02c_two_modifs",
])
def test_modified_zoom_grids_match_reference(tmp_path, test):
    """Constraints on zoom hierarchies (one and two zooms, zeroed levels, a
    filtered multi-level variance modification): every level's dumped grid
    as compare.py compare_grids checks it, 4 decimals."""
    import glob
    d, _ = _run(test, tmp_path)
    refs = sorted(glob.glob(os.path.join(golden(test, "reference_grid"), "grid-?.npy")))
    assert refs
    for ref in refs:
        name = os.path.basename(ref)
        assert_decimal(np.load(os.path.join(d, name)), np.load(ref), 4, test + " " + name)


def test_modified_overdensity_array_and_grids(tmp_path):
    """test_02's dump_tipsy array (compare.py: 5 decimals) and test_20's grids."""
    d, _ = _run("test_02_simple_modif", tmp_path / "a")
    got = read_tipsy_array(os.path.join(d, "test_2.tipsy.overdensity"))
    ref = read_tipsy_array(golden("test_02_simple_modif", "reference_output.overdensity"))
    assert_decimal(got, ref, 5, "overdensity")
    d, _ = _run("test_20_velmodif_zoom", tmp_path / "b")
    for lev in (0, 1):
        g = np.load(os.path.join(d, "grid-%d.npy" % lev))
        r = np.load(golden("test_20_velmodif_zoom", "reference_grid", "grid-%d.npy" % lev))
        assert_decimal(g, r, 4, "grid-%d" % lev)


_BASE = """Om 0.279
Ol 0.721
s8 0.817
zin 99
random_seed_real_space 8896131
camb ../camb_transfer_kmax40_z0.dat
outname out
outdir ./
outformat tipsy
basegrid 50.0 16
"""


def test_expand_flagged_region_is_a_one_cell_box_dilation(tmp_path):
    """expand_flagged_region 1 (grid.hpp:282-300): neighbours along z, then
    y, then x join the flags, which adds up to every cell within one cell
    (Chebyshev distance, wrapped) of a flagged one."""
    text = _BASE + "center 3 3 3\nselect_sphere 4\ndump_id_file before.txt\nexpand_flagged_region 1\n" \
                   "dump_id_file after.txt\n"
    d = stage_test("expand_case", tmp_path, text)
    rc, err = run_frontend(d)
    assert rc == 0, err
    assert "Expand flagged region by 1 cells" in err
    before = np.loadtxt(os.path.join(d, "before.txt"), dtype=np.int64, ndmin=1)
    after = np.loadtxt(os.path.join(d, "after.txt"), dtype=np.int64, ndmin=1)
    n = 16
    m = np.zeros((n, n, n), bool)
    m.ravel()[before] = True
    want = np.zeros_like(m)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                want |= np.roll(m, (dx, dy, dz), axis=(0, 1, 2))
    np.testing.assert_array_equal(np.sort(after), np.flatnonzero(want.ravel()))
    assert "increased number of flagged cells by %d (now %d)" % (len(after) - len(before), len(after)) in err


def test_num_neighbours_and_centre_output_are_accepted(tmp_path):
    """num_neighbours (ic.hpp:1268-1271) sets the SPH smoothing scale, which
    only the HDF5 writer reads: the tipsy softening stays cell size x
    epsNorm (0.01075, ic.hpp:101; zeldovich.hpp:116-118), in the tipsy length
    unit of the box size."""
    d = stage_test("neighbours_case", tmp_path, _BASE + "num_neighbours 32\ncentre_output\ndone\n")
    rc, err = run_frontend(d)
    assert rc == 0, err
    t = read_tipsy(os.path.join(d, "out.tipsy"))
    np.testing.assert_allclose(t["dark"]["eps"], 0.01075 / 16, rtol=1e-6)


def test_dump_vx_is_the_zeldovich_offset_of_the_dumped_grid(tmp_path):
    """dump_vx 0 (ic.hpp:920-923): the x offset field, whose transform is
    i k_x delta_k / k^2 of dump_grid's delta (zeldovich.hpp:151-181), zero at
    k = 0 and at the x Nyquist plane."""
    d = stage_test("vx_case", tmp_path, _BASE + "done\ndump_grid 0\ndump_vx 0\n")
    rc, err = run_frontend(d)
    assert rc == 0, err
    delta = np.load(os.path.join(d, "grid-0.npy"))
    vx = np.load(os.path.join(d, "vx-0.npy"))
    n, box = 16, 50.0
    f = np.fft.fftfreq(n, 1.0 / n)
    kx, ky, kz = np.meshgrid(f, f, f, indexing="ij")
    kw = 2 * np.pi / box
    k2 = (kx ** 2 + ky ** 2 + kz ** 2) * kw * kw
    with np.errstate(divide="ignore", invalid="ignore"):
        op = np.where(k2 > 0, 1j * kx * kw / k2, 0.0)
    op[n // 2, :, :] = 0
    want = np.fft.ifftn(op * np.fft.fftn(delta)).real
    assert_decimal(vx, want, 5, "vx-0")
    assert open(os.path.join(d, "vx-info-0.txt")).read().split("\n")[0].split() == ["0", "0", "0", "50"]


def test_wrapping_double_zoom_matches_reference(tmp_path):
    """test_01e_wrapping_zoom: two nested zooms around a point near the box
    corner, so both windows wrap; particles as compare.py checks them and
    the dump_tipsy overdensity array to 5 decimals (compare.py:50-51)."""
    test = "test_01e_wrapping_zoom"
    d, _ = _run(test, tmp_path)
    compare_particles(particle_file(d), golden(test, "reference_output"))
    got = read_tipsy_array(os.path.join(d, "test_1.tipsy.overdensity"))
    want = read_tipsy_array(golden(test, "reference_output.overdensity"))
    assert got.shape == want.shape
    np.testing.assert_almost_equal(got, want, decimal=5)


def test_gadget_flagged_particle_type(tmp_path):
    """test_10d (run-only in the reference): the particles an id file flags
    on the finest level leave their block and follow every other particle as
    gadget type 0, with the same positions and velocities as the same run
    without the command, new ids at the end of the mapper and the finest
    level's mass (ic.hpp:706-714,1180-1190)."""
    test = "test_10d_gadget_flagged_particles"
    d, _ = _run(test, tmp_path / "flagged")
    got = read_gadget(particle_file(d))
    plain = open(golden(test, "paramfile.txt")).read().replace("gadget_flagged_particle_type 0", "")
    d2 = stage_test(test, tmp_path / "plain", paramfile_text=plain)
    rc, err = run_frontend(d2)
    assert rc == 0, err
    ref = read_gadget(particle_file(d2))
    n = 64 ** 3
    flagged = np.unique(np.loadtxt(golden(test, "zoomtest.txt"), dtype=np.int64))
    nf = flagged.size
    assert got["npart"][:2] == (nf, n - nf) and sum(got["npart"][2:]) == 0
    assert ref["npart"][1] == n
    rest = np.setdiff1d(np.arange(n), flagged)
    # file order is by type: the flagged block (type 0) first, then type 1
    np.testing.assert_array_equal(got["ids"][:nf], np.arange(n - nf, n))
    np.testing.assert_array_equal(got["ids"][nf:], np.arange(n - nf))
    np.testing.assert_array_equal(got["pos"][:nf], ref["pos"][flagged])
    np.testing.assert_array_equal(got["vel"][:nf], ref["vel"][flagged])
    np.testing.assert_array_equal(got["pos"][nf:], ref["pos"][rest])
    np.testing.assert_array_equal(got["vel"][nf:], ref["vel"][rest])
    assert got["mass"][0] == got["mass"][1] == ref["mass"][1]


@pytest.mark.parametrize("test", ["test_01h_zoom_grid_pos_offset", "test_01i_zoom_pos_offset_wrap"])
def test_centred_grafic_zoom_runs(tmp_path, test):
    """test_01h / 01i (run-only in the reference's compare.py): a zoom near a
    box corner, wrapping in 01i, written as centred grafic output -- the run
    succeeds and writes one complete directory per level."""
    import glob
    d, _ = _run(test, tmp_path)
    dirs = sorted(glob.glob(os.path.join(d, "*.grafic_*")))
    assert len(dirs) >= 2
    for g in dirs:
        for name in ("ic_velcx", "ic_poscx", "ic_deltab", "ic_refmap", "ic_particle_ids"):
            assert os.path.getsize(os.path.join(g, name)) > 44, (g, name)


@pytest.mark.parametrize("test", [
    "test_04a_supersampling",
    "test_04b_subsampling",
    "test_04d_gas_supersampling",
    "test_10e_gadget_zoom_ptype_resample",
])
def test_resampled_particle_output_matches_reference(tmp_path, test):
    """Particles laid on sub/supersampled grids (04a: DM supersampled beside
    unsampled gas; 04b: base subsampled and zoom supersampled; 04d: gas and
    DM supersampled by different factors; 10e: gadget types kept through
    both) at compare.py's tolerances; 04b's dump_tipsy overdensity array,
    evaluated on the same virtual grids, to 5 decimals (compare.py:50-51)."""
    d, _ = _run(test, tmp_path)
    compare_particles(particle_file(d), golden(test, "reference_output"))
    if test == "test_04b_subsampling":
        got = read_tipsy_array(os.path.join(d, "test_4.tipsy.overdensity"))
        want = read_tipsy_array(golden(test, "reference_output.overdensity"))
        assert got.shape == want.shape
        np.testing.assert_almost_equal(got, want, decimal=5)


def test_double_zoom_supersampling_keeps_the_coarse_levels(tmp_path):
    """test_04c (run-only in the reference): supersampling the finest of two
    zoom levels by 2 leaves the particles of the levels above unchanged and
    gives each finest cell 8 particles of an eighth of its mass with its
    softening, whose mean displacement follows the cell's own."""
    test = "test_04c_doublezoom_supersampling"
    d, _ = _run(test, tmp_path)
    plain = os.path.join(str(tmp_path), "plain")
    text = open(os.path.join(d, "paramfile.txt")).read()
    assert "supersample 2" in text
    dp = stage_test(test, plain, text.replace("supersample 2", ""))
    rc, err = run_frontend(dp)
    assert rc == 0, err
    a = read_tipsy(os.path.join(d, "test_1.tipsy"))["dark"]
    b = read_tipsy(os.path.join(dp, "test_1.tipsy"))["dark"]
    mfine = b["mass"].min()
    nfine = int(np.sum(b["mass"] == mfine))
    ncoarse = len(b) - nfine
    assert len(a) == ncoarse + 8 * nfine
    for k in ("pos", "vel", "eps"):
        np.testing.assert_array_equal(a[k][:ncoarse], b[k][:ncoarse])
    # tipsy masses are normalised to sum to Om: compare ratios
    np.testing.assert_allclose(a["mass"][ncoarse:] / a["mass"][0], b["mass"][-1] / b["mass"][0] / 8, rtol=1e-6)
    np.testing.assert_array_equal(a["eps"][ncoarse:], b["eps"][-1])
    # velocities are proportional to displacements: the supersampled level
    # interpolates the same field, so its mean velocity over the level and
    # its rms stay within a few per cent of the unsampled level's
    fa, fb = a["vel"][ncoarse:], b["vel"][ncoarse:]
    np.testing.assert_allclose(fa.mean(axis=0), fb.mean(axis=0), atol=0.02 * fb.std())
    np.testing.assert_allclose(fa.std(axis=0), fb.std(axis=0), rtol=0.05)


@pytest.mark.parametrize("test", [
    "test_06a_stray_particles",
    "test_06b_stray_supersample",
    "test_06c_stray_subsampling",
    "test_06d_stray_wrapping_zoom",
])
def test_stray_particles_match_reference(tmp_path, test):
    """strays_on: zoom particles outside the zoom window are kept, laid on
    the zoom level's box-covering grid and interpolated from the base level
    (06a; 06b supersampled; 06c with a subsampled base, gadget; 06d a zoom
    window that wraps), at compare.py's tolerances, and the dump_tipsy
    overdensity arrays of 06a and 06b to 5 decimals."""
    d, err = _run(test, tmp_path)
    # 06c's flagged region sits far from the zoom window's edges: no strays
    assert ("will be interpolated from LR grid (stray particle mode)" in err) == (test != "test_06c_stray_subsampling")
    compare_particles(particle_file(d), golden(test, "reference_output"))
    if os.path.exists(golden(test, "reference_output.overdensity")):
        got = read_tipsy_array(os.path.join(d, "test_6.tipsy.overdensity"))
        want = read_tipsy_array(golden(test, "reference_output.overdensity"))
        assert got.shape == want.shape
        np.testing.assert_almost_equal(got, want, decimal=5)


@pytest.mark.parametrize("test", [
    "test_10i_gadget_zoom_ptype_autopad",
    "test_18_intermediate_levels",
])
def test_autopad_matches_reference(tmp_path, test):
    """autopad: a layer of the coarse level supersampled by 2 around each zoom
    that refines by more than 2 (10i: gadget types and gas; 18: two zooms by 4
    under a subsampled base, with adapt_mask), at compare.py's tolerances."""
    d, err = _run(test, tmp_path)
    assert "Added a padding region of effective resolution" in err
    compare_particles(particle_file(d), golden(test, "reference_output"))


def test_adapt_mask_two_zooms(tmp_path):
    """test_01f (run-only in the reference): adapt_mask 1 redraws the base
    level's zoom cells as the parents of the second level's zoomed cells plus
    one cell; the finest level's particles are those of the same run without
    it, the two coarser levels trade particles, and the masses keep their
    ratios 8 and 27."""
    test = "test_01f_two_zooms_adapt_mask"
    d, _ = _run(test, tmp_path)
    text = open(os.path.join(d, "paramfile.txt")).read()
    assert "adapt_mask 1" in text
    dp = stage_test(test, os.path.join(str(tmp_path), "plain"), text.replace("adapt_mask 1", ""))
    rc, err = run_frontend(dp)
    assert rc == 0, err
    a = read_tipsy(os.path.join(d, "test_1.tipsy"))["dark"]
    b = read_tipsy(os.path.join(dp, "test_1.tipsy"))["dark"]
    ma, mb = np.unique(a["mass"]), np.unique(b["mass"])
    assert len(ma) == 3 and len(mb) == 3
    np.testing.assert_allclose(ma[2] / ma[1], 27, rtol=1e-5)
    np.testing.assert_allclose(ma[1] / ma[0], 8, rtol=1e-5)
    fa, fb = a[a["mass"] == ma[0]], b[b["mass"] == mb[0]]
    assert len(fa) == len(fb)
    for k in ("pos", "vel", "eps"):
        np.testing.assert_array_equal(fa[k], fb[k])
    assert np.sum(a["mass"] == ma[2]) != np.sum(b["mass"] == mb[2])
