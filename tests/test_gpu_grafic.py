"""Grafic output and grafic masks of the paramfile front-end against the
reference's fixtures (reference_grafic, reference_mask).  Tolerances as
tools/compare.py: positions, velocities and deltab to 4 decimals (float
output), particle IDs exactly, masks to 4 decimals; the headers and the
refmap / pvar files are compared exactly."""
import glob
import os

import numpy as np
import pytest

from helpers import assert_decimal, golden, golden_array, run_frontend, stage_test

pytestmark = pytest.mark.gpu

NAMES = ["ic_velcx", "ic_velcy", "ic_velcz", "ic_poscx", "ic_poscy", "ic_poscz", "ic_deltab", "ic_refmap",
         "ic_pvar_00001", "ic_particle_ids"]


def read_grafic(path):
    """A grafic file: the 44-byte header record, then one record per z plane
    (x fastest); returns (header bytes, array [z, y, x])."""
    raw = open(path, "rb").read()
    n0 = np.frombuffer(raw[:4], np.int32)[0]
    assert n0 == 44
    hdr = raw[4:48]
    assert np.frombuffer(raw[48:52], np.int32)[0] == 44
    nx, ny, nz = np.frombuffer(hdr[:12], np.int32)
    pos = 52
    planes = []
    dt = np.uint64 if path.endswith("ids") else np.float32
    for _ in range(nz):
        ln = np.frombuffer(raw[pos:pos + 4], np.int32)[0]
        assert ln == nx * ny * np.dtype(dt).itemsize
        planes.append(np.frombuffer(raw[pos + 4:pos + 4 + ln], dt).reshape(ny, nx))
        assert np.frombuffer(raw[pos + 4 + ln:pos + 8 + ln], np.int32)[0] == ln
        pos += ln + 8
    assert pos == len(raw)
    return hdr, np.array(planes)


def compare_grafic(test_dir, ref_dir):
    refs = sorted(glob.glob(os.path.join(ref_dir, "*.grafic_*")))
    gots = sorted(glob.glob(os.path.join(test_dir, "*.grafic_*")))
    assert [os.path.basename(p) for p in refs] == [os.path.basename(p) for p in gots]
    for r, g in zip(refs, gots):
        for name in NAMES:
            hr, ar = read_grafic(os.path.join(r, name))
            hg, ag = read_grafic(os.path.join(g, name))
            assert hr[:12] == hg[:12], (g, name)
            np.testing.assert_allclose(np.frombuffer(hg[12:], np.float32), np.frombuffer(hr[12:], np.float32),
                                       rtol=1e-6, err_msg=g + " " + name)
            what = os.path.basename(g) + "/" + name
            if name == "ic_particle_ids":
                np.testing.assert_array_equal(ag, ar, err_msg=what)
            elif name in ("ic_refmap", "ic_pvar_00001"):
                np.testing.assert_array_equal(ag, ar, err_msg=what)
            else:
                assert_decimal(ag, ar, 4, what)


def _run(test, tmp_path):
    print("grafic case", test, flush=True)
    d = stage_test(test, tmp_path)
    rc, err = run_frontend(d)
    assert rc == 0, err
    return d


@pytest.mark.parametrize("test", ["test_12_grafic", "test_08e_mask_multiple_virtual_grids",
                                  "test_08f_mask_multiple_supersample"])
def test_grafic_output_matches_reference(tmp_path, test):
    """test_12: subsample 2 and a x4 zoom (levels 4, 8, 16, 32, the 16 an
    average of the zoom grid).  test_08e: subsample 4, a x4 zoom and the
    output centred on (0, 0, 0).  test_08f: supersample 4 (two levels
    interpolated from the zoom grid)."""
    d = _run(test, tmp_path)
    compare_grafic(d, golden(test, "reference_grafic"))


@pytest.mark.parametrize("test,nlev", [
    ("test_08a_mask_lagrange_volume", 3),
    ("test_08c_centering_masks", 3),
    ("test_08d_mask_flag_interaction", 3),
    ("test_08e_mask_multiple_virtual_grids", 5),
    ("test_08b_mask_virtual_architecture", 6),
    ("test_08f_mask_multiple_supersample", 4),
])
def test_grafic_masks_match_reference(tmp_path, test, nlev):
    d = _run(test, tmp_path)
    for lev in range(nlev):
        got = np.load(os.path.join(d, "mask-%d.npy" % lev))
        ref = golden_array(golden(test, "reference_mask", "mask-%d.npy" % lev))
        if ref is not None:
            assert_decimal(got, ref, 4, "%s mask-%d" % (test, lev))
        else:
            # the finest level carries no flags: every cell is in the mask
            assert (got == 1.0).all()
