"""The reference's particle-mapper tests (genetIC/tests/mapper_test_*, run by
run_mapper_tests.sh): an id file is loaded through the output format's
particle mapper and dump_idfile must write exactly the reference's list.
mapper_test_01a_zoom goes through the two-level mapper of tipsy output
(twolevelmapper.hpp:357-513); the 02 cases through the grafic hierarchy,
a subsampled copy of the base grid and a subsampled intermediate of the
zoom included (graficmapper.hpp:75-184).  The cases that replay a second
paramfile (mapper_relative_to) and stray particles are not covered."""
import os
import shutil

import numpy as np
import pytest

from helpers import golden, run_frontend, stage_test

pytestmark = pytest.mark.gpu


def _ids(path):
    return np.loadtxt(path, dtype=np.int64, ndmin=1)


@pytest.mark.parametrize("test", ["mapper_test_01a_zoom", "mapper_test_02a_grafic",
                                  "mapper_test_02b_grafic_zoom_to_coarse"])
def test_dump_idfile_matches_reference(test, tmp_path):
    print("mapper case", test, flush=True)
    d = stage_test(test, tmp_path)
    # 01a reads the reference list back as its second id file
    shutil.copyfile(golden(test, "reference.txt"), os.path.join(d, "reference.txt"))
    rc, err = run_frontend(d)
    assert rc == 0, err
    got, want = _ids(os.path.join(d, "output.txt")), _ids(golden(test, "reference.txt"))
    np.testing.assert_array_equal(got, want)


def test_grafic_ids_out_of_range_are_refused(tmp_path):
    test = "mapper_test_02a_grafic"
    print("mapper ids out of range", flush=True)
    d = stage_test(test, tmp_path)
    # the hierarchy holds 32^3 + 64^3 + 32^3 + 64^3 cells
    with open(os.path.join(d, "test_ids.txt"), "w") as f:
        f.write("1\n%d\n" % (2 * 32 ** 3 + 2 * 64 ** 3))
    rc, err = run_frontend(d)
    assert rc != 0
    assert "Ran out of grids when interpreting grafic cell IDs" in err
