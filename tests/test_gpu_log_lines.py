"""Reference tests whose expected output is lines of the log
(genetIC/tests/<test>/reference.txt, checked as tools/compare.py
compare_outputlogfile does: every reference line appears in the log, time
stamps stripped): the zoom corner of a second zoom opened next to the first
one's edge (ic.hpp:580-588), the border-safety warning (ic.hpp:462-473) and
the gadget particle-type reuse error with its paramfile line, and the
resampled two-level mapper's error when a subsampled base grid's zoomed cells
reach past the zoom grid (twolevelmapper.hpp:600-605), reported at the
subsample command."""
import pytest

from helpers import golden, run_frontend, stage_test

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("test,must_succeed", [
    ("test_01j_two_zooms_edge", True),
    ("test_09a_border_safety", True),
    ("test_10j_impossible_gadget_ptype_request_error", False),
    ("test_15_edge_subsampling_error", False),
])
def test_log_contains_reference_lines(test, must_succeed, tmp_path):
    print("log case", test, flush=True)
    d = stage_test(test, tmp_path)
    rc, err = run_frontend(d)
    if must_succeed:
        assert rc == 0, err
    lines = err.splitlines()
    for want in open(golden(test, "reference.txt")).read().splitlines():
        assert want in lines, "line %r missing from the log:\n%s" % (want, err)

