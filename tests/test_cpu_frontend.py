"""The paramfile front-end's parser contract (tools/parser.hpp:40-211):
errors are reported as 'Error "<what>" on line N ("<line>")' with exit
status 1, before any command that needs the GPU runs."""
import pytest

from helpers import run_frontend, stage_test


@pytest.mark.parametrize("text,line,what", [
    ("outdir ./\nOm 0.3\nno_such_command 1\n", 3, "Unknown command"),
    ("Om\n", 1, "Insufficient arguments"),
    ("Om 0.3 0.4\n", 1, "Too many arguments"),
    ("Om 0.3 # a comment\n% also a comment\nOm abc\n", 3, "Error while parsing arguments"),
    ("centre 1 2\n", 1, "Insufficient arguments"),
    ("centre 1 2 \n", 1, "Error while parsing arguments"),
    ("OUTFORMAT lemon\n", 1, "Error while parsing arguments"),
    ("outformat 9\n", 1, "Unknown output format"),
    ("select_sphere 3\n", 1, "No grid has yet been initialised"),
    ("gadget_particle_type 7\n", 1, "Gadget particle type must be between 0 and 5"),
    ("gadget_flagged_particle_type 6\n", 1, "Gadget particle type must be between 0 and 5"),
    ("camb does_not_exist.dat\n", 1, "File does_not_exist.dat not found"),
])
def test_parser_errors(tmp_path, text, line, what):
    d = stage_test("parser_case", tmp_path, text)
    rc, err = run_frontend(d)
    assert rc == 1
    bad = text.splitlines()[line - 1]
    assert 'Error "%s" on line %d ("%s")' % (what, line, bad) in err, err


def test_comments_and_blank_lines_are_ignored(tmp_path):
    d = stage_test("parser_ok", tmp_path, "# header\n\nOm 0.3 # trailing\n%done\noutdir ./\n")
    rc, err = run_frontend(d)
    assert rc == 0, err
    echoed = open(d + "/IC_output.params").read().split("\n")
    assert echoed[0].split() == ["om", "0.3"]
    assert echoed[1].split() == ["outdir", "./"]
