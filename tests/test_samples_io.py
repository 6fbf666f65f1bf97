"""Text sample files in the reference's format (rd/utils/samples_set.hpp:342-372, 444-462)."""

import numpy as np

from ridgedetection_amd import samples, synth


def test_write_format_and_round_trip(tmp_path):
    P = np.array([[1.0, -2.5], [123.45678, 0.00004]], dtype=np.float64)
    path = tmp_path / "s.txt"
    samples.save_samples(str(path), P, comment="-dim=2 -np=2 -a=1 -b=2 -s=0.1")
    lines = path.read_text().split("\n")
    assert lines[0] == "%-dim=2 -np=2 -a=1 -b=2 -s=0.1"
    # std::fixed, precision 4, setw(10), right-aligned, then two spaces
    assert lines[1] == "    1.0000     -2.5000  "
    assert lines[2] == "  123.4568      0.0000  "
    Q = samples.load_samples(str(path), dtype=np.float64)
    np.testing.assert_array_equal(Q, np.round(P, 4))
    d = samples.read_descriptor(str(path))
    assert d == {"dim": 2.0, "np": 2.0, "a": 1.0, "b": 2.0, "s": 0.1}


def test_round_trip_is_within_half_a_unit_of_the_fourth_decimal(tmp_path):
    P = synth.helix3d(500, seed=3, dtype=np.float64)
    path = tmp_path / "h.txt"
    samples.save_samples(str(path), P)
    Q = samples.load_samples(str(path), dtype=np.float64)
    assert Q.shape == P.shape
    assert np.abs(Q - P).max() <= 0.5e-4 + 1e-12


def test_blank_lines_and_comments_are_skipped(tmp_path):
    path = tmp_path / "c.txt"
    path.write_text("%comment\n\n 1.0 2.0 3.0\n\n4 5 6\n")
    Q = samples.load_samples(str(path))
    np.testing.assert_array_equal(Q, np.array([[1, 2, 3], [4, 5, 6]], dtype=np.float32))
