"""Text sample files against the reference's own writer and readers.

tests/golden/samples_{f32,f64}.txt (with the reference generators' descriptor line) and
samples_plain_{f32,f64}.txt (samples only) were written by rd::Samples<T>::saveToFile, and
samples_io.npz holds the values rd::Samples<T>::loadGeneratedSamples / loadFromFile read back
from them (rd/utils/samples_set.hpp:147-291, 342-372, 474-503; tests/golden/make_golden.py)."""

import os

import numpy as np
import pytest

from conftest import GOLDEN
from ridgedetection_amd import samples, synth

G = np.load(os.path.join(GOLDEN, "samples_io.npz"))
DESC = " --dim=3 --np=50 --a=31.4159 --b=0.15  --s=0.05\n"


def _text(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


@pytest.mark.parametrize("sfx", ["f32", "f64"])
def test_writer_is_byte_identical_to_the_reference(tmp_path, sfx):
    P = G[f"P_{sfx}"]
    samples.save_samples(str(tmp_path / "a.txt"), P, comment=DESC)
    assert (tmp_path / "a.txt").read_text() == _text(f"samples_{sfx}.txt")
    samples.save_samples(str(tmp_path / "b.txt"), P)
    assert (tmp_path / "b.txt").read_text() == _text(f"samples_plain_{sfx}.txt")


@pytest.mark.parametrize("sfx,dt", [("f32", np.float32), ("f64", np.float64)])
def test_readers_match_the_reference(sfx, dt):
    P, d = samples.load_generated_samples(os.path.join(GOLDEN, f"samples_{sfx}.txt"), dtype=dt)
    np.testing.assert_array_equal(P, G[f"loadgen_{sfx}"])
    assert d == {"dim": 3.0, "np": 50.0, "a": 31.4159, "b": 0.15, "s": 0.05}
    P2, _ = samples.load_generated_samples(os.path.join(GOLDEN, f"samples_{sfx}.txt"), dtype=dt, downgrade_to_dim=2)
    np.testing.assert_array_equal(P2, G[f"loadgen_dim2_{sfx}"])
    np.testing.assert_array_equal(samples.load_samples(os.path.join(GOLDEN, f"samples_plain_{sfx}.txt"), dtype=dt),
                                  G[f"load_plain_{sfx}"])
    # a descriptor file read as plain samples gives the same values
    np.testing.assert_array_equal(samples.load_samples(os.path.join(GOLDEN, f"samples_{sfx}.txt"), dtype=dt), P)


def test_format_lines():
    lines = _text("samples_plain_f32.txt").split("\n")
    assert lines[0] == "    5.5677     -3.8297     25.5547  "     # std::fixed, precision 4, setw(10), "  "
    lines = _text("samples_plain_f64.txt").split("\n")
    assert lines[0] == "  5.567689060248   -3.829747390247   25.554715082365  "   # precision 12, setw(16)


@pytest.mark.parametrize("dt,tol", [(np.float32, 0.5e-4), (np.float64, 0.5e-12)])
def test_round_trip_error_is_half_a_unit_of_the_last_decimal(tmp_path, dt, tol):
    P = synth.helix3d(500, seed=3, dtype=dt)
    path = str(tmp_path / "h.txt")
    samples.save_samples(path, P)
    Q = samples.load_samples(path, dtype=dt)
    assert Q.shape == P.shape
    err = np.abs(Q.astype(np.float64) - P.astype(np.float64)).max()
    # printing rounds to the last decimal; reading rounds once more to the value type
    ulp = np.spacing(np.abs(P).max().astype(dt)).astype(np.float64)
    assert err <= tol + ulp


def test_float_parsing_rounds_once():
    """1 + 2^-24 is the midpoint of the floats 1 and 1 + 2^-23.  A decimal just above it rounds
    (once) up to 1 + 2^-23; through a double it lands on the midpoint and ties-to-even gives 1."""
    tok = "1.000000059604644775390625000000001"
    got = samples._parse([tok], np.float32)[0]
    assert got == np.nextafter(np.float32(1.0), np.float32(2.0))
    assert np.float32(float(tok)) == np.float32(1.0)
    below = "1.000000059604644775390624999999999"
    assert samples._parse([below], np.float32)[0] == np.float32(1.0)


def test_descriptor_errors_follow_the_reference(tmp_path):
    path = tmp_path / "d.txt"
    path.write_text("% --np=2 --a=1 --s=0.1\n 1.0 2.0\n 3.0 4.0\n")
    with pytest.raises(ValueError, match='Missing "dim"'):
        samples.load_generated_samples(str(path))
    path.write_text("% --dim=2 --np=2 --a=1 --s=0.1\n 1.0 2.0\n 3.0 4.0\n")
    with pytest.raises(ValueError, match='Missing "b"'):
        samples.load_generated_samples(str(path))
    path.write_text("% --dim=2 --np=3 --a=1 --b=2 --s=0.1\n 1.0 2.0\n 3.0 4.0\n")
    with pytest.raises(ValueError, match="Incorrect points count"):
        samples.load_generated_samples(str(path))


def test_blank_lines_and_comments_are_skipped(tmp_path):
    path = tmp_path / "c.txt"
    path.write_text("%comment\n\n 1.0 2.0 3.0\n\n4 5 6\n")
    Q = samples.load_samples(str(path))
    np.testing.assert_array_equal(Q, np.array([[1, 2, 3], [4, 5, 6]], dtype=np.float32))
