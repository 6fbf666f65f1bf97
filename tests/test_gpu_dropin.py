"""The header-only drop-in classes (include/rdhip/ridge_detection.hpp), driven from C++ the
way the reference's README and benchmark drive rd::gpu::{bruteForce,tiled}::RidgeDetection,
against the CPU oracle."""

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "dropin_usage")


def _read(path):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
        return np.frombuffer(f.read(), dtype=np.float32).reshape(n, 3)


def test_dropin_classes(oracle, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/dropin_usage is not built (run __graft_entry__.build())"
    P = synth.helix3d(30000, seed=77)
    r1, r2 = 0.2, 0.4
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([len(P), 3], dtype=np.int64).tobytes())
        f.write(np.array([r1, r2], dtype=np.float64).tobytes())
        f.write(P.tobytes())
    print("case dropin helix30k", flush=True)
    r = subprocess.run([EXE, str(inp), str(tmp_path / "out")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    brute = _read(tmp_path / "out_brute.bin")
    tiled = _read(tmp_path / "out_tiled.bin")
    ref_brute, _ = oracle.ridge_detection(P, r1, r2)
    np.testing.assert_array_equal(brute, ref_brute)
    np.testing.assert_array_equal(tiled, oracle.tiled(5, P, r1, r2, 2500, [4, 4, 4]))
    # SamplesGenerator<float>::spiral3D into dP_: the library generator's samples (sigma
    # scaled by 1/sqrt(3), seed 1234), and ridge detection on them equals the oracle's
    from ridgedetection_amd import _lib as L
    gen = _read(tmp_path / "out_genpts.bin")
    exp = L.generate(L.GEN_SPIRAL3D, 20000, 3, params=[np.float32(4.0) * np.float32(3.14159265358979), np.float32(0.3)],
                     sigma=float(np.float32(0.05)) / np.sqrt(3.0), seed=1234)
    np.testing.assert_array_equal(gen, exp)
    np.testing.assert_array_equal(_read(tmp_path / "out_genrd.bin"), oracle.ridge_detection(gen, 0.3, 0.6)[0])
