"""The header-only drop-ins (include/rdhip/ridge_detection.hpp, include/rdhip/cpu_ridge_detection.hpp),
driven from C++ the way the reference's README, benchmarks and kernel tests drive
rd::gpu::{bruteForce,tiled}::RidgeDetection, the DeviceChoose / DeviceEvolve / DeviceDecimate
facades and the CPU entry points rd::RidgeDetection<T> / rd::tiled::tiledRidgeDetection, against
the CPU oracle.  Layouts ROW_MAJOR and (pitched) COL_MAJOR on both sides."""

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "dropin_usage")


def _read(prefix, case):
    with open(f"{prefix}_{case}.bin", "rb") as f:
        n = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
        return np.frombuffer(f.read(), dtype=np.float32).reshape(n, 3)


def _raw(prefix, case):
    with open(f"{prefix}_{case}.bin", "rb") as f:
        n = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
        return n, f.read()


def test_dropin_classes(oracle, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/dropin_usage is not built (run __graft_entry__.build())"
    P = synth.helix3d(30000, seed=77)
    r1, r2 = 0.2, 0.4
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([len(P), 3], dtype=np.int64).tobytes())
        f.write(np.array([r1, r2], dtype=np.float64).tobytes())
        f.write(P.tobytes())
    out = str(tmp_path / "out")
    print("case dropin helix30k", flush=True)
    r = subprocess.run([EXE, str(inp), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)

    # GPU classes, ROW_MAJOR and pitched COL_MAJOR buffers
    ref_brute, _ = oracle.ridge_detection(P, r1, r2)
    for case in ("brute", "brute_col", "brute_rowcol"):
        np.testing.assert_array_equal(_read(out, case), ref_brute, err_msg=case)
    ref_tiled = oracle.tiled(5, P, r1, r2, 2500, [4, 4, 4])
    np.testing.assert_array_equal(_read(out, "tiled"), ref_tiled)
    np.testing.assert_array_equal(_read(out, "tiled_col"), ref_tiled)   # host ROW_MAJOR either way
    assert _raw(out, "soa_throws")[0] == 3   # SOA results and SOA input both rejected

    # device facades
    S0, _ = oracle.choose(P, r1)
    np.testing.assert_array_equal(_read(out, "dchoose"), S0)
    SE, _ = oracle.evolve(P, S0, r1)
    np.testing.assert_array_equal(_read(out, "devolve"), SE)
    K, raw = _raw(out, "devolve_stats")
    counts = np.frombuffer(raw[:4 * K], dtype=np.int32)
    sums = np.frombuffer(raw[4 * K:], dtype=np.float32).reshape(K, 3)
    _, _, rc, rs, _, _ = oracle.evolve_iter(P, SE, r1)   # sums and counts at the fixed point
    np.testing.assert_array_equal(counts, rc)
    np.testing.assert_array_equal(sums, rs)
    np.testing.assert_array_equal(_read(out, "devolve_nb"), SE)   # own + neighbour halves = whole cloud
    SD, _ = oracle.decimate(SE, r2)
    np.testing.assert_array_equal(_read(out, "ddecimate"), SD)
    np.testing.assert_array_equal(_read(out, "ddecimate_col"), SD)

    # CPU entry points
    np.testing.assert_array_equal(_read(out, "cpu_rd"), ref_brute)
    m, raw = _raw(out, "cpu_chains")
    words = np.frombuffer(raw, dtype=np.int64)
    offsets, order = words[:m + 1], words[m + 1:]
    ref_order, ref_off = oracle.order(ref_brute, r2)
    np.testing.assert_array_equal(offsets, ref_off)
    np.testing.assert_array_equal(order, ref_order)
    ref_own, _ = oracle.ridge_detection_own(P, len(P) // 2, r1, r2)
    np.testing.assert_array_equal(_read(out, "cpu_rd_own"), ref_own)
    np.testing.assert_array_equal(_read(out, "cpu_tiled"), oracle.tiled(5, P, r1, r2, 2500, [4, 4, 4], refinement=True))

    # SamplesGenerator<float>::spiral3D into dP_: the library generator's samples (sigma
    # scaled by 1/sqrt(3), seed 1234), and ridge detection on them equals the oracle's
    from ridgedetection_amd import _lib as L
    gen = _read(out, "genpts")
    exp = L.generate(L.GEN_SPIRAL3D, 20000, 3, params=[np.float32(4.0) * np.float32(3.14159265358979), np.float32(0.3)],
                     sigma=float(np.float32(0.05)) / np.sqrt(3.0), seed=1234)
    np.testing.assert_array_equal(gen, exp)
    np.testing.assert_array_equal(_read(out, "genrd"), oracle.ridge_detection(gen, 0.3, 0.6)[0])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ridge_detection_own_matches_oracle(oracle, dtype):
    """Choose over the first sample set only, evolve over all sets: the sample-list overload
    of rd::RidgeDetection<T>::ridgeDetection through the C ABI."""
    from ridgedetection_amd._lib import Context
    P = synth.spiral2d(12000, seed=5, dtype=dtype)
    print(f"case ridge detection own {np.dtype(dtype).name}", flush=True)
    for n_own in (4000, 12000):
        with Context(2, dtype) as c:
            c.set_points(P)
            S = c.ridge_detection_own(n_own, 0.45, 0.9)
        np.testing.assert_array_equal(S, oracle.ridge_detection_own(P, n_own, 0.45, 0.9)[0])
