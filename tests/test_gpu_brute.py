"""GPU parity of the brute-force path through the C ABI, bit-exact against the CPU oracle
(the restatement pinned to the reference by tests/test_oracle_golden.py)."""

import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN
from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu

BRUTE = sorted(glob.glob(os.path.join(GOLDEN, "brute_*.npz")))


def _name(p):
    return os.path.basename(p)[:-4]


def _ctx(P):
    from ridgedetection_amd._lib import Context
    c = Context(P.shape[1], P.dtype)
    c.set_points(P)
    return c


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("path", BRUTE, ids=_name)
def test_phases_match_golden(oracle, path):
    g = np.load(path)
    P, r1, r2 = g["P"], float(g["r1"]), float(g["r2"])
    _say(f"case {_name(path)} n={len(P)} d={P.shape[1]} {P.dtype}")
    with _ctx(P) as c:
        S0, idx = c.choose(r1)
        np.testing.assert_array_equal(idx, g["choose_idx"])
        np.testing.assert_array_equal(S0, g["choose_S"])
        S1, flag, counts, sums, labels, moved = c.evolve_iteration(r1)
        oS1, oflag, ocounts, osums, olabels, omoved = oracle.evolve_iter(P, S0, r1)
        np.testing.assert_array_equal(labels, olabels)
        np.testing.assert_array_equal(counts, g["iter_counts"])
        np.testing.assert_array_equal(sums, g["iter_sums"])
        np.testing.assert_array_equal(S1, g["iter_S"])
        assert flag == int(g["iter_flag"])
        mv = (moved[:, None] >> np.arange(P.shape[1])[None, :]) & 1
        np.testing.assert_array_equal(mv.astype(np.uint8), omoved)
        c.set_chosen(S0)
        SE, it = c.evolve(r1)
        np.testing.assert_array_equal(SE, g["evolve_S"])
        SD, keep = c.decimate(r2)
        np.testing.assert_array_equal(SD, g["decimate_S"])
        np.testing.assert_array_equal(np.nonzero(keep)[0], g["decimate_kept"])
        SR = c.ridge_detection(r1, r2)
        np.testing.assert_array_equal(SR, g["rd_S"])


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_large_cloud_choose_and_iteration(oracle, dt):
    """>32768 points takes the chunked multi-CU choose path."""
    P = synth.helix3d(120000, seed=21, dtype=dt)
    _say(f"case helix120k {np.dtype(dt).name}")
    with _ctx(P) as c:
        S0, idx = c.choose(0.2)
        oS0, oidx = oracle.choose(P, 0.2)
        np.testing.assert_array_equal(idx, oidx)
        np.testing.assert_array_equal(S0, oS0)
        S1, flag, counts, sums, labels, moved = c.evolve_iteration(0.2)
        oS1, oflag, ocounts, osums, olabels, _ = oracle.evolve_iter(P, S0, 0.2)
        np.testing.assert_array_equal(labels, olabels)
        np.testing.assert_array_equal(counts, ocounts)
        np.testing.assert_array_equal(sums, osums)
        np.testing.assert_array_equal(S1, oS1)


@pytest.mark.parametrize("dim", [1, 2, 4, 5, 7, 12])
def test_dims(oracle, dim):
    g = np.random.Generator(np.random.PCG64(dim))
    P = synth.segment_chain(6000, dim, seed=dim, sigma=0.05)
    r1 = 0.15 + 0.05 * dim
    _say(f"case segment_chain d={dim}")
    with _ctx(P) as c:
        SR = c.ridge_detection(r1, 2 * r1)
    oSR, _ = oracle.ridge_detection(P, r1, 2 * r1)
    np.testing.assert_array_equal(SR, oSR)


def test_c1_spiral_capped(oracle):
    """C1: 20k-point 2-D spiral, r=0.45, max 100 evolve iterations per round."""
    P = synth.spiral2d(20000, seed=1)
    _say("case C1 spiral20k capped")
    with _ctx(P) as c:
        SR = c.ridge_detection(0.45, 0.9, max_evolve_iters=100)
        t = c.timers()
    oSR, rounds = oracle.ridge_detection(P, 0.45, 0.9, max_evolve_iter=100)
    np.testing.assert_array_equal(SR, oSR)
    assert t["rounds"] == rounds


def test_decimate_nocopy_rule(oracle):
    g = np.load(os.path.join(GOLDEN, "nocopy_spiral_f32.npz"))
    S = g["S"]
    P = synth.spiral2d(2000, seed=3)
    _say("case decimate no-copy")
    with _ctx(P) as c:
        c.set_chosen(S)
        SD, keep = c.decimate(float(g["r2"]), rule=1)
    np.testing.assert_array_equal(keep, g["alive"])
    assert len(SD) == int(g["count"])


def test_scan_shapes_and_autotuner(oracle):
    """Every evolve-scan shape (samples per lane x threads per block) gives the oracle's
    result; the autotuner picks one of them and reports it."""
    from ridgedetection_amd._lib import Context, OPT_PPT, OPT_SCAN_THREADS
    P = synth.helix3d(30000, seed=12)
    ref, _ = oracle.ridge_detection(P, 0.2, 0.4)
    for ppt in (1, 2, 4):
        for threads in (128, 256, 512):
            _say(f"case scan shape ppt={ppt} threads={threads}")
            with Context(3, np.float32) as c:
                c.set_option(OPT_PPT, ppt)
                c.set_option(OPT_SCAN_THREADS, threads)
                c.set_points(P)
                np.testing.assert_array_equal(c.ridge_detection(0.2, 0.4), ref)
                t = c.timers()
                assert (t["scan_ppt"], t["scan_threads"]) == (ppt, threads)
    _say("case autotuned scan shape")
    with Context(3, np.float32) as c:
        c.set_points(P)
        np.testing.assert_array_equal(c.ridge_detection(0.2, 0.4), ref)
        t = c.timers()
        assert t["scan_ppt"] in (1, 2, 4) and t["scan_threads"] in (128, 256, 512)


@pytest.mark.parametrize("path", BRUTE, ids=_name)
def test_order_chains_match_golden(path):
    """Chain ordering (rd/cpu/brute_force/order.hpp) of the ridge and of the chosen set,
    against the reference's chains stored in the fixtures."""
    g = np.load(path)
    P, r1, r2 = g["P"], float(g["r1"]), float(g["r2"])
    _say(f"case order {_name(path)}")
    with _ctx(P) as c:
        c.choose(r1)
        idx, off = c.order_chains(r1)
        np.testing.assert_array_equal(idx, g["order0_idx"])
        np.testing.assert_array_equal(off, g["order0_off"])
        S = c.ridge_detection(r1, r2)
        np.testing.assert_array_equal(S, g["rd_S"])
        idx, off = c.order_chains(r2)
        np.testing.assert_array_equal(idx, g["order_idx"])
        np.testing.assert_array_equal(off, g["order_off"])


@pytest.mark.parametrize("path", BRUTE, ids=_name)
def test_quality_matches_golden(path):
    """Hausdorff distance and distance statistics against a pattern
    (rd/utils/assessment_quality.hpp), against the reference's values in the fixtures."""
    g = np.load(path)
    if "quality_stats" not in g.files:
        pytest.skip("empty ridge")
    P, S = g["P"], g["rd_S"]
    ideal = np.ascontiguousarray(P[::7])
    _say(f"case quality {_name(path)}")
    with _ctx(P) as c:
        assert c.hausdorff(S, ideal) == float(g["quality_hausdorff"])
        np.testing.assert_array_equal(c.distance_stats(S, ideal), g["quality_stats"].astype(np.float64))
