"""FP32 results against FP64 on the same coordinates.

Every FP32 ridge is bit-identical to the FP32 oracle (the other GPU tests); these tests state
how far that FP32 result is from the FP64 one:

* end to end (point counts and the Hausdorff distance between the FP32 and the FP64 ridges,
  rd::RDAssessmentQuality::hausdorffDistance, rd/utils/assessment_quality.hpp:91-107, with the
  FP32 ridge as the test set and the FP64 ridge as the pattern) for the 20k-point spiral,
  reduced 3-D helix and 6-D Bezier tilings and a sample of batched clouds.  The ridge is a
  discontinuous function of the input (a near tie in the nearest-centre test or a count at the
  decimation threshold can flip), so this distance is reported, not bounded: the FP32 and FP64
  ridges of the reference itself differ by as much;
* at frozen memberships (the evolve map is then affine, so rounding is the only difference):
  20 FP32 iterations on the GPU against the same iterations replayed in FP64 by the oracle with
  the GPU's own labels and moved coordinates, max |S32 - S64| <= 1e-5 * extent."""

import numpy as np
import pytest

from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _ctx(P):
    from ridgedetection_amd._lib import Context
    c = Context(P.shape[1], P.dtype)
    c.set_points(P)
    return c


def _extent(P):
    return float(np.max(P.max(axis=0) - P.min(axis=0)))


def _report(name, P32, S32, S64, oracle, ref32, ref64):
    np.testing.assert_array_equal(S32, ref32)
    np.testing.assert_array_equal(S64, ref64)
    with _ctx(S64) as c:
        h = c.hausdorff(S32.astype(np.float64), S64) if len(S32) and len(S64) else -1.0
    ext = _extent(P32)
    _say(f"  {name}: FP32 {len(S32)} points, FP64 {len(S64)} points, "
         f"Hausdorff(S32, S64) = {h:.3e} = {h / ext:.3e} x extent")
    return h


def test_spiral20k_brute(oracle):
    P = synth.spiral2d(20000, seed=1)
    _say("case precision spiral20k brute force")
    out = []
    for dt in (np.float32, np.float64):
        with _ctx(P.astype(dt)) as c:
            out.append(c.ridge_detection(0.45, 0.9, max_evolve_iters=100))
    ref32, _ = oracle.ridge_detection(P, 0.45, 0.9, max_evolve_iter=100)
    ref64, _ = oracle.ridge_detection(P.astype(np.float64), 0.45, 0.9, max_evolve_iter=100)
    _report("spiral20k", P, out[0], out[1], oracle, ref32, ref64)


@pytest.mark.parametrize("alg", [0, 5, 6])
def test_helix200k_tiled(oracle, alg):
    P = synth.helix3d(200000, seed=9)
    _say(f"case precision helix200k tiled alg {alg}")
    out = []
    for dt in (np.float32, np.float64):
        with _ctx(P.astype(dt)) as c:
            out.append(c.tiled(alg, 0.2, 0.4, 2500, [8, 8, 8]))
    ref32 = oracle.tiled(alg, P, 0.2, 0.4, 2500, [8, 8, 8])
    ref64 = oracle.tiled(alg, P.astype(np.float64), 0.2, 0.4, 2500, [8, 8, 8])
    _report(f"helix200k alg {alg}", P, out[0], out[1], oracle, ref32, ref64)


@pytest.mark.parametrize("alg", [0, 5])
def test_bezier6d_tiled(oracle, alg):
    P = synth.bezier_union(60000, dim=6, seed=3)
    _say(f"case precision bezier6d 60k tiled alg {alg}")
    out = []
    for dt in (np.float32, np.float64):
        with _ctx(P.astype(dt)) as c:
            out.append(c.tiled(alg, 0.08, 0.16, 1000, [4] * 6))
    ref32 = oracle.tiled(alg, P, 0.08, 0.16, 1000, [4] * 6)
    ref64 = oracle.tiled(alg, P.astype(np.float64), 0.08, 0.16, 1000, [4] * 6)
    _report(f"bezier6d alg {alg}", P, out[0], out[1], oracle, ref32, ref64)


def test_batched_clouds_sample(oracle):
    """Eight clouds of the batched workload's kinds (r = 4 sigma), one FP32 and one FP64 batch."""
    from ridgedetection_amd._lib import Context
    clouds = synth.batched_clouds(8, seed=4, max_pts=40000)
    _say("case precision batched clouds")
    res = {}
    for dt in (np.float32, np.float64):
        with Context(3, dt) as c:
            res[dt] = c.batch([p.astype(dt) for p, _, _ in clouds], [r1 for _, r1, _ in clouds],
                              [r2 for _, _, r2 in clouds])
    for i, (P, r1, r2) in enumerate(clouds):
        ref32, _ = oracle.ridge_detection(P, r1, r2)
        ref64, _ = oracle.ridge_detection(P.astype(np.float64), r1, r2)
        _report(f"cloud {i} ({len(P)} x {P.shape[1]}, r {r1:.3g})", P, res[np.float32][i], res[np.float64][i],
                oracle, ref32, ref64)


@pytest.mark.parametrize("case", ["helix200k", "spiral20k", "bezier6d"])
def test_frozen_membership_evolve(oracle, case):
    """Rounding alone: the FP32 iterations against FP64 arithmetic on the same memberships."""
    if case == "helix200k":
        P, r = synth.helix3d(200000, seed=11), 0.2
    elif case == "spiral20k":
        P, r = synth.spiral2d(20000, seed=1), 0.45
    else:
        P, r = synth.bezier_union(40000, dim=6, seed=5), 0.08
    T = 20
    _say(f"case precision frozen-membership evolve {case}, {T} iterations")
    with _ctx(P) as c:
        S0, _ = c.choose(r)
        labels, moved = [], []
        for _ in range(T):
            S, _, _, _, lab, mv = c.evolve_iteration(r)
            labels.append(lab)
            moved.append(((mv[:, None] >> np.arange(P.shape[1])[None, :]) & 1).astype(np.uint8))
    S64 = oracle.evolve_replay(P.astype(np.float64), S0.astype(np.float64), np.stack(labels), np.stack(moved))
    ext = _extent(P)
    err = float(np.max(np.abs(S.astype(np.float64) - S64))) / ext
    _say(f"  K = {len(S0)}: max |S32 - S64| = {err:.3e} x extent (bound 1e-5)")
    assert err <= 1e-5
