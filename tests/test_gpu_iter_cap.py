"""GPU parity when the evolve iteration cap binds. The reference's evolve loop
(rd/cpu/brute_force/evolve.hpp:292-342) runs until no chosen point moves; the cap (0 = none, the
reference behaviour) stops it after max_evolve_iters iterations whether or not points still move,
as the 20k-point spiral workload asks with its cap of 100 (which that cloud never reaches: its
first round converges in 44). Here: caps below the iteration count the clouds need to converge, through the segment-local and the scan
paths, FP32 and FP64, batches, tilings and the owned-sample variant, bit-exact against the
oracle run with the same cap. Each case first checks on the oracle that its cap changes the
result, so a cap that no longer binds fails here instead of passing unnoticed."""

import numpy as np
import pytest

from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _ctx(P, local=None):
    from ridgedetection_amd._lib import Context, OPT_LOCAL
    c = Context(P.shape[1], P.dtype)
    if local is not None:
        c.set_option(OPT_LOCAL, local)
    c.set_points(P)
    return c


def _binds(a, b):
    return len(a) != len(b) or not np.array_equal(a, b)


@pytest.mark.parametrize("local", [None, 0], ids=["default", "scan"])
@pytest.mark.parametrize("cap", [1, 2, 5, 17])
def test_spiral_capped(oracle, cap, local):
    """20k-point 2-D spiral at r = 0.45 (its first round needs 44 iterations)."""
    P = synth.spiral2d(20000, seed=1)
    free, _ = oracle.ridge_detection(P, 0.45, 0.9)
    oSR, rounds = oracle.ridge_detection(P, 0.45, 0.9, max_evolve_iter=cap)
    assert _binds(oSR, free)
    _say(f"case spiral20k cap={cap} local={local}")
    with _ctx(P, local) as c:
        SR = c.ridge_detection(0.45, 0.9, max_evolve_iters=cap)
        assert c.timers()["rounds"] == rounds
    np.testing.assert_array_equal(SR, oSR)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_helix_capped(oracle, dtype):
    """60k-point helix at r = 0.2 with a cap of 3, in the scan path."""
    P = synth.helix3d(60000, seed=8, dtype=dtype)
    free, _ = oracle.ridge_detection(P, 0.2, 0.4)
    oSR, _ = oracle.ridge_detection(P, 0.2, 0.4, max_evolve_iter=3)
    assert _binds(oSR, free)
    _say(f"case helix60k cap=3 {np.dtype(dtype).name}")
    with _ctx(P, 0) as c:
        np.testing.assert_array_equal(c.ridge_detection(0.2, 0.4, max_evolve_iters=3), oSR)


def test_batch_capped(oracle):
    """A batch of clouds with one cap: every segment stops at its own count of iterations."""
    from ridgedetection_amd._lib import Context, OPT_LOCAL
    g = np.random.default_rng(515)
    clouds = []
    for i in range(8):
        n = int(g.integers(3000, 25000))
        P = (synth.spiral2d(n, seed=600 + i) * float(g.uniform(0.5, 4.0))).astype(np.float32)
        r1 = float(g.uniform(0.015, 0.03)) * float(np.max(P.max(0) - P.min(0)))
        clouds.append((P, r1, 2 * r1))
    expected = [oracle.ridge_detection(P, r1, r2, max_evolve_iter=4)[0] for P, r1, r2 in clouds]
    assert any(_binds(e, oracle.ridge_detection(P, r1, r2)[0]) for e, (P, r1, r2) in zip(expected, clouds))
    for local in (None, 0):
        _say(f"case batch cap=4 local={local}")
        with Context(2, np.float32) as c:
            if local is not None:
                c.set_option(OPT_LOCAL, local)
            got = c.batch([p for p, _, _ in clouds], [a for _, a, _ in clouds], [b for _, _, b in clouds],
                          max_evolve_iters=4)
        for G, E in zip(got, expected):
            np.testing.assert_array_equal(G, E)


@pytest.mark.parametrize("alg", [6, 1])
def test_tiled_capped(oracle, alg):
    """Tiled 30k-point helix with a cap of 2 per tile."""
    P = synth.helix3d(30000, seed=35)
    free = oracle.tiled(alg, P, 0.2, 0.4, 3000, [4, 4, 4])
    expected = oracle.tiled(alg, P, 0.2, 0.4, 3000, [4, 4, 4], max_evolve_iter=2)
    assert _binds(expected, free)
    _say(f"case tiled alg={alg} cap=2")
    with _ctx(P) as c:
        np.testing.assert_array_equal(c.tiled(alg, 0.2, 0.4, 3000, [4, 4, 4], max_evolve_iters=2), expected)


def test_own_capped(oracle):
    """Owned samples (the first half of the cloud) with a cap of 3."""
    P = synth.spiral2d(16000, seed=12)
    n_own = len(P) // 2
    free, _ = oracle.ridge_detection_own(P, n_own, 0.45, 0.9)
    expected, _ = oracle.ridge_detection_own(P, n_own, 0.45, 0.9, max_evolve_iter=3)
    assert _binds(expected, free)
    _say("case own cap=3")
    with _ctx(P) as c:
        np.testing.assert_array_equal(c.ridge_detection_own(n_own, 0.45, 0.9, max_evolve_iters=3), expected)
