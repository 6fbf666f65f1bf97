"""GPU parity of the expanded-form evolve screens (RDHIP_OPT_SCREEN 1: |s-o|^2 - 2 (s-o).(p-o) with a
rigorous error bound and an exact pass over the near-tie window; 3: the same values from the matrix
cores, FP32) against the oracle and against the direct screen (RDHIP_OPT_SCREEN 0), bit for bit:
every scan shape, FP32 and FP64, dimensions 1 to 12, clouds far from the origin, many near ties (a
small radius over a large cloud), exact ties (duplicated chosen points: the reference's first index
wins), batches and tilings, and the autotuned form."""

import numpy as np
import pytest

from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu

SHAPES = [(1, 128), (2, 256), (4, 256), (1, 512), (2, 512), (4, 512)]
SCREENS = [1, 3]


def _form(screen, dtype):
    """screen_form the timers report for an RDHIP_OPT_SCREEN value (the matrix cores are FP32 only)."""
    return {0: 0, 1: 1, 3: 2 if np.dtype(dtype) == np.float32 else 1}[screen]


def _say(msg):
    print(msg, flush=True)


def _ctx(P, screen, shape=None):
    from ridgedetection_amd._lib import Context, OPT_PPT, OPT_SCAN_THREADS, OPT_SCREEN
    c = Context(P.shape[1], P.dtype)
    c.set_option(OPT_SCREEN, screen)
    if shape:
        c.set_option(OPT_PPT, shape[0])
        c.set_option(OPT_SCAN_THREADS, shape[1])
    c.set_points(P)
    return c


def _check_iterations(oracle, c, P, S, r, steps):
    """steps iterations from the chosen set S, each against the oracle; returns the labels."""
    c.set_chosen(S)
    out = []
    for _ in range(steps):
        oS, oflag, ocounts, osums, olabels, omoved = oracle.evolve_iter(P, S, r)
        S1, flag, counts, sums, labels, moved = c.evolve_iteration(r)
        np.testing.assert_array_equal(labels, olabels)
        np.testing.assert_array_equal(counts, ocounts)
        np.testing.assert_array_equal(sums, osums)
        np.testing.assert_array_equal(S1, oS)
        assert flag == oflag
        out.append(labels)
        S = S1
    return out


@pytest.mark.parametrize("screen", SCREENS)
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_every_scan_shape(oracle, shape, dtype, screen):
    P = synth.helix3d(30011, seed=5, dtype=dtype)
    _say(f"case screen {screen} shape {shape} {np.dtype(dtype).name}")
    S, _ = oracle.choose(P, 0.2)
    with _ctx(P, screen, shape) as c:
        _check_iterations(oracle, c, P, S, 0.2, 3)
        assert c.timers()["screen_form"] == _form(screen, dtype)


def _cloud(dim, n, seed, dtype):
    if dim == 1:
        rng = np.random.default_rng(seed)
        return np.sort(rng.uniform(0, 10, (n, 1))).astype(dtype)
    if dim == 2:
        return synth.spiral2d(n, seed=seed).astype(dtype)
    if dim == 3:
        return synth.helix3d(n, seed=seed, dtype=dtype)
    return synth.bezier_union(n, dim=dim, seed=seed).astype(dtype)


@pytest.mark.parametrize("dim", [1, 2, 3, 4, 6, 12])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dims(oracle, dim, dtype):
    """Checked iterations, then whole ridge detections: both expanded screens == direct screen ==
    oracle."""
    P = _cloud(dim, 16000 if dim <= 4 else 8000, 70 + dim, dtype)
    r1 = {1: 0.05, 2: 0.45, 3: 0.2, 4: 0.25, 6: 0.3, 12: 0.45}[dim]
    S, _ = oracle.choose(P, r1)
    SO = oracle.ridge_detection(P, r1, 2 * r1)[0]
    for screen in SCREENS:
        _say(f"case screen {screen} dim={dim} {np.dtype(dtype).name}")
        with _ctx(P, screen) as c:
            _check_iterations(oracle, c, P, S, r1, 4)
            assert c.timers()["screen_form"] == _form(screen, dtype)
            # (a cloud this small evolves in the segment-local path: one workgroup, no scan)
            np.testing.assert_array_equal(c.ridge_detection(r1, 2 * r1), SO)
    with _ctx(P, 0) as g:
        np.testing.assert_array_equal(g.ridge_detection(r1, 2 * r1), SO)


@pytest.mark.parametrize("screen", SCREENS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_far_from_origin(oracle, dtype, screen):
    """A cloud offset by 10^4 in every coordinate: the centre keeps the screen's bound small."""
    P = (synth.helix3d(20000, seed=21, dtype=np.float64) + 1.0e4).astype(dtype)
    _say(f"case screen {screen} offset cloud {np.dtype(dtype).name}")
    S, _ = oracle.choose(P, 0.2)
    with _ctx(P, screen) as c:
        _check_iterations(oracle, c, P, S, 0.2, 3)


@pytest.mark.parametrize("screen", SCREENS)
def test_many_near_ties(oracle, screen):
    """A uniform cloud 200 radii wide with a small radius: the bound is wide against the gaps,
    so many samples take the exact window pass."""
    rng = np.random.default_rng(33)
    P = rng.uniform(-20.0, 20.0, (40000, 2)).astype(np.float32)
    _say(f"case screen {screen} many near ties")
    S, _ = oracle.choose(P, 0.2)
    with _ctx(P, screen) as c:
        _check_iterations(oracle, c, P, S, 0.2, 2)


@pytest.mark.parametrize("screen", SCREENS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_exact_ties(oracle, dtype, screen):
    """Chosen points repeated (each twice, and one three times): equal screened values, so the
    window pass must return the reference's first index."""
    P = synth.spiral2d(12000, seed=44).astype(dtype)
    _say(f"case screen {screen} exact ties {np.dtype(dtype).name}")
    S, _ = oracle.choose(P, 0.3)
    S = np.concatenate([S, S[::-1], S[:1]]).astype(dtype)
    with _ctx(P, screen) as c:
        labels = _check_iterations(oracle, c, P, S, 0.3, 1)[0]
    assert (labels[labels >= 0] < len(S) // 2).all()


@pytest.mark.parametrize("dim", [2, 3])
def test_batches(oracle, dim):
    """A batch of clouds of different extents and offsets, one centre per cloud."""
    from ridgedetection_amd._lib import Context, OPT_LOCAL, OPT_SCREEN
    g = np.random.default_rng(900 + dim)
    clouds = []
    for i in range(12):
        n = int(g.integers(2000, 30000))
        P = synth.helix3d(n, seed=300 + i)[:, :dim] * float(g.uniform(0.5, 20.0)) + float(g.uniform(-50, 50))
        ext = float(np.max(P.max(0) - P.min(0)))
        r1 = float(g.uniform(0.02, 0.06)) * ext
        clouds.append((P.astype(np.float32), r1, 2 * r1))
    _say(f"case screen batch d={dim}")
    res = {}
    for screen in (1, 3, 0):
        with Context(dim, np.float32) as c:
            c.set_option(OPT_SCREEN, screen)
            c.set_option(OPT_LOCAL, 0)
            res[screen] = c.batch([p for p, _, _ in clouds], [a for _, a, _ in clouds], [b for _, _, b in clouds])
            assert c.timers()["screen_form"] == _form(screen, np.float32)
    for (P, r1, r2), S, G, M in zip(clouds, res[1], res[0], res[3]):
        np.testing.assert_array_equal(S, G)
        np.testing.assert_array_equal(M, G)
        np.testing.assert_array_equal(S, oracle.ridge_detection(P, r1, r2)[0])


@pytest.mark.parametrize("alg", [6, 1])
def test_tiles(oracle, alg):
    """Tilings (one centre per leaf; neighbour samples read through the position index)."""
    from ridgedetection_amd._lib import Context, OPT_LOCAL, OPT_SCREEN
    P = synth.helix3d(30000, seed=35)
    _say(f"case screen tiled alg={alg}")
    res = {}
    for screen in (1, 3, 0):
        with Context(3, np.float32) as c:
            c.set_option(OPT_SCREEN, screen)
            c.set_option(OPT_LOCAL, 0)
            c.set_points(P)
            res[screen] = c.tiled(alg, 0.2, 0.4, 3000, [4, 4, 4])
    np.testing.assert_array_equal(res[1], res[0])
    np.testing.assert_array_equal(res[3], res[0])
    np.testing.assert_array_equal(res[1], oracle.tiled(alg, P, 0.2, 0.4, 3000, [4, 4, 4]))


def test_autotuned_form(oracle):
    """The default (RDHIP_OPT_SCREEN 2) times both forms and reports the one it keeps; results
    equal the oracle's either way."""
    from ridgedetection_amd._lib import Context
    P = synth.helix3d(60000, seed=8)
    _say("case screen autotuned")
    with Context(3, np.float32) as c:
        c.set_points(P)
        SR = c.ridge_detection(0.2, 0.4)
        assert c.timers()["screen_form"] in (0, 1, 2)
        _say(f"  picked form {c.timers()['screen_form']}")
    np.testing.assert_array_equal(SR, oracle.ridge_detection(P, 0.2, 0.4)[0])


def test_option_values():
    """RDHIP_OPT_SCREEN takes 0..3 and RDHIP_OPT_CHAIN 0..2; other values are refused with a
    message and leave the context usable."""
    from ridgedetection_amd._lib import Context, OPT_CHAIN, OPT_SCREEN, RDError
    P = synth.helix3d(5000, seed=3)
    with Context(3, np.float32) as c:
        for bad in (-1, 4):
            with pytest.raises(RDError, match="screen option"):
                c.set_option(OPT_SCREEN, bad)
        with pytest.raises(RDError, match="chain option"):
            c.set_option(OPT_CHAIN, 3)
        for ok in (0, 1, 2, 3):
            c.set_option(OPT_SCREEN, ok)
        c.set_points(P)
        assert len(c.ridge_detection(0.2, 0.4)) > 0
