"""GPU parity of the single-cloud iteration path of rd_chain.hip (label histogram in the scan,
one member scatter, one lane per chain sum) against the oracle and against the general sort +
chain-sum path (RDHIP_OPT_CHAIN = 0), bit for bit: every scan shape (so every scatter block
size), FP32 and FP64, dimensions 1 to 4, chains longer than many unrolled chunks, chains
shorter than one, empty chains and samples outside every ball, and chosen counts at the path's
2048-key limit and past it."""

import numpy as np
import pytest

from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu

SHAPES = [(1, 128), (2, 128), (4, 128), (1, 256), (2, 256), (4, 256), (1, 512), (2, 512), (4, 512)]


def _say(msg):
    print(msg, flush=True)


def _ctx(P, chain=True, shape=None):
    from ridgedetection_amd._lib import Context, OPT_CHAIN, OPT_PPT, OPT_SCAN_THREADS
    c = Context(P.shape[1], P.dtype)
    c.set_option(OPT_CHAIN, 2 if chain else 0)   # 2: the path whenever the key limit allows
    if shape:
        c.set_option(OPT_PPT, shape[0])
        c.set_option(OPT_SCAN_THREADS, shape[1])
    c.set_points(P)
    return c


def _cloud(dim, n, seed, dtype):
    if dim == 1:
        rng = np.random.default_rng(seed)
        return np.sort(rng.uniform(0, 10, (n, 1))).astype(dtype)
    if dim == 2:
        return synth.spiral2d(n, seed=seed).astype(dtype)
    if dim == 3:
        return synth.helix3d(n, seed=seed, dtype=dtype)
    return synth.bezier_union(n, dim=4, seed=seed).astype(dtype)


def _iterations(oracle, c, P, r, steps):
    """steps iterations through evolve_iteration, each checked against the oracle from the same
    chosen set; returns the final chosen set."""
    S, _ = c.choose(r)
    for _ in range(steps):
        oS, oflag, ocounts, osums, olabels, omoved = oracle.evolve_iter(P, S, r)
        S1, flag, counts, sums, labels, moved = c.evolve_iteration(r)
        np.testing.assert_array_equal(labels, olabels)
        np.testing.assert_array_equal(counts, ocounts)
        np.testing.assert_array_equal(sums, osums)
        np.testing.assert_array_equal(S1, oS)
        mv = ((moved[:, None] >> np.arange(P.shape[1])[None, :]) & 1).astype(np.uint8)
        np.testing.assert_array_equal(mv, omoved)
        assert flag == oflag
        S = S1
    return S


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_every_scan_shape(oracle, shape, dtype):
    P = synth.helix3d(30011, seed=5, dtype=dtype)
    _say(f"case chain shape {shape} {np.dtype(dtype).name}")
    with _ctx(P, shape=shape) as c:
        _iterations(oracle, c, P, 0.2, 3)
        assert c.timers()["chain_path"] == 1


@pytest.mark.parametrize("dim", [1, 2, 3, 4])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dims_and_evolve_loop(oracle, dim, dtype):
    """Five checked iterations, then the graph-replayed evolve loop and a whole ridge detection
    against the oracle and the general path."""
    P = _cloud(dim, 24000, 7 + dim, dtype)
    r1 = {1: 0.05, 2: 0.45, 3: 0.2, 4: 0.25}[dim]
    _say(f"case chain dim={dim} {np.dtype(dtype).name}")
    with _ctx(P) as c:
        _iterations(oracle, c, P, r1, 5)
        S0, _ = c.choose(r1)
        SE, it = c.evolve(r1, 60)
        oSE, oit = oracle.evolve(P, S0, r1, 60)
        np.testing.assert_array_equal(SE, oSE)
        assert it == oit
        SR = c.ridge_detection(r1, 2 * r1)
    np.testing.assert_array_equal(SR, oracle.ridge_detection(P, r1, 2 * r1)[0])
    with _ctx(P, chain=False) as g:
        np.testing.assert_array_equal(g.ridge_detection(r1, 2 * r1), SR)
        assert g.timers()["chain_path"] == 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_one_chosen_point(oracle, dtype):
    """A radius larger than the cloud: one chosen point, one chain of every sample (hundreds of
    unrolled chunks in one lane)."""
    P = synth.helix3d(50000, seed=9, dtype=dtype)
    _say(f"case chain K=1 {np.dtype(dtype).name}")
    with _ctx(P) as c:
        S, _ = c.choose(100.0)
        assert len(S) == 1
        _iterations(oracle, c, P, 100.0, 2)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_short_chains_and_uncovered_samples(oracle, dtype):
    """A small radius: chains of a few members (no whole chunk), and after the first shifts
    some samples lie outside every ball and some chosen points have no members."""
    P = synth.spiral2d(12000, seed=13).astype(dtype)
    _say(f"case chain short chains {np.dtype(dtype).name}")
    with _ctx(P) as c:
        S, _ = c.choose(0.15)
        assert 300 < len(S) <= 2048
        _iterations(oracle, c, P, 0.15, 6)


@pytest.mark.parametrize("target", ["at_limit", "past_limit"])
def test_key_limit(oracle, target):
    """Chosen counts just under the path's 2048-key limit (taken) and above it (general path):
    the same results either way."""
    P = synth.spiral2d(60000, seed=17)
    r = 0.124 if target == "at_limit" else 0.122    # 2032 and 2098 chosen points
    _say(f"case chain key limit {target}")
    with _ctx(P) as c:
        S, _ = c.choose(r)
        K = len(S)
        _say(f"  K = {K}")
        assert (K <= 2048) == (target == "at_limit")
        _iterations(oracle, c, P, r, 2)
        assert c.timers()["chain_path"] == (1 if K <= 2048 else 0)


def _batch_clouds(dim, count, seed):
    g = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(count):
        n = int(np.exp(g.uniform(np.log(2000), np.log(40000))))
        P = synth.helix3d(n, seed=seed + i, sigma=0.05)[:, :dim] if i % 2 else synth.segment_chain(n, dim, seed=seed + i, sigma=0.03)
        ext = float(np.max(P.max(0) - P.min(0)))
        r1 = float(g.uniform(0.02, 0.06)) * ext
        out.append((P.astype(np.float32), r1, 2 * r1))
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_batches_take_the_path(oracle, dim):
    """Batch-wide iterations of many clouds (the segment-local path off, so every cloud iterates
    batch-wide) order members per cloud without the sort: equal to the general path and to the
    oracle's single-cloud runs."""
    from ridgedetection_amd._lib import Context, OPT_CHAIN, OPT_LOCAL
    clouds = _batch_clouds(dim, 16, 500 + dim)
    _say(f"case chain batch d={dim} clouds={len(clouds)}")
    res = {}
    for chain in (2, 0):
        with Context(dim, np.float32) as c:
            c.set_option(OPT_CHAIN, chain)
            c.set_option(OPT_LOCAL, 0)
            res[chain] = c.batch([p for p, _, _ in clouds], [a for _, a, _ in clouds], [b for _, _, b in clouds])
            assert c.timers()["chain_path"] == (1 if chain else 0)
    for (P, r1, r2), S, G in zip(clouds, res[2], res[0]):
        np.testing.assert_array_equal(S, G)
        np.testing.assert_array_equal(S, oracle.ridge_detection(P, r1, r2)[0])


@pytest.mark.parametrize("alg", [6, 1, 5])
def test_tiles_take_the_path(oracle, alg):
    """Tiled runs (leaves with neighbour samples, indexed positions) through the path: equal to
    the general path and to the oracle."""
    from ridgedetection_amd._lib import Context, OPT_CHAIN, OPT_LOCAL
    P = synth.helix3d(30000, seed=33)
    _say(f"case chain tiled alg={alg}")
    res = {}
    for chain in (2, 0):
        with Context(3, np.float32) as c:
            c.set_option(OPT_CHAIN, chain)
            c.set_option(OPT_LOCAL, 0)
            c.set_points(P)
            res[chain] = c.tiled(alg, 0.2, 0.4, 3000, [4, 4, 4])
            assert c.timers()["chain_path"] == (1 if chain else 0)
    np.testing.assert_array_equal(res[2], res[0])
    np.testing.assert_array_equal(res[2], oracle.tiled(alg, P, 0.2, 0.4, 3000, [4, 4, 4]))


def test_six_dimensional_tiles(oracle):
    """A 6-D Bezier union tiled with the grouped tree (mixed RD) through the path, against the
    general path and the oracle."""
    from ridgedetection_amd._lib import Context, OPT_CHAIN, OPT_LOCAL
    P = synth.bezier_union(20000, dim=6, seed=41)
    _say("case chain tiled 6-D grouped")
    res = {}
    for chain in (2, 0):
        with Context(6, np.float32) as c:
            c.set_option(OPT_CHAIN, chain)
            c.set_option(OPT_LOCAL, 0)
            c.set_points(P)
            res[chain] = c.tiled(1, 0.08, 0.16, 2000, [3] * 6)
            assert c.timers()["chain_path"] == (1 if chain else 0)
    np.testing.assert_array_equal(res[2], res[0])
    np.testing.assert_array_equal(res[2], oracle.tiled(1, P, 0.08, 0.16, 2000, [3] * 6))
