"""GPU parity at the benchmark workloads' own shapes, bit-exact against the CPU oracle.

* 3-D helix, 1,000,000 points (sigma 0.05), r1 = 0.2, r2 = 0.4, 8 x 8 x 8 initial tiles, tiles
  split at 10,000 points: the three trees (membership, neighbours, bounds, capacities) and the
  tiled drivers 0, 1, 5 and 6, in FP32 and FP64; brute-force choose and evolve iterations on
  the whole cloud.
* 6-D union of four cubic Bezier curves with 5% uniform outliers, r1 = 0.08, r2 = 0.16, 4 tiles
  per dimension: every tiled driver with and without end-phase refinement in FP32 and FP64 on
  60,000 points with tiles split at 1,000 points (trees up to 13 levels deep below the 4^6 grid),
  and the three trees of the full 2,000,000-point cloud at 10,000 points per tile.

The reference paths are rd/cpu/tiled/rd_tiled.hpp:756-1438 (drivers),
rd/cpu/tiled/local_tile_tree.inl:48-364 and grouped_tile_tree.inl:86-624 (trees)."""

import time

import numpy as np
import pytest

from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu

HELIX_N, HELIX_R1, HELIX_R2, HELIX_TPD, HELIX_CAP = 1_000_000, 0.2, 0.4, [8, 8, 8], 10000
BEZ_R1, BEZ_R2 = 0.08, 0.16


def _say(msg):
    print(msg, flush=True)


def _ctx(P):
    from ridgedetection_amd._lib import Context
    c = Context(P.shape[1], P.dtype)
    c.set_points(P)
    return c


def _same_tree(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x["own"], y["own"])
        np.testing.assert_array_equal(x["nb"], y["nb"])
        assert x["cap"] == y["cap"]
        np.testing.assert_array_equal(x["bounds"], y["bounds"])


_clouds = {}


def _helix(dtype):
    key = ("helix", np.dtype(dtype).name)
    if key not in _clouds:
        _clouds[key] = synth.helix3d(HELIX_N, seed=2, sigma=0.05, turns=5.0, dtype=dtype)
    return _clouds[key]


def _bezier(n, dtype):
    key = ("bezier", n, np.dtype(dtype).name)
    if key not in _clouds:
        _clouds[key] = synth.bezier_union(n, dim=6, seed=3, dtype=dtype)
    return _clouds[key]


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_helix1m_trees(oracle, kind):
    P = _helix(np.float32)
    _say(f"case helix1M tree kind {kind}")
    with _ctx(P) as c:
        got = c.tree(kind, HELIX_R1, HELIX_CAP, HELIX_TPD)
    _same_tree(got, oracle.tree(kind, P, HELIX_R1, HELIX_CAP, HELIX_TPD))


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("alg", [0, 1, 5, 6])
def test_helix1m_tiled(oracle, alg, dtype):
    P = _helix(dtype)
    _say(f"case helix1M tiled alg {alg} {np.dtype(dtype).name}")
    t0 = time.time()
    with _ctx(P) as c:
        S = c.tiled(alg, HELIX_R1, HELIX_R2, HELIX_CAP, HELIX_TPD)
    t1 = time.time()
    ref = oracle.tiled(alg, P, HELIX_R1, HELIX_R2, HELIX_CAP, HELIX_TPD)
    _say(f"  {len(S)} ridge points; gpu {t1 - t0:.2f} s, oracle {time.time() - t1:.2f} s")
    np.testing.assert_array_equal(S, ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_helix1m_brute_iterations(oracle, dtype):
    """Brute-force choose on the whole 1M-point cloud, then three evolve iterations: chosen
    indices, labels, counts, sums and moved coordinates bit for bit."""
    P = _helix(dtype)
    _say(f"case helix1M brute choose + 3 iterations {np.dtype(dtype).name}")
    with _ctx(P) as c:
        S, idx = c.choose(HELIX_R1)
        S_ref, idx_ref = oracle.choose(P, HELIX_R1)
        np.testing.assert_array_equal(idx, idx_ref)
        np.testing.assert_array_equal(S, S_ref)
        _say(f"  K = {len(S)}")
        for it in range(3):
            S, flag, counts, sums, labels, moved = c.evolve_iteration(HELIX_R1)
            S_ref, flag_ref, c_ref, s_ref, l_ref, m_ref = oracle.evolve_iter(P, S_ref, HELIX_R1)
            np.testing.assert_array_equal(labels, l_ref)
            np.testing.assert_array_equal(counts, c_ref)
            np.testing.assert_array_equal(sums, s_ref)
            mv = (moved[:, None] >> np.arange(P.shape[1])[None, :]) & 1
            np.testing.assert_array_equal(mv.astype(np.uint8), m_ref)
            np.testing.assert_array_equal(S, S_ref)
            assert flag == flag_ref


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("alg", [0, 1, 2, 4, 5, 6])
def test_bezier6d_tiled(oracle, alg, dtype):
    P = _bezier(60000, dtype)
    for refine in (False, True):
        _say(f"case bezier6d 60k tiled alg {alg} refine {int(refine)} {np.dtype(dtype).name}")
        with _ctx(P) as c:
            S = c.tiled(alg, BEZ_R1, BEZ_R2, 1000, [4] * 6, refinement=refine)
        ref = oracle.tiled(alg, P, BEZ_R1, BEZ_R2, 1000, [4] * 6, refinement=refine)
        _say(f"  {len(S)} ridge points")
        np.testing.assert_array_equal(S, ref)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_bezier6d_60k_trees(oracle, kind):
    P = _bezier(60000, np.float32)
    _say(f"case bezier6d 60k tree kind {kind}")
    with _ctx(P) as c:
        got = c.tree(kind, BEZ_R1, 1000, [4] * 6)
    ref = oracle.tree(kind, P, BEZ_R1, 1000, [4] * 6)
    _same_tree(got, ref)
    if kind == 1:   # cell bounds: depth below the 4^6 grid from the halved widths
        ext = P.max(axis=0) - P.min(axis=0)
        w = np.array([l["bounds"][:, 1] - l["bounds"][:, 0] for l in ref])
        depth = np.round(np.log2((ext / 4)[None, :] / w)).sum(axis=1)
        assert depth.max() >= 3


@pytest.mark.timeout(600)
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_bezier6d_2m_trees(oracle, kind):
    """The full 2M-point cloud at 10,000 points per tile (about 4,700 leaves)."""
    P = _bezier(2_000_000, np.float32)
    _say(f"case bezier6d 2M tree kind {kind}")
    t0 = time.time()
    with _ctx(P) as c:
        got = c.tree(kind, BEZ_R1, 10000, [4] * 6)
    t1 = time.time()
    ref = oracle.tree(kind, P, BEZ_R1, 10000, [4] * 6)
    _say(f"  {len(got)} leaves; gpu {t1 - t0:.2f} s, oracle {time.time() - t1:.2f} s")
    _same_tree(got, ref)
