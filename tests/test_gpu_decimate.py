"""Decimation of large chosen sets (a group of at least 2048 points decides in dependency
rounds on the device): kept points and their order must equal the CPU oracle's serial walk
(rd/cpu/brute_force/decimate.hpp:40-127), including the passes where the group-wide count
rule (stop below 3 points / keep at least 3) acts."""

import numpy as np
import pytest

from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _ctx(dim, dtype=np.float32):
    from ridgedetection_amd._lib import Context
    return Context(dim, dtype)


def _decimate(S, r2, rule):
    with _ctx(S.shape[1], S.dtype) as c:
        c.set_points(S)
        c.set_chosen(S)
        return c.decimate(r2, rule=rule)


def _noisy_curve(K, dim, seed, sigma):
    g = np.random.Generator(np.random.PCG64(seed))
    t = g.uniform(0.0, 6.0 * np.pi, K)
    base = np.stack([np.cos(t), np.sin(t), 0.1 * t, np.cos(2 * t)][:dim], axis=1)
    return np.ascontiguousarray((base + g.normal(0.0, sigma, base.shape)).astype(np.float32))


@pytest.mark.parametrize("K,dim,r2", [(2048, 2, 0.05), (6000, 3, 0.04), (9000, 3, 0.15), (4000, 4, 0.1)])
def test_large_group_matches_oracle(oracle, K, dim, r2):
    S = _noisy_curve(K, dim, seed=K + dim, sigma=0.03)
    _say(f"case decimate K={K} d={dim} r2={r2}")
    SD, keep = _decimate(S, r2, rule=0)
    ref, kept = oracle.decimate(S, r2)
    np.testing.assert_array_equal(SD, ref)
    np.testing.assert_array_equal(np.flatnonzero(keep), kept)


@pytest.mark.parametrize("K", [1, 2, 3, 5, 9, 17, 33, 65, 257])
def test_small_groups_match_oracle(oracle, K):
    """Groups that end inside a pruning chunk or a wave: the count pass reads candidates past
    the last element (never used) and lanes past the end stay out of every row."""
    g = np.random.Generator(np.random.PCG64(K))
    S = np.ascontiguousarray(g.normal(0.0, 0.05, (K, 3)).astype(np.float32))
    _say(f"case decimate small K={K}")
    SD, keep = _decimate(S, 0.04, rule=0)
    ref, kept = oracle.decimate(S, 0.04)
    np.testing.assert_array_equal(SD, ref)
    np.testing.assert_array_equal(np.flatnonzero(keep), kept)


@pytest.mark.parametrize("K,dim,r2", [(6000, 3, 0.04), (5000, 12, 0.2)])
def test_large_group_fp64_matches_oracle(oracle, K, dim, r2):
    """FP64 groups: the count pass, the provisional and walked rows and the rounds in double."""
    S = _noisy_curve(K, min(dim, 4), seed=3 * K + dim, sigma=0.03).astype(np.float64)
    if dim > S.shape[1]:   # the remaining coordinates: small noise around 0
        g = np.random.Generator(np.random.PCG64(dim))
        S = np.concatenate([S, g.normal(0.0, 0.01, (K, dim - S.shape[1]))], axis=1)
    S = np.ascontiguousarray(S)
    _say(f"case decimate fp64 K={K} d={dim} r2={r2}")
    with _ctx(dim, np.float64) as c:
        c.set_points(S)
        c.set_chosen(S)
        SD, keep = c.decimate(r2, rule=0)
    ref, kept = oracle.decimate(S, r2)
    np.testing.assert_array_equal(SD, ref)
    np.testing.assert_array_equal(np.flatnonzero(keep), kept)


@pytest.mark.parametrize("K,r2", [(3000, 0.05), (8000, 0.12)])
def test_large_group_nocopy_matches_oracle(oracle, K, r2):
    S = _noisy_curve(K, 3, seed=7 * K, sigma=0.03)
    _say(f"case decimate no-copy K={K} r2={r2}")
    SD, keep = _decimate(S, r2, rule=1)
    ref, kept = oracle.decimate(S, r2, nocopy=True)
    np.testing.assert_array_equal(SD, ref)
    np.testing.assert_array_equal(np.flatnonzero(keep), kept)


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("layout", ["crowded", "isolated"])
def test_count_rule_binds(oracle, rule, layout):
    """Every point passes the removal test: the count rule decides where the pass stops."""
    g = np.random.Generator(np.random.PCG64(5))
    K = 2500
    if layout == "crowded":    # all within r2 of each other: at least 4 neighbours
        S = g.uniform(0.0, 0.01, (K, 3)).astype(np.float32)
    else:                      # nothing within 2 r2: fewer than 3
        S = (np.arange(K, dtype=np.float32)[:, None] * np.float32(10.0)) * np.ones((1, 3), np.float32)
    r2 = 1.0
    _say(f"case decimate count rule {layout} rule={rule}")
    SD, keep = _decimate(np.ascontiguousarray(S), r2, rule=rule)
    if rule == 0:
        ref, kept = oracle.decimate(S, r2)
        np.testing.assert_array_equal(SD, ref)
        np.testing.assert_array_equal(np.flatnonzero(keep), kept)
    else:
        ref, kept = oracle.decimate(S, r2, nocopy=True)
        np.testing.assert_array_equal(SD, ref)
        np.testing.assert_array_equal(np.flatnonzero(keep), kept)
        assert len(kept) == 3


def test_batch_of_large_groups(oracle):
    """Several clouds whose chosen sets exceed 2048 points decide in the same rounds."""
    clouds = []
    for i, n in enumerate((14000, 9000, 16000)):
        P = synth.spiral2d(n, seed=40 + i, sigma=0.02)
        clouds.append((P, 0.045, 0.09))
    _say("case decimate batch of large groups")
    with _ctx(2) as c:
        res = c.batch([p for p, _, _ in clouds], [a for _, a, _ in clouds], [b for _, _, b in clouds])
        K0 = []
        for P, r1, _ in clouds:
            K0.append(len(oracle.choose(P, r1)[0]))
    assert max(K0) >= 2048, K0
    for (P, r1, r2), S in zip(clouds, res):
        ref, _ = oracle.ridge_detection(P, r1, r2)
        np.testing.assert_array_equal(S, ref)
