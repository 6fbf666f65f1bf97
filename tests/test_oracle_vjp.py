"""Pins the oracle's differentiated evolve (frozen memberships) without the GPU.

The reference has no derivative, so the oracle defines the map: with the labels, counts
and moved flags of a recorded run held fixed, each iteration is the affine shift of
rd/cpu/brute_force/evolve.hpp:129-179 applied to member sums (evolve.hpp:32-77).  Checks:
  * the recording forward equals the oracle's evolve (itself pinned to the reference by
    the golden fixtures) bit for bit;
  * replaying the recorded memberships from S_0 reproduces S_T bit for bit, so the replay
    is the same function as the real forward pass at these inputs;
  * the reverse-mode gradient equals central finite differences (h = 1e-4, FP64) of the
    replay composed with the helix loss, on random coordinates and on the samples the
    chosen set was drawn from (the seed path);
  * closed forms: no moved coordinate -> zero point gradient and G passed through; one
    iteration -> dL/dP(n) = m G(l(n)) / c(l(n)).
"""

import numpy as np
import pytest

from ridgedetection_amd import synth


def _run(oracle, n, T, dtype, seed=3, r=0.2):
    P = synth.helix3d(n, seed=seed, dtype=dtype)
    S0, idx = oracle.choose(P, r)
    ST, lab, cnt, mv = oracle.evolve_tape(P, S0, r, T)
    return P, S0, idx, ST, lab, cnt, mv


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tape_and_replay_match_evolve(oracle, dtype):
    P, S0, idx, ST, lab, cnt, mv = _run(oracle, 6000, 7, dtype)
    ref, _ = oracle.evolve(P, S0, 0.2, max_iter=7)
    np.testing.assert_array_equal(ST, ref)
    np.testing.assert_array_equal(oracle.evolve_replay(P, S0, lab, mv), ST)
    # counts are the label histogram of every iteration
    for t in range(lab.shape[0]):
        h = np.bincount(lab[t][lab[t] >= 0], minlength=len(S0))
        np.testing.assert_array_equal(h, cnt[t])
    assert 0 < mv.mean() < 1


def test_vjp_matches_central_differences(oracle):
    P, S0, idx, ST, lab, cnt, mv = _run(oracle, 8000, 6, np.float64)
    _, G = synth.helix_loss(ST)
    gp, gs0 = oracle.evolve_vjp(lab, cnt, mv, G)
    gp[idx] += gs0

    def loss(Pp):
        return synth.helix_loss(oracle.evolve_replay(Pp, Pp[idx], lab, mv))[0]

    rng = np.random.default_rng(11)
    coords = list(zip(rng.integers(0, len(P), 40).tolist(), rng.integers(0, 3, 40).tolist()))
    coords += [(int(idx[k]), k % 3) for k in rng.integers(0, len(idx), 10)]
    h = 1e-4
    scale = np.abs(gp).max()
    for i, d in coords:
        Pp = P.copy()
        Pp[i, d] += h
        Pm = P.copy()
        Pm[i, d] -= h
        fd = (loss(Pp) - loss(Pm)) / (2 * h)
        assert abs(fd - gp[i, d]) <= 1e-7 * scale, (i, d, fd, gp[i, d])


def test_vjp_closed_forms(oracle):
    P, S0, idx, ST, lab, cnt, mv = _run(oracle, 3000, 1, np.float64)
    rng = np.random.default_rng(5)
    G = rng.normal(size=S0.shape)
    gp, gs0 = oracle.evolve_vjp(lab, cnt, mv, G)
    expect = np.zeros_like(P)
    for n in range(len(P)):
        k = lab[0, n]
        if k >= 0:
            expect[n] = mv[0, k] * G[k] / cnt[0, k]
    np.testing.assert_allclose(gp, expect, rtol=1e-15, atol=0)
    np.testing.assert_array_equal(gs0, np.where(mv[0] == 1, 0.0, G))
    # nothing moves: S_T = S_0, no gradient reaches the points
    gp, gs0 = oracle.evolve_vjp(lab, cnt, np.zeros_like(mv), G)
    assert not gp.any()
    np.testing.assert_array_equal(gs0, G)
