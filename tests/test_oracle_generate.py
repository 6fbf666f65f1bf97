"""The oracle's restatement of the seeded generator: Philox4x32-10 against the published
known-answer vectors of its authors (Random123, Salmon et al. SC'11), the uniform and normal
streams by their statistics, every shape by its closed form (noise-free samples lie on
the curve / surface), and counter-based determinism (a sample depends only on its index,
stream and seed)."""

import numpy as np
import pytest

GEN_CIRCLE, GEN_SPHERE, GEN_SPIRAL2D, GEN_SPIRAL3D, GEN_SEGMENT, GEN_HELIX, GEN_POLYLINE, GEN_UNIFORM = range(8)

# (counter, key, output) -- Philox4x32-10 known-answer vectors published with Random123
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(oracle, ctr, key, out):
    np.testing.assert_array_equal(oracle.philox(ctr, key), np.array(out, dtype=np.uint32))


def test_philox_plants_fail(oracle):
    """A dropped round or a swapped multiplier changes the output (the vectors pin it)."""
    got = oracle.philox(KAT[2][0], KAT[2][1])
    assert not np.array_equal(got, oracle.philox(KAT[2][0], (KAT[2][1][1], KAT[2][1][0])))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_uniform_stream(oracle, dtype):
    U = oracle.generate(GEN_UNIFORM, 200000, 5, seed=11, dtype=dtype)
    assert U.min() > 0 and U.max() <= 1
    np.testing.assert_allclose(U.mean(axis=0), 0.5, atol=4e-3)
    np.testing.assert_allclose(U.var(axis=0), 1 / 12, rtol=2e-2)
    if dtype == np.float32:   # multiples of 2^-24
        np.testing.assert_array_equal(U * np.float32(2 ** 24), np.round(U * np.float32(2 ** 24)))
    # columns are independent streams
    c = np.corrcoef(U.T)
    assert np.max(np.abs(c - np.eye(5))) < 0.01


def test_normal_noise(oracle):
    """Segment of length 0: the samples are the noise alone, N(0, sigma^2) per coordinate."""
    X = oracle.generate(GEN_SEGMENT, 400000, 4, params=[0.0], sigma=0.3, seed=3, dtype=np.float64)
    np.testing.assert_allclose(X.mean(axis=0), 0.0, atol=3e-3)
    np.testing.assert_allclose(X.std(axis=0), 0.3, rtol=5e-3)
    kurt = ((X / X.std(axis=0)) ** 4).mean(axis=0)
    np.testing.assert_allclose(kurt, 3.0, atol=0.05)
    assert np.max(np.abs(np.corrcoef(X.T) - np.eye(4))) < 0.01


@pytest.mark.parametrize("dtype,tol", [(np.float32, 2e-5), (np.float64, 1e-12)])
def test_shapes_closed_form(oracle, dtype, tol):
    n = 5000
    C = oracle.generate(GEN_CIRCLE, n, 2, params=[1.0, -2.0, 3.0], seed=1, dtype=dtype).astype(np.float64)
    np.testing.assert_allclose(np.hypot(C[:, 0] - 1.0, C[:, 1] + 2.0), 3.0, rtol=tol * 4)
    S = oracle.generate(GEN_SPHERE, n, 3, params=[0.5, 0.5, 0.5, 2.0], seed=2, dtype=dtype).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(S - 0.5, axis=1), 2.0, rtol=tol * 4)
    a, b = 6 * np.pi, 0.5
    P = oracle.generate(GEN_SPIRAL2D, n, 2, params=[a, b], seed=3, dtype=dtype).astype(np.float64)
    ph = np.linalg.norm(P, axis=1) / b      # |(b p cos p, b p sin p)| = b p
    assert ph.max() <= a * (1 + tol * 8) + tol
    np.testing.assert_allclose(P[:, 0], b * ph * np.cos(ph), atol=tol * 200)
    Q = oracle.generate(GEN_SPIRAL3D, n, 3, params=[a, b], seed=4, dtype=dtype).astype(np.float64)
    np.testing.assert_allclose(np.hypot(Q[:, 0], Q[:, 1]), Q[:, 2], rtol=tol * 8, atol=tol * 8)
    H = oracle.generate(GEN_HELIX, n, 3, params=[5.0, 0.15], seed=5, dtype=dtype).astype(np.float64)
    np.testing.assert_allclose(np.hypot(H[:, 0], H[:, 1]), 1.0, rtol=tol * 4)
    t = H[:, 2] / 0.15
    np.testing.assert_allclose(H[:, 0], np.cos(t), atol=tol * 400)
    assert t.min() >= 0 and t.max() <= 2 * np.pi * 5 * (1 + tol)
    G = oracle.generate(GEN_SEGMENT, n, 4, params=[2.5], seed=6, dtype=dtype).astype(np.float64)
    assert np.all(G[:, 1:] == 0) and G[:, 0].min() > 0 and G[:, 0].max() <= 2.5


def test_polyline_points_on_segments(oracle):
    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 2.0, 0.0], [1.0, 2.0, -1.0]])
    X = oracle.generate(GEN_POLYLINE, 3000, 3, params=[3] + V.ravel().tolist(), seed=9, dtype=np.float64)
    d = []
    for s in range(3):
        a, b = V[s], V[s + 1]
        u = np.clip((X - a) @ (b - a) / ((b - a) @ (b - a)), 0, 1)
        d.append(np.linalg.norm(X - (a + u[:, None] * (b - a)), axis=1))
    assert np.min(d, axis=0).max() < 1e-12
    # segment s is picked with probability 1/3
    counts = [(np.min(d, axis=0) == d[s]).sum() for s in range(3)]
    assert min(counts) > 800


def test_counter_based_determinism(oracle):
    A = oracle.generate(GEN_HELIX, 1000, 3, params=[5.0, 0.15], sigma=0.05, seed=77, dtype=np.float32)
    B = oracle.generate(GEN_HELIX, 400, 3, params=[5.0, 0.15], sigma=0.05, seed=77, dtype=np.float32)
    np.testing.assert_array_equal(A[:400], B)
    Cs = oracle.generate(GEN_HELIX, 400, 3, params=[5.0, 0.15], sigma=0.05, seed=77, stream=1, dtype=np.float32)
    assert not np.array_equal(B, Cs)
    D = oracle.generate(GEN_HELIX, 400, 3, params=[5.0, 0.15], sigma=0.05, seed=78, dtype=np.float32)
    assert not np.array_equal(B, D)
