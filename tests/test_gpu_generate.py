"""The device generator against the oracle's restatement: the uniform stream (Philox4x32-10,
no transcendentals) bit for bit; the shapes within the last bits the device and host sin /
cos / log may differ by; generated points, generated batches and their ridge detection
against the same clouds uploaded from the host and the oracle."""

import numpy as np
import pytest

from ridgedetection_amd import _lib as L

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dim", [1, 2, 5, 12])
def test_uniform_stream_bit_exact(oracle, dtype, dim):
    _say(f"case generator uniform d={dim} {np.dtype(dtype).name}")
    G = L.generate(L.GEN_UNIFORM, 70000, dim, seed=0x1234567890ABCDEF, stream_id=3, dtype=dtype)
    O = oracle.generate(L.GEN_UNIFORM, 70000, dim, seed=0x1234567890ABCDEF, stream=3, dtype=dtype)
    np.testing.assert_array_equal(G, O)


SHAPES = [
    (L.GEN_CIRCLE, 2, [1.0, -2.0, 3.0]),
    (L.GEN_SPHERE, 3, [0.5, 0.5, 0.5, 2.0]),
    (L.GEN_SPIRAL2D, 2, [6 * np.pi, 0.5]),
    (L.GEN_SPIRAL3D, 3, [6 * np.pi, 0.5]),
    (L.GEN_SEGMENT, 7, [2.5]),
    (L.GEN_HELIX, 3, [5.0, 0.15]),
    (L.GEN_POLYLINE, 4, [2, 0, 0, 0, 0, 1, 1, 0, 0, -1, 2, 3, 1]),
]


@pytest.mark.parametrize("dtype,rtol", [(np.float32, 2e-6), (np.float64, 1e-13)])
@pytest.mark.parametrize("kind,dim,params", SHAPES)
def test_shapes_match_oracle(oracle, kind, dim, params, dtype, rtol):
    _say(f"case generator kind={kind} d={dim} {np.dtype(dtype).name}")
    G = L.generate(kind, 50000, dim, params=params, sigma=0.05, seed=99, dtype=dtype)
    O = oracle.generate(kind, 50000, dim, params=params, sigma=0.05, seed=99, dtype=dtype)
    scale = float(np.max(np.abs(O)))
    np.testing.assert_allclose(G, O, rtol=rtol * 16, atol=rtol * 64 * scale)


def test_generate_points_ridge_detection(oracle):
    """Points generated into the context equal the generator's host output; ridge detection
    on them equals the oracle's."""
    params, n = [5.0, 0.15], 60000
    P = L.generate(L.GEN_HELIX, n, 3, params=params, sigma=0.05, seed=2024)
    _say("case generated points helix 60k")
    with L.Context(3, np.float32) as c:
        c.generate_points(L.GEN_HELIX, n, params=params, sigma=0.05, seed=2024)
        S = c.ridge_detection(0.2, 0.4)
    ref, _ = oracle.ridge_detection(P, 0.2, 0.4)
    np.testing.assert_array_equal(S, ref)


def test_batch_generate_matches_uploaded_batch(oracle):
    """A batch generated on the device (cloud c = stream c) gives the same results as the
    same clouds uploaded from the host, and as the oracle per cloud."""
    kinds = [L.GEN_SPIRAL2D, L.GEN_HELIX, L.GEN_POLYLINE, L.GEN_HELIX, L.GEN_SPIRAL2D]
    dims = [2, 3, 3, 2, 2]
    sizes = [12000, 30000, 8000, 15000, 4000]
    params = [[6 * np.pi, 0.5], [5.0, 0.15], [2, -1, -1, 0, 1, 0, 1, 0, 1, -1], [4.0, 0.2], [4 * np.pi, 0.3]]
    sigmas = [0.08, 0.05, 0.03, 0.04, 0.05]
    seeds = [7, 8, 9, 10, 11]
    r1 = [0.4, 0.2, 0.15, 0.18, 0.25]
    r2 = [2 * r for r in r1]
    clouds = [L.generate(k, n, d, params=p, sigma=s, seed=sd, stream_id=c)
              for c, (k, n, d, p, s, sd) in enumerate(zip(kinds, sizes, dims, params, sigmas, seeds))]
    _say("case batch generated on the device")
    with L.Context(3, np.float32) as c:
        c.batch_generate(kinds, dims, sizes, params, sigmas, seeds)
        dev = c.batch_run(r1, r2)
        up = c.batch(clouds, r1, r2)
    for i, (a, b) in enumerate(zip(dev, up)):
        np.testing.assert_array_equal(a, b)
        if i in (0, 2, 3):
            np.testing.assert_array_equal(a, oracle.ridge_detection(clouds[i], r1[i], r2[i])[0])
