"""GPU parity of the split ordered sums of long chains: a chain of at least 1024 members is summed
in four runs at once, each run from 64 speculative starts (consecutive values around the
correctly rounded sum before it), and a run whose exact start is not among them is summed again
from it. These cases force that second pass: one chain's members are a, -a, b, -b (four runs of
equal length), so the exact sum before the third run is 0 while the serial sum holds only its
rounding residue, far outside any window around 0. Its other coordinate and a second chain of
positive members take the first pass. A chain of 40 000 members covers runs that start after
long prefixes (10 000 to 30 000 members), where the serial sum strays furthest from the
correctly rounded one. Bit-exact against the oracle for every member ordering
(RDHIP_OPT_CHAIN 0, 1, 2), FP32 and FP64."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _cloud(dtype, q=3000, seed=61):
    rng = np.random.default_rng(seed)
    a = rng.uniform(50.0, 100.0, q)
    b = rng.uniform(50.0, 100.0, q)
    x = np.concatenate([a, -a, b, -b])
    y = rng.uniform(0.0, 1.0, 4 * q)
    far = np.stack([rng.uniform(1000.0, 1100.0, 5000), rng.uniform(0.0, 1.0, 5000)], 1)
    P = np.concatenate([np.stack([x, y], 1), far]).astype(dtype)
    S = np.array([[0.0, 0.5], [1050.0, 0.5]], dtype=dtype)
    return P, S, 4 * q


def _residue(P, m):
    """Serial sum (in the cloud's type) of the first chain's first coordinate after two runs."""
    return np.cumsum(P[: m // 2, 0], dtype=P.dtype)[-1]


@pytest.mark.parametrize("chain", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_second_pass(oracle, dtype, chain):
    from ridgedetection_amd._lib import Context, OPT_CHAIN
    P, S, m = _cloud(dtype)
    # the exact sum before the third run is 0; the serial one is not, so no start matches it
    assert _residue(P, m) != 0
    _say(f"case chain parts {np.dtype(dtype).name} chain option {chain}")
    with Context(2, dtype) as c:
        c.set_option(OPT_CHAIN, chain)
        c.set_points(P)
        c.set_chosen(S)
        for _ in range(2):
            oS, oflag, ocounts, osums, olabels, omoved = oracle.evolve_iter(P, S, 200.0)
            S1, flag, counts, sums, labels, moved = c.evolve_iteration(200.0)
            np.testing.assert_array_equal(labels, olabels)
            np.testing.assert_array_equal(counts, ocounts)
            np.testing.assert_array_equal(sums, osums)
            np.testing.assert_array_equal(S1, oS)
            assert flag == oflag
            S = S1
    assert ocounts.min() >= 1024


@pytest.mark.parametrize("chain", [0, 1])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_long_prefix_windows(oracle, dtype, chain):
    """One chain of 40 000 members: runs of 10 000 members after prefixes of up to 30 000."""
    from ridgedetection_amd._lib import Context, OPT_CHAIN
    rng = np.random.default_rng(62)
    P = rng.uniform(4.0, 6.0, (40000, 2)).astype(dtype)
    S = np.array([[5.0, 5.0]], dtype=dtype)
    _say(f"case long prefix {np.dtype(dtype).name} chain option {chain}")
    with Context(2, dtype) as c:
        c.set_option(OPT_CHAIN, chain)
        c.set_points(P)
        c.set_chosen(S)
        for _ in range(2):
            oS, oflag, ocounts, osums, olabels, omoved = oracle.evolve_iter(P, S, 2.0)
            S1, flag, counts, sums, labels, moved = c.evolve_iteration(2.0)
            np.testing.assert_array_equal(counts, ocounts)
            np.testing.assert_array_equal(sums, osums)
            np.testing.assert_array_equal(S1, oS)
            assert flag == oflag
            S = S1
    assert ocounts[0] == 40000
