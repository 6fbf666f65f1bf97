"""GPU parity of the tiled path through the C ABI: tile trees (membership, neighbours,
bounds, capacities) and every tiled driver, bit-exact against the golden fixtures of the
reference's CPU implementation and against the CPU oracle on further clouds."""

import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN
from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu

TILED = sorted(glob.glob(os.path.join(GOLDEN, "tiled_*.npz")))


def _name(p):
    return os.path.basename(p)[:-4]


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


@pytest.mark.parametrize("path", TILED, ids=_name)
def test_trees_match_golden(path):
    g = np.load(path)
    P, r1, tpd, cap = g["P"], float(g["r1"]), g["tpd"], int(g["cap"])
    print(f"case {_name(path)} trees", flush=True)
    with _ctx(P) as c:
        for kind in (0, 1, 2):
            leaves = c.tree(kind, r1, cap, tpd)
            np.testing.assert_array_equal([len(l["own"]) for l in leaves], g[f"tree{kind}_own_counts"])
            np.testing.assert_array_equal(np.concatenate([l["own"] for l in leaves]), g[f"tree{kind}_own"])
            np.testing.assert_array_equal([len(l["nb"]) for l in leaves], g[f"tree{kind}_nb_counts"])
            np.testing.assert_array_equal(np.concatenate([l["nb"] for l in leaves]), g[f"tree{kind}_nb"])
            np.testing.assert_array_equal([l["cap"] for l in leaves], g[f"tree{kind}_cap"])
            np.testing.assert_array_equal(np.stack([l["bounds"] for l in leaves]), g[f"tree{kind}_bounds"])


@pytest.mark.parametrize("path", TILED, ids=_name)
@pytest.mark.parametrize("alg", [0, 1, 2, 3, 4, 5])
def test_tiled_drivers_match_golden(path, alg):
    g = np.load(path)
    P, r1, r2, tpd, cap = g["P"], float(g["r1"]), float(g["r2"]), g["tpd"], int(g["cap"])
    print(f"case {_name(path)} alg {alg}", flush=True)
    with _ctx(P) as c:
        for refine in (0, 1):
            S = c.tiled(alg, r1, r2, cap, tpd, refinement=bool(refine))
            np.testing.assert_array_equal(S, g[f"alg{alg}_ref{refine}_S"])


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("case", ["helix3d", "bezier4d"])
def test_trees_match_oracle(oracle, kind, case):
    if case == "helix3d":
        P, r1, tpd, cap = synth.helix3d(60000, seed=31), 0.2, [4, 4, 4], 700
    else:
        P, r1, tpd, cap = synth.bezier_union(30000, dim=4, seed=32), 0.15, [3, 3, 3, 3], 900
    print(f"case tree {case} kind {kind}", flush=True)
    with _ctx(P) as c:
        _same_tree(c.tree(kind, r1, cap, tpd), oracle.tree(kind, P, r1, cap, tpd))


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_trees_fp64_match_oracle(oracle, kind):
    """FP64 trees: split bounds, cell tables and the grouped neighbour matrix in double."""
    P = synth.bezier_union(30000, dim=4, seed=33, dtype=np.float64)
    r1, tpd, cap = 0.15, [3, 3, 3, 3], 900
    print(f"case tree fp64 kind {kind}", flush=True)
    with _ctx(P) as c:
        _same_tree(c.tree(kind, r1, cap, tpd), oracle.tree(kind, P, r1, cap, tpd))


def test_grouped_tree_many_leaves(oracle):
    """10,000 grouped leaves: the device's all-pairs neighbour bit matrix is filled and read
    back in more than one row block."""
    rng = np.random.default_rng(77)
    P = rng.random((200000, 2)).astype(np.float32)
    r1, cap, tpd = 0.01, 100000, [100, 100]
    print("case grouped tree 10k leaves", flush=True)
    with _ctx(P) as c:
        got = c.tree(1, r1, cap, tpd)
    assert len(got) == 10000
    _same_tree(got, oracle.tree(1, P, r1, cap, tpd))


@pytest.mark.parametrize("alg", [0, 2, 5, 6])
def test_tiled_drivers_match_oracle(oracle, alg):
    P = synth.helix3d(50000, seed=41, dtype=np.float64)
    print(f"case tiled helix50k alg {alg}", flush=True)
    with _ctx(P) as c:
        S = c.tiled(alg, 0.2, 0.4, 2500, [4, 4, 4])
    np.testing.assert_array_equal(S, oracle.tiled(alg, P, 0.2, 0.4, 2500, [4, 4, 4]))


@pytest.mark.parametrize("alg", [1, 5, 6])
def test_position_copy_off_matches(oracle, alg):
    """Tiled runs read their samples from a position-ordered copy by default; reading them
    through the position index instead (RDHIP_OPT_POSITION_COPY = 0) gives the same result."""
    from ridgedetection_amd._lib import OPT_POSITION_COPY
    P = synth.bezier_union(40000, dim=4, seed=12, dtype=np.float32)
    print(f"case tiled position copy bezier4d alg {alg}", flush=True)
    out = []
    for copy in (1, 0):
        with _ctx(P) as c:
            c.set_option(OPT_POSITION_COPY, copy)
            out.append(c.tiled(alg, 0.1, 0.2, 3000, [3, 3, 3, 3]))
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], oracle.tiled(alg, P, 0.1, 0.2, 3000, [3, 3, 3, 3]))
