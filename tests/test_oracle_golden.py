"""The CPU oracle against golden fixtures written by the compiled reference CPU path
(tests/golden/make_golden.py).  Every comparison is bit-exact."""

import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN

BRUTE = sorted(glob.glob(os.path.join(GOLDEN, "brute_*.npz")))
TILED = sorted(glob.glob(os.path.join(GOLDEN, "tiled_*.npz")))
NOCOPY = sorted(glob.glob(os.path.join(GOLDEN, "nocopy_*.npz")))


def _name(p):
    return os.path.basename(p)[:-4]


@pytest.mark.parametrize("path", BRUTE, ids=_name)
def test_brute_phases(oracle, path):
    g = np.load(path)
    P, r1, r2 = g["P"], float(g["r1"]), float(g["r2"])
    S0, idx = oracle.choose(P, r1)
    np.testing.assert_array_equal(S0, g["choose_S"])
    np.testing.assert_array_equal(idx, g["choose_idx"])
    S1, flag, counts, sums, labels, _ = oracle.evolve_iter(P, S0, r1)
    np.testing.assert_array_equal(S1, g["iter_S"])
    assert flag == int(g["iter_flag"])
    np.testing.assert_array_equal(counts, g["iter_counts"])
    np.testing.assert_array_equal(sums, g["iter_sums"])
    SE, _ = oracle.evolve(P, S0, r1)
    np.testing.assert_array_equal(SE, g["evolve_S"])
    SD, kept = oracle.decimate(SE, r2)
    np.testing.assert_array_equal(SD, g["decimate_S"])
    np.testing.assert_array_equal(kept, g["decimate_kept"])
    SR, _ = oracle.ridge_detection(P, r1, r2)
    np.testing.assert_array_equal(SR, g["rd_S"])
    # quality against a pattern (rd/utils/assessment_quality.hpp)
    ideal = np.ascontiguousarray(P[::7])
    if len(SR):
        assert oracle.hausdorff(SR, ideal) == g["quality_hausdorff"]
        np.testing.assert_array_equal(oracle.distance_stats(SR, ideal), g["quality_stats"])
    # chain ordering (rd/cpu/brute_force/order.hpp) of the ridge and of the chosen set
    for S, r, key in ((SR, r2, "order"), (S0, r1, "order0")):
        idx, off = oracle.order(S, r)
        np.testing.assert_array_equal(idx, g[f"{key}_idx"])
        np.testing.assert_array_equal(off, g[f"{key}_off"])


@pytest.mark.parametrize("path", NOCOPY, ids=_name)
def test_decimate_nocopy(oracle, path):
    g = np.load(path)
    S = g["S"]

    import ctypes as C
    from oracle.oracle import _p
    S2 = np.array(S, copy=True)
    kept = np.empty(len(S), dtype=np.int64)
    fn = oracle._fn("decimate", S.dtype)
    k = fn(_p(S2), C.c_long(len(S)), C.c_int(S.shape[1]), oracle._scalar(S.dtype, float(g["r2"])),
           C.c_int(1), _p(kept))
    alive = np.zeros(len(S), dtype=bool)
    alive[kept[:k]] = True
    assert k == int(g["count"])
    np.testing.assert_array_equal(alive, g["alive"])


@pytest.mark.parametrize("path", TILED, ids=_name)
def test_trees(oracle, path):
    g = np.load(path)
    P, r1, tpd, cap = g["P"], float(g["r1"]), g["tpd"], int(g["cap"])
    for kind in (0, 1, 2):
        leaves = oracle.tree(kind, P, r1, cap, tpd)
        np.testing.assert_array_equal([len(l["own"]) for l in leaves], g[f"tree{kind}_own_counts"])
        np.testing.assert_array_equal(np.concatenate([l["own"] for l in leaves]), g[f"tree{kind}_own"])
        np.testing.assert_array_equal([len(l["nb"]) for l in leaves], g[f"tree{kind}_nb_counts"])
        nb = [l["nb"] for l in leaves]
        np.testing.assert_array_equal(np.concatenate(nb) if nb else [], g[f"tree{kind}_nb"])
        np.testing.assert_array_equal([l["cap"] for l in leaves], g[f"tree{kind}_cap"])
        np.testing.assert_array_equal(np.stack([l["bounds"] for l in leaves]), g[f"tree{kind}_bounds"])


@pytest.mark.parametrize("path", TILED, ids=_name)
@pytest.mark.parametrize("alg", [0, 1, 2, 3, 4, 5])
def test_tiled_drivers(oracle, path, alg):
    g = np.load(path)
    P, r1, r2, tpd, cap = g["P"], float(g["r1"]), float(g["r2"]), g["tpd"], int(g["cap"])
    for refine in (0, 1):
        S = oracle.tiled(alg, P, r1, r2, cap, tpd, refinement=bool(refine))
        np.testing.assert_array_equal(S, g[f"alg{alg}_ref{refine}_S"])
