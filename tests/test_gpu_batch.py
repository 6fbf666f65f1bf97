"""Batched ridge detection (many clouds, one launch per phase) and the launch forms
(persistent work-queue scan vs one block per item, hipGraph replay vs eager launches):
every cloud's result must equal the CPU oracle's single-cloud result bit for bit."""

import numpy as np
import pytest

from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu


def _clouds(dim, count, seed):
    g = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(count):
        n = int(np.exp(g.uniform(np.log(300), np.log(20000))))
        kind = i % 3
        if kind == 0 and dim == 2:
            P = synth.spiral2d(n, seed=seed + i, sigma=0.08)
        elif (kind == 0 or kind == 1) and dim <= 3:
            P = synth.helix3d(n, seed=seed + i, sigma=0.05)[:, :dim]
        else:
            P = synth.segment_chain(n, dim, seed=seed + i, sigma=0.03)
        ext = float(np.max(P.max(0) - P.min(0)))
        r1 = float(g.uniform(0.02, 0.06)) * ext
        out.append((P, r1, 2 * r1))
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_batch_matches_single_cloud_oracle(oracle, dim):
    clouds = _clouds(dim, 24, 100 + dim)
    print(f"case batch d={dim} clouds={len(clouds)}", flush=True)
    from ridgedetection_amd._lib import Context
    with Context(dim, np.float32) as c:
        res = c.batch([p for p, _, _ in clouds], [a for _, a, _ in clouds], [b for _, _, b in clouds])
    for (P, r1, r2), S in zip(clouds, res):
        ref, _ = oracle.ridge_detection(P, r1, r2)
        np.testing.assert_array_equal(S, ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mixed_dimension_batch(oracle, dtype):
    """Clouds of dimensions 1..4 in one batch (zero-padded to 4 on upload): every cloud's
    result equals its own-dimension oracle run bit for bit."""
    mix = []
    for d in (2, 3, 4, 1):
        mix += _clouds(d, 5, 300 + d)
    g = np.random.Generator(np.random.PCG64(3))
    order = g.permutation(len(mix))
    mix = [mix[i] for i in order]
    print(f"case mixed-dimension batch {np.dtype(dtype).name} clouds={len(mix)}", flush=True)
    from ridgedetection_amd._lib import Context
    with Context(4, dtype) as c:
        res = c.batch([p.astype(dtype) for p, _, _ in mix], [a for _, a, _ in mix], [b for _, _, b in mix])
    for (P, r1, r2), S in zip(mix, res):
        assert S.shape[1] == P.shape[1]
        ref, _ = oracle.ridge_detection(P.astype(dtype), r1, r2)
        np.testing.assert_array_equal(S, ref)


@pytest.mark.parametrize("persistent,graph", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_launch_forms_agree(oracle, persistent, graph):
    from ridgedetection_amd._lib import Context, OPT_GRAPH, OPT_PERSISTENT, OPT_ITERS_PER_SYNC
    P = synth.helix3d(40000, seed=5)
    print(f"case launch form persistent={persistent} graph={graph}", flush=True)
    with Context(3, np.float32) as c:
        c.set_option(OPT_PERSISTENT, persistent)
        c.set_option(OPT_GRAPH, graph)
        c.set_option(OPT_ITERS_PER_SYNC, 4)
        c.set_points(P)
        S = c.ridge_detection(0.2, 0.4)
        # the same context again (graphs re-captured, buffers reused)
        S2 = c.ridge_detection(0.2, 0.4)
        clouds = _clouds(3, 8, 7)
        B = c.batch([p for p, _, _ in clouds], [a for _, a, _ in clouds], [b for _, _, b in clouds])
    ref, _ = oracle.ridge_detection(P, 0.2, 0.4)
    np.testing.assert_array_equal(S, ref)
    np.testing.assert_array_equal(S2, ref)
    for (Pc, r1, r2), Sc in zip(clouds, B):
        np.testing.assert_array_equal(Sc, oracle.ridge_detection(Pc, r1, r2)[0])


def test_fixed_iterations_graph_matches_eager():
    """evolve_fixed replayed from a graph (several replays, eager work and syncs in between)
    against the same iterations launched eagerly."""
    from ridgedetection_amd._lib import Context, OPT_GRAPH, OPT_ITERS_PER_SYNC
    P = synth.helix3d(60000, seed=9)
    out = []
    for graph in (0, 1):
        with Context(3, np.float32) as c:
            c.set_option(OPT_GRAPH, graph)
            c.set_option(OPT_ITERS_PER_SYNC, 3)
            c.set_points(P)
            c.choose(0.2)
            snaps = []
            for _ in range(4):
                c.evolve_fixed(0.2, 6)
                snaps.append(c.points())
            out.append(snaps)
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("local,segmented", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_segment_local_and_batchwide_iterations_agree(oracle, local, segmented):
    """Small segments evolve inside one workgroup each (every iteration in LDS), batches of
    small segments iterate as a batch-wide scan plus per-segment sort and sums, unless the
    options are off; every form must give the oracle's result, on a batch mixing clouds
    below and above the size limits and on tiled runs."""
    from ridgedetection_amd._lib import Context, OPT_LOCAL, OPT_SEGMENTED
    g = np.random.Generator(np.random.PCG64(77))
    clouds = []
    for i, n in enumerate([900, 5000, 16384, 16385, 30000, 2500, 70000]):
        P = synth.spiral2d(n, seed=300 + i, sigma=0.08) if i % 2 else synth.segment_chain(n, 2, seed=300 + i, sigma=0.03)
        ext = float(np.max(P.max(0) - P.min(0)))
        r1 = float(g.uniform(0.02, 0.05)) * ext
        clouds.append((P, r1, 2 * r1))
    print(f"case segment-local={local} segmented={segmented} batch", flush=True)
    with Context(2, np.float32) as c:
        c.set_option(OPT_LOCAL, local)
        c.set_option(OPT_SEGMENTED, segmented)
        res = c.batch([p for p, _, _ in clouds], [a for _, a, _ in clouds], [b for _, _, b in clouds])
    for (P, r1, r2), S in zip(clouds, res):
        np.testing.assert_array_equal(S, oracle.ridge_detection(P, r1, r2)[0])
    P = synth.helix3d(30000, seed=31)
    for alg in (6, 1, 0):
        print(f"case segment-local={local} segmented={segmented} tiled alg={alg}", flush=True)
        with Context(3, np.float32) as c:
            c.set_option(OPT_LOCAL, local)
            c.set_option(OPT_SEGMENTED, segmented)
            c.set_points(P)
            S = c.tiled(alg, 0.2, 0.4, 3000, [4, 4, 4])
        np.testing.assert_array_equal(S, oracle.tiled(alg, P, 0.2, 0.4, 3000, [4, 4, 4]))


def test_large_clouds_batch_rule(oracle):
    """Clouds of 100k-300k points drawn by the batched workload's own rule (spiral / helix /
    segment chains, noise sigma ~ U[0.01, 0.1] x extent, r1 = 4 sigma, r2 = 2 r1), two 2-D and
    two 3-D, in one mixed-dimension batch with the autotuned batch scan shape: each cloud's
    ridge equals the oracle's single-cloud result."""
    from ridgedetection_amd._lib import Context
    clouds = synth.batched_clouds(count=4, seed=11, min_pts=100000, max_pts=300000)
    print("case batch of 4 clouds, " + ", ".join(f"{len(p)} x {p.shape[1]}" for p, _, _ in clouds), flush=True)
    with Context(3, np.float32) as c:
        res = c.batch([p for p, _, _ in clouds], [a for _, a, _ in clouds], [b for _, _, b in clouds])
        t = c.timers()
    print(f"  scan shape {t['scan_ppt']} x {t['scan_threads']} persistent={t['scan_persistent']}", flush=True)
    for (P, r1, r2), S in zip(clouds, res):
        np.testing.assert_array_equal(S, oracle.ridge_detection(P, r1, r2)[0])
