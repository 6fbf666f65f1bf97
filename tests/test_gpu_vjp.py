"""GPU parity of the evolve forward/backward pair (memberships frozen) through the C ABI.

The forward pass is the ordinary evolve, so S_T is bit-exact against the oracle.  The
backward pass adds, per sample, the step coefficients G/c of its chosen point newest step
first, the oracle's order, so in the same precision the gradient is bit-exact as well;
the FP32 gradient is also compared with the FP64 oracle gradient (on the same frozen
memberships) and the FP64 gradient with central finite differences of the FP64 oracle
replay (h = 1e-4), which the CPU tests pin to the reference evolve.
"""

import numpy as np
import pytest

from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _ctx(P):
    from ridgedetection_amd._lib import Context
    c = Context(P.shape[1], P.dtype)
    c.set_points(P)
    return c


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("every", [0, 3])
def test_fwd_bwd_bit_exact(oracle, dtype, every):
    n, T, r = 20000, 8, 0.2
    _say(f"case fwd/bwd n={n} T={T} every={every} {np.dtype(dtype).name}")
    P = synth.helix3d(n, seed=3, dtype=dtype)
    S0o, idx = oracle.choose(P, r)
    ST_o, lab, cnt, mv = oracle.evolve_tape(P, S0o, r, T)
    with _ctx(P) as c:
        S0, gidx = c.choose(r)
        np.testing.assert_array_equal(gidx, idx)
        c.evolve_fwd(r, T, every)
        ST = c.points()
        np.testing.assert_array_equal(ST, ST_o)
        _, G = synth.helix_loss(ST)
        G = G.astype(dtype)
        gp, gs0 = c.evolve_bwd(G, seed_to_points=True)
        ogp, ogs0 = oracle.evolve_vjp(lab, cnt, mv, G, dtype=dtype)
        np.testing.assert_array_equal(gs0, ogs0)
        ogp[idx] += ogs0
        np.testing.assert_array_equal(gp, ogp)
        # the chosen set is left as the forward produced it; a second backward is identical
        np.testing.assert_array_equal(c.points(), ST)
        gp2, gs02 = c.evolve_bwd(G, seed_to_points=True)
        np.testing.assert_array_equal(gp2, gp)
        np.testing.assert_array_equal(gs02, gs0)
        # FP32 error against the FP64 gradient on the same memberships
        if dtype == np.float32:
            fgp, fgs0 = oracle.evolve_vjp(lab, cnt, mv, G.astype(np.float64), dtype=np.float64)
            fgp[idx] += fgs0
            err = np.abs(gp - fgp).max() / np.abs(fgp).max()
            _say(f"  fp32 gradient rel. error vs fp64: {err:.2e}")
            assert err <= 1e-5


def test_checkpointing_is_exact():
    n, T, r = 30000, 10, 0.2
    P = synth.helix3d(n, seed=8, dtype=np.float32)
    outs = []
    for every in (0, 1, 4, 10):
        _say(f"case checkpoint every={every}")
        with _ctx(P) as c:
            c.choose(r)
            c.evolve_fwd(r, T, every)
            _, G = synth.helix_loss(c.points())
            outs.append(c.evolve_bwd(G.astype(np.float32)))
    for gp, gs0 in outs[1:]:
        np.testing.assert_array_equal(gp, outs[0][0])
        np.testing.assert_array_equal(gs0, outs[0][1])


def test_automatic_checkpoint_interval():
    """checkpoint_every < 0: every iteration's memberships when they fit the memory budget,
    else the longest run that fits; the gradient does not change."""
    from ridgedetection_amd._lib import OPT_VJP_MEMORY_MB
    n, T, r = 200000, 20, 0.2
    P = synth.helix3d(n, seed=8, dtype=np.float32)
    outs, intervals = [], []
    # 20 x 200000 x 4 bytes of labels fit 1 GiB; with 16 MiB a run of 14 iterations and two
    # 2.4 MB checkpoints do, 15 do not
    for budget_mb in (1024, 16):
        _say(f"case automatic checkpoints, budget {budget_mb} MiB")
        with _ctx(P) as c:
            c.set_option(OPT_VJP_MEMORY_MB, budget_mb)
            c.choose(r)
            c.evolve_fwd(r, T, -1)
            intervals.append(c.timers()["vjp_checkpoint_every"])
            _, G = synth.helix_loss(c.points())
            outs.append(c.evolve_bwd(G.astype(np.float32)))
    assert intervals[0] == T
    assert 1 < intervals[1] < T and intervals[1] * n * 4 <= 16 << 20
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_helix_c5_fd_and_fp64(oracle):
    """200 000-point helix, T = 20, checkpoints every 4 iterations: FP64 gradient against
    central differences on 1000 random coordinates; FP32 gradient against the FP64 one."""
    n, T, r, every = 200000, 20, 0.2, 4
    _say(f"case C5 n={n} T={T} every={every}")
    P64 = synth.helix3d(n, seed=2024, dtype=np.float64)
    with _ctx(P64) as c:
        _, idx = c.choose(r)
        c.evolve_fwd(r, T, every)
        ST = c.points()
        L, G = synth.helix_loss(ST)
        gp, _ = c.evolve_bwd(G)
    S0 = P64[idx]
    ST_o, lab, cnt, mv = oracle.evolve_tape(P64, S0, r, T)
    np.testing.assert_array_equal(ST, ST_o)
    _say(f"  K={len(idx)} loss={L:.6f}")

    def partial_loss(S, rows):
        return synth.helix_loss(S[rows])[0]

    rng = np.random.default_rng(7)
    rows = rng.integers(0, n, 1000)
    dims = rng.integers(0, 3, 1000)
    h = 1e-4
    scale = np.abs(gp).max()
    worst = 0.0
    for i, d in zip(rows.tolist(), dims.tolist()):
        Pp = P64.copy()
        Pp[i, d] += h
        Pm = P64.copy()
        Pm[i, d] -= h
        Sp = oracle.evolve_replay(Pp, Pp[idx], lab, mv)
        Sm = oracle.evolve_replay(Pm, Pm[idx], lab, mv)
        changed = np.flatnonzero(np.any(Sp != ST, axis=1) | np.any(Sm != ST, axis=1))
        fd = 0.0 if len(changed) == 0 else (partial_loss(Sp, changed) - partial_loss(Sm, changed)) / (2 * h)
        worst = max(worst, abs(fd - gp[i, d]))
    _say(f"  max |fd - grad| = {worst:.2e} (max |grad| {scale:.2e})")
    assert worst <= 1e-7 * scale

    P32 = P64.astype(np.float32)
    with _ctx(P32) as c:
        _, idx32 = c.choose(r)
        c.evolve_fwd(r, T, every)
        ST32 = c.points()
        _, G32 = synth.helix_loss(ST32)
        gp32, _ = c.evolve_bwd(G32.astype(np.float32))
    ST_f, lab32, cnt32, mv32 = oracle.evolve_tape(P32, P32[idx32], r, T)
    np.testing.assert_array_equal(ST32, ST_f)
    fgp, fgs0 = oracle.evolve_vjp(lab32, cnt32, mv32, G32, dtype=np.float64)
    fgp[idx32] += fgs0
    err = np.abs(gp32 - fgp).max() / np.abs(fgp).max()
    _say(f"  fp32 gradient rel. error vs fp64 oracle: {err:.2e}")
    assert err <= 1e-5


def test_autograd_function_wrapper():
    """Optional torch.autograd.Function over the pair (tests only).  The tensors stay on the
    host: torch ships its own HIP runtime, and a second runtime in the process cannot open
    the GPU once the library's runtime holds it, so only the library touches the device."""
    torch = pytest.importorskip("torch")
    from ridgedetection_amd._lib import Context

    class Evolve(torch.autograd.Function):
        @staticmethod
        def forward(fctx, P, r, iters, every):
            c = Context(P.shape[1], np.float32)
            c.set_points(P.detach().numpy())
            c.choose(r)
            c.evolve_fwd(r, iters, every)
            fctx.c = c
            return torch.from_numpy(c.points())

        @staticmethod
        def backward(fctx, gS):
            upstream.append(gS.detach().contiguous().numpy().copy())
            gp, _ = fctx.c.evolve_bwd(upstream[-1])
            fctx.c.close()
            return torch.from_numpy(gp), None, None, None

    upstream = []
    n, T, r = 20000, 6, 0.2
    _say(f"case autograd n={n} T={T}")
    P = torch.from_numpy(synth.helix3d(n, seed=4)).requires_grad_(True)
    S = Evolve.apply(P, r, T, 2)
    _, G = synth.helix_loss(S.detach().numpy())
    H = torch.from_numpy((S.detach().numpy() - G / 2).astype(np.float32))   # nearest helix points
    loss = ((S - H) ** 2).sum()
    loss.backward()
    Pn = P.detach().numpy()
    with _ctx(Pn) as c:
        c.choose(r)
        c.evolve_fwd(r, T, 2)
        np.testing.assert_array_equal(c.points(), S.detach().numpy())
        gp, _ = c.evolve_bwd(upstream[0])
    # torch delivered 2 (S - H), the helix loss gradient up to the FP32 rounding of H
    ulp = float(np.spacing(np.float32(np.abs(S.detach().numpy()).max())))
    np.testing.assert_allclose(upstream[0], G, rtol=0, atol=4 * ulp)
    np.testing.assert_array_equal(P.grad.numpy(), gp)
