"""Failure handling through the C ABI, forced from the host so that every kernel still ends
normally: the evolve iteration guard (RDHIP_OPT_EVOLVE_GUARD) and device allocations that fail
(RDHIP_OPT_INJECT_ALLOC_FAILURE).  A failed call returns its error code and message; a HIP
error leaves the context refusing further work until it is destroyed; other contexts, and new
ones, keep giving the oracle's results."""

import numpy as np
import pytest

from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("local", [0, 1])
def test_iteration_guard(oracle, local):
    """local = 1: the cloud (12k points) evolves inside one workgroup (device-side guard);
    local = 0: batch-wide iterations (host-side guard)."""
    from ridgedetection_amd._lib import OPT_EVOLVE_GUARD, OPT_LOCAL, Context, RDError
    P = synth.spiral2d(12000, seed=8)
    r = 0.45
    S0, _ = oracle.choose(P, r)
    SE, iters = oracle.evolve(P, S0, r)
    assert iters > 10
    _say(f"case iteration guard local={local} ({iters} iterations to converge)")
    with Context(2, np.float32) as c:
        c.set_option(OPT_LOCAL, local)
        c.set_points(P)
        c.set_chosen(S0)
        c.set_option(OPT_EVOLVE_GUARD, 3)
        with pytest.raises(RDError, match="iteration guard") as e:
            c.evolve(r)
        assert "-3" in str(e.value)   # RDHIP_ERR_RUNTIME
        # not a HIP error: the context goes on; the evolve map is deterministic, so resuming
        # from the interrupted set reaches the same fixed point
        c.set_option(OPT_EVOLVE_GUARD, 0)
        S, _ = c.evolve(r)
        np.testing.assert_array_equal(S, SE)
        # the guard also ends a whole ridge detection
        c.set_option(OPT_EVOLVE_GUARD, 2)
        with pytest.raises(RDError, match="iteration guard"):
            c.ridge_detection(r, 2 * r)
        c.set_option(OPT_EVOLVE_GUARD, 100000)
        np.testing.assert_array_equal(c.ridge_detection(r, 2 * r), oracle.ridge_detection(P, r, 2 * r)[0])


@pytest.mark.parametrize("case", ["brute", "tiled"])
def test_allocation_failures(oracle, case):
    """The k-th device allocation fails, for every k until a run needs fewer allocations."""
    from ridgedetection_amd._lib import OPT_INJECT_ALLOC_FAILURE, Context, RDError
    if case == "brute":
        P = synth.helix3d(20000, seed=13)
        run = lambda c: c.ridge_detection(0.2, 0.4)   # noqa: E731
        ref = oracle.ridge_detection(P, 0.2, 0.4)[0]
    else:
        P = synth.helix3d(30000, seed=14)
        run = lambda c: c.tiled(5, 0.2, 0.4, 2000, [3, 3, 3])   # noqa: E731
        ref = oracle.tiled(5, P, 0.2, 0.4, 2000, [3, 3, 3])
    bystander = Context(3, np.float32)   # a context that never sees a failure
    bystander.set_points(P)
    failures = 0
    try:
        for k in range(1, 400):
            c = Context(3, np.float32)
            try:
                c.set_points(P)
                c.set_option(OPT_INJECT_ALLOC_FAILURE, k)
                try:
                    S = run(c)
                except RDError as e:
                    assert "-2" in str(e) and "out of memory" in str(e), str(e)
                    failures += 1
                    with pytest.raises(RDError, match="unusable"):
                        c.count()
                    continue
                finally:
                    c.set_option(OPT_INJECT_ALLOC_FAILURE, 0)   # accepted by a broken context too
                np.testing.assert_array_equal(S, ref)   # fewer than k allocations: a normal run
                break
            finally:
                c.close()
        else:
            pytest.fail("every run hit an injected failure")
        _say(f"case allocation failures {case}: {failures} failing runs, then a clean run")
        assert failures >= 3
        fresh = Context(3, np.float32)
        fresh.set_points(P)
        np.testing.assert_array_equal(run(fresh), ref)
        fresh.close()
        np.testing.assert_array_equal(run(bystander), ref)
    finally:
        bystander.close()
