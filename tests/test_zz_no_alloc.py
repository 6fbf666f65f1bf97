"""rdhip_choose_noalloc (the choose on caller-owned device buffers, with the library's fixed
scratch array) against rdhip_choose (the context entry point) through the C ABI: chosen
indices, coordinates and counts bit for bit, on two calls in a row -- the second with a
smaller input -- for the single-block path (<= 32768 points) and the chunked path above it, in
FP32 and FP64; sizes beyond the limit are rejected before anything is launched."""

import ctypes as C

import numpy as np
import pytest

from ridgedetection_amd import synth

pytestmark = pytest.mark.gpu

_hip = None


def _hip_rt():
    """The HIP runtime librdhip.so is linked against (device buffers for the test)."""
    global _hip
    if _hip is None:
        from ridgedetection_amd import _lib
        _lib.load()
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert _hip_rt().hipMalloc(C.byref(self.p), max(int(nbytes), 1)) == 0
        self.nbytes = nbytes

    def up(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert _hip_rt().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def down(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert _hip_rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            _hip_rt().hipFree(self.p)
            self.p = C.c_void_p()


def _noalloc(P, r):
    """Greedy choose through rdhip_choose_noalloc: points uploaded SoA with a padded stride,
    chosen points written SoA with another; returns (S (K, dim), indices)."""
    from ridgedetection_amd import _lib
    lib = _lib.load()
    n, dim = P.shape
    dt = P.dtype
    stride, cstride = n + 37, n + 5
    soa = np.zeros((dim, stride), dtype=dt)
    soa[:, :n] = P.T
    dP, dS, dI, dK = DevBuf(soa.nbytes), DevBuf(dim * cstride * dt.itemsize), DevBuf(4 * n), DevBuf(4)
    try:
        dP.up(soa)
        rc = lib.rdhip_choose_noalloc(_lib.RDHIP_F32 if dt == np.float32 else _lib.RDHIP_F64, dim, dP.p, n, stride,
                                      float(r), dS.p, cstride, dI.p, dK.p, None)
        assert rc == 0, lib.rdhip_last_error().decode()
        assert _hip_rt().hipDeviceSynchronize() == 0
        K = int(dK.down(1, np.int32)[0])
        S = dS.down((dim, cstride), dt)[:, :K].T.copy()
        idx = dI.down(n, np.int32)[:K].copy()
        return S, idx
    finally:
        for b in (dP, dS, dI, dK):
            b.free()


def _context(P, r):
    from ridgedetection_amd._lib import Context
    with Context(P.shape[1], P.dtype) as c:
        c.set_points(P)
        return c.choose(r)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_noalloc_choose_matches_choose(dtype):
    # a chunked-path call (120k points, 3-D), then a smaller single-block call (9k, 2-D), then
    # a smaller chunked one (40k, 3-D)
    cases = [(synth.helix3d(120000, seed=61, dtype=dtype), 0.2), (synth.spiral2d(9000, seed=62, dtype=dtype), 0.45),
             (synth.helix3d(40000, seed=63, dtype=dtype), 0.2)]
    for P, r in cases:
        print(f"case choose_noalloc n={len(P)} d={P.shape[1]} {np.dtype(dtype).name}", flush=True)
        S, idx = _noalloc(P, r)
        S_ref, idx_ref = _context(P, r)
        np.testing.assert_array_equal(idx, idx_ref)
        np.testing.assert_array_equal(S, S_ref)
        assert len(idx) > 1


def test_noalloc_rejects_bad_sizes():
    from ridgedetection_amd import _lib
    lib = _lib.load()
    fake = C.c_void_p(256)   # never dereferenced: every call below fails validation first
    assert lib.rdhip_choose_noalloc(0, 3, fake, (1 << 24) + 1, 0, 0.1, fake, 0, fake, fake, None) == -1
    assert b"RDHIP_NOALLOC_MAX_POINTS" in lib.rdhip_last_error()
    assert lib.rdhip_choose_noalloc(0, 3, fake, 0, 0, 0.1, fake, 0, fake, fake, None) == -1
    assert lib.rdhip_choose_noalloc(0, 13, fake, 10, 0, 0.1, fake, 0, fake, fake, None) == -1
    assert lib.rdhip_choose_noalloc(0, 3, fake, 10, 5, 0.1, fake, 0, fake, fake, None) == -1   # stride < n
    assert lib.rdhip_choose_noalloc(0, 3, None, 10, 0, 0.1, fake, 0, fake, fake, None) == -1
    assert lib.rdhip_choose_noalloc(0, 3, fake, 10, 0, -1.0, fake, 0, fake, fake, None) == -1
