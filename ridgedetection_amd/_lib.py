"""ctypes binding of librdhip.so (include/rdhip.h).

This is plumbing for the tests and the benchmark: the product is the C ABI.  Loading
fails loudly when the library has not been built -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librdhip.so")

RDHIP_F32, RDHIP_F64 = 0, 1
ROW_MAJOR, COL_MAJOR, SOA = 0, 1, 2
DECIMATE, DECIMATE_NO_COPY = 0, 1
TILED_LOCAL_TREE_MIXED_RD = 0
TILED_GROUPED_TREE_MIXED_RD = 1
TILED_LOCAL_TREE_LOCAL_RD = 2
TILED_LOCAL_LIST = 3
TILED_GROUPED_TREE_LOCAL_RD = 4
EXT_TILE_TREE_MIXED_RD = 5
EXT_TILE_TREE_LOCAL_RD = 6

EXPORTS = [
    "rdhip_version", "rdhip_last_error", "rdhip_create", "rdhip_destroy", "rdhip_set_points",
    "rdhip_ridge_detection", "rdhip_tiled_ridge_detection", "rdhip_get_count", "rdhip_get_points",
    "rdhip_get_timers", "rdhip_choose", "rdhip_set_chosen", "rdhip_evolve", "rdhip_evolve_iteration",
    "rdhip_decimate", "rdhip_set_option", "rdhip_evolve_fixed", "rdhip_build_tree", "rdhip_get_tree",
    "rdhip_batch_ridge_detection", "rdhip_evolve_fwd", "rdhip_evolve_bwd", "rdhip_order_chains",
    "rdhip_set_distances", "rdhip_hausdorff", "rdhip_distance_stats", "rdhip_batch_ridge_detection_dims",
    "rdhip_generate", "rdhip_generate_points", "rdhip_batch_generate", "rdhip_batch_run",
    "rdhip_set_chosen_ex", "rdhip_ridge_detection_own", "rdhip_choose_noalloc",
]
TREE_LOCAL, TREE_GROUPED, TREE_EXTENDED = 0, 1, 2
OPT_PPT, OPT_ITERS_PER_SYNC, OPT_PROFILE, OPT_PERSISTENT, OPT_GRAPH, OPT_LOCAL = 0, 1, 2, 3, 4, 5
OPT_LOCAL_MAX_SAMPLES, OPT_LOCAL_MAX_WORK, OPT_SCAN_THREADS, OPT_AUTOTUNE, OPT_SEGMENTED = 6, 7, 8, 9, 10
OPT_POSITION_COPY, OPT_EVOLVE_GUARD, OPT_INJECT_ALLOC_FAILURE, OPT_VJP_MEMORY_MB, OPT_SUM_MODE = 11, 12, 13, 14, 15
OPT_CHAIN = 16
OPT_SCREEN = 17
GEN_CIRCLE, GEN_SPHERE, GEN_SPIRAL2D, GEN_SPIRAL3D, GEN_SEGMENT, GEN_HELIX, GEN_POLYLINE, GEN_UNIFORM = range(8)


class Timers(C.Structure):
    _fields_ = [("whole_ms", C.c_double), ("build_ms", C.c_double), ("choose_ms", C.c_double),
                ("evolve_ms", C.c_double), ("decimate_ms", C.c_double), ("refine_ms", C.c_double),
                ("evolve_iterations", C.c_int64), ("rounds", C.c_int64),
                ("scan_kernel_ms", C.c_double), ("scan_launches", C.c_int64), ("iteration_ms", C.c_double),
                ("forward_ms", C.c_double), ("backward_ms", C.c_double), ("evolve_bytes", C.c_double),
                ("scan_ppt", C.c_int64), ("scan_threads", C.c_int64), ("tune_ms", C.c_double),
                ("evolve_pairs", C.c_double), ("scan_persistent", C.c_int64),
                ("vjp_checkpoint_every", C.c_int64), ("chain_path", C.c_int64),
                ("screen_form", C.c_int64)]


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m ridgedetection_amd.build`")
    lib = C.CDLL(path)
    vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int
    sig = {
        "rdhip_version": (i32, []),
        "rdhip_last_error": (C.c_char_p, []),
        "rdhip_create": (i32, [C.POINTER(vp), i32, i32, vp]),
        "rdhip_destroy": (i32, [vp]),
        "rdhip_set_points": (i32, [vp, vp, i64, i32, i64, i32]),
        "rdhip_ridge_detection": (i32, [vp, C.c_double, C.c_double, i64, C.POINTER(i64)]),
        "rdhip_tiled_ridge_detection": (i32, [vp, i32, C.c_double, C.c_double, i64, vp, i32, i64,
                                              C.POINTER(i64)]),
        "rdhip_get_count": (i32, [vp, C.POINTER(i64)]),
        "rdhip_get_points": (i32, [vp, vp, i32, i64, i32]),
        "rdhip_get_timers": (i32, [vp, C.POINTER(Timers)]),
        "rdhip_choose": (i32, [vp, C.c_double, C.POINTER(i64), vp]),
        "rdhip_set_chosen": (i32, [vp, vp, i64, i32]),
        "rdhip_set_chosen_ex": (i32, [vp, vp, i64, i32, i64, i32]),
        "rdhip_ridge_detection_own": (i32, [vp, i64, C.c_double, C.c_double, i64, C.POINTER(i64)]),
        "rdhip_choose_noalloc": (i32, [i32, i32, vp, i64, i64, C.c_double, vp, i64, vp, vp, vp]),
        "rdhip_evolve": (i32, [vp, C.c_double, i64, C.POINTER(i64)]),
        "rdhip_evolve_iteration": (i32, [vp, C.c_double, vp, vp, vp, vp, C.POINTER(i32)]),
        "rdhip_decimate": (i32, [vp, C.c_double, i32, vp, C.POINTER(i64)]),
        "rdhip_set_option": (i32, [vp, i32, i64]),
        "rdhip_evolve_fixed": (i32, [vp, C.c_double, i64]),
        "rdhip_build_tree": (i32, [vp, i32, C.c_double, i64, vp, C.POINTER(i64)]),
        "rdhip_get_tree": (i32, [vp, vp, vp, vp, vp]),
        "rdhip_batch_ridge_detection": (i32, [vp, i64, vp, vp, vp, vp, i64, vp]),
        "rdhip_batch_ridge_detection_dims": (i32, [vp, i64, vp, vp, vp, vp, vp, i64, vp]),
        "rdhip_generate": (i32, [C.c_int, C.c_int, i64, C.c_int, vp, i64, C.c_double, C.c_uint64, C.c_uint32, vp,
                                 C.c_int, i64, C.c_int]),
        "rdhip_generate_points": (i32, [vp, C.c_int, i64, vp, i64, C.c_double, C.c_uint64]),
        "rdhip_batch_generate": (i32, [vp, i64, vp, vp, vp, vp, vp, vp, vp]),
        "rdhip_batch_run": (i32, [vp, vp, vp, i64, vp]),
        "rdhip_evolve_fwd": (i32, [vp, C.c_double, i64, i64]),
        "rdhip_order_chains": (i32, [vp, C.c_double, vp, vp, C.POINTER(i64)]),
        "rdhip_set_distances": (i32, [vp, vp, i64, vp, i64, vp]),
        "rdhip_hausdorff": (i32, [vp, vp, i64, vp, i64, C.POINTER(C.c_double)]),
        "rdhip_distance_stats": (i32, [vp, vp, i64, vp, i64, vp]),
        "rdhip_evolve_bwd": (i32, [vp, vp, vp, vp, i32, i32, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class RDError(RuntimeError):
    pass


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def generate(kind, n, dim, params=(), sigma=0.0, seed=0, stream_id=0, dtype=np.float32):
    """n seeded samples of a generator kind (GEN_*), computed on the GPU, as an (n, dim)
    host array."""
    lib = load()
    dt = np.dtype(dtype)
    out = np.empty((max(int(n), 1), dim), dtype=dt)
    prm = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
    rc = lib.rdhip_generate(RDHIP_F32 if dt == np.float32 else RDHIP_F64, int(kind), int(n), int(dim),
                            _ptr(prm) if prm.size else None, int(prm.size), float(sigma), int(seed), int(stream_id),
                            _ptr(out), ROW_MAJOR, 0, 0)
    if rc != 0:
        raise RDError(f"rdhip error {rc}: {lib.rdhip_last_error().decode()}")
    return out[:int(n)]


class Context:
    """One rdhip context (a cloud on the GPU).  dtype: np.float32 or np.float64."""

    def __init__(self, dim: int, dtype=np.float32, stream: int | None = None):
        self.lib = load()
        self.dim = dim
        self.dtype = np.dtype(dtype)
        h = C.c_void_p()
        code = RDHIP_F32 if self.dtype == np.float32 else RDHIP_F64
        self._check(self.lib.rdhip_create(C.byref(h), code, dim, C.c_void_p(stream) if stream else None))
        self.h = h
        self.n = 0

    def _check(self, rc):
        if rc != 0:
            raise RDError(f"rdhip error {rc}: {self.lib.rdhip_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.rdhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- input / output ------------------------------------------------------------
    def set_points(self, P, layout=ROW_MAJOR, stride=0):
        P = np.ascontiguousarray(P, dtype=self.dtype)
        self.n = P.shape[0] if layout == ROW_MAJOR else P.shape[-1]
        self._keep = P
        self._check(self.lib.rdhip_set_points(self.h, _ptr(P), self.n, layout, stride, 0))

    def set_points_device(self, ptr: int, n: int, layout=ROW_MAJOR, stride=0):
        self.n = n
        self._check(self.lib.rdhip_set_points(self.h, C.c_void_p(ptr), n, layout, stride, 1))

    def count(self) -> int:
        k = C.c_int64()
        self._check(self.lib.rdhip_get_count(self.h, C.byref(k)))
        return k.value

    def points(self, layout=ROW_MAJOR) -> np.ndarray:
        k = self.count()
        out = np.empty((k, self.dim) if layout == ROW_MAJOR else (self.dim, k), dtype=self.dtype)
        if k:
            self._check(self.lib.rdhip_get_points(self.h, _ptr(out), layout, 0, 0))
        return out

    def timers(self) -> dict:
        t = Timers()
        self._check(self.lib.rdhip_get_timers(self.h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in Timers._fields_}

    # -- drivers -------------------------------------------------------------------
    def ridge_detection(self, r1, r2, max_evolve_iters=0):
        k = C.c_int64()
        self._check(self.lib.rdhip_ridge_detection(self.h, float(r1), float(r2), int(max_evolve_iters),
                                                   C.byref(k)))
        return self.points()

    def ridge_detection_own(self, n_own, r1, r2, max_evolve_iters=0):
        """Ridge detection choosing from the first n_own points, evolving over all of them."""
        k = C.c_int64()
        self._check(self.lib.rdhip_ridge_detection_own(self.h, int(n_own), float(r1), float(r2),
                                                       int(max_evolve_iters), C.byref(k)))
        return self.points()

    def batch(self, clouds, r1, r2, max_evolve_iters=0):
        """Batched ridge detection over clouds (list of (n_c, dim) arrays) with per-cloud radii.
        Returns the list of per-cloud ridge points."""
        arrs = [np.ascontiguousarray(c, dtype=self.dtype) for c in clouds]
        m = len(arrs)
        ptrs = (C.c_void_p * m)(*[a.ctypes.data for a in arrs])
        sizes = np.array([len(a) for a in arrs], dtype=np.int64)
        dims = np.array([a.shape[1] for a in arrs], dtype=np.int64)
        vr1 = np.ascontiguousarray(r1, dtype=np.float64).reshape(m)
        vr2 = np.ascontiguousarray(r2, dtype=np.float64).reshape(m)
        counts = np.zeros(m, dtype=np.int64)
        if np.all(dims == self.dim):
            self._check(self.lib.rdhip_batch_ridge_detection(self.h, m, ptrs, _ptr(sizes), _ptr(vr1), _ptr(vr2),
                                                             int(max_evolve_iters), _ptr(counts)))
        else:   # clouds of lower dimensions run zero-padded to the context's
            self._check(self.lib.rdhip_batch_ridge_detection_dims(self.h, m, ptrs, _ptr(dims), _ptr(sizes),
                                                                  _ptr(vr1), _ptr(vr2), int(max_evolve_iters),
                                                                  _ptr(counts)))
        self.n = int(sizes.sum())
        allp = self.points()
        out, o = [], 0
        for k, dc in zip(counts, dims):
            out.append(np.ascontiguousarray(allp[o:o + k, :dc]))
            o += k
        return out

    def generate_points(self, kind, n, params=(), sigma=0.0, seed=0):
        """n generated samples of the context's dimension as its points (stream 0)."""
        prm = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        self._check(self.lib.rdhip_generate_points(self.h, int(kind), int(n), _ptr(prm) if prm.size else None,
                                                   int(prm.size), float(sigma), int(seed)))
        self.n = int(n)

    def batch_generate(self, kinds, dims, sizes, params, sigmas, seeds):
        """Clouds generated on the device as the resident batch: params is a list of
        per-cloud parameter sequences."""
        m = len(kinds)
        off = np.zeros(m, dtype=np.int64)
        flat = []
        for c, p in enumerate(params):
            off[c] = len(flat)
            flat.extend(float(x) for x in p)
        prm = np.ascontiguousarray(flat if flat else [0.0], dtype=np.float64)
        self._batch_dims = np.ascontiguousarray(dims, dtype=np.int64).reshape(m)
        self._batch_sizes = np.ascontiguousarray(sizes, dtype=np.int64).reshape(m)
        self._check(self.lib.rdhip_batch_generate(
            self.h, m, _ptr(np.ascontiguousarray(kinds, dtype=np.int32)), _ptr(self._batch_dims),
            _ptr(self._batch_sizes), _ptr(prm), _ptr(off), _ptr(np.ascontiguousarray(sigmas, dtype=np.float64)),
            _ptr(np.ascontiguousarray(seeds, dtype=np.uint64))))
        self.n = int(self._batch_sizes.sum())

    def batch_run(self, r1, r2, max_evolve_iters=0):
        """Ridge detection of the resident batch; per-cloud results (first dims[c] coordinates)."""
        m = len(self._batch_sizes)
        vr1 = np.ascontiguousarray(r1, dtype=np.float64).reshape(m)
        vr2 = np.ascontiguousarray(r2, dtype=np.float64).reshape(m)
        counts = np.zeros(m, dtype=np.int64)
        self._check(self.lib.rdhip_batch_run(self.h, _ptr(vr1), _ptr(vr2), int(max_evolve_iters), _ptr(counts)))
        allp = self.points()
        out, o = [], 0
        for k, dc in zip(counts, self._batch_dims):
            out.append(np.ascontiguousarray(allp[o:o + k, :dc]))
            o += k
        return out

    def tiled(self, algorithm, r1, r2, max_tile_capacity, tiles_per_dim, refinement=False, max_evolve_iters=0):
        tpd = np.ascontiguousarray(tiles_per_dim, dtype=np.int64)
        k = C.c_int64()
        self._check(self.lib.rdhip_tiled_ridge_detection(self.h, int(algorithm), float(r1), float(r2),
                                                         int(max_tile_capacity), _ptr(tpd), int(bool(refinement)),
                                                         int(max_evolve_iters), C.byref(k)))
        return self.points()

    def tree(self, kind, r1, max_tile_capacity, tiles_per_dim):
        """Builds a tile tree on the device; returns its leaves in preorder as dicts
        (own, nb, cap, bounds) like the oracle's tree dump."""
        tpd = np.ascontiguousarray(tiles_per_dim, dtype=np.int64)
        L = C.c_int64()
        self._check(self.lib.rdhip_build_tree(self.h, int(kind), float(r1), int(max_tile_capacity), _ptr(tpd),
                                              C.byref(L)))
        L = L.value
        info = np.zeros((max(L, 1), 4), dtype=np.int64)
        self._check(self.lib.rdhip_get_tree(self.h, _ptr(info), None, None, None))
        own = np.zeros(max(int(info[:L, 0].sum()), 1), dtype=np.int32)
        nb = np.zeros(max(int(info[:L, 1].sum()), 1), dtype=np.int32)
        bounds = np.zeros((max(L, 1), 2 * self.dim), dtype=self.dtype)
        self._check(self.lib.rdhip_get_tree(self.h, _ptr(info), _ptr(own), _ptr(nb), _ptr(bounds)))
        leaves, o, b = [], 0, 0
        for l in range(L):
            no, nn, cap, _ = info[l]
            leaves.append(dict(own=own[o:o + no].astype(np.int64), nb=nb[b:b + nn].astype(np.int64), cap=int(cap),
                               bounds=bounds[l].reshape(self.dim, 2).copy()))
            o += no
            b += nn
        return leaves

    # -- phases --------------------------------------------------------------------
    def choose(self, r):
        k = C.c_int64()
        idx = np.empty(max(self.n, 1), dtype=np.int32)
        self._check(self.lib.rdhip_choose(self.h, float(r), C.byref(k), _ptr(idx)))
        return self.points(), idx[:k.value].copy()

    def set_chosen(self, S, layout=ROW_MAJOR, stride=0):
        """Chosen points from host memory: (K, dim) ROW_MAJOR, or (dim, K) COL_MAJOR / SOA."""
        S = np.ascontiguousarray(S, dtype=self.dtype)
        K = S.shape[0] if layout == ROW_MAJOR else S.shape[-1]
        if layout == ROW_MAJOR and stride == 0:
            self._check(self.lib.rdhip_set_chosen(self.h, _ptr(S), K, ROW_MAJOR))
        else:
            self._check(self.lib.rdhip_set_chosen_ex(self.h, _ptr(S), K, layout, stride, 0))

    def evolve(self, r, max_iters=0):
        it = C.c_int64()
        self._check(self.lib.rdhip_evolve(self.h, float(r), int(max_iters), C.byref(it)))
        return self.points(), it.value

    def set_option(self, option, value):
        self._check(self.lib.rdhip_set_option(self.h, int(option), int(value)))

    def evolve_fixed(self, r, iters):
        self._check(self.lib.rdhip_evolve_fixed(self.h, float(r), int(iters)))

    def set_distances(self, A, B):
        A = np.ascontiguousarray(A, dtype=self.dtype)
        B = np.ascontiguousarray(B, dtype=self.dtype)
        out = np.empty(len(A), dtype=self.dtype)
        self._check(self.lib.rdhip_set_distances(self.h, _ptr(A), len(A), _ptr(B), len(B), _ptr(out)))
        return out

    def hausdorff(self, A, B):
        A = np.ascontiguousarray(A, dtype=self.dtype)
        B = np.ascontiguousarray(B, dtype=self.dtype)
        out = C.c_double()
        self._check(self.lib.rdhip_hausdorff(self.h, _ptr(A), len(A), _ptr(B), len(B), C.byref(out)))
        return out.value

    def distance_stats(self, A, B):
        A = np.ascontiguousarray(A, dtype=self.dtype)
        B = np.ascontiguousarray(B, dtype=self.dtype)
        out = np.zeros(4, dtype=np.float64)
        self._check(self.lib.rdhip_distance_stats(self.h, _ptr(A), len(A), _ptr(B), len(B), _ptr(out)))
        return out

    def order_chains(self, r2):
        """Chain ordering of the current chosen points: (order, offsets) as the oracle's."""
        K = self.count()
        order = np.empty(max(K, 1), dtype=np.int32)
        off = np.empty(K + 1, dtype=np.int64)
        nc = C.c_int64()
        self._check(self.lib.rdhip_order_chains(self.h, float(r2), _ptr(order), _ptr(off), C.byref(nc)))
        return order[:K].astype(np.int64), off[:nc.value + 1].copy()

    def evolve_fwd(self, r, iters, checkpoint_every=0):
        """Exactly `iters` evolve iterations, keeping what evolve_bwd needs."""
        self._check(self.lib.rdhip_evolve_fwd(self.h, float(r), int(iters), int(checkpoint_every)))

    def evolve_bwd(self, grad_chosen, seed_to_points=True):
        """VJP of the last evolve_fwd with memberships frozen: grad_chosen = dL/dS_T (K x dim,
        host).  Returns (dL/dP [n x dim], dL/dS_0 [K x dim])."""
        g = np.ascontiguousarray(grad_chosen, dtype=self.dtype)
        gp = np.empty((self.n, self.dim), dtype=self.dtype)
        gs = np.empty_like(g)
        self._check(self.lib.rdhip_evolve_bwd(self.h, _ptr(g), _ptr(gp), _ptr(gs), ROW_MAJOR, 0,
                                              1 if seed_to_points else 0))
        return gp, gs

    def evolve_bwd_device(self, grad_chosen_ptr: int, grad_points_ptr: int, grad_seeds_ptr: int = 0,
                          seed_to_points=True):
        """Device form (dense SoA buffers: dim x K in, dim x n out)."""
        self._check(self.lib.rdhip_evolve_bwd(self.h, C.c_void_p(grad_chosen_ptr), C.c_void_p(grad_points_ptr),
                                              C.c_void_p(grad_seeds_ptr) if grad_seeds_ptr else None, SOA, 1,
                                              1 if seed_to_points else 0))

    def evolve_iteration(self, r):
        K = self.count()
        labels = np.empty(max(self.n, 1), dtype=np.int32)
        counts = np.empty(max(K, 1), dtype=np.int32)
        sums = np.empty((max(K, 1), self.dim), dtype=self.dtype)
        moved = np.empty(max(K, 1), dtype=np.uint16)
        flag = C.c_int()
        self._check(self.lib.rdhip_evolve_iteration(self.h, float(r), _ptr(labels), _ptr(counts), _ptr(sums),
                                                    _ptr(moved), C.byref(flag)))
        return (self.points(), flag.value, counts[:K].copy(), sums[:K].copy(), labels[:self.n].copy(),
                moved[:K].copy())

    def decimate(self, r2, rule=DECIMATE):
        K = self.count()
        keep = np.empty(max(K, 1), dtype=np.uint8)
        k = C.c_int64()
        self._check(self.lib.rdhip_decimate(self.h, float(r2), int(rule), _ptr(keep), C.byref(k)))
        return self.points(), keep[:K].astype(bool)
