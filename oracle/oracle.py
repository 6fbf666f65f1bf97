"""ctypes bindings for the CPU oracle (test infrastructure only).

`Oracle()` loads oracle/liboracle.so -- the C++ restatement of the reference's serial CPU
path (rd_oracle.cpp).  `Oracle(kind="ref")` loads oracle/_ref/libref.so, the reference's
own headers behind ref_driver.cpp (only where the reference sources were present at build
time).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

LIBS = {
    "oracle": os.path.join(HERE, "liboracle.so"),
    "oracle_omp": os.path.join(HERE, "liboracle_omp.so"),
    "ref": os.path.join(HERE, "_ref", "libref.so"),
    "ref_omp": os.path.join(HERE, "_ref", "libref_omp.so"),
}

# Tiled algorithm ids (reference rd/cpu/tiled/rd_tiled.hpp:102-110, + local RD on extended tiles)
TILED_LOCAL_TREE_MIXED_RD = 0
TILED_GROUPED_TREE_MIXED_RD = 1
TILED_LOCAL_TREE_LOCAL_RD = 2
TILED_LOCAL_LIST = 3
TILED_GROUPED_TREE_LOCAL_RD = 4
EXT_TILE_TREE_MIXED_RD = 5
EXT_TILE_TREE_LOCAL_RD = 6

TREE_LOCAL, TREE_GROUPED, TREE_EXTENDED = 0, 1, 2


def available(kind: str) -> bool:
    return os.path.exists(LIBS[kind])


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Oracle:
    def __init__(self, kind: str = "oracle"):
        path = LIBS[kind]
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
        self.kind = kind
        self.lib = C.CDLL(path)
        self.prefix = "ref_" if kind.startswith("ref") else "oracle_"

    def _fn(self, name, dt):
        suf = "f32" if np.dtype(dt) == np.float32 else "f64"
        f = getattr(self.lib, f"{self.prefix}{name}_{suf}")
        f.restype = C.c_long
        return f

    @staticmethod
    def _scalar(dt, v):
        return C.c_float(v) if np.dtype(dt) == np.float32 else C.c_double(v)

    def _check(self, rc, what):
        if rc < 0:
            msg = ""
            if self.prefix == "oracle_":
                self.lib.oracle_last_error.restype = C.c_char_p
                msg = self.lib.oracle_last_error().decode()
            raise RuntimeError(f"{what} failed: {msg}")
        return rc

    # -- phases ------------------------------------------------------------------
    def choose(self, P, r):
        P = np.ascontiguousarray(P)
        n, d = P.shape
        S = np.empty_like(P)
        idx = np.empty(n, dtype=np.int64)
        k = self._check(self._fn("choose", P.dtype)(_p(P), C.c_long(n), C.c_int(d),
                                                      self._scalar(P.dtype, r), _p(S), _p(idx)), "choose")
        return S[:k].copy(), idx[:k].copy()

    def evolve_iter(self, P, S, r):
        """One evolve iteration.  Returns (S', flag, counts, sums, labels, moved);
        labels/moved are None for the reference (it does not expose them)."""
        P = np.ascontiguousarray(P)
        S = np.array(S, dtype=P.dtype, copy=True, order="C")
        n, d = P.shape
        K = S.shape[0]
        counts = np.zeros(max(K, 1), dtype=np.int32)
        sums = np.zeros((max(K, 1), d), dtype=P.dtype)
        if self.prefix == "oracle_":
            labels = np.empty(n, dtype=np.int32)
            moved = np.zeros((max(K, 1), d), dtype=np.uint8)
            flag = self._fn("evolve_iter", P.dtype)(_p(P), C.c_long(n), _p(S), C.c_long(K), C.c_int(d),
                                                     self._scalar(P.dtype, r), _p(labels), _p(counts),
                                                     _p(sums), _p(moved))
            return S, int(flag), counts[:K], sums[:K], labels, moved[:K]
        flag = self._fn("evolve_iter", P.dtype)(_p(P), C.c_long(n), _p(S), C.c_long(K), C.c_int(d),
                                                 self._scalar(P.dtype, r), _p(counts), _p(sums))
        return S, int(flag), counts[:K], sums[:K], None, None

    def evolve_omp_iter(self, P, S, r):
        P = np.ascontiguousarray(P)
        S = np.array(S, dtype=P.dtype, copy=True, order="C")
        n, d = P.shape
        K = S.shape[0]
        counts = np.zeros(max(K, 1), dtype=np.int32)
        sums = np.zeros((max(K, 1), d), dtype=P.dtype)
        flag = self._fn("evolve_omp_iter", P.dtype)(_p(P), C.c_long(n), _p(S), C.c_long(K), C.c_int(d),
                                                     self._scalar(P.dtype, r), _p(counts), _p(sums))
        return S, int(flag)

    def evolve(self, P, S, r, max_iter=0):
        P = np.ascontiguousarray(P)
        S = np.array(S, dtype=P.dtype, copy=True, order="C")
        n, d = P.shape
        if self.prefix == "oracle_":
            it = self._fn("evolve", P.dtype)(_p(P), C.c_long(n), _p(S), C.c_long(S.shape[0]), C.c_int(d),
                                              self._scalar(P.dtype, r), C.c_long(max_iter))
        else:
            it = self._fn("evolve", P.dtype)(_p(P), C.c_long(n), _p(S), C.c_long(S.shape[0]), C.c_int(d),
                                              self._scalar(P.dtype, r))
        return S, int(it)

    def order(self, S, r2):
        """Chain ordering of a ridge (rd::orderSamples with the r2 condition): returns
        (order, offsets) -- point indices chain after chain, and the chain boundaries."""
        S = np.ascontiguousarray(S)
        K, d = S.shape
        order = np.empty(max(K, 1), dtype=np.int64)
        offsets = np.empty(K + 1, dtype=np.int64)
        nc = self._check(self._fn("order", S.dtype)(_p(S), C.c_long(K), C.c_int(d), self._scalar(S.dtype, r2),
                                                     _p(order), _p(offsets)), "order")
        return order[:K].copy(), offsets[:nc + 1].copy()

    # -- quality against an ideal pattern (rd/utils/assessment_quality.hpp) -----------
    def hausdorff(self, A, B):
        A = np.ascontiguousarray(A)
        B = np.ascontiguousarray(B, dtype=A.dtype)
        out = np.zeros(1, dtype=A.dtype)
        self._check(self._fn("hausdorff", A.dtype)(_p(A), C.c_long(len(A)), _p(B), C.c_long(len(B)),
                                                    C.c_int(A.shape[1]), _p(out)), "hausdorff")
        return out[0]

    def distance_stats(self, A, B):
        """(median, average, min, max) of the distances from every point of A to B."""
        A = np.ascontiguousarray(A)
        B = np.ascontiguousarray(B, dtype=A.dtype)
        out = np.zeros(4, dtype=A.dtype)
        self._check(self._fn("distance_stats", A.dtype)(_p(A), C.c_long(len(A)), _p(B), C.c_long(len(B)),
                                                         C.c_int(A.shape[1]), _p(out)), "distance_stats")
        return out

    def set_distances(self, A, B):
        A = np.ascontiguousarray(A)
        B = np.ascontiguousarray(B, dtype=A.dtype)
        out = np.zeros(len(A), dtype=A.dtype)
        self._check(self._fn("set_distances", A.dtype)(_p(A), C.c_long(len(A)), _p(B), C.c_long(len(B)),
                                                        C.c_int(A.shape[1]), _p(out)), "set_distances")
        return out

    # -- differentiated evolve (frozen memberships; restatement only) ---------------
    def evolve_tape(self, P, S, r, iters):
        """`iters` evolve iterations from S recording what the frozen map needs.
        Returns (S_T, labels [iters, n] int32, counts [iters, K] int32, moved [iters, K, d] uint8)."""
        P = np.ascontiguousarray(P)
        S = np.array(S, dtype=P.dtype, copy=True, order="C")
        n, d = P.shape
        K = S.shape[0]
        labels = np.empty((iters, n), dtype=np.int32)
        counts = np.empty((iters, K), dtype=np.int32)
        moved = np.empty((iters, K, d), dtype=np.uint8)
        self._check(self._fn("evolve_tape", P.dtype)(_p(P), C.c_long(n), _p(S), C.c_long(K), C.c_int(d),
                                                      self._scalar(P.dtype, r), C.c_long(iters), _p(labels),
                                                      _p(counts), _p(moved)), "evolve_tape")
        return S, labels, counts, moved

    def evolve_replay(self, P, S0, labels, moved):
        """The frozen-membership evolve map evaluated at (P, S0): returns S_T."""
        P = np.ascontiguousarray(P)
        S = np.array(S0, dtype=P.dtype, copy=True, order="C")
        n, d = P.shape
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        moved = np.ascontiguousarray(moved, dtype=np.uint8)
        self._check(self._fn("evolve_replay", P.dtype)(_p(P), C.c_long(n), _p(S), C.c_long(S.shape[0]), C.c_int(d),
                                                        C.c_long(labels.shape[0]), _p(labels), _p(moved)),
                    "evolve_replay")
        return S

    def evolve_vjp(self, labels, counts, moved, G, dtype=np.float64):
        """Reverse mode of the frozen map: returns (dL/dP [n, d], dL/dS_0 [K, d]) for
        G = dL/dS_T [K, d], in `dtype`."""
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        moved = np.ascontiguousarray(moved, dtype=np.uint8)
        iters, n = labels.shape
        K, d = G.shape
        g = np.array(G, dtype=dtype, copy=True, order="C")
        gp = np.empty((n, d), dtype=dtype)
        self._check(self._fn("evolve_vjp", dtype)(C.c_long(n), C.c_long(K), C.c_int(d), C.c_long(iters),
                                                   _p(labels), _p(counts), _p(moved), _p(g), _p(gp)),
                    "evolve_vjp")
        return gp, g

    def decimate(self, S, r2, nocopy=False):
        """decimate (or, nocopy=True, decimateNoCopy over one list); returns the kept points
        in order and their indices."""
        S = np.array(S, copy=True, order="C")
        K, d = S.shape
        kept = np.empty(max(K, 1), dtype=np.int64)
        if self.prefix == "oracle_":
            k = self._fn("decimate", S.dtype)(_p(S), C.c_long(K), C.c_int(d), self._scalar(S.dtype, r2),
                                               C.c_int(1 if nocopy else 0), _p(kept))
        else:
            k = self._fn("decimate", S.dtype)(_p(S), C.c_long(K), C.c_int(d), self._scalar(S.dtype, r2),
                                               _p(kept))
        return S[:k].copy(), kept[:k].copy()

    def decimate_nocopy(self, S, r2, list_sizes):
        """Reference-only: decimateNoCopy over consecutive lists; returns the alive mask."""
        S = np.array(S, copy=True, order="C")
        K, d = S.shape
        sizes = np.asarray(list_sizes, dtype=np.int64)
        alive = np.zeros(max(K, 1), dtype=np.uint8)
        k = self._fn("decimate_nocopy", S.dtype)(_p(S), C.c_long(K), C.c_int(d), self._scalar(S.dtype, r2),
                                                  C.c_int(len(sizes)), _p(sizes), _p(alive))
        return int(k), alive[:K].astype(bool)

    def philox(self, ctr, key):
        """Philox4x32-10 of a 4-word counter and a 2-word key (oracle only)."""
        c = np.ascontiguousarray(ctr, dtype=np.uint32)
        k = np.ascontiguousarray(key, dtype=np.uint32)
        out = np.zeros(4, dtype=np.uint32)
        self.lib.oracle_philox4x32_10(_p(c), _p(k), _p(out))
        return out

    def generate(self, kind, n, dim, params=(), sigma=0.0, seed=0, stream=0, dtype=np.float32):
        """The seeded generator's n samples (oracle only), (n, dim) ROW_MAJOR."""
        out = np.empty((max(int(n), 1), dim), dtype=dtype)
        prm = np.ascontiguousarray(list(params) + [0.0], dtype=np.float64)
        self._check(self._fn("generate", np.dtype(dtype))(C.c_int(kind), C.c_long(n), C.c_int(dim), _p(prm),
                                                            C.c_double(sigma), C.c_uint64(seed), C.c_uint32(stream),
                                                            _p(out)), "generate")
        return out[:int(n)]

    def ridge_detection(self, P, r1, r2, max_evolve_iter=0, omp=False):
        P = np.ascontiguousarray(P)
        n, d = P.shape
        S = np.empty_like(P)
        if self.prefix == "oracle_":
            rounds = C.c_long(0)
            k = self._check(self._fn("rd", P.dtype)(_p(P), C.c_long(n), C.c_int(d), self._scalar(P.dtype, r1),
                                                     self._scalar(P.dtype, r2), C.c_long(max_evolve_iter),
                                                     _p(S), C.byref(rounds)), "rd")
            return S[:k].copy(), int(rounds.value)
        k = self._fn("rd", P.dtype)(_p(P), C.c_long(n), C.c_int(d), self._scalar(P.dtype, r1),
                                     self._scalar(P.dtype, r2), _p(S), C.c_int(1 if omp else 0))
        return S[:k].copy(), -1

    def ridge_detection_own(self, P, n_own, r1, r2, max_evolve_iter=0):
        """The sample-list overload: choose over P[:n_own], evolve over all of P (restatement only)."""
        P = np.ascontiguousarray(P)
        n, d = P.shape
        S = np.empty_like(P)
        rounds = C.c_long(0)
        k = self._check(self._fn("rd_own", P.dtype)(_p(P), C.c_long(n), C.c_long(n_own), C.c_int(d),
                                                     self._scalar(P.dtype, r1), self._scalar(P.dtype, r2),
                                                     C.c_long(max_evolve_iter), _p(S), C.byref(rounds)), "rd_own")
        return S[:k].copy(), int(rounds.value)

    def tiled(self, alg, P, r1, r2, max_tile_capacity, tiles_per_dim, refinement=False,
              max_evolve_iter=0, threads=1):
        P = np.ascontiguousarray(P)
        n, d = P.shape
        tpd = np.asarray(tiles_per_dim, dtype=np.int64)
        cap = 4 * n + 16
        S = np.empty((cap, d), dtype=P.dtype)
        if self.prefix == "oracle_":
            k = self._check(self._fn("tiled", P.dtype)(C.c_int(alg), _p(P), C.c_long(n), C.c_int(d),
                                                        self._scalar(P.dtype, r1), self._scalar(P.dtype, r2),
                                                        C.c_long(max_tile_capacity), _p(tpd),
                                                        C.c_int(int(refinement)), C.c_long(max_evolve_iter),
                                                        _p(S), C.c_long(cap)), "tiled")
            return S[:k].copy()
        times = np.zeros(4, dtype=np.float32)
        k = self._fn("tiled", P.dtype)(C.c_int(alg), _p(P), C.c_long(n), C.c_int(d),
                                        self._scalar(P.dtype, r1), self._scalar(P.dtype, r2),
                                        C.c_long(max_tile_capacity), _p(tpd), C.c_int(int(refinement)),
                                        C.c_int(threads), _p(S), C.c_long(cap), _p(times))
        if k < 0:
            raise RuntimeError("reference tiled run failed")
        self.last_times_ms = times
        return S[:k].copy()

    def tree(self, kind, P, r1, max_tile_capacity, tiles_per_dim):
        """Returns a list of leaves (preorder): dict(own, nb, cap, bounds)."""
        P = np.ascontiguousarray(P)
        n, d = P.shape
        tpd = np.asarray(tiles_per_dim, dtype=np.int64)
        f = self._fn("tree", P.dtype)
        args = (C.c_int(kind), _p(P), C.c_long(n), C.c_int(d), self._scalar(P.dtype, r1),
                C.c_long(max_tile_capacity), _p(tpd))
        L = self._check(f(*args, None, C.c_long(0), None, C.c_long(0), None, C.c_long(0), None), "tree")
        info = np.zeros((max(L, 1), 4), dtype=np.int64)
        own = np.zeros(n + 1, dtype=np.int64)
        nbcap = 64 * n + 1024
        nb = np.zeros(nbcap, dtype=np.int64)
        bounds = np.zeros((max(L, 1), 2 * d), dtype=P.dtype)
        L = self._check(f(*args, _p(info), C.c_long(L), _p(own), C.c_long(n + 1), _p(nb), C.c_long(nbcap),
                          _p(bounds)), "tree")
        leaves = []
        o = b = 0
        for l in range(L):
            no, nn, cap, _ = info[l]
            leaves.append(dict(own=own[o:o + no].copy(), nb=nb[b:b + nn].copy(), cap=int(cap),
                               bounds=bounds[l].reshape(d, 2).copy()))
            o += no
            b += nn
        return leaves
