// rd_tree.hip -- level-synchronous construction of the tile trees on the device.
//
// Every level handles all the tiles being split at once:
//   1. bin      -- the child of every own sample (histogram map for LOCAL, half-open cell
//                  search for GROUPED / EXTENDED), then a stable sort by child id;
//   2. nearby   -- EXTENDED only: (child cell, sample) pairs for every sample of
//                  (parent own, parent neighbours) that is near the cell, counted, emitted
//                  and stably sorted by child id;
//   3. bounds   -- LOCAL only: tight bounds of every child's own samples;
// after the last level, GROUPED leaves take their neighbour lists from a device bit matrix
// of every pair's closed-box test;
// and the host keeps the skeleton (counts, bounds, leaf/split decisions).  Cell bounds are
// computed on the host with the reference's expressions (min + c * step, last cell ending at
// the parent max) and uploaded as per-dimension interval tables, so device and host agree
// bit for bit.
#include "rd_tree.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

namespace rdhip {

namespace {

template <typename T>
struct ParentDev {
    const int* idx;      // own samples: idx[off .. off+cnt)
    int64_t off, cnt;
    int64_t outOff;      // position of its own samples in the level's key array
    const int* nbIdx;    // EXTENDED neighbour samples
    int64_t nbOff, nbCnt;
    int64_t posOff;      // position of (own, nb) in the level's nearby-count array
    int childBase;
    int ivOff;           // first interval of dimension 0 in the interval tables
    int nbins[kMaxDim];
    int stride[kMaxDim];
    T mn[kMaxDim];
    T dist[kMaxDim];
};

struct RangeCopy {
    const int* src;
    int64_t srcOff, cnt, dstOff;
};

template <typename T>
__device__ int bin_local(const T* x, const ParentDev<T>& pd, int dim) {
    const T eps = FpTraits<T>::eps;
    int64_t id = 0, total = 1;
    for (int d = 0; d < dim; ++d) {
        const int n = pd.nbins[d];
        int64_t b = 0;
        if (!(pd.dist[d] < eps)) {
            const T normCord = fabs(x[d] - pd.mn[d]);
            const T step = pd.dist[d] / T(n);
            if (fabs(normCord - pd.dist[d]) <= eps || normCord - pd.dist[d] > 0)
                b = n - 1;
            else
                b = (int64_t)floor(normCord / step);
        }
        id += b * pd.stride[d];
        total *= n;
    }
    return (int)(id < total ? id : total - 1);
}

// index of the half-open interval [lo, hi) holding v, or -1
template <typename T>
__device__ int find_interval(const T* lo, const T* hi, int n, T v) {
    // last c with lo[c] <= v (lo is non-decreasing)
    int a = 0, b = n - 1, c = -1;
    while (a <= b) {
        const int m = (a + b) >> 1;
        if (lo[m] <= v) {
            c = m;
            a = m + 1;
        } else {
            b = m - 1;
        }
    }
    if (c >= 0 && v < hi[c]) return c;
    return -1;
}

template <typename T>
__device__ int bin_cell(const T* x, const ParentDev<T>& pd, const T* ivLo, const T* ivHi, int dim) {
    int id = 0, o = pd.ivOff;
    for (int d = 0; d < dim; ++d) {
        const int c = find_interval(ivLo + o, ivHi + o, pd.nbins[d], x[d]);
        if (c < 0) return -1;
        id += c * pd.stride[d];
        o += pd.nbins[d];
    }
    return id;
}

// per-dimension range of cells whose r-expanded interval holds v: [c0, c1]
template <typename T>
__device__ void near_range(const T* lo, const T* hi, int n, T v, T r, int& c0, int& c1) {
    // first c with !(v >= hi[c] + r)  (hi non-decreasing)
    int a = 0, b = n - 1, f = n;
    while (a <= b) {
        const int m = (a + b) >> 1;
        if (!(v >= hi[m] + r)) {
            f = m;
            b = m - 1;
        } else {
            a = m + 1;
        }
    }
    // last c with !(v < lo[c] - r)
    a = 0;
    b = n - 1;
    int l = -1;
    while (a <= b) {
        const int m = (a + b) >> 1;
        if (!(v < lo[m] - r)) {
            l = m;
            a = m + 1;
        } else {
            b = m - 1;
        }
    }
    c0 = f;
    c1 = l;
}

template <typename T, bool COUNT>
__device__ int near_cells(const T* x, const ParentDev<T>& pd, const T* ivLo, const T* ivHi, int dim, T r,
                          int* outKeys) {
    int c0[kMaxDim], c1[kMaxDim], c[kMaxDim];
    int o = pd.ivOff;
    int inside = 1;
    int own[kMaxDim];
    for (int d = 0; d < dim; ++d) {
        near_range(ivLo + o, ivHi + o, pd.nbins[d], x[d], r, c0[d], c1[d]);
        if (c0[d] > c1[d]) return 0;
        own[d] = find_interval(ivLo + o, ivHi + o, pd.nbins[d], x[d]);
        inside &= own[d] >= 0;
        c[d] = c0[d];
        o += pd.nbins[d];
    }
    int emitted = 0;
    for (;;) {
        bool isOwn = inside;
        int id = 0;
        for (int d = 0; d < dim; ++d) {
            isOwn = isOwn && c[d] == own[d];
            id += c[d] * pd.stride[d];
        }
        if (!isOwn) {
            if (!COUNT) outKeys[emitted] = pd.childBase + id;
            ++emitted;
        }
        int d = 0;
        for (; d < dim; ++d) {
            if (++c[d] <= c1[d]) break;
            c[d] = c0[d];
        }
        if (d == dim) break;
    }
    return emitted;
}

template <typename T>
__global__ __launch_bounds__(256) void iota_kernel(int* __restrict__ a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) a[i] = (int)i;
}

template <typename T>
__global__ __launch_bounds__(256) void bin_kernel(const T* __restrict__ P, int64_t ps, int dim, int local,
                                                  const ParentDev<T>* __restrict__ parents,
                                                  const WorkItem* __restrict__ items, const T* __restrict__ ivLo,
                                                  const T* __restrict__ ivHi, int* __restrict__ keys,
                                                  int* __restrict__ vals) {
    const WorkItem it = items[blockIdx.x];
    const ParentDev<T>& pd = parents[it.seg];
    for (int i = it.start + threadIdx.x; i < it.end; i += 256) {
        const int p = pd.idx[pd.off + i];
        T x[kMaxDim];
        for (int d = 0; d < dim; ++d) x[d] = P[d * ps + p];
        const int b = local ? bin_local(x, pd, dim) : bin_cell(x, pd, ivLo, ivHi, dim);
        keys[pd.outOff + i] = b >= 0 ? pd.childBase + b : -1;
        vals[pd.outOff + i] = p;
    }
}

// EXTENDED: positions of a parent are its own samples then its neighbour samples
template <typename T, bool COUNT>
__global__ __launch_bounds__(256) void nearby_kernel(const T* __restrict__ P, int64_t ps, int dim, T r,
                                                     const ParentDev<T>* __restrict__ parents,
                                                     const WorkItem* __restrict__ items, const T* __restrict__ ivLo,
                                                     const T* __restrict__ ivHi, int* __restrict__ counts,
                                                     int* __restrict__ keys, int* __restrict__ vals) {
    const WorkItem it = items[blockIdx.x];
    const ParentDev<T>& pd = parents[it.seg];
    for (int i = it.start + threadIdx.x; i < it.end; i += 256) {
        const int p = i < pd.cnt ? pd.idx[pd.off + i] : pd.nbIdx[pd.nbOff + (i - pd.cnt)];
        T x[kMaxDim];
        for (int d = 0; d < dim; ++d) x[d] = P[d * ps + p];
        const int64_t pos = pd.posOff + i;
        if (COUNT) {
            counts[pos] = near_cells<T, true>(x, pd, ivLo, ivHi, dim, r, nullptr);
        } else {
            const int o = counts[pos];
            const int m = near_cells<T, false>(x, pd, ivLo, ivHi, dim, r, keys + o);
            for (int q = 0; q < m; ++q) vals[o + q] = p;
        }
    }
}

// Tight bounds of the whole cloud P[:, 0..n): per-block partial min/max (order-free, so
// exact), then one block combines the partials.
template <typename T>
__device__ __forceinline__ void block_minmax(T& mn, T& mx, T* smn, T* smx) {
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const T a = smn[threadIdx.x + h], b = smx[threadIdx.x + h];
            smn[threadIdx.x] = (a < smn[threadIdx.x]) ? a : smn[threadIdx.x];
            smx[threadIdx.x] = (b > smx[threadIdx.x]) ? b : smx[threadIdx.x];
        }
        __syncthreads();
    }
    mn = smn[0];
    mx = smx[0];
    __syncthreads();
}

// Tight bounds of the kept children of a level, one block per item of up to kBoundItem
// sorted positions: per-item partial min / max with the reference's comparisons.  Min and max
// are order-free, so folding the partials equals one pass over the child, and a large child
// is reduced by many blocks.
constexpr int kBoundItem = 4096;

template <typename T>
__global__ __launch_bounds__(256) void item_bounds_kernel(const T* __restrict__ P, int64_t ps, int dim,
                                                          const int* __restrict__ idx,
                                                          const WorkItem* __restrict__ items, T* __restrict__ partial) {
    __shared__ T smn[256], smx[256];
    const WorkItem it = items[blockIdx.x];
    T mn[kMaxDim], mx[kMaxDim];
#pragma unroll
    for (int d = 0; d < kMaxDim; ++d) {
        mn[d] = FpTraits<T>::big;
        mx[d] = -FpTraits<T>::big;
    }
    for (int i = it.start + threadIdx.x; i < it.end; i += 256) {
        const int p = idx[i];
#pragma unroll
        for (int d = 0; d < kMaxDim; ++d) {
            if (d >= dim) break;
            const T v = P[d * ps + p];
            mn[d] = (v < mn[d]) ? v : mn[d];
            mx[d] = (v > mx[d]) ? v : mx[d];
        }
    }
#pragma unroll
    for (int d = 0; d < kMaxDim; ++d) {
        if (d >= dim) break;   // dim is uniform over the block: every thread meets the same barriers
        block_minmax(mn[d], mx[d], smn, smx);
        if (threadIdx.x == 0) {
            partial[(int64_t)blockIdx.x * 2 * dim + 2 * d] = mn[d];
            partial[(int64_t)blockIdx.x * 2 * dim + 2 * d + 1] = mx[d];
        }
    }
}

// GROUPED neighbours: the closed boxes of leaves l and j overlap when, in every dimension,
// an endpoint of one lies within the other's [mn, mx].  Bit j of row l is that test for
// j != l (rows row0 .. row0 + rows, `words` 64-bit words a row).  A wave evaluates the 64
// candidates of one word and writes its ballot; words * 64 threads a row keep every wave
// inside one row.
template <typename T>
__global__ __launch_bounds__(256) void leaf_overlap_kernel(const T* __restrict__ mn, const T* __restrict__ mx, int L,
                                                           int dim, int row0, int rows, int words,
                                                           unsigned long long* __restrict__ bits) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t perRow = (int64_t)words * 64;
    const int64_t r = t / perRow;
    if (r >= rows) return;
    const int j = (int)(t - r * perRow);
    const int l = row0 + (int)r;
    bool hit = j < L && j != l;
    for (int d = 0; d < dim && hit; ++d) {
        const T amn = mn[(int64_t)l * dim + d], amx = mx[(int64_t)l * dim + d];
        const T bmn = mn[(int64_t)j * dim + d], bmx = mx[(int64_t)j * dim + d];
        hit = (bmn >= amn && bmn <= amx) || (bmx >= amn && bmx <= amx) || (amn >= bmn && amn <= bmx) ||
              (amx >= bmn && amx <= bmx);
    }
    const unsigned long long w = __ballot(hit);
    if (lane_id() == 0) bits[r * words + j / 64] = w;
}

template <typename T>
__global__ __launch_bounds__(256) void cloud_bounds_partial_kernel(const T* __restrict__ P, int64_t ps, int dim,
                                                                   int64_t n, T* __restrict__ partial) {
    __shared__ T smn[256], smx[256];
    for (int d = 0; d < dim; ++d) {
        T mn = FpTraits<T>::big, mx = -FpTraits<T>::big;
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
            const T v = P[d * ps + i];
            mn = (v < mn) ? v : mn;
            mx = (v > mx) ? v : mx;
        }
        block_minmax(mn, mx, smn, smx);
        if (threadIdx.x == 0) {
            partial[((int64_t)blockIdx.x * dim + d) * 2] = mn;
            partial[((int64_t)blockIdx.x * dim + d) * 2 + 1] = mx;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void cloud_bounds_final_kernel(const T* __restrict__ partial, int nb, int dim,
                                                                 T* __restrict__ out) {
    __shared__ T smn[256], smx[256];
    for (int d = 0; d < dim; ++d) {
        T mn = FpTraits<T>::big, mx = -FpTraits<T>::big;
        for (int b = threadIdx.x; b < nb; b += 256) {
            const T a = partial[((int64_t)b * dim + d) * 2], c = partial[((int64_t)b * dim + d) * 2 + 1];
            mn = (a < mn) ? a : mn;
            mx = (c > mx) ? c : mx;
        }
        block_minmax(mn, mx, smn, smx);
        if (threadIdx.x == 0) {
            out[2 * d] = mn;
            out[2 * d + 1] = mx;
        }
    }
}

__global__ __launch_bounds__(256) void copy_ranges_kernel(const RangeCopy* __restrict__ rc, int* __restrict__ dst) {
    const RangeCopy r = rc[blockIdx.x];
    for (int64_t i = threadIdx.x; i < r.cnt; i += 256) dst[r.dstOff + i] = r.src[r.srcOff + i];
}

__global__ void zero_kernel(int* __restrict__ a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) a[i] = 0;
}

inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

template <typename T>
void calc_dist(const std::vector<T>& mn, const std::vector<T>& mx, std::vector<T>& dist) {
    dist.resize(mn.size());
    for (size_t k = 0; k < mn.size(); ++k) {
        dist[k] = std::abs(mx[k] - mn[k]);
        if (dist[k] <= std::numeric_limits<T>::epsilon()) dist[k] = 0;
    }
}

template <typename T>
size_t count_spheres(const std::vector<T>& dist, T r) {
    size_t cnt = 1;
    for (T dv : dist) {
        size_t dc = static_cast<size_t>(std::ceil(dv / r));
        cnt *= dc ? dc + 1 : 1;
    }
    return cnt;
}

}  // namespace

template <typename T>
TileTreeBuilder<T>::~TileTreeBuilder() = default;

template <typename T>
void TileTreeBuilder<T>::preorder(int t) {
    if (tiles_[t].leaf)
        leafTiles_.push_back(t);
    else
        for (int c : tiles_[t].children) preorder(c);
}

template <typename T>
void TileTreeBuilder<T>::build(const T* P, int64_t ps, int64_t n, int kind, T r1, int64_t cap,
                               const std::vector<int64_t>& tpd) {
    kind_ = kind;
    r1_ = r1;
    cap_ = cap;
    tiles_.clear();
    rootChildren_.clear();
    bufs_.clear();
    leafTiles_.clear();
    leaves_.clear();
    const int dim = dim_;
    const bool local = kind == TREE_LOCAL;
    const bool ext = kind == TREE_EXTENDED;
    if ((int)tpd.size() != dim) throw std::invalid_argument("tiles_per_dim needs one entry per dimension");

    // root: identity list + global tight bounds
    bufs_.emplace_back(new DevBuf<int>());
    int* ident = bufs_[0]->reserve(std::max<int64_t>(n, 1));
    hipLaunchKernelGGL((iota_kernel<T>), dim3(grid_of(n)), dim3(256), 0, st_, ident, n);
    HTile root;
    {
        const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 4095) / 4096));
        DevBuf<T> out, partial;
        out.reserve(2 * dim);
        partial.reserve((size_t)nb * 2 * dim);
        hipLaunchKernelGGL((cloud_bounds_partial_kernel<T>), dim3(nb), dim3(256), 0, st_, P, ps, dim, n,
                           partial.get());
        hipLaunchKernelGGL((cloud_bounds_final_kernel<T>), dim3(1), dim3(256), 0, st_, partial.get(), nb, dim,
                           out.get());
        std::vector<T> hb(2 * dim);
        check(hipMemcpyAsync(hb.data(), out.get(), sizeof(T) * 2 * dim, hipMemcpyDeviceToHost, st_), "bounds");
        check(hipStreamSynchronize(st_), "bounds");
        root.mn.resize(dim);
        root.mx.resize(dim);
        for (int d = 0; d < dim; ++d) {
            root.mn[d] = hb[2 * d];
            root.mx[d] = hb[2 * d + 1];
        }
        calc_dist(root.mn, root.mx, root.dist);
    }
    root.level = 0;
    root.ownBuf = 0;
    root.ownOff = 0;
    root.ownCnt = n;
    tiles_.push_back(root);

    std::vector<int> parents{0};
    levels_ = 0;
    DevBuf<int> zeroFlag;
    zeroFlag.reserve(1);
    {
        int z = 0;
        check(hipMemcpyAsync(zeroFlag.get(), &z, sizeof(int), hipMemcpyHostToDevice, st_), "flag");
    }
    // Level scratch, reused by every level (each level's sorted own lists stay with its tiles
    // in bufs_).
    DevBuf<ParentDev<T>> dPar;
    DevBuf<WorkItem> dOwnItems, dNearItems, dBoundItems;
    DevBuf<T> dLo, dHi, dPartial;
    DevBuf<int> keys, vals, sortedKeys, cOff, cEnd, counts, partials, pk, pv, sk;
    DevBuf<int64_t> dN, dM, pairTotal, pn, pm;
    DevBuf<unsigned char> ws, ws2;
    while (!parents.empty()) {
        if (levels_ > 64) throw std::runtime_error("tile tree exceeds 64 levels (identical points?)");
        // ---- parent descriptors and cell interval tables
        std::vector<ParentDev<T>> pdev(parents.size());
        std::vector<T> ivLo, ivHi;
        std::vector<WorkItem> ownItems, nearItems;
        std::vector<int> childBase(parents.size()), childCount(parents.size());
        std::vector<std::vector<int>> nbinsOf(parents.size());
        int64_t ownTotal = 0, posTotal = 0;
        int totalChildren = 0;
        for (size_t q = 0; q < parents.size(); ++q) {
            const HTile& t = tiles_[parents[q]];
            std::vector<int> nb(dim, 1);
            if (parents[q] == 0)
                for (int d = 0; d < dim; ++d) nb[d] = (int)tpd[d];
            else
                nb[t.level % dim] = 2;
            nbinsOf[q] = nb;
            ParentDev<T>& pd = pdev[q];
            std::memset(&pd, 0, sizeof(pd));
            pd.idx = bufs_[t.ownBuf]->get();
            pd.off = t.ownOff;
            pd.cnt = t.ownCnt;
            pd.outOff = ownTotal;
            pd.nbIdx = t.nbBuf >= 0 ? bufs_[t.nbBuf]->get() : nullptr;
            pd.nbOff = t.nbOff;
            pd.nbCnt = ext ? t.nbCnt : 0;
            pd.posOff = posTotal;
            pd.childBase = totalChildren;
            pd.ivOff = (int)ivLo.size();
            int stride = 1;
            int nch = 1;
            for (int d = 0; d < dim; ++d) {
                pd.nbins[d] = nb[d];
                pd.stride[d] = stride;
                stride *= nb[d];
                nch *= nb[d];
                pd.mn[d] = t.mn[d];
                pd.dist[d] = t.dist[d];
            }
            if (!local) {
                // allocTiles (grouped_tile_tree.inl:144-211): [min + c*step, min + (c+1)*step),
                // the last cell of a dimension ending at the parent max
                for (int d = 0; d < dim; ++d) {
                    const T step = t.dist[d] / static_cast<T>(nb[d]);
                    for (int c = 0; c < nb[d]; ++c) {
                        ivLo.push_back(t.mn[d] + c * step);
                        ivHi.push_back(c == nb[d] - 1 ? t.mx[d] : t.mn[d] + (c + 1) * step);
                    }
                }
            }
            childBase[q] = totalChildren;
            childCount[q] = nch;
            totalChildren += nch;
            for (int64_t s = 0; s < t.ownCnt; s += 2048)
                ownItems.push_back({(int)q, (int)s, (int)std::min<int64_t>(t.ownCnt, s + 2048), 0});
            const int64_t np = t.ownCnt + (ext ? t.nbCnt : 0);
            for (int64_t s = 0; s < np; s += 2048)
                nearItems.push_back({(int)q, (int)s, (int)std::min<int64_t>(np, s + 2048), 0});
            ownTotal += t.ownCnt;
            posTotal += np;
        }
        if (ownTotal >= (int64_t(1) << 31) || posTotal >= (int64_t(1) << 31))
            throw std::runtime_error("tile level too large");
        dPar.reserve(pdev.size());
        check(hipMemcpyAsync(dPar.get(), pdev.data(), sizeof(ParentDev<T>) * pdev.size(), hipMemcpyHostToDevice, st_),
              "parents");
        dOwnItems.reserve(std::max<size_t>(ownItems.size(), 1));
        if (!ownItems.empty())
            check(hipMemcpyAsync(dOwnItems.get(), ownItems.data(), sizeof(WorkItem) * ownItems.size(),
                                 hipMemcpyHostToDevice, st_),
                  "items");
        dLo.reserve(std::max<size_t>(ivLo.size(), 1));
        dHi.reserve(std::max<size_t>(ivHi.size(), 1));
        if (!ivLo.empty()) {
            check(hipMemcpyAsync(dLo.get(), ivLo.data(), sizeof(T) * ivLo.size(), hipMemcpyHostToDevice, st_), "iv");
            check(hipMemcpyAsync(dHi.get(), ivHi.data(), sizeof(T) * ivHi.size(), hipMemcpyHostToDevice, st_), "iv");
        }

        // ---- own samples: bin + stable sort by child
        keys.reserve(std::max<int64_t>(ownTotal, 1));
        vals.reserve(std::max<int64_t>(ownTotal, 1));
        if (!ownItems.empty())
            hipLaunchKernelGGL((bin_kernel<T>), dim3((unsigned)ownItems.size()), dim3(256), 0, st_, P, ps, dim,
                               local ? 1 : 0, dPar.get(), dOwnItems.get(), dLo.get(), dHi.get(), keys.get(),
                               vals.get());
        check(hipGetLastError(), "bin");
        const int levelOwnBuf = (int)bufs_.size();
        bufs_.emplace_back(new DevBuf<int>());
        int* sortedVals = bufs_.back()->reserve(std::max<int64_t>(ownTotal, 1));
        sortedKeys.reserve(std::max<int64_t>(ownTotal, 1));
        dN.reserve(1);
        dM.reserve(1);
        check(hipMemcpyAsync(dN.get(), &ownTotal, sizeof(int64_t), hipMemcpyHostToDevice, st_), "n");
        ws.reserve(sort_workspace_bytes(std::max<int64_t>(ownTotal, 1), std::max(totalChildren, 2)));
        if (ownTotal > 0)
            check(stable_sort_by_key(keys.get(), vals.get(), dN.get(), ownTotal, std::max(totalChildren, 2),
                                     sortedKeys.get(), sortedVals, dM.get(), ws.get(), st_),
                  "sort own");
        cOff.reserve(totalChildren);
        cEnd.reserve(totalChildren);
        hipLaunchKernelGGL(zero_kernel, dim3(grid_of(totalChildren)), dim3(256), 0, st_, cOff.get(),
                           (int64_t)totalChildren);
        hipLaunchKernelGGL(zero_kernel, dim3(grid_of(totalChildren)), dim3(256), 0, st_, cEnd.get(),
                           (int64_t)totalChildren);
        if (ownTotal > 0)
            check(launch_chain_bounds(sortedKeys.get(), dM.get(), ownTotal, cOff.get(), cEnd.get(), nullptr, 0, zeroFlag.get(), st_),
                  "runs");
        std::vector<int> hOff(totalChildren), hEnd(totalChildren);
        check(hipMemcpyAsync(hOff.data(), cOff.get(), sizeof(int) * totalChildren, hipMemcpyDeviceToHost, st_), "runs");
        check(hipMemcpyAsync(hEnd.data(), cEnd.get(), sizeof(int) * totalChildren, hipMemcpyDeviceToHost, st_), "runs");

        // ---- EXTENDED: nearby samples of every child cell
        std::vector<int> nOff(totalChildren, 0), nEnd(totalChildren, 0);
        int levelNbBuf = -1;
        if (ext && posTotal > 0) {
            counts.reserve(posTotal + 1);
            dNearItems.reserve(nearItems.size());
            check(hipMemcpyAsync(dNearItems.get(), nearItems.data(), sizeof(WorkItem) * nearItems.size(),
                                 hipMemcpyHostToDevice, st_),
                  "items");
            hipLaunchKernelGGL(zero_kernel, dim3(grid_of(posTotal + 1)), dim3(256), 0, st_, counts.get(), posTotal + 1);
            hipLaunchKernelGGL((nearby_kernel<T, true>), dim3((unsigned)nearItems.size()), dim3(256), 0, st_, P, ps,
                               dim, r1, dPar.get(), dNearItems.get(), dLo.get(), dHi.get(), counts.get(), nullptr,
                               nullptr);
            partials.reserve(scan_partials_count(posTotal + 1));
            pairTotal.reserve(1);
            check(exclusive_scan_inplace(counts.get(), posTotal + 1, partials.get(), pairTotal.get(), st_), "scan");
            int64_t pairs = 0;
            check(hipMemcpyAsync(&pairs, pairTotal.get(), sizeof(int64_t), hipMemcpyDeviceToHost, st_), "pairs");
            check(hipStreamSynchronize(st_), "pairs");
            if (pairs >= (int64_t(1) << 31)) throw std::runtime_error("too many neighbour samples");
            levelNbBuf = (int)bufs_.size();
            bufs_.emplace_back(new DevBuf<int>());
            int* nbVals = bufs_.back()->reserve(std::max<int64_t>(pairs, 1));
            if (pairs > 0) {
                pk.reserve(pairs);
                pv.reserve(pairs);
                sk.reserve(pairs);
                hipLaunchKernelGGL((nearby_kernel<T, false>), dim3((unsigned)nearItems.size()), dim3(256), 0, st_, P,
                                   ps, dim, r1, dPar.get(), dNearItems.get(), dLo.get(), dHi.get(), counts.get(),
                                   pk.get(), pv.get());
                pn.reserve(1);
                pm.reserve(1);
                check(hipMemcpyAsync(pn.get(), &pairs, sizeof(int64_t), hipMemcpyHostToDevice, st_), "n");
                ws2.reserve(sort_workspace_bytes(pairs, std::max(totalChildren, 2)));
                check(stable_sort_by_key(pk.get(), pv.get(), pn.get(), pairs, std::max(totalChildren, 2), sk.get(), nbVals,
                                         pm.get(), ws2.get(), st_),
                      "sort nb");
                hipLaunchKernelGGL(zero_kernel, dim3(grid_of(totalChildren)), dim3(256), 0, st_, cOff.get(),
                                   (int64_t)totalChildren);
                hipLaunchKernelGGL(zero_kernel, dim3(grid_of(totalChildren)), dim3(256), 0, st_, cEnd.get(),
                                   (int64_t)totalChildren);
                check(launch_chain_bounds(sk.get(), pm.get(), pairs, cOff.get(), cEnd.get(), nullptr, 0, zeroFlag.get(), st_),
                      "runs");
                check(hipMemcpyAsync(nOff.data(), cOff.get(), sizeof(int) * totalChildren, hipMemcpyDeviceToHost, st_),
                      "runs");
                check(hipMemcpyAsync(nEnd.data(), cEnd.get(), sizeof(int) * totalChildren, hipMemcpyDeviceToHost, st_),
                      "runs");
            }
        }
        check(hipStreamSynchronize(st_), "level");

        // ---- LOCAL: tight bounds of the kept children
        std::vector<int> keep;
        for (int c = 0; c < totalChildren; ++c)
            if (hEnd[c] - hOff[c] >= 2) keep.push_back(c);
        std::vector<T> tb;
        if (local && !keep.empty()) {
            std::vector<WorkItem> bi;
            for (size_t k = 0; k < keep.size(); ++k)
                for (int s = hOff[keep[k]]; s < hEnd[keep[k]]; s += kBoundItem)
                    bi.push_back({(int)k, s, std::min(hEnd[keep[k]], s + kBoundItem), 0});
            dBoundItems.reserve(bi.size());
            check(hipMemcpyAsync(dBoundItems.get(), bi.data(), sizeof(WorkItem) * bi.size(), hipMemcpyHostToDevice,
                                 st_),
                  "items");
            dPartial.reserve(bi.size() * 2 * dim);
            hipLaunchKernelGGL((item_bounds_kernel<T>), dim3((unsigned)bi.size()), dim3(256), 0, st_, P, ps, dim,
                               sortedVals, dBoundItems.get(), dPartial.get());
            check(hipGetLastError(), "bounds");
            std::vector<T> part(bi.size() * 2 * dim);
            check(hipMemcpyAsync(part.data(), dPartial.get(), sizeof(T) * part.size(), hipMemcpyDeviceToHost, st_),
                  "bounds");
            check(hipStreamSynchronize(st_), "bounds");
            tb.resize(keep.size() * 2 * dim);
            for (size_t k = 0; k < keep.size(); ++k)
                for (int d = 0; d < dim; ++d) {
                    tb[k * 2 * dim + 2 * d] = FpTraits<T>::big;
                    tb[k * 2 * dim + 2 * d + 1] = -FpTraits<T>::big;
                }
            for (size_t b = 0; b < bi.size(); ++b) {
                T* t = &tb[(size_t)bi[b].seg * 2 * dim];
                const T* q = &part[b * 2 * dim];
                for (int d = 0; d < dim; ++d) {
                    t[2 * d] = (q[2 * d] < t[2 * d]) ? q[2 * d] : t[2 * d];
                    t[2 * d + 1] = (q[2 * d + 1] > t[2 * d + 1]) ? q[2 * d + 1] : t[2 * d + 1];
                }
            }
        }

        // ---- children (in each parent's linear order), leaf / split decisions
        std::vector<int> next;
        size_t kp = 0;
        for (size_t q = 0; q < parents.size(); ++q) {
            const int pt = parents[q];
            const int lvl = tiles_[pt].level;
            const std::vector<int>& nb = nbinsOf[q];
            for (int b = 0; b < childCount[q]; ++b) {
                const int c = childBase[q] + b;
                const int64_t cnt = hEnd[c] - hOff[c];
                if (cnt < 2) continue;
                HTile ch;
                ch.level = lvl + 1;
                ch.mn.resize(dim);
                ch.mx.resize(dim);
                if (local) {
                    for (int d = 0; d < dim; ++d) {
                        ch.mn[d] = tb[kp * 2 * dim + 2 * d];
                        ch.mx[d] = tb[kp * 2 * dim + 2 * d + 1];
                    }
                    ++kp;
                } else {
                    int rem = b, o = pdev[q].ivOff;
                    for (int d = 0; d < dim; ++d) {
                        const int cd = rem % nb[d];
                        rem /= nb[d];
                        ch.mn[d] = ivLo[o + cd];
                        ch.mx[d] = ivHi[o + cd];
                        o += nb[d];
                    }
                }
                calc_dist(ch.mn, ch.mx, ch.dist);
                ch.ownBuf = levelOwnBuf;
                ch.ownOff = hOff[c];
                ch.ownCnt = cnt;
                if (ext) {
                    ch.nbBuf = levelNbBuf;
                    ch.nbOff = nOff[c];
                    ch.nbCnt = nEnd[c] - nOff[c];
                }
                ch.leaf = cnt < cap;
                tiles_.push_back(ch);
                const int id = (int)tiles_.size() - 1;
                if (pt == 0)
                    rootChildren_.push_back(id);
                else
                    tiles_[pt].children.push_back(id);
                if (!ch.leaf) next.push_back(id);
            }
        }
        for (int pt : parents)
            if (pt != 0 && tiles_[pt].children.empty())
                throw std::runtime_error("a split tile lost all its samples");
        parents.swap(next);
        ++levels_;
    }

    // ---- leaves in preorder
    for (int c : rootChildren_) preorder(c);
    const int L = (int)leafTiles_.size();
    leaves_.resize(L);
    totalPositions_ = 0;
    for (int l = 0; l < L; ++l) {
        const HTile& t = tiles_[leafTiles_[l]];
        LeafInfo<T>& li = leaves_[l];
        li.ownCnt = t.ownCnt;
        li.mn = t.mn;
        li.mx = t.mx;
        li.sphereCap = count_spheres(t.dist, r1);
        li.nbCnt = ext ? t.nbCnt : 0;
    }
    if (kind == TREE_GROUPED && L > 0) {
        // every pair tested on the device, rows of the bit matrix decoded in increasing j
        std::vector<T> hmn((size_t)L * dim), hmx((size_t)L * dim);
        for (int l = 0; l < L; ++l)
            for (int d = 0; d < dim; ++d) {
                hmn[(size_t)l * dim + d] = leaves_[l].mn[d];
                hmx[(size_t)l * dim + d] = leaves_[l].mx[d];
            }
        DevBuf<T> dmn, dmx;
        dmn.reserve(hmn.size());
        dmx.reserve(hmx.size());
        check(hipMemcpyAsync(dmn.get(), hmn.data(), sizeof(T) * hmn.size(), hipMemcpyHostToDevice, st_), "boxes");
        check(hipMemcpyAsync(dmx.get(), hmx.data(), sizeof(T) * hmx.size(), hipMemcpyHostToDevice, st_), "boxes");
        const int words = (L + 63) / 64;
        const int rowsPer = (int)std::max<int64_t>(1, std::min<int64_t>(L, (int64_t(1) << 20) / words));   // 8 MB a block
        DevBuf<unsigned long long> dbits;
        dbits.reserve((size_t)rowsPer * words);
        std::vector<unsigned long long> hbits((size_t)rowsPer * words);
        for (int row0 = 0; row0 < L; row0 += rowsPer) {
            const int rows = std::min(rowsPer, L - row0);
            const int64_t threads = (int64_t)rows * words * 64;
            hipLaunchKernelGGL((leaf_overlap_kernel<T>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st_,
                               dmn.get(), dmx.get(), L, dim, row0, rows, words, dbits.get());
            check(hipGetLastError(), "overlap");
            check(hipMemcpyAsync(hbits.data(), dbits.get(), sizeof(unsigned long long) * rows * words,
                                 hipMemcpyDeviceToHost, st_),
                  "overlap");
            check(hipStreamSynchronize(st_), "overlap");
            for (int r = 0; r < rows; ++r) {
                LeafInfo<T>& li = leaves_[row0 + r];
                for (int w = 0; w < words; ++w)
                    for (unsigned long long m = hbits[(size_t)r * words + w]; m; m &= m - 1) {
                        const int j = w * 64 + __builtin_ctzll(m);
                        li.nbLeaves.push_back(j);
                        li.nbCnt += leaves_[j].ownCnt;
                    }
            }
        }
    }
    for (auto& li : leaves_) totalPositions_ += li.ownCnt + li.nbCnt;
    if (totalPositions_ >= (int64_t(1) << 31)) throw std::runtime_error("tiled layout exceeds 2^31 samples");
}

template <typename T>
void TileTreeBuilder<T>::writeIndex(int* dIndex) {
    const int L = (int)leaves_.size();
    // every leaf's own range, compact (a grouped layout repeats each one for every neighbour)
    std::vector<RangeCopy> own(L);
    size_t ranges = (size_t)L;
    for (int l = 0; l < L; ++l) {
        const HTile& t = tiles_[leafTiles_[l]];
        own[l] = {bufs_[t.ownBuf]->get(), t.ownOff, t.ownCnt, 0};
        ranges += kind_ == TREE_GROUPED ? leaves_[l].nbLeaves.size() : 1;
    }
    std::vector<RangeCopy> rc;
    rc.reserve(ranges);
    int64_t dst = 0;
    for (int l = 0; l < L; ++l) {
        rc.push_back({own[l].src, own[l].srcOff, own[l].cnt, dst});
        dst += own[l].cnt;
        if (kind_ == TREE_EXTENDED) {
            const HTile& t = tiles_[leafTiles_[l]];
            if (t.nbCnt > 0) {
                rc.push_back({bufs_[t.nbBuf]->get(), t.nbOff, t.nbCnt, dst});
                dst += t.nbCnt;
            }
        }
        if (kind_ == TREE_GROUPED)
            for (int j : leaves_[l].nbLeaves) {
                rc.push_back({own[j].src, own[j].srcOff, own[j].cnt, dst});
                dst += own[j].cnt;
            }
    }
    if (rc.empty()) return;
    DevBuf<RangeCopy> d;
    d.reserve(rc.size());
    check(hipMemcpyAsync(d.get(), rc.data(), sizeof(RangeCopy) * rc.size(), hipMemcpyHostToDevice, st_), "ranges");
    hipLaunchKernelGGL(copy_ranges_kernel, dim3((unsigned)rc.size()), dim3(256), 0, st_, d.get(), dIndex);
    check(hipGetLastError(), "copy_ranges");
    check(hipStreamSynchronize(st_), "writeIndex");
}

template class TileTreeBuilder<float>;
template class TileTreeBuilder<double>;

}  // namespace rdhip
