// rd_order.hip -- chain ordering of a ridge (rd/cpu/brute_force/order.hpp:42-142 with the
// r2 condition of rd/cpu/brute_force/ridge_detection.inl:148-168).
//
// The reference walks a list: every unchained point (input order) seeds a chain, which
// repeatedly takes the first unchained point after the seed within r2 of its front
// (tested first) or of its back.  Every test it makes is a neighbour test within r2, so
// the device builds the r2-neighbour lists of all points (CSR, ascending indices, the
// reference's distance in its operation order) in two O(K^2) passes, and the walk itself
// -- order-dependent, O(K x neighbours) -- runs on the host over those lists: the first
// qualifying point is the smallest unchained index after the seed in the front's or the
// back's list.
#include "rd_engine.hpp"
#include "rd_sort.hpp"

#include <algorithm>
#include <deque>

namespace rdhip {

namespace {

template <typename T, int DIM>
__device__ __forceinline__ bool near_pair(const T* __restrict__ S, int64_t ss, const T (&p)[DIM], int j, T r2Sqr) {
    T s[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) s[d] = S[d * ss + j];
    return exact_sqdist<T, DIM>(p, s) <= r2Sqr;
}

// pass 0: neighbour counts; pass 1: neighbour lists (ascending j) at rowOff[i]
template <typename T, int DIM, int PASS>
__global__ __launch_bounds__(256) void order_neighbours_kernel(const T* __restrict__ S, int64_t ss, int K, T r2Sqr,
                                                                int* __restrict__ cnt, const int* __restrict__ rowOff,
                                                                int* __restrict__ nbr) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    T p[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = i < K ? S[d * ss + i] : T(0);
    int c = 0;
    int w = (PASS == 1 && i < K) ? rowOff[i] : 0;
    for (int j = 0; j < K; ++j) {   // uniform loop: the chosen coordinates come in as scalars
        const bool hit = i < K && j != i && near_pair<T, DIM>(S, ss, p, j, r2Sqr);
        if (PASS == 0) {
            c += hit ? 1 : 0;
        } else if (hit) {
            nbr[w++] = j;
        }
    }
    if (PASS == 0 && i < K) cnt[i] = c;
}

}  // namespace

template <typename T>
int64_t Engine<T>::orderChains(T r2, std::vector<int>& order, std::vector<int64_t>& offsets) {
    const int64_t K = totalCount();
    order.clear();
    offsets.assign(1, 0);
    if (K == 0) return 0;
    if (K > 0x7fffffff / 2) throw std::invalid_argument("too many points to order");
    DevBuf<T> dense;
    dense.reserve((size_t)K * dim_);
    getChosen(dense.get(), LAYOUT_SOA, true, K);
    DevBuf<int> rowOff, parts, nbr;
    DevBuf<int64_t> total;
    rowOff.reserve((size_t)K + 1);
    parts.reserve((size_t)scan_partials_count(K + 1));
    total.reserve(1);
    const T r2Sqr = r2 * r2;
    const unsigned blocks = (unsigned)((K + 255) / 256);
    bool ok = false;
    RD_DISPATCH_DIM(dim_, {
        hipLaunchKernelGGL((order_neighbours_kernel<T, DIM, 0>), dim3(blocks), dim3(256), 0, st_, dense.get(), K,
                           (int)K, r2Sqr, rowOff.get(), nullptr, nullptr);
        ok = true;
    });
    if (!ok) throw std::invalid_argument("unsupported dimension");
    check(hipMemsetAsync(rowOff.get() + K, 0, sizeof(int), st_), "order");
    check(exclusive_scan_inplace(rowOff.get(), K + 1, parts.get(), total.get(), st_), "order scan");
    int64_t nnz = 0;
    check(hipMemcpyAsync(&nnz, total.get(), sizeof(int64_t), hipMemcpyDeviceToHost, st_), "order");
    check(hipStreamSynchronize(st_), "order");
    nbr.reserve((size_t)std::max<int64_t>(nnz, 1));
    RD_DISPATCH_DIM(dim_, {
        hipLaunchKernelGGL((order_neighbours_kernel<T, DIM, 1>), dim3(blocks), dim3(256), 0, st_, dense.get(), K,
                           (int)K, r2Sqr, nullptr, rowOff.get(), nbr.get());
    });
    check(hipGetLastError(), "order_neighbours");
    std::vector<int> hOff((size_t)K + 1), hNbr((size_t)std::max<int64_t>(nnz, 1));
    check(hipMemcpyAsync(hOff.data(), rowOff.get(), sizeof(int) * (K + 1), hipMemcpyDeviceToHost, st_), "order");
    if (nnz > 0)
        check(hipMemcpyAsync(hNbr.data(), nbr.get(), sizeof(int) * nnz, hipMemcpyDeviceToHost, st_), "order");
    check(hipStreamSynchronize(st_), "order");

    // the walk
    std::vector<char> unchained((size_t)K, 1);
    // first unchained neighbour of v after `seed`, or K
    auto firstAfter = [&](int v, int seed) {
        for (int e = hOff[v]; e < hOff[v + 1]; ++e) {
            const int j = hNbr[e];
            if (j > seed && unchained[j]) return j;
        }
        return (int)K;
    };
    order.reserve((size_t)K);
    for (int seed = 0; seed < (int)K; ++seed) {
        if (!unchained[seed]) continue;
        std::deque<int> chain{seed};
        for (;;) {
            const int jf = firstAfter(chain.front(), seed);
            const int jb = firstAfter(chain.back(), seed);
            const int j = std::min(jf, jb);
            if (j >= (int)K) break;
            unchained[j] = 0;
            if (j == jf)
                chain.push_front(j);
            else
                chain.push_back(j);
        }
        order.insert(order.end(), chain.begin(), chain.end());
        offsets.push_back((int64_t)order.size());
    }
    return (int64_t)offsets.size() - 1;
}

template int64_t Engine<float>::orderChains(float, std::vector<int>&, std::vector<int64_t>&);
template int64_t Engine<double>::orderChains(double, std::vector<int>&, std::vector<int64_t>&);

}  // namespace rdhip
