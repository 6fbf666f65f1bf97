// rd_quality.hip -- quality of a ridge against an ideal pattern (rd/utils/assessment_quality.hpp:
// 91-220): for every point of A the distance to the nearest point of B, the Hausdorff distance
// and the distance statistics.  The O(|A| |B|) nearest-distance search runs on the device
// (the reference distance in its operation order, strict < from the largest value, so the
// minima are the reference's); the square roots, the sort and the in-order average run on
// the host, as the reference does them.
#include "rd_quality.hpp"

#include "rd_engine.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace rdhip {

namespace {

template <typename T, int DIM>
__global__ __launch_bounds__(256) void nearest_sqdist_kernel(const T* __restrict__ A, int64_t nA,
                                                             const T* __restrict__ B, int64_t nB,
                                                             T* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    T p[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = i < nA ? A[d * nA + i] : T(0);
    T best = FpTraits<T>::big;
    for (int64_t j = 0; j < nB; ++j) {   // uniform loop: B's coordinates come in as scalars
        T s[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) s[d] = B[d * nB + j];
        const T e = exact_sqdist<T, DIM>(p, s);
        if (e < best) best = e;
    }
    if (i < nA) out[i] = best;
}

template <typename T>
std::vector<T> soa(const T* rows, int64_t n, int dim) {
    std::vector<T> v((size_t)std::max<int64_t>(n, 1) * dim);
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < dim; ++d) v[(size_t)d * n + i] = rows[i * dim + d];
    return v;
}

}  // namespace

template <typename T>
std::vector<T> set_distances(const T* A, int64_t nA, const T* B, int64_t nB, int dim, hipStream_t st) {
    std::vector<T> out((size_t)nA);
    if (nA == 0) return out;
    if (dim < 1 || dim > kMaxDim) throw std::invalid_argument("dimension must be in [1, 12]");
    const std::vector<T> a = soa(A, nA, dim), b = soa(B, nB, dim);
    DevBuf<T> dA, dB, dO;
    dA.reserve(a.size());
    dB.reserve(b.size());
    dO.reserve((size_t)nA);
    check(hipMemcpyAsync(dA.get(), a.data(), sizeof(T) * a.size(), hipMemcpyHostToDevice, st), "quality");
    check(hipMemcpyAsync(dB.get(), b.data(), sizeof(T) * b.size(), hipMemcpyHostToDevice, st), "quality");
    const unsigned blocks = (unsigned)((nA + 255) / 256);
    RD_DISPATCH_DIM(dim, {
        hipLaunchKernelGGL((nearest_sqdist_kernel<T, DIM>), dim3(blocks), dim3(256), 0, st, dA.get(), nA, dB.get(),
                           nB, dO.get());
    });
    check(hipGetLastError(), "nearest_sqdist");
    check(hipMemcpyAsync(out.data(), dO.get(), sizeof(T) * nA, hipMemcpyDeviceToHost, st), "quality");
    check(hipStreamSynchronize(st), "quality");
    for (T& v : out) v = std::sqrt(v);
    return out;
}

template <typename T>
T hausdorff_distance(const T* A, int64_t nA, const T* B, int64_t nB, int dim, hipStream_t st) {
    if (nA == 0) return T(-1);
    const std::vector<T> ab = set_distances(A, nA, B, nB, dim, st);
    const std::vector<T> ba = set_distances(B, nB, A, nA, dim, st);
    T m = *std::max_element(ab.begin(), ab.end());
    if (!ba.empty()) m = std::max(m, *std::max_element(ba.begin(), ba.end()));
    return m;
}

template <typename T>
void distance_stats(const T* A, int64_t nA, const T* B, int64_t nB, int dim, hipStream_t st, T* out4) {
    if (nA == 0) throw std::invalid_argument("distance statistics of an empty set");
    std::vector<T> d = set_distances(A, nA, B, nB, dim, st);
    std::sort(d.begin(), d.end());
    out4[0] = d[d.size() / 2];
    T avg = 0;
    for (T v : d) avg += v;
    out4[1] = avg / (T)d.size();
    out4[2] = d.front();
    out4[3] = d.back();
}

template std::vector<float> set_distances<float>(const float*, int64_t, const float*, int64_t, int, hipStream_t);
template std::vector<double> set_distances<double>(const double*, int64_t, const double*, int64_t, int, hipStream_t);
template float hausdorff_distance<float>(const float*, int64_t, const float*, int64_t, int, hipStream_t);
template double hausdorff_distance<double>(const double*, int64_t, const double*, int64_t, int, hipStream_t);
template void distance_stats<float>(const float*, int64_t, const float*, int64_t, int, hipStream_t, float*);
template void distance_stats<double>(const double*, int64_t, const double*, int64_t, int, hipStream_t, double*);

}  // namespace rdhip
