// rd_vjp.hip -- reverse mode of the evolve iterations with frozen memberships.
//
// With the labels l_t(n), counts c_t(k) and moved flags m_t(k,d) of a forward run held
// fixed, one evolve iteration (the shift of rd/cpu/brute_force/evolve.hpp:129-179 after
// the member sums of evolve.hpp:32-77) is the affine map
//     S_{t+1}(k,d) = m_t(k,d) * sum_{n: l_t(n)=k} P(n,d) / c_t(k) + (1 - m_t(k,d)) * S_t(k,d).
// Its vector-Jacobian product for an upstream G = dL/dS_{t+1}:
//     dL/dP(n,d)  += A_t(l_t(n), d)        with A_t(k,d) = m_t(k,d) * G(k,d) / c_t(k)
//     dL/dS_t(k,d) = (1 - m_t(k,d)) * G(k,d).
// The coefficient pass walks the logged steps newest first, one lane per chosen slot; the
// member pass gives every sample its own sum of A over the steps (a gather through its
// labels), so no two lanes write the same gradient: no atomics, and the order of the
// additions is fixed (steps newest first), which makes the result reproducible.
#include "rd_vjp.hpp"

#include <algorithm>

namespace rdhip {

namespace {

template <typename T, int DIM>
__global__ __launch_bounds__(256) void vjp_coef_kernel(VjpArgs<T> a) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= a.K) return;
    T g[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) g[d] = a.G[d * a.gStride + k];
    for (int u = a.steps - 1; u >= 0; --u) {
        const int c = a.counts[u * a.cStride + k];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            T coef = T(0);
            if (c > 0 && a.moved[((int64_t)u * DIM + d) * a.cStride + k]) {
                coef = g[d] / (T)c;
                g[d] = T(0);
            }
            a.A[((int64_t)u * DIM + d) * a.aStride + k] = coef;
        }
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) a.G[d * a.gStride + k] = g[d];
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void vjp_members_kernel(VjpArgs<T> a) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * 256) {
        T acc[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) acc[d] = a.accumulate ? a.gradP[d * a.gpStride + i] : T(0);
        for (int u = a.steps - 1; u >= 0; --u) {
            const int l = a.labels[u * a.lStride + i];
            if (l < 0) continue;
#pragma unroll
            for (int d = 0; d < DIM; ++d) acc[d] = acc[d] + a.A[((int64_t)u * DIM + d) * a.aStride + l];
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d) a.gradP[d * a.gpStride + i] = acc[d];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void vjp_seed_kernel(const T* __restrict__ G, int64_t gStride,
                                                       const int* __restrict__ idx, int K, int dim,
                                                       T* __restrict__ gradP, int64_t gpStride) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= K) return;
    const int64_t p = idx[k];
    for (int d = 0; d < dim; ++d) gradP[d * gpStride + p] = gradP[d * gpStride + p] + G[d * gStride + k];
}

}  // namespace

template <typename T>
hipError_t launch_vjp_run(const VjpArgs<T>& a, int dim, hipStream_t st) {
    if (a.steps <= 0) return hipSuccess;
    const unsigned cb = (unsigned)std::max(1, (a.K + 255) / 256);
    const unsigned pb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((a.n + 255) / 256, 16384));
    bool ok = false;
    RD_DISPATCH_DIM(dim, {
        if (a.K > 0) hipLaunchKernelGGL((vjp_coef_kernel<T, DIM>), dim3(cb), dim3(256), 0, st, a);
        if (a.n > 0) hipLaunchKernelGGL((vjp_members_kernel<T, DIM>), dim3(pb), dim3(256), 0, st, a);
        ok = true;
    });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_vjp_seed(const T* G, int64_t gStride, const int* idx, int K, int dim, T* gradP, int64_t gpStride,
                           hipStream_t st) {
    if (K <= 0) return hipSuccess;
    hipLaunchKernelGGL((vjp_seed_kernel<T>), dim3((unsigned)((K + 255) / 256)), dim3(256), 0, st, G, gStride, idx, K,
                       dim, gradP, gpStride);
    return hipGetLastError();
}

template hipError_t launch_vjp_run<float>(const VjpArgs<float>&, int, hipStream_t);
template hipError_t launch_vjp_run<double>(const VjpArgs<double>&, int, hipStream_t);
template hipError_t launch_vjp_seed<float>(const float*, int64_t, const int*, int, int, float*, int64_t, hipStream_t);
template hipError_t launch_vjp_seed<double>(const double*, int64_t, const int*, int, int, double*, int64_t,
                                            hipStream_t);

}  // namespace rdhip
