// rd_generate.hip -- seeded synthetic samples on the device.
//
// One thread per sample of a cloud (grid: x over samples, y over clouds).  Sample i of a
// cloud draws Philox blocks (i, stream, j) for j = 0, 1, ...: block 0 gives the shape's
// uniforms (GEN_UNIFORM: coordinates 2e, 2e + 1 from block e), block 1 + e the Box-Muller
// pair of the noise of coordinates 2e and 2e + 1.  Coordinates are written SoA (a row per coordinate, coalesced over samples)
// or row-major.
#include "rd_generate.hpp"

namespace rdhip {

namespace {

// Philox4x32-10 (SC'11): multipliers and Weyl key increments from the paper.
constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return c;
}

// uniform in (0, 1]: 24 (float) or 53 (double, from two words) random bits, plus one ulp
template <typename T>
__device__ __forceinline__ T uniform01(uint32_t a, uint32_t b);
template <>
__device__ __forceinline__ float uniform01<float>(uint32_t a, uint32_t) {
    return (float)((a >> 8) + 1u) * 5.9604644775390625e-08f;   // 2^-24
}
template <>
__device__ __forceinline__ double uniform01<double>(uint32_t a, uint32_t b) {
    const uint64_t m = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);   // 53 bits
    return (double)(m + 1u) * 1.1102230246251565404e-16;             // 2^-53
}

template <typename T>
__device__ __forceinline__ void box_muller(T u1, T u2, T& z0, T& z1) {
    constexpr T kTwoPi = T(6.283185307179586476925286766559);
    const T r = sqrt(T(-2) * log(u1));
    const T t = kTwoPi * u2;
    z0 = r * cos(t);
    z1 = r * sin(t);
}

template <typename T>
__global__ __launch_bounds__(256) void generate_kernel(const GenCloud<T>* __restrict__ clouds, int numClouds) {
    const int c = blockIdx.y;
    if (c >= numClouds) return;
    const GenCloud<T> g = clouds[c];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    constexpr T kPi = T(3.1415926535897932384626433832795);
    const uint32_t k0 = (uint32_t)g.seed, k1 = (uint32_t)(g.seed >> 32);
    const uint32_t ilo = (uint32_t)i, ihi = (uint32_t)(i >> 32);
    auto draw = [&](uint32_t block) { return philox4x32_10(U4{ilo, ihi, g.stream, block}, k0, k1); };
    const U4 b0 = draw(0);
    const T u1 = uniform01<T>(b0.x, b0.y), u2 = uniform01<T>(b0.z, b0.w);
    T base[kMaxDim];
#pragma unroll
    for (int d = 0; d < kMaxDim; ++d) base[d] = T(0);
    const double* p = g.params;
    switch (g.kind) {
        case GEN_CIRCLE: {
            const T ph = T(2) * kPi * u1;
            base[0] = (T)p[0] + (T)p[2] * cos(ph);
            base[1] = (T)p[1] + (T)p[2] * sin(ph);
            break;
        }
        case GEN_SPHERE: {
            const T ph = kPi * u1, th = T(2) * kPi * u2;
            base[0] = (T)p[0] + (T)p[3] * cos(th) * sin(ph);
            base[1] = (T)p[1] + (T)p[3] * sin(th) * sin(ph);
            base[2] = (T)p[2] + (T)p[3] * cos(ph);
            break;
        }
        case GEN_SPIRAL2D:
        case GEN_SPIRAL3D: {
            const T ph = (T)p[0] * u1, bp = (T)p[1] * ph;
            base[0] = bp * cos(ph);
            base[1] = bp * sin(ph);
            base[2] = g.kind == GEN_SPIRAL3D ? bp : T(0);
            break;
        }
        case GEN_SEGMENT:
            base[0] = (T)p[0] * u1;
            break;
        case GEN_HELIX: {
            const T t = T(2) * kPi * (T)p[0] * u1;
            base[0] = cos(t);
            base[1] = sin(t);
            base[2] = (T)p[1] * t;
            break;
        }
        case GEN_POLYLINE: {
            const int m = (int)p[0];
            const int sg = min((int)((T)m * u1), m - 1);
            const double* v0 = p + 1 + (int64_t)sg * g.dim;
            const double* v1 = v0 + g.dim;
#pragma unroll
            for (int d = 0; d < kMaxDim; ++d)
                if (d < g.dim) base[d] = (T)v0[d] + u2 * ((T)v1[d] - (T)v0[d]);
            break;
        }
        case GEN_UNIFORM: {
            // coordinates 2e, 2e + 1: the two uniforms of block e
            base[0] = u1;
            base[1] = u2;
#pragma unroll
            for (int d = 2; d < kMaxDim; d += 2) {
                if (d >= g.dim) break;
                const U4 b = draw((uint32_t)(d >> 1));
                base[d] = uniform01<T>(b.x, b.y);
                base[d + 1] = uniform01<T>(b.z, b.w);
            }
            break;
        }
        default:
            break;
    }
    // noise: coordinates 2e and 2e + 1 take the Box-Muller pair of block 1 + e
    const T sigma = (T)g.sigma;
    if (g.kind != GEN_UNIFORM && sigma != T(0)) {
#pragma unroll
        for (int d = 0; d < kMaxDim; d += 2) {
            if (d >= g.dim) break;
            const U4 b = draw(1u + (uint32_t)(d >> 1));
            T z0, z1;
            box_muller<T>(uniform01<T>(b.x, b.y), uniform01<T>(b.z, b.w), z0, z1);
            base[d] = base[d] + sigma * z0;
            base[d + 1] = base[d + 1] + sigma * z1;
        }
    }
#pragma unroll
    for (int d = 0; d < kMaxDim; ++d) {
        if (d >= g.padDim) break;
        const T v = d < g.dim ? base[d] : T(0);
        if (g.rowMajor)
            g.out[i * g.ld + d] = v;
        else
            g.out[(int64_t)d * g.ld + i] = v;
    }
}

}  // namespace

int gen_kind_dim(int kind) {
    switch (kind) {
        case GEN_CIRCLE: return 2;
        case GEN_SPHERE: return 3;
        case GEN_SPIRAL2D: return 2;
        case GEN_SPIRAL3D: return 3;
        default: return 0;
    }
}

int gen_kind_params(int kind, int dim, const double* params, int64_t numParams) {
    if (dim < 1 || dim > kMaxDim) return -1;
    switch (kind) {
        case GEN_CIRCLE: return 3;
        case GEN_SPHERE: return 4;
        case GEN_SPIRAL2D:
        case GEN_SPIRAL3D: return 2;
        case GEN_SEGMENT: return 1;
        case GEN_HELIX: return dim <= 3 ? 2 : -1;
        case GEN_POLYLINE: {
            if (!params || numParams < 1) return -1;
            const double m = params[0];
            if (!(m >= 1 && m <= (1 << 20)) || m != (double)(int64_t)m) return -1;
            return 1 + (int)(m + 1) * dim;
        }
        case GEN_UNIFORM: return 0;
        default: return -1;
    }
}

template <typename T>
hipError_t launch_generate(const GenCloud<T>* clouds, int numClouds, int64_t maxN, hipStream_t st) {
    if (numClouds <= 0 || maxN <= 0) return hipSuccess;
    const int64_t bx = (maxN + 255) / 256;
    if (bx > 0x7fffffff || numClouds > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL((generate_kernel<T>), dim3((unsigned)bx, (unsigned)numClouds), dim3(256), 0, st, clouds,
                       numClouds);
    return hipGetLastError();
}

template hipError_t launch_generate<float>(const GenCloud<float>*, int, int64_t, hipStream_t);
template hipError_t launch_generate<double>(const GenCloud<double>*, int, int64_t, hipStream_t);

}  // namespace rdhip
