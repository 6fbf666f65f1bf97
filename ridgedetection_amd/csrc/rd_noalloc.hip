// rd_noalloc.hip -- rdhip_choose_noalloc: the greedy choose of rdhip_choose on caller-owned
// device buffers, for callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// The choose kernels (rd_choose.hip) need, beyond the points (input) and the chosen points,
// indices and count (outputs), a few device words (the segment description, its radius, the
// running count, an overflow flag), the work items of the chunked path and one "covered" byte
// per sample.  Here they live in `scratch`, a fixed device array of this library, instead of
// buffers a context allocates.  Every call on a device shares that device's copy of the array:
// calls that overlap on one device (different streams, or threads) race and must be ordered by
// the caller.  A call rewrites all of the array it reads before reading it, so nothing is
// carried over from an earlier call.  The entry point only launches kernels (no allocation,
// copy or synchronisation), so it may also be captured into a graph.
#include "../../include/rdhip.h"
#include "rd_choose.hpp"

#include <mutex>
#include <string>

namespace rdhip {
namespace {

constexpr int64_t kHeaderBytes = 256;                  // SegDesc, count, flags, radius
constexpr int64_t kItemsBytes = 64 * sizeof(WorkItem); // one chunk's filter items (<= 64 x 1024 samples)
constexpr int64_t kCoveredOff = 4096;                  // then one byte per sample
constexpr int64_t kScratchBytes = kCoveredOff + RDHIP_NOALLOC_MAX_POINTS;

static_assert(sizeof(SegDesc) + 16 + sizeof(double) <= kHeaderBytes, "header layout");
static_assert(kHeaderBytes + kItemsBytes <= kCoveredOff, "items layout");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct Header {
    SegDesc seg;
    int segK;
    int overflow;
    int segLarge;
    int pad;
    double rSqr;   // read through segRSqr as the value type (float in the first 4 bytes)
};

template <typename T>
__global__ void noalloc_setup_kernel(unsigned char* base, int n, T rSqr, int large) {
    if (threadIdx.x != 0) return;
    Header* h = reinterpret_cast<Header*>(base);
    SegDesc s{};
    s.pOff = 0;
    s.nP = n;
    s.nOwn = n;
    s.sOff = 0;
    s.sCap = n;   // the kernels write slots [0, sCap) only
    s.group = 0;
    h->seg = s;
    h->segK = 0;
    h->overflow = 0;
    h->segLarge = large;
    *reinterpret_cast<T*>(&h->rSqr) = rSqr;
}

// the filter items of chunk [c0, c1): 1024 samples each
__global__ void noalloc_items_kernel(WorkItem* items, int c0, int c1) {
    const int i = (int)threadIdx.x;
    const int st = c0 + i * 1024;
    if (st < c1) items[i] = WorkItem{0, st, min(c1, st + 1024), 0};
}

__global__ void noalloc_finish_kernel(const unsigned char* base, int32_t* count) {
    if (threadIdx.x == 0) *count = reinterpret_cast<const Header*>(base)->segK;
}

// hipGetSymbolAddress once per device, with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

hipError_t scratch_for_current_device(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

template <typename T>
hipError_t choose_noalloc(int dim, const T* points, int64_t n, int64_t stride, T r, T* chosen, int64_t chosenStride,
                          int32_t* indices, int32_t* count, hipStream_t st) {
    unsigned char* base = nullptr;
    hipError_t e = scratch_for_current_device(&base);
    if (e != hipSuccess) return e;
    Header* h = reinterpret_cast<Header*>(base);
    WorkItem* items = reinterpret_cast<WorkItem*>(base + kHeaderBytes);
    constexpr int kLarge = 32768;   // above this, the chunked path (as Engine::choose)
    const int nn = (int)n;
    const int large = nn > kLarge ? 1 : 0;
    hipLaunchKernelGGL((noalloc_setup_kernel<T>), dim3(1), dim3(64), 0, st, base, nn, r * r, large);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    ChooseArgs<T> a{points, stride, nullptr, chosen, chosenStride, &h->seg, 1, &h->segK,
                    reinterpret_cast<const T*>(&h->rSqr), base + kCoveredOff, indices, &h->overflow, &h->segLarge};
    if ((e = launch_choose_block<T>(a, dim, -1, 0, st)) != hipSuccess) return e;
    if (large) {
        if ((e = launch_choose_block<T>(a, dim, 0, 0, st)) != hipSuccess) return e;
        for (int c = 1; choose_chunk_start(c) < nn; ++c) {
            const int c0 = choose_chunk_start(c), c1 = std::min(choose_chunk_start(c + 1), nn);
            const int numItems = (c1 - c0 + 1023) / 1024;
            hipLaunchKernelGGL(noalloc_items_kernel, dim3(1), dim3(64), 0, st, items, c0, c1);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            if ((e = launch_choose_filter<T>(a, dim, items, numItems, st)) != hipSuccess) return e;
            if ((e = launch_choose_block<T>(a, dim, c, 1, st)) != hipSuccess) return e;
        }
    }
    hipLaunchKernelGGL(noalloc_finish_kernel, dim3(1), dim3(64), 0, st, base, count);
    return hipGetLastError();
}

int fail(int code, const char* msg) {
    set_last_error(msg);
    return code;
}

}  // namespace
}  // namespace rdhip

extern "C" int rdhip_choose_noalloc(int dtype, int dim, const void* points, int64_t n, int64_t stride, double r,
                                    void* chosen, int64_t chosen_stride, int32_t* indices, int32_t* count,
                                    void* hip_stream) {
    using namespace rdhip;
    if (dtype != RDHIP_F32 && dtype != RDHIP_F64) return fail(RDHIP_ERR_INVALID, "dtype must be RDHIP_F32 or RDHIP_F64");
    if (dim < 1 || dim > kMaxDim) return fail(RDHIP_ERR_INVALID, "dimension must be in [1, 12]");
    if (n < 1 || n > RDHIP_NOALLOC_MAX_POINTS)
        return fail(RDHIP_ERR_INVALID, "rdhip_choose_noalloc: n must be in [1, RDHIP_NOALLOC_MAX_POINTS = 2^24]");
    if (!points || !chosen || !indices || !count) return fail(RDHIP_ERR_INVALID, "null device buffer");
    if (stride != 0 && stride < n) return fail(RDHIP_ERR_INVALID, "point stride smaller than the point count");
    if (chosen_stride != 0 && chosen_stride < n) return fail(RDHIP_ERR_INVALID, "chosen stride smaller than the point count");
    if (!(r >= 0)) return fail(RDHIP_ERR_INVALID, "radius must be non-negative");
    const int64_t ps = stride ? stride : n, cs = chosen_stride ? chosen_stride : n;
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    const hipError_t e =
        dtype == RDHIP_F32
            ? choose_noalloc<float>(dim, static_cast<const float*>(points), n, ps, (float)r, static_cast<float*>(chosen),
                                    cs, indices, count, st)
            : choose_noalloc<double>(dim, static_cast<const double*>(points), n, ps, r, static_cast<double*>(chosen), cs,
                                     indices, count, st);
    if (e != hipSuccess) {
        set_last_error((std::string("rdhip_choose_noalloc: ") + hipGetErrorString(e)).c_str());
        return RDHIP_ERR_HIP;
    }
    return RDHIP_OK;
}
