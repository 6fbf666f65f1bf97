// rd_api.hip -- the C ABI (include/rdhip.h) over the device engine.
#include "../../include/rdhip.h"
#include "rd_engine.hpp"
#include "rd_generate.hpp"
#include "rd_quality.hpp"
#include "rd_tiled.hpp"
#include "rd_tree.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

using namespace rdhip;

struct rdhip_ctx {
    int dtype = RDHIP_F32;
    int dim = 0;
    hipStream_t stream = nullptr;
    bool ownStream = false;
    std::unique_ptr<Engine<float>> f32;
    std::unique_ptr<Engine<double>> f64;
    bool hasPoints = false;
    bool hasChosen = false;
    Timers timers;
    std::unique_ptr<TileTreeBuilder<float>> tree32;
    std::unique_ptr<TileTreeBuilder<double>> tree64;
    int treeKind = -1;
    std::vector<int64_t> batchSizes;   // clouds of the resident batch (rdhip_batch_generate)
    bool broken = false;               // a HIP error left device state undefined

    std::unique_ptr<TileTreeBuilder<float>>& tree(float) { return tree32; }
    std::unique_ptr<TileTreeBuilder<double>>& tree(double) { return tree64; }
};

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        return fail(RDHIP_ERR_HIP, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(RDHIP_ERR_INVALID, e.what());
    } catch (const std::logic_error& e) {
        return fail(RDHIP_ERR_STATE, e.what());
    } catch (const std::bad_alloc&) {
        return fail(RDHIP_ERR_RUNTIME, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(RDHIP_ERR_RUNTIME, e.what());
    }
}

// Runs `body(engine)` on the context's engine of the right value type.  After a HIP error
// (failed allocation, launch or copy) the context's device buffers may be half set up, so
// every later call on it fails until it is destroyed.
template <typename F>
int withEngine(rdhip_ctx* ctx, F&& body) {
    if (!ctx) return fail(RDHIP_ERR_INVALID, "null context");
    if (ctx->broken)
        return fail(RDHIP_ERR_STATE, "an earlier HIP error left this context unusable; destroy it and create another");
    const int rc = guarded([&]() -> int {
        if (ctx->dtype == RDHIP_F32) return body(*ctx->f32, (float)0);
        return body(*ctx->f64, (double)0);
    });
    if (rc == RDHIP_ERR_HIP) ctx->broken = true;
    return rc;
}

}  // namespace

namespace rdhip {
void set_last_error(const char* msg) { g_error = msg; }
}  // namespace rdhip

extern "C" {

int rdhip_version(void) { return 100; }

const char* rdhip_last_error(void) { return g_error.c_str(); }

int rdhip_create(rdhip_ctx** out, int dtype, int dim, void* hip_stream) {
    if (!out) return fail(RDHIP_ERR_INVALID, "null output pointer");
    *out = nullptr;
    if (dtype != RDHIP_F32 && dtype != RDHIP_F64) return fail(RDHIP_ERR_INVALID, "dtype must be RDHIP_F32 or RDHIP_F64");
    if (dim < 1 || dim > kMaxDim) return fail(RDHIP_ERR_INVALID, "dimension must be in [1, 12]");
    return guarded([&]() -> int {
        std::unique_ptr<rdhip_ctx> c(new rdhip_ctx());
        c->dtype = dtype;
        c->dim = dim;
        if (hip_stream) {
            c->stream = static_cast<hipStream_t>(hip_stream);
        } else {
            check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
            c->ownStream = true;
        }
        if (dtype == RDHIP_F32)
            c->f32.reset(new Engine<float>(dim, c->stream));
        else
            c->f64.reset(new Engine<double>(dim, c->stream));
        *out = c.release();
        return RDHIP_OK;
    });
}

int rdhip_destroy(rdhip_ctx* ctx) {
    if (!ctx) return RDHIP_OK;
    return guarded([&]() -> int {
        hipStream_t s = ctx->stream;
        bool own = ctx->ownStream;
        (void)hipStreamSynchronize(s);
        ctx->f32.reset();
        ctx->f64.reset();
        if (own) (void)hipStreamDestroy(s);
        delete ctx;
        return RDHIP_OK;
    });
}

int rdhip_set_points(rdhip_ctx* ctx, const void* points, int64_t n, int layout, int64_t stride, int on_device) {
    if (ctx && n > 0 && !points) return fail(RDHIP_ERR_INVALID, "null points");
    if (n < 0 || n >= (int64_t(1) << 31)) return fail(RDHIP_ERR_INVALID, "point count out of range");
    if (layout < RDHIP_ROW_MAJOR || layout > RDHIP_SOA) return fail(RDHIP_ERR_INVALID, "unknown layout");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (layout == RDHIP_ROW_MAJOR && stride != 0 && stride < ctx->dim)
            return fail(RDHIP_ERR_INVALID, "row stride smaller than the dimension");
        if (layout != RDHIP_ROW_MAJOR && stride != 0 && stride < n)
            return fail(RDHIP_ERR_INVALID, "column stride smaller than the point count");
        e.setPoints(static_cast<const T*>(points), n, layout, stride, on_device != 0);
        ctx->hasPoints = true;
        ctx->hasChosen = false;
        ctx->batchSizes.clear();
        return RDHIP_OK;
    });
}

int rdhip_ridge_detection(rdhip_ctx* ctx, double r1, double r2, int64_t max_evolve_iters, int64_t* out_count) {
    return rdhip_ridge_detection_own(ctx, -1, r1, r2, max_evolve_iters, out_count);
}

int rdhip_ridge_detection_own(rdhip_ctx* ctx, int64_t n_own, double r1, double r2, int64_t max_evolve_iters,
                              int64_t* out_count) {
    if (!(r1 >= 0) || !(r2 >= 0)) return fail(RDHIP_ERR_INVALID, "radii must be non-negative");
    if (max_evolve_iters < 0) return fail(RDHIP_ERR_INVALID, "negative iteration limit");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasPoints || e.numPoints() < 1) return fail(RDHIP_ERR_STATE, "no input points");
        if (n_own == 0 || n_own > e.numPoints())
            return fail(RDHIP_ERR_INVALID, "own sample count must be in [1, number of points] (or negative for all)");
        e.timers = Timers();
        e.setSingleSegment((T)r1, (T)r2, n_own);
        RunOptions opt = e.options;
        opt.maxEvolveIters = max_evolve_iters;
        e.ridgeDetectionLocal(opt);
        ctx->hasChosen = true;
        if (out_count) *out_count = e.totalCount();
        return RDHIP_OK;
    });
}

}  // extern "C"

namespace {

// Runs the batch whose clouds (sizes) are the context's resident points, in order.
template <typename E, typename T>
int run_batch(rdhip_ctx* ctx, E& e, T, const std::vector<int64_t>& sizes, const double* r1, const double* r2,
              int64_t max_evolve_iters, int64_t* out_counts) {
    const int64_t num = (int64_t)sizes.size();
    std::vector<SegDesc> segs((size_t)num);
    std::vector<int64_t> caps((size_t)num);
    std::vector<T> vr1((size_t)num), vr2((size_t)num);
    int64_t off = 0;
    for (int64_t c = 0; c < num; ++c) {
        SegDesc s{};
        s.pOff = off;
        s.nP = (int32_t)sizes[c];
        s.nOwn = (int32_t)sizes[c];
        s.group = (int32_t)c;
        segs[c] = s;
        caps[c] = sizes[c];
        vr1[c] = (T)r1[c];
        vr2[c] = (T)r2[c];
        off += sizes[c];
    }
    e.setSegments(segs, caps, vr1, vr2, false);
    RunOptions opt = e.options;
    opt.maxEvolveIters = max_evolve_iters;
    e.ridgeDetectionLocal(opt);
    ctx->hasChosen = true;
    if (out_counts) {
        std::vector<int> K = e.countsHost();
        for (int64_t c = 0; c < num; ++c) out_counts[c] = K[c];
    }
    return RDHIP_OK;
}

int check_batch_radii(int64_t num, const double* r1, const double* r2) {
    for (int64_t c = 0; c < num; ++c)
        if (!(r1[c] >= 0) || !(r2[c] >= 0)) return fail(RDHIP_ERR_INVALID, "radii must be non-negative");
    return RDHIP_OK;
}

// Batched ridge detection; dims (optional) gives every cloud's own dimension <= the
// context's: its rows are padded with zero coordinates, which leave every distance (the
// padded terms come last and add +0), every sum, shift and decision unchanged.
int batch_rd(rdhip_ctx* ctx, int64_t num_clouds, const void* const* clouds, const int64_t* dims,
             const int64_t* sizes, const double* r1, const double* r2, int64_t max_evolve_iters,
             int64_t* out_counts) {
    if (num_clouds < 1 || num_clouds > (int64_t(1) << 24)) return fail(RDHIP_ERR_INVALID, "cloud count out of range");
    if (ctx && (!clouds || !sizes || !r1 || !r2)) return fail(RDHIP_ERR_INVALID, "null batch arrays");
    if (max_evolve_iters < 0) return fail(RDHIP_ERR_INVALID, "negative iteration limit");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        const int dim = ctx->dim;
        int64_t total = 0;
        for (int64_t c = 0; c < num_clouds; ++c) {
            if (sizes[c] < 1 || !clouds[c]) return fail(RDHIP_ERR_INVALID, "every cloud needs at least one point");
            if (dims && (dims[c] < 1 || dims[c] > dim))
                return fail(RDHIP_ERR_INVALID, "cloud dimensions must be in [1, context dimension]");
            total += sizes[c];
        }
        if (int rc = check_batch_radii(num_clouds, r1, r2)) return rc;
        if (total >= (int64_t(1) << 31)) return fail(RDHIP_ERR_INVALID, "batch exceeds 2^31 points");
        std::vector<T> all((size_t)total * dim, T(0));
        int64_t off = 0;
        for (int64_t c = 0; c < num_clouds; ++c) {
            const int dc = dims ? (int)dims[c] : dim;
            const T* src = static_cast<const T*>(clouds[c]);
            if (dc == dim) {
                std::memcpy(all.data() + off * dim, src, sizeof(T) * sizes[c] * dim);
            } else {
                for (int64_t i = 0; i < sizes[c]; ++i)
                    std::memcpy(all.data() + (off + i) * dim, src + i * dc, sizeof(T) * dc);
            }
            off += sizes[c];
        }
        e.timers = Timers();
        e.setPoints(all.data(), total, RDHIP_ROW_MAJOR, 0, false);
        ctx->hasPoints = true;
        ctx->batchSizes.clear();
        return run_batch(ctx, e, zero, std::vector<int64_t>(sizes, sizes + num_clouds), r1, r2, max_evolve_iters,
                         out_counts);
    });
}

// Clouds generated on the device into dst (SoA rows of stride ld, padded to padDim
// coordinates), cloud c at position offsets[c]; params are the host parameter lists.
template <typename T>
void generate_clouds(hipStream_t st, T* dst, int64_t ld, int padDim, const std::vector<GenCloud<T>>& proto,
                     const std::vector<double>& params, const std::vector<int64_t>& paramOff,
                     const std::vector<int64_t>& offsets) {
    const int num = (int)proto.size();
    DevBuf<double> dParams;
    dParams.reserve(std::max<size_t>(params.size(), 1));
    if (!params.empty())
        check(hipMemcpyAsync(dParams.get(), params.data(), sizeof(double) * params.size(), hipMemcpyHostToDevice, st),
              "generator parameters");
    std::vector<GenCloud<T>> gc(proto);
    int64_t maxN = 0;
    for (int c = 0; c < num; ++c) {
        gc[c].out = dst + offsets[c];
        gc[c].ld = ld;
        gc[c].padDim = padDim;
        gc[c].rowMajor = 0;
        gc[c].params = dParams.get() + paramOff[c];
        maxN = std::max(maxN, gc[c].n);
    }
    DevBuf<GenCloud<T>> dClouds;
    dClouds.reserve(gc.size());
    check(hipMemcpyAsync(dClouds.get(), gc.data(), sizeof(GenCloud<T>) * gc.size(), hipMemcpyHostToDevice, st),
          "generator clouds");
    for (int c0 = 0; c0 < num; c0 += 65535)
        check(launch_generate<T>(dClouds.get() + c0, std::min(num - c0, 65535), maxN, st), "generate");
    check(hipStreamSynchronize(st), "generate");
}

// Validates one generator request; returns its parameter count (or a negative error).
int gen_request(int kind, int64_t n, int dim, const double* params, int64_t num_params, double sigma) {
    if (n < 0) return fail(RDHIP_ERR_INVALID, "negative sample count");
    if (!(sigma >= 0) || !std::isfinite(sigma)) return fail(RDHIP_ERR_INVALID, "sigma must be finite and >= 0");
    const int kd = gen_kind_dim(kind);
    if (kd != 0 && kd != dim) return fail(RDHIP_ERR_INVALID, "this generator kind has a fixed dimension");
    const int np = gen_kind_params(kind, dim, params, num_params);
    if (np < 0) return fail(RDHIP_ERR_INVALID, "unknown generator kind, bad dimension or bad parameters");
    if (np > num_params || (np > 0 && !params)) return fail(RDHIP_ERR_INVALID, "too few generator parameters");
    for (int k = 0; k < np; ++k)
        if (!std::isfinite(params[k])) return fail(RDHIP_ERR_INVALID, "generator parameters must be finite");
    return np;
}

}  // namespace

extern "C" {

int rdhip_batch_ridge_detection(rdhip_ctx* ctx, int64_t num_clouds, const void* const* clouds, const int64_t* sizes,
                                const double* r1, const double* r2, int64_t max_evolve_iters, int64_t* out_counts) {
    return batch_rd(ctx, num_clouds, clouds, nullptr, sizes, r1, r2, max_evolve_iters, out_counts);
}

int rdhip_batch_ridge_detection_dims(rdhip_ctx* ctx, int64_t num_clouds, const void* const* clouds,
                                     const int64_t* dims, const int64_t* sizes, const double* r1, const double* r2,
                                     int64_t max_evolve_iters, int64_t* out_counts) {
    if (!dims) return fail(RDHIP_ERR_INVALID, "null dimension array");
    return batch_rd(ctx, num_clouds, clouds, dims, sizes, r1, r2, max_evolve_iters, out_counts);
}

int rdhip_generate(int dtype, int kind, int64_t n, int dim, const double* params, int64_t num_params, double sigma,
                   uint64_t seed, uint32_t stream_id, void* out, int layout, int64_t ld, int out_on_device) {
    if (dtype != RDHIP_F32 && dtype != RDHIP_F64) return fail(RDHIP_ERR_INVALID, "dtype must be RDHIP_F32 or RDHIP_F64");
    const int np = gen_request(kind, n, dim, params, num_params, sigma);
    if (np < 0) return np;
    if (n > 0 && !out) return fail(RDHIP_ERR_INVALID, "null output");
    if (layout < RDHIP_ROW_MAJOR || layout > RDHIP_SOA) return fail(RDHIP_ERR_INVALID, "unknown layout");
    const bool row = layout == RDHIP_ROW_MAJOR;
    if (ld == 0) ld = row ? dim : n;
    if (row ? ld < dim : ld < n) return fail(RDHIP_ERR_INVALID, "leading dimension too small");
    if (n == 0) return RDHIP_OK;
    auto body = [&](auto zero) -> int {
        using T = decltype(zero);
        hipStream_t st = nullptr;
        check(hipStreamCreate(&st), "hipStreamCreate");
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{st};
        const int64_t elems = row ? n * ld : (int64_t)dim * ld;
        DevBuf<T> staging;
        T* dst = static_cast<T*>(out);
        if (!out_on_device) dst = staging.reserve((size_t)elems);
        DevBuf<double> dParams;
        dParams.reserve((size_t)std::max(np, 1));
        if (np > 0)
            check(hipMemcpyAsync(dParams.get(), params, sizeof(double) * np, hipMemcpyHostToDevice, st),
                  "generator parameters");
        GenCloud<T> g{kind, dim, dim, row ? 1 : 0, n, ld, dst, dParams.get(), sigma, seed, stream_id};
        DevBuf<GenCloud<T>> dg;
        dg.reserve(1);
        check(hipMemcpyAsync(dg.get(), &g, sizeof(g), hipMemcpyHostToDevice, st), "generator cloud");
        check(launch_generate<T>(dg.get(), 1, n, st), "generate");
        if (!out_on_device)
            check(hipMemcpyAsync(out, dst, sizeof(T) * elems, hipMemcpyDeviceToHost, st), "download");
        check(hipStreamSynchronize(st), "generate");
        return RDHIP_OK;
    };
    return guarded([&]() -> int { return dtype == RDHIP_F32 ? body(0.0f) : body(0.0); });
}

int rdhip_generate_points(rdhip_ctx* ctx, int kind, int64_t n, const double* params, int64_t num_params,
                          double sigma, uint64_t seed) {
    if (!ctx) return fail(RDHIP_ERR_INVALID, "null context");
    if (n < 0 || n >= (int64_t(1) << 31)) return fail(RDHIP_ERR_INVALID, "point count out of range");
    const int np = gen_request(kind, n, ctx->dim, params, num_params, sigma);
    if (np < 0) return np;
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        T* P = e.reservePoints(n);
        GenCloud<T> g{kind, ctx->dim, ctx->dim, 0, n, 0, nullptr, nullptr, sigma, seed, 0};
        generate_clouds<T>(e.stream(), P, e.pointStride(), ctx->dim, {g},
                           std::vector<double>(params, params + np), {0}, {0});
        ctx->hasPoints = true;
        ctx->hasChosen = false;
        ctx->batchSizes.clear();
        return RDHIP_OK;
    });
}

int rdhip_batch_generate(rdhip_ctx* ctx, int64_t num_clouds, const int* kinds, const int64_t* dims,
                         const int64_t* sizes, const double* params, const int64_t* param_offsets,
                         const double* sigmas, const uint64_t* seeds) {
    if (!ctx) return fail(RDHIP_ERR_INVALID, "null context");
    if (num_clouds < 1 || num_clouds > (int64_t(1) << 24)) return fail(RDHIP_ERR_INVALID, "cloud count out of range");
    if (!kinds || !dims || !sizes || !param_offsets || !sigmas || !seeds)
        return fail(RDHIP_ERR_INVALID, "null batch arrays");
    int64_t total = 0;
    std::vector<double> allParams;
    std::vector<int64_t> paramOff((size_t)num_clouds), offsets((size_t)num_clouds);
    for (int64_t c = 0; c < num_clouds; ++c) {
        if (sizes[c] < 1) return fail(RDHIP_ERR_INVALID, "every cloud needs at least one point");
        if (dims[c] < 1 || dims[c] > ctx->dim)
            return fail(RDHIP_ERR_INVALID, "cloud dimensions must be in [1, context dimension]");
        if (param_offsets[c] < 0) return fail(RDHIP_ERR_INVALID, "negative parameter offset");
        const double* pc = params ? params + param_offsets[c] : nullptr;
        const int np = gen_request(kinds[c], sizes[c], (int)dims[c], pc, pc ? (int64_t(1) << 30) : 0, sigmas[c]);
        if (np < 0) return np;
        paramOff[c] = (int64_t)allParams.size();
        allParams.insert(allParams.end(), pc, pc + np);
        offsets[c] = total;
        total += sizes[c];
    }
    if (total >= (int64_t(1) << 31)) return fail(RDHIP_ERR_INVALID, "batch exceeds 2^31 points");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        std::vector<GenCloud<T>> proto((size_t)num_clouds);
        for (int64_t c = 0; c < num_clouds; ++c)
            proto[c] = GenCloud<T>{kinds[c], (int)dims[c], ctx->dim, 0, sizes[c], 0, nullptr, nullptr, sigmas[c],
                                   seeds[c], (uint32_t)c};
        T* P = e.reservePoints(total);
        generate_clouds<T>(e.stream(), P, e.pointStride(), ctx->dim, proto, allParams, paramOff, offsets);
        ctx->hasPoints = true;
        ctx->hasChosen = false;
        ctx->batchSizes.assign(sizes, sizes + num_clouds);
        return RDHIP_OK;
    });
}

int rdhip_batch_run(rdhip_ctx* ctx, const double* r1, const double* r2, int64_t max_evolve_iters,
                    int64_t* out_counts) {
    if (!ctx) return fail(RDHIP_ERR_INVALID, "null context");
    if (ctx->batchSizes.empty()) return fail(RDHIP_ERR_STATE, "no batch generated on the device");
    if (!r1 || !r2) return fail(RDHIP_ERR_INVALID, "null radii");
    if (max_evolve_iters < 0) return fail(RDHIP_ERR_INVALID, "negative iteration limit");
    if (int rc = check_batch_radii((int64_t)ctx->batchSizes.size(), r1, r2)) return rc;
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        e.timers = Timers();
        return run_batch(ctx, e, zero, ctx->batchSizes, r1, r2, max_evolve_iters, out_counts);
    });
}

int rdhip_tiled_ridge_detection(rdhip_ctx* ctx, int algorithm, double r1, double r2, int64_t max_tile_capacity,
                                const int64_t* tiles_per_dim, int refinement, int64_t max_evolve_iters,
                                int64_t* out_count) {
    if (!(r1 > 0) || !(r2 >= 0)) return fail(RDHIP_ERR_INVALID, "r1 must be positive and r2 non-negative");
    if (algorithm < RDHIP_TILED_LOCAL_TREE_MIXED_RD || algorithm > RDHIP_EXT_TILE_TREE_LOCAL_RD)
        return fail(RDHIP_ERR_INVALID, "unknown tiled algorithm");
    if (max_tile_capacity < 3) return fail(RDHIP_ERR_INVALID, "max_tile_capacity must be at least 3");
    if (max_evolve_iters < 0) return fail(RDHIP_ERR_INVALID, "negative iteration limit");
    if (ctx && !tiles_per_dim) return fail(RDHIP_ERR_INVALID, "null tiles_per_dim");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasPoints || e.numPoints() < 1) return fail(RDHIP_ERR_STATE, "no input points");
        std::vector<int64_t> tpd(tiles_per_dim, tiles_per_dim + ctx->dim);
        for (int64_t v : tpd)
            if (v < 1 || v > 4096) return fail(RDHIP_ERR_INVALID, "tiles per dimension must be in [1, 4096]");
        e.timers = Timers();
        TiledParams<T> p;
        p.algorithm = algorithm;
        p.r1 = (T)r1;
        p.r2 = (T)r2;
        p.maxTileCapacity = max_tile_capacity;
        p.tilesPerDim = tpd;
        p.refinement = refinement != 0;
        p.maxEvolveIters = max_evolve_iters;
        tiled_ridge_detection<T>(e, p);
        ctx->hasChosen = true;
        if (out_count) *out_count = e.totalCount();
        return RDHIP_OK;
    });
}

int rdhip_get_count(rdhip_ctx* ctx, int64_t* out_count) {
    if (!out_count) return fail(RDHIP_ERR_INVALID, "null output");
    return withEngine(ctx, [&](auto& e, auto) -> int {
        *out_count = ctx->hasChosen ? e.totalCount() : 0;
        return RDHIP_OK;
    });
}

int rdhip_get_points(rdhip_ctx* ctx, void* out, int layout, int64_t stride, int on_device) {
    if (layout < RDHIP_ROW_MAJOR || layout > RDHIP_SOA) return fail(RDHIP_ERR_INVALID, "unknown layout");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasChosen) return fail(RDHIP_ERR_STATE, "no result yet");
        const int64_t K = e.totalCount();
        if (K == 0) return RDHIP_OK;
        if (!out) return fail(RDHIP_ERR_INVALID, "null output");
        if (layout == RDHIP_ROW_MAJOR && stride != 0 && stride < ctx->dim)
            return fail(RDHIP_ERR_INVALID, "row stride smaller than the dimension");
        if (layout != RDHIP_ROW_MAJOR && stride != 0 && stride < K)
            return fail(RDHIP_ERR_INVALID, "column stride smaller than the point count");
        e.getChosen(static_cast<T*>(out), layout, on_device != 0, stride);
        return RDHIP_OK;
    });
}

int rdhip_get_timers(rdhip_ctx* ctx, rdhip_timers* out) {
    if (!out) return fail(RDHIP_ERR_INVALID, "null output");
    return withEngine(ctx, [&](auto& e, auto) -> int {
        out->whole_ms = e.timers.wholeMs;
        out->build_ms = e.timers.buildMs;
        out->choose_ms = e.timers.chooseMs;
        out->evolve_ms = e.timers.evolveMs;
        out->decimate_ms = e.timers.decimateMs;
        out->refine_ms = e.timers.refineMs;
        out->evolve_iterations = e.timers.evolveIterations;
        out->rounds = e.timers.rounds;
        out->scan_kernel_ms = e.timers.scanKernelMs;
        out->scan_launches = e.timers.scanLaunches;
        out->iteration_ms = e.timers.iterationMs;
        out->forward_ms = e.timers.forwardMs;
        out->backward_ms = e.timers.backwardMs;
        out->evolve_bytes = e.timers.evolveBytes;
        out->scan_ppt = e.timers.scanPpt;
        out->scan_threads = e.timers.scanThreads;
        out->tune_ms = e.timers.tuneMs;
        out->evolve_pairs = e.timers.evolvePairs;
        out->scan_persistent = e.timers.scanPersistent;
        out->vjp_checkpoint_every = e.timers.vjpCheckpointEvery;
        out->chain_path = e.timers.chainPath;
        out->screen_form = e.timers.screenForm;
        return RDHIP_OK;
    });
}

int rdhip_build_tree(rdhip_ctx* ctx, int kind, double r1, int64_t max_tile_capacity, const int64_t* tiles_per_dim,
                     int64_t* out_leaves) {
    if (kind < RDHIP_TREE_LOCAL || kind > RDHIP_TREE_EXTENDED) return fail(RDHIP_ERR_INVALID, "unknown tree kind");
    if (!(r1 > 0)) return fail(RDHIP_ERR_INVALID, "r1 must be positive");
    if (max_tile_capacity < 3) return fail(RDHIP_ERR_INVALID, "max_tile_capacity must be at least 3");
    if (ctx && !tiles_per_dim) return fail(RDHIP_ERR_INVALID, "null tiles_per_dim");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasPoints || e.numPoints() < 1) return fail(RDHIP_ERR_STATE, "no input points");
        std::vector<int64_t> tpd(tiles_per_dim, tiles_per_dim + ctx->dim);
        for (int64_t v : tpd)
            if (v < 1 || v > 4096) return fail(RDHIP_ERR_INVALID, "tiles per dimension must be in [1, 4096]");
        auto& tb = ctx->tree(zero);
        tb.reset(new TileTreeBuilder<T>(ctx->dim, ctx->stream));
        tb->build(e.devicePoints(), e.pointStride(), e.numPoints(), kind, (T)r1, max_tile_capacity, tpd);
        ctx->treeKind = kind;
        if (out_leaves) *out_leaves = (int64_t)tb->leaves().size();
        return RDHIP_OK;
    });
}

int rdhip_get_tree(rdhip_ctx* ctx, int64_t* leaf_info, int32_t* own_idx, int32_t* nb_idx, void* bounds) {
    return withEngine(ctx, [&](auto&, auto zero) -> int {
        using T = decltype(zero);
        auto& tb = ctx->tree(zero);
        if (!tb) return fail(RDHIP_ERR_STATE, "no tree (call rdhip_build_tree first)");
        const auto& leaves = tb->leaves();
        const int64_t total = tb->totalPositions();
        std::vector<int> idx((size_t)std::max<int64_t>(total, 1));
        if (total > 0) {
            DevBuf<int> d;
            d.reserve(total);
            tb->writeIndex(d.get());
            check(hipMemcpy(idx.data(), d.get(), sizeof(int) * total, hipMemcpyDeviceToHost), "tree index");
        }
        int64_t pos = 0, o = 0, b = 0;
        for (size_t l = 0; l < leaves.size(); ++l) {
            const auto& li = leaves[l];
            const int64_t nbShown = ctx->treeKind == RDHIP_TREE_GROUPED ? (int64_t)li.nbLeaves.size() : li.nbCnt;
            if (leaf_info) {
                leaf_info[4 * l] = li.ownCnt;
                leaf_info[4 * l + 1] = nbShown;
                leaf_info[4 * l + 2] = (int64_t)li.sphereCap;
                leaf_info[4 * l + 3] = 0;
            }
            if (own_idx)
                for (int64_t i = 0; i < li.ownCnt; ++i) own_idx[o + i] = idx[pos + i];
            o += li.ownCnt;
            if (nb_idx) {
                if (ctx->treeKind == RDHIP_TREE_GROUPED)
                    for (size_t j = 0; j < li.nbLeaves.size(); ++j) nb_idx[b + j] = li.nbLeaves[j];
                else
                    for (int64_t i = 0; i < li.nbCnt; ++i) nb_idx[b + i] = idx[pos + li.ownCnt + i];
            }
            b += nbShown;
            if (bounds) {
                T* bo = static_cast<T*>(bounds) + 2 * ctx->dim * l;
                for (int d = 0; d < ctx->dim; ++d) {
                    bo[2 * d] = li.mn[d];
                    bo[2 * d + 1] = li.mx[d];
                }
            }
            pos += li.ownCnt + li.nbCnt;
        }
        return RDHIP_OK;
    });
}

int rdhip_set_option(rdhip_ctx* ctx, int option, int64_t value) {
    if (ctx && option == RDHIP_OPT_INJECT_ALLOC_FAILURE) {   // no device state: also on a broken context
        if (value < 0) return fail(RDHIP_ERR_INVALID, "negative allocation count");
        t_allocFailCountdown = value;
        return RDHIP_OK;
    }
    return withEngine(ctx, [&](auto& e, auto) -> int {
        switch (option) {
            case RDHIP_OPT_PPT:
                if (value != 1 && value != 2 && value != 4) return fail(RDHIP_ERR_INVALID, "ppt must be 1, 2 or 4");
                e.options.autotune = false;
                if (e.options.ppt != (int)value) {
                    e.options.ppt = (int)value;
                    e.rebuildWorkItems();
                }
                return RDHIP_OK;
            case RDHIP_OPT_SCAN_THREADS:
                if (value != 128 && value != 256 && value != 512)
                    return fail(RDHIP_ERR_INVALID, "scan threads must be 128, 256 or 512");
                e.options.autotune = false;
                if (e.options.scanThreads != (int)value) {
                    e.options.scanThreads = (int)value;
                    e.rebuildWorkItems();
                }
                return RDHIP_OK;
            case RDHIP_OPT_AUTOTUNE:
                e.options.autotune = value != 0;
                return RDHIP_OK;
            case RDHIP_OPT_SEGMENTED:
                e.options.segmented = value != 0;
                return RDHIP_OK;
            case RDHIP_OPT_POSITION_COPY:
                e.options.positionCopy = value != 0;
                return RDHIP_OK;
            case RDHIP_OPT_ITERS_PER_SYNC:
                if (value < 1 || value > 1024) return fail(RDHIP_ERR_INVALID, "iterations per sync must be in [1, 1024]");
                e.options.itersPerSync = (int)value;
                return RDHIP_OK;
            case RDHIP_OPT_PROFILE:
                e.options.profile = value != 0;
                return RDHIP_OK;
            case RDHIP_OPT_PERSISTENT:
                e.options.persistent = value != 0;
                e.options.autotune = false;
                return RDHIP_OK;
            case RDHIP_OPT_GRAPH:
                e.options.graph = value != 0;
                return RDHIP_OK;
            case RDHIP_OPT_LOCAL:
                e.options.local = value != 0;
                return RDHIP_OK;
            case RDHIP_OPT_LOCAL_MAX_SAMPLES:
                if (value < 0 || value > kLocalMaxP) return fail(RDHIP_ERR_INVALID, "local sample limit must be in [0, 65536]");
                e.options.localMaxP = (int)value;
                return RDHIP_OK;
            case RDHIP_OPT_LOCAL_MAX_WORK:
                if (value < 0) return fail(RDHIP_ERR_INVALID, "negative local work limit");
                e.options.localMaxWork = value;
                return RDHIP_OK;
            case RDHIP_OPT_CHAIN:
                if (value < 0 || value > 2) return fail(RDHIP_ERR_INVALID, "chain option must be 0, 1 or 2");
                e.options.chain = (int)value;
                return RDHIP_OK;
            case RDHIP_OPT_SCREEN:
                if (value < 0 || value > 3) return fail(RDHIP_ERR_INVALID, "screen option must be 0, 1, 2 or 3");
                e.options.screen = (int)value;
                return RDHIP_OK;
            case RDHIP_OPT_SUM_MODE:
                if (value != 0 && value != 1) return fail(RDHIP_ERR_INVALID, "sum mode must be 0 (exact) or 1 (tolerance)");
                e.options.tolerance = value == 1;
                return RDHIP_OK;
            case RDHIP_OPT_VJP_MEMORY_MB:
                if (value < 1) return fail(RDHIP_ERR_INVALID, "the VJP memory budget must be at least 1 MiB");
                e.options.vjpBudgetBytes = value << 20;
                return RDHIP_OK;
            case RDHIP_OPT_EVOLVE_GUARD:
                if (value < 0) return fail(RDHIP_ERR_INVALID, "negative iteration guard");
                e.options.evolveGuard = value;
                return RDHIP_OK;
        }
        return fail(RDHIP_ERR_INVALID, "unknown option");
    });
}

int rdhip_evolve_fixed(rdhip_ctx* ctx, double r, int64_t iters) {
    if (!(r >= 0)) return fail(RDHIP_ERR_INVALID, "radius must be non-negative");
    if (iters < 0) return fail(RDHIP_ERR_INVALID, "negative iteration count");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasChosen) return fail(RDHIP_ERR_STATE, "no chosen points (call rdhip_choose first)");
        e.setRadii({(T)r}, {(T)r});
        e.evolveFixed(iters);
        return RDHIP_OK;
    });
}

int rdhip_choose(rdhip_ctx* ctx, double r, int64_t* out_count, int32_t* out_indices) {
    if (!(r >= 0)) return fail(RDHIP_ERR_INVALID, "radius must be non-negative");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasPoints || e.numPoints() < 1) return fail(RDHIP_ERR_STATE, "no input points");
        e.setSingleSegment((T)r, (T)r);
        e.choose();
        ctx->hasChosen = true;
        const int64_t K = e.totalCount();
        if (out_count) *out_count = K;
        if (out_indices) {
            std::vector<int> idx = e.chosenIndicesHost(0);
            for (size_t i = 0; i < idx.size(); ++i) out_indices[i] = idx[i];
        }
        return RDHIP_OK;
    });
}

int rdhip_set_chosen(rdhip_ctx* ctx, const void* chosen, int64_t count, int layout) {
    return rdhip_set_chosen_ex(ctx, chosen, count, layout, 0, 0);
}

int rdhip_set_chosen_ex(rdhip_ctx* ctx, const void* chosen, int64_t count, int layout, int64_t stride, int on_device) {
    if (count < 0) return fail(RDHIP_ERR_INVALID, "negative count");
    if (count > 0 && !chosen) return fail(RDHIP_ERR_INVALID, "null chosen points");
    if (layout < RDHIP_ROW_MAJOR || layout > RDHIP_SOA) return fail(RDHIP_ERR_INVALID, "unknown layout");
    if (stride < 0) return fail(RDHIP_ERR_INVALID, "negative stride");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasPoints) return fail(RDHIP_ERR_STATE, "no input points");
        if (count > e.numPoints()) return fail(RDHIP_ERR_INVALID, "more chosen points than input points");
        if (layout == RDHIP_ROW_MAJOR && stride != 0 && stride < ctx->dim)
            return fail(RDHIP_ERR_INVALID, "row stride smaller than the dimension");
        if (layout != RDHIP_ROW_MAJOR && stride != 0 && stride < count)
            return fail(RDHIP_ERR_INVALID, "column stride smaller than the point count");
        e.setSingleSegment(T(0), T(0));
        e.setChosen(0, static_cast<const T*>(chosen), count, layout, stride, on_device != 0);
        ctx->hasChosen = true;
        return RDHIP_OK;
    });
}

int rdhip_evolve(rdhip_ctx* ctx, double r, int64_t max_iters, int64_t* out_iters) {
    if (!(r >= 0)) return fail(RDHIP_ERR_INVALID, "radius must be non-negative");
    if (max_iters < 0) return fail(RDHIP_ERR_INVALID, "negative iteration limit");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasChosen) return fail(RDHIP_ERR_STATE, "no chosen points (call rdhip_choose first)");
        e.setRadii({(T)r}, {(T)r});
        long long it = e.evolve(max_iters);
        if (out_iters) *out_iters = it;
        return RDHIP_OK;
    });
}

int rdhip_evolve_iteration(rdhip_ctx* ctx, double r, int32_t* labels, int32_t* counts, void* sums, uint16_t* moved,
                           int* out_flag) {
    if (!(r >= 0)) return fail(RDHIP_ERR_INVALID, "radius must be non-negative");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasChosen) return fail(RDHIP_ERR_STATE, "no chosen points (call rdhip_choose first)");
        e.setRadii({(T)r}, {(T)r});
        std::vector<int> lab, cnt;
        std::vector<T> sm;
        std::vector<unsigned short> mv;
        int flag = e.evolveIteration(labels ? &lab : nullptr, counts ? &cnt : nullptr, sums ? &sm : nullptr,
                                     moved ? &mv : nullptr);
        const int64_t K = e.totalCount();
        const int64_t n = e.numPoints();
        if (labels)
            for (int64_t i = 0; i < n; ++i) labels[i] = lab[i];
        if (counts)
            for (int64_t k = 0; k < K; ++k) counts[k] = cnt[k];
        if (sums) {
            T* out = static_cast<T*>(sums);
            const int64_t stride = (int64_t)sm.size() / ctx->dim;
            for (int64_t k = 0; k < K; ++k)
                for (int d = 0; d < ctx->dim; ++d) out[k * ctx->dim + d] = sm[d * stride + k];
        }
        if (moved)
            for (int64_t k = 0; k < K; ++k) moved[k] = mv[k];
        if (out_flag) *out_flag = flag;
        return RDHIP_OK;
    });
}

int rdhip_set_distances(rdhip_ctx* ctx, const void* A, int64_t nA, const void* B, int64_t nB, void* out) {
    if (nA < 0 || nB < 0 || (nA > 0 && (!A || !out)) || (nB > 0 && !B)) return fail(RDHIP_ERR_INVALID, "bad point sets");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        std::vector<T> d = set_distances<T>(static_cast<const T*>(A), nA, static_cast<const T*>(B), nB, ctx->dim,
                                            e.stream());
        std::copy(d.begin(), d.end(), static_cast<T*>(out));
        return RDHIP_OK;
    });
}

int rdhip_hausdorff(rdhip_ctx* ctx, const void* A, int64_t nA, const void* B, int64_t nB, double* out) {
    if (nA < 0 || nB < 0 || (nA > 0 && !A) || (nB > 0 && !B) || !out) return fail(RDHIP_ERR_INVALID, "bad point sets");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        *out = (double)hausdorff_distance<T>(static_cast<const T*>(A), nA, static_cast<const T*>(B), nB, ctx->dim,
                                             e.stream());
        return RDHIP_OK;
    });
}

int rdhip_distance_stats(rdhip_ctx* ctx, const void* A, int64_t nA, const void* B, int64_t nB, double* out4) {
    if (nA < 1 || nB < 0 || !A || (nB > 0 && !B) || !out4) return fail(RDHIP_ERR_INVALID, "bad point sets");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        T s[4];
        distance_stats<T>(static_cast<const T*>(A), nA, static_cast<const T*>(B), nB, ctx->dim, e.stream(), s);
        for (int i = 0; i < 4; ++i) out4[i] = (double)s[i];
        return RDHIP_OK;
    });
}

int rdhip_order_chains(rdhip_ctx* ctx, double r2, int32_t* order, int64_t* chain_offsets, int64_t* out_chains) {
    if (!(r2 >= 0)) return fail(RDHIP_ERR_INVALID, "radius must be non-negative");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasChosen) return fail(RDHIP_ERR_STATE, "no chosen points");
        std::vector<int> ord;
        std::vector<int64_t> off;
        const int64_t nc = e.orderChains((T)r2, ord, off);
        if (order)
            for (size_t i = 0; i < ord.size(); ++i) order[i] = ord[i];
        if (chain_offsets)
            for (size_t i = 0; i < off.size(); ++i) chain_offsets[i] = off[i];
        if (out_chains) *out_chains = nc;
        return RDHIP_OK;
    });
}

int rdhip_evolve_fwd(rdhip_ctx* ctx, double r, int64_t iters, int64_t checkpoint_every) {
    if (!(r >= 0)) return fail(RDHIP_ERR_INVALID, "radius must be non-negative");
    if (iters < 1) return fail(RDHIP_ERR_INVALID, "iteration count must be positive");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasChosen) return fail(RDHIP_ERR_STATE, "no chosen points (call rdhip_choose first)");
        e.setRadii({(T)r}, {(T)r});
        e.evolveForward(iters, (int)std::max<int64_t>(std::min<int64_t>(checkpoint_every, iters), -1));
        return RDHIP_OK;
    });
}

int rdhip_evolve_bwd(rdhip_ctx* ctx, const void* grad_chosen, void* grad_points, void* grad_seeds, int layout,
                     int on_device, int seed_to_points) {
    if (!grad_chosen || !grad_points) return fail(RDHIP_ERR_INVALID, "null gradient buffer");
    if (layout != RDHIP_ROW_MAJOR && layout != RDHIP_SOA && layout != RDHIP_COL_MAJOR)
        return fail(RDHIP_ERR_INVALID, "unknown layout");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        const int dim = ctx->dim;
        const int64_t K = e.totalCount();
        const int64_t n = e.numPoints();
        const bool soa = layout != RDHIP_ROW_MAJOR;
        if (on_device && soa) {
            e.evolveBackward(static_cast<const T*>(grad_chosen), std::max<int64_t>(K, 1), static_cast<T*>(grad_points),
                             std::max<int64_t>(n, 1), static_cast<T*>(grad_seeds), std::max<int64_t>(K, 1),
                             seed_to_points != 0);
            return RDHIP_OK;
        }
        // staged: dense SoA device buffers, converted on the host
        const hipMemcpyKind up = on_device ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
        std::vector<T> g((size_t)K * dim), gHost((size_t)K * dim);
        if (K > 0) check(hipMemcpy(gHost.data(), grad_chosen, sizeof(T) * K * dim, up), "gradient");
        for (int64_t k = 0; k < K; ++k)
            for (int d = 0; d < dim; ++d) g[(size_t)d * K + k] = soa ? gHost[(size_t)d * K + k] : gHost[(size_t)k * dim + d];
        DevBuf<T> dG, dP, dS0;
        dG.reserve((size_t)std::max<int64_t>(K, 1) * dim);
        dP.reserve((size_t)std::max<int64_t>(n, 1) * dim);
        if (grad_seeds) dS0.reserve((size_t)std::max<int64_t>(K, 1) * dim);
        if (K > 0) check(hipMemcpy(dG.get(), g.data(), sizeof(T) * K * dim, hipMemcpyHostToDevice), "gradient");
        e.evolveBackward(dG.get(), std::max<int64_t>(K, 1), dP.get(), std::max<int64_t>(n, 1),
                         grad_seeds ? dS0.get() : nullptr, std::max<int64_t>(K, 1), seed_to_points != 0);
        auto down = [&](T* dst, const T* src, int64_t cnt) {
            std::vector<T> h((size_t)cnt * dim), o((size_t)cnt * dim);
            if (cnt > 0) check(hipMemcpy(h.data(), src, sizeof(T) * cnt * dim, hipMemcpyDeviceToHost), "gradient");
            for (int64_t i = 0; i < cnt; ++i)
                for (int d = 0; d < dim; ++d)
                    o[soa ? (size_t)d * cnt + i : (size_t)i * dim + d] = h[(size_t)d * cnt + i];
            if (cnt > 0)
                check(hipMemcpy(dst, o.data(), sizeof(T) * cnt * dim, on_device ? hipMemcpyHostToDevice : hipMemcpyHostToHost),
                      "gradient");
        };
        down(static_cast<T*>(grad_points), dP.get(), n);
        if (grad_seeds) down(static_cast<T*>(grad_seeds), dS0.get(), K);
        return RDHIP_OK;
    });
}

int rdhip_decimate(rdhip_ctx* ctx, double r2, int rule, uint8_t* keep, int64_t* out_count) {
    if (!(r2 >= 0)) return fail(RDHIP_ERR_INVALID, "radius must be non-negative");
    if (rule != RDHIP_DECIMATE && rule != RDHIP_DECIMATE_NO_COPY) return fail(RDHIP_ERR_INVALID, "unknown rule");
    return withEngine(ctx, [&](auto& e, auto zero) -> int {
        using T = decltype(zero);
        if (!ctx->hasChosen) return fail(RDHIP_ERR_STATE, "no chosen points");
        const int64_t K0 = e.totalCount();
        e.setRadii({(T)r2}, {(T)r2});
        std::vector<unsigned char> kp;
        e.decimate(rule == RDHIP_DECIMATE_NO_COPY, keep ? &kp : nullptr);
        if (keep)
            for (int64_t k = 0; k < K0; ++k) keep[k] = kp[k];
        if (out_count) *out_count = e.totalCount();
        return RDHIP_OK;
    });
}

}  // extern "C"
