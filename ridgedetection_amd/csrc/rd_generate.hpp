// rd_generate.hpp -- seeded synthetic samples generated on the device (rd_generate.hip).
//
// The device counterpart of rd/gpu/device/samples_generator.cuh:27-384 (circle, sphere,
// spiral2D, spiral3D, segmentND), plus the helix and polyline shapes of the benchmark
// clouds.  Random numbers come from a counter-based generator (Philox4x32-10; Salmon,
// Moraes, Dror and Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11) keyed by the
// seed, with the counter (sample index, stream, block): every sample is independent of the
// launch shape, and a cloud of a batch (stream = cloud index) is the same cloud whatever
// the rest of the batch.  Normals are Box-Muller pairs.
#pragma once

#include "rd_common.hpp"

namespace rdhip {

enum GenKind : int {
    GEN_CIRCLE = 0,    // params {x0, y0, r}: (x0 + r cos p, y0 + r sin p), p = 2 pi U
    GEN_SPHERE = 1,    // params {x0, y0, z0, r}: p = pi U1, q = 2 pi U2,
                       //   (x0 + r cos q sin p, y0 + r sin q sin p, z0 + r cos p)
    GEN_SPIRAL2D = 2,  // params {a, b}: p = a U, (b p cos p, b p sin p)
    GEN_SPIRAL3D = 3,  // params {a, b}: p = a U, (b p cos p, b p sin p, b p)
    GEN_SEGMENT = 4,   // params {length}: (length U, 0, ..., 0)
    GEN_HELIX = 5,     // params {turns, pitch}: t = 2 pi turns U, the first dim of (cos t, sin t, pitch t)
    GEN_POLYLINE = 6,  // params {m, v_0 .. v_m (dim values each)}: segment s = floor(m U1), v_s + U2 (v_{s+1} - v_s)
    GEN_UNIFORM = 7,   // params {}: every coordinate U (the generator's uniform stream, no noise)
    GEN_KINDS = 8
};

// Dimension a kind produces (0: any).
int gen_kind_dim(int kind);
// Number of parameters a kind takes for dimension dim (-1: unknown kind / bad dim).
int gen_kind_params(int kind, int dim, const double* params, int64_t numParams);

// One cloud to generate: n samples of `kind` (dimension dim) with per-coordinate noise
// sigma, written to out[d * ld + i] (SoA) or out[i * ld + d] (rowMajor); coordinates
// dim .. padDim - 1 are written as 0.  params points to device memory.
template <typename T>
struct GenCloud {
    int kind;
    int dim;
    int padDim;
    int rowMajor;
    int64_t n;
    int64_t ld;
    T* out;
    const double* params;
    double sigma;
    uint64_t seed;
    uint32_t stream;
};

// Generates every cloud of the list (device-resident GenCloud array, one launch).
template <typename T>
hipError_t launch_generate(const GenCloud<T>* clouds, int numClouds, int64_t maxN, hipStream_t st);

}  // namespace rdhip
