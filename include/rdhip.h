/* rdhip.h -- C ABI of the MI355X (gfx950) ridge-detection library (librdhip.so).
 *
 * The entry points replace the reference's GPU classes and device-level facades of
 * aosewski/RidgeDetection (paths relative to ridge_detection/):
 *   rdhip_create/_set_points/_ridge_detection/_get_count/_get_points
 *       <- rd::gpu::bruteForce::RidgeDetection<T,DIM,IN,OUT>
 *          (rd/gpu/device/brute_force/simulation.cuh:35-112, simulation.inl:40-384)
 *   rdhip_tiled_ridge_detection/_get_timers
 *       <- rd::gpu::tiled::RidgeDetection<DIM,IN,OUT,ALG,POLICY,TILE,T>
 *          (rd/gpu/device/tiled/simulation.cuh:33-241) and the CPU
 *          rd::tiled::tiledRidgeDetection<ALG,T> (rd/cpu/tiled/rd_tiled.hpp:1763-1798)
 *   rdhip_choose   <- DeviceChoose::choose     (rd/gpu/device/device_choose.cuh:55-82)
 *   rdhip_evolve   <- DeviceEvolve::evolve     (rd/gpu/device/device_evolve.cuh:29-97)
 *   rdhip_evolve_iteration
 *                  <- DeviceClosestSphereKernel + DeviceShiftSphereKernel, one pass
 *                     (rd/gpu/device/dispatch/dispatch_evolve.cuh:42-85)
 *   rdhip_decimate <- DeviceDecimate::decimate (rd/gpu/device/device_decimate.cuh:50-117) and
 *                     tiled::DeviceDecimate::globalDecimate (decimateNoCopy semantics)
 *
 * Conventions
 *   - Every function returns RDHIP_OK (0) or a negative error code; the message of the
 *     last failure on the calling thread is returned by rdhip_last_error().  Nothing
 *     aborts the process (the reference exits through checkCudaErrors).
 *   - A context owns all device memory it allocates and runs on the HIP stream given at
 *     creation (NULL = a stream it creates).  A context is not thread-safe; distinct
 *     contexts are independent (no process-global device state).
 *   - Point layouts follow rd::DataMemoryLayout (rd/utils/memory.h:22-27):
 *     ROW_MAJOR: point i at p[i*stride + d] (stride >= dim, 0 = dim);
 *     COL_MAJOR / SOA: point i at p[d*stride + i] (stride >= n, 0 = n).
 *   - Radii: r1 is the choose/evolve radius, r2 the decimate radius (reference names).
 *   - max_evolve_iters = 0 runs every evolve to convergence, as the reference does, up to an
 *     iteration guard (RDHIP_OPT_EVOLVE_GUARD, 100000 by default, 0 = none) past which the call
 *     fails with RDHIP_ERR_RUNTIME rather than looping on.
 *   - After RDHIP_ERR_HIP (a failed allocation, launch or copy) a context is unusable: every
 *     later call on it returns RDHIP_ERR_STATE until rdhip_destroy.
 */
#ifndef RDHIP_H
#define RDHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RDHIP_OK 0
#define RDHIP_ERR_INVALID (-1)   /* bad argument */
#define RDHIP_ERR_HIP (-2)       /* HIP runtime error (allocation, launch, ...) */
#define RDHIP_ERR_RUNTIME (-3)   /* algorithmic limit (capacity, tree depth, ...) */
#define RDHIP_ERR_STATE (-4)     /* call out of order (e.g. evolve before choose) */

enum rdhip_dtype { RDHIP_F32 = 0, RDHIP_F64 = 1 };
enum rdhip_layout { RDHIP_ROW_MAJOR = 0, RDHIP_COL_MAJOR = 1, RDHIP_SOA = 2 };

/* Tiled algorithms: values of rd::tiled::TiledRDAlgorithm (rd/cpu/tiled/rd_tiled.hpp:102-110)
 * plus local RD over extended tiles (the reference GPU's RD_LOCAL x RD_EXTENDED_TILE,
 * rd/gpu/agent/agent_tiled_ridge_detection.cuh:44-88). */
enum rdhip_tiled_algorithm {
    RDHIP_TILED_LOCAL_TREE_MIXED_RD = 0,
    RDHIP_TILED_GROUPED_TREE_MIXED_RD = 1,
    RDHIP_TILED_LOCAL_TREE_LOCAL_RD = 2,
    RDHIP_TILED_LOCAL_LIST = 3,
    RDHIP_TILED_GROUPED_TREE_LOCAL_RD = 4,
    RDHIP_EXT_TILE_TREE_MIXED_RD = 5,
    RDHIP_EXT_TILE_TREE_LOCAL_RD = 6
};

/* Tile trees: rd::tiled::TiledLocalTree, TiledGroupTree<FIND_NEIGHBOURS>,
 * TiledGroupTree<EXTENDED_TILE> (rd/cpu/tiled/{local,grouped}_tile_tree.hpp). */
enum rdhip_tree_kind { RDHIP_TREE_LOCAL = 0, RDHIP_TREE_GROUPED = 1, RDHIP_TREE_EXTENDED = 2 };

/* Decimation rules: rd::decimate and rd::decimateNoCopy (rd/cpu/brute_force/decimate.hpp). */
enum rdhip_decimate_rule { RDHIP_DECIMATE = 0, RDHIP_DECIMATE_NO_COPY = 1 };

typedef struct rdhip_ctx rdhip_ctx;

/* Counterpart of TiledRidgeDetectionTimers (rd/gpu/agent/agent_tiled_ridge_detection.cuh:96-102)
 * and of the (whole, build, rdTiles, refinement) tuple of rd_tiled.hpp:1763-1798. */
typedef struct rdhip_timers {
    double whole_ms;
    double build_ms;
    double choose_ms;
    double evolve_ms;
    double decimate_ms;
    double refine_ms;
    int64_t evolve_iterations;
    int64_t rounds;
    double scan_kernel_ms;   /* HIP-event time of the evolve scan kernel (RDHIP_OPT_PROFILE) */
    int64_t scan_launches;
    double iteration_ms;     /* HIP-event time of whole evolve iterations (RDHIP_OPT_PROFILE) */
    double forward_ms;       /* rdhip_evolve_fwd (wall time, stream synchronised) */
    double backward_ms;      /* rdhip_evolve_bwd without the staging of host buffers */
    double evolve_bytes;     /* algorithmic bytes of the evolve iterations: per segment and
                                iteration n*dim*sizeof(T) + K*(3*dim*sizeof(T) + 4) */
    int64_t scan_ppt;        /* evolve scan shape in use: samples per lane ... */
    int64_t scan_threads;    /* ... and threads per block (after autotuning) */
    double tune_ms;          /* time spent timing candidate scan shapes */
    double evolve_pairs;     /* (sample, chosen point) pairs the evolve iterations evaluated:
                                per segment and iteration, its samples x its chosen points */
    int64_t scan_persistent; /* 1: the scan's blocks pull work items from a device queue */
    int64_t vjp_checkpoint_every;  /* iterations per recorded run of the last rdhip_evolve_fwd */
    int64_t chain_path;      /* 1: the last evolve iteration enqueued ordered its members with the
                                single-cloud path (RDHIP_OPT_CHAIN), 0: with the general sort */
    int64_t screen_form;     /* screen of the last evolve scan enqueued (RDHIP_OPT_SCREEN):
                                0 direct, 1 expanded, 2 expanded on the matrix cores */
} rdhip_timers;

/* Tuning / instrumentation options (rdhip_set_option). */
enum rdhip_option {
    RDHIP_OPT_PPT = 0,             /* samples per lane of the evolve scan: 1, 2 or 4 (setting it
                                      turns the autotuner off) */
    RDHIP_OPT_ITERS_PER_SYNC = 1,  /* evolve iterations queued between convergence checks (8) */
    RDHIP_OPT_PROFILE = 2,         /* 1: time the scan kernel and iterations with HIP events */
    RDHIP_OPT_PERSISTENT = 3,      /* batches: 1 = persistent scan blocks fed by a device work
                                      queue, 0 = one block per work item (setting it turns the
                                      autotuner off; by default the autotuner picks) */
    RDHIP_OPT_GRAPH = 4,           /* 1 (default): replay captured batches of evolve iterations
                                      (hipGraph) until the device convergence flag is set */
    RDHIP_OPT_LOCAL = 5,           /* 1 (default): small segments (at most 1024 chosen points and
                                      the next two limits) evolve inside one workgroup each, every
                                      iteration in LDS -- all active segments when all fit the
                                      sample limit, else those of at most 2048 samples;
                                      0: batch-wide iterations only */
    RDHIP_OPT_LOCAL_MAX_SAMPLES = 6,  /* sample limit of that path (default 16384, at most 65536) */
    RDHIP_OPT_LOCAL_MAX_WORK = 7,     /* limit on samples x chosen points (default 2^22) */
    RDHIP_OPT_SCAN_THREADS = 8,    /* threads per evolve-scan block: 128, 256 or 512 (setting it
                                      turns the autotuner off) */
    RDHIP_OPT_AUTOTUNE = 9,        /* 1 (default): on first use for a (value type, dimension,
                                      log2 samples, log2 chosen points) class, a single cloud's
                                      evolve times every (samples per lane x threads per block)
                                      scan shape on the data and keeps the fastest; a batch of
                                      segments (clouds, tiles) does the same per (value type,
                                      dimension, log2 samples, log2 segments) class over nine
                                      (samples per lane x threads x persistent queue or not)
                                      shapes.  Cached per process */
    RDHIP_OPT_SEGMENTED = 10,      /* 1: a batch whose active segments all have at most 65536
                                      samples and 1024 chosen points iterates as a batch-wide scan
                                      plus one workgroup per segment for the member sort, ordered
                                      sums and shift; 0 (default, faster on the measured tiled and
                                      batched workloads): batch-wide sort and sums */
    RDHIP_OPT_POSITION_COPY = 11,  /* 1 (default): tiled runs copy the samples into segment-position
                                      order once per tiling, so the evolve and choose passes read them
                                      without the position -> point index; 0: read through the index */
    RDHIP_OPT_EVOLVE_GUARD = 12,   /* an evolve that still moves after this many iterations stops
                                      with RDHIP_ERR_RUNTIME instead of running on (default 100000;
                                      0 = no limit, the reference's behaviour, which never returns
                                      on an input that does not converge) */
    RDHIP_OPT_INJECT_ALLOC_FAILURE = 13, /* test hook: the value-th device allocation the calling
                                      thread makes from now on (in any context) fails with
                                      hipErrorOutOfMemory; 0 = off */
    RDHIP_OPT_VJP_MEMORY_MB = 14,  /* memory budget (MiB, default 1024) of rdhip_evolve_fwd with
                                      checkpoint_every < 0 */
    RDHIP_OPT_SUM_MODE = 15,       /* 0 (default): every centroid sum in the reference's serial
                                      sample order, results bit-identical to the reference CPU
                                      path; 1 (tolerance mode): batch-wide iterations sum each chosen
                                      point's members in parallel in a wider type (double for float
                                      data, double-double for double data) in a fixed order and take
                                      the mean of that sum -- no serial floor on long chains,
                                      deterministic, but means differ from the reference's by its
                                      own rounding of the serial sum */
    RDHIP_OPT_CHAIN = 16,          /* member ordering of the evolve iterations: 1 (default): when
                                      every segment (cloud, tile) has at most 2048 chosen points and
                                      the segments have fewer chosen points than a scan block has
                                      samples (on average), the scan counts the labels per block and
                                      one scatter writes the members in chain order; 2: that path
                                      whenever the 2048 limit allows; 0: the general stable sort.
                                      Same results either way */
    RDHIP_OPT_SCREEN = 17          /* how the evolve scan screens (sample, chosen point) pairs
                                      before its exact decision: 2 (default): the fastest form as
                                      timed by the autotuner (0 where it does not run); 1:
                                      expanded (|s-o|^2 - 2 (s-o).(p-o) about a
                                      per-segment centre o, DIM fused multiply-adds per pair, a
                                      rigorous bound and an exact pass for near ties); 3: the
                                      expanded form on the matrix cores (FP32; FP64 uses 1); 0:
                                      direct (|s-p|^2 with a relative bound).  Same results
                                      whichever form runs */
};

int rdhip_version(void);
const char* rdhip_last_error(void);

/* Context: value type (RDHIP_F32/F64), point dimension in [1, 12], HIP stream (hipStream_t
 * or NULL). */
int rdhip_create(rdhip_ctx** out, int dtype, int dim, void* hip_stream);
int rdhip_destroy(rdhip_ctx* ctx);

/* Input cloud.  on_device != 0: `points` is a device pointer (copied, not retained). */
int rdhip_set_points(rdhip_ctx* ctx, const void* points, int64_t n, int layout, int64_t stride,
                     int on_device);

/* Brute-force ridge detection: choose, then rounds of evolve + decimate until the number of
 * chosen points is stable (rd/cpu/brute_force/ridge_detection.inl:67-146). */
int rdhip_ridge_detection(rdhip_ctx* ctx, double r1, double r2, int64_t max_evolve_iters,
                          int64_t* out_count);

/* Brute-force ridge detection over a list of sample sets, concatenated in the context's
 * points: choose takes the first n_own points (the first set) only, evolve takes all of them,
 * own points first -- rd::RidgeDetection<T>::ridgeDetection(std::vector<std::pair<T const*,
 * size_t>> samples, ...) (rd/cpu/brute_force/ridge_detection.inl:171-286).  n_own < 0: all
 * points (= rdhip_ridge_detection). */
int rdhip_ridge_detection_own(rdhip_ctx* ctx, int64_t n_own, double r1, double r2, int64_t max_evolve_iters,
                              int64_t* out_count);

/* Batched brute-force ridge detection: `num_clouds` independent clouds of the context's
 * dimension (host ROW_MAJOR pointers, sizes[c] points, radii r1[c] / r2[c]) each run the
 * loop of rdhip_ridge_detection, all clouds of a phase in one launch.  out_counts[c]
 * receives each cloud's number of ridge points; rdhip_get_points then returns all clouds'
 * points concatenated in cloud order.  No reference counterpart: the reference runs one
 * cloud per RidgeDetection object. */
int rdhip_batch_ridge_detection(rdhip_ctx* ctx, int64_t num_clouds, const void* const* clouds,
                                const int64_t* sizes, const double* r1, const double* r2,
                                int64_t max_evolve_iters, int64_t* out_counts);

/* The same for clouds of different dimensions: cloud c has dims[c] <= the context's
 * dimension values per point (host ROW_MAJOR, n x dims[c]) and runs as that many
 * coordinates followed by zeros -- the zero coordinates add +0 to every distance and never
 * move, so each cloud's result equals its own-dimension run bit for bit.  rdhip_get_points
 * returns rows of the context's dimension; cloud c's points are the first dims[c] values
 * of its rows. */
int rdhip_batch_ridge_detection_dims(rdhip_ctx* ctx, int64_t num_clouds, const void* const* clouds,
                                     const int64_t* dims, const int64_t* sizes, const double* r1,
                                     const double* r2, int64_t max_evolve_iters, int64_t* out_counts);

/* ---- seeded synthetic samples on the device ------------------------------------------
 * Replaces rd::gpu::SamplesGenerator<T> (rd/gpu/device/samples_generator.cuh:291-384:
 * circle, sphere, spiral2D, spiral3D, segmentND, writing into a device buffer) with a
 * counter-based generator (Philox4x32-10) keyed by `seed`: sample i of stream `stream_id`
 * is the same whatever n, the launch shape or the other clouds.  Every coordinate gets
 * N(0, sigma^2) noise (the reference scales its sigma by 1/sqrt(dim) first; pass the scaled
 * value to reproduce its spread).  Kinds and their params: */
enum rdhip_gen_kind {
    RDHIP_GEN_CIRCLE = 0,    /* dim 2, {x0, y0, r}: p = 2 pi U, (x0 + r cos p, y0 + r sin p) */
    RDHIP_GEN_SPHERE = 1,    /* dim 3, {x0, y0, z0, r}: p = pi U1, q = 2 pi U2,
                                (x0 + r cos q sin p, y0 + r sin q sin p, z0 + r cos p) */
    RDHIP_GEN_SPIRAL2D = 2,  /* dim 2, {a, b}: p = a U, (b p cos p, b p sin p) */
    RDHIP_GEN_SPIRAL3D = 3,  /* dim 3, {a, b}: p = a U, (b p cos p, b p sin p, b p) */
    RDHIP_GEN_SEGMENT = 4,   /* any dim, {length}: (length U, 0, ..., 0) */
    RDHIP_GEN_HELIX = 5,     /* dim <= 3, {turns, pitch}: t = 2 pi turns U, first dim of (cos t, sin t, pitch t) */
    RDHIP_GEN_POLYLINE = 6,  /* any dim, {m, v_0, ..., v_m (dim values each)}: s = floor(m U1),
                                v_s + U2 (v_{s+1} - v_s) */
    RDHIP_GEN_UNIFORM = 7    /* any dim, {}: every coordinate uniform in (0, 1], no noise */
};

/* n samples into `out` (device memory when out_on_device, else host memory staged through
 * the device), ROW_MAJOR (row stride ld >= dim) or COL_MAJOR / SOA (column stride ld >= n;
 * 0 = packed).  Synchronous. */
int rdhip_generate(int dtype, int kind, int64_t n, int dim, const double* params, int64_t num_params,
                   double sigma, uint64_t seed, uint32_t stream_id, void* out, int layout, int64_t ld,
                   int out_on_device);

/* The same straight into the context's point buffer (its dimension, stream 0): the device
 * form of rdhip_set_points for generated clouds (the reference tests fill dP_ this way). */
int rdhip_generate_points(rdhip_ctx* ctx, int kind, int64_t n, const double* params, int64_t num_params,
                          double sigma, uint64_t seed);

/* No reference counterpart: a batch of clouds generated on the device as the context's
 * resident points (cloud c = stream c, dims[c] <= the context's dimension, zero-padded as in
 * rdhip_batch_ridge_detection_dims; its params start at params[param_offsets[c]]), then
 * run with rdhip_batch_run -- batched benchmarks without a host upload. */
int rdhip_batch_generate(rdhip_ctx* ctx, int64_t num_clouds, const int* kinds, const int64_t* dims,
                         const int64_t* sizes, const double* params, const int64_t* param_offsets,
                         const double* sigmas, const uint64_t* seeds);
int rdhip_batch_run(rdhip_ctx* ctx, const double* r1, const double* r2, int64_t max_evolve_iters,
                    int64_t* out_counts);

/* Tiled ridge detection (tree build, per-tile choose/evolve, local or global decimation,
 * optional end-phase refinement: decimate + evolve over all points). tiles_per_dim has dim
 * entries; max_tile_capacity: a tile with at least this many own points is split. */
int rdhip_tiled_ridge_detection(rdhip_ctx* ctx, int algorithm, double r1, double r2,
                                int64_t max_tile_capacity, const int64_t* tiles_per_dim,
                                int refinement, int64_t max_evolve_iters, int64_t* out_count);

/* Builds a tile tree (enum rdhip_tree_kind) over the current points and keeps it for
 * rdhip_get_tree: the device counterpart of TiledLocalTree / TiledGroupTree::buildTree
 * (rd/cpu/tiled/local_tile_tree.inl:48-76, grouped_tile_tree.inl:86-115, 528-548).
 * out_leaves receives the number of leaves. */
int rdhip_build_tree(rdhip_ctx* ctx, int kind, double r1, int64_t max_tile_capacity,
                     const int64_t* tiles_per_dim, int64_t* out_leaves);

/* Leaves of the last built tree in preorder (host outputs, optional):
 *   leaf_info[4*l..4*l+3] = {own samples, neighbours, chosen capacity, 0}; neighbours counts
 *     samples for EXTENDED trees and leaves for GROUPED trees;
 *   own_idx: every leaf's own input indices, concatenated (in the tile's sample order);
 *   nb_idx: EXTENDED: neighbour input indices; GROUPED: neighbour leaf positions;
 *   bounds: 2*dim values per leaf, [min_0, max_0, min_1, max_1, ...] in the value type. */
int rdhip_get_tree(rdhip_ctx* ctx, int64_t* leaf_info, int32_t* own_idx, int32_t* nb_idx, void* bounds);

/* Sets a tuning / instrumentation option (enum rdhip_option). */
int rdhip_set_option(rdhip_ctx* ctx, int option, int64_t value);

/* Results of the last run (or phase). */
int rdhip_get_count(rdhip_ctx* ctx, int64_t* out_count);
int rdhip_get_points(rdhip_ctx* ctx, void* out, int layout, int64_t stride, int on_device);
int rdhip_get_timers(rdhip_ctx* ctx, rdhip_timers* out);

/* ---- phase-level entry points on the whole cloud (parity tests, custom drivers) ---- */

/* Greedy choose with radius r; out_indices (host, optional, capacity n) receives the
 * chosen points' input indices in ascending order. */
int rdhip_choose(rdhip_ctx* ctx, double r, int64_t* out_count, int32_t* out_indices);

/* Replaces the chosen set (host pointer, `count` points in `layout`, packed).  At most as many
 * points as the cloud holds. */
int rdhip_set_chosen(rdhip_ctx* ctx, const void* chosen, int64_t count, int layout);

/* The same from host or device memory (on_device) with a stride (0 = packed): ROW_MAJOR stride
 * >= dim between points, COL_MAJOR / SOA stride >= count between coordinates.  The input form
 * of DeviceEvolve::evolve's d_chosenSamples and DeviceDecimate::decimate's d_chosenPoints
 * (rd/gpu/device/device_evolve.cuh:35-62, device_decimate.cuh:55-70). */
int rdhip_set_chosen_ex(rdhip_ctx* ctx, const void* chosen, int64_t count, int layout, int64_t stride,
                        int on_device);

/* Evolve the chosen set with radius r until convergence (or max_iters > 0 iterations). */
int rdhip_evolve(rdhip_ctx* ctx, double r, int64_t max_iters, int64_t* out_iters);

/* Exactly `iters` evolve iterations with radius r, ignoring convergence (a converged set
 * stays where it is).  The benchmark's unit of work and the forward pass of the evolve VJP. */
int rdhip_evolve_fixed(rdhip_ctx* ctx, double r, int64_t iters);

/* One evolve iteration.  Optional host outputs: labels[n] (chosen index of every point, or
 * -1 outside every r-ball), counts[K], sums[K*dim] (ROW_MAJOR, value type of the context),
 * moved[K] (bit d set when coordinate d moved).  out_flag receives the reference's
 * contFlag. */
int rdhip_evolve_iteration(rdhip_ctx* ctx, double r, int32_t* labels, int32_t* counts, void* sums,
                           uint16_t* moved, int* out_flag);

/* Quality against an ideal pattern (rd::RDAssessmentQuality, rd/utils/assessment_quality.hpp:
 * 91-220).  A and B are host ROW_MAJOR point sets of the context's value type and dimension.
 *   rdhip_set_distances: out[i] (value type) = distance from A[i] to the nearest point of B;
 *   rdhip_hausdorff: max of the two directed set distances, -1 for an empty A;
 *   rdhip_distance_stats: out4 = (median, average, min, max) of the distances from A to B. */
int rdhip_set_distances(rdhip_ctx* ctx, const void* A, int64_t nA, const void* B, int64_t nB, void* out);
int rdhip_hausdorff(rdhip_ctx* ctx, const void* A, int64_t nA, const void* B, int64_t nB, double* out);
int rdhip_distance_stats(rdhip_ctx* ctx, const void* A, int64_t nA, const void* B, int64_t nB, double* out4);

/* Chain ordering of the current chosen points (e.g. after rdhip_ridge_detection), the
 * reference CPU class's `order_` step: rd::orderSamples (rd/cpu/brute_force/order.hpp:42-142)
 * with the condition |p - q|^2 <= r2^2 of rd/cpu/brute_force/ridge_detection.inl:148-168.
 * order (capacity K) receives point indices (rdhip_get_points order) chain after chain;
 * chain_offsets (capacity K + 1) the chain boundaries; out_chains the number of chains. */
int rdhip_order_chains(rdhip_ctx* ctx, double r2, int32_t* order, int64_t* chain_offsets, int64_t* out_chains);

/* ---- differentiated evolve (no reference counterpart: the reference has no derivative) ----
 *
 * rdhip_evolve_fwd runs exactly `iters` evolve iterations with radius r from the current
 * chosen set (as rdhip_evolve_fixed) and keeps what rdhip_evolve_bwd needs: the chosen set
 * at the start of every run of `checkpoint_every` iterations, and the memberships of the
 * last run.  checkpoint_every = 0 keeps the memberships of every iteration (iters x n
 * int32 of device memory); otherwise earlier runs are recomputed from their checkpoints
 * during the backward pass (about one more forward pass of work).  checkpoint_every < 0
 * picks the interval: every iteration's memberships when iters x (4n + 4K + dim K) bytes plus
 * the checkpoints fit RDHIP_OPT_VJP_MEMORY_MB, else the longest run that does
 * (rdhip_timers.vjp_checkpoint_every reports it).
 *
 * rdhip_evolve_bwd is the vector-Jacobian product of that forward pass with the
 * memberships (labels, counts and moved coordinates of every iteration) held fixed:
 *   grad_chosen (in):  dL/dS_T, K x dim (the K chosen points after the forward pass);
 *   grad_points (out): dL/dP, n x dim;
 *   grad_seeds (out, optional): dL/dS_0, K x dim;
 *   seed_to_points != 0: S_0 came from rdhip_choose, so dL/dS_0 is also added to the
 *     samples it was drawn from (an error when the chosen set was set explicitly).
 * Buffers are host or device (on_device) arrays in ROW_MAJOR or SOA layout (dense).  The
 * chosen set is left as the forward pass produced it.  Any call that changes the points or
 * the chosen set in between invalidates the forward pass (RDHIP_ERR_STATE). */
int rdhip_evolve_fwd(rdhip_ctx* ctx, double r, int64_t iters, int64_t checkpoint_every);
int rdhip_evolve_bwd(rdhip_ctx* ctx, const void* grad_chosen, void* grad_points, void* grad_seeds, int layout,
                     int on_device, int seed_to_points);

/* ---- the choose without allocation -------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the greedy
 * choose of rdhip_choose (rd/cpu/brute_force/choose.hpp:36-70) on caller-owned device buffers,
 * with no context.  It allocates nothing, copies nothing and does not synchronise: it only
 * enqueues kernels on hip_stream (NULL = the null stream), which must belong to the current
 * device.  Its work memory (about 16 MiB) is a fixed device array of the library, one per
 * device, held until the process ends; calls that overlap on one device race.
 *   points:  n points (device), COL_MAJOR / SOA: coordinate d of point i at points[d*stride + i]
 *            (stride >= n, 0 = n); n in [1, RDHIP_NOALLOC_MAX_POINTS];
 *   chosen:  device, capacity n per coordinate: chosen point k at chosen[d*chosen_stride + k]
 *            (chosen_stride >= n, 0 = n);
 *   indices: device int32[n]: the chosen points' input indices, ascending;
 *   count:   device int32: the number of chosen points.
 * Results equal rdhip_choose's. */
#define RDHIP_NOALLOC_MAX_POINTS (1 << 24)
int rdhip_choose_noalloc(int dtype, int dim, const void* points, int64_t n, int64_t stride, double r, void* chosen,
                         int64_t chosen_stride, int32_t* indices, int32_t* count, void* hip_stream);

/* Decimate the chosen set with radius r2 and `rule`; keep (host, optional, capacity K)
 * receives 1 for kept and 0 for removed points, in the pre-decimation order. */
int rdhip_decimate(rdhip_ctx* ctx, double r2, int rule, uint8_t* keep, int64_t* out_count);

#ifdef __cplusplus
}
#endif

#endif /* RDHIP_H */
