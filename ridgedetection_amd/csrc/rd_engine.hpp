// rd_engine.hpp -- host-side engine: device state and phase orchestration for one
// "segment set" (a whole cloud, the leaves of a tile tree, or a batch of clouds).
#pragma once

#include "rd_choose.hpp"
#include "rd_common.hpp"
#include "rd_decimate.hpp"
#include "rd_chain.hpp"
#include "rd_evolve.hpp"
#include "rd_local.hpp"
#include "rd_sort.hpp"
#include "rd_vjp.hpp"

#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>

namespace rdhip {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string& where)
        : std::runtime_error(where + ": " + hipGetErrorString(c)), code(c) {}
};

inline void check(hipError_t e, const char* where) {
    if (e != hipSuccess) throw HipError(e, where);
}

// Grow-only device buffer.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T* reserve(size_t count) {
        if (count <= n && p) return p;
        release();
        size_t c = count ? count : 1;
        check(deviceMalloc((void**)&p, sizeof(T) * c), "hipMalloc");
        n = c;
        return p;
    }
    T* get() const { return p; }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
};

enum Layout { LAYOUT_ROW_MAJOR = 0, LAYOUT_COL_MAJOR = 1, LAYOUT_SOA = 2 };

struct Timers {
    double wholeMs = 0, buildMs = 0, chooseMs = 0, evolveMs = 0, decimateMs = 0, refineMs = 0;
    long long evolveIterations = 0, rounds = 0;
    double scanKernelMs = 0;      // HIP-event time of the evolve scan kernel (when profiling)
    long long scanLaunches = 0;
    double iterationMs = 0;       // HIP-event time of whole evolve iterations (when profiling)
    double forwardMs = 0;         // evolveForward calls (stream synchronised at the end)
    double backwardMs = 0;        // evolveBackward calls, device buffers in and out
    // algorithmic bytes of the evolve iterations run: per segment and iteration, its samples
    // read once (nP * dim * sizeof(T)) and its chosen points read, summed and updated
    // (K * (3 * dim * sizeof(T) + 4))
    double evolveBytes = 0;
    double evolvePairs = 0;             // samples x chosen points, summed over segments and iterations
    int scanPpt = 0, scanThreads = 0;   // scan shape in use (after autotuning)
    int scanPersistent = 0;             // ... and whether its blocks pull items from a queue
    long long vjpCheckpointEvery = 0;   // iterations per recorded run of the last VJP forward
    int chainPath = 0;                  // the last iteration enqueued took the rd_chain.hip path
    int screenForm = 0;                 // screen of the last scan enqueued: 0 direct, 1 expanded,
                                        // 2 expanded on the matrix cores
    double tuneMs = 0;                  // time spent timing candidate shapes
};

struct RunOptions {
    long long maxEvolveIters = 0;   // per evolve call; 0 = until convergence (reference)
    long long maxRounds = 0;        // 0 = until the count is stable (reference)
    int ppt = 4;                    // samples per lane in the evolve scan
    int scanThreads = 256;          // threads per scan block (128, 256, 512)
    bool autotune = true;           // pick ppt x scanThreads by timing the scan on first use
    int itersPerSync = 8;           // evolve iterations enqueued between convergence checks
    bool profile = false;           // record HIP events around the scan kernel / iterations
    bool persistent = false;        // scan as persistent blocks fed by a device work queue
    bool graph = true;              // replay captured batches of iterations (hipGraph)
    bool local = true;              // evolve small segments inside one workgroup each
    int localMaxP = 16384;          // ... when every active segment has at most this many samples
    int localSmallP = 2048;         // ... otherwise only those with at most this many
    int64_t localMaxWork = 1 << 22; // ... and samples x chosen points at most this
    bool segmented = false;         // batches whose active segments all have <= segmentedMaxP
    int segmentedMaxP = 65536;      // samples: batch-wide scan + per-segment sort and sums
    bool positionCopy = true;       // indexed segments: evolve reads a position-ordered copy
    int64_t vjpBudgetBytes = int64_t(1) << 30;   // automatic checkpoint interval of the VJP forward
    bool tolerance = false;         // batch-wide sums in parallel in a wider type (not the
                                    // reference's serial order); segment-local ones stay exact
    long long evolveGuard = 100000; // an evolve still moving after this many iterations fails
                                    // (0 = no limit, as in the reference)
    int chain = 1;                  // member ordering without the sort (rd_chain.hip) for
                                    // segments of at most kChainMaxKeys chosen points: 1 where it
                                    // pays (fewer keys per segment than samples per block), 2
                                    // always, 0 never
    int screen = 2;                 // evolve-scan screen (rd_nearest.hpp): 2 the fastest form as
                                    // timed by the autotuner (direct where it does not run),
                                    // 1 expanded, 3 expanded on the matrix cores (FP32; FP64:
                                    // expanded), 0 direct
};

template <typename T>
class Engine {
public:
    Engine(int dim, hipStream_t stream);
    ~Engine();

    int dim() const { return dim_; }
    hipStream_t stream() const { return st_; }

    // -- points --------------------------------------------------------------------
    // Uploads n points (host or device pointer) given in ROW_MAJOR, COL_MAJOR (stride =
    // elements between rows of the same coordinate) or SOA layout into the SoA buffer.
    void setPoints(const T* src, int64_t n, int layout, int64_t stride, bool onDevice);
    // point buffer for n points filled on the device by the caller (SoA, pointStride())
    T* reservePoints(int64_t n);
    int64_t numPoints() const { return nPoints_; }
    const T* devicePoints() const { return P_.get(); }
    int64_t pointStride() const { return pStride_; }

    // -- segments ------------------------------------------------------------------
    // Defines the segments over the SoA buffer (pOff/nP/nOwn/group must be set; sOff and
    // sCap are assigned here from capacities).  Group g = consecutive segments with
    // group == g.  r1/r2 are per segment.
    // indexed: segment positions map to points through the index list filled via
    // reserveIndex() (tiles whose neighbour samples repeat other tiles' points).
    void setSegments(const std::vector<SegDesc>& segs, const std::vector<int64_t>& capacities,
                     const std::vector<T>& r1, const std::vector<T>& r2, bool indexed = false);
    int* reserveIndex(size_t positions) { return pIndexBuf_.reserve(positions); }
    // one segment over all points; choose takes the first nOwn of them (< 0: all)
    void setSingleSegment(T r1, T r2, int64_t nOwn = -1);
    // Replaces the radii of the current segments without touching their chosen points.
    void setRadii(const std::vector<T>& r1, const std::vector<T>& r2);
    int numSegments() const { return (int)segs_.size(); }
    void rebuildWorkItems() {
        buildWorkItems();
        check(hipStreamSynchronize(st_), "rebuildWorkItems");
    }

    // -- phases --------------------------------------------------------------------
    void choose();
    // Re-lays the chosen slots so that every segment's capacity is its live count (rounded
    // to a screening tile): counts only shrink after choose, and a dense slot range keeps
    // the per-iteration sort to few radix passes when many segments are batched.
    void compactSlots();
    // Runs evolve iterations on every segment whose flag in `active` is set (all when
    // empty) until each converges or reaches maxIter.  Returns the number of iterations.
    long long evolve(long long maxIter, const std::vector<char>& active = {});
    // Exactly `iters` evolve iterations on every segment, ignoring convergence (benchmarks,
    // frozen-membership forward passes).
    void evolveFixed(long long iters);
    // One evolve iteration on all segments, recording labels/counts/sums/moved on host.
    int evolveIteration(std::vector<int>* labels, std::vector<int>* counts, std::vector<T>* sums,
                        std::vector<unsigned short>* moved);
    void decimate(bool noCopyRule, std::vector<unsigned char>* keep = nullptr);

    // -- differentiated evolve (one segment) ------------------------------------------
    // Forward: exactly `iters` iterations (as evolveFixed), keeping what the backward needs:
    // the chosen set at the start of every run of `every` iterations (checkpoints) and the
    // labels / counts / moved flags of the last run.  every <= 0 or >= iters keeps the logs
    // of all iterations (no recomputation).
    void evolveForward(long long iters, int every);
    // Backward of the last forward with memberships frozen: gS = dL/dS_T (SoA, stride
    // gsStride, K columns), gP = dL/dP (SoA, stride gpStride, n columns, overwritten),
    // gS0 = dL/dS_0 (optional, SoA, stride gs0Stride).  Earlier runs are recomputed from
    // their checkpoints.  seedToPoints: S_0 came from choose, so dL/dS_0 is added to the
    // samples it was drawn from.  All pointers are device pointers.
    void evolveBackward(const T* gS, int64_t gsStride, T* gP, int64_t gpStride, T* gS0, int64_t gs0Stride,
                        bool seedToPoints);
    bool chosenFromPoints() const { return chosenFromPoints_; }

    // Chain ordering of the current chosen points (all segments, in getChosen order) with
    // the reference's r2 condition: point indices chain after chain into `order`, chain
    // boundaries into `offsets` (size chains + 1).  Returns the number of chains.
    int64_t orderChains(T r2, std::vector<int>& order, std::vector<int64_t>& offsets);

    // -- drivers -------------------------------------------------------------------
    // Brute force over every segment independently (each its own group): choose, then
    // rounds of evolve + decimate until each segment's count is stable.
    void ridgeDetectionLocal(const RunOptions& opt);
    // Mixed: choose per segment, then rounds of evolve per segment + one decimateNoCopy over
    // all segments (single group) until the total count is stable.
    void ridgeDetectionMixed(const RunOptions& opt);

    // -- results -------------------------------------------------------------------
    std::vector<int> countsHost();
    int64_t totalCount();
    // Chosen points of all segments in segment order, into `out` with `layout`.
    void getChosen(T* out, int layout, bool onDevice, int64_t outStride = 0);
    std::vector<int> chosenIndicesHost(int seg);
    // chosen points of segment `seg` from host or device memory: ROW_MAJOR (stride >= dim
    // between points) or COL_MAJOR / SOA (stride >= K between coordinates); 0 = packed
    void setChosen(int seg, const T* S, int64_t K, int layout, int64_t stride = 0, bool onDevice = false);

    Timers timers;
    RunOptions options;
    std::string guardMessage() const {
        return "evolve did not converge within the iteration guard (" + std::to_string(options.evolveGuard) +
               " iterations; RDHIP_OPT_EVOLVE_GUARD raises it, 0 removes it)";
    }
    int64_t lastEvolveIters = 0;

private:
    int dim_;
    hipStream_t st_;
    // points
    DevBuf<T> P_;
    int64_t nPoints_ = 0, pStride_ = 0;
    DevBuf<int> pIndexBuf_;
    bool useIndex_ = false;
    const int* pIndexPtr() const { return useIndex_ ? pIndexBuf_.get() : nullptr; }
    // samples as the evolve passes read them: the points through the position index, or
    // (indexed segments, positionCopy) a copy in position order, written once per points /
    // segments setting, read without the index
    struct EvolveSource {
        const T* P;
        int64_t stride;
        const int* index;
    };
    DevBuf<T> Pseg_;
    int64_t psegStride_ = 0;
    bool psegValid_ = false;
    void preparePositions();
    DevBuf<int> rowSlots_;
    int64_t numRowSlots_ = 0;
    uint64_t rowSlotsVersion_ = ~uint64_t(0);
    DevBuf<int> denseOf_;     // batch sort key of every live slot (-1: none)
    bool denseKeys_ = false;
    void prepareRows(const std::vector<int>& done);
    EvolveSource evolveSource() const {
        if (useIndex_ && psegValid_ && options.positionCopy) return {Pseg_.get(), psegStride_, nullptr};
        return {P_.get(), pStride_, pIndexPtr()};
    }
    // segments
    std::vector<SegDesc> segs_;
    std::vector<int> groupBegin_, groupEnd_;
    std::vector<T> r1_, r2_;
    DevBuf<SegDesc> dSegs_;
    DevBuf<int> dSegK_, dSegDone_, dSegIter_, dSegMoved_, dGroupBegin_, dGroupEnd_, dSlotSeg_;
    DevBuf<T> dSegR1Sqr_, dGroupR2Sqr_;
    DevBuf<WorkItem> dItems_;
    int numItems_ = 0;
    int64_t totalSlots_ = 0, sStride_ = 0, totalP_ = 0;
    DevBuf<T> S_;
    DevBuf<int> chosenIdx_;
    // evolve scratch
    DevBuf<int> labels_, keys_, members_, chainOff_, chainEnd_, counts_;
    DevBuf<int> longList_;   // wave-summed chains (ShiftArgs::longList)
    DevBuf<T> sums_, sortedP_;
    hipGraphExec_t graphExec_ = nullptr;   // captured evolve iterations (captureIterations)
    // member ordering of rd_chain.hip: per-segment count tables, key totals and the layout
    // arrays of ChainArgs, sized for the chosen counts of countsEpoch_ == chainEpoch_
    DevBuf<int> chainHist_, chainTotals_, chainSegKeyCap_, chainSegRowOff_, chainLongIdx_, chainLongCount_;
    DevBuf<int64_t> chainSegHistOff_, chainSegMemberOff_, chainTotalPos_;
    int64_t chainMStride_ = 0;
    int chainTotalRows_ = 0, chainKeyCapMax_ = 0;
    double chainMeanKeys_ = 0;   // keys per segment, averaged over samples
    bool chainReady_ = false;
    long long chainEpoch_ = -1, countsEpoch_ = 0;   // countsEpoch_: bumped when chosen counts change
    int itemsPer_ = 0;                              // samples per work item of buildWorkItems
    // expanded-form screen: table rows (dim + 1) x sStride_ and per-segment centre and bound
    // (rd_screen.hip), rebuilt before every scan that uses them; tunedForm_: the form the
    // autotuner picked
    DevBuf<T> xTable_, segX_;
    static constexpr int kMatrixForm = sizeof(T) == 4 ? 2 : 1;   // the matrix-core form is FP32
    int tunedForm_ = 0;
    // screen form of the next scans (0 direct, 1 expanded, 2 expanded on the matrix cores)
    int screenForm() const {
        switch (options.screen) {
            case 0: return 0;
            case 1: return 1;
            case 3: return kMatrixForm;
            default: return tunedForm_;
        }
    }
    // the forms the tuner may pick from under the current option (the batch tuner times its
    // shapes in the first one, then the best shape in the others)
    std::vector<int> screenForms() const {
        if (options.screen != 2) return {screenForm()};
        if (kMatrixForm == 2) return {0, 2, 1};
        return {0, 1};
    }
    // enqueues the table for the scan when form > 0, and points sa at it (else clears those fields)
    void prepScreen(ScanArgs<T>& sa, int form);
    int64_t keyBound_ = 0;   // 1 + highest live chosen slot (host copy, refreshed per phase)
    void updateKeyBound();
    DevBuf<unsigned char> moved_;   // [dim][totalSlots]
    DevBuf<unsigned char> sortWs_;
    DevBuf<int64_t> dN_, dM_;
    DevBuf<int> flags_;   // [0] allDone, [1] totalIters, [2] overflow
    // choose scratch
    DevBuf<unsigned char> covered_;
    std::vector<DevBuf<WorkItem>*> chooseItems_;
    // decimate scratch
    DevBuf<int> decInts_;
    DevBuf<int64_t> decI64_;
    DecWorkspace decWs_;
    const int* segSkip_ = nullptr;   // optional per-segment skip mask for decimate

    void buildWorkItems();
    void orderItemsByCost();
    std::vector<WorkItem> items_;
    // compact mode (more than one segment): the scan writes (label, position) pairs of the
    // work items of active segments only, into a dense range the sort runs on
    bool compact() const { return segs_.size() > 1; }
    DevBuf<int> cKeys_, cPos_;
    int64_t activeWork_ = 0;
    void rebuildActive(const std::vector<int>& segDone);
    int persistentBlocks_ = 0;
    int persistentBlocks();
    // captures `batch` iterations (+ finalize when `finalize`) into an executable graph;
    // returns false when the stream cannot be captured
    bool captureIterations(int batch, bool finalize, int capIter, hipGraphExec_t* exec);
    void ensureEvolveScratch();
    // per-iteration log targets of a recorded forward (labels of every position; counts and
    // per-coordinate moved flags of every slot, the flags at moved[d * movedStride + slot])
    struct IterLog {
        int* labels;
        int* counts;
        unsigned char* moved;
        int64_t movedStride;
    };
    // sizes the rd_chain.hip scratch for the current segments and chosen counts (outside graph
    // capture); chainUsable: the next iteration may take that path
    void ensureChainScratch();
    void enqueueChainOrder(int itemSize, const int* labels);
    bool chainUsable(int itemSize) const;
    // the general member ordering: stable sort of the scan's keys over totalP positions (keys
    // below kb), member coordinates into sortedP_, chain ranges and the wave list; returns the
    // wave list's shift
    int enqueueSortOrder(const int* labels, int64_t totalP, int64_t kb);
    // finalize: also the per-segment convergence bookkeeping (iteration cap capIter)
    void autotuneBatchScan();
    void enqueueIteration(bool record, int slot = -1, const IterLog* log = nullptr, bool finalize = false,
                          int capIter = 0);
    // state of the last evolveForward
    struct Tape {
        long long iters = 0;
        int every = 0;
        int runs = 0;
        int loggedRun = -1;      // run whose logs the buffers hold
        int K = 0;
        int64_t positions = 0;
        T r = 0;
        uint64_t version = 0;    // engine state the tape belongs to
        DevBuf<T> ckpt;          // [runs][dim][sStride]
        DevBuf<T> finalS;        // [dim][sStride]
        DevBuf<int> labels;      // [every][positions]
        DevBuf<int> counts;      // [every][K]
        DevBuf<unsigned char> moved;   // [every][dim][K]
        DevBuf<T> G, A;          // upstream gradient and step coefficients
        bool valid = false;
    } tape_;
    uint64_t version_ = 0;       // bumped by every change of points, segments or chosen set
    bool chosenFromPoints_ = false;
    void recordRun(int run);
    void readEvents(int used);
    std::vector<hipEvent_t> events_;   // per batch slot: scan start, scan end, iteration end
    void fillSentinels();
    // Single-segment runs: times the scan kernel for every (samples per lane, block size)
    // candidate on the current cloud and chosen set and keeps the fastest; choices are
    // cached per (value type, dimension, log2 samples, log2 chosen points).
    void autotuneScan();
    void addEvolveBytes(const std::vector<int>& itersPerSeg);
    // Evolves the small active segments with the segment-local kernel and marks them done;
    // returns the largest iteration count among them (0 when none).
    int evolveLocal(long long maxIter, std::vector<int>& done);
    // Whether the active segments of a batch iterate as batch-wide scan + per-segment sort
    // and sums (sets the LDS layout of that kernel).
    bool chooseSegmented(const std::vector<int>& done);
    bool segmentedNow_ = false;
    int segKStride_ = 16, segChunk_ = 1;
    DevBuf<int> dLocalList_;
};

}  // namespace rdhip
