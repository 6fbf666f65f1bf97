// rd_engine.hip -- host-side orchestration of the ridge-detection phases on one GPU.
//
// An Engine owns a SoA point buffer and a set of segments over it.  A segment is the
// unit the reference runs its loop on: the whole cloud (brute force,
// rd/gpu/device/brute_force/simulation.inl:183-384), one leaf tile (tiled RD,
// rd/cpu/tiled/rd_tiled.hpp:756-1732) or one cloud of a batch.  Every phase runs on all
// segments in one launch per kernel; all counts stay on the device and the host reads
// them back once per round (or once every few evolve iterations to test convergence).
#include "rd_engine.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>

namespace rdhip {

namespace {

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

__global__ void fill_int_kernel(int* __restrict__ a, int64_t n, int v, const int* __restrict__ skip) {
    if (skip && *skip) return;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) a[i] = v;
}

// out[idx[i]] = i for i < n
__global__ void scatter_rank_kernel(const int* __restrict__ idx, int64_t n, int* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        out[idx[i]] = (int)i;
}

// a[i] = b[i] = 0 for i < n (one launch for the chain bounds of an iteration)
__global__ void zero_pair_kernel(int* __restrict__ a, int* __restrict__ b, int64_t n, const int* __restrict__ skip) {
    if (skip && *skip) return;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        a[i] = 0;
        b[i] = 0;
    }
}

// ROW_MAJOR (n x dim, row stride `stride`) -> SoA (dim x pStride)
template <typename T>
__global__ void row_to_soa_kernel(const T* __restrict__ src, int64_t n, int dim, int64_t stride,
                                  T* __restrict__ dst, int64_t pStride) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        for (int d = 0; d < dim; ++d) dst[d * pStride + i] = src[i * stride + d];
}

// rows[p][0, DIM) = P[0, DIM)[p]: the points row-major, so that a gather finds a point's
// coordinates in one cache line
template <typename T, int DIM>
__global__ void points_to_rows_kernel(const T* __restrict__ P, int64_t pStride, int64_t n, T* __restrict__ rows) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
#pragma unroll
        for (int d = 0; d < DIM; ++d) rows[i * DIM + d] = P[d * pStride + i];
}

// out[d][i] = rows[pIndex[i]][d] for the n positions
template <typename T, int DIM>
__global__ void gather_positions_kernel(const T* __restrict__ rows, const int* __restrict__ pIndex, int64_t n,
                                        T* __restrict__ out, int64_t outStride) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t src = pIndex[i];
        T v[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] = rows[src * DIM + d];
#pragma unroll
        for (int d = 0; d < DIM; ++d) out[d * outStride + i] = v[d];
    }
}

// gathers the live chosen points of every segment into a dense output
template <typename T>
__global__ void gather_chosen_kernel(const T* __restrict__ S, int64_t sStride, const SegDesc* __restrict__ segs,
                                     const int* __restrict__ segK, const int64_t* __restrict__ outOff, int dim,
                                     T* __restrict__ out, int layout, int64_t outStride) {
    const int s = blockIdx.x;
    const SegDesc sd = segs[s];
    const int K = segK[s];
    const int64_t o = outOff[s];
    for (int k = threadIdx.x; k < K; k += 256)
        for (int d = 0; d < dim; ++d) {
            T v = S[d * sStride + sd.sOff + k];
            if (layout == LAYOUT_ROW_MAJOR)
                out[(o + k) * outStride + d] = v;
            else
                out[d * outStride + o + k] = v;
        }
}

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

// copies every segment's first `cnt` slots (coordinates and chosen indices) to a new layout
template <typename T>
__global__ void relayout_slots_kernel(const T* __restrict__ S, int64_t sStride, const int* __restrict__ idx,
                                      const int64_t* __restrict__ oldOff, const int64_t* __restrict__ newOff,
                                      const int* __restrict__ cnt, int dim, T* __restrict__ S2, int64_t s2Stride,
                                      int* __restrict__ idx2) {
    const int s = blockIdx.x;
    const int64_t o = oldOff[s], n = newOff[s];
    for (int k = threadIdx.x; k < cnt[s]; k += 256) {
        for (int d = 0; d < dim; ++d) S2[d * s2Stride + n + k] = S[d * sStride + o + k];
        idx2[n + k] = idx[o + k];
    }
}

}  // namespace

template <typename T>
Engine<T>::Engine(int dim, hipStream_t stream) : dim_(dim), st_(stream) {
    if (dim < 1 || dim > kMaxDim) throw std::invalid_argument("dimension must be in [1, 12]");
    flags_.reserve(10);  // [0] all done, [1] iterations, [2] overflow, [3] queue, [4] active items,
                         // [5] [6] local evolve, [8] blocks done (fused convergence bookkeeping)
    dN_.reserve(1);
    dM_.reserve(1);
}

template <typename T>
Engine<T>::~Engine() {
    if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
    for (auto ev : events_) (void)hipEventDestroy(ev);
    if (decWs_.nbr) (void)hipFree(decWs_.nbr);
    if (decWs_.par) (void)hipFree(decWs_.par);
    if (decWs_.boxes) (void)hipFree(decWs_.boxes);
    for (auto* b : chooseItems_) delete b;
}

template <typename T>
T* Engine<T>::reservePoints(int64_t n) {
    if (n < 0) throw std::invalid_argument("negative point count");
    ++version_;
    psegValid_ = false;
    nPoints_ = n;
    pStride_ = round_up(std::max<int64_t>(n, 1), 64);
    P_.reserve((size_t)pStride_ * dim_);
    return P_.get();
}

template <typename T>
void Engine<T>::setPoints(const T* src, int64_t n, int layout, int64_t stride, bool onDevice) {
    if (n < 0) throw std::invalid_argument("negative point count");
    ++version_;
    psegValid_ = false;
    nPoints_ = n;
    pStride_ = round_up(std::max<int64_t>(n, 1), 64);
    P_.reserve((size_t)pStride_ * dim_);
    if (n == 0) return;
    if (layout == LAYOUT_ROW_MAJOR) {
        if (stride <= 0) stride = dim_;
        const T* dsrc = src;
        DevBuf<T> staging;
        if (!onDevice) {
            staging.reserve((size_t)n * stride);
            check(hipMemcpyAsync(staging.get(), src, sizeof(T) * n * stride, hipMemcpyHostToDevice, st_), "upload");
            dsrc = staging.get();
        }
        hipLaunchKernelGGL((row_to_soa_kernel<T>), dim3(grid_for(n)), dim3(256), 0, st_, dsrc, n, dim_, stride,
                           P_.get(), pStride_);
        check(hipGetLastError(), "row_to_soa");
        check(hipStreamSynchronize(st_), "setPoints");
    } else if (layout == LAYOUT_COL_MAJOR || layout == LAYOUT_SOA) {
        if (stride <= 0) stride = n;
        check(hipMemcpy2DAsync(P_.get(), sizeof(T) * pStride_, src, sizeof(T) * stride, sizeof(T) * n, dim_,
                               onDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st_),
              "upload");
        check(hipStreamSynchronize(st_), "setPoints");
    } else {
        throw std::invalid_argument("unsupported memory layout");
    }
}

template <typename T>
void Engine<T>::setSingleSegment(T r1, T r2, int64_t nOwn) {
    if (nOwn < 0 || nOwn > nPoints_) nOwn = nPoints_;
    SegDesc s{};
    s.pOff = 0;
    s.nP = (int32_t)nPoints_;
    s.nOwn = (int32_t)nOwn;
    s.group = 0;
    setSegments({s}, {std::max<int64_t>(nOwn, 1)}, {r1}, {r2}, false);
}

template <typename T>
void Engine<T>::setSegments(const std::vector<SegDesc>& segs, const std::vector<int64_t>& capacities,
                            const std::vector<T>& r1, const std::vector<T>& r2, bool indexed) {
    const int ns = (int)segs.size();
    useIndex_ = indexed;
    psegValid_ = false;
    if (capacities.size() != segs.size() || r1.size() != segs.size() || r2.size() != segs.size())
        throw std::invalid_argument("segment arrays differ in length");
    ++version_;
    segs_ = segs;
    r1_ = r1;
    r2_ = r2;
    int64_t off = 0;
    for (int s = 0; s < ns; ++s) {
        segs_[s].sOff = off;
        int64_t cap = round_up(std::max<int64_t>(capacities[s], 1), kChosenTile);
        if (cap > 0x7fffffffLL) throw std::invalid_argument("segment capacity too large");
        segs_[s].sCap = (int32_t)cap;
        off += cap;
    }
    totalSlots_ = off;
    sStride_ = round_up(std::max<int64_t>(totalSlots_, 1), 64);
    if (totalSlots_ >= (int64_t(1) << 30)) throw std::invalid_argument("too many chosen slots");
    // groups: consecutive runs of equal `group`
    groupBegin_.clear();
    groupEnd_.clear();
    for (int s = 0; s < ns; ++s) {
        if (s == 0 || segs_[s].group != segs_[s - 1].group) {
            if (s > 0) groupEnd_.push_back(s);
            groupBegin_.push_back(s);
        }
        segs_[s].group = (int32_t)groupBegin_.size() - 1;
    }
    if (ns > 0) groupEnd_.push_back(ns);
    const int ng = (int)groupBegin_.size();
    std::vector<T> r1sq(ns), r2sq(ng);
    for (int s = 0; s < ns; ++s) r1sq[s] = r1_[s] * r1_[s];
    for (int g = 0; g < ng; ++g) r2sq[g] = r2_[groupBegin_[g]] * r2_[groupBegin_[g]];
    std::vector<int> slotSeg((size_t)totalSlots_);
    for (int s = 0; s < ns; ++s)
        std::fill(slotSeg.begin() + segs_[s].sOff, slotSeg.begin() + segs_[s].sOff + segs_[s].sCap, s);

    dSegs_.reserve(std::max(ns, 1));
    dSegK_.reserve(std::max(ns, 1));
    dSegDone_.reserve(std::max(ns, 1));
    dSegIter_.reserve(std::max(ns, 1));
    dSegMoved_.reserve(std::max(ns, 1));
    dSegR1Sqr_.reserve(std::max(ns, 1));
    dGroupR2Sqr_.reserve(std::max(ng, 1));
    dGroupBegin_.reserve(std::max(ng, 1));
    dGroupEnd_.reserve(std::max(ng, 1));
    dSlotSeg_.reserve(std::max<int64_t>(totalSlots_, 1));
    S_.reserve((size_t)sStride_ * dim_);
    chosenIdx_.reserve((size_t)std::max<int64_t>(totalSlots_, 1));
    if (ns > 0) {
        check(hipMemcpyAsync(dSegs_.get(), segs_.data(), sizeof(SegDesc) * ns, hipMemcpyHostToDevice, st_), "segs");
        check(hipMemcpyAsync(dSegR1Sqr_.get(), r1sq.data(), sizeof(T) * ns, hipMemcpyHostToDevice, st_), "r1");
        check(hipMemcpyAsync(dGroupR2Sqr_.get(), r2sq.data(), sizeof(T) * ng, hipMemcpyHostToDevice, st_), "r2");
        check(hipMemcpyAsync(dGroupBegin_.get(), groupBegin_.data(), sizeof(int) * ng, hipMemcpyHostToDevice, st_), "g");
        check(hipMemcpyAsync(dGroupEnd_.get(), groupEnd_.data(), sizeof(int) * ng, hipMemcpyHostToDevice, st_), "g");
        check(hipMemcpyAsync(dSlotSeg_.get(), slotSeg.data(), sizeof(int) * totalSlots_, hipMemcpyHostToDevice, st_),
              "slots");
        std::vector<int> zeros(ns, 0);
        check(hipMemcpyAsync(dSegK_.get(), zeros.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "K");
    }
    buildWorkItems();
    ensureEvolveScratch();
    check(hipStreamSynchronize(st_), "setSegments");
}

template <typename T>
void Engine<T>::setRadii(const std::vector<T>& r1, const std::vector<T>& r2) {
    const int ns = (int)segs_.size();
    if ((int)r1.size() != ns || (int)r2.size() != ns) throw std::invalid_argument("radius arrays differ in length");
    r1_ = r1;
    r2_ = r2;
    const int ng = (int)groupBegin_.size();
    std::vector<T> r1sq(ns), r2sq(ng);
    for (int s = 0; s < ns; ++s) r1sq[s] = r1_[s] * r1_[s];
    for (int g = 0; g < ng; ++g) r2sq[g] = r2_[groupBegin_[g]] * r2_[groupBegin_[g]];
    if (ns > 0) {
        check(hipMemcpyAsync(dSegR1Sqr_.get(), r1sq.data(), sizeof(T) * ns, hipMemcpyHostToDevice, st_), "r1");
        check(hipMemcpyAsync(dGroupR2Sqr_.get(), r2sq.data(), sizeof(T) * ng, hipMemcpyHostToDevice, st_), "r2");
        check(hipStreamSynchronize(st_), "setRadii");
    }
}

template <typename T>
void Engine<T>::buildWorkItems() {
    std::vector<WorkItem>& items = items_;
    items.clear();
    const int per = std::max(64, options.scanThreads) * std::max(1, options.ppt);
    itemsPer_ = per;
    for (int s = 0; s < (int)segs_.size(); ++s)
        for (int st = 0; st < segs_[s].nP; st += per) items.push_back({s, st, std::min(segs_[s].nP, st + per), 0});
    numItems_ = (int)items.size();
    dItems_.reserve(std::max<size_t>(items.size(), 1));
    if (!items.empty())
        check(hipMemcpyAsync(dItems_.get(), items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice, st_),
              "items");
    // choose filter items for the chunked (large segment) path
    for (auto* b : chooseItems_) delete b;
    chooseItems_.clear();
}

// Longest-first order for the persistent scan's work queue: cost ~ samples x live chosen.
// Active work items (segments not done), in cost order, with their dense output offsets;
// the count and the total live on the device so a captured graph stays valid.
template <typename T>
void Engine<T>::rebuildActive(const std::vector<int>& segDone) {
    std::vector<WorkItem> act;
    act.reserve(items_.size());
    int64_t off = 0;
    for (WorkItem it : items_) {
        if (segDone[it.seg]) continue;
        it.pad = (int32_t)off;
        off += it.end - it.start;
        act.push_back(it);
    }
    const int n = (int)act.size();
    if (n > 0)
        check(hipMemcpyAsync(dItems_.get(), act.data(), sizeof(WorkItem) * n, hipMemcpyHostToDevice, st_), "items");
    check(hipMemcpyAsync(flags_.get() + 4, &n, sizeof(int), hipMemcpyHostToDevice, st_), "items");
    check(hipMemcpyAsync(dN_.get(), &off, sizeof(int64_t), hipMemcpyHostToDevice, st_), "n");
    check(hipStreamSynchronize(st_), "rebuildActive");
    activeWork_ = off;
}

template <typename T>
void Engine<T>::orderItemsByCost() {
    if (items_.size() < 2) return;
    std::vector<int> K = countsHost();
    std::stable_sort(items_.begin(), items_.end(), [&](const WorkItem& a, const WorkItem& b) {
        return (int64_t)(a.end - a.start) * K[a.seg] > (int64_t)(b.end - b.start) * K[b.seg];
    });
    check(hipMemcpyAsync(dItems_.get(), items_.data(), sizeof(WorkItem) * items_.size(), hipMemcpyHostToDevice, st_),
          "items");
    check(hipStreamSynchronize(st_), "items");
}

template <typename T>
int Engine<T>::persistentBlocks() {
    if (!options.persistent) return 0;
    if (persistentBlocks_ == 0) {
        int dev = 0, cus = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
        persistentBlocks_ = std::max(1, cus) * 8;
    }
    return persistentBlocks_;
}

// Slots of the segments still evolving, when far fewer than the slot range (tiles keep one
// slot per own sample): the sum rows of this evolve call visit only those.
template <typename T>
void Engine<T>::prepareRows(const std::vector<int>& done) {
    numRowSlots_ = 0;
    const std::vector<int> K = countsHost();
    std::vector<int> rows;
    for (size_t s = 0; s < segs_.size(); ++s)
        if (!done[s])
            for (int k = 0; k < K[s]; ++k) rows.push_back((int)(segs_[s].sOff + k));
    denseKeys_ = false;
    if (rows.empty()) return;
    // a single cloud keeps its slot keys (already dense); segment sets always list
    if (!compact() && 4 * (int64_t)rows.size() > std::max<int64_t>(keyBound_, 1)) return;
    rowSlots_.reserve(rows.size());
    check(hipMemcpyAsync(rowSlots_.get(), rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, st_),
          "row slots");
    numRowSlots_ = (int64_t)rows.size();
    rowSlotsVersion_ = version_;
    if (compact()) {
        // sort keys of a batch: a slot's index in the list (the active clouds' live chosen
        // points only, so the sort spans fewer key bits as clouds converge)
        const int64_t kb = std::max<int64_t>(keyBound_, 1);
        denseOf_.reserve(kb);
        hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(kb)), dim3(256), 0, st_, denseOf_.get(), kb, -1, nullptr);
        hipLaunchKernelGGL(scatter_rank_kernel, dim3(grid_for(numRowSlots_)), dim3(256), 0, st_, rowSlots_.get(),
                           numRowSlots_, denseOf_.get());
        check(hipGetLastError(), "dense keys");
        denseKeys_ = true;
    }
}

template <typename T>
void Engine<T>::preparePositions() {
    if (!useIndex_ || psegValid_ || !options.positionCopy) return;
    int64_t n = 0;
    for (auto& s : segs_) n = std::max<int64_t>(n, s.pOff + s.nP);
    if (n <= 0) return;
    psegStride_ = round_up(n, 64);
    Pseg_.reserve((size_t)psegStride_ * dim_);
    const int64_t np = std::max<int64_t>(nPoints_, 1);
    DevBuf<T> rows;
    rows.reserve((size_t)np * dim_);
    bool ok = false;
    RD_DISPATCH_DIM(dim_, {
        hipLaunchKernelGGL((points_to_rows_kernel<T, DIM>), dim3(grid_for(nPoints_)), dim3(256), 0, st_, P_.get(),
                           pStride_, nPoints_, rows.get());
        hipLaunchKernelGGL((gather_positions_kernel<T, DIM>), dim3(grid_for(n)), dim3(256), 0, st_, rows.get(),
                           pIndexBuf_.get(), n, Pseg_.get(), psegStride_);
        ok = true;
    });
    if (!ok) throw std::invalid_argument("unsupported dimension");
    check(hipGetLastError(), "gather_positions");
    check(hipStreamSynchronize(st_), "gather_positions");   // rows is released on return
    psegValid_ = true;
}

template <typename T>
void Engine<T>::ensureEvolveScratch() {
    int64_t totalP = 0;
    for (auto& s : segs_) totalP = std::max<int64_t>(totalP, s.pOff + s.nP);
    totalP = std::max<int64_t>(totalP, 1);
    totalP_ = totalP;
    labels_.reserve(totalP);
    keys_.reserve(totalP);
    members_.reserve(totalP);
    if (segs_.size() > 1) {
        cKeys_.reserve(totalP);
        cPos_.reserve(totalP);
    }
    covered_.reserve(totalP);
    const int64_t slots = std::max<int64_t>(totalSlots_, 1);
    chainOff_.reserve(slots);
    chainEnd_.reserve(slots);
    longList_.reserve(long_list_entries(totalP, kLongMinShift));
    counts_.reserve(slots);
    moved_.reserve((size_t)slots * dim_);
    sums_.reserve((size_t)sStride_ * dim_);
    sortWs_.reserve(sort_workspace_bytes(totalP, (int)std::max<int64_t>(totalSlots_, 2)));
    sortedP_.reserve((size_t)totalP * dim_);
    xTable_.reserve((size_t)(dim_ + 1) * std::max<int64_t>(sStride_, 1));
    segX_.reserve((size_t)std::max<size_t>(segs_.size(), 1) * kScreenStride);
    check(hipMemcpyAsync(dN_.get(), &totalP, sizeof(int64_t), hipMemcpyHostToDevice, st_), "n");
    // decimate scratch: elemBase, groupOff/Cnt/Left, elemSlot, elemGroup, c1, c2, rowOff, alive
    const int ns = std::max<int>((int)segs_.size(), 1);
    const int ng = std::max<int>((int)groupBegin_.size(), 1);
    decInts_.reserve((size_t)ns + 3 * ng + 5 * slots + (slots + 1) + scan_partials_count(std::max<int64_t>(slots + 1, ns)));
    decI64_.reserve(2);
}

// Layout of the rd_chain.hip member ordering for the current segments and chosen counts: per
// segment its key capacity (its live count), its count table sized for the smallest scan item
// (128 samples, so every scan shape fits), its rows of key totals and its member positions.
template <typename T>
void Engine<T>::ensureChainScratch() {
    chainReady_ = false;
    const int ns = (int)segs_.size();
    if (!options.chain || ns == 0 || dim_ > kChainMaxDim) return;
    const std::vector<int> K = countsHost();
    std::vector<int> cap(ns), rowOff(ns + 1);
    std::vector<int64_t> histOff(ns), memOff(ns + 1);
    int rows = 0, capMax = 0;
    int64_t h = 0, m = 0;
    double keySamples = 0, samples = 0;
    for (int s = 0; s < ns; ++s) {
        cap[s] = std::max(K[s], 0);
        keySamples += (double)cap[s] * segs_[s].nP;
        samples += segs_[s].nP;
        capMax = std::max(capMax, cap[s]);
        rowOff[s] = rows;
        rows += cap[s];
        histOff[s] = h;
        h += (int64_t)cap[s] * ((segs_[s].nP + 127) / 128);
        memOff[s] = m;
        m += chain_region(segs_[s].nP, cap[s]);
    }
    rowOff[ns] = rows;
    memOff[ns] = m;
    if (capMax > kChainMaxKeys || rows == 0 || m >= (int64_t(1) << 31)) return;
    chainSegKeyCap_.reserve(ns);
    chainSegRowOff_.reserve(ns + 1);
    chainSegHistOff_.reserve(ns);
    chainSegMemberOff_.reserve(ns + 1);
    chainTotalPos_.reserve(1);
    chainHist_.reserve((size_t)std::max<int64_t>(h, 1));
    chainTotals_.reserve((size_t)rows);
    if (sortedP_.n < (size_t)(m * dim_)) sortedP_.reserve((size_t)(m * dim_));
    const int64_t entries = kLongWaves;   // the length-based wave list (at most kLongWaves / dim used)
    if (longList_.n < (size_t)entries) longList_.reserve((size_t)entries);
    chainLongIdx_.reserve((size_t)std::max<int64_t>(totalSlots_, 1));
    chainLongCount_.reserve(1);
    check(hipMemcpyAsync(chainSegKeyCap_.get(), cap.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "chain");
    check(hipMemcpyAsync(chainSegRowOff_.get(), rowOff.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice, st_),
          "chain");
    check(hipMemcpyAsync(chainSegHistOff_.get(), histOff.data(), sizeof(int64_t) * ns, hipMemcpyHostToDevice, st_),
          "chain");
    check(hipMemcpyAsync(chainSegMemberOff_.get(), memOff.data(), sizeof(int64_t) * (ns + 1), hipMemcpyHostToDevice,
                         st_),
          "chain");
    check(hipMemcpyAsync(chainTotalPos_.get(), &m, sizeof(int64_t), hipMemcpyHostToDevice, st_), "chain");
    check(hipStreamSynchronize(st_), "chain");   // the host vectors above are released on return
    chainMStride_ = m;
    chainTotalRows_ = rows;
    chainKeyCapMax_ = std::max(capMax, 1);
    chainMeanKeys_ = samples > 0 ? keySamples / samples : 0;
    chainEpoch_ = countsEpoch_;
    chainReady_ = true;
}

template <typename T>
bool Engine<T>::chainUsable(int itemSize) const {
    if (!chainReady_ || chainEpoch_ != countsEpoch_ || !options.chain || options.tolerance || segmentedNow_)
        return false;
    // items of itemSize samples from the starts of their segments (the count tables' columns);
    // every block writes and the prefix pass rewrites a count per key of its segment, so the
    // path pays only while a block holds more samples than its segment has keys (on average,
    // weighted by samples): many keys per block make the general sort the cheaper ordering
    return itemsPer_ == itemSize && chain_iteration_supported(dim_, itemSize, chainKeyCapMax_) &&
           (options.chain == 2 || chainMeanKeys_ <= itemSize);
}

// rd_chain.hip's prefix and scatter after a histogram scan into labels.  Chains of at least
// 2^kLongMinShift members are listed for the wave-summed sums, up to kLongWaves waves.
template <typename T>
void Engine<T>::enqueueChainOrder(int itemSize, const int* labels) {
    const EvolveSource es = evolveSource();
    const ChainArgs<T> ca{es.P, es.stride, es.index, dItems_.get(), compact() ? flags_.get() + 4 : nullptr,
                          dSegs_.get(), dSegK_.get(), dSegDone_.get(), (int)segs_.size(), labels, chainHist_.get(),
                          chainSegHistOff_.get(), chainSegKeyCap_.get(), chainSegRowOff_.get(), itemSize,
                          chainTotalRows_, chainTotals_.get(), sortedP_.get(), chainMStride_,
                          chainSegMemberOff_.get(), flags_.get(), chainOff_.get(), chainEnd_.get(), longList_.get(),
                          chainLongCount_.get(), chainLongIdx_.get(), 1 << kLongMinShift, kLongWaves / dim_};
    check(launch_chain_order<T>(ca, dim_, numItems_, chainKeyCapMax_, st_), "chain order");
}

template <typename T>
void Engine<T>::fillSentinels() {
    launch_fill_sentinel<T>(S_.get(), sStride_, dim_, dSegs_.get(), dSegK_.get(), (int)segs_.size(), st_);
    check(hipGetLastError(), "fill_sentinel");
}

template <typename T>
void Engine<T>::compactSlots() {
    const int ns = (int)segs_.size();
    if (ns == 0) return;
    ++version_;
    std::vector<int> K = countsHost();
    std::vector<int64_t> oldOff(ns), newOff(ns);
    std::vector<int> copyCnt(ns);
    int64_t off = 0;
    for (int s = 0; s < ns; ++s) {
        oldOff[s] = segs_[s].sOff;
        newOff[s] = off;
        const int64_t cap = round_up(std::max<int64_t>(K[s], 1), kChosenTile);
        copyCnt[s] = (int)std::min<int64_t>(cap, segs_[s].sCap);   // live points + sentinel tail
        segs_[s].sOff = off;
        segs_[s].sCap = (int32_t)cap;
        off += cap;
    }
    if (off == totalSlots_) return;
    const int64_t newStride = round_up(std::max<int64_t>(off, 1), 64);
    DevBuf<T> S2;
    DevBuf<int> idx2;
    S2.reserve((size_t)newStride * dim_);
    idx2.reserve((size_t)std::max<int64_t>(off, 1));
    DevBuf<int64_t> dOld, dNew;
    DevBuf<int> dCnt;
    dOld.reserve(ns);
    dNew.reserve(ns);
    dCnt.reserve(ns);
    check(hipMemcpyAsync(dOld.get(), oldOff.data(), sizeof(int64_t) * ns, hipMemcpyHostToDevice, st_), "relayout");
    check(hipMemcpyAsync(dNew.get(), newOff.data(), sizeof(int64_t) * ns, hipMemcpyHostToDevice, st_), "relayout");
    check(hipMemcpyAsync(dCnt.get(), copyCnt.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "relayout");
    hipLaunchKernelGGL((relayout_slots_kernel<T>), dim3(ns), dim3(256), 0, st_, S_.get(), sStride_, chosenIdx_.get(),
                       dOld.get(), dNew.get(), dCnt.get(), dim_, S2.get(), newStride, idx2.get());
    check(hipGetLastError(), "relayout");
    check(hipStreamSynchronize(st_), "relayout");
    S_.swap(S2);
    chosenIdx_.swap(idx2);
    totalSlots_ = off;
    sStride_ = newStride;
    std::vector<int> slotSeg((size_t)totalSlots_);
    for (int s = 0; s < ns; ++s)
        std::fill(slotSeg.begin() + segs_[s].sOff, slotSeg.begin() + segs_[s].sOff + segs_[s].sCap, s);
    check(hipMemcpyAsync(dSegs_.get(), segs_.data(), sizeof(SegDesc) * ns, hipMemcpyHostToDevice, st_), "segs");
    check(hipMemcpyAsync(dSlotSeg_.get(), slotSeg.data(), sizeof(int) * totalSlots_, hipMemcpyHostToDevice, st_),
          "slots");
    sums_.reserve((size_t)sStride_ * dim_);
    xTable_.reserve((size_t)(dim_ + 1) * sStride_);
    check(hipStreamSynchronize(st_), "relayout");
    updateKeyBound();
}

template <typename T>
void Engine<T>::updateKeyBound() {
    std::vector<int> K = countsHost();
    int64_t kb = 0;
    for (size_t s = 0; s < K.size(); ++s)
        if (K[s] > 0) kb = std::max<int64_t>(kb, segs_[s].sOff + K[s]);
    keyBound_ = kb;
    ++countsEpoch_;
}

// ------------------------------------------------------------------------------------
// choose
// ------------------------------------------------------------------------------------
template <typename T>
void Engine<T>::choose() {
    const int ns = (int)segs_.size();
    if (ns == 0) return;
    preparePositions();
    const EvolveSource es = evolveSource();
    ++version_;
    chosenFromPoints_ = true;
    auto t0 = std::chrono::steady_clock::now();
    constexpr int kLarge = 32768;
    std::vector<int> large(ns);
    int maxOwn = 0;
    bool anyLarge = false;
    for (int s = 0; s < ns; ++s) {
        large[s] = segs_[s].nOwn > kLarge ? 1 : 0;
        anyLarge |= large[s] != 0;
        if (large[s]) maxOwn = std::max(maxOwn, segs_[s].nOwn);
    }
    DevBuf<int> dLarge;
    dLarge.reserve(ns);
    check(hipMemcpyAsync(dLarge.get(), large.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "large");
    int zero = 0;
    check(hipMemcpyAsync(flags_.get() + 2, &zero, sizeof(int), hipMemcpyHostToDevice, st_), "overflow");
    ChooseArgs<T> a{es.P, es.stride, es.index, S_.get(), sStride_, dSegs_.get(), ns, dSegK_.get(), dSegR1Sqr_.get(),
                    covered_.get(), chosenIdx_.get(), flags_.get() + 2, dLarge.get()};
    // small segments: one block each, whole segment
    check(launch_choose_block<T>(a, dim_, -1, 0, st_), "choose_block");
    if (anyLarge) {
        // chunk 0 sequentially, then filter + resolve per chunk
        check(launch_choose_block<T>(a, dim_, 0, 0, st_), "choose_block");
        std::vector<DevBuf<WorkItem>> bufs;
        for (int c = 1; choose_chunk_start(c) < maxOwn; ++c) {
            std::vector<WorkItem> items;
            const int c0 = choose_chunk_start(c), c1 = choose_chunk_start(c + 1);
            for (int s = 0; s < ns; ++s) {
                if (!large[s]) continue;
                const int e = std::min(c1, segs_[s].nOwn);
                for (int st = c0; st < e; st += 1024) items.push_back({s, st, std::min(e, st + 1024), 0});
            }
            if (items.empty()) continue;
            DevBuf<WorkItem> buf;
            buf.reserve(items.size());
            check(hipMemcpyAsync(buf.get(), items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice, st_),
                  "items");
            check(launch_choose_filter<T>(a, dim_, buf.get(), (int)items.size(), st_), "choose_filter");
            check(launch_choose_block<T>(a, dim_, c, 1, st_), "choose_block");
            check(hipStreamSynchronize(st_), "choose chunk");
        }
    }
    fillSentinels();
    int ovf = 0;
    check(hipMemcpyAsync(&ovf, flags_.get() + 2, sizeof(int), hipMemcpyDeviceToHost, st_), "overflow");
    check(hipStreamSynchronize(st_), "choose");
    if (ovf) throw std::runtime_error("choose: chosen points exceed a segment's capacity");
    updateKeyBound();
    timers.chooseMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------------------------
// evolve
// ------------------------------------------------------------------------------------
template <typename T>
void Engine<T>::enqueueIteration(bool record, int slot, const IterLog* log, bool finalize, int capIter) {
    const EvolveSource es = evolveSource();
    const bool ev = options.profile && slot >= 0;
    if (ev) {
        while ((int)events_.size() < 3 * (slot + 1)) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            events_.push_back(e);
        }
        check(hipEventRecord(events_[3 * slot], st_), "event");
    }
    if (segmentedNow_ && !log && !record) {
        // batch-wide scan into labels_, then one workgroup per segment sorts and sums
        ScanArgs<T> sa{es.P, es.stride, es.index, S_.get(), sStride_, dItems_.get(), dSegs_.get(),
                       dSegK_.get(), dSegDone_.get(), dSegR1Sqr_.get(), flags_.get(), labels_.get(), flags_.get() + 3,
                       nullptr, nullptr, nullptr};
        prepScreen(sa, screenForm());
        check(launch_evolve_scan<T>(sa, dim_, numItems_, options.ppt, options.scanThreads, 0, st_), "evolve_scan");
        if (ev) check(hipEventRecord(events_[3 * slot + 1], st_), "event");
        LocalArgs<T> la{es.P, es.stride, es.index, S_.get(), sStride_, dSegs_.get(), dSegK_.get(),
                        dSegR1Sqr_.get(), nullptr, (int)segs_.size(), segKStride_, segChunk_, 0, 0, nullptr, nullptr,
                        nullptr, labels_.get(), dSegDone_.get(), dSegMoved_.get(), flags_.get()};
        check(launch_segment_sums<T>(la, dim_, (int)segs_.size(), st_), "segment_sums");
        if (finalize)
            check(launch_evolve_finalize((int)segs_.size(), dSegDone_.get(), dSegMoved_.get(), dSegIter_.get(),
                                         capIter, flags_.get(), flags_.get() + 1, st_),
                  "finalize");
        if (ev) check(hipEventRecord(events_[3 * slot + 2], st_), "event");
        return;
    }
    int* labels = log ? log->labels : labels_.get();
    const int itemSize = std::max(64, options.scanThreads) * std::max(1, options.ppt);
    timers.chainPath = chainUsable(itemSize) ? 1 : 0;
    if (timers.chainPath) {
        // label histogram in the scan (one block per item in use), rd_chain.hip's prefix and
        // scatter, then the ordered sums over the chain layout
        ScanArgs<T> sa{es.P, es.stride, es.index, S_.get(), sStride_, dItems_.get(), dSegs_.get(), dSegK_.get(),
                       dSegDone_.get(), dSegR1Sqr_.get(), flags_.get(), labels, flags_.get() + 3,
                       nullptr, nullptr, compact() ? flags_.get() + 4 : nullptr, nullptr, chainHist_.get(),
                       chainSegHistOff_.get(), chainSegKeyCap_.get(), itemSize};
        prepScreen(sa, screenForm());
        check(launch_evolve_scan<T>(sa, dim_, numItems_, options.ppt, options.scanThreads, 0, st_), "evolve_scan");
        if (ev) check(hipEventRecord(events_[3 * slot + 1], st_), "event");
        enqueueChainOrder(itemSize, labels);
        // rows over the live slots of the evolve call (batches), else every slot below the bound
        const bool rowsValid = rowSlotsVersion_ == version_ && numRowSlots_ > 0;
        ShiftArgs<T> ha{es.P, es.stride, es.index, S_.get(), sStride_,
                        rowsValid ? numRowSlots_ : std::max<int64_t>(keyBound_, 1), dSlotSeg_.get(), dSegs_.get(),
                        dSegK_.get(), dSegDone_.get(), dSegMoved_.get(), nullptr, chainTotalPos_.get(),
                        sortedP_.get(), chainMStride_, chainOff_.get(), chainEnd_.get(), flags_.get(),
                        log ? log->counts : (record ? counts_.get() : nullptr), record ? sums_.get() : nullptr,
                        log ? log->moved : (record ? moved_.get() : nullptr),
                        log ? log->movedStride : totalSlots_, longList_.get(), 0,
                        rowsValid ? rowSlots_.get() : nullptr};
        ha.longCount = chainLongCount_.get();
        ha.longIdx = chainLongIdx_.get();
        ha.longMin = 1 << kLongMinShift;
        ha.longCap = kLongWaves / dim_;
        // the convergence bookkeeping in a launch of its own, not in the sum launch's last
        // block: that form's device-scope fence and completion count in each of the launch's
        // ~2000 blocks cost ~100 us per iteration of a 322k-point cloud against ~5 us here
        check(launch_chain_sum_shift<T>(ha, dim_, chainMStride_, false, st_), "chain_sum_shift");
        if (finalize)
            check(launch_evolve_finalize((int)segs_.size(), dSegDone_.get(), dSegMoved_.get(), dSegIter_.get(),
                                         capIter, flags_.get(), flags_.get() + 1, st_),
                  "finalize");
        if (ev) check(hipEventRecord(events_[3 * slot + 2], st_), "event");
        return;
    }
    const bool rowsValid = rowSlotsVersion_ == version_ && numRowSlots_ > 0;
    const bool dense = rowsValid && denseKeys_ && compact() && !log && !record;
    ScanArgs<T> sa{es.P, es.stride, es.index, S_.get(), sStride_, dItems_.get(), dSegs_.get(), dSegK_.get(),
                   dSegDone_.get(), dSegR1Sqr_.get(), flags_.get(), labels, flags_.get() + 3,
                   compact() ? cKeys_.get() : nullptr, compact() ? cPos_.get() : nullptr,
                   compact() ? flags_.get() + 4 : nullptr, dense ? denseOf_.get() : nullptr};
    prepScreen(sa, screenForm());
    // one block per work item (batches: the active-item list, or persistent blocks on its
    // queue when that form is selected)
    check(launch_evolve_scan<T>(sa, dim_, numItems_, options.ppt, options.scanThreads,
                                compact() ? persistentBlocks() : 0, st_),
          "evolve_scan");
    if (ev) check(hipEventRecord(events_[3 * slot + 1], st_), "event");
    // positions the sort and the member passes cover: the dense active range of a batch
    // (it only shrinks; graphs are captured again when it does), else every position
    const int64_t totalP = compact() ? std::max<int64_t>(activeWork_, 1) : totalP_;
    // labels are live slots < keyBound_ (counts only shrink between refreshes), or in a
    // batch their indices in the live-slot list
    const int64_t kb = dense ? numRowSlots_ : std::max<int64_t>(keyBound_, 1);
    const int longShift = enqueueSortOrder(labels, totalP, kb);
    ShiftArgs<T> ha{es.P, es.stride, es.index, S_.get(), sStride_, kb, dSlotSeg_.get(), dSegs_.get(),
                    dSegK_.get(), dSegDone_.get(), dSegMoved_.get(), members_.get(), dM_.get(), sortedP_.get(),
                    totalP_, chainOff_.get(), chainEnd_.get(), flags_.get(),
                    log ? log->counts : (record ? counts_.get() : nullptr), record ? sums_.get() : nullptr,
                    log ? log->moved : (record ? moved_.get() : nullptr),
                    log ? log->movedStride : totalSlots_, longList_.get(), longShift, nullptr};
    ha.tolerance = options.tolerance ? 1 : 0;
    if (rowsValid) {
        // rows over the live slots of the evolve call only
        ha.totalSlots = numRowSlots_;
        ha.rowSlots = rowSlots_.get();
        ha.denseChains = dense ? 1 : 0;
    }
    // the convergence bookkeeping in a launch of its own (as above)
    check(launch_chain_sum_shift<T>(ha, dim_, totalP, false, st_), "chain_sum_shift");
    if (finalize)
        check(launch_evolve_finalize((int)segs_.size(), dSegDone_.get(), dSegMoved_.get(), dSegIter_.get(), capIter,
                                     flags_.get(), flags_.get() + 1, st_),
              "finalize");
    if (ev) check(hipEventRecord(events_[3 * slot + 2], st_), "event");
}

template <typename T>
void Engine<T>::prepScreen(ScanArgs<T>& sa, int form) {
    sa.xTable = nullptr;
    sa.segX = nullptr;
    sa.mfma = form == 2 ? 1 : 0;
    timers.screenForm = form;
    if (form == 0) return;
    check(launch_expand_chosen<T>(S_.get(), sStride_, dSegs_.get(), dSegK_.get(), dSegDone_.get(), (int)segs_.size(),
                                  dim_, xTable_.get(), segX_.get(), flags_.get(), st_),
          "expand_chosen");
    sa.xTable = xTable_.get();
    sa.segX = segX_.get();
}

template <typename T>
int Engine<T>::enqueueSortOrder(const int* labels, int64_t totalP, int64_t kb) {
    const EvolveSource es = evolveSource();
    // members sorted by (label, position); the last pass writes their coordinates in
    // chain order straight into sortedP_
    const SortGather<T> sg{es.P, es.stride, es.index, sortedP_.get(), totalP_, dim_};
    check(stable_sort_by_key<T>(compact() ? cKeys_.get() : labels, compact() ? cPos_.get() : nullptr, dN_.get(),
                                totalP, (int)std::max<int64_t>(kb, 2), keys_.get(), members_.get(), dM_.get(),
                                sortWs_.get(), st_, flags_.get(), sg),
          "sort");
    const int longShift = long_chain_shift(dim_, totalP);
    const int sortKeys = (int)std::max<int64_t>(kb, 2);
    if (sort_single_pass(sortKeys)) {
        // every chosen slot's member run straight from the one-pass sort's scanned histogram
        check(launch_key_ranges_from_sort(sortWs_.get(), totalP, sortKeys, dM_.get(), chainOff_.get(),
                                          chainEnd_.get(), longList_.get(), longShift, flags_.get(), st_),
              "key ranges");
    } else {
        hipLaunchKernelGGL(zero_pair_kernel, dim3(grid_for(kb)), dim3(256), 0, st_, chainOff_.get(), chainEnd_.get(),
                           kb, flags_.get());
        check(launch_chain_bounds(keys_.get(), dM_.get(), totalP, chainOff_.get(), chainEnd_.get(), longList_.get(),
                                  longShift, flags_.get(), st_),
              "chain_bounds");
    }
    return longShift;
}

template <typename T>
void Engine<T>::readEvents(int used) {
    for (int b = 0; b < used; ++b) {
        float a = 0, w = 0;
        check(hipEventElapsedTime(&a, events_[3 * b], events_[3 * b + 1]), "elapsed");
        check(hipEventElapsedTime(&w, events_[3 * b], events_[3 * b + 2]), "elapsed");
        timers.scanKernelMs += a;
        timers.iterationMs += w;
        timers.scanLaunches += 1;
    }
}

namespace {
std::mutex g_tuneMutex;
struct ScanShape {
    int ppt;
    int threads;
    int form;   // screen form (Engine::screenForm)
};
std::map<std::tuple<int, int, int, int, int, int>, ScanShape> g_tuned;
struct BatchShape {
    int ppt;
    int threads;
    bool persistent;
    int form;
};
std::map<std::tuple<int, int, int, int, int, int>, BatchShape> g_tunedBatch;
inline int ilog2(int64_t v) {
    int l = 0;
    while (v > 1) {
        v >>= 1;
        ++l;
    }
    return l;
}
}  // namespace

// Batches of segments: the compact scan over every segment's work items, timed for each
// (samples per lane x threads per block x persistent queue or one block per item) shape,
// three rounds, best round kept; cached per (value type, dimension, log2 samples, log2 chosen
// points, log2 segments).  Items then follow the picked shape (its size is also the
// granularity of the active-item list).
template <typename T>
void Engine<T>::autotuneBatchScan() {
    auto t0 = std::chrono::steady_clock::now();
    const std::vector<int> K = countsHost();
    int64_t kTotal = 0;
    for (int k : K) kTotal += k;
    const int ns = (int)segs_.size();
    const auto key = std::make_tuple((int)sizeof(T), dim_, ilog2(std::max<int64_t>(totalP_, 1)), 0, 1 + ilog2(ns),
                                     options.screen);
    const std::vector<int> forms = screenForms();
    BatchShape best{options.ppt, options.scanThreads, options.persistent, forms[0]};
    bool cached = false;
    {
        std::lock_guard<std::mutex> g(g_tuneMutex);
        auto it = g_tunedBatch.find(key);
        if (it != g_tunedBatch.end()) {
            best = it->second;
            cached = true;
        }
    }
    if (!cached) {
        // shapes in the first screen form allowed (the direct one where the tuner picks the
        // form), then the best shape in each other form allowed
        static const BatchShape shapes[] = {{4, 256, true, 0},  {2, 256, true, 0},  {1, 256, true, 0},
                                            {4, 256, false, 0}, {2, 256, false, 0}, {1, 256, false, 0},
                                            {2, 512, false, 0}, {1, 512, false, 0}, {4, 128, false, 0}};
        constexpr int kShapes = (int)(sizeof(shapes) / sizeof(shapes[0]));
        BatchShape cands[kShapes + 2];
        for (int c = 0; c < kShapes; ++c) {
            cands[c] = shapes[c];
            cands[c].form = forms[0];
        }
        const int kCands = kShapes + (int)forms.size() - 1;
        const EvolveSource es = evolveSource();
        const std::vector<int> none(ns, 0);
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        float score[kShapes + 2];
        for (int c = 0; c < kCands; ++c) score[c] = 1e30f;
        // two rounds of one untimed and one timed launch per shape (one round when a launch
        // takes more than 0.2 ms, long enough to time alone: the tuning then costs about 20 scans
        // of the batch)
        int rounds = 2;
        for (int round = 0; round < rounds; ++round) {
            for (int c = 0; c < kCands; ++c) {
                if (c >= kShapes) {
                    // another form at the best shape of the first round (timed in that round only)
                    if (round > 0) continue;
                    int b = 0;
                    for (int i = 1; i < kShapes; ++i)
                        if (score[i] < score[b]) b = i;
                    cands[c] = cands[b];
                    cands[c].form = forms[c - kShapes + 1];
                }
                options.ppt = cands[c].ppt;
                options.scanThreads = cands[c].threads;
                options.persistent = cands[c].persistent;
                buildWorkItems();
                orderItemsByCost();
                rebuildActive(none);
                ScanArgs<T> sa{es.P, es.stride, es.index, S_.get(), sStride_, dItems_.get(), dSegs_.get(), dSegK_.get(),
                               dSegDone_.get(), dSegR1Sqr_.get(), flags_.get(), labels_.get(), flags_.get() + 3,
                               cKeys_.get(), cPos_.get(), flags_.get() + 4, nullptr};
                const int pb = persistentBlocks();
                prepScreen(sa, cands[c].form);
                check(launch_evolve_scan<T>(sa, dim_, numItems_, options.ppt, options.scanThreads, pb, st_), "tune");
                check(hipEventRecord(e0, st_), "event");
                prepScreen(sa, cands[c].form);
                check(launch_evolve_scan<T>(sa, dim_, numItems_, options.ppt, options.scanThreads, pb, st_), "tune");
                check(hipEventRecord(e1, st_), "event");
                check(hipEventSynchronize(e1), "event");
                float ms = 0;
                check(hipEventElapsedTime(&ms, e0, e1), "event");
                score[c] = std::min(score[c], ms);
                if (round == 0 && c == 0 && ms > 0.2f) rounds = 1;
            }
        }
        float bestMs = 1e30f;
        for (int c = 0; c < kCands; ++c)
            if (score[c] < bestMs) {
                bestMs = score[c];
                best = cands[c];
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::lock_guard<std::mutex> g(g_tuneMutex);
        g_tunedBatch[key] = best;
    }
    if (!cached || options.ppt != best.ppt || options.scanThreads != best.threads ||
        options.persistent != best.persistent) {
        options.ppt = best.ppt;
        options.scanThreads = best.threads;
        options.persistent = best.persistent;
        buildWorkItems();
        orderItemsByCost();
        check(hipStreamSynchronize(st_), "autotune");
    }
    if (options.screen == 2) tunedForm_ = best.form;
    timers.scanPpt = options.ppt;
    timers.scanThreads = options.scanThreads;
    timers.scanPersistent = options.persistent ? 1 : 0;
    timers.tuneMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

template <typename T>
void Engine<T>::autotuneScan() {
    ensureChainScratch();
    const EvolveSource es = evolveSource();
    if (options.autotune && segs_.size() > 1) {
        autotuneBatchScan();
        return;
    }
    if (!options.autotune || segs_.size() != 1) {
        timers.scanPpt = options.ppt;
        timers.scanThreads = options.scanThreads;
        timers.scanPersistent = options.persistent ? 1 : 0;
        return;
    }
    auto t0 = std::chrono::steady_clock::now();
    const std::vector<int> K = countsHost();
    int64_t kTotal = 0;
    for (int k : K) kTotal += k;
    const auto key = std::make_tuple((int)sizeof(T), dim_, ilog2(std::max<int64_t>(totalP_, 1)),
                                     ilog2(kTotal + 1), options.chain != 0 && !options.tolerance ? -1 : 0,
                                     options.screen);
    const std::vector<int> forms = screenForms();
    ScanShape best{options.ppt, options.scanThreads, forms[0]};
    bool cached = false;
    {
        std::lock_guard<std::mutex> g(g_tuneMutex);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            best = it->second;
            cached = true;
        }
    }
    if (!cached) {
        static const int shapes[][2] = {{4, 256}, {2, 256}, {1, 256}, {4, 128}, {2, 128}, {1, 128},
                                        {4, 512}, {2, 512}, {1, 512}};
        constexpr int kShapes = (int)(sizeof(shapes) / sizeof(shapes[0]));
        // every shape in every screen form allowed
        ScanShape cands[3 * kShapes];
        int kCands = 0;
        for (int f : forms)
            for (int c = 0; c < kShapes; ++c) cands[kCands++] = {shapes[c][0], shapes[c][1], f};
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        ScanArgs<T> sa{es.P, es.stride, es.index, S_.get(), sStride_, dItems_.get(), dSegs_.get(), dSegK_.get(),
                       dSegDone_.get(), dSegR1Sqr_.get(), flags_.get(), labels_.get(), flags_.get() + 3,
                       nullptr, nullptr, nullptr};
        // three rounds over all candidates (so that no shape always runs first), each
        // launch pair timed after one untimed launch; a shape scores its best round
        float score[3 * kShapes];
        for (int c = 0; c < kCands; ++c) score[c] = 1e30f;
        for (int round = 0; round < 3; ++round) {
            for (int c = 0; c < kCands; ++c) {
                options.ppt = cands[c].ppt;
                options.scanThreads = cands[c].threads;
                buildWorkItems();
                sa.items = dItems_.get();
                // the scan and the member ordering the iterations will run with this shape: the
                // label histogram and rd_chain.hip's scatter (whose cost depends on the item
                // size) where that path is usable, else the sort; the ordered sums after either
                // cost about the same
                const int itemSize = cands[c].ppt * cands[c].threads;
                const bool hist = chainUsable(itemSize);
                sa.blockHist = hist ? chainHist_.get() : nullptr;
                sa.histSegOff = hist ? chainSegHistOff_.get() : nullptr;
                sa.histKeyCap = hist ? chainSegKeyCap_.get() : nullptr;
                sa.histItem = hist ? itemSize : 0;
                auto pass = [&] {
                    prepScreen(sa, cands[c].form);
                    check(launch_evolve_scan<T>(sa, dim_, numItems_, cands[c].ppt, cands[c].threads, 0, st_), "tune");
                    if (hist)
                        enqueueChainOrder(itemSize, labels_.get());
                    else
                        (void)enqueueSortOrder(labels_.get(), totalP_, std::max<int64_t>(keyBound_, 1));
                };
                pass();
                check(hipEventRecord(e0, st_), "event");
                for (int r = 0; r < 2; ++r) pass();
                check(hipEventRecord(e1, st_), "event");
                check(hipEventSynchronize(e1), "event");
                float ms = 0;
                check(hipEventElapsedTime(&ms, e0, e1), "event");
                score[c] = std::min(score[c], ms);
            }
        }
        float bestMs = 1e30f;
        for (int c = 0; c < kCands; ++c)
            if (score[c] < bestMs) {
                bestMs = score[c];
                best = cands[c];
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::lock_guard<std::mutex> g(g_tuneMutex);
        g_tuned[key] = best;
    }
    if (options.ppt != best.ppt || options.scanThreads != best.threads || !cached) {
        options.ppt = best.ppt;
        options.scanThreads = best.threads;
        buildWorkItems();
        if (segs_.size() > 1) orderItemsByCost();
        check(hipStreamSynchronize(st_), "autotune");
    }
    if (options.screen == 2) tunedForm_ = best.form;
    timers.scanPpt = options.ppt;
    timers.scanThreads = options.scanThreads;
    timers.tuneMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

template <typename T>
void Engine<T>::addEvolveBytes(const std::vector<int>& itersPerSeg) {
    std::vector<int> K = countsHost();
    double b = 0, pairs = 0;
    for (size_t s = 0; s < segs_.size(); ++s) {
        b += (double)itersPerSeg[s] *
             ((double)segs_[s].nP * dim_ * sizeof(T) + (double)K[s] * (3.0 * dim_ * sizeof(T) + 4.0));
        pairs += (double)itersPerSeg[s] * (double)segs_[s].nP * (double)K[s];
    }
    timers.evolveBytes += b;
    timers.evolvePairs += pairs;
}

template <typename T>
void Engine<T>::evolveFixed(long long iters) {
    const int ns = (int)segs_.size();
    if (ns == 0 || iters <= 0) return;
    ++version_;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int> zeros(ns, 0);
    int f[2] = {0, 0};
    check(hipMemcpyAsync(dSegDone_.get(), zeros.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "done");
    check(hipMemcpyAsync(flags_.get(), f, sizeof(int) * 2, hipMemcpyHostToDevice, st_), "flags");
    autotuneScan();
    if (compact()) rebuildActive(zeros);
    const int batch = std::max(1, options.itersPerSync);
    long long done = 0;
    hipGraphExec_t exec = nullptr;
    const bool useGraph = options.graph && !options.profile && iters >= batch &&
                          captureIterations(batch, false, 0, &exec);
    // no convergence check between batches: they are all enqueued and the stream is synchronised
    // once (per batch only when profiling, whose events are read back per batch)
    while (done < iters) {
        const int nb = (int)std::min<long long>(batch, iters - done);
        if (useGraph && nb == batch)
            check(hipGraphLaunch(exec, st_), "hipGraphLaunch");
        else
            for (int b = 0; b < nb; ++b) enqueueIteration(false, b);
        if (options.profile) {
            check(hipStreamSynchronize(st_), "evolveFixed");
            readEvents(nb);
        }
        done += nb;
    }
    check(hipStreamSynchronize(st_), "evolveFixed");
    addEvolveBytes(std::vector<int>(ns, (int)std::min<long long>(iters, 0x7fffffff)));
    timers.evolveIterations += iters;
    timers.evolveMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Captures `batch` iterations into a graph.  The engine keeps one executable graph: a new
// capture of the same shape only updates its kernel arguments (hipGraphExecUpdate), so repeated
// evolve calls do not pay for an instantiation each; another shape replaces it.  Every caller
// has synchronised the stream after its last launch, so the graph is idle when updated.
template <typename T>
bool Engine<T>::captureIterations(int batch, bool finalize, int capIter, hipGraphExec_t* exec) {
    *exec = nullptr;
    (void)persistentBlocks();   // device queries stay outside the capture
    if (hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    bool ok = true;
    try {
        for (int b = 0; b < batch; ++b) enqueueIteration(false, -1, nullptr, finalize, capIter);
    } catch (const HipError&) {
        ok = false;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st_, &g);
    if (!ok || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return false;
    }
    if (graphExec_) {
        hipGraphNode_t errNode = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(graphExec_, g, &errNode, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            (void)hipGraphDestroy(g);
            *exec = graphExec_;
            return true;
        }
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(graphExec_);
        graphExec_ = nullptr;
    }
    e = hipGraphInstantiate(&graphExec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        graphExec_ = nullptr;
        (void)hipGetLastError();
        return false;
    }
    *exec = graphExec_;
    return true;
}

template <typename T>
bool Engine<T>::chooseSegmented(const std::vector<int>& done) {
    if (!options.segmented || !compact()) return false;
    const std::vector<int> K = countsHost();
    int maxK = 0, maxP = 0;
    for (size_t s = 0; s < segs_.size(); ++s) {
        if (done[s]) continue;
        if (segs_[s].nP > options.segmentedMaxP || K[s] > kLocalMaxK) return false;
        maxK = std::max(maxK, K[s]);
        maxP = std::max(maxP, (int)segs_[s].nP);
    }
    segKStride_ = (int)round_up(std::max(maxK, 1), kChosenTile);
    segChunk_ = std::min(std::max(maxP, 1), kLocalChunk);
    return local_lds_bytes(dim_, sizeof(T), segKStride_, segChunk_) <= kLocalLdsBudget;
}

template <typename T>
int Engine<T>::evolveLocal(long long maxIter, std::vector<int>& done) {
    const int ns = (int)segs_.size();
    const EvolveSource es = evolveSource();
    std::vector<int> K = countsHost();
    // Taking every active segment off the batch-wide iterations removes their fixed
    // per-iteration cost entirely; taking only some keeps it, so then only the smallest go.
    auto eligible = [&](int s, int maxP) {
        return segs_[s].nP <= std::min(kLocalMaxP, maxP) && K[s] <= kLocalMaxK &&
               (int64_t)segs_[s].nP * K[s] <= options.localMaxWork;
    };
    bool all = true;
    for (int s = 0; s < ns; ++s) all = all && (done[s] || eligible(s, options.localMaxP));
    const int limitP = all ? options.localMaxP : std::min(options.localMaxP, options.localSmallP);
    std::vector<int> cand;
    for (int s = 0; s < ns; ++s)
        if (!done[s] && eligible(s, limitP)) cand.push_back(s);
    if (cand.empty()) return 0;
    // the group shares one LDS layout sized by its largest K (and the chunk, bounded by the
    // largest sample count): take segments by increasing K while the layout fits
    std::stable_sort(cand.begin(), cand.end(), [&](int x, int y) { return K[x] < K[y]; });
    std::vector<int> list;
    int maxK = 0, maxP = 0;
    for (int s : cand) {
        const int k2 = std::max(maxK, K[s]), p2 = std::max(maxP, (int)segs_[s].nP);
        const int ks = (int)round_up(std::max(k2, 1), kChosenTile);
        if (local_lds_bytes(dim_, sizeof(T), ks, std::min(std::max(p2, 1), kLocalChunk)) > kLocalLdsBudget) break;
        maxK = k2;
        maxP = p2;
        list.push_back(s);
    }
    if (list.empty()) return 0;
    // longest first: cost ~ samples x chosen points
    std::stable_sort(list.begin(), list.end(), [&](int x, int y) {
        return (int64_t)segs_[x].nP * std::max(K[x], 1) > (int64_t)segs_[y].nP * std::max(K[y], 1);
    });
    dLocalList_.reserve(list.size());
    check(hipMemcpyAsync(dLocalList_.get(), list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, st_),
          "local list");
    int zero2[2] = {0, 0};
    check(hipMemcpyAsync(flags_.get() + 5, zero2, sizeof(int) * 2, hipMemcpyHostToDevice, st_), "local flags");
    LocalArgs<T> a{es.P, es.stride, es.index, S_.get(), sStride_, dSegs_.get(), dSegK_.get(), dSegR1Sqr_.get(),
                   dLocalList_.get(), (int)list.size(), (int)round_up(std::max(maxK, 1), kChosenTile),
                   std::min(std::max(maxP, 1), kLocalChunk), maxIter > 0 ? (int)std::min<long long>(maxIter, 0x7fffffff) : 0,
                   options.evolveGuard > 0 ? (int)std::min<long long>(options.evolveGuard, 0x7fffffff) : 0x7fffffff,
                   dSegIter_.get(), flags_.get() + 5, flags_.get() + 6,
                   nullptr, nullptr, nullptr, nullptr};
    check(launch_local_evolve<T>(a, dim_, st_), "local_evolve");
    int res[2] = {0, 0};
    check(hipMemcpyAsync(res, flags_.get() + 5, sizeof(int) * 2, hipMemcpyDeviceToHost, st_), "local flags");
    check(hipStreamSynchronize(st_), "local_evolve");
    if (res[1]) throw std::runtime_error(guardMessage());
    for (int s : list) done[s] = 1;
    return res[0];
}

template <typename T>
long long Engine<T>::evolve(long long maxIter, const std::vector<char>& active) {
    const int ns = (int)segs_.size();
    if (ns == 0) return 0;
    preparePositions();
    ++version_;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int> done(ns, 0), zeros(ns, 0);
    int anyActive = 0;
    for (int s = 0; s < ns; ++s) {
        done[s] = (!active.empty() && !active[s]) ? 1 : 0;
        anyActive |= !done[s];
    }
    if (!anyActive) return 0;
    int f[2] = {0, 0};
    check(hipMemcpyAsync(dSegIter_.get(), zeros.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "iter");
    // small segments iterate inside one workgroup each; the rest share batch-wide iterations
    const int localIters = options.local ? evolveLocal(maxIter, done) : 0;
    bool anyGlobal = false;
    for (int s = 0; s < ns; ++s) anyGlobal |= !done[s];
    if (!anyGlobal) {
        std::vector<int> segIter(ns);
        check(hipMemcpy(segIter.data(), dSegIter_.get(), sizeof(int) * ns, hipMemcpyDeviceToHost), "iterations");
        addEvolveBytes(segIter);
        lastEvolveIters = localIters;
        timers.evolveIterations += localIters;
        timers.evolveMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return localIters;
    }
    check(hipMemcpyAsync(dSegDone_.get(), done.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "done");
    check(hipMemcpyAsync(dSegMoved_.get(), zeros.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "moved");
    check(hipMemcpyAsync(flags_.get(), f, sizeof(int) * 2, hipMemcpyHostToDevice, st_), "flags");
    autotuneScan();
    segmentedNow_ = chooseSegmented(done);
    if (segmentedNow_) {
        // the scan runs over every item (done segments' items return at once)
        check(hipMemcpyAsync(dItems_.get(), items_.data(), sizeof(WorkItem) * items_.size(), hipMemcpyHostToDevice, st_),
              "items");
    } else if (compact()) {
        rebuildActive(done);
    }
    prepareRows(done);
    const int capIter = maxIter > 0 ? (int)std::min<long long>(maxIter, 0x7fffffff) : 0;
    long long iters = 0;
    const int batch = std::max(1, options.itersPerSync);
    std::vector<int> doneNow(ns);
    // the iterations after convergence in a batch return at once (device all-done flag)
    hipGraphExec_t exec = nullptr;
    bool useGraph = options.graph && !options.profile && captureIterations(batch, true, capIter, &exec);
    for (;;) {
        if (useGraph) {
            check(hipGraphLaunch(exec, st_), "hipGraphLaunch");
        } else {
            for (int b = 0; b < batch; ++b) enqueueIteration(false, b, nullptr, true, capIter);
        }
        check(hipMemcpyAsync(f, flags_.get(), sizeof(int) * 2, hipMemcpyDeviceToHost, st_), "flags");
        check(hipStreamSynchronize(st_), "evolve");
        if (options.profile) readEvents(batch);
        iters = f[1];
        if (f[0]) break;
        if (options.evolveGuard > 0 && iters >= options.evolveGuard) throw std::runtime_error(guardMessage());
        if (compact() && !segmentedNow_) {
            // shrink the active work list once half of it belongs to converged segments
            check(hipMemcpy(doneNow.data(), dSegDone_.get(), sizeof(int) * ns, hipMemcpyDeviceToHost), "done");
            int64_t remaining = 0;
            for (const WorkItem& it : items_)
                if (!doneNow[it.seg]) remaining += it.end - it.start;
            if (2 * remaining <= activeWork_) {
                rebuildActive(doneNow);
                prepareRows(doneNow);
                if (exec) useGraph = captureIterations(batch, true, capIter, &exec);
            }
        }
    }
    {
        std::vector<int> segIter(ns);
        check(hipMemcpy(segIter.data(), dSegIter_.get(), sizeof(int) * ns, hipMemcpyDeviceToHost), "iterations");
        addEvolveBytes(segIter);
    }
    iters = std::max<long long>(iters, localIters);
    segmentedNow_ = false;
    lastEvolveIters = iters;
    timers.evolveIterations += iters;
    timers.evolveMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return iters;
}

template <typename T>
int Engine<T>::evolveIteration(std::vector<int>* labels, std::vector<int>* counts, std::vector<T>* sums,
                               std::vector<unsigned short>* moved) {
    ++version_;
    const int ns = (int)segs_.size();
    std::vector<int> zeros(ns, 0);
    int f[2] = {0, 0};
    check(hipMemcpyAsync(dSegDone_.get(), zeros.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "done");
    check(hipMemcpyAsync(dSegIter_.get(), zeros.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "iter");
    check(hipMemcpyAsync(dSegMoved_.get(), zeros.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "moved");
    check(hipMemcpyAsync(flags_.get(), f, sizeof(int) * 2, hipMemcpyHostToDevice, st_), "flags");
    hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(totalSlots_)), dim3(256), 0, st_, counts_.get(), totalSlots_, 0,
                       nullptr);
    ensureChainScratch();
    enqueueIteration(true);
    std::vector<int> moved0(ns);
    check(hipMemcpyAsync(moved0.data(), dSegMoved_.get(), sizeof(int) * ns, hipMemcpyDeviceToHost, st_), "moved");
    const int64_t totalP = segs_.empty() ? 0 : segs_.back().pOff + segs_.back().nP;
    if (labels) {
        labels->resize((size_t)totalP);
        check(hipMemcpyAsync(labels->data(), labels_.get(), sizeof(int) * totalP, hipMemcpyDeviceToHost, st_), "labels");
    }
    if (counts) {
        counts->resize((size_t)totalSlots_);
        check(hipMemcpyAsync(counts->data(), counts_.get(), sizeof(int) * totalSlots_, hipMemcpyDeviceToHost, st_),
              "counts");
    }
    if (sums) {
        sums->resize((size_t)sStride_ * dim_);
        check(hipMemcpyAsync(sums->data(), sums_.get(), sizeof(T) * sStride_ * dim_, hipMemcpyDeviceToHost, st_),
              "sums");
    }
    std::vector<unsigned char> mv;
    if (moved) {
        mv.resize((size_t)totalSlots_ * dim_);
        check(hipMemcpyAsync(mv.data(), moved_.get(), (size_t)totalSlots_ * dim_, hipMemcpyDeviceToHost, st_), "moved");
    }
    check(hipStreamSynchronize(st_), "evolveIteration");
    if (moved) {
        moved->assign((size_t)totalSlots_, 0);
        for (int d = 0; d < dim_; ++d)
            for (int64_t k = 0; k < totalSlots_; ++k)
                if (mv[(size_t)d * totalSlots_ + k]) (*moved)[k] |= (unsigned short)(1u << d);
    }
    int any = 0;
    for (int v : moved0) any |= v;
    return any;
}

// ------------------------------------------------------------------------------------
// differentiated evolve (memberships frozen)
// ------------------------------------------------------------------------------------
template <typename T>
void Engine<T>::recordRun(int run) {
    const long long first = (long long)run * tape_.every;
    const int steps = (int)std::min<long long>(tape_.every, tape_.iters - first);
    const int64_t kc = std::max(tape_.K, 1);
    for (int u = 0; u < steps; ++u) {
        IterLog lg{tape_.labels.get() + (int64_t)u * tape_.positions, tape_.counts.get() + (int64_t)u * kc,
                   tape_.moved.get() + (int64_t)u * dim_ * kc, kc};
        enqueueIteration(false, -1, &lg);
    }
}

template <typename T>
void Engine<T>::evolveForward(long long iters, int every) {
    if (segs_.size() != 1 || useIndex_)
        throw std::invalid_argument("the evolve forward/backward pair runs on one whole cloud");
    if (iters <= 0 || iters > (1LL << 24)) throw std::invalid_argument("iteration count must be in [1, 2^24]");
    ++version_;
    auto t0 = std::chrono::steady_clock::now();
    tape_.valid = false;
    if (every < 0) {
        // automatic: the longest run whose memberships (labels, counts, moved flags) plus the
        // chosen-set checkpoints fit the memory budget; every iteration's when they all fit
        const double perIter = 4.0 * (double)totalP_ + (4.0 + dim_) * std::max(countsHost()[0], 1);
        const double ckpt = (double)sizeof(T) * (double)sStride_ * dim_;
        const double budget = (double)options.vjpBudgetBytes;
        every = 1;
        for (long long e = iters; e >= 1; e = e > 64 ? e - (e + 63) / 64 : e - 1) {
            const double need = (double)e * perIter + (double)((iters + e - 1) / e) * ckpt;
            if (need <= budget) {
                every = (int)e;
                break;
            }
        }
    }
    if (every <= 0 || every > iters) every = (int)iters;
    timers.vjpCheckpointEvery = every;
    tape_.iters = iters;
    tape_.every = every;
    tape_.runs = (int)((iters + every - 1) / every);
    tape_.K = countsHost()[0];
    tape_.positions = totalP_;
    tape_.r = r1_[0];
    const int64_t kc = std::max(tape_.K, 1);
    const size_t sBytes = sizeof(T) * (size_t)sStride_ * dim_;
    tape_.ckpt.reserve((size_t)tape_.runs * sStride_ * dim_);
    tape_.finalS.reserve((size_t)sStride_ * dim_);
    tape_.labels.reserve((size_t)every * tape_.positions);
    tape_.counts.reserve((size_t)every * kc);
    tape_.moved.reserve((size_t)every * dim_ * kc);
    int zero = 0, f[2] = {0, 0};
    check(hipMemcpyAsync(dSegDone_.get(), &zero, sizeof(int), hipMemcpyHostToDevice, st_), "done");
    check(hipMemcpyAsync(flags_.get(), f, sizeof(int) * 2, hipMemcpyHostToDevice, st_), "flags");
    autotuneScan();
    for (int run = 0; run < tape_.runs; ++run) {
        check(hipMemcpyAsync(tape_.ckpt.get() + (size_t)run * sStride_ * dim_, S_.get(), sBytes,
                             hipMemcpyDeviceToDevice, st_),
              "checkpoint");
        recordRun(run);
    }
    check(hipMemcpyAsync(tape_.finalS.get(), S_.get(), sBytes, hipMemcpyDeviceToDevice, st_), "checkpoint");
    check(hipStreamSynchronize(st_), "evolveForward");
    tape_.loggedRun = tape_.runs - 1;
    tape_.version = version_;
    tape_.valid = true;
    timers.evolveIterations += iters;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    timers.evolveMs += ms;
    timers.forwardMs += ms;
}

template <typename T>
void Engine<T>::evolveBackward(const T* gS, int64_t gsStride, T* gP, int64_t gpStride, T* gS0, int64_t gs0Stride,
                               bool seedToPoints) {
    if (!tape_.valid || tape_.version != version_)
        throw std::logic_error("no evolve forward pass to differentiate (the chosen set changed or none ran)");
    if (seedToPoints && !chosenFromPoints_)
        throw std::invalid_argument("the chosen set was not drawn from the points (set explicitly)");
    auto t0 = std::chrono::steady_clock::now();
    const int K = tape_.K;
    const int64_t kc = std::max(K, 1);
    tape_.G.reserve((size_t)dim_ * kc);
    tape_.A.reserve((size_t)tape_.every * dim_ * kc);
    if (K > 0)
        check(hipMemcpy2DAsync(tape_.G.get(), sizeof(T) * kc, gS, sizeof(T) * gsStride, sizeof(T) * K, dim_,
                               hipMemcpyDeviceToDevice, st_),
              "upstream");
    // recomputed runs replay the forward's iterations from their checkpoints
    const T r1sq = tape_.r * tape_.r;
    check(hipMemcpyAsync(dSegR1Sqr_.get(), &r1sq, sizeof(T), hipMemcpyHostToDevice, st_), "r1");
    int zero = 0, f[2] = {0, 0};
    check(hipMemcpyAsync(dSegDone_.get(), &zero, sizeof(int), hipMemcpyHostToDevice, st_), "done");
    check(hipMemcpyAsync(flags_.get(), f, sizeof(int) * 2, hipMemcpyHostToDevice, st_), "flags");
    const size_t sBytes = sizeof(T) * (size_t)sStride_ * dim_;
    for (int run = tape_.runs - 1; run >= 0; --run) {
        if (run != tape_.loggedRun) {
            check(hipMemcpyAsync(S_.get(), tape_.ckpt.get() + (size_t)run * sStride_ * dim_, sBytes,
                                 hipMemcpyDeviceToDevice, st_),
                  "restore");
            recordRun(run);
            tape_.loggedRun = run;
        }
        const long long first = (long long)run * tape_.every;
        VjpArgs<T> a{tape_.labels.get(), tape_.positions, tape_.counts.get(), tape_.moved.get(), kc,
                     (int)std::min<long long>(tape_.every, tape_.iters - first), K, tape_.G.get(), kc,
                     tape_.A.get(), kc, gP, gpStride, nPoints_, run != tape_.runs - 1 ? 1 : 0};
        check(launch_vjp_run<T>(a, dim_, st_), "vjp");
    }
    if (seedToPoints) check(launch_vjp_seed<T>(tape_.G.get(), kc, chosenIdx_.get(), K, dim_, gP, gpStride, st_), "seed");
    if (gS0 && K > 0)
        check(hipMemcpy2DAsync(gS0, sizeof(T) * gs0Stride, tape_.G.get(), sizeof(T) * kc, sizeof(T) * K, dim_,
                               hipMemcpyDeviceToDevice, st_),
              "seed gradient");
    check(hipMemcpyAsync(S_.get(), tape_.finalS.get(), sBytes, hipMemcpyDeviceToDevice, st_), "restore");
    check(hipStreamSynchronize(st_), "evolveBackward");
    timers.backwardMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------------------------
// decimate
// ------------------------------------------------------------------------------------
template <typename T>
void Engine<T>::decimate(bool noCopyRule, std::vector<unsigned char>* keep) {
    const int ns = (int)segs_.size();
    if (ns == 0) return;
    ++version_;
    auto t0 = std::chrono::steady_clock::now();
    const int ng = (int)groupBegin_.size();
    const int64_t slots = std::max<int64_t>(totalSlots_, 1);
    int* p = decInts_.get();
    DecArgs a{};
    a.segs = dSegs_.get();
    a.numSegs = ns;
    a.segK = dSegK_.get();
    a.groupSegBegin = dGroupBegin_.get();
    a.groupSegEnd = dGroupEnd_.get();
    a.numGroups = ng;
    a.noCopyRule = noCopyRule ? 1 : 0;
    a.elemBase = p; p += ns;
    a.groupOff = p; p += ng;
    a.groupCnt = p; p += ng;
    a.groupLeft = p; p += ng;
    a.elemSlot = p; p += slots;
    a.elemGroup = p; p += slots;
    a.c1 = p; p += slots;
    a.c2 = p; p += slots;
    a.alive = p; p += slots;
    a.rowOff = p; p += slots + 1;
    decWs_.partials = p;
    a.totalElems = decI64_.get();
    decWs_.nbrTotal = decI64_.get() + 1;
    a.segSkip = segSkip_;
    DevBuf<unsigned char> keepBuf;
    if (keep) keepBuf.reserve(slots);
    check(decimate_groups<T>(a, S_.get(), sStride_, dim_, dGroupR2Sqr_.get(), chosenIdx_.get(),
                             keep ? keepBuf.get() : nullptr, totalSlots_, decWs_, st_),
          "decimate");
    if (keep) {
        keep->resize((size_t)totalSlots_);
        check(hipMemcpyAsync(keep->data(), keepBuf.get(), totalSlots_, hipMemcpyDeviceToHost, st_), "keep");
    }
    check(hipStreamSynchronize(st_), "decimate");
    updateKeyBound();
    timers.decimateMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------------------------
// drivers
// ------------------------------------------------------------------------------------
template <typename T>
std::vector<int> Engine<T>::countsHost() {
    std::vector<int> k(segs_.size());
    if (!k.empty()) {
        check(hipMemcpyAsync(k.data(), dSegK_.get(), sizeof(int) * k.size(), hipMemcpyDeviceToHost, st_), "K");
        check(hipStreamSynchronize(st_), "K");
    }
    return k;
}

template <typename T>
int64_t Engine<T>::totalCount() {
    auto k = countsHost();
    return std::accumulate(k.begin(), k.end(), int64_t(0));
}

template <typename T>
void Engine<T>::ridgeDetectionLocal(const RunOptions& opt) {
    options = opt;
    auto t0 = std::chrono::steady_clock::now();
    const int ns = (int)segs_.size();
    choose();
    compactSlots();
    orderItemsByCost();
    std::vector<int> K = countsHost();
    std::vector<char> active(ns);
    std::vector<int> skip(ns);
    for (int s = 0; s < ns; ++s) active[s] = K[s] > 0;
    DevBuf<int> dSkip;
    dSkip.reserve(std::max(ns, 1));
    long long rounds = 0;
    for (;;) {
        bool any = false;
        for (int s = 0; s < ns; ++s) any |= active[s] != 0;
        if (!any) break;
        if (opt.maxRounds > 0 && rounds >= opt.maxRounds) break;
        evolve(opt.maxEvolveIters, active);
        for (int s = 0; s < ns; ++s) skip[s] = active[s] ? 0 : 1;
        check(hipMemcpyAsync(dSkip.get(), skip.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st_), "skip");
        segSkip_ = dSkip.get();
        decimate(false);
        segSkip_ = nullptr;
        std::vector<int> K2 = countsHost();
        for (int s = 0; s < ns; ++s)
            if (active[s] && K2[s] == K[s]) active[s] = 0;   // count stable: this segment is done
        K = K2;
        ++rounds;
    }
    timers.rounds += rounds;
    timers.wholeMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

template <typename T>
void Engine<T>::ridgeDetectionMixed(const RunOptions& opt) {
    options = opt;
    auto t0 = std::chrono::steady_clock::now();
    const int ns = (int)segs_.size();
    choose();
    compactSlots();
    orderItemsByCost();
    std::vector<int> K = countsHost();
    int64_t total = std::accumulate(K.begin(), K.end(), int64_t(0));
    int64_t oldTotal = 0;
    std::vector<char> active(ns, 1);
    long long rounds = 0;
    while (oldTotal != total) {
        if (opt.maxRounds > 0 && rounds >= opt.maxRounds) break;
        oldTotal = total;
        evolve(opt.maxEvolveIters, active);
        decimate(true);
        std::vector<int> K2 = countsHost();
        total = std::accumulate(K2.begin(), K2.end(), int64_t(0));
        // a leaf whose chosen set did not change is a fixed point of an uncapped evolve
        for (int s = 0; s < ns; ++s) active[s] = (opt.maxEvolveIters > 0 || K2[s] != K[s]) ? 1 : 0;
        K = K2;
        ++rounds;
    }
    timers.rounds += rounds;
    timers.wholeMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------------------------
// results
// ------------------------------------------------------------------------------------
template <typename T>
void Engine<T>::getChosen(T* out, int layout, bool onDevice, int64_t outStride) {
    const int ns = (int)segs_.size();
    std::vector<int> K = countsHost();
    std::vector<int64_t> off(ns);
    int64_t total = 0;
    for (int s = 0; s < ns; ++s) {
        off[s] = total;
        total += K[s];
    }
    if (total == 0) return;
    if (layout == LAYOUT_ROW_MAJOR && outStride <= 0) outStride = dim_;
    if (layout != LAYOUT_ROW_MAJOR && outStride <= 0) outStride = total;
    DevBuf<int64_t> dOff;
    dOff.reserve(ns);
    check(hipMemcpyAsync(dOff.get(), off.data(), sizeof(int64_t) * ns, hipMemcpyHostToDevice, st_), "off");
    const size_t bytes = sizeof(T) * (layout == LAYOUT_ROW_MAJOR ? total * outStride : outStride * dim_);
    DevBuf<T> staging;
    T* dst = out;
    if (!onDevice) dst = staging.reserve(bytes / sizeof(T));
    hipLaunchKernelGGL((gather_chosen_kernel<T>), dim3(ns), dim3(256), 0, st_, S_.get(), sStride_, dSegs_.get(),
                       dSegK_.get(), dOff.get(), dim_, dst, layout == LAYOUT_ROW_MAJOR ? 0 : 1, outStride);
    check(hipGetLastError(), "gather");
    if (!onDevice) check(hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, st_), "download");
    check(hipStreamSynchronize(st_), "getChosen");
}

template <typename T>
std::vector<int> Engine<T>::chosenIndicesHost(int seg) {
    std::vector<int> K = countsHost();
    std::vector<int> idx((size_t)K.at(seg));
    if (!idx.empty()) {
        check(hipMemcpyAsync(idx.data(), chosenIdx_.get() + segs_[seg].sOff, sizeof(int) * idx.size(),
                             hipMemcpyDeviceToHost, st_),
              "idx");
        check(hipStreamSynchronize(st_), "idx");
    }
    return idx;
}

template <typename T>
void Engine<T>::setChosen(int seg, const T* S, int64_t K, int layout, int64_t stride, bool onDevice) {
    ++version_;
    chosenFromPoints_ = false;
    if (seg < 0 || seg >= (int)segs_.size()) throw std::invalid_argument("segment out of range");
    if (K < 0 || K > segs_[seg].sCap) throw std::invalid_argument("too many chosen points for the segment");
    if (layout != LAYOUT_ROW_MAJOR && layout != LAYOUT_COL_MAJOR && layout != LAYOUT_SOA)
        throw std::invalid_argument("unsupported memory layout");
    const bool row = layout == LAYOUT_ROW_MAJOR;
    if (stride <= 0) stride = row ? dim_ : K;
    T* dst = S_.get() + segs_[seg].sOff;
    if (K > 0 && onDevice && row) {
        hipLaunchKernelGGL((row_to_soa_kernel<T>), dim3(grid_for(K)), dim3(256), 0, st_, S, K, dim_, stride, dst,
                           sStride_);
        check(hipGetLastError(), "row_to_soa");
    } else if (K > 0 && onDevice) {
        check(hipMemcpy2DAsync(dst, sizeof(T) * sStride_, S, sizeof(T) * stride, sizeof(T) * K, dim_,
                               hipMemcpyDeviceToDevice, st_),
              "setChosen");
    } else if (K > 0) {
        std::vector<T> soa((size_t)K * dim_);
        for (int64_t k = 0; k < K; ++k)
            for (int d = 0; d < dim_; ++d)
                soa[(size_t)d * K + k] = row ? S[k * stride + d] : S[(size_t)d * stride + k];
        check(hipMemcpy2DAsync(dst, sizeof(T) * sStride_, soa.data(), sizeof(T) * K, sizeof(T) * K, dim_,
                               hipMemcpyHostToDevice, st_),
              "setChosen");
    }
    int k32 = (int)K;
    check(hipMemcpyAsync(dSegK_.get() + seg, &k32, sizeof(int), hipMemcpyHostToDevice, st_), "K");
    std::vector<int> idx((size_t)std::max<int64_t>(K, 1));
    std::iota(idx.begin(), idx.end(), 0);
    if (K > 0)
        check(hipMemcpyAsync(chosenIdx_.get() + segs_[seg].sOff, idx.data(), sizeof(int) * K, hipMemcpyHostToDevice,
                             st_),
              "idx");
    fillSentinels();
    check(hipStreamSynchronize(st_), "setChosen");
    updateKeyBound();
}

template class Engine<float>;
template class Engine<double>;

}  // namespace rdhip
