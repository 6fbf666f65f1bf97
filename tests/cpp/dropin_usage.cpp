// Uses the drop-in headers the way the reference's callers use its classes and facades:
//   rd::gpu::tiled::RidgeDetection           -- README.md:146-180
//   rd::gpu::bruteForce::RidgeDetection      -- tests/test_rd_gpu_simulation/brute_force/benchmark.cu:143-158
//   DeviceChoose / DeviceEvolve / DeviceDecimate -- tests/test_choose_kernel, test_evolve_brute_force,
//                                               test_decimate_kernel
//   rd::RidgeDetection<T>, rd::tiled::tiledRidgeDetection -- tests/test_rd_cpu*, test_rd_tiled*
// Usage: dropin_usage <input.bin> <output-prefix>
//   input.bin: int64 n, int64 dim(=3), double r1, double r2, then n*3 float ROW_MAJOR points.
// Writes <prefix>_<case>.bin files: int64 count, then count*3 floats (ROW_MAJOR), except
// <prefix>_devolve_stats.bin (int64 K, K int32 counts, K*3 float sums) and <prefix>_cpu_chains.bin
// (int64 m, m int64 chain offsets + 1 more, then the chained point indices as int64).
#include "rdhip/cpu_ridge_detection.hpp"
#include "rdhip/ridge_detection.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

static std::string g_prefix;

static void writeRaw(const std::string& name, const void* data, size_t bytes, int64_t count) {
    FILE* f = std::fopen((g_prefix + "_" + name + ".bin").c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + name);
    std::fwrite(&count, sizeof count, 1, f);
    if (bytes) std::fwrite(data, 1, bytes, f);
    std::fclose(f);
}

static void writePts(const std::string& name, const std::vector<float>& rowMajor, int64_t count) {
    writeRaw(name, rowMajor.data(), sizeof(float) * (size_t)count * 3, count);
}

#define HIP_OK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_) + \
                                                       " / " + rdhip_last_error());        \
    } while (0)

// device COL_MAJOR (3 rows of `pitchBytes`) -> host ROW_MAJOR
static std::vector<float> colToRows(const float* d, size_t pitchBytes, size_t count) {
    std::vector<float> col(count * 3), row(count * 3);
    if (count)
        HIP_OK(hipMemcpy2D(col.data(), sizeof(float) * count, d, pitchBytes, sizeof(float) * count, 3,
                           hipMemcpyDeviceToHost));
    for (size_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) row[i * 3 + k] = col[k * count + i];
    return row;
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t n = 0, dim = 0;
    double r1 = 0, r2 = 0;
    if (std::fread(&n, 8, 1, f) != 1 || std::fread(&dim, 8, 1, f) != 1 || std::fread(&r1, 8, 1, f) != 1 ||
        std::fread(&r2, 8, 1, f) != 1 || dim != 3)
        return 2;
    std::vector<float> P((size_t)n * 3);
    if (std::fread(P.data(), sizeof(float), P.size(), f) != P.size()) return 2;
    std::fclose(f);
    g_prefix = argv[2];
    std::vector<float> Pcol((size_t)n * 3);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) Pcol[(size_t)k * n + i] = P[(size_t)i * 3 + k];
    const float fr1 = (float)r1, fr2 = (float)r2;
    try {
        // ---- brute force: the caller writes the samples into the public device buffer dP_ ----
        {
            rd::gpu::bruteForce::RidgeDetection<float, 3, rd::ROW_MAJOR, rd::ROW_MAJOR> bf((size_t)n, fr1, fr2);
            HIP_OK(hipMemcpy(bf.dP_, P.data(), sizeof(float) * P.size(), hipMemcpyHostToDevice));
            bf.ridgeDetection();
            std::vector<float> S(bf.ns_ * 3);
            HIP_OK(hipMemcpy(S.data(), bf.dS_, sizeof(float) * S.size(), hipMemcpyDeviceToHost));
            const size_t ns = bf.ns_;
            bf.ns_ = 0;
            bf.getChosenSamplesCount();   // from dNs_
            if (bf.ns_ != ns) throw std::runtime_error("dNs_ disagrees with ns_");
            writePts("brute", S, (int64_t)ns);
        }
        {   // pitched COL_MAJOR samples and results
            rd::gpu::bruteForce::RidgeDetection<float, 3, rd::COL_MAJOR, rd::COL_MAJOR> bf((size_t)n, fr1, fr2);
            HIP_OK(hipMemcpy2D(bf.dP_, bf.pPitch_, Pcol.data(), sizeof(float) * n, sizeof(float) * n, 3,
                               hipMemcpyHostToDevice));
            bf.ridgeDetection();
            writePts("brute_col", colToRows(bf.dS_, bf.sPitch_, bf.ns_), (int64_t)bf.ns_);
        }
        {   // ROW_MAJOR samples, pitched COL_MAJOR results
            rd::gpu::bruteForce::RidgeDetection<float, 3, rd::ROW_MAJOR, rd::COL_MAJOR> bf((size_t)n, fr1, fr2);
            HIP_OK(hipMemcpy(bf.dP_, P.data(), sizeof(float) * P.size(), hipMemcpyHostToDevice));
            bf.ridgeDetection();
            writePts("brute_rowcol", colToRows(bf.dS_, bf.sPitch_, bf.ns_), (int64_t)bf.ns_);
        }

        // ---- tiled, extended tiles, mixed policy (README example) ----
        using namespace rd::gpu::tiled;
        int dimTiles[3] = {4, 4, 4};
        {
            RidgeDetection<3, rd::ROW_MAJOR, rd::ROW_MAJOR, rd::gpu::RD_BRUTE_FORCE, RD_MIXED, RD_EXTENDED_TILE, float>
                rdGpu((int)n, P.data(), true, false);
            rdGpu(fr1, fr2, 2500, dimTiles, false);
            int k = rdGpu.getChosenPointsNum();
            std::vector<float> T((size_t)k * 3);
            rdGpu.getChosenPoints(T.data());
            writePts("tiled", T, k);
        }
        {   // COL_MAJOR on the device; getChosenPoints still returns host ROW_MAJOR
            RidgeDetection<3, rd::COL_MAJOR, rd::COL_MAJOR, rd::gpu::RD_BRUTE_FORCE, RD_MIXED, RD_EXTENDED_TILE, float>
                rdGpu((int)n, P.data());
            rdGpu.ridgeDetection(fr1, fr2, 2500, dimTiles, false);
            int k = rdGpu.getChosenPointsNum();
            std::vector<float> T((size_t)k * 3);
            rdGpu.getChosenPoints(T.data());
            writePts("tiled_col", T, k);
        }
        int soaThrows = 0;
        {   // SOA results are rejected, as the reference rejects them
            RidgeDetection<3, rd::ROW_MAJOR, rd::SOA, rd::gpu::RD_BRUTE_FORCE, RD_MIXED, RD_EXTENDED_TILE, float> rdGpu(
                (int)n, P.data());
            rdGpu(fr1, fr2, 2500, dimTiles, false);
            std::vector<float> T((size_t)rdGpu.getChosenPointsNum() * 3 + 3);
            try {
                rdGpu.getChosenPoints(T.data());
            } catch (const std::runtime_error&) {
                soaThrows += 1;
            }
        }
        try {
            RidgeDetection<3, rd::SOA, rd::ROW_MAJOR, rd::gpu::RD_BRUTE_FORCE, RD_MIXED, RD_EXTENDED_TILE, float> bad(
                (int)n, P.data());
        } catch (const std::runtime_error&) {
            soaThrows += 2;
        }
        writeRaw("soa_throws", nullptr, 0, soaThrows);

        // ---- device facades ----
        using rd::gpu::bruteForce::DeviceChoose;
        using rd::gpu::bruteForce::DeviceDecimate;
        using rd::gpu::bruteForce::DeviceEvolve;
        float *dP = nullptr, *dPcol = nullptr, *dS = nullptr, *dScol = nullptr, *dSums = nullptr;
        int *dCount = nullptr, *dCounts = nullptr;
        size_t pPitch = 0, sPitch = 0;
        HIP_OK(hipMalloc((void**)&dP, sizeof(float) * n * 3));
        HIP_OK(hipMemcpy(dP, P.data(), sizeof(float) * n * 3, hipMemcpyHostToDevice));
        HIP_OK(hipMallocPitch((void**)&dPcol, &pPitch, sizeof(float) * n, 3));
        HIP_OK(hipMemcpy2D(dPcol, pPitch, Pcol.data(), sizeof(float) * n, sizeof(float) * n, 3, hipMemcpyHostToDevice));
        HIP_OK(hipMalloc((void**)&dS, sizeof(float) * n * 3));
        HIP_OK(hipMallocPitch((void**)&dScol, &sPitch, sizeof(float) * n, 3));
        HIP_OK(hipMalloc((void**)&dSums, sizeof(float) * n * 3));
        HIP_OK(hipMalloc((void**)&dCount, sizeof(int)));
        HIP_OK(hipMalloc((void**)&dCounts, sizeof(int) * n));
        hipStream_t st;
        HIP_OK(hipStreamCreate(&st));

        // choose: ROW_MAJOR samples -> pitched COL_MAJOR chosen set
        HIP_OK((DeviceChoose::choose<3, rd::ROW_MAJOR, rd::COL_MAJOR, float>(
            dP, dScol, (int)n, dCount, fr1, 3, (int)(sPitch / sizeof(float)), st)));
        int K = 0;
        HIP_OK(hipMemcpy(&K, dCount, sizeof(int), hipMemcpyDeviceToHost));
        writePts("dchoose", colToRows(dScol, sPitch, (size_t)K), K);

        // evolve: pitched COL_MAJOR samples, the chosen set copied to ROW_MAJOR
        std::vector<float> S0 = colToRows(dScol, sPitch, (size_t)K);
        HIP_OK(hipMemcpy(dS, S0.data(), sizeof(float) * K * 3, hipMemcpyHostToDevice));
        HIP_OK((DeviceEvolve::evolve<3, rd::COL_MAJOR, rd::ROW_MAJOR, float>(
            dPcol, dS, dSums, dCounts, (int)n, K, fr1, (int)(pPitch / sizeof(float)), 3, 3, st)));
        std::vector<float> SE((size_t)K * 3), sums((size_t)K * 3);
        std::vector<int> counts((size_t)K);
        HIP_OK(hipMemcpy(SE.data(), dS, sizeof(float) * K * 3, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(sums.data(), dSums, sizeof(float) * K * 3, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(counts.data(), dCounts, sizeof(int) * K, hipMemcpyDeviceToHost));
        writePts("devolve", SE, K);
        {
            std::vector<char> buf(sizeof(int) * K + sizeof(float) * K * 3);
            std::memcpy(buf.data(), counts.data(), sizeof(int) * K);
            std::memcpy(buf.data() + sizeof(int) * K, sums.data(), sizeof(float) * K * 3);
            writeRaw("devolve_stats", buf.data(), buf.size(), K);
        }

        // evolve over own + neighbour samples: the first half and the second half of the cloud
        HIP_OK(hipMemcpy(dS, S0.data(), sizeof(float) * K * 3, hipMemcpyHostToDevice));
        const int nOwn = (int)(n / 2);
        HIP_OK((DeviceEvolve::evolve<3, rd::ROW_MAJOR, rd::ROW_MAJOR, float>(
            dP, dP + (size_t)nOwn * 3, dS, dSums, dCounts, nOwn, (int)n - nOwn, K, fr1, 3, 3, 3, 3, st)));
        std::vector<float> SN((size_t)K * 3);
        HIP_OK(hipMemcpy(SN.data(), dS, sizeof(float) * K * 3, hipMemcpyDeviceToHost));
        writePts("devolve_nb", SN, K);

        // decimate the evolved set: ROW_MAJOR in place, and pitched COL_MAJOR via decimateDistMtx
        HIP_OK(hipMemcpy(dS, SE.data(), sizeof(float) * K * 3, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dCount, &K, sizeof(int), hipMemcpyHostToDevice));
        HIP_OK((DeviceDecimate::decimate<3, rd::ROW_MAJOR, float>(dS, dCount, fr2, 3, st)));
        int KD = 0;
        HIP_OK(hipMemcpy(&KD, dCount, sizeof(int), hipMemcpyDeviceToHost));
        std::vector<float> SD((size_t)KD * 3);
        HIP_OK(hipMemcpy(SD.data(), dS, sizeof(float) * KD * 3, hipMemcpyDeviceToHost));
        writePts("ddecimate", SD, KD);
        {
            std::vector<float> seCol((size_t)K * 3);
            for (int i = 0; i < K; ++i)
                for (int k = 0; k < 3; ++k) seCol[(size_t)k * K + i] = SE[(size_t)i * 3 + k];
            HIP_OK(hipMemcpy2D(dScol, sPitch, seCol.data(), sizeof(float) * K, sizeof(float) * K, 3,
                               hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(dCount, &K, sizeof(int), hipMemcpyHostToDevice));
            HIP_OK((DeviceDecimate::decimateDistMtx<3, rd::COL_MAJOR, float>(dScol, dCount, (int)(sPitch / sizeof(float)),
                                                                             nullptr, 0, nullptr, fr2, st)));
            HIP_OK(hipMemcpy(&KD, dCount, sizeof(int), hipMemcpyDeviceToHost));
            writePts("ddecimate_col", colToRows(dScol, sPitch, (size_t)KD), KD);
        }
        (void)hipStreamDestroy(st);
        (void)hipFree(dP);
        (void)hipFree(dPcol);
        (void)hipFree(dS);
        (void)hipFree(dScol);
        (void)hipFree(dSums);
        (void)hipFree(dCount);
        (void)hipFree(dCounts);

        // ---- CPU API ----
        {
            rd::RidgeDetection<float> cpu;
            cpu.order_ = true;
            cpu.noOMP();
            std::vector<float> S((size_t)n * 3);
            cpu.ridgeDetection(P.data(), (size_t)n, S.data(), fr1, fr2, 3);
            writePts("cpu_rd", S, (int64_t)cpu.ns_);
            std::vector<int64_t> out;
            out.push_back(0);
            std::vector<int64_t> idx;
            for (auto& chain : cpu.chainList_) {
                for (float const* p : chain) idx.push_back((int64_t)((p - S.data()) / 3));
                out.push_back((int64_t)idx.size());
            }
            out.insert(out.end(), idx.begin(), idx.end());
            writeRaw("cpu_chains", out.data(), sizeof(int64_t) * out.size(), (int64_t)cpu.chainList_.size());
        }
        {   // sample-list overload: choose over the first half, evolve over both halves
            rd::RidgeDetection<float> cpu;
            std::vector<float> S((size_t)n * 3);
            std::vector<std::pair<float const*, size_t>> sets{{P.data(), (size_t)nOwn},
                                                              {P.data() + (size_t)nOwn * 3, (size_t)(n - nOwn)}};
            cpu.ridgeDetection(sets, S.data(), fr1, fr2, 3);
            writePts("cpu_rd_own", S, (int64_t)cpu.ns_);
        }
        {
            rd::tiled::rdTiledData<float> dp;
            dp.dim = 3;
            dp.np = (size_t)n;
            dp.P = P.data();
            dp.r1 = fr1;
            dp.r2 = fr2;
            dp.maxTileCapacity = 2500;
            dp.nTilesPerDim = {4, 4, 4};
            auto t = rd::tiled::tiledRidgeDetection<rd::tiled::EXT_TILE_TREE_MIXED_RD>(dp, 4, true);
            if (std::get<0>(t) <= 0.f) throw std::runtime_error("tiledRidgeDetection returned no whole time");
            writePts("cpu_tiled", dp.S, (int64_t)dp.ns);
        }

        // ---- samples generated straight into the device buffer ----
        const size_t ng = 20000;
        rd::gpu::bruteForce::RidgeDetection<float, 3, rd::ROW_MAJOR, rd::ROW_MAJOR> bg(ng, 0.3f, 0.6f);
        rd::gpu::SamplesGenerator<float>::spiral3D(ng, 4.0f * 3.14159265358979f, 0.3f, 0.05f, bg.dP_);
        std::vector<float> G(ng * 3);
        HIP_OK(hipMemcpy(G.data(), bg.dP_, sizeof(float) * G.size(), hipMemcpyDeviceToHost));
        writePts("genpts", G, (int64_t)ng);
        bg.ridgeDetection();
        std::vector<float> R(bg.ns_ * 3);
        HIP_OK(hipMemcpy(R.data(), bg.dS_, sizeof(float) * R.size(), hipMemcpyDeviceToHost));
        writePts("genrd", R, (int64_t)bg.ns_);
        std::printf("drop-in cases written (choose K = %d)\n", K);
    } catch (const rd::gpu::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
