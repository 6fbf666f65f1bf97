// Uses the drop-in classes the way the reference's README (README.md:146-180) and its
// brute-force benchmark (tests/test_rd_gpu_simulation/brute_force/benchmark.cu:143-158) use
// rd::gpu::tiled::RidgeDetection and rd::gpu::bruteForce::RidgeDetection.
// Usage: dropin_usage <input.bin> <output-prefix>
//   input.bin: int64 n, int64 dim(=3), double r1, double r2, then n*3 float ROW_MAJOR points.
//   writes <prefix>_brute.bin and <prefix>_tiled.bin: int64 count, then count*3 floats;
//   then, as the reference's generator-driven tests do, fills dP_ with
//   SamplesGenerator<float>::spiral3D (20000 samples, a 4 pi, b 0.3, sigma 0.05) and writes
//   the samples (<prefix>_genpts.bin) and their ridge points (<prefix>_genrd.bin).
#include "rdhip/ridge_detection.hpp"

#include <cstdio>
#include <cstdint>
#include <vector>

static void write(const char* path, const std::vector<float>& v, long long n) {
    FILE* f = std::fopen(path, "wb");
    int64_t c = n;
    std::fwrite(&c, sizeof c, 1, f);
    std::fwrite(v.data(), sizeof(float), v.size(), f);
    std::fclose(f);
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
    try {
        // brute force: the caller writes the samples into the public device buffer dP_
        rd::gpu::bruteForce::RidgeDetection<float, 3, rd::ROW_MAJOR, rd::ROW_MAJOR> bf((size_t)n, (float)r1,
                                                                                           (float)r2);
        if (hipMemcpy(bf.dP_, P.data(), sizeof(float) * P.size(), hipMemcpyHostToDevice) != hipSuccess) return 3;
        bf.ridgeDetection();
        std::vector<float> S(bf.ns_ * 3);
        if (hipMemcpy(S.data(), bf.dS_, sizeof(float) * S.size(), hipMemcpyDeviceToHost) != hipSuccess) return 3;
        write((std::string(argv[2]) + "_brute.bin").c_str(), S, (long long)bf.ns_);

        // tiled, extended tiles, mixed policy (README example)
        using namespace rd::gpu::tiled;
        RidgeDetection<3, rd::ROW_MAJOR, rd::ROW_MAJOR, rd::gpu::RD_BRUTE_FORCE, RD_MIXED, RD_EXTENDED_TILE, float>
            rdGpu((int)n, P.data(), true, false);
        int dimTiles[3] = {4, 4, 4};
        rdGpu((float)r1, (float)r2, 2500, dimTiles, false);
        int k = rdGpu.getChosenPointsNum();
        std::vector<float> T((size_t)k * 3);
        rdGpu.getChosenPoints(T.data());
        write((std::string(argv[2]) + "_tiled.bin").c_str(), T, k);
        std::printf("brute %zu tiled %d\n", bf.ns_, k);

        // samples generated straight into the device buffer
        const size_t ng = 20000;
        rd::gpu::bruteForce::RidgeDetection<float, 3, rd::ROW_MAJOR, rd::ROW_MAJOR> bg(ng, 0.3f, 0.6f);
        rd::gpu::SamplesGenerator<float>::spiral3D(ng, 4.0f * 3.14159265358979f, 0.3f, 0.05f, bg.dP_);
        std::vector<float> G(ng * 3);
        if (hipMemcpy(G.data(), bg.dP_, sizeof(float) * G.size(), hipMemcpyDeviceToHost) != hipSuccess) return 3;
        write((std::string(argv[2]) + "_genpts.bin").c_str(), G, (long long)ng);
        bg.ridgeDetection();
        std::vector<float> R(bg.ns_ * 3);
        if (hipMemcpy(R.data(), bg.dS_, sizeof(float) * R.size(), hipMemcpyDeviceToHost) != hipSuccess) return 3;
        write((std::string(argv[2]) + "_genrd.bin").c_str(), R, (long long)bg.ns_);
    } catch (const rd::gpu::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
