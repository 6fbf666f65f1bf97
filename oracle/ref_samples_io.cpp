// ref_samples_io.cpp -- the reference's sample text writer and readers as a program.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the reference headers where
// they lie (`/root/reference/ridge_detection`), into oracle/_ref/.  The reference resolves file
// names against the directory of the running executable (rd/utils/utilities.hpp findPath), so
// this is a program of its own rather than part of the ctypes driver, and its files live in
// oracle/_ref/.  tests/golden/make_golden.py runs it to write the text-format fixtures that pin
// ridgedetection_amd/samples.py.
//
//   ref_samples_io save <in.bin> <name> [comment]
//       in.bin: int64 n, int64 dim, int64 is_double, then n*dim values of that type.
//       rd::Samples<T>::saveToFile(name, comment) (rd/utils/samples_set.hpp:342-372, and the
//       double specialisation :474-503) -> oracle/_ref/<name>.txt
//   ref_samples_io load <name.txt> <is_double>
//       rd::Samples<T>::loadFromFile (samples_set.hpp:254-291) -> oracle/_ref/<name.txt>.bin
//   ref_samples_io loadgen <name.txt> <is_double> [downgradeToDim]
//       rd::Samples<T>::loadGeneratedSamples (samples_set.hpp:147-241) -> oracle/_ref/<name.txt>.bin
//   Outputs: int64 n, int64 dim, then n*dim doubles (the values the reader produced, widened).

#include "rd/utils/samples_set.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::string exeDir() { return rd::findPath("", "", true); }

template <typename T>
int save(const std::vector<char>& blob, int64_t n, int64_t dim, const std::string& name, const std::string& comment) {
    rd::Samples<T> s;
    s.dim_ = (size_t)dim;
    s.size_ = (size_t)n;
    s.samples_ = new T[(size_t)(n * dim)];
    std::memcpy(s.samples_, blob.data() + 24, sizeof(T) * (size_t)(n * dim));
    s.saveToFile(name, comment);
    return 0;
}

template <typename T>
int dump(const rd::Samples<T>& s, const std::string& name) {
    std::FILE* f = std::fopen((exeDir() + name + ".bin").c_str(), "wb");
    if (!f) return 3;
    int64_t hdr[2] = {(int64_t)s.size_, (int64_t)s.dim_};
    std::fwrite(hdr, sizeof hdr, 1, f);
    for (size_t i = 0; i < s.size_ * s.dim_; ++i) {
        double v = (double)s.samples_[i];
        std::fwrite(&v, sizeof v, 1, f);
    }
    std::fclose(f);
    return 0;
}

template <typename T>
int load(const std::string& name) {
    rd::Samples<T> s;
    s.loadFromFile(name);
    return dump(s, name);
}

template <typename T>
int loadgen(const std::string& name, size_t downgrade) {
    rd::Samples<T> s;
    rd::RDParams<T> p;
    rd::RDSpiralParams<T> sp;
    sp.file = name;
    s.loadGeneratedSamples(p, sp, std::vector<std::string>{""}, downgrade);
    return dump(s, name);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    try {
        if (mode == "save" && argc >= 4) {
            std::FILE* f = std::fopen(argv[2], "rb");
            if (!f) return 2;
            std::vector<char> blob;
            char buf[65536];
            size_t r;
            while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) blob.insert(blob.end(), buf, buf + r);
            std::fclose(f);
            if (blob.size() < 24) return 2;
            int64_t hdr[3];
            std::memcpy(hdr, blob.data(), sizeof hdr);
            const std::string comment = argc >= 5 ? argv[4] : "";
            return hdr[2] ? save<double>(blob, hdr[0], hdr[1], argv[3], comment)
                          : save<float>(blob, hdr[0], hdr[1], argv[3], comment);
        }
        if (mode == "load" && argc >= 4) return std::atoi(argv[3]) ? load<double>(argv[2]) : load<float>(argv[2]);
        if (mode == "loadgen" && argc >= 4) {
            const size_t down = argc >= 5 ? (size_t)std::atol(argv[4]) : 0;
            return std::atoi(argv[3]) ? loadgen<double>(argv[2], down) : loadgen<float>(argv[2], down);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 2;
}
