// Drop-in use of the C++ surface (include/faz/sz.hpp): a program written
// against the reference's SZ_compress<float> / SZ_decompress<float> over a
// QoZ::Config (include/QoZ/api/sz.hpp:54-225), built against this library.
//
//   sz_roundtrip <in.f32> <out.faz> <out.f32> <REL bound> <interp|qoz|lorenzo> d0 [d1 [d2 [d3]]]
//
// Dimensions slowest first.  Compresses the field, writes the stream and the
// decompressed field, and prints "bytes=<stream bytes> max_err=<max |x - x'|>".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "faz/sz.hpp"

int main(int argc, char **argv) {
    if (argc < 7 || argc > 10) {
        fprintf(stderr, "usage: sz_roundtrip in.f32 out.faz out.f32 rel interp|qoz|lorenzo d0 [d1 [d2 [d3]]]\n");
        return 2;
    }
    std::vector<size_t> dims;
    for (int i = 6; i < argc; i++) dims.push_back(strtoull(argv[i], nullptr, 10));
    size_t num = 1;
    for (size_t d: dims) num *= d;
    std::vector<float> data(num);
    {
        std::ifstream f(argv[1], std::ios::binary);
        if (!f.read(reinterpret_cast<char *>(data.data()), (std::streamsize) (num * sizeof(float)))) {
            fprintf(stderr, "cannot read %zu floats from %s\n", num, argv[1]);
            return 2;
        }
    }
    QoZ::Config conf;
    conf.setDims(dims.begin(), dims.end());
    conf.errorBoundMode = QoZ::EB_REL;
    conf.relErrorBound = atof(argv[4]);
    const std::string mode = argv[5];
    if (mode == "interp") {
        conf.cmprAlgo = QoZ::ALGO_INTERP;
        conf.interpAlgo = QoZ::INTERP_ALGO_CUBIC;
        conf.maxStep = dims.size() == 2 ? 64 : 32;
    } else if (mode == "qoz") {
        conf.cmprAlgo = QoZ::ALGO_INTERP_LORENZO;
        conf.QoZ = 1;
    } else if (mode == "lorenzo") {
        conf.cmprAlgo = QoZ::ALGO_LORENZO_REG;
    } else {
        fprintf(stderr, "unknown mode %s\n", argv[5]);
        return 2;
    }
    try {
        size_t outSize = 0;
        char *stream = SZ_compress<float>(conf, data.data(), outSize);
        std::ofstream(argv[2], std::ios::binary).write(stream, (std::streamsize) outSize);
        QoZ::Config dconf;
        float *dec = SZ_decompress<float>(dconf, stream, outSize);
        std::ofstream(argv[3], std::ios::binary).write(reinterpret_cast<char *>(dec),
                                                      (std::streamsize) (dconf.num * sizeof(float)));
        double err = 0;
        for (size_t i = 0; i < num; i++) err = std::max(err, std::fabs((double) dec[i] - (double) data[i]));
        printf("bytes=%zu max_err=%.9g num=%zu\n", outSize, err, (size_t) dconf.num);
        delete[] dec;
        delete[] stream;
    } catch (const faz::Error &e) {
        fprintf(stderr, "faz::Error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
