// C++ compression API of the MI355X build, mirroring the reference's
// include/QoZ/api/sz.hpp:54-225: SZ_compress / SZ_decompress over a
// QoZ::Config.  The interpolation engine, quantizer and Huffman coder run on
// the GPU through the C-ABI in faz_gpu.h; zstd and the Config trailer stay on
// the host, so the byte stream is the reference's.
//
// Differences from the reference surface:
//   * errors throw faz::Error instead of printing and calling exit(0);
//   * SZ_decompress does not write the reference's debug file
//     "waved.qoz.dec.final2" (sz.hpp:198);
//   * the GPU engines serve float data: ALGO_INTERP; ALGO_LORENZO_REG (the
//     3-D fast frontend, and the general frontend for other ranks with the
//     Lorenzo predictors only); ALGO_INTERP_LORENZO with conf.QoZ set (the
//     `faz -q` tuner).  Other algorithms, engine modes and types throw
//     faz::Error (FAZ_ERR_UNSUPPORTED).
#pragma once

#include <stdexcept>
#include <string>

#include "faz/config.hpp"

namespace faz {
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};
}  // namespace faz

// Compresses conf.num values of `data` (host memory).  Returns a new[]
// buffer (delete[] it) of outSize bytes: zstd(engine buffer), Config trailer,
// int trailer size (sz.hpp:77-91).
template<class T>
char *SZ_compress(QoZ::Config &config, const T *data, size_t &outSize);

// Decompresses into decData (allocated with new[] when nullptr); the stream's
// Config is loaded into `config` (sz.hpp:159-205).
template<class T>
void SZ_decompress(QoZ::Config &config, char *cmpData, size_t cmpSize, T *&decData);

template<class T>
T *SZ_decompress(QoZ::Config &config, char *cmpData, size_t cmpSize) {
    T *decData = nullptr;
    SZ_decompress<T>(config, cmpData, cmpSize, decData);
    return decData;
}

template<>
char *SZ_compress<float>(QoZ::Config &config, const float *data, size_t &outSize);
template<>
void SZ_decompress<float>(QoZ::Config &config, char *cmpData, size_t cmpSize, float *&decData);
