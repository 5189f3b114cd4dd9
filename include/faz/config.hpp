// QoZ::Config for the MI355X build: the configuration object of the
// reference's C++ surface (include/QoZ/utils/Config.hpp:62-508), with the same
// field names, defaults, INI keys and binary trailer layout, so streams and
// config files move between the two builds unchanged.
#pragma once

#include <array>
#include <climits>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>
#include <vector>

namespace QoZ {

typedef unsigned int uint;
typedef unsigned char uchar;

enum EB { EB_ABS, EB_REL, EB_PSNR, EB_L2NORM, EB_ABS_AND_REL, EB_ABS_OR_REL };
constexpr const char *EB_STR[] = {"ABS", "REL", "PSNR", "NORM", "ABS_AND_REL", "ABS_OR_REL"};

enum ALGO { ALGO_LORENZO_REG, ALGO_INTERP_LORENZO, ALGO_INTERP, ALGO_INTERP_BLOCKED };
constexpr const char *ALGO_STR[] = {"ALGO_LORENZO_REG", "ALGO_INTERP_LORENZO", "ALGO_INTERP", "ALGO_INTERP_BLOCKED"};

enum INTERP_ALGO { INTERP_ALGO_LINEAR, INTERP_ALGO_CUBIC };
constexpr const char *INTERP_ALGO_STR[] = {"INTERP_ALGO_LINEAR", "INTERP_ALGO_CUBIC"};

enum TUNING_TARGET { TUNING_TARGET_RD, TUNING_TARGET_CR, TUNING_TARGET_SSIM, TUNING_TARGET_AC };
constexpr const char *TUNING_TARGET_STR[] = {"TUNING_TARGET_RD", "TUNING_TARGET_CR", "TUNING_TARGET_SSIM",
                                             "TUNING_TARGET_AC"};

// Raw little-endian field serialisation (include/QoZ/utils/MemoryUtil.hpp:11-55).
template<class T1>
inline void write(T1 const var, uchar *&p) {
    memcpy(p, &var, sizeof(T1));
    p += sizeof(T1);
}
template<class T1>
inline void write(T1 const *arr, size_t n, uchar *&p) {
    memcpy(p, arr, n * sizeof(T1));
    p += n * sizeof(T1);
}
template<class T1>
inline void read(T1 &var, uchar const *&p) {
    memcpy(&var, p, sizeof(T1));
    p += sizeof(T1);
}
template<class T1>
inline void read(T1 *arr, size_t n, uchar const *&p) {
    memcpy(arr, p, n * sizeof(T1));
    p += n * sizeof(T1);
}

class Config {
public:
    template<class... Dims>
    Config(Dims... args) {
        dims = std::vector<size_t>{static_cast<size_t>(args)...};
        init_dims();
    }

    template<class Iter>
    size_t setDims(Iter begin, Iter end) {
        dims = std::vector<size_t>(begin, end);
        init_dims();
        return num;
    }

    // INI configuration (Config.hpp:91-233).  Returns false when the file
    // cannot be parsed (the reference exits instead).
    bool loadcfg(const std::string &cfgpath);

    static size_t size_est() {
        return sizeof(size_t) * 10 + sizeof(double) * 8 + sizeof(bool) * 10 + sizeof(uint8_t) * 12 +
               sizeof(int) * 10 + 200;
    }

    // Binary trailer (Config.hpp:240-309).
    void save(uchar *&c) const {
        write(N, c);
        write(dims.data(), dims.size(), c);
        write(num, c);
        write(cmprAlgo, c);
        write(errorBoundMode, c);
        write(tuningTarget, c);
        write(absErrorBound, c);
        write(relErrorBound, c);
        write(alpha, c);
        write(beta, c);
        write(lorenzo, c);
        write(lorenzo2, c);
        write(regression, c);
        write(regression2, c);
        write(interpAlgo, c);
        write(interpDirection, c);
        write(interpBlockSize, c);
        write(lossless, c);
        write(encoder, c);
        write(quantbinCnt, c);
        write(blockSize, c);
        write(levelwisePredictionSelection, c);
        write(blockwiseTuning, c);
        write(stride, c);
        write(maxStep, c);
        write(pred_dim, c);
        write(openmp, c);
        write(fixBlockSize, c);
        write(blockwiseSampleBlockSize, c);
        write(crossBlock, c);
        write(wavelet, c);
        write(firstSize, c);
        write(offsetPredictor, c);
        write(coeffs_num, c);
        if (coeffs_num > 0) {
            size_t k = coeffs_dims.size();
            write(k, c);
            write(coeffs_dims.data(), k, c);
        }
        write(pid, c);
        write(blockOrder, c);
        write(sperr, c);
        write(sperrWithoutWave, c);
        write(pyBind, c);
        if (pyBind) {
            write(metadata.size(), c);
            write(metadata.data(), metadata.size(), c);
        }
        write(conditioning, c);
        if (conditioning > 0) write(meta.data(), 17, c);
    }

    // Config.hpp:311-388.  The reference reads without bounds; load(c, end)
    // checks every read against `end` and the field values a stream cannot
    // hold (N outside 1..4, counts beyond the bytes left) and returns false on
    // a truncated or corrupt trailer.
    void load(const uchar *&c) { load_from(c, nullptr); }
    bool load(const uchar *&c, const uchar *end) { return end && load_from(c, end); }

private:
    bool load_from(const uchar *&c, const uchar *end) {
        bool ok = true;
        auto left = [&]() { return end ? (size_t) (end - c) : SIZE_MAX; };
        auto rd = [&](auto &v) {
            if (ok && sizeof(v) <= left()) read(v, c);
            else ok = false;
        };
        auto rda = [&](auto *a, size_t n) {
            if (ok && n <= left() / sizeof(*a)) read(a, n, c);
            else ok = false;
        };
        rd(N);
        if (!ok || (end && (N < 1 || N > 4))) return false;
        dims.resize(N);
        rda(dims.data(), (size_t) N);
        rd(num);
        rd(cmprAlgo);
        rd(errorBoundMode);
        rd(tuningTarget);
        rd(absErrorBound);
        rd(relErrorBound);
        rd(alpha);
        rd(beta);
        rd(lorenzo);
        rd(lorenzo2);
        rd(regression);
        rd(regression2);
        rd(interpAlgo);
        rd(interpDirection);
        rd(interpBlockSize);
        rd(lossless);
        rd(encoder);
        rd(quantbinCnt);
        rd(blockSize);
        rd(levelwisePredictionSelection);
        rd(blockwiseTuning);
        rd(stride);
        rd(maxStep);
        rd(pred_dim);
        rd(openmp);
        rd(fixBlockSize);
        rd(blockwiseSampleBlockSize);
        rd(crossBlock);
        rd(wavelet);
        rd(firstSize);
        rd(offsetPredictor);
        rd(coeffs_num);
        if (ok && coeffs_num > 0) {
            size_t k = 0;
            rd(k);
            if (!ok || k > left() / sizeof(size_t)) return false;
            coeffs_dims.resize(k);
            rda(coeffs_dims.data(), k);
        }
        rd(pid);
        rd(blockOrder);
        rd(sperr);
        rd(sperrWithoutWave);
        rd(pyBind);
        if (ok && pyBind) {
            size_t m = 0;
            rd(m);
            if (!ok || m > left()) return false;
            metadata.resize(m);
            if (m) rda(&metadata[0], m);
        }
        rd(conditioning);
        if (ok && conditioning > 0) rda(meta.data(), 17);
        return ok;
    }

public:
    char N = 0;
    std::vector<size_t> dims;
    size_t num = 0;
    uint8_t cmprAlgo = ALGO_INTERP_LORENZO;
    uint8_t errorBoundMode = EB_ABS;
    uint8_t tuningTarget = TUNING_TARGET_RD;
    double absErrorBound = 0;
    double relErrorBound = -1.0;
    double psnrErrorBound = 0;
    double l2normErrorBound = 0;
    double rng = -1;
    double alpha = -1;
    double beta = -1;
    double autoTuningRate = 0.0;
    double predictorTuningRate = 0.0;
    bool lorenzo = true;
    bool lorenzo2 = false;
    bool regression = true;
    bool regression2 = false;
    bool openmp = false;
    uint8_t lossless = 1;
    uint8_t encoder = 1;
    uint8_t interpAlgo = INTERP_ALGO_CUBIC;
    uint8_t interpDirection = 0;
    int levelwisePredictionSelection = 0;
    std::vector<uint8_t> interpAlgo_list;
    std::vector<uint8_t> interpDirection_list;
    size_t maxStep = 0;
    int interpBlockSize = 32;
    int quantbinCnt = 65536;
    int blockSize = 6;
    int testLorenzo = 0;
    std::vector<int> quant_bins;
    double decomp_square_error = 0;
    std::vector<size_t> quant_bin_counts;
    int sampleBlockSize = 0;
    bool blockwiseTuning = 0;
    int stride = 6;
    int pred_dim = 3;
    int profiling = 0;
    int SSIMBlockSize = 8;
    int fixBlockSize = 0;
    int blockwiseSampleBlockSize = 0;
    bool verbose = 1;
    bool QoZ = 0;
    int abList = 0;
    int crossBlock = 0;
    int sampleBlockSampleBlockSize = 0;
    bool peTracking = 0;
    int wavelet = 0;
    std::vector<size_t> coeffs_dims;
    size_t coeffs_num = 0;
    double wavelet_rel_coeff = 0.75;
    size_t firstSize = 0;  // the reference leaves this uninitialised when wavelet == 0
    int coeffTracking = 0;
    int pid = 0;
    int offsetPredictor = 0;
    int trimToZero = 0;
    double preTrim = 0.0;
    int blockOrder = 0;
    int waveletTest = 1;
    double waveletTuningRate = 0.0;
    int waveletAutoTuning = 0;
    double waveletBrFix = 1.0;
    double waveletMseFix = 1.0;
    double waveletBrFix2 = 1.0;
    double waveletMseFix2 = 1.0;
    double lorenzoBrFix = 1.0;
    int var_first = 0;
    size_t profStride = 0;
    int sperr = -1;
    int waveAutoFix = 1;
    int conditioning = 0;
    std::array<uint8_t, 17> meta{};
    int fixWave = -1;
    bool sperrWithoutWave = false;
    bool pyBind = true;
    std::string metadata;
    bool pybind_activated = false;
    bool FZ = false;

private:
    void init_dims() {
        coeffs_dims = dims;
        N = (char) dims.size();
        num = std::accumulate(dims.begin(), dims.end(), (size_t) 1, std::multiplies<size_t>());
        coeffs_num = num;
        blockSize = (N == 1 ? 128 : (N == 2 ? 16 : 6));
        pred_dim = N;
        stride = blockSize;
    }
};

}  // namespace QoZ
