// Golden-vector driver around the reference FAZ engine classes.
//
// Test infrastructure only: this file is compiled (by oracle/Makefile) against
// the reference headers where they lie and writes its binary into oracle/_ref/.
// It composes the reference classes exactly as the reference drivers do and
// dumps what the product must reproduce bit-for-bit:
//   * SZ_compress_Interp (include/QoZ/api/impl/SZInterp.hpp:387-414) +
//     SZ_compress trailer (include/QoZ/api/sz.hpp:54-91), restated here because
//     SZInterp.hpp pulls in PolyRegressionPredictor.hpp whose
//     PolyRegressionCoeffAux.hpp is absent from the reference tree;
//   * the pre-zstd engine buffer (same engine, Lossless_bypass);
//   * codes decoded from that buffer with the reference HuffmanEncoder;
//   * tuning modes 1/2 of SZInterpolationCompressor::compress (SIC:328-476);
//   * SZ_decompress for ALGO_INTERP (sz.hpp:159-205, SZInterp.hpp:417-431);
//   * the Huffman coder alone (HuffmanEncoder.hpp:95-227);
//   * the Lorenzo/regression compressors (SZLorenzoReg.hpp:27-114), fast and
//     general frontends, and CompressTest's Lorenzo variant.
#include "QoZ/utils/Config.hpp"
#include "QoZ/utils/Statistic.hpp"  // QoZ::factorial, used by the engine header below
#include "QoZ/compressor/SZInterpolationCompressor.hpp"
#include "QoZ/compressor/SZGeneralCompressor.hpp"
#include "QoZ/frontend/SZFastFrontend.hpp"
#include "QoZ/frontend/SZGeneralFrontend.hpp"
#include "QoZ/predictor/ComposedPredictor.hpp"
#include "QoZ/predictor/LorenzoPredictor.hpp"
#include "QoZ/predictor/RegressionPredictor.hpp"
#include "QoZ/quantizer/IntegerQuantizer.hpp"
#include "QoZ/encoder/HuffmanEncoder.hpp"
#include "QoZ/lossless/Lossless_zstd.hpp"
#include "QoZ/lossless/Lossless_bypass.hpp"
#include "QoZ/compressor/deprecated/SZBlockInterpolationCompressor.hpp"
#include "QoZ/utils/Extraction.hpp"
#include "QoZ/utils/QuantOptimization.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

typedef std::map<std::string, std::string> Args;

static Args parse_args(int argc, char **argv, int first) {
    Args a;
    for (int i = first; i < argc; i++) {
        std::string s(argv[i]);
        size_t eq = s.find('=');
        if (eq == std::string::npos) {
            fprintf(stderr, "bad argument %s\n", argv[i]);
            exit(2);
        }
        a[s.substr(0, eq)] = s.substr(eq + 1);
    }
    return a;
}

static std::string get(const Args &a, const char *k, const char *def = nullptr) {
    auto it = a.find(k);
    if (it == a.end()) {
        if (!def) {
            fprintf(stderr, "missing argument %s\n", k);
            exit(2);
        }
        return def;
    }
    return it->second;
}

static std::vector<size_t> parse_list(const std::string &s) {
    std::vector<size_t> v;
    size_t p = 0;
    while (p < s.size()) {
        size_t q = s.find(',', p);
        if (q == std::string::npos) q = s.size();
        if (q > p) v.push_back(strtoull(s.substr(p, q - p).c_str(), nullptr, 10));
        p = q + 1;
    }
    return v;
}

template<class T>
static std::vector<T> read_vec(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) { perror(path.c_str()); exit(2); }
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<T> v(sz / sizeof(T));
    if (!v.empty() && fread(v.data(), sizeof(T), v.size(), f) != v.size()) { perror("read"); exit(2); }
    fclose(f);
    return v;
}

static void write_bytes(const std::string &path, const void *p, size_t n) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(2); }
    if (n && fwrite(p, 1, n, f) != n) { perror("write"); exit(2); }
    fclose(f);
}

static QoZ::Config make_config(const Args &a) {
    std::vector<size_t> dims = parse_list(get(a, "dims"));
    QoZ::Config conf;
    conf.setDims(dims.begin(), dims.end());
    conf.cmprAlgo = QoZ::ALGO_INTERP;
    std::string mode = get(a, "mode", "abs");
    if (mode == "abs") {
        conf.errorBoundMode = QoZ::EB_ABS;
        conf.absErrorBound = atof(get(a, "eb").c_str());
    } else {
        conf.errorBoundMode = QoZ::EB_REL;
        conf.relErrorBound = atof(get(a, "eb").c_str());
    }
    conf.interpAlgo = (uint8_t) atoi(get(a, "algo", "1").c_str());
    conf.interpDirection = (uint8_t) atoi(get(a, "dir", "0").c_str());
    conf.alpha = atof(get(a, "alpha", "-1").c_str());
    conf.beta = atof(get(a, "beta", "-1").c_str());
    conf.maxStep = strtoull(get(a, "maxstep", "0").c_str(), nullptr, 10);
    conf.interpBlockSize = atoi(get(a, "bs", "32").c_str());
    conf.quantbinCnt = atoi(get(a, "qbins", "65536").c_str());
    std::vector<size_t> ops = parse_list(get(a, "ops", ""));
    std::vector<size_t> dirs = parse_list(get(a, "dirs", ""));
    conf.interpAlgo_list.assign(ops.begin(), ops.end());
    conf.interpDirection_list.assign(dirs.begin(), dirs.end());
    conf.levelwisePredictionSelection = (int) ops.size();
    conf.firstSize = 0;  // never initialised by the reference when wavelet==0
    conf.meta.fill(0);
    return conf;
}

// Header length of the interp engine buffer (SIC:489-512).
template<QoZ::uint N>
static size_t interp_header_len(size_t levels) {
    return N * sizeof(size_t) + sizeof(QoZ::uint) + 2 * sizeof(int) + 2 * sizeof(double) + sizeof(size_t) +
           sizeof(int) + sizeof(bool) + sizeof(int) + sizeof(size_t) + 2 * sizeof(int) + 2 * levels;
}

template<QoZ::uint N>
static int run_interp(const Args &a) {
    typedef float T;
    std::string out = get(a, "out");
    std::vector<T> input = read_vec<T>(get(a, "in"));
    QoZ::Config conf = make_config(a);
    if (input.size() != conf.num) { fprintf(stderr, "size mismatch\n"); return 2; }
    int tuning = atoi(get(a, "tuning", "0").c_str());
    int start = atoi(get(a, "start", "0").c_str());
    int end = atoi(get(a, "end", "0").c_str());

    if (tuning) {
        QoZ::calAbsErrorBound(conf, input.data());
        std::vector<T> data(input);
        auto sz = QoZ::SZInterpolationCompressor<T, N, QoZ::LinearQuantizer<T>, QoZ::HuffmanEncoder<int>, QoZ::Lossless_bypass>(
                QoZ::LinearQuantizer<T>(conf.absErrorBound, conf.quantbinCnt / 2), QoZ::HuffmanEncoder<int>(),
                QoZ::Lossless_bypass());
        size_t outSize = 0;
        conf.verbose = 0;
        QoZ::uchar *buf = sz.compress(conf, data.data(), outSize, tuning, start, end);
        delete[] buf;
        write_bytes(out + ".codes", conf.quant_bins.data(), conf.quant_bins.size() * sizeof(int));
        write_bytes(out + ".recon", data.data(), data.size() * sizeof(T));
        write_bytes(out + ".loss", &conf.decomp_square_error, sizeof(double));
        write_bytes(out + ".levelcounts", conf.quant_bin_counts.data(), conf.quant_bin_counts.size() * sizeof(size_t));
        return 0;
    }

    // Pre-zstd engine buffer: same engine with the bypass lossless stage.
    {
        QoZ::Config c2(conf);
        QoZ::calAbsErrorBound(c2, input.data());
        std::vector<T> data(input);
        auto sz = QoZ::SZInterpolationCompressor<T, N, QoZ::LinearQuantizer<T>, QoZ::HuffmanEncoder<int>, QoZ::Lossless_bypass>(
                QoZ::LinearQuantizer<T>(c2.absErrorBound, c2.quantbinCnt / 2), QoZ::HuffmanEncoder<int>(),
                QoZ::Lossless_bypass());
        size_t rawSize = 0;
        c2.verbose = 0;
        QoZ::uchar *raw = sz.compress(c2, data.data(), rawSize);
        write_bytes(out + ".raw", raw, rawSize);
        write_bytes(out + ".recon", data.data(), data.size() * sizeof(T));
        write_bytes(out + ".eb", &c2.absErrorBound, sizeof(double));
        // codes and unpredictable list, decoded with the reference classes
        size_t L = c2.interpAlgo_list.size();
        const QoZ::uchar *pos = raw + interp_header_len<N>(L);
        size_t U;
        memcpy(&U, pos + 1 + sizeof(double) + sizeof(int), sizeof(size_t));
        const QoZ::uchar *up = pos + 1 + sizeof(double) + sizeof(int) + sizeof(size_t);
        write_bytes(out + ".unpred", up, U * sizeof(T));
        pos = up + U * sizeof(T);
        size_t remaining = rawSize - (pos - raw);
        QoZ::HuffmanEncoder<int> enc;
        enc.load(pos, remaining);
        std::vector<int> codes = enc.decode(pos, c2.num);
        write_bytes(out + ".codes", codes.data(), codes.size() * sizeof(int));
        delete[] raw;
    }

    // Full stream as SZ_compress writes it: zstd(engine buffer) + Config trailer + int.
    {
        QoZ::Config c3(conf);
        std::vector<T> data(input);
        QoZ::calAbsErrorBound(c3, data.data());
        c3.verbose = 0;
        auto sz = QoZ::SZInterpolationCompressor<T, N, QoZ::LinearQuantizer<T>, QoZ::HuffmanEncoder<int>, QoZ::Lossless_zstd>(
                QoZ::LinearQuantizer<T>(c3.absErrorBound, c3.quantbinCnt / 2), QoZ::HuffmanEncoder<int>(),
                QoZ::Lossless_zstd());
        size_t outSize = 0;
        char *cmpData = (char *) sz.compress(c3, data.data(), outSize);
        QoZ::uchar *cmpDataPos = (QoZ::uchar *) cmpData + outSize;
        c3.save(cmpDataPos);
        size_t newSize = (char *) cmpDataPos - cmpData;
        QoZ::write(int(newSize - outSize), cmpDataPos);
        outSize = (char *) cmpDataPos - cmpData;
        write_bytes(out + ".stream", cmpData, outSize);
        delete[] cmpData;
    }
    return 0;
}

template<QoZ::uint N>
static int run_decompress_n(QoZ::Config &conf, char *cmpData, size_t cmpSize, float *dec) {
    auto sz = QoZ::SZInterpolationCompressor<float, N, QoZ::LinearQuantizer<float>, QoZ::HuffmanEncoder<int>, QoZ::Lossless_zstd>(
            QoZ::LinearQuantizer<float>(), QoZ::HuffmanEncoder<int>(), QoZ::Lossless_zstd());
    sz.decompress((QoZ::uchar *) cmpData, cmpSize, dec);
    return 0;
}

static int run_decompress(const Args &a) {
    std::vector<char> stream = read_vec<char>(get(a, "in"));
    int confSize;
    memcpy(&confSize, stream.data() + stream.size() - sizeof(int), sizeof(int));
    const QoZ::uchar *pos = (const QoZ::uchar *) stream.data() + (stream.size() - sizeof(int) - confSize);
    QoZ::Config conf;
    conf.load(pos);
    if (conf.cmprAlgo != QoZ::ALGO_INTERP) { fprintf(stderr, "only ALGO_INTERP streams\n"); return 2; }
    std::vector<float> dec(conf.num);
    size_t cmpSize = stream.size() - sizeof(int) - confSize;
    switch (conf.N) {
        case 1: run_decompress_n<1>(conf, stream.data(), cmpSize, dec.data()); break;
        case 2: run_decompress_n<2>(conf, stream.data(), cmpSize, dec.data()); break;
        case 3: run_decompress_n<3>(conf, stream.data(), cmpSize, dec.data()); break;
        case 4: run_decompress_n<4>(conf, stream.data(), cmpSize, dec.data()); break;
        default: return 2;
    }
    write_bytes(get(a, "out"), dec.data(), dec.size() * sizeof(float));
    return 0;
}

// One SZ_compress-equivalent ALGO_INTERP compression (range, engine, zstd,
// trailer) timed with a steady clock; prints seconds and the stream size.
template<QoZ::uint N>
static int run_time(const Args &a) {
    typedef float T;
    std::vector<T> input = read_vec<T>(get(a, "in"));
    QoZ::Config conf = make_config(a);
    if (input.size() != conf.num) { fprintf(stderr, "size mismatch\n"); return 2; }
    int reps = atoi(get(a, "reps", "1").c_str());
    double best = 1e30;
    size_t outSize = 0;
    for (int r = 0; r < reps; r++) {
        QoZ::Config c3(conf);
        c3.verbose = 0;
        auto t0 = std::chrono::steady_clock::now();
        std::vector<T> data(input);  // SZ_compress copies its input (sz.hpp:57)
        QoZ::calAbsErrorBound(c3, data.data());
        auto sz = QoZ::SZInterpolationCompressor<T, N, QoZ::LinearQuantizer<T>, QoZ::HuffmanEncoder<int>, QoZ::Lossless_zstd>(
                QoZ::LinearQuantizer<T>(c3.absErrorBound, c3.quantbinCnt / 2), QoZ::HuffmanEncoder<int>(),
                QoZ::Lossless_zstd());
        char *cmpData = (char *) sz.compress(c3, data.data(), outSize);
        QoZ::uchar *pos = (QoZ::uchar *) cmpData + outSize;
        c3.save(pos);
        size_t newSize = (char *) pos - cmpData;
        QoZ::write(int(newSize - outSize), pos);
        outSize = (char *) pos - cmpData;
        double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        delete[] cmpData;
        if (secs < best) best = secs;
    }
    printf("{\"seconds\": %.6f, \"bytes\": %zu, \"points\": %zu}\n", best, outSize, conf.num);
    return 0;
}

// Lorenzo/regression path for 3-D float fields as SZ_compress runs it with
// cmprAlgo = ALGO_LORENZO_REG (SZDispatcher.hpp:13-20, SZLorenzoReg.hpp:95-114):
// calAbsErrorBound, then SZGeneralCompressor over SZFastFrontend with a
// LinearQuantizer of radius quantbinCnt/2, HuffmanEncoder<int> and zstd.
// Dumps the frontend codes, the pre-zstd buffer (bypass lossless stage), the
// reconstruction decoded from it, the SZ_compress stream and the bound.
static int run_lorenzo(const Args &a) {
    typedef float T;
    std::string out = get(a, "out");
    std::vector<T> input = read_vec<T>(get(a, "in"));
    QoZ::Config conf = make_config(a);
    if (conf.N != 3 || input.size() != conf.num) { fprintf(stderr, "3-D field of dims size expected\n"); return 2; }
    conf.cmprAlgo = QoZ::ALGO_LORENZO_REG;
    conf.blockSize = atoi(get(a, "block", "6").c_str());
    conf.pred_dim = atoi(get(a, "pd", "3").c_str());
    conf.lorenzo = atoi(get(a, "lo", "1").c_str()) != 0;
    conf.lorenzo2 = atoi(get(a, "lo2", "0").c_str()) != 0;
    conf.regression = atoi(get(a, "reg", "1").c_str()) != 0;
    conf.regression2 = false;
    conf.verbose = 0;
    QoZ::calAbsErrorBound(conf, input.data());
    write_bytes(out + ".eb", &conf.absErrorBound, sizeof(double));
    {
        QoZ::Config c(conf);
        std::vector<T> data(input);
        auto fe = QoZ::make_sz_fast_frontend<T, 3>(c, QoZ::LinearQuantizer<T>(c.absErrorBound, c.quantbinCnt / 2));
        std::vector<int> codes = fe.compress(data.data());
        write_bytes(out + ".codes", codes.data(), codes.size() * sizeof(int));
    }
    {
        QoZ::Config c(conf);
        std::vector<T> data(input);
        auto *sz = QoZ::make_sz_general_compressor<T, 3>(
                QoZ::make_sz_fast_frontend<T, 3>(c, QoZ::LinearQuantizer<T>(c.absErrorBound, c.quantbinCnt / 2)),
                QoZ::HuffmanEncoder<int>(), QoZ::Lossless_bypass());
        size_t rawSize = 0;
        QoZ::uchar *raw = sz->compress(c, data.data(), rawSize);
        write_bytes(out + ".raw", raw, rawSize);
        delete sz;
        QoZ::Config c2(conf);
        QoZ::LinearQuantizer<T> q;
        auto *dz = QoZ::make_sz_general_compressor<T, 3>(QoZ::make_sz_fast_frontend<T, 3>(c2, q),
                                                          QoZ::HuffmanEncoder<int>(), QoZ::Lossless_bypass());
        std::vector<T> dec(conf.num);
        dz->decompress(raw, rawSize, dec.data());
        write_bytes(out + ".recon", dec.data(), dec.size() * sizeof(T));
        delete dz;
        delete[] raw;
    }
    {
        QoZ::Config c(conf);
        std::vector<T> data(input);
        auto *sz = QoZ::make_sz_general_compressor<T, 3>(
                QoZ::make_sz_fast_frontend<T, 3>(c, QoZ::LinearQuantizer<T>(c.absErrorBound, c.quantbinCnt / 2)),
                QoZ::HuffmanEncoder<int>(), QoZ::Lossless_zstd());
        size_t outSize = 0;
        char *cmpData = (char *) sz->compress(c, data.data(), outSize);
        QoZ::uchar *pos = (QoZ::uchar *) cmpData + outSize;
        c.save(pos);
        size_t newSize = (char *) pos - cmpData;
        QoZ::write(int(newSize - outSize), pos);
        outSize = (char *) pos - cmpData;
        write_bytes(out + ".stream", cmpData, outSize);
        delete sz;
        delete[] cmpData;
    }
    return 0;
}

// CompressTest for ALGO_INTERP with the RD target (SZInterp.hpp:1131-1392),
// composed from the reference classes it uses: one SZInterpolationCompressor
// (LinearQuantizer(eb), HuffmanEncoder<int>, Lossless_zstd) compresses every
// block in tuning mode 1, the codes are interleaved level by level across the
// blocks and encoding_lossless (SIC:784-809) gives the compressed size.
// blocks: a file of int64 start triples; prints bytes and the squared error.
template<QoZ::uint N>
static int run_ctest(const Args &a) {
    typedef float T;
    std::vector<T> input = read_vec<T>(get(a, "in"));
    std::vector<size_t> fdims = parse_list(get(a, "fdims"));
    std::vector<int64_t> starts = read_vec<int64_t>(get(a, "starts"));
    size_t sbs = strtoull(get(a, "sbs").c_str(), nullptr, 10);
    QoZ::Config conf = make_config(a);  // dims = block shape
    conf.verbose = 0;
    size_t per = conf.num;
    std::vector<std::vector<T>> blocks;
    for (size_t b = 0; b + 2 < starts.size(); b += 3) {
        std::vector<T> blk;
        if (N == 2) {
            for (size_t i = 0; i <= sbs; i++)
                for (size_t j = 0; j <= sbs; j++) blk.push_back(input[(starts[b] + i) * fdims[1] + starts[b + 1] + j]);
        } else {
            for (size_t i = 0; i <= sbs; i++)
                for (size_t j = 0; j <= sbs; j++)
                    for (size_t k = 0; k <= sbs; k++)
                        blk.push_back(input[((starts[b] + i) * fdims[1] + starts[b + 1] + j) * fdims[2] + starts[b + 2] + k]);
        }
        blocks.push_back(blk);
    }
    auto sz = QoZ::SZInterpolationCompressor<T, N, QoZ::LinearQuantizer<T>, QoZ::HuffmanEncoder<int>, QoZ::Lossless_zstd>(
            QoZ::LinearQuantizer<T>(conf.absErrorBound), QoZ::HuffmanEncoder<int>(), QoZ::Lossless_zstd());
    double square_error = 0;
    std::vector<std::vector<int>> bq;
    for (auto &blk: blocks) {
        std::vector<T> cur(blk);
        size_t outSize = 0;
        QoZ::uchar *d = sz.compress(conf, cur.data(), outSize, 1);
        delete[] d;
        bq.push_back(conf.quant_bins);
        square_error += conf.decomp_square_error;
    }
    std::vector<size_t> counts = conf.quant_bin_counts;
    std::vector<int> q;
    size_t last = 0;
    for (int k = (int) counts.size() - 1; k >= 0; k--) {
        for (size_t l = 0; l < bq.size(); l++)
            for (size_t m = last; m < counts[k]; m++) q.push_back(bq[l][m]);
        last = counts[k];
    }
    size_t total = 0;
    QoZ::uchar *c = sz.encoding_lossless(total, q);
    if (a.count("raw")) {  // the pre-zstd buffer, recovered from the zstd frame
        size_t len = total;
        QoZ::Lossless_zstd z;
        QoZ::uchar *raw = z.decompress(c, len);
        write_bytes(get(a, "raw"), raw, len);
        delete[] raw;
    }
    delete[] c;
    printf("{\"bytes\": %zu, \"square_error\": %.17g, \"points\": %zu}\n", total, square_error, per * blocks.size());
    return 0;
}

// make_lorenzo_regression_compressor (SZLorenzoReg.hpp:27-91) without the
// second-order regression: the one enabled predictor alone, else a
// ComposedPredictor over the enabled ones in the order 1-layer Lorenzo,
// 2-layer Lorenzo, RegressionPredictor, in an SZGeneralFrontend.
template<QoZ::uint N, class Lossless>
static QoZ::concepts::CompressorInterface<float> *general_compressor(const QoZ::Config &conf, int radius,
                                                                      Lossless lossless) {
    typedef float T;
    auto q = QoZ::LinearQuantizer<T>(conf.absErrorBound, radius);
    const int count = conf.lorenzo + conf.lorenzo2 + conf.regression;
    if (count > 1) {
        std::vector<std::shared_ptr<QoZ::concepts::PredictorInterface<T, N>>> ps;
        if (conf.lorenzo) ps.push_back(std::make_shared<QoZ::LorenzoPredictor<T, N, 1>>(conf.absErrorBound));
        if (conf.lorenzo2) ps.push_back(std::make_shared<QoZ::LorenzoPredictor<T, N, 2>>(conf.absErrorBound));
        if (conf.regression)
            ps.push_back(std::make_shared<QoZ::RegressionPredictor<T, N>>(conf.blockSize, conf.absErrorBound));
        return QoZ::make_sz_general_compressor<T, N>(
                QoZ::make_sz_general_frontend<T, N>(conf, QoZ::ComposedPredictor<T, N>(ps), q),
                QoZ::HuffmanEncoder<int>(), lossless);
    }
    if (conf.regression)
        return QoZ::make_sz_general_compressor<T, N>(
                QoZ::make_sz_general_frontend<T, N>(
                        conf, QoZ::RegressionPredictor<T, N>(conf.blockSize, conf.absErrorBound), q),
                QoZ::HuffmanEncoder<int>(), lossless);
    if (conf.lorenzo2)
        return QoZ::make_sz_general_compressor<T, N>(
                QoZ::make_sz_general_frontend<T, N>(conf, QoZ::LorenzoPredictor<T, N, 2>(conf.absErrorBound), q),
                QoZ::HuffmanEncoder<int>(), lossless);
    return QoZ::make_sz_general_compressor<T, N>(
            QoZ::make_sz_general_frontend<T, N>(conf, QoZ::LorenzoPredictor<T, N, 1>(conf.absErrorBound), q),
            QoZ::HuffmanEncoder<int>(), lossless);
}

// SZ_compress_LorenzoReg through the general frontend (LR:95-114 else-branch):
// codes, pre-zstd buffer, reconstruction decoded from it, SZ_compress stream.
template<QoZ::uint N>
static int run_general(const Args &a) {
    typedef float T;
    std::string out = get(a, "out");
    std::vector<T> input = read_vec<T>(get(a, "in"));
    QoZ::Config conf = make_config(a);
    if (input.size() != conf.num) { fprintf(stderr, "size mismatch\n"); return 2; }
    conf.cmprAlgo = QoZ::ALGO_LORENZO_REG;
    if (a.count("block")) conf.blockSize = atoi(get(a, "block").c_str());
    conf.lorenzo = atoi(get(a, "lo", "1").c_str()) != 0;
    conf.lorenzo2 = atoi(get(a, "lo2", "1").c_str()) != 0;
    conf.regression = atoi(get(a, "reg", "0").c_str()) != 0;
    conf.regression2 = false;
    conf.verbose = 0;
    QoZ::calAbsErrorBound(conf, input.data());
    write_bytes(out + ".eb", &conf.absErrorBound, sizeof(double));
    const int radius = conf.quantbinCnt / 2;
    {
        QoZ::Config c(conf);
        std::vector<T> data(input);
        auto *sz = general_compressor<N>(c, radius, QoZ::Lossless_bypass());
        size_t rawSize = 0;
        QoZ::uchar *raw = sz->compress(c, data.data(), rawSize);
        write_bytes(out + ".raw", raw, rawSize);
        write_bytes(out + ".overwrite", data.data(), data.size() * sizeof(T));
        delete sz;
        QoZ::Config c2(conf);
        auto *dz = general_compressor<N>(c2, radius, QoZ::Lossless_bypass());
        std::vector<T> dec(conf.num);
        dz->decompress(raw, rawSize, dec.data());
        write_bytes(out + ".recon", dec.data(), dec.size() * sizeof(T));
        delete dz;
        delete[] raw;
    }
    {
        QoZ::Config c(conf);
        std::vector<T> data(input);
        auto *sz = general_compressor<N>(c, radius, QoZ::Lossless_zstd());
        size_t outSize = 0;
        char *cmpData = (char *) sz->compress(c, data.data(), outSize);
        QoZ::uchar *pos = (QoZ::uchar *) cmpData + outSize;
        c.save(pos);
        size_t newSize = (char *) pos - cmpData;
        QoZ::write(int(newSize - outSize), pos);
        outSize = (char *) pos - cmpData;
        write_bytes(out + ".stream", cmpData, outSize);
        delete sz;
        delete[] cmpData;
    }
    return 0;
}

// CompressTest for ALGO_LORENZO_REG with useFast = false (SZI:1136-1170,
// 1254-1256, 1283-1294, 1353-1392): one general compressor (lorenzo +
// lorenzo2, blockSize 5, quantbinCnt 131072) compresses every sample block in
// tuning mode, squared error from the overwritten blocks, then
// encoding_lossless over the accumulated codes.
template<QoZ::uint N>
static int run_ctest_lorenzo(const Args &a) {
    typedef float T;
    std::vector<T> input = read_vec<T>(get(a, "in"));
    std::vector<size_t> fdims = parse_list(get(a, "fdims"));
    std::vector<int64_t> starts = read_vec<int64_t>(get(a, "starts"));
    size_t sbs = strtoull(get(a, "sbs").c_str(), nullptr, 10);
    QoZ::Config conf = make_config(a);  // dims = block shape
    conf.verbose = 0;
    conf.cmprAlgo = QoZ::ALGO_LORENZO_REG;
    conf.lorenzo = true;
    conf.lorenzo2 = true;
    conf.regression = false;
    conf.regression2 = false;
    conf.openmp = false;
    conf.blockSize = 5;
    conf.quantbinCnt = 65536 * 2;
    size_t per = conf.num;
    std::vector<std::vector<T>> blocks;
    for (size_t b = 0; b + 2 < starts.size(); b += 3) {
        std::vector<T> blk;
        if (N == 2) {
            for (size_t i = 0; i <= sbs; i++)
                for (size_t j = 0; j <= sbs; j++) blk.push_back(input[(starts[b] + i) * fdims[1] + starts[b + 1] + j]);
        } else {
            for (size_t i = 0; i <= sbs; i++)
                for (size_t j = 0; j <= sbs; j++)
                    for (size_t k = 0; k <= sbs; k++)
                        blk.push_back(input[((starts[b] + i) * fdims[1] + starts[b + 1] + j) * fdims[2] + starts[b + 2] + k]);
        }
        blocks.push_back(blk);
    }
    auto *sz = general_compressor<N>(conf, conf.quantbinCnt / 2, QoZ::Lossless_zstd());
    double square_error = 0;
    for (auto &blk: blocks) {
        std::vector<T> cur(blk);
        size_t outSize = 0;
        QoZ::uchar *d = sz->compress(conf, cur.data(), outSize, 1);
        delete[] d;
        for (size_t j = 0; j < per; j++) {
            T value = blk[j] - cur[j];
            square_error += value * value;
        }
    }
    size_t total = 0;
    QoZ::uchar *c = sz->encoding_lossless(total);
    if (a.count("raw")) {
        size_t len = total;
        QoZ::Lossless_zstd z;
        QoZ::uchar *raw = z.decompress(c, len);
        write_bytes(get(a, "raw"), raw, len);
        delete[] raw;
    }
    delete[] c;
    delete sz;
    printf("{\"bytes\": %zu, \"square_error\": %.17g, \"points\": %zu}\n", total, square_error, per * blocks.size());
    return 0;
}

// sampling<T,N> (Extraction.hpp): the sample, its shape and block.
template<QoZ::uint N>
static int run_sampling(const Args &a) {
    typedef float T;
    std::vector<T> input = read_vec<T>(get(a, "in"));
    std::vector<size_t> dims = parse_list(get(a, "dims"));
    size_t num = 0, block = 0;
    std::vector<size_t> sdims(N);
    std::vector<T> s = QoZ::sampling<T, N>(input.data(), dims, num, sdims, block);
    write_bytes(get(a, "out"), s.data(), s.size() * sizeof(T));
    printf("{\"sampled\": %d, \"sampling_block\": %zu, \"sample_num\": %zu, \"sample_dims\": [", s.empty() ? 0 : 1,
           block, num);
    for (QoZ::uint i = 0; i < N; i++) printf("%s%zu", i ? ", " : "", sdims[i]);
    printf("]}\n");
    return 0;
}

// do_not_use_this_interp_compress_block_test (SZInterp.hpp:558-576),
// composed from the same classes: the block interpolation compressor's
// buffer (Lossless_bypass) and its zstd-framed size.
template<QoZ::uint N>
static int run_blockinterp(const Args &a) {
    typedef float T;
    std::vector<T> input = read_vec<T>(get(a, "in"));
    std::vector<size_t> dims = parse_list(get(a, "dims"));
    double eb = atof(get(a, "eb").c_str());
    QoZ::Config conf;
    conf.absErrorBound = eb;
    conf.setDims(dims.begin(), dims.end());
    conf.blockSize = atoi(get(a, "block").c_str());
    conf.interpAlgo = (uint8_t) atoi(get(a, "algo").c_str());
    conf.interpDirection = (uint8_t) atoi(get(a, "dir").c_str());
    size_t bytes = 0;
    {
        std::vector<T> d(input);
        auto sz = QoZ::SZBlockInterpolationCompressor<T, N, QoZ::LinearQuantizer<T>, QoZ::HuffmanEncoder<int>,
                                                      QoZ::Lossless_zstd>(QoZ::LinearQuantizer<T>(eb),
                                                                          QoZ::HuffmanEncoder<int>(),
                                                                          QoZ::Lossless_zstd());
        QoZ::uchar *c = sz.compress(conf, d.data(), bytes);
        delete[] c;
    }
    if (a.count("raw")) {
        std::vector<T> d(input);
        auto sz = QoZ::SZBlockInterpolationCompressor<T, N, QoZ::LinearQuantizer<T>, QoZ::HuffmanEncoder<int>,
                                                      QoZ::Lossless_bypass>(QoZ::LinearQuantizer<T>(eb),
                                                                            QoZ::HuffmanEncoder<int>(),
                                                                            QoZ::Lossless_bypass());
        size_t n = 0;
        QoZ::uchar *c = sz.compress(conf, d.data(), n);
        write_bytes(get(a, "raw"), c, n);
        delete[] c;
    }
    printf("{\"bytes\": %zu}\n", bytes);
    return 0;
}

static int run_quantinvl(const Args &a) {
    std::vector<float> input = read_vec<float>(get(a, "in"));
    std::vector<size_t> dims = parse_list(get(a, "dims"));
    double eb = atof(get(a, "eb").c_str());
    int q = QoZ::optimize_quant_invl_3d<float>(input.data(), dims[0], dims[1], dims[2], eb);
    printf("{\"quantbin\": %d}\n", q);
    return 0;
}

static int run_huffman(const Args &a) {
    std::vector<int> codes = read_vec<int>(get(a, "in"));
    QoZ::HuffmanEncoder<int> enc;
    enc.preprocess_encode(codes, 0);
    std::vector<QoZ::uchar> buf(64 + enc.size_est() + codes.size() * 8 + 64);
    QoZ::uchar *pos = buf.data();
    enc.save(pos);
    enc.encode(codes, pos);
    enc.postprocess_encode();
    write_bytes(get(a, "out"), buf.data(), pos - buf.data());
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: ref_harness interp|decompress|huffman key=value...\n");
        return 2;
    }
    std::string cmd(argv[1]);
    Args a = parse_args(argc, argv, 2);
    if (cmd == "interp") {
        size_t N = parse_list(get(a, "dims")).size();
        switch (N) {
            case 1: return run_interp<1>(a);
            case 2: return run_interp<2>(a);
            case 3: return run_interp<3>(a);
            case 4: return run_interp<4>(a);
            default: return 2;
        }
    }
    if (cmd == "time") {
        switch (parse_list(get(a, "dims")).size()) {
            case 1: return run_time<1>(a);
            case 2: return run_time<2>(a);
            case 3: return run_time<3>(a);
            case 4: return run_time<4>(a);
            default: return 2;
        }
    }
    if (cmd == "decompress") return run_decompress(a);
    if (cmd == "huffman") return run_huffman(a);
    if (cmd == "lorenzo") return run_lorenzo(a);
    if (cmd == "ctest") {
        bool lor = get(a, "ctalgo", "interp") == "lorenzo";
        if (parse_list(get(a, "dims")).size() == 2) return lor ? run_ctest_lorenzo<2>(a) : run_ctest<2>(a);
        return lor ? run_ctest_lorenzo<3>(a) : run_ctest<3>(a);
    }
    if (cmd == "sampling" || cmd == "blockinterp") {
        const bool s = cmd == "sampling";
        switch (parse_list(get(a, "dims")).size()) {
            case 1: return s ? run_sampling<1>(a) : run_blockinterp<1>(a);
            case 2: return s ? run_sampling<2>(a) : run_blockinterp<2>(a);
            case 3: return s ? run_sampling<3>(a) : run_blockinterp<3>(a);
            default: return 2;
        }
    }
    if (cmd == "quantinvl") return run_quantinvl(a);
    if (cmd == "general") {
        switch (parse_list(get(a, "dims")).size()) {
            case 1: return run_general<1>(a);
            case 2: return run_general<2>(a);
            case 3: return run_general<3>(a);
            case 4: return run_general<4>(a);
            default: return 2;
        }
    }
    fprintf(stderr, "unknown command %s\n", argv[1]);
    return 2;
}
