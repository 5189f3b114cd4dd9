// Stream layer of the MI355X build: SZ_compress / SZ_decompress
// (include/QoZ/api/sz.hpp:54-205), the ALGO_INTERP driver
// (SZInterp.hpp:387-431), calAbsErrorBound (Statistic.hpp:34-66), the zstd
// wrapper (Lossless_zstd.hpp:22-47) and the Config trailer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "faz/config.hpp"
#include "faz/sz.hpp"
#include "faz_gpu.h"
#include "raw_api.h"

extern "C" {
// libzstd (system shared library libzstd.so.1); prototypes as published in zstd.h
size_t ZSTD_compress(void *dst, size_t dstCapacity, const void *src, size_t srcSize, int compressionLevel);
size_t ZSTD_decompress(void *dst, size_t dstCapacity, const void *src, size_t compressedSize);
size_t ZSTD_compressBound(size_t srcSize);
unsigned ZSTD_isError(size_t code);
}

namespace {

constexpr int kZstdLevel = 3;  // Lossless_zstd::compression_level

// zstd stage: size_t raw length, then the zstd frame (Lossless_zstd.hpp:22-31)
bool zstd_wrap(const std::vector<uint8_t> &raw, std::vector<uint8_t> &out, size_t reserve_tail) {
    out.resize(8 + ZSTD_compressBound(raw.size()) + reserve_tail);
    uint64_t len = raw.size();
    memcpy(out.data(), &len, 8);
    size_t z = ZSTD_compress(out.data() + 8, out.size() - 8, raw.data(), raw.size(), kZstdLevel);
    if (ZSTD_isError(z)) return false;
    out.resize(8 + z);
    return true;
}

bool zstd_unwrap(const uint8_t *p, size_t n, std::vector<uint8_t> &raw) {
    if (n < 8) return false;
    uint64_t len;
    memcpy(&len, p, 8);
    if (len > (1ull << 40)) return false;
    raw.resize(len);
    size_t z = ZSTD_decompress(raw.data(), len, p + 8, n - 8);
    return !ZSTD_isError(z) && z == len;
}

void append_trailer(const QoZ::Config &conf, std::vector<uint8_t> &out) {
    size_t before = out.size();
    out.resize(before + QoZ::Config::size_est() + conf.metadata.size() + 8 * conf.coeffs_dims.size() + 64);
    QoZ::uchar *w = out.data() + before;
    conf.save(w);
    int confSize = (int) (w - (out.data() + before));
    memcpy(w, &confSize, 4);
    w += 4;
    out.resize(w - out.data());
}

// The Config as SZ_compress stores it for an ALGO_INTERP field.
QoZ::Config interp_config(const faz_interp_params *p, double absEB, double relEB) {
    QoZ::Config c;
    c.setDims(p->dims, p->dims + p->N);
    c.cmprAlgo = QoZ::ALGO_INTERP;
    c.errorBoundMode = QoZ::EB_ABS;
    c.absErrorBound = absEB;
    c.relErrorBound = relEB;
    c.alpha = p->alpha;
    c.beta = p->beta;
    c.interpAlgo = (uint8_t) p->interp_algo;
    c.interpDirection = (uint8_t) p->direction;
    c.interpBlockSize = p->block_size;
    c.quantbinCnt = p->radius * 2;
    c.levelwisePredictionSelection = p->levelwise;
    c.interpAlgo_list.assign(p->ops, p->ops + p->levelwise);
    c.interpDirection_list.assign(p->dirs, p->dirs + p->levelwise);
    c.maxStep = p->max_step;
    return c;
}

faz_interp_params params_from_config(const QoZ::Config &c) {
    faz_interp_params p;
    memset(&p, 0, sizeof(p));
    p.N = c.N;
    for (int i = 0; i < c.N && i < 4; i++) p.dims[i] = c.dims[i];
    p.eb = c.absErrorBound;
    p.interp_algo = c.interpAlgo;
    p.direction = c.interpDirection;
    p.alpha = c.alpha;
    p.beta = c.beta;
    p.max_step = c.maxStep;
    p.block_size = c.interpBlockSize;
    p.radius = c.quantbinCnt / 2;
    p.levelwise = (int) std::min<size_t>(c.interpAlgo_list.size(), 32);
    for (int i = 0; i < p.levelwise; i++) {
        p.ops[i] = c.interpAlgo_list[i];
        p.dirs[i] = i < (int) c.interpDirection_list.size() ? c.interpDirection_list[i] : 0;
    }
    return p;
}

// calAbsErrorBound (Statistic.hpp:34-66) with the value range from the GPU
int cal_abs_error_bound(faz_gpu_ctx *ctx, QoZ::Config &conf, const float *d_data) {
    if (conf.errorBoundMode == QoZ::EB_ABS) return FAZ_OK;
    auto range = [&](double &rng) -> int {
        float mn, mx;
        int rc = faz_gpu_range_f32(ctx, d_data, conf.num, &mn, &mx);
        if (rc) return rc;
        rng = (double) (float) (mx - mn);
        conf.rng = rng;
        return FAZ_OK;
    };
    double rng;
    int rc;
    switch (conf.errorBoundMode) {
        case QoZ::EB_REL:
            conf.errorBoundMode = QoZ::EB_ABS;
            if ((rc = range(rng))) return rc;
            conf.absErrorBound = conf.relErrorBound * rng;
            return FAZ_OK;
        case QoZ::EB_PSNR: {
            conf.errorBoundMode = QoZ::EB_ABS;
            if ((rc = range(rng))) return rc;
            double v1 = conf.psnrErrorBound + 10 * std::log10(1 - 2.0 / 3.0 * 0.99);
            double v2 = v1 / (-20);
            double v3 = std::pow(10, v2);
            conf.absErrorBound = rng * v3;
            conf.relErrorBound = conf.absErrorBound / rng;
            return FAZ_OK;
        }
        case QoZ::EB_L2NORM:
            conf.errorBoundMode = QoZ::EB_ABS;
            conf.absErrorBound = std::sqrt(3.0 / conf.num) * conf.l2normErrorBound;
            return FAZ_OK;
        case QoZ::EB_ABS_AND_REL:
            conf.errorBoundMode = QoZ::EB_ABS;
            if ((rc = range(rng))) return rc;
            conf.absErrorBound = std::min(conf.absErrorBound, conf.relErrorBound * rng);
            return FAZ_OK;
        case QoZ::EB_ABS_OR_REL:
            conf.errorBoundMode = QoZ::EB_ABS;
            if ((rc = range(rng))) return rc;
            conf.absErrorBound = std::max(conf.absErrorBound, conf.relErrorBound * rng);
            return FAZ_OK;
        default:
            return FAZ_ERR_INVALID;
    }
}

// Stages host data on the device (no copy when already there).
const float *to_device(faz_gpu_ctx *ctx, const float *data, size_t n, bool on_device, int &rc) {
    rc = FAZ_OK;
    if (on_device) return data;
    float *d = faz_ctx_work_in(ctx, n);
    if (!d) {
        rc = FAZ_ERR_HIP;
        return nullptr;
    }
    if (hipMemcpyAsync(d, data, n * sizeof(float), hipMemcpyHostToDevice, faz_ctx_stream(ctx)) != hipSuccess ||
        hipStreamSynchronize(faz_ctx_stream(ctx)) != hipSuccess) {
        rc = FAZ_ERR_HIP;
        return nullptr;
    }
    return d;
}

// Everything SZ_compress does for one ALGO_INTERP field after the Config copy.
int compress_interp_stream(faz_gpu_ctx *ctx, QoZ::Config &conf, const float *d_data, std::vector<uint8_t> &out) {
    int rc = cal_abs_error_bound(ctx, conf, d_data);
    if (rc) return rc;
    faz_interp_params p = params_from_config(conf);
    if (conf.fixBlockSize || conf.blockwiseTuning || conf.crossBlock || conf.trimToZero || conf.blockOrder) {
        faz_ctx_set_error(ctx, "engine mode outside the GPU path (fixBlockSize/blockwiseTuning/crossBlock/trimToZero/blockOrder)");
        return FAZ_ERR_UNSUPPORTED;
    }
    std::vector<uint8_t> raw;
    rc = faz_compress_raw_vec(ctx, d_data, nullptr, &p, raw);
    if (rc) return rc;
    if (!zstd_wrap(raw, out, QoZ::Config::size_est())) {
        faz_ctx_set_error(ctx, "zstd compression failed");
        return FAZ_ERR_INVALID;
    }
    append_trailer(conf, out);
    return FAZ_OK;
}

faz_lorenzo_params lorenzo_params_from_config(const QoZ::Config &c) {
    faz_lorenzo_params p;
    memset(&p, 0, sizeof(p));
    for (int i = 0; i < 3; i++) p.dims[i] = c.dims[i];
    p.eb = c.absErrorBound;
    p.block_size = c.blockSize;
    p.pred_dim = c.pred_dim;
    p.lorenzo = c.lorenzo;
    p.lorenzo2 = c.lorenzo2;
    p.regression = c.regression;
    p.radius = c.quantbinCnt / 2;
    return p;
}

// SZ_compress for ALGO_LORENZO_REG on a 3-D field: calAbsErrorBound
// (SZDispatcher.hpp:13-20), SZ_compress_LorenzoReg's fast frontend path
// (SZLorenzoReg.hpp:95-110), zstd, trailer.
faz_general_params general_params_from_config(const QoZ::Config &c) {
    faz_general_params p;
    memset(&p, 0, sizeof(p));
    p.N = c.N;
    for (int i = 0; i < c.N; i++) p.dims[i] = c.dims[i];
    p.eb = c.absErrorBound;
    p.block_size = c.blockSize;
    p.lorenzo = c.lorenzo;
    p.lorenzo2 = c.lorenzo2;
    p.regression = c.regression;
    p.radius = c.quantbinCnt / 2;
    return p;
}

int compress_lorenzo_stream(faz_gpu_ctx *ctx, QoZ::Config &conf, const float *d_data, std::vector<uint8_t> &out) {
    // SZ_compress_LorenzoReg (LR:95-114): the fast frontend for 3-D fields
    // without regression2, else the general frontend
    const bool fast = conf.N == 3 && !conf.regression2;
    if (conf.regression2) {
        faz_ctx_set_error(ctx, "the PolyRegressionPredictor (regression2) is outside the GPU path");
        return FAZ_ERR_UNSUPPORTED;
    }
    if (conf.wavelet) {
        faz_ctx_set_error(ctx, "wavelet preprocessing is outside the GPU path");
        return FAZ_ERR_UNSUPPORTED;
    }
    int rc = cal_abs_error_bound(ctx, conf, d_data);
    if (rc) return rc;
    std::vector<uint8_t> raw;
    if (fast) {
        faz_lorenzo_params p = lorenzo_params_from_config(conf);
        rc = faz_lorenzo_raw_vec(ctx, d_data, nullptr, &p, raw);
    } else {
        faz_general_params p = general_params_from_config(conf);
        rc = faz_general_raw_vec(ctx, d_data, nullptr, &p, raw);
    }
    if (rc) return rc;
    if (!zstd_wrap(raw, out, QoZ::Config::size_est())) {
        faz_ctx_set_error(ctx, "zstd compression failed");
        return FAZ_ERR_INVALID;
    }
    append_trailer(conf, out);
    return FAZ_OK;
}

// SZ_compress_Interp_lorenzo with conf.QoZ set (SZInterp.hpp:2562-2899): the
// exhaustive QoZ tuner (Tuning, SZI:1589-2557) on the GPU, then the
// interpolation engine with the tuned Config.
// SZ_compress_Interp_lorenzo in the default mode (SZI:2562-2899 with
// predictorTuningRate = autoTuningRate = 0): the SZ3 tuning on the GPU
// (faz_gpu_sz3_tune_f32), then SZ_compress_Interp or, when Lorenzo wins,
// SZ_compress_LorenzoReg with the Config the Lorenzo branch leaves (its
// setDims calls restore blockSize and pred_dim to the defaults for N; the bin
// count of its tests stays).
int compress_sz3_stream(faz_gpu_ctx *ctx, QoZ::Config &conf, const float *d_data, std::vector<uint8_t> &out) {
    if (conf.N > 3 || conf.wavelet > 0 || conf.autoTuningRate > 0 || conf.predictorTuningRate > 0) {
        faz_ctx_set_error(ctx, "the default tuning mode runs on 1- to 3-D fields without wavelets and with zero "
                               "tuning rates (set conf.QoZ for the QoZ tuner)");
        return FAZ_ERR_UNSUPPORTED;
    }
    int rc = cal_abs_error_bound(ctx, conf, d_data);
    if (rc) return rc;
    if (conf.rng < 0) {  // data_range (Statistic.hpp:11-21)
        float mn, mx;
        if ((rc = faz_gpu_range_f32(ctx, d_data, conf.num, &mn, &mx))) return rc;
        conf.rng = (double) (float) (mx - mn);
    }
    if (conf.relErrorBound <= 0) conf.relErrorBound = conf.absErrorBound / conf.rng;
    size_t dims[4] = {1, 1, 1, 1};
    for (int i = 0; i < conf.N; i++) dims[i] = conf.dims[i];
    faz_sz3_out o;
    if ((rc = faz_gpu_sz3_tune_f32(ctx, d_data, conf.N, dims, conf.absErrorBound, conf.relErrorBound, &o))) return rc;
    conf.interpAlgo = (uint8_t) o.interp_algo;
    conf.interpDirection = (uint8_t) o.direction;
    if (o.use_interp) {
        conf.cmprAlgo = QoZ::ALGO_INTERP;
        return compress_interp_stream(ctx, conf, d_data, out);
    }
    conf.cmprAlgo = QoZ::ALGO_LORENZO_REG;
    conf.lorenzo = true;
    conf.lorenzo2 = true;
    conf.regression = false;
    conf.regression2 = false;
    conf.openmp = false;
    conf.blockSize = conf.N == 1 ? 128 : (conf.N == 2 ? 16 : 6);
    conf.pred_dim = conf.N;
    conf.quantbinCnt = o.quantbin;
    return compress_lorenzo_stream(ctx, conf, d_data, out);
}

int compress_qoz_stream(faz_gpu_ctx *ctx, QoZ::Config &conf, const float *d_data, std::vector<uint8_t> &out) {
    if (!conf.QoZ) {
        if (conf.FZ) {
            faz_ctx_set_error(ctx, "the FZ mode (wavelet tuning) is outside the GPU path");
            return FAZ_ERR_UNSUPPORTED;
        }
        return compress_sz3_stream(ctx, conf, d_data, out);
    }
    if ((conf.N != 2 && conf.N != 3) || conf.wavelet > 0 || conf.tuningTarget != QoZ::TUNING_TARGET_RD) {
        faz_ctx_set_error(ctx, "the GPU QoZ tuner serves 2-D and 3-D fields, no wavelets, the PSNR target");
        return FAZ_ERR_UNSUPPORTED;
    }
    int rc = cal_abs_error_bound(ctx, conf, d_data);
    if (rc) return rc;
    if (conf.rng < 0) {  // data_range (Statistic.hpp:11-21)
        float mn, mx;
        if ((rc = faz_gpu_range_f32(ctx, d_data, conf.num, &mn, &mx))) return rc;
        conf.rng = (double) (float) (mx - mn);
    }
    if (conf.relErrorBound <= 0) conf.relErrorBound = conf.absErrorBound / conf.rng;
    // the QoZ settings Tuning applies (SZI:1594-1613)
    const int N = conf.N;
    if (conf.autoTuningRate <= 0) conf.autoTuningRate = N == 2 ? 0.01 : 0.005;
    if (conf.predictorTuningRate <= 0) conf.predictorTuningRate = N == 2 ? 0.01 : 0.005;
    if (conf.maxStep <= 0) conf.maxStep = N == 2 ? 64 : 32;
    if (conf.levelwisePredictionSelection <= 0) conf.levelwisePredictionSelection = N == 2 ? 6 : 4;
    if (conf.sampleBlockSize <= 0) conf.sampleBlockSize = N == 2 ? 64 : 32;
    if (conf.autoTuningRate != (N == 2 ? 0.01 : 0.005) || conf.predictorTuningRate != conf.autoTuningRate) {
        faz_ctx_set_error(ctx, "the GPU QoZ tuner uses the QoZ default tuning rates");
        return FAZ_ERR_UNSUPPORTED;
    }
    conf.profiling = 0;
    conf.var_first = 0;
    conf.testLorenzo = 0;
    conf.waveletAutoTuning = 0;
    conf.waveletTest = 0;
    conf.waveAutoFix = 0;
    conf.sperr = -1;
    conf.conditioning = 0;
    conf.pyBind = 0;
    conf.fixWave = -1;
    conf.sperrWithoutWave = false;
    faz_tune_in tin;
    memset(&tin, 0, sizeof(tin));
    tin.N = N;
    for (int i = 0; i < N; i++) tin.dims[i] = conf.dims[i];
    tin.abs_eb = conf.absErrorBound;
    tin.rel_eb = conf.relErrorBound;
    tin.rng = conf.rng;
    tin.interp_block_size = conf.interpBlockSize;
    tin.max_step = conf.maxStep;
    tin.levelwise = conf.levelwisePredictionSelection;
    tin.sample_block_size = conf.sampleBlockSize;
    tin.ab_list = conf.abList;
    std::unique_ptr<faz_tune_out> tout(new faz_tune_out);
    if ((rc = faz_gpu_tune_qoz_f32(ctx, d_data, &tin, tout.get()))) return rc;
    conf.levelwisePredictionSelection = tout->levelwise;
    conf.sampleBlockSize = tout->sample_block_size;
    conf.interpAlgo = (uint8_t) tout->interp_algo;
    conf.interpDirection = (uint8_t) tout->direction;
    conf.interpAlgo_list.assign(tout->ops, tout->ops + tout->levelwise);
    conf.interpDirection_list.assign(tout->dirs, tout->dirs + tout->levelwise);
    conf.alpha = tout->alpha;
    conf.beta = tout->beta;
    conf.wavelet_rel_coeff = 1;
    conf.wavelet = 0;
    conf.cmprAlgo = QoZ::ALGO_INTERP;
    return compress_interp_stream(ctx, conf, d_data, out);
}

int decompress_stream(faz_gpu_ctx *ctx, const uint8_t *stream, size_t n, QoZ::Config &conf, float *d_out,
                      size_t num) {
    if (n < 4) return FAZ_ERR_FORMAT;
    int confSize;
    memcpy(&confSize, stream + n - 4, 4);
    if (confSize <= 0 || (size_t) confSize + 4 > n) return FAZ_ERR_FORMAT;
    const QoZ::uchar *c = stream + (n - 4 - confSize);
    if (!conf.load(c, stream + n - 4)) {
        faz_ctx_set_error(ctx, "Config trailer truncated or corrupt");
        return FAZ_ERR_FORMAT;
    }
    if (conf.N < 1 || conf.N > 4 || conf.num != num) {
        faz_ctx_set_error(ctx, "stream shape does not match the output buffer");
        return FAZ_ERR_FORMAT;
    }
    // SZ_decompress_LorenzoReg (LR:117-135): fast frontend for 3-D without
    // regression2, else the general frontend
    const bool lorenzo = conf.cmprAlgo == QoZ::ALGO_LORENZO_REG && conf.N == 3 && !conf.regression2;
    const bool general = conf.cmprAlgo == QoZ::ALGO_LORENZO_REG && !lorenzo && !conf.regression2;
    if (!(conf.cmprAlgo == QoZ::ALGO_INTERP || lorenzo || general) || conf.wavelet != 0) {
        faz_ctx_set_error(ctx, "only ALGO_INTERP and ALGO_LORENZO_REG streams (without regression2) without "
                               "wavelets decode on the GPU path");
        return FAZ_ERR_UNSUPPORTED;
    }
    std::vector<uint8_t> raw;
    if (!zstd_unwrap(stream, n - 4 - confSize, raw)) {
        faz_ctx_set_error(ctx, "zstd frame malformed");
        return FAZ_ERR_FORMAT;
    }
    if (lorenzo) {
        size_t dims[3] = {conf.dims[0], conf.dims[1], conf.dims[2]};
        return faz_gpu_lorenzo_decompress_raw(ctx, raw.data(), raw.size(), dims, d_out);
    }
    if (general)
        return faz_general_decompress_raw_vec(ctx, raw.data(), raw.size(), conf.N, conf.lorenzo, conf.lorenzo2,
                                              conf.regression, d_out,
                                              num);
    return faz_gpu_interp_decompress_raw(ctx, raw.data(), raw.size(), conf.N, d_out, num);
}

// Process-wide context of the C++ API (one per host thread).
faz_gpu_ctx *api_ctx() {
    thread_local std::unique_ptr<faz_gpu_ctx, void (*)(faz_gpu_ctx *)> ctx(nullptr, faz_gpu_ctx_destroy);
    if (!ctx) {
        const char *dev = getenv("FAZ_DEVICE");
        faz_gpu_ctx *c = nullptr;
        int rc = faz_gpu_ctx_create(dev ? atoi(dev) : 0, &c);
        if (rc) throw faz::Error(rc, "no HIP device available for the FAZ GPU engine");
        ctx.reset(c);
    }
    return ctx.get();
}

[[noreturn]] void fail(faz_gpu_ctx *ctx, int rc, const char *what) {
    std::string m = std::string(what) + ": " + faz_gpu_last_error(ctx);
    throw faz::Error(rc, m);
}

}  // namespace

// ---------------------------------------------------------------------------
// C++ surface (include/faz/sz.hpp)
// ---------------------------------------------------------------------------
template<>
char *SZ_compress<float>(QoZ::Config &config, const float *data, size_t &outSize) {
    QoZ::Config conf(config);
    if (conf.N < 1 || conf.N > 4) throw faz::Error(FAZ_ERR_INVALID, "Data dimension higher than 4 is not supported.");
    if (conf.cmprAlgo != QoZ::ALGO_INTERP && conf.cmprAlgo != QoZ::ALGO_LORENZO_REG &&
        conf.cmprAlgo != QoZ::ALGO_INTERP_LORENZO)
        throw faz::Error(FAZ_ERR_UNSUPPORTED, std::string("GPU engine does not serve ") + QoZ::ALGO_STR[conf.cmprAlgo]);
    faz_gpu_ctx *ctx = api_ctx();
    int rc;
    const float *d = to_device(ctx, data, conf.num, false, rc);
    if (rc) fail(ctx, rc, "staging input");
    std::vector<uint8_t> out;
    if (conf.cmprAlgo == QoZ::ALGO_LORENZO_REG)
        rc = compress_lorenzo_stream(ctx, conf, d, out);
    else if (conf.cmprAlgo == QoZ::ALGO_INTERP_LORENZO)
        rc = compress_qoz_stream(ctx, conf, d, out);
    else
        rc = compress_interp_stream(ctx, conf, d, out);
    if (rc) fail(ctx, rc, "SZ_compress");
    char *buf = new char[out.size()];
    memcpy(buf, out.data(), out.size());
    outSize = out.size();
    return buf;
}

template<>
void SZ_decompress<float>(QoZ::Config &config, char *cmpData, size_t cmpSize, float *&decData) {
    QoZ::Config conf(config);
    if (cmpSize < 4) throw faz::Error(FAZ_ERR_FORMAT, "stream too short");
    int confSize;
    memcpy(&confSize, cmpData + cmpSize - 4, 4);
    if (confSize <= 0 || (size_t) confSize + 4 > cmpSize) throw faz::Error(FAZ_ERR_FORMAT, "bad Config trailer");
    const QoZ::uchar *c = (const QoZ::uchar *) cmpData + (cmpSize - 4 - confSize);
    if (!conf.load(c, (const QoZ::uchar *) cmpData + cmpSize - 4))
        throw faz::Error(FAZ_ERR_FORMAT, "Config trailer truncated or corrupt");
    faz_gpu_ctx *ctx = api_ctx();
    float *dev = faz_ctx_work_in(ctx, conf.num);
    if (!dev) fail(ctx, FAZ_ERR_HIP, "device buffer");
    QoZ::Config loaded;
    int rc = decompress_stream(ctx, (const uint8_t *) cmpData, cmpSize, loaded, dev, conf.num);
    if (rc) fail(ctx, rc, "SZ_decompress");
    if (decData == nullptr) decData = new float[conf.num];
    if (hipMemcpy(decData, dev, conf.num * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        throw faz::Error(FAZ_ERR_HIP, "copying the reconstruction to the host");
    config = loaded;
}

// ---------------------------------------------------------------------------
// INI configuration (Config.hpp:91-233)
// ---------------------------------------------------------------------------
namespace {
std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}
bool as_bool(const std::string &v, bool def) {
    std::string l;
    for (char ch: v) l += (char) tolower(ch);
    if (l == "true" || l == "yes" || l == "on" || l == "1") return true;
    if (l == "false" || l == "no" || l == "off" || l == "0") return false;
    return def;
}
}  // namespace

bool QoZ::Config::loadcfg(const std::string &cfgpath) {
    std::ifstream f(cfgpath);
    if (!f) return false;
    std::string section, line;
    std::vector<std::pair<std::string, std::string>> kv;  // "section.key" → value
    while (std::getline(f, line)) {
        std::string t = trim(line);
        if (t.empty() || t[0] == ';' || t[0] == '#') continue;
        if (t[0] == '[') {
            size_t e = t.find(']');
            if (e == std::string::npos) return false;
            section = trim(t.substr(1, e - 1));
            continue;
        }
        size_t eq = t.find_first_of("=:");
        if (eq == std::string::npos) return false;
        std::string val = trim(t.substr(eq + 1));
        size_t cmt = val.find(" ;");
        if (cmt != std::string::npos) val = trim(val.substr(0, cmt));
        std::string key = section + "." + trim(t.substr(0, eq));
        for (auto &ch: key) ch = (char) tolower(ch);
        kv.push_back({key, val});
    }
    auto get = [&](const char *sec, const char *key, std::string &out) {
        std::string k = std::string(sec) + "." + key;
        for (auto &ch: k) ch = (char) tolower(ch);
        bool found = false;
        for (auto &p: kv)
            if (p.first == k) {
                out = p.second;
                found = true;
            }
        return found;
    };
    std::string v;
    if (get("GlobalSettings", "CmprAlgo", v))
        for (int i = 0; i < 4; i++)
            if (v == ALGO_STR[i]) cmprAlgo = (uint8_t) i;
    if (get("GlobalSettings", "ErrorBoundMode", v))
        for (int i = 0; i < 6; i++)
            if (v == EB_STR[i]) errorBoundMode = (uint8_t) i;
    if (get("GlobalSettings", "tuningTarget", v))
        for (int i = 0; i < 4; i++)
            if (v == TUNING_TARGET_STR[i]) tuningTarget = (uint8_t) i;
    auto real = [&](const char *sec, const char *key, double &dst) {
        if (get(sec, key, v)) dst = strtod(v.c_str(), nullptr);
    };
    auto integer = [&](const char *sec, const char *key, auto &dst) {
        if (get(sec, key, v)) dst = (std::remove_reference_t<decltype(dst)>) strtol(v.c_str(), nullptr, 0);
    };
    auto boolean = [&](const char *sec, const char *key, bool &dst) {
        if (get(sec, key, v)) dst = as_bool(v, dst);
    };
    real("GlobalSettings", "AbsErrorBound", absErrorBound);
    real("GlobalSettings", "RelErrorBound", relErrorBound);
    real("GlobalSettings", "PSNRErrorBound", psnrErrorBound);
    real("GlobalSettings", "L2NormErrorBound", l2normErrorBound);
    real("AlgoSettings", "alpha", alpha);
    real("AlgoSettings", "beta", beta);
    real("AlgoSettings", "autoTuningRate", autoTuningRate);
    real("AlgoSettings", "predictorTuningRate", predictorTuningRate);
    real("AlgoSettings", "wavelet_rel_coeff", wavelet_rel_coeff);
    real("AlgoSettings", "preTrim", preTrim);
    real("AlgoSettings", "waveletTuningRate", waveletTuningRate);
    real("AlgoSettings", "waveletBrFix", waveletBrFix);
    real("AlgoSettings", "waveletBrFix2", waveletBrFix2);
    real("AlgoSettings", "waveletMseFix", waveletMseFix);
    real("AlgoSettings", "waveletMseFix2", waveletMseFix2);
    real("AlgoSettings", "lorenzoBrFix", lorenzoBrFix);
    boolean("GlobalSettings", "OpenMP", openmp);
    boolean("AlgoSettings", "Lorenzo", lorenzo);
    boolean("AlgoSettings", "Lorenzo2ndOrder", lorenzo2);
    boolean("AlgoSettings", "Regression", regression);
    boolean("AlgoSettings", "Regression2ndOrder", regression2);
    if (get("AlgoSettings", "InterpolationAlgo", v)) {
        if (v == INTERP_ALGO_STR[0]) interpAlgo = INTERP_ALGO_LINEAR;
        if (v == INTERP_ALGO_STR[1]) interpAlgo = INTERP_ALGO_CUBIC;
    }
    integer("AlgoSettings", "InterpDirection", interpDirection);
    integer("AlgoSettings", "InterpBlockSize", interpBlockSize);
    integer("AlgoSettings", "BlockSize", blockSize);
    integer("AlgoSettings", "QuantizationBinTotal", quantbinCnt);
    integer("AlgoSettings", "maxStep", maxStep);
    integer("AlgoSettings", "sampleBlockSize", sampleBlockSize);
    integer("AlgoSettings", "levelwisePredictionSelection", levelwisePredictionSelection);
    integer("AlgoSettings", "testLorenzo", testLorenzo);
    integer("AlgoSettings", "profiling", profiling);
    integer("AlgoSettings", "SSIMBlockSize", SSIMBlockSize);
    integer("AlgoSettings", "fixBlockSize", fixBlockSize);
    boolean("AlgoSettings", "verbose", verbose);
    boolean("AlgoSettings", "QoZ", QoZ);
    boolean("AlgoSettings", "FZ", FZ);
    boolean("AlgoSettings", "sperrWithoutWave", sperrWithoutWave);
    boolean("AlgoSettings", "pyBind", pyBind);
    integer("AlgoSettings", "blockwiseSampleBlockSize", blockwiseSampleBlockSize);
    integer("AlgoSettings", "abList", abList);
    integer("AlgoSettings", "crossBlock", crossBlock);
    integer("AlgoSettings", "sampleBlockSampleBlockSize", sampleBlockSampleBlockSize);
    boolean("AlgoSettings", "peTracking", peTracking);
    integer("AlgoSettings", "wavelet", wavelet);
    integer("AlgoSettings", "offsetPredictor", offsetPredictor);
    integer("AlgoSettings", "trimToZero", trimToZero);
    integer("AlgoSettings", "pid", pid);
    integer("AlgoSettings", "blockOrder", blockOrder);
    integer("AlgoSettings", "coeffTracking", coeffTracking);
    integer("AlgoSettings", "waveletTest", waveletTest);
    integer("AlgoSettings", "waveletAutoTuning", waveletAutoTuning);
    integer("AlgoSettings", "var_first", var_first);
    integer("AlgoSettings", "profStride", profStride);
    integer("AlgoSettings", "sperr", sperr);
    integer("AlgoSettings", "waveAutoFix", waveAutoFix);
    integer("AlgoSettings", "conditioning", conditioning);
    integer("AlgoSettings", "fixWave", fixWave);
    return true;
}

// ---------------------------------------------------------------------------
// C-ABI forms (faz_gpu.h)
// ---------------------------------------------------------------------------
extern "C" {

faz_status faz_sz_compress_f32(faz_gpu_ctx *ctx, const faz_interp_params *p, int rel_mode, double eb,
                               const float *data, int data_on_device, uint8_t **out, size_t *out_size) {
    if (!ctx || !p || !data || !out || !out_size || p->N < 1 || p->N > 4) return FAZ_ERR_INVALID;
    size_t num = 1;
    for (int i = 0; i < p->N; i++) num *= p->dims[i];
    QoZ::Config conf = interp_config(p, 0.0, -1.0);
    if (rel_mode) {
        conf.errorBoundMode = QoZ::EB_REL;
        conf.relErrorBound = eb;
    } else {
        conf.errorBoundMode = QoZ::EB_ABS;
        conf.absErrorBound = eb;
    }
    int rc;
    const float *d = to_device(ctx, data, num, data_on_device != 0, rc);
    if (rc) return rc;
    std::vector<uint8_t> buf;
    rc = compress_interp_stream(ctx, conf, d, buf);
    if (rc) return rc;
    uint8_t *m = (uint8_t *) malloc(buf.size());
    if (!m) return FAZ_ERR_INVALID;
    memcpy(m, buf.data(), buf.size());
    *out = m;
    *out_size = buf.size();
    return FAZ_OK;
}

faz_status faz_sz_decompress_f32(faz_gpu_ctx *ctx, const uint8_t *stream, size_t n, float *out, size_t num,
                                 int out_on_device) {
    if (!ctx || !stream || !out) return FAZ_ERR_INVALID;
    float *d = out;
    if (!out_on_device) {
        d = faz_ctx_work_in(ctx, num);
        if (!d) return FAZ_ERR_HIP;
    }
    QoZ::Config conf;
    int rc;
    try {  // no C++ exception crosses the C boundary
        rc = decompress_stream(ctx, stream, n, conf, d, num);
    } catch (const std::exception &e) {
        faz_ctx_set_error(ctx, e.what());
        return FAZ_ERR_FORMAT;
    }
    if (rc) return rc;
    if (!out_on_device && hipMemcpy(out, d, num * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return FAZ_ERR_HIP;
    return FAZ_OK;
}

// The Config of a stream's trailer without decoding it (host only): N, the
// dimensions and cmprAlgo, for callers that size their output first.
faz_status faz_stream_info(const uint8_t *stream, size_t n, int *N, size_t *dims, int *cmpr_algo) {
    if (!stream || !N || !dims || !cmpr_algo) return FAZ_ERR_INVALID;
    if (n < 4) return FAZ_ERR_FORMAT;
    int confSize;
    memcpy(&confSize, stream + n - 4, 4);
    if (confSize <= 0 || (size_t) confSize + 4 > n) return FAZ_ERR_FORMAT;
    const QoZ::uchar *c = stream + (n - 4 - confSize);
    QoZ::Config conf;
    try {
        if (!conf.load(c, stream + n - 4)) return FAZ_ERR_FORMAT;
    } catch (const std::exception &) {
        return FAZ_ERR_FORMAT;
    }
    *N = conf.N;
    for (int i = 0; i < conf.N; i++) dims[i] = conf.dims[i];
    *cmpr_algo = conf.cmprAlgo;
    return FAZ_OK;
}

faz_status faz_sz_compress_lorenzo_f32(faz_gpu_ctx *ctx, const faz_lorenzo_params *p, int rel_mode, double eb,
                                       const float *data, int data_on_device, uint8_t **out, size_t *out_size) {
    if (!ctx || !p || !data || !out || !out_size) return FAZ_ERR_INVALID;
    QoZ::Config conf(p->dims[0], p->dims[1], p->dims[2]);
    conf.cmprAlgo = QoZ::ALGO_LORENZO_REG;
    conf.blockSize = p->block_size;
    conf.pred_dim = p->pred_dim;
    conf.lorenzo = p->lorenzo != 0;
    conf.lorenzo2 = p->lorenzo2 != 0;
    conf.regression = p->regression != 0;
    conf.quantbinCnt = p->radius * 2;
    if (rel_mode) {
        conf.errorBoundMode = QoZ::EB_REL;
        conf.relErrorBound = eb;
    } else {
        conf.errorBoundMode = QoZ::EB_ABS;
        conf.absErrorBound = eb;
    }
    int rc;
    const float *d = to_device(ctx, data, conf.num, data_on_device != 0, rc);
    if (rc) return rc;
    std::vector<uint8_t> buf;
    rc = compress_lorenzo_stream(ctx, conf, d, buf);
    if (rc) return rc;
    uint8_t *m = (uint8_t *) malloc(buf.size());
    if (!m) return FAZ_ERR_INVALID;
    memcpy(m, buf.data(), buf.size());
    *out = m;
    *out_size = buf.size();
    return FAZ_OK;
}

faz_status faz_sz_compress_batch_f32(faz_batch *b, const int *rel_mode, const double *eb, uint8_t **out,
                                     size_t *out_size) {
    if (!b || !rel_mode || !eb || !out || !out_size) return FAZ_ERR_INVALID;
    faz_batch_stats s;
    faz_gpu_batch_stats(b, &s, nullptr);
    int n = 0;
    for (faz_interp_params p; faz_gpu_batch_field(b, n, &p) == FAZ_OK;) n++;
    std::vector<size_t> sizes(n);
    std::vector<double> aeb(n), rng(n);
    int rc = faz_gpu_batch_compress(b, rel_mode, eb, sizes.data(), aeb.data(), rng.data());
    if (rc) return rc;
    std::vector<std::vector<uint8_t>> raw(n), streams(n);
    for (int i = 0; i < n; i++) {
        raw[i].resize(sizes[i]);
        if ((rc = faz_gpu_batch_fetch_raw(b, i, raw[i].data(), sizes[i]))) return rc;
    }
    // zstd + Config trailer per field, as compress_interp_stream writes them
    std::vector<int> ok(n, 0);
    auto work = [&](int t, int nt) {
        for (int i = t; i < n; i += nt) {
            faz_interp_params p;
            faz_gpu_batch_field(b, i, &p);
            QoZ::Config conf = interp_config(&p, 0.0, -1.0);
            if (rel_mode[i]) {
                conf.errorBoundMode = QoZ::EB_ABS;  // cal_abs_error_bound after EB_REL
                conf.relErrorBound = eb[i];
                conf.rng = rng[i];
                conf.absErrorBound = aeb[i];
            } else {
                conf.errorBoundMode = QoZ::EB_ABS;
                conf.absErrorBound = eb[i];
            }
            if (!zstd_wrap(raw[i], streams[i], QoZ::Config::size_est())) continue;
            append_trailer(conf, streams[i]);
            ok[i] = 1;
        }
    };
    const int nt = std::max(1, std::min(n, 16));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++) pool.emplace_back(work, t, nt);
    for (auto &th: pool) th.join();
    for (int i = 0; i < n; i++)
        if (!ok[i]) return FAZ_ERR_INVALID;
    for (int i = 0; i < n; i++) {
        uint8_t *m = (uint8_t *) malloc(streams[i].size());
        if (!m) {
            for (int j = 0; j < i; j++) free(out[j]);
            return FAZ_ERR_INVALID;
        }
        memcpy(m, streams[i].data(), streams[i].size());
        out[i] = m;
        out_size[i] = streams[i].size();
    }
    return FAZ_OK;
}

faz_status faz_sz_compress_general_f32(faz_gpu_ctx *ctx, const faz_general_params *p, int rel_mode, double eb,
                                       const float *data, int data_on_device, uint8_t **out, size_t *out_size) {
    if (!ctx || !p || !data || !out || !out_size || p->N < 1 || p->N > 4) return FAZ_ERR_INVALID;
    if (p->N == 3) {
        faz_ctx_set_error(ctx, "SZ_compress takes the fast frontend for 3-D Lorenzo (faz_sz_compress_lorenzo_f32)");
        return FAZ_ERR_INVALID;
    }
    QoZ::Config conf;
    conf.setDims(p->dims, p->dims + p->N);
    conf.cmprAlgo = QoZ::ALGO_LORENZO_REG;
    conf.blockSize = p->block_size;
    conf.lorenzo = p->lorenzo != 0;
    conf.lorenzo2 = p->lorenzo2 != 0;
    conf.regression = p->regression != 0;
    conf.regression2 = false;
    conf.quantbinCnt = p->radius * 2;
    if (rel_mode) {
        conf.errorBoundMode = QoZ::EB_REL;
        conf.relErrorBound = eb;
    } else {
        conf.errorBoundMode = QoZ::EB_ABS;
        conf.absErrorBound = eb;
    }
    int rc;
    const float *d = to_device(ctx, data, conf.num, data_on_device != 0, rc);
    if (rc) return rc;
    std::vector<uint8_t> buf;
    rc = compress_lorenzo_stream(ctx, conf, d, buf);
    if (rc) return rc;
    uint8_t *m = (uint8_t *) malloc(buf.size());
    if (!m) return FAZ_ERR_INVALID;
    memcpy(m, buf.data(), buf.size());
    *out = m;
    *out_size = buf.size();
    return FAZ_OK;
}

faz_status faz_sz_compress_sz3_f32(faz_gpu_ctx *ctx, int N, const size_t *dims, int rel_mode, double eb,
                                   const float *data, int data_on_device, uint8_t **out, size_t *out_size) {
    if (!ctx || !dims || !data || !out || !out_size || N < 1 || N > 4) return FAZ_ERR_INVALID;
    QoZ::Config conf;
    conf.setDims(dims, dims + N);
    conf.cmprAlgo = QoZ::ALGO_INTERP_LORENZO;
    if (rel_mode) {
        conf.errorBoundMode = QoZ::EB_REL;
        conf.relErrorBound = eb;
    } else {
        conf.errorBoundMode = QoZ::EB_ABS;
        conf.absErrorBound = eb;
    }
    int rc;
    const float *d = to_device(ctx, data, conf.num, data_on_device != 0, rc);
    if (rc) return rc;
    std::vector<uint8_t> buf;
    rc = compress_sz3_stream(ctx, conf, d, buf);
    if (rc) return rc;
    uint8_t *m = (uint8_t *) malloc(buf.size());
    if (!m) return FAZ_ERR_INVALID;
    memcpy(m, buf.data(), buf.size());
    *out = m;
    *out_size = buf.size();
    return FAZ_OK;
}

faz_status faz_sz_compress_qoz_f32(faz_gpu_ctx *ctx, int N, const size_t *dims, int rel_mode, double eb,
                                   const float *data, int data_on_device, uint8_t **out, size_t *out_size) {
    if (!ctx || !dims || !data || !out || !out_size || N < 1 || N > 4) return FAZ_ERR_INVALID;
    QoZ::Config conf;
    conf.setDims(dims, dims + N);
    conf.cmprAlgo = QoZ::ALGO_INTERP_LORENZO;
    conf.QoZ = 1;
    if (rel_mode) {
        conf.errorBoundMode = QoZ::EB_REL;
        conf.relErrorBound = eb;
    } else {
        conf.errorBoundMode = QoZ::EB_ABS;
        conf.absErrorBound = eb;
    }
    int rc;
    const float *d = to_device(ctx, data, conf.num, data_on_device != 0, rc);
    if (rc) return rc;
    std::vector<uint8_t> buf;
    rc = compress_qoz_stream(ctx, conf, d, buf);
    if (rc) return rc;
    uint8_t *m = (uint8_t *) malloc(buf.size());
    if (!m) return FAZ_ERR_INVALID;
    memcpy(m, buf.data(), buf.size());
    *out = m;
    *out_size = buf.size();
    return FAZ_OK;
}

void faz_free(void *p) { free(p); }

}  // extern "C"
