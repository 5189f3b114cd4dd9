// Context of the C-ABI (faz_gpu.h): device scratch, stream, events and the
// engine buffer waiting to be fetched.  Shared by the host orchestration of
// the interpolation engine (faz_gpu.hip) and of the Lorenzo/regression engine
// (faz_lorenzo.hip).  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <cstdio>
#include <string>
#include <vector>

#include "faz_gpu.h"
#include "../host/engine_format.h"
#include "faz_internal.h"
#include "../host/huffman_host.h"

namespace faz {

// Page-locked host array (transfers without a staging copy).
template<class T>
struct HostPinned {
    T *p = nullptr;
    bool pinned = false;
    explicit HostPinned(size_t n, unsigned flags = 0) {
        if (hipHostMalloc((void **) &p, n * sizeof(T), flags) == hipSuccess) {
            pinned = true;
        } else {
            (void) hipGetLastError();
            p = new T[n];
        }
    }
    ~HostPinned() {
        if (pinned) (void) hipHostFree(p);
        else delete[] p;
    }
    HostPinned(const HostPinned &) = delete;
    HostPinned &operator=(const HostPinned &) = delete;
    T *data() { return p; }
};

// Page-locked host array that grows on demand (contents not kept across
// growth); a pageable source would make hipMemcpyAsync wait for its copy.
template<class T>
struct PinnedVec {
    T *p = nullptr;
    size_t cap = 0;
    bool pinned = false;
    T *ensure(size_t n) {
        if (n <= cap && p) return p;
        release();
        const size_t want = std::max<size_t>(n, 1);
        if (hipHostMalloc((void **) &p, want * sizeof(T), 0) == hipSuccess) {
            pinned = true;
        } else {
            (void) hipGetLastError();
            p = new T[want];
            pinned = false;
        }
        cap = want;
        return p;
    }
    void release() {
        if (p) {
            if (pinned) (void) hipHostFree(p);
            else delete[] p;
        }
        p = nullptr;
        cap = 0;
    }
    ~PinnedVec() { release(); }
    T &operator[](size_t i) { return p[i]; }
    T *data() { return p; }
};

template<class T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n) {
        if (n <= cap && p) return hipSuccess;
        if (p) (void) hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(n, 1);
        hipError_t e = hipMalloc((void **) &p, want * sizeof(T));
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void) hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Engine buffer waiting in a context: everything but the device arrays.
struct PendingRaw {
    faz::EngineHeader h;
    double eb = 0;
    int radius = 0;
    uint64_t n_unpred = 0;
    faz::HuffTable tab;
    size_t nbytes = 0;
    bool valid = false;
};

// Waits for a stream by polling it: the waits of the compression path are
// short, and polling avoids the sleep and wake-up of a blocking wait.  After
// about a second of polling it falls back to the blocking wait.
inline hipError_t spin_wait(hipStream_t st) {
    for (int i = 0; i < 2000000; i++) {
        const hipError_t e = hipStreamQuery(st);
        if (e != hipErrorNotReady) return e;
    }
    return hipStreamSynchronize(st);
}

// The same polling wait on one event (work enqueued after it keeps running).
inline hipError_t spin_wait_event(hipEvent_t ev) {
    for (int i = 0; i < 2000000; i++) {
        const hipError_t e = hipEventQuery(ev);
        if (e != hipErrorNotReady) return e;
    }
    return hipEventSynchronize(ev);
}

}  // namespace faz

struct faz_gpu_ctx {
    int device = -1;
    hipStream_t stream = nullptr;
    std::string err;
    faz::DevBuf<uint16_t> codes;
    faz::DevBuf<float> outval, unpred, work_in, work_out;
    // compact lattices handed from one whole-tile level to the next
    faz::DevBuf<float> lat[2];
    faz::DevBuf<uint32_t> freq, first, chunk_zeros, err_flag, minmax, words;
    faz::DevBuf<uint32_t> hist_pack;  // one readback block after the histogram
    faz::DevBuf<uint32_t> enc_head;   // one-pass encode: bits of each chunk's first, shared word
    // FAZ_ENCODE=onepass selects the decoupled look-back encoder over the
    // sizes + scan + pack one (slower at 512^3: the 512^3 step takes about
    // 40 us longer, 43 rounds of look-back latency against one more read of
    // the codes)
    bool enc_onepass = false;
    bool dec_run = true;  // the decoder skips the search inside the anchors' run of zero codes
    bool dev_bounds = true;  // REL compressions compute the level bounds on the device
    faz::DevBuf<faz::LevelEB> lvl;  // level bounds of the last REL compression (entry 0: the base bound)
    faz::DevBuf<uint64_t> chunk_base, scalars, prefix, chunk_bits, chunk_off, enc_code, tiles;
    uint64_t prefix_key = 0;  // identity of the block prefix table in `prefix`
    const uint64_t *prefix_at = nullptr;  // the allocation that table was uploaded to
    // engine plan of the last parameter set (see cached_plan in faz_gpu.hip)
    std::shared_ptr<void> plan_cache;
    std::vector<uint8_t> plan_key;
    // decoder of a device-resident buffer: its head and its Huffman tree on the host
    std::vector<uint8_t> h_stage, h_stage2;
    hipEvent_t ev[8] = {};
    double timings[10] = {};
    // histogram of the codes last produced by interp_compress
    const uint16_t *hist_codes = nullptr;
    size_t hist_n = 0;
    std::vector<uint32_t> h_freq = std::vector<uint32_t>(65536), h_first = std::vector<uint32_t>(65536);
    // pinned: histogram and first positions, then 8 words of scalars
    faz::HostPinned<uint32_t> hp_hist{2 * 65536 + 8};
    // coherent pinned words that kernels write directly: the histogram block
    // and base bound, then (at kMailReport) the encode's bit total and error
    // flag, then (at kMailDecode) the decompression's checks
    static constexpr size_t kMailReport = faz::kHistPackWords + 4;
    static constexpr size_t kMailDecode = kMailReport + 4;
    faz::HostPinned<uint32_t> mailbox{kMailDecode + 8, hipHostMallocCoherent};
    faz::PinnedVec<uint64_t> h_tab;  // packed code table of the last encode (pinned: asynchronous upload)
    faz::PinnedVec<uint32_t> h_dec;  // the decoder's lookup tables and tree, staged for upload
    faz::PendingRaw pending;
    // timing of the interpolation kernels (whole-tile launches timed as one
    // span per run of consecutive launches); phase_events: the compression's
    // phase timings (each event record costs the stream a few microseconds)
    bool profile = false;
    bool phase_events = false;
    struct ProfBucket {
        double ms = 0, bytes = 0;  // bytes: algorithmic HBM bytes of the launches
        uint64_t launches = 0, points = 0;
    } prof[3];
    std::vector<hipEvent_t> prof_ev;
    // launches of the last compression whose event times are still to be read:
    // bucket, event pair index, points, algorithmic bytes
    struct ProfRec {
        int bucket;
        uint32_t ev;
        uint64_t points;
        double bytes;
        uint32_t launches;
    };
    std::vector<ProfRec> prof_pending;
    // 3-D fields: one fused launch per level (0 = per-pass launches)
    bool fused3d = true;
    // levels that qualify run on the whole-tile kernel (persistent, tile_grid
    // workgroups: one per compute unit, a multiple of 8 XCDs)
    bool tile3d = true;
    uint32_t tile_grid = 256;
    // levels with fewer blocks than this run as per-pass launches
    uint64_t fused_min_blocks = 64;
    // Lorenzo/regression engine: per-block selector, coefficients, regression
    // ranks, the in-block hyperplane order table and its offsets
    faz::DevBuf<uint8_t> lr_ind;
    faz::DevBuf<float4> lr_coef, lr_regc;
    faz::DevBuf<uint32_t> lr_flag, lr_diag;
    faz::DevBuf<uint64_t> lr_rank;
    faz::DevBuf<uint16_t> lr_tab;
    uint32_t lr_tab_bs = 0;  // block size the table was built for
    // general-frontend Lorenzo engine: 32-bit codes, per-block selection,
    // gathered sample blocks, the in-block hyperplane table for (N, bs)
    faz::DevBuf<uint32_t> gf_codes, gf_tab, gf_diag;
    faz::DevBuf<uint8_t> gf_sel;
    faz::DevBuf<float> gf_fields, gf_recon;
    faz::DevBuf<float> gf_reg_raw, gf_reg_deq;  // RegressionPredictor planes, 5 floats per block
    faz::DevBuf<uint32_t> gf_reg_code;          // 1-D: coefficient codes per block
    // the tuner's device buffers (owned by the tuner's source file)
    void *tune_cache = nullptr;
    void (*tune_cache_free)(void *) = nullptr;
    uint32_t gf_tab_key = 0;  // N << 16 | bs of the table
    // Huffman decoder: the bitstream as words, the lookup table and tree,
    // per-chunk start and symbol offset, symbol counts + error flag, the
    // per-chunk entry maps (two buffers) and counts per entry
    faz::DevBuf<uint32_t> dec_words, dec_tab, dec_cnt;
    faz::DevBuf<uint8_t> dec_leaf, dec_x, dec_g;
    faz::DevBuf<uint16_t> dec_n;
    // per chunk and entry offset: first codeword start past each inner quarter boundary, codewords before it
    faz::DevBuf<uint32_t> dec_qx;
    faz::DevBuf<uint64_t> dec_qn;
    faz::DevBuf<uint64_t> dec_wtab;  // multi-codeword output table
    bool hist_cleared = false;       // freq/first already cleared on the stream for this compression
    bool err_cleared = false;        // err_flag already cleared on the stream for this compression
    faz::DevBuf<uint64_t> dec_pos;
    uint32_t dec_rounds = 0;  // chunk groups of the last decode
    hipStream_t ustream = nullptr;  // uploads of the decoder (default priority)
    hipEvent_t dec_ev[4] = {};
};

#define CTX_FAIL(code, ...)                                          \
    do {                                                             \
        char _m[512];                                                \
        snprintf(_m, sizeof(_m), __VA_ARGS__);                       \
        ctx->err = _m;                                               \
        return code;                                                 \
    } while (0)

#define HIPCHECK(x)                                                                               \
    do {                                                                                          \
        hipError_t _e = (x);                                                                      \
        if (_e != hipSuccess) CTX_FAIL(FAZ_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(_e)); \
    } while (0)

