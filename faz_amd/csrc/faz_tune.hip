// Exhaustive QoZ sampling tuner on gfx950 (`faz -q`: Tuning() with conf.QoZ,
// include/QoZ/api/impl/SZInterp.hpp:1589-2557).
//
// The tuner runs the interpolation engine many times on small sample blocks
// ((sbs+1)^N values, 33^3 for 3-D fields) in the reference's tuning modes:
// mode 2 (sum of |prediction error| over a level range, SIC:955-965) to pick
// the per-level interpolator and direction, mode 1 (codes + squared error,
// SIC:944-950) inside CompressTest to rank alpha/beta pairs.  A sample block
// fits in LDS whole, so one workgroup runs a whole engine call (every level,
// block, pass and line) on its own, and one launch runs every (candidate,
// sample block) pair at once.  Within a pass every line is independent; a
// thread walks one line in the reference's emission order, so its error sum
// and code ranks follow the reference; the per-line sums are then added in
// line order by one thread, which reproduces the reference's FP64 sums bit for
// bit (line -> block accumulator -> running total, SIC:438-457, 1306-1417).
#include "faz_gpu.h"
#include "faz_ctx.h"
#include "faz_device.h"
#include "faz_internal.h"
#include "../host/raw_api.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#pragma clang fp contract(off)

extern "C" {
// libzstd (system shared library); prototypes as published in zstd.h
size_t ZSTD_compress(void *dst, size_t dstCapacity, const void *src, size_t srcSize, int compressionLevel);
size_t ZSTD_compressBound(size_t srcSize);
unsigned ZSTD_isError(size_t code);
}

namespace faz {
namespace {

constexpr int kTuneMaxLevels = 16;
constexpr uint32_t kTuneThreads = 512;
constexpr uint32_t kTuneMaxPoints = 33 * 33 * 33;  // largest sample block staged in LDS
constexpr uint32_t kTuneMaxLines = 33 * 33;         // most lines of one pass

// Shape of the sample blocks and where they lie in the field.
struct TuneGeom {
    int N;
    uint32_t bd[3];      // block extent per axis ((sbs+1) each)
    uint32_t bstr[3];    // row-major strides inside the block
    uint64_t fstr[3];    // strides of the field
    int levels;          // interpolation_level of a block (SIC:820-852)
    int anchor;          // maxStep > 0
    uint32_t max_step;
    uint32_t ibs;        // interpBlockSize
    int radius;
    uint32_t ncodes;     // codes of one mode-1 run
    uint32_t max_lines;  // most lines of one pass of one interpolation block
    int perm[6][3];      // dimension_sequences
};

// One engine call: tuning mode, level range and per-level quantizer and
// interpolator (SIC:384-422).
struct TuneCand {
    int mode;            // 1 or 2
    int start, end;      // levels start..end+1 (already clamped as SIC:360-366)
    int top;             // start == levels: anchors / top-level handling
    double eb0, recip0;  // base bound (first point when there are no anchors)
    float ebf0;
    double eb[kTuneMaxLevels], recip[kTuneMaxLevels];
    float ebf[kTuneMaxLevels];
    int8_t linear[kTuneMaxLevels], dir[kTuneMaxLevels];
};

struct TuneItem {
    uint32_t cand;
    uint32_t block;      // index into the block origins
};

__device__ __forceinline__ float line_pred(const float *v, int f, int S, uint32_t i, uint32_t n, int linear) {
    return stencil_pred(i, n, linear, [&](int k) { return v[f + k * S]; });
}

// Emission order of the odd positions of a line of n points (SIC:1316-1414).
// Returns the k-th position.
__device__ __forceinline__ uint32_t emit_pos(uint32_t k, uint32_t n, int linear) {
    if (linear || n < 5) return 2 * k + 1;
    const uint32_t m = (n - 3) / 2 - 1;  // interior cubic points 3, 5, ...
    if (k < m) return 3 + 2 * k;
    if (k == m) return 1;
    if (k == m + 1) return 3 + 2 * m;
    return n - 1;
}

// One workgroup per item: the whole engine call on one sample block in LDS.
// codes/outval: per item, ncodes slots in the reference's push order (mode 1).
__global__ void __launch_bounds__(kTuneThreads) tune_kernel(TuneGeom G, const TuneCand *__restrict__ cands,
                                                            const TuneItem *__restrict__ items,
                                                            const uint64_t *__restrict__ origins,
                                                            const float *__restrict__ field,
                                                            uint16_t *__restrict__ codes, float *__restrict__ outval,
                                                            double *__restrict__ loss) {
    __shared__ double s_line[kTuneMaxLines];  // per-line sums of one pass
    __shared__ float s_v[kTuneMaxPoints];      // the block
    const TuneItem it = items[blockIdx.x];
    const TuneCand &C = cands[it.cand];
    const uint32_t npts = G.bd[0] * G.bd[1] * (G.N == 3 ? G.bd[2] : 1);
    __shared__ double s_total;
    __shared__ uint32_t s_seq;
    const uint32_t tid = threadIdx.x;
    const uint64_t org = origins[it.block];
    // stage the block
    for (uint32_t q = tid; q < npts; q += blockDim.x) {
        uint32_t r = q;
        uint64_t f = org;
#pragma unroll
        for (int a = 2; a >= 0; a--) {
            if (a >= G.N) continue;
            f += (uint64_t) (r % G.bd[a]) * G.fstr[a];
            r /= G.bd[a];
        }
        s_v[q] = field[f];
    }
    __syncthreads();
    const uint64_t cbase = (uint64_t) blockIdx.x * G.ncodes;
    if (tid == 0) {
        s_total = 0.0;
        uint32_t seq = 0;
        if (!G.anchor) {
            // SIC:372-374: the first value, predicted by 0 with the base bound
            float rec;
            const int code = quantize_point(s_v[0], 0.0f, C.eb0, C.recip0, G.radius, &rec);
            if (C.mode == 1) {
                codes[cbase] = (uint16_t) code;
                if (code == 0) outval[cbase] = s_v[0];
            }
            s_v[0] = rec;
            seq = 1;
        } else if (C.top && C.mode == 1) {
            // build_grid in tuning mode 1 (SIC:854-900): anchors from maxStep on
            const uint32_t m = G.max_step;
            if (G.N == 2) {
                for (uint32_t a = m; a < G.bd[0]; a += m)
                    for (uint32_t b = m; b < G.bd[1]; b += m) {
                        codes[cbase + seq] = 0;
                        outval[cbase + seq] = s_v[a * G.bstr[0] + b];
                        seq++;
                    }
            } else {
                for (uint32_t a = m; a < G.bd[0]; a += m)
                    for (uint32_t b = m; b < G.bd[1]; b += m)
                        for (uint32_t c = m; c < G.bd[2]; c += m) {
                            codes[cbase + seq] = 0;
                            outval[cbase + seq] = s_v[a * G.bstr[0] + b * G.bstr[1] + c];
                            seq++;
                        }
            }
        }
        s_seq = seq;
    }
    __syncthreads();
    uint32_t seq = s_seq;
    for (int level = C.start; level > C.end; level--) {
        const int li = level - 1;
        const uint32_t s = 1u << li;
        const uint32_t B = G.ibs * s;
        const int linear = C.linear[li];
        const int *perm = G.perm[C.dir[li]];
        const double eb = C.eb[li], recip = C.recip[li];
        const float ebf = C.ebf[li];
        // per-axis arrays are indexed by unrolled natural axes only (pick3
        // selects by a run-time axis), so they stay in registers
        auto pick3 = [](const uint32_t *arr, int ax) { return ax == 0 ? arr[0] : (ax == 1 ? arr[1] : arr[2]); };
        uint32_t nb[3] = {1, 1, 1};
#pragma unroll
        for (int a = 0; a < 3; a++)
            if (a < G.N) nb[a] = (G.bd[a] - 1) / B + 1;
        const uint32_t nblk = nb[0] * nb[1] * nb[2];
        for (uint32_t blk = 0; blk < nblk; blk++) {
            uint32_t kb[3] = {0, 0, 0}, r = blk;
#pragma unroll
            for (int a = 2; a >= 0; a--) {
                if (a >= G.N) continue;
                kb[a] = r % nb[a];
                r /= nb[a];
            }
            uint32_t b[3] = {0, 0, 0}, nn[3] = {1, 1, 1}, g1[3] = {1, 1, 1}, g2[3] = {1, 1, 1};
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if (a >= G.N) continue;
                b[a] = kb[a] * B;
                const uint32_t e = min(b[a] + B, G.bd[a] - 1);
                nn[a] = (e - b[a]) / s + 1;
                g1[a] = b[a] ? (e - b[a]) / s : e / s + 1;
                g2[a] = b[a] ? (e - b[a]) / (2 * s) : e / (2 * s) + 1;
            }
            double blk_err = 0.0;  // thread 0's block accumulator (SIC:3143-3185)
            for (int P = 0; P < G.N; P++) {
                const int ax = perm[P];
                const uint32_t n = pick3(nn, ax), per_line = n / 2;
                // outer axes in permutation order: G1 (stride s) before the
                // pass axis, G2 (stride 2s) after it (SIC:3153-3176)
                uint32_t nlines = 1;
#pragma unroll
                for (int Q = 0; Q < 3; Q++) {
                    if (Q >= G.N || Q == P) continue;
                    nlines *= Q < P ? pick3(g1, perm[Q]) : pick3(g2, perm[Q]);
                }
                for (uint32_t line = tid; line < nlines; line += blockDim.x) {
                    uint32_t rr = line;
                    int f = 0;
#pragma unroll
                    for (int Q = 2; Q >= 0; Q--) {  // earlier permutation positions outer
                        if (Q >= G.N || Q == P) continue;
                        const int oa = perm[Q];
                        const uint32_t size = Q < P ? pick3(g1, oa) : pick3(g2, oa);
                        const uint32_t step = Q < P ? s : 2 * s;
                        const uint32_t idx = rr % size;
                        rr /= size;
                        const uint32_t bo = pick3(b, oa);
                        const uint32_t xc = bo ? bo + (idx + 1) * step : idx * step;
                        f += (int) (xc * G.bstr[oa]);
                    }
                    f += (int) (pick3(b, ax) * G.bstr[ax]);
                    const int S = (int) (s * G.bstr[ax]);
                    double err = 0.0;
                    const uint64_t lbase = cbase + seq + (uint64_t) line * per_line;
                    for (uint32_t k = 0; k < per_line; k++) {
                        const uint32_t i = emit_pos(k, n, linear);
                        const int at = f + (int) i * S;
                        const float pred = line_pred(s_v, at, S, i, n, linear);
                        const float orig = s_v[at];
                        float rec;
                        const int code = quantize_select(orig, pred, eb, recip, ebf, G.radius, &rec);
                        if (C.mode == 1) {
                            const float d = rec - orig;
                            err += (double) (d * d);
                            if (lbase + k < cbase + G.ncodes) {
                                codes[lbase + k] = (uint16_t) code;
                                if (code == 0) outval[lbase + k] = orig;
                            }
                        } else {
                            err += (double) fabsf(orig - pred);
                        }
                        s_v[at] = rec;
                    }
                    s_line[line] = err;
                }
                __syncthreads();
                if (tid == 0)
                    for (uint32_t l = 0; l < nlines; l++) blk_err += s_line[l];
                seq += nlines * per_line;
                __syncthreads();
            }
            if (tid == 0) s_total += blk_err;
        }
    }
    if (tid == 0) loss[blockIdx.x] = s_total;
}

}  // namespace
}  // namespace faz

namespace faz {
namespace {

// Reorders one candidate's block-major codes into CompressTest's order
// (SZI:1346-1356): level segments from the top level down, within a segment
// every block's codes of that level in block order.
struct QSeg {
    uint32_t qstart;     // first position in the interleaved order
    uint32_t m0, len;    // the segment inside one block's sequence
};

__global__ void interleave_kernel(const uint16_t *__restrict__ src, uint16_t *__restrict__ dst, uint32_t nblocks,
                                  uint32_t ncodes, const QSeg *__restrict__ segs, int nsegs) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nblocks * ncodes) return;
    int k = 0;
    while (k + 1 < nsegs && q >= segs[k + 1].qstart) k++;
    const uint32_t r = q - segs[k].qstart;
    const uint32_t l = r / segs[k].len, m = segs[k].m0 + r % segs[k].len;
    dst[q] = src[(uint64_t) l * ncodes + m];
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

// init (SIC:820-852) for a sample block, and the codes each level pushes.
struct TunePlan {
    TuneGeom G;
    std::vector<uint32_t> level_codes;  // [level-1]
    uint32_t head = 0;                  // codes before the first level (anchors / first point)
};

void make_tune_plan(int N, uint32_t edge, const uint64_t *fstr, size_t max_step, int ibs, TunePlan &pl) {
    TuneGeom &G = pl.G;
    memset(&G, 0, sizeof(G));
    G.N = N;
    for (int a = 0; a < 3; a++) G.bd[a] = a < N ? edge : 1;
    G.bstr[N - 1] = 1;
    for (int a = N - 2; a >= 0; a--) G.bstr[a] = G.bstr[a + 1] * G.bd[a + 1];
    for (int a = 0; a < N; a++) G.fstr[a] = fstr[a];
    int levels = -1;
    for (int i = 0; i < N; i++)
        if (levels < std::ceil(std::log2((double) edge))) levels = (unsigned) std::ceil(std::log2((double) edge));
    G.anchor = max_step > 0;
    if (G.anchor) {
        levels += 1;
        int ml = (unsigned) std::log2((double) max_step) + 1;
        if (ml <= levels) levels = ml;
    }
    G.levels = levels;
    G.max_step = (uint32_t) max_step;
    G.ibs = (uint32_t) ibs;
    G.radius = 32768;  // LinearQuantizer<T>(eb): default radius
    int seq[3] = {0, 1, 2}, np = 0;
    do {
        for (int a = 0; a < 3; a++) G.perm[np][a] = seq[a];
        np++;
    } while (std::next_permutation(seq, seq + N));
    // codes per level (direction independent) and the widest pass
    pl.level_codes.assign(levels, 0);
    G.max_lines = 1;
    for (int level = levels - (G.anchor ? 1 : 0); level >= 1; level--) {
        const uint32_t s = 1u << (level - 1), B = G.ibs * s;
        uint32_t nb[3] = {1, 1, 1};
        for (int a = 0; a < N; a++) nb[a] = (G.bd[a] - 1) / B + 1;
        uint64_t total = 0;
        for (uint32_t blk = 0; blk < nb[0] * nb[1] * nb[2]; blk++) {
            uint32_t kb[3] = {0, 0, 0}, r = blk;
            for (int a = N - 1; a >= 0; a--) {
                kb[a] = r % nb[a];
                r /= nb[a];
            }
            uint32_t nn[3], g1[3], g2[3];
            for (int a = 0; a < N; a++) {
                const uint32_t b = kb[a] * B, e = std::min(b + B, G.bd[a] - 1);
                nn[a] = (e - b) / s + 1;
                g1[a] = b ? (e - b) / s : e / s + 1;
                g2[a] = b ? (e - b) / (2 * s) : e / (2 * s) + 1;
            }
            for (int p = 0; p < 6 && p < np; p++)
                for (int P = 0; P < N; P++) {
                    uint64_t lines = 1;
                    for (int Q = 0; Q < N; Q++)
                        if (Q != P) lines *= Q < P ? g1[G.perm[p][Q]] : g2[G.perm[p][Q]];
                    G.max_lines = std::max<uint32_t>(G.max_lines, (uint32_t) lines);
                    if (p == 0) total += lines * (nn[G.perm[0][P]] / 2);
                }
        }
        pl.level_codes[level - 1] = (uint32_t) total;
    }
    uint32_t head = 0;
    if (!G.anchor) {
        head = 1;
    } else {
        uint32_t c = 1;
        for (int a = 0; a < N; a++) c *= (G.bd[a] - 1) / G.max_step;  // anchors from maxStep on
        head = c;
    }
    pl.head = head;
    uint64_t n = head;
    for (auto c: pl.level_codes) n += c;
    G.ncodes = (uint32_t) n;
}

// Level-wise bound and operator of level `level` (SIC:386-422).
void fill_levels(TuneCand &c, int levels, double eb, double alpha, double beta, int interp, int dir,
                 const std::vector<uint8_t> &ops, const std::vector<uint8_t> &dirs) {
    c.eb0 = eb;
    c.recip0 = 1.0 / eb;
    c.ebf0 = float_floor(eb);
    for (int level = 1; level <= levels && level <= kTuneMaxLevels; level++) {
        double e;
        if (alpha < 0) {
            e = level >= 3 ? eb * 0.5 : eb;
        } else if (alpha >= 1) {
            double r = std::pow(alpha, level - 1);
            if (r > beta) r = beta;
            e = eb / r;
        } else {
            double r = 1 - (level - 1) * alpha;
            if (r < beta) r = beta;
            e = eb * r;
        }
        c.eb[level - 1] = e;
        c.recip[level - 1] = 1.0 / e;
        c.ebf[level - 1] = float_floor(e);
        int op = interp, d = dir;
        const int L = (int) ops.size();
        if (L > 0) {
            const int k = level - 1 < L ? level - 1 : L - 1;
            op = ops[k];
            d = dirs[k];
        }
        c.linear[level - 1] = (int8_t) (op == 0);
        c.dir[level - 1] = (int8_t) d;
    }
}

// The reference clamps the level range (SIC:360-380).
void set_range(TuneCand &c, const TuneGeom &G, int mode, int start, int end) {
    c.mode = mode;
    if (start <= 0 || start > G.levels) start = G.levels;
    if (end >= start || end < 0) end = 0;
    c.top = start == G.levels;
    if (G.anchor && c.top) start--;
    c.start = start;
    c.end = end;
}

// Device buffers of the tuner, kept by the context across calls (they only
// grow), so a tuning call allocates nothing once the sizes have been seen.
struct TuneBufs {
    DevBuf<TuneCand> d_cands;
    DevBuf<TuneItem> d_items;
    DevBuf<uint64_t> d_origins;
    DevBuf<double> d_loss;
    DevBuf<uint16_t> d_codes, d_q;
    DevBuf<float> d_outval;
    DevBuf<QSeg> d_segs;
    DevBuf<float> d_feat;
    void release() {
        d_cands.release();
        d_items.release();
        d_origins.release();
        d_loss.release();
        d_codes.release();
        d_q.release();
        d_outval.release();
        d_segs.release();
        d_feat.release();
    }
};

TuneBufs &tune_bufs(faz_gpu_ctx *ctx) {
    if (!ctx->tune_cache) {
        ctx->tune_cache = new TuneBufs();
        ctx->tune_cache_free = [](void *p) {
            TuneBufs *b = static_cast<TuneBufs *>(p);
            b->release();
            delete b;
        };
    }
    return *static_cast<TuneBufs *>(ctx->tune_cache);
}

// One tuning run over the sample blocks at `origins` (one at a time per context).
struct Tuner {
    faz_gpu_ctx *ctx;
    TunePlan pl;
    std::vector<uint64_t> origins;  // field offsets of the sample blocks
    const float *d_field = nullptr;
    double rng = 0;
    DevBuf<TuneCand> &d_cands;
    DevBuf<TuneItem> &d_items;
    DevBuf<uint64_t> &d_origins;
    DevBuf<double> &d_loss;
    DevBuf<uint16_t> &d_codes, &d_q;
    DevBuf<float> &d_outval;
    DevBuf<QSeg> &d_segs;
    std::vector<double> h_loss;

    Tuner(faz_gpu_ctx *c, TuneBufs *b)
        : ctx(c), d_cands(b->d_cands), d_items(b->d_items), d_origins(b->d_origins), d_loss(b->d_loss),
          d_codes(b->d_codes), d_q(b->d_q), d_outval(b->d_outval), d_segs(b->d_segs) {}

    // Runs every candidate on every block; h_loss[c * nblocks + b].
    int run(const std::vector<TuneCand> &cands, bool mode1) {
        hipStream_t st = ctx->stream;
        const uint32_t nb = (uint32_t) origins.size();
        std::vector<TuneItem> items;
        for (uint32_t c = 0; c < cands.size(); c++)
            for (uint32_t b = 0; b < nb; b++) items.push_back(TuneItem{c, b});
        HIPCHECK(d_cands.ensure(cands.size()));
        HIPCHECK(d_items.ensure(items.size()));
        HIPCHECK(d_loss.ensure(items.size()));
        HIPCHECK(hipMemcpyAsync(d_cands.p, cands.data(), cands.size() * sizeof(TuneCand), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(d_items.p, items.data(), items.size() * sizeof(TuneItem), hipMemcpyHostToDevice, st));
        uint16_t *codes = nullptr;
        float *outval = nullptr;
        if (mode1) {
            HIPCHECK(d_codes.ensure(items.size() * (size_t) pl.G.ncodes));
            HIPCHECK(d_outval.ensure(items.size() * (size_t) pl.G.ncodes));
            codes = d_codes.p;
            outval = d_outval.p;
        }
        hipLaunchKernelGGL(tune_kernel, dim3((uint32_t) items.size()), dim3(kTuneThreads), 0, st, pl.G,
                           d_cands.p, d_items.p, d_origins.p, d_field, codes, outval, d_loss.p);
        HIPCHECK(hipGetLastError());
        h_loss.resize(items.size());
        HIPCHECK(hipMemcpyAsync(h_loss.data(), d_loss.p, items.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return FAZ_OK;
    }

    // CompressTest (SZI:1131-1392) for one mode-1 candidate already run as
    // candidate 0: outliers in block order, codes interleaved by level,
    // encoding_lossless (SIC:784-809) through zstd.
    int compress_test(double eb, double &bitrate, double &psnr, size_t &bytes,
                      std::vector<uint8_t> *raw_out = nullptr, double *sq_out = nullptr, uint32_t cand = 0) {
        hipStream_t st = ctx->stream;
        const uint32_t nb = (uint32_t) origins.size();
        const uint64_t n = (uint64_t) nb * pl.G.ncodes;
        double square_error = 0;
        for (uint32_t b = 0; b < nb; b++) square_error += h_loss[(size_t) cand * nb + b];
        const uint16_t *codes = d_codes.p + (size_t) cand * n;
        const float *outval = d_outval.p + (size_t) cand * n;
        // outliers: code-0 slots in block-major order (the one quantizer of
        // CompressTest accumulates them block by block)
        const uint32_t nchunks = (uint32_t) ((n + kHistChunk - 1) / kHistChunk);
        HIPCHECK(ctx->chunk_zeros.ensure(nchunks));
        HIPCHECK(ctx->chunk_base.ensure(nchunks));
        HIPCHECK(ctx->freq.ensure(kHistSlots));
        HIPCHECK(ctx->first.ensure(65536));
        HIPCHECK(ctx->scalars.ensure(4));
        HIPCHECK(ctx->tiles.ensure(scan_tiles(nchunks)));
        launch_hist(codes, n, ctx->freq.p, ctx->first.p, ctx->chunk_zeros.p, 32768, st);
        launch_zero_scan(ctx->chunk_zeros.p, ctx->chunk_base.p, nchunks, ctx->tiles.p, ctx->scalars.p, st);
        uint64_t U = 0;
        HIPCHECK(hipMemcpyAsync(&U, ctx->scalars.p, 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(ctx->unpred.ensure(U));
        launch_unpred_gather(codes, n, outval, ctx->chunk_zeros.p, ctx->chunk_base.p, ctx->unpred.p, st);
        // codes in CompressTest order
        std::vector<QSeg> segs;
        uint32_t last = 0, q = 0;
        std::vector<uint32_t> counts(pl.G.levels, 0);  // quant_bin_counts after each level
        {
            uint32_t run = pl.head;
            if (pl.G.anchor) counts[pl.G.levels - 1] = 0;
            for (int level = pl.G.levels - (pl.G.anchor ? 1 : 0); level >= 1; level--) {
                run += pl.level_codes[level - 1];
                counts[level - 1] = run;
            }
        }
        for (int k = pl.G.levels - 1; k >= 0; k--) {
            const uint32_t len = counts[k] - last;
            if (len) {
                segs.push_back(QSeg{q, last, len});
                q += len * nb;
            }
            last = counts[k];
        }
        HIPCHECK(d_segs.ensure(segs.size()));
        HIPCHECK(d_q.ensure(n));
        HIPCHECK(hipMemcpyAsync(d_segs.p, segs.data(), segs.size() * sizeof(QSeg), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(interleave_kernel, dim3((uint32_t) ((n + 255) / 256)), dim3(256), 0, st, codes, d_q.p,
                           nb, pl.G.ncodes, d_segs.p, (int) segs.size());
        launch_hist(d_q.p, n, ctx->freq.p, ctx->first.p, ctx->chunk_zeros.p, 32768, st);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(ctx->h_freq.data(), ctx->freq.p, 65536 * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(ctx->h_first.data(), ctx->first.p, 65536 * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        ctx->hist_codes = nullptr;
        HuffTable tab;
        if (build_huffman(ctx->h_freq.data(), ctx->h_first.data(), 65536, tab))
            CTX_FAIL(FAZ_ERR_UNSUPPORTED, "tuner: Huffman tree could not be built");
        faz_huff_table ht{tab.offset, tab.n_states, tab.code.data(), tab.len.data()};
        uint64_t maxbits = 0;
        for (auto l: tab.len) maxbits = std::max<uint64_t>(maxbits, l);
        HIPCHECK(ctx->words.ensure((size_t) ((n * maxbits + 31) / 32 + 2)));
        size_t nbytes = 0;
        int rc = faz_gpu_huffman_encode(ctx, d_q.p, n, &ht, reinterpret_cast<uint8_t *>(ctx->words.p),
                                        ctx->words.cap * 4, &nbytes);
        if (rc) return rc;
        std::vector<uint8_t> raw(1 + 8 + 4 + 8 + U * 4 + tab.tree.size() + 8 + nbytes);
        uint8_t *w = raw.data();
        *w++ = 2;
        memcpy(w, &eb, 8), w += 8;
        const int32_t radius = 32768;
        memcpy(w, &radius, 4), w += 4;
        memcpy(w, &U, 8), w += 8;
        HIPCHECK(hipMemcpyAsync(w, ctx->unpred.p, U * 4, hipMemcpyDeviceToHost, st));
        w += U * 4;
        memcpy(w, tab.tree.data(), tab.tree.size()), w += tab.tree.size();
        const uint64_t nb64 = nbytes;
        memcpy(w, &nb64, 8), w += 8;
        HIPCHECK(hipMemcpyAsync(w, ctx->words.p, nbytes, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        std::vector<uint8_t> z(ZSTD_compressBound(raw.size()));
        const size_t zs = ZSTD_compress(z.data(), z.size(), raw.data(), raw.size(), 3);
        if (ZSTD_isError(zs)) CTX_FAIL(FAZ_ERR_INVALID, "tuner: zstd failed");
        bytes = 8 + zs;
        if (raw_out) raw_out->swap(raw);
        if (sq_out) *sq_out = square_error;
        size_t per = 1;
        for (int a = 0; a < pl.G.N; a++) per *= pl.G.bd[a];
        const size_t ele = origins.size() * per;
        bitrate = 8 * double(bytes) / ele;
        const double mse = square_error / ele;
        psnr = 20 * std::log10(rng) - 10 * std::log10(mse);
        return FAZ_OK;
    }
};

// setABwithRelBound, configuration 0 (SZI:1394-1414)
void set_ab(double rel, double &a, double &b) {
    if (rel >= 0.01) a = 2, b = 2;
    else if (rel >= 0.007) a = 1.75, b = 2;
    else if (rel >= 0.004) a = 1.5, b = 2;
    else if (rel > 0.001) a = 1.25, b = 1.5;
    else a = 1, b = 1;
}

}  // namespace

struct LearnedModel {
    faz_mlp *net;
    const float *mu, *sigma;  // input standardisation (FAZ_LEARNED_INPUTS each)
};

int learned_predict(faz_gpu_ctx *ctx, const LearnedModel &lm, const std::vector<float> &rows, size_t nrows,
                    std::vector<float> &y);
void learned_row(const float *block_feats, size_t nblocks, double eb_norm, double rel, double alpha, double beta,
                 double eb_scale, int levelwise, const uint8_t *ops, const uint8_t *dirs, const double *level_res,
                 int lorenzo, float *row);
int gf_compress_test(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *fdims, const size_t *starts,
                     size_t nblocks, size_t sbs, double eb, std::vector<uint8_t> &raw, size_t *bytes,
                     double *square_error);
int block_features(faz_gpu_ctx *ctx, const TuneGeom &G, const uint64_t *d_origins, size_t nblocks,
                   const float *d_field, float vmin, float rng, float eb_norm, std::vector<float> &feats);

// Tuning() with conf.QoZ set, RD target, no wavelets (SZI:1589-2557).  With a
// learned model the CompressTest calls of the alpha/beta search are replaced
// by the MLP's predictions of the same bitrate and PSNR.
int tune_qoz(faz_gpu_ctx *ctx, const float *d_in, const faz_tune_in *in, faz_tune_out *out,
             const LearnedModel *lm = nullptr) {
    const int N = in->N;
    if (N != 2 && N != 3) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "the QoZ tuner samples blocks of 2-D and 3-D fields only");
    memset(out, 0, sizeof(*out));
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(hipEventRecord(ctx->ev[6], ctx->stream));
    size_t num = 1;
    for (int i = 0; i < N; i++) num *= in->dims[i];
    const float rngf = (float) in->rng;
    const double rel = in->rel_eb > 0 ? in->rel_eb : in->abs_eb / rngf;
    const double rate = N == 2 ? 0.01 : 0.005;
    size_t max_step = in->max_step > 0 ? in->max_step : (N == 2 ? 64 : 32);
    int levelwise = in->levelwise > 0 ? in->levelwise : (N == 2 ? 6 : 4);
    long sbs = in->sample_block_size > 0 ? in->sample_block_size : (N == 2 ? 64 : 32);
    int max_interp_level = -1;
    for (int i = 0; i < N; i++)
        if (max_interp_level < std::ceil(std::log2((double) in->dims[i])))
            max_interp_level = (unsigned) std::ceil(std::log2((double) in->dims[i]));
    if (max_step > 0) {
        int t = (unsigned) std::log2((double) max_step);
        if (t <= max_interp_level) max_interp_level = t;
        if (levelwise > max_interp_level) levelwise = max_interp_level;
    }
    size_t shortest = in->dims[0];
    for (int i = 0; i < N; i++) shortest = std::min(shortest, in->dims[i]);
    const long minimum_sbs = 16;
    if (sbs < minimum_sbs) sbs = minimum_sbs;
    while ((size_t) sbs >= shortest) sbs /= 2;
    while (rate > 0 && sbs >= 2 * minimum_sbs && (std::pow((double) sbs, N) / (double) num) > 1.5 * rate) sbs /= 2;
    if (sbs < 1) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "field too small for sample blocks");
    if ((N == 3 && sbs + 1 > 33) || (N == 2 && (sbs + 1) * (sbs + 1) > (long) kTuneMaxPoints))
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "sample blocks of edge %ld exceed the LDS-resident tuner", sbs + 1);
    if (levelwise > kTuneMaxLevels || levelwise > 32) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "too many levels");
    uint64_t fstr[3] = {0, 0, 0};
    fstr[N - 1] = 1;
    for (int a = N - 2; a >= 0; a--) fstr[a] = fstr[a + 1] * in->dims[a + 1];
    Tuner T(ctx, &tune_bufs(ctx));
    T.d_field = d_in;
    T.rng = in->rng;
    make_tune_plan(N, (uint32_t) sbs + 1, fstr, max_step, in->interp_block_size, T.pl);
    if (T.pl.G.levels > kTuneMaxLevels) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "too many levels");
    if (T.pl.G.max_lines > kTuneMaxLines) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "sample block passes exceed the LDS line table");
    // sampleBlocks, profiling = 0, var_first = 0 (SZI:1056-1086)
    size_t stride = (size_t) (1.0 / rate);
    if (stride <= 0) stride = 1;
    size_t idx = 0;
    auto take = [&](size_t x, size_t y, size_t z) {
        if (out->n_blocks < 4096) {
            out->block_starts[out->n_blocks][0] = x;
            out->block_starts[out->n_blocks][1] = y;
            out->block_starts[out->n_blocks][2] = z;
        }
        out->n_blocks++;
        T.origins.push_back(x * fstr[0] + y * fstr[1] + (N == 3 ? z : 0));
    };
    if (N == 2) {
        for (size_t x = 0; x < in->dims[0] - sbs; x += sbs)
            for (size_t y = 0; y < in->dims[1] - sbs; y += sbs) {
                if (idx % stride == 0) take(x, y, 0);
                idx++;
            }
    } else {
        for (size_t x = 0; x < in->dims[0] - sbs; x += sbs)
            for (size_t y = 0; y < in->dims[1] - sbs; y += sbs)
                for (size_t z = 0; z < in->dims[2] - sbs; z += sbs) {
                    if (idx % stride == 0) take(x, y, z);
                    idx++;
                }
    }
    if (T.origins.empty()) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "no sample blocks");
    HIPCHECK(T.d_origins.ensure(T.origins.size()));
    HIPCHECK(hipMemcpyAsync(T.d_origins.p, T.origins.data(), T.origins.size() * 8, hipMemcpyHostToDevice,
                            ctx->stream));
    const uint32_t nb = (uint32_t) T.origins.size();
    double a0, b0;
    set_ab(rel, a0, b0);
    const int dir_cands[2] = {0, N == 2 ? 1 : 5};
    // predictor tuning (SZI:1938-2000): every (level, op, direction) on every
    // block in one launch, mode 2
    std::vector<TuneCand> cands;
    const std::vector<uint8_t> none;
    for (int level = levelwise; level > 0; level--)
        for (int op = 0; op < 2; op++)
            for (int di = 0; di < 2; di++) {
                TuneCand c;
                memset(&c, 0, sizeof(c));
                fill_levels(c, T.pl.G.levels, in->abs_eb, a0, b0, op, dir_cands[di], none, none);
                set_range(c, T.pl.G, 2, level == levelwise ? 9999 : level, level - 1);
                cands.push_back(c);
            }
    int rc = T.run(cands, false);
    if (rc) return rc;
    std::vector<uint8_t> ops(levelwise), dirs(levelwise);
    size_t ci = 0;
    for (int level = levelwise; level > 0; level--) {
        int best_op = 1, best_dir = 0;
        double best = std::numeric_limits<double>::max();
        for (int op = 0; op < 2; op++)
            for (int di = 0; di < 2; di++, ci++) {
                double loss = 0;
                for (uint32_t b = 0; b < nb; b++) loss += T.h_loss[ci * nb + b];
                out->pred_loss[level - 1][op * 2 + di] = loss;
                if (loss < best) {
                    best = loss;
                    best_op = op;
                    best_dir = dir_cands[di];
                }
            }
        ops[level - 1] = (uint8_t) best_op;
        dirs[level - 1] = (uint8_t) best_dir;
    }
    out->interp_algo = 1;  // the last candidate tried stays in the Config
    out->direction = dir_cands[1];
    // mean |mode-2 prediction error| per point of each tuned level for its
    // chosen interpolator, in units of the bound: the top tuned level's run
    // also covered every level above it
    {
        const int top = T.pl.G.levels;
        auto lattice = [&](int sh) {
            uint64_t c = 1;
            for (int a = 0; a < N; a++) c *= sh >= 63 ? 1 : (uint64_t) ((T.pl.G.bd[a] - 1) >> sh) + 1;
            return c;
        };
        for (int level = 1; level <= levelwise; level++) {
            const int k = ops[level - 1] * 2 + (dirs[level - 1] != 0);
            const uint64_t pts = lattice(level - 1) - lattice(level == levelwise ? top : level);
            const double cnt = (double) pts * nb;
            out->level_res[level - 1] = cnt > 0 ? out->pred_loss[level - 1][k] / cnt / in->abs_eb : 0.0;
        }
    }
    // alpha / beta candidates (init_alphalist / init_betalist, SZI:869-933)
    std::vector<double> alphas = {1, 1.25, 1.5, 1.75, 2};
    if (in->ab_list == 1) alphas.insert(alphas.end(), {2.25, 2.5});
    if (in->ab_list >= 2) alphas.insert(alphas.end(), {2.25, 2.5, 2.75, 3});
    const double betas[4] = {1.5, 2, 3, 4};
    double best_a = 1, best_b = 1, bestb = 9999, bestm = 0;
    out->use_interp = 1;
    out->learned_path = lm ? 0 : -1;
    const double eb_norm = in->abs_eb / in->rng;
    size_t per = 1;
    for (int a = 0; a < N; a++) per *= (size_t) sbs + 1;
    std::vector<size_t> starts3;
    for (size_t b = 0; b < out->n_blocks && b < 4096; b++)
        for (int a = 0; a < 3; a++) starts3.push_back(out->block_starts[b][a]);
    auto record = [&](double a, double b, double eb, double br, double ps, size_t bytes) {
        const int t = out->n_tests;
        if (t < 128) {
            out->test_alpha[t] = a, out->test_beta[t] = b, out->test_eb[t] = eb;
            out->test_bitrate[t] = br, out->test_psnr[t] = ps, out->test_bytes[t] = bytes;
        }
        out->n_tests++;
    };
    // CompressTest with ALGO_INTERP (SZI:1131-1392) of one candidate
    auto test = [&](double a, double b, double eb, double &br, double &ps) -> int {
        TuneCand c;
        memset(&c, 0, sizeof(c));
        fill_levels(c, T.pl.G.levels, eb, a, b, 1, dir_cands[1], ops, dirs);
        set_range(c, T.pl.G, 1, 0, 0);
        int r = T.run(std::vector<TuneCand>{c}, true);
        if (r) return r;
        size_t bytes = 0;
        r = T.compress_test(eb, br, ps, bytes);
        if (r) return r;
        record(a, b, eb, br, ps, bytes);
        return FAZ_OK;
    };
    // CompressTest with ALGO_LORENZO_REG, useFast = false (SZI:1136-1170,
    // 1283-1294, 1353-1392; lorenzoBrFix = 1), recorded with alpha = beta = -1
    auto lorenzo_test = [&](double eb, double &br, double &ps) -> int {
        if (out->n_blocks > 4096) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "the Lorenzo test takes at most 4096 sample blocks");
        std::vector<uint8_t> raw;
        size_t bytes = 0;
        double sq = 0;
        int r = gf_compress_test(ctx, d_in, N, in->dims, starts3.data(), out->n_blocks, (size_t) sbs, eb, raw,
                                 &bytes, &sq);
        if (r) return r;
        const double ele = (double) per * out->n_blocks;
        br = 8 * double(bytes) / ele;
        ps = 20 * std::log10(in->rng) - 10 * std::log10(sq / ele);
        record(-1, -1, eb, br, ps, bytes);
        return FAZ_OK;
    };
    bool exhaustive = lm == nullptr;
    if (lm) {
        // every candidate predicted in one batch; the lowest predicted cost
        // (bits/value at equal PSNR along 6.02 dB per bit) wins unless the
        // gates below ask for real CompressTests
        std::vector<float> feats;
        float mn, mx;
        if ((rc = faz_gpu_range_f32(ctx, d_in, num, &mn, &mx))) return rc;
        if ((rc = block_features(ctx, T.pl.G, T.d_origins.p, nb, d_in, mn, (float) in->rng, (float) eb_norm, feats)))
            return rc;
        struct LC {
            double a, b;
            int lorenzo;
            double br, ps, cost;
        };
        // the candidates the exhaustive search can reach: per alpha, the betas
        // up to the first one its stop rule ends on (larger betas give the
        // same bounds on every level)
        std::vector<LC> lc;
        for (double a: alphas)
            for (double b: betas) {
                if (a >= 1 && a > b) continue;
                lc.push_back(LC{a, b, 0, 0, 0, 0});
                if (a >= 1 && std::pow(a, max_interp_level - 1) <= b) break;
            }
        if (in->test_lorenzo) lc.push_back(LC{-1, -1, 1, 0, 0, 0});
        std::vector<float> rows(lc.size() * FAZ_LEARNED_INPUTS);
        for (size_t c = 0; c < lc.size(); c++)
            learned_row(feats.data(), nb, eb_norm, rel, lc[c].lorenzo ? 0.0 : lc[c].a, lc[c].lorenzo ? 0.0 : lc[c].b,
                        1.0, levelwise, ops.data(), dirs.data(), out->level_res, lc[c].lorenzo,
                        rows.data() + c * FAZ_LEARNED_INPUTS);
        bool ood = false;
        if (in->learned_lo && in->learned_hi)
            for (size_t k = 0; k < rows.size(); k++) {
                const int i = (int) (k % FAZ_LEARNED_INPUTS);
                const float z = (rows[k] - lm->mu[i]) / lm->sigma[i];
                if (!(z >= in->learned_lo[i] && z <= in->learned_hi[i])) ood = true;
            }
        if (ood) {
            out->learned_path = 2;
            exhaustive = true;
        } else {
            std::vector<float> y;
            if ((rc = learned_predict(ctx, *lm, rows, lc.size(), y))) return rc;
            for (size_t c = 0; c < lc.size(); c++) {
                lc[c].br = std::exp2((double) y[2 * c]);
                lc[c].ps = 5.0 * y[2 * c + 1] - 20 * std::log10(eb_norm);
                lc[c].cost = lc[c].br - lc[c].ps / 6.02;
                record(lc[c].a, lc[c].b, in->abs_eb, lc[c].br, lc[c].ps, 0);
            }
            // interpolation candidates by predicted cost; the Lorenzo candidate apart
            std::vector<size_t> order;
            int64_t lor = -1;
            for (size_t c = 0; c < lc.size(); c++) {
                if (lc[c].lorenzo) lor = (int64_t) c;
                else order.push_back(c);
            }
            std::stable_sort(order.begin(), order.end(), [&](size_t i, size_t j) { return lc[i].cost < lc[j].cost; });
            const int kmax = in->learned_verify > 0 ? in->learned_verify : 4;
            std::vector<size_t> ver;
            if (in->learned_margin >= 0)
                for (size_t c: order)
                    if ((int) ver.size() < kmax && lc[c].cost <= lc[order[0]].cost + in->learned_margin) ver.push_back(c);
            // the predictor choice is verified when the two predicted costs
            // are close; a clear Lorenzo win needs no interpolation test
            const bool lor_close = lor >= 0 && in->learned_margin_lorenzo >= 0 &&
                                   std::fabs(lc[lor].cost - lc[order[0]].cost) <= in->learned_margin_lorenzo;
            const bool lor_better = lor >= 0 && lc[lor].cost < lc[order[0]].cost;
            size_t pick = lor_better ? (size_t) lor : order[0];
            if (!(lor_better && !lor_close) && (ver.size() >= 2 || lor_close)) {
                // the closest candidates by real CompressTests (interpolation
                // candidates in one launch) and the lowest real cost among them
                out->learned_path = 1;
                if (ver.empty()) ver.push_back(order[0]);
                if (lor_close) ver.push_back((size_t) lor);
                std::vector<TuneCand> cands;
                std::vector<size_t> which;
                for (size_t c: ver) {
                    if (lc[c].lorenzo) continue;
                    TuneCand tc;
                    memset(&tc, 0, sizeof(tc));
                    fill_levels(tc, T.pl.G.levels, in->abs_eb, lc[c].a, lc[c].b, 1, dir_cands[1], ops, dirs);
                    set_range(tc, T.pl.G, 1, 0, 0);
                    cands.push_back(tc);
                    which.push_back(c);
                }
                double best_cost = std::numeric_limits<double>::max();
                if (!cands.empty()) {
                    if ((rc = T.run(cands, true))) return rc;
                    for (size_t k = 0; k < cands.size(); k++) {
                        double br, ps;
                        size_t bytes = 0;
                        if ((rc = T.compress_test(in->abs_eb, br, ps, bytes, nullptr, nullptr, (uint32_t) k))) return rc;
                        record(lc[which[k]].a, lc[which[k]].b, in->abs_eb, br, ps, bytes);
                        out->n_verified++;
                        const double cost = br - ps / 6.02;
                        if (cost < best_cost) best_cost = cost, pick = which[k];
                    }
                }
                if (lor_close) {
                    double br, ps;
                    if ((rc = lorenzo_test(in->abs_eb, br, ps))) return rc;
                    out->n_verified++;
                    const double cost = br - ps / 6.02;
                    if (cost < best_cost) best_cost = cost, pick = (size_t) lor;
                }
            }
            if (lc[pick].lorenzo) {
                out->use_interp = 0;
                best_a = best_b = -1;
            } else {
                best_a = lc[pick].a;
                best_b = lc[pick].b;
            }
        }
    }
    if (exhaustive) {
        // alpha / beta (SZI:2340-2420)
        for (size_t i = 0; i < alphas.size(); i++) {
            for (int j = 0; j < 4; j++) {
                const double alpha = alphas[i], beta = betas[j];
                if (alpha >= 1 && alpha > beta) continue;
                double bitrate, metric;
                if ((rc = test(alpha, beta, in->abs_eb, bitrate, metric))) return rc;
                const bool stop = (alpha >= 1 && std::pow(alpha, max_interp_level - 1) <= beta) ||
                                  (alpha < 1 && alpha * (max_interp_level - 1) <= beta);
                if (metric >= bestm && bitrate <= bestb) {
                    best_a = alpha, best_b = beta, bestb = bitrate, bestm = metric;
                } else if (metric <= bestm && bitrate >= bestb) {
                    if (stop) break;
                    continue;
                } else {
                    const double fix = bitrate / bestb;
                    double br_r, m_r;
                    if ((rc = test(alpha, beta, in->abs_eb * fix, br_r, m_r))) return rc;
                    const double a = (metric - m_r) / (bitrate - br_r);
                    const double b = metric - a * bitrate;
                    const double reg = a * bestb + b;
                    if (reg > bestm) best_a = alpha, best_b = beta, bestb = bitrate, bestm = metric;
                }
                if (stop) break;
            }
        }
        // the Lorenzo test (SZI:2431-2493)
        if (in->test_lorenzo) {
            double bitrate, metric;
            if ((rc = lorenzo_test(in->abs_eb, bitrate, metric))) return rc;
            if (metric >= bestm && bitrate <= bestb) {
                bestb = bitrate, bestm = metric, best_a = best_b = -1, out->use_interp = 0;
            } else if (metric <= bestm && bitrate >= bestb) {
                out->use_interp = 1;
            } else {
                const double fix = bitrate / bestb;
                double br_r, m_r;
                if ((rc = lorenzo_test(in->abs_eb * fix, br_r, m_r))) return rc;
                const double a = (metric - m_r) / (bitrate - br_r);
                if (a > 0) {
                    const double b = metric - a * bitrate;
                    if (a * bestb + b > bestm) bestb = bitrate, bestm = metric, best_a = best_b = -1, out->use_interp = 0;
                }
            }
        }
    }
    out->alpha = best_a;
    out->beta = best_b;
    out->levelwise = levelwise;
    memcpy(out->ops, ops.data(), levelwise);
    memcpy(out->dirs, dirs.data(), levelwise);
    out->max_step = max_step;
    out->sample_block_size = (int) sbs;
    HIPCHECK(hipEventRecord(ctx->ev[7], ctx->stream));
    HIPCHECK(hipEventSynchronize(ctx->ev[7]));
    out->ms = elapsed_ms(ctx->ev[6], ctx->ev[7]);
    return FAZ_OK;
}

}  // namespace faz

namespace faz {

// CompressTest alone on blocks at the given starts.
int compress_test_at(faz_gpu_ctx *ctx, const float *d_in, const size_t *fdims, const faz_interp_params *p,
                     const size_t *starts, size_t nblocks, size_t sbs, double rng, std::vector<uint8_t> &raw,
                     size_t &bytes, double &sq) {
    const int N = p->N;
    if (N != 2 && N != 3) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "CompressTest samples 2-D and 3-D fields only");
    if (sbs < 1 || (N == 3 && sbs + 1 > 33) || (N == 2 && (sbs + 1) * (sbs + 1) > kTuneMaxPoints))
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "sample blocks of edge %zu exceed the LDS-resident tuner", sbs + 1);
    if (nblocks == 0) CTX_FAIL(FAZ_ERR_INVALID, "no blocks");
    HIPCHECK(hipSetDevice(ctx->device));
    uint64_t fstr[3] = {0, 0, 0};
    fstr[N - 1] = 1;
    for (int a = N - 2; a >= 0; a--) fstr[a] = fstr[a + 1] * fdims[a + 1];
    Tuner T(ctx, &tune_bufs(ctx));
    T.d_field = d_in;
    T.rng = rng;
    make_tune_plan(N, (uint32_t) sbs + 1, fstr, p->max_step, p->block_size, T.pl);
    if (T.pl.G.levels > kTuneMaxLevels || T.pl.G.max_lines > kTuneMaxLines)
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "sample block outside the tuner's LDS tables");
    for (size_t b = 0; b < nblocks; b++) {
        for (int a = 0; a < N; a++)
            if (starts[3 * b + a] + sbs >= fdims[a]) CTX_FAIL(FAZ_ERR_INVALID, "block %zu outside the field", b);
        T.origins.push_back(starts[3 * b] * fstr[0] + starts[3 * b + 1] * fstr[1] + (N == 3 ? starts[3 * b + 2] : 0));
    }
    HIPCHECK(T.d_origins.ensure(T.origins.size()));
    HIPCHECK(hipMemcpyAsync(T.d_origins.p, T.origins.data(), T.origins.size() * 8, hipMemcpyHostToDevice,
                            ctx->stream));
    TuneCand c;
    memset(&c, 0, sizeof(c));
    std::vector<uint8_t> ops(p->ops, p->ops + p->levelwise), dirs(p->dirs, p->dirs + p->levelwise);
    fill_levels(c, T.pl.G.levels, p->eb, p->alpha, p->beta, p->interp_algo, p->direction, ops, dirs);
    set_range(c, T.pl.G, 1, 0, 0);
    int rc = T.run(std::vector<TuneCand>{c}, true);
    if (rc) return rc;
    double br, ps;
    return T.compress_test(p->eb, br, ps, bytes, &raw, &sq);
}

}  // namespace faz

namespace faz {

// CompressTest of many candidates on the same blocks: one tuning-mode-1
// launch for all (candidate, block) pairs, then each candidate's entropy stage.
int compress_test_batch(faz_gpu_ctx *ctx, const float *d_in, const size_t *fdims, const faz_interp_params *ps,
                        size_t ncand, const size_t *starts, size_t nblocks, size_t sbs, double rng, size_t *bytes,
                        double *sq) {
    if (!ncand) return FAZ_OK;
    const int N = ps[0].N;
    if (N != 2 && N != 3) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "CompressTest samples 2-D and 3-D fields only");
    if (sbs < 1 || (N == 3 && sbs + 1 > 33) || (N == 2 && (sbs + 1) * (sbs + 1) > kTuneMaxPoints))
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "sample blocks of edge %zu exceed the LDS-resident tuner", sbs + 1);
    if (nblocks == 0) CTX_FAIL(FAZ_ERR_INVALID, "no blocks");
    for (size_t c = 1; c < ncand; c++)
        if (ps[c].N != N || ps[c].max_step != ps[0].max_step || ps[c].block_size != ps[0].block_size)
            CTX_FAIL(FAZ_ERR_INVALID, "batched candidates must share N, maxStep and interpBlockSize");
    HIPCHECK(hipSetDevice(ctx->device));
    uint64_t fstr[3] = {0, 0, 0};
    fstr[N - 1] = 1;
    for (int a = N - 2; a >= 0; a--) fstr[a] = fstr[a + 1] * fdims[a + 1];
    Tuner T(ctx, &tune_bufs(ctx));
    T.d_field = d_in;
    T.rng = rng;
    make_tune_plan(N, (uint32_t) sbs + 1, fstr, ps[0].max_step, ps[0].block_size, T.pl);
    if (T.pl.G.levels > kTuneMaxLevels || T.pl.G.max_lines > kTuneMaxLines)
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "sample block outside the tuner's LDS tables");
    for (size_t b = 0; b < nblocks; b++) {
        for (int a = 0; a < N; a++)
            if (starts[3 * b + a] + sbs >= fdims[a]) CTX_FAIL(FAZ_ERR_INVALID, "block %zu outside the field", b);
        T.origins.push_back(starts[3 * b] * fstr[0] + starts[3 * b + 1] * fstr[1] + (N == 3 ? starts[3 * b + 2] : 0));
    }
    HIPCHECK(T.d_origins.ensure(T.origins.size()));
    HIPCHECK(hipMemcpyAsync(T.d_origins.p, T.origins.data(), T.origins.size() * 8, hipMemcpyHostToDevice,
                            ctx->stream));
    std::vector<TuneCand> cands(ncand);
    for (size_t c = 0; c < ncand; c++) {
        memset(&cands[c], 0, sizeof(TuneCand));
        std::vector<uint8_t> ops(ps[c].ops, ps[c].ops + ps[c].levelwise),
                dirs(ps[c].dirs, ps[c].dirs + ps[c].levelwise);
        fill_levels(cands[c], T.pl.G.levels, ps[c].eb, ps[c].alpha, ps[c].beta, ps[c].interp_algo, ps[c].direction,
                    ops, dirs);
        set_range(cands[c], T.pl.G, 1, 0, 0);
    }
    int rc = T.run(cands, true);
    if (rc) return rc;
    for (size_t c = 0; c < ncand; c++) {
        double br, psnr;
        if ((rc = T.compress_test(ps[c].eb, br, psnr, bytes[c], nullptr, &sq[c], (uint32_t) c))) return rc;
    }
    return FAZ_OK;
}

}  // namespace faz

extern "C" faz_status faz_gpu_compress_test_batch_f32(faz_gpu_ctx *ctx, const float *d_in, const size_t *fdims,
                                                      const faz_interp_params *ps, size_t ncand, const size_t *starts,
                                                      size_t nblocks, size_t sbs, double rng, size_t *bytes,
                                                      double *square_error) {
    if (!ctx || !d_in || !fdims || !ps || !starts || !bytes || !square_error) return FAZ_ERR_INVALID;
    return faz::compress_test_batch(ctx, d_in, fdims, ps, ncand, starts, nblocks, sbs, rng, bytes, square_error);
}

extern "C" faz_status faz_gpu_compress_test_f32(faz_gpu_ctx *ctx, const float *d_in, const size_t *fdims,
                                                const faz_interp_params *p, const size_t *starts, size_t nblocks,
                                                size_t sbs, double rng, uint8_t *h_raw, size_t cap, size_t *raw_size,
                                                size_t *bytes, double *square_error) {
    if (!ctx || !d_in || !fdims || !p || !starts || !raw_size || !bytes || !square_error) return FAZ_ERR_INVALID;
    std::vector<uint8_t> raw;
    int rc = faz::compress_test_at(ctx, d_in, fdims, p, starts, nblocks, sbs, rng, raw, *bytes, *square_error);
    if (rc) return rc;
    *raw_size = raw.size();
    if (!h_raw || cap < raw.size()) {
        ctx->err = "raw buffer too small";
        return FAZ_ERR_CAPACITY;
    }
    memcpy(h_raw, raw.data(), raw.size());
    return FAZ_OK;
}

extern "C" faz_status faz_gpu_tune_qoz_f32(faz_gpu_ctx *ctx, const float *d_in, const faz_tune_in *in,
                                           faz_tune_out *out) {
    if (!ctx || !d_in || !in || !out) return FAZ_ERR_INVALID;
    return faz::tune_qoz(ctx, d_in, in, out);
}

// ---------------------------------------------------------------------------
// Learned tuner: block statistics, then the MLP surrogate in place of the
// CompressTest search over alpha/beta.
// ---------------------------------------------------------------------------
namespace faz {
namespace {

constexpr int kFeat = 40;
constexpr int kFeatThreads = 1024;
constexpr int kFeatWaves = kFeatThreads / 64;
constexpr int kFeatAcc = 36;  // double accumulators per thread

// Sum over the wave in a fixed butterfly pattern (the same for every run).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Statistics of one sample block u = (v - vmin) / rng (see faz_gpu.h for the
// list).  Per-thread partial sums over a fixed point assignment, reduced in a
// fixed order (butterflies within each wave, then the waves in order): the
// result does not depend on scheduling.
__global__ void __launch_bounds__(kFeatThreads) block_features_kernel(TuneGeom G, const uint64_t *__restrict__ origins,
                                                                      const float *__restrict__ field, float vmin,
                                                                      float rng, float eb_norm,
                                                                      float *__restrict__ out) {
    __shared__ float s_u[kTuneMaxPoints];
    __shared__ double s_wpart[kFeatAcc + 1][kFeatWaves];
    __shared__ float s_wmax[5][kFeatWaves];
    __shared__ double s_res[kFeatAcc + 8];
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t tid = threadIdx.x;
    const uint64_t org = origins[blockIdx.x];
    const int N = G.N;
    const uint32_t d0 = G.bd[0], d1 = G.bd[1], d2 = N == 3 ? G.bd[2] : 1;
    const uint32_t npts = d0 * d1 * d2;
    for (uint32_t q = tid; q < npts; q += blockDim.x) {
        const uint32_t z = q % d2, r = q / d2, y = r % d1, x = r / d1;
        const uint64_t f = org + (uint64_t) x * G.fstr[0] + (uint64_t) y * G.fstr[1] + (N == 3 ? z : 0);
        s_u[q] = (field[f] - vmin) / rng;
    }
    __syncthreads();
    const uint32_t dd[3] = {d0, d1, d2};
    const uint32_t st[3] = {d1 * d2, d2, 1};
    double acc[kFeatAcc];
#pragma unroll
    for (int i = 0; i < kFeatAcc; i++) acc[i] = 0.0;
    float mn = 3.0e38f, mx = -3.0e38f, lmax[3] = {0.f, 0.f, 0.f};
    // acc: 0 sum, 1 sum2, [2 + 6*si + a] sum|d|, [5 + 6*si + a] sum d^2 (stride si),
    // 20 + 2*si sum|L|, 21 + 2*si sum L^2, 26 + a lag1 cross, 29 + a lag2 cross,
    // 32 sum L1^4, 33 + a count |d1| < eb
    const int strides[3] = {1, 2, 4};
    for (uint32_t q = tid; q < npts; q += blockDim.x) {
        const uint32_t c[3] = {q / (d1 * d2), (q / d2) % d1, q % d2};
        const float u = s_u[q];
        acc[0] += u;
        acc[1] += (double) u * u;
        mn = fminf(mn, u);
        mx = fmaxf(mx, u);
#pragma unroll
        for (int si = 0; si < 3; si++) {
            const uint32_t s = strides[si];
            bool interior = true;
            float L = 0.f;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if (a >= N) continue;
                if (c[a] + s < dd[a]) {
                    const float d = s_u[q + s * st[a]] - u;
                    acc[2 + 6 * si + a] += fabsf(d);
                    acc[5 + 6 * si + a] += (double) d * d;
                    if (si == 0 && fabsf(d) < eb_norm) acc[33 + a] += 1.0;
                }
                if (c[a] >= s && c[a] + s < dd[a])
                    L += s_u[q + s * st[a]] - 2.f * u + s_u[q - s * st[a]];
                else
                    interior = false;
            }
            if (interior) {
                acc[20 + 2 * si] += fabsf(L);
                acc[21 + 2 * si] += (double) L * L;
                lmax[si] = fmaxf(lmax[si], fabsf(L));
                if (si == 0) acc[32] += (double) L * L * L * L;
            }
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (a >= N) continue;
            if (c[a] + 1 < dd[a]) acc[26 + a] += (double) u * s_u[q + st[a]];
            if (c[a] + 2 < dd[a]) acc[29 + a] += (double) u * s_u[q + 2 * st[a]];
        }
    }
    // deterministic reductions: each wave's butterfly sum, then the waves in order
#pragma unroll
    for (int i = 0; i < kFeatAcc; i++) {
        const double w = wave_sum(acc[i]);
        if (lane == 0) s_wpart[i][wave] = w;
    }
    {
        const float mxs[5] = {-mn, mx, lmax[0], lmax[1], lmax[2]};
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const float w = wave_max(mxs[i]);
            if (lane == 0) s_wmax[i][wave] = w;
        }
    }
    __syncthreads();
    if (tid < kFeatAcc) {
        double t = 0;
        for (int k = 0; k < kFeatWaves; k++) t += s_wpart[tid][k];
        s_res[tid] = t;
    } else if (tid >= 64 && tid < 69) {
        const int i = tid - 64;
        float t = s_wmax[i][0];
        for (int k = 1; k < kFeatWaves; k++) t = fmaxf(t, s_wmax[i][k]);
        s_res[kFeatAcc + i] = i == 0 ? -t : t;
    }
    __syncthreads();
    // second pass: share of |L1| above twice its RMS
    uint64_t n_int1 = 1;
    for (int a = 0; a < N; a++) n_int1 *= dd[a] - 2;
    const double rms1 = n_int1 ? sqrt(s_res[21] / (double) n_int1) : 0.0;
    double big = 0;
    for (uint32_t q = tid; q < npts; q += blockDim.x) {
        const uint32_t c[3] = {q / (d1 * d2), (q / d2) % d1, q % d2};
        bool interior = true;
        float L = 0.f;
        for (int a = 0; a < N; a++) {
            if (c[a] >= 1 && c[a] + 1 < dd[a])
                L += s_u[q + st[a]] - 2.f * s_u[q] + s_u[q - st[a]];
            else
                interior = false;
        }
        if (interior && fabsf(L) > 2.0 * rms1) big += 1.0;
    }
    {
        const double w = wave_sum(big);
        if (lane == 0) s_wpart[kFeatAcc][wave] = w;
    }
    __syncthreads();
    if (tid == 0) {
        double t = 0;
        for (int k = 0; k < kFeatWaves; k++) t += s_wpart[kFeatAcc][k];
        const double n = (double) npts;
        const double mean = s_res[0] / n, var = fmax(s_res[1] / n - mean * mean, 0.0);
        float *o = out + (uint64_t) blockIdx.x * kFeat;
        o[0] = (float) mean;
        o[1] = (float) sqrt(var);
        o[2] = (float) s_res[kFeatAcc + 0];
        o[3] = (float) s_res[kFeatAcc + 1];
        for (int si = 0; si < 3; si++) {
            const uint32_t s = strides[si];
            for (int a = 0; a < 3; a++) {
                double cnt = 0;
                if (a < N && dd[a] > s) cnt = n / dd[a] * (dd[a] - s);
                o[4 + 9 * si + a] = cnt > 0 ? (float) (s_res[2 + 6 * si + a] / cnt) : 0.f;
                o[7 + 9 * si + a] = cnt > 0 ? (float) sqrt(s_res[5 + 6 * si + a] / cnt) : 0.f;
            }
            double ni = 1;
            for (int a = 0; a < N; a++) ni *= dd[a] > 2 * s ? (double) (dd[a] - 2 * s) : 0.0;
            o[10 + 9 * si] = ni > 0 ? (float) (s_res[20 + 2 * si] / ni) : 0.f;
            o[11 + 9 * si] = ni > 0 ? (float) sqrt(s_res[21 + 2 * si] / ni) : 0.f;
            o[12 + 9 * si] = (float) s_res[kFeatAcc + 2 + si];
        }
        for (int a = 0; a < 3; a++) {
            for (int lag = 1; lag <= 2; lag++) {
                double cnt = 0, cr = 0;
                if (a < N && dd[a] > (uint32_t) lag) {
                    cnt = n / dd[a] * (dd[a] - lag);
                    cr = var > 0 ? (s_res[(lag == 1 ? 26 : 29) + a] / cnt - mean * mean) / var : 0.0;
                }
                o[(lag == 1 ? 31 : 34) + a] = (float) cr;
            }
        }
        const double ni1 = (double) n_int1;
        const double m2 = ni1 > 0 ? s_res[21] / ni1 : 0.0;
        o[37] = m2 > 0 ? (float) ((s_res[32] / ni1) / (m2 * m2)) : 0.f;
        o[38] = ni1 > 0 ? (float) (t / ni1) : 0.f;
        double fr = 0;
        int na = 0;
        for (int a = 0; a < N; a++) {
            if (dd[a] > 1) {
                fr += s_res[33 + a] / (n / dd[a] * (dd[a] - 1));
                na++;
            }
        }
        o[39] = na ? (float) (fr / na) : 0.f;
    }
}

}  // namespace
}  // namespace faz

namespace faz {

int block_features(faz_gpu_ctx *ctx, const TuneGeom &G, const uint64_t *d_origins, size_t nblocks,
                   const float *d_field, float vmin, float rng, float eb_norm, std::vector<float> &feats) {
    DevBuf<float> &d_out = tune_bufs(ctx).d_feat;
    HIPCHECK(d_out.ensure(nblocks * kFeat));
    hipLaunchKernelGGL(block_features_kernel, dim3((uint32_t) nblocks), dim3(kFeatThreads), 0, ctx->stream, G,
                       d_origins, d_field, vmin, rng, eb_norm, d_out.p);
    HIPCHECK(hipGetLastError());
    feats.resize(nblocks * kFeat);
    HIPCHECK(hipMemcpyAsync(feats.data(), d_out.p, feats.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return FAZ_OK;
}

// Input row of the MLP (FAZ_LEARNED_INPUTS values): the block statistics
// averaged over the sample blocks, magnitudes as log2 of their ratio to the
// normalised bound, then the candidate: its alpha/beta and interpolators, the
// bound, the prediction error of levels 1..4 over that level's bound of the
// candidate, the Lorenzo flag.
void learned_row(const float *block_feats, size_t nblocks, double eb_norm, double rel, double alpha, double beta,
                 double eb_scale, int levelwise, const uint8_t *ops, const uint8_t *dirs, const double *level_res,
                 int lorenzo, float *row) {
    for (int i = 0; i < kFeat; i++) {
        double s = 0;
        for (size_t b = 0; b < nblocks; b++) {
            double v = block_feats[b * kFeat + i];
            const bool magnitude = i == 1 || (i >= 4 && i <= 30);
            if (magnitude) v = std::log2(std::max(v, 1e-30) / eb_norm);
            if (i == 37) v = std::log2(std::max(v, 1e-30));
            s += v;
        }
        row[i] = (float) (s / (double) nblocks);
    }
    double mop = 0, mdir = 0;
    for (int l = 0; l < levelwise; l++) {
        mop += ops[l];
        mdir += dirs[l] != 0;
    }
    row[40] = lorenzo ? 0.f : (float) (alpha - 1);
    row[41] = lorenzo ? 0.f : (float) (beta / 4);
    row[42] = lorenzo ? 0.f : (float) std::log2(std::max(beta, 1e-6));
    row[43] = (float) ops[0];
    row[44] = (float) (dirs[0] != 0);
    row[45] = (float) (mop / levelwise);
    row[46] = (float) (mdir / levelwise);
    row[47] = (float) std::log10(rel);
    row[48] = (float) std::log2(eb_scale);
    // level l's bound is eb * scale / min(alpha^(l-1), beta) (SIC:392-398)
    for (int l = 1; l <= 4; l++) {
        const int k = std::min(l, levelwise) - 1;
        double ratio = 1.0;
        if (!lorenzo && alpha >= 1) ratio = std::min(std::pow(alpha, l - 1), beta);
        const double r = level_res ? level_res[k] : 0.0;
        row[48 + l] = (float) std::log2(std::max(r * ratio / eb_scale, 1e-6));
    }
    row[53] = lorenzo ? 1.f : 0.f;
}

int learned_predict(faz_gpu_ctx *ctx, const LearnedModel &lm, const std::vector<float> &rows, size_t nrows,
                    std::vector<float> &y) {
    std::vector<float> x(rows.size());
    for (size_t r = 0; r < nrows; r++)
        for (int i = 0; i < FAZ_LEARNED_INPUTS; i++) {
            const size_t k = r * FAZ_LEARNED_INPUTS + i;
            x[k] = (rows[k] - lm.mu[i]) / lm.sigma[i];
        }
    DevBuf<float> dx, dy;
    HIPCHECK(dx.ensure(x.size()));
    HIPCHECK(dy.ensure(nrows * 2));
    HIPCHECK(hipMemcpyAsync(dx.p, x.data(), x.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    int rc = faz_mlp_forward(ctx, lm.net, dx.p, (int) nrows, dy.p);
    if (rc) return rc;
    y.resize(nrows * 2);
    HIPCHECK(hipMemcpy(y.data(), dy.p, y.size() * 4, hipMemcpyDeviceToHost));
    dx.release();
    dy.release();
    return FAZ_OK;
}

}  // namespace faz

extern "C" {

faz_status faz_gpu_block_features_f32(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims,
                                      const size_t *starts, size_t nblocks, size_t sbs, float vmin, float rng,
                                      double eb_norm, float *h_feats) {
    if (!ctx || !d_in || !dims || !starts || !h_feats || (N != 2 && N != 3) || nblocks == 0) return FAZ_ERR_INVALID;
    if ((N == 3 && sbs + 1 > 33) || (N == 2 && (sbs + 1) * (sbs + 1) > faz::kTuneMaxPoints))
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "sample blocks of edge %zu exceed the LDS-resident kernels", sbs + 1);
    HIPCHECK(hipSetDevice(ctx->device));
    uint64_t fstr[3] = {0, 0, 0};
    fstr[N - 1] = 1;
    for (int a = N - 2; a >= 0; a--) fstr[a] = fstr[a + 1] * dims[a + 1];
    faz::TunePlan pl;
    faz::make_tune_plan(N, (uint32_t) sbs + 1, fstr, 0, 32, pl);
    std::vector<uint64_t> org;
    for (size_t b = 0; b < nblocks; b++) {
        for (int a = 0; a < N; a++)
            if (starts[3 * b + a] + sbs >= dims[a]) CTX_FAIL(FAZ_ERR_INVALID, "block %zu outside the field", b);
        org.push_back(starts[3 * b] * fstr[0] + starts[3 * b + 1] * fstr[1] + (N == 3 ? starts[3 * b + 2] : 0));
    }
    faz::DevBuf<uint64_t> d_org;
    HIPCHECK(d_org.ensure(org.size()));
    HIPCHECK(hipMemcpy(d_org.p, org.data(), org.size() * 8, hipMemcpyHostToDevice));
    std::vector<float> f;
    int rc = faz::block_features(ctx, pl.G, d_org.p, nblocks, d_in, vmin, rng, (float) eb_norm, f);
    d_org.release();
    if (rc) return rc;
    memcpy(h_feats, f.data(), f.size() * 4);
    return FAZ_OK;
}

faz_status faz_learned_input(const float *block_feats, size_t nblocks, double eb_norm, double rel_eb, double alpha,
                             double beta, double eb_scale, int levelwise, const uint8_t *ops, const uint8_t *dirs,
                             const double *level_res, int lorenzo, float *row) {
    if (!block_feats || !nblocks || !ops || !dirs || !row || levelwise < 1 || !(eb_norm > 0) || !(rel_eb > 0) ||
        !(eb_scale > 0))
        return FAZ_ERR_INVALID;
    faz::learned_row(block_feats, nblocks, eb_norm, rel_eb, alpha, beta, eb_scale, levelwise, ops, dirs, level_res,
                     lorenzo, row);
    return FAZ_OK;
}

faz_status faz_learned_target(double bitrate, double psnr, double eb_norm, double eb_scale, float *t) {
    if (!t || !(eb_norm > 0) || !(eb_scale > 0) || !(bitrate > 0)) return FAZ_ERR_INVALID;
    t[0] = (float) std::log2(bitrate);
    t[1] = (float) ((psnr + 20 * std::log10(eb_norm * eb_scale)) / 5);
    return FAZ_OK;
}

faz_status faz_gpu_tune_learned_f32(faz_gpu_ctx *ctx, const float *d_in, const faz_tune_in *in, faz_mlp *net,
                                    const float *mu, const float *sigma, faz_tune_out *out) {
    if (!ctx || !d_in || !in || !net || !mu || !sigma || !out) return FAZ_ERR_INVALID;
    faz::LearnedModel lm{net, mu, sigma};
    return faz::tune_qoz(ctx, d_in, in, out, &lm);
}

}  // extern "C"
