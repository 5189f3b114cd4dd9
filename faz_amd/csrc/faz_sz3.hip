// The default (SZ3) tuning mode on gfx950: Tuning with predictorTuningRate
// = 0 (include/QoZ/api/impl/SZInterp.hpp:2096-2180) and the Lorenzo branch of
// SZ_compress_Interp_lorenzo (SZI:2783-2898).
//
//   * sampling (utils/Extraction.hpp:17-212): a gather kernel;
//   * the interpolation test (do_not_use_this_interp_compress_block_test,
//     SZI:558-576) runs the deprecated block interpolation compressor
//     (compressor/deprecated/SZBlockInterpolationCompressor.hpp:110-458):
//     blocks of the sampling block edge, each with its own first point and
//     levels and no shared faces, so a level's pass over every block is one
//     launch with one thread per line; codes land at closed-form positions
//     (block, level, pass, line, (i-1)/2) because a line emits its odd points
//     in increasing order;
//   * the Lorenzo test reuses the 3-D fast-frontend and general-frontend
//     engines on the sample;
//   * optimize_quant_invl_3d (utils/QuantOptimization.hpp:89-147): the host
//     replays the reference's pointer walk for the sample positions, a kernel
//     histograms the 1-layer Lorenzo errors there.
#include "faz_gpu.h"
#include "faz_ctx.h"
#include "faz_device.h"
#include "faz_internal.h"
#include "../host/raw_api.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#pragma clang fp contract(off)

extern "C" {
// libzstd (system shared library); prototypes as published in zstd.h
size_t ZSTD_compress(void *dst, size_t dstCapacity, const void *src, size_t srcSize, int compressionLevel);
size_t ZSTD_compressBound(size_t srcSize);
unsigned ZSTD_isError(size_t code);
}

using namespace faz;

namespace faz {
namespace {

struct BIGeom {
    int N;
    uint32_t dims[3];
    uint64_t st[3];
    uint32_t bs;
    uint32_t nb[3];
    uint64_t nblocks;
};

// One pass of a level: interpolated axis, outer line axis and its step
// factor, inner line axis and its step factor (-1: none).
struct BIPass {
    int a, u, su, w, sw;
};

// block_interpolation (SBI:263-458): the passes of direction dir for N.
__host__ __device__ inline int bi_passes(int N, int dir, BIPass *P) {
    if (N == 1) {
        P[0] = {0, -1, 1, -1, 1};
        return 1;
    }
    if (N == 2) {
        if (dir == 0) {
            P[0] = {0, 1, 2, -1, 1};
            P[1] = {1, 0, 1, -1, 1};
        } else {
            P[0] = {1, 0, 2, -1, 1};
            P[1] = {0, 1, 1, -1, 1};
        }
        return 2;
    }
    const int T[6][3][5] = {
        {{0, 1, 2, 2, 2}, {1, 0, 1, 2, 2}, {2, 0, 1, 1, 1}}, {{0, 1, 2, 2, 2}, {2, 0, 1, 1, 2}, {1, 0, 1, 2, 1}},
        {{1, 0, 2, 2, 2}, {0, 1, 1, 2, 2}, {2, 0, 1, 1, 1}}, {{1, 0, 2, 2, 2}, {2, 0, 2, 1, 1}, {0, 1, 1, 2, 1}},
        {{2, 0, 2, 1, 2}, {0, 1, 2, 2, 1}, {1, 0, 1, 2, 1}}, {{2, 0, 2, 1, 2}, {1, 0, 2, 2, 1}, {0, 1, 1, 2, 1}},
    };
    for (int p = 0; p < 3; p++) P[p] = {T[dir][p][0], T[dir][p][1], T[dir][p][2], T[dir][p][3], T[dir][p][4]};
    return 3;
}

__host__ __device__ inline uint32_t bi_len(uint32_t b, uint32_t bs, uint32_t d) {
    return b * bs + bs > d ? d - b * bs : bs;
}

__host__ __device__ inline uint32_t bi_count(uint32_t extent, uint32_t step) { return (extent - 1) / step + 1; }

// Codes of one pass of a block with extents bd at stride s.
__host__ __device__ inline uint64_t bi_pass_codes(const uint32_t *bd, const BIPass &p, uint32_t s) {
    const uint32_t n = bi_count(bd[p.a], s);
    const uint64_t lines = (uint64_t) (p.u >= 0 ? bi_count(bd[p.u], s * p.su) : 1) *
                           (p.w >= 0 ? bi_count(bd[p.w], s * p.sw) : 1);
    return lines * (n / 2);
}

__host__ __device__ inline uint32_t bi_levels(int N, const uint32_t *bd) {
    uint32_t m = 1;
    for (int d = 0; d < N; d++) {
        uint32_t l = 0;
        while ((1u << l) < bd[d]) l++;  // ceil(log2(bd)), 0 for bd = 1
        m = m > l ? m : l;
    }
    return m;
}

struct BIQuant {
    double eb, recip;
    float ebf;
    int radius;
};

// Block coordinates, extents and first code of block b.
__device__ inline void bi_block(const BIGeom &G, uint64_t b, uint32_t *B, uint32_t *bd, uint64_t &base) {
    uint32_t k[3] = {0, 0, 0};
    uint64_t r = b;
    for (int d = G.N - 1; d >= 0; d--) {
        k[d] = (uint32_t) (r % G.nb[d]);
        r /= G.nb[d];
    }
    base = 0;
    uint64_t pre = 1;
    for (int d = 0; d < G.N; d++) {
        B[d] = k[d] * G.bs;
        bd[d] = bi_len(k[d], G.bs, G.dims[d]);
        uint64_t post = 1;
        for (int e = d + 1; e < G.N; e++) post *= G.dims[e];
        base += pre * ((uint64_t) B[d]) * post;
        pre *= bd[d];
    }
}

// The first point of every block, predicted by 0 (SBI:139).
__global__ void bi_first_kernel(BIGeom G, BIQuant Q, const float *__restrict__ in, float *__restrict__ rec,
                                uint16_t *__restrict__ codes, float *__restrict__ outval) {
    const uint64_t b = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= G.nblocks) return;
    uint32_t B[3] = {0, 0, 0}, bd[3] = {1, 1, 1};
    uint64_t base;
    bi_block(G, b, B, bd, base);
    uint64_t at = 0;
    for (int d = 0; d < G.N; d++) at += (uint64_t) B[d] * G.st[d];
    float v;
    const float x = in[at];
    const int code = quantize_select(x, 0.0f, Q.eb, Q.recip, Q.ebf, Q.radius, &v);
    codes[base] = (uint16_t) code;
    if (code == 0) outval[base] = x;
    rec[at] = v;
}

// Pass `pi` of level `level` over every block: one thread per line
// (blockIdx.y = block), block_interpolation_1d (SBI:205-262) along it.
__global__ void __launch_bounds__(256) bi_pass_kernel(BIGeom G, BIQuant Q, int dir, int pi, uint32_t level,
                                                      int linear, const float *__restrict__ in,
                                                      float *__restrict__ rec, uint16_t *__restrict__ codes,
                                                      float *__restrict__ outval, uint32_t *__restrict__ err) {
    const uint64_t b = blockIdx.y;
    uint32_t B[3] = {0, 0, 0}, bd[3] = {1, 1, 1};
    uint64_t base;
    bi_block(G, b, B, bd, base);
    BIPass P[3];
    const int np = bi_passes(G.N, dir, P);
    const uint32_t top = bi_levels(G.N, bd);
    if (level > top) return;
    const uint32_t s = 1u << (level - 1);
    const BIPass p = P[pi];
    const uint32_t nu = p.u >= 0 ? bi_count(bd[p.u], s * p.su) : 1;
    const uint32_t nw = p.w >= 0 ? bi_count(bd[p.w], s * p.sw) : 1;
    const uint64_t line = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= (uint64_t) nu * nw) return;
    const uint32_t n = bi_count(bd[p.a], s);
    if (n <= 1) return;
    // codes of the block before this pass: its first point, coarser levels, earlier passes
    uint64_t off = base + 1;
    for (uint32_t l = top; l > level; l--)
        for (int q = 0; q < np; q++) off += bi_pass_codes(bd, P[q], 1u << (l - 1));
    for (int q = 0; q < pi; q++) off += bi_pass_codes(bd, P[q], s);
    off += line * (n / 2);
    const uint32_t iu = (uint32_t) (line / nw), iw = (uint32_t) (line % nw);
    uint64_t at = 0;
    for (int d = 0; d < G.N; d++) {
        uint32_t c = B[d];
        if (d == p.u) c += iu * s * p.su;
        if (d == p.w) c += iw * s * p.sw;
        at += (uint64_t) c * G.st[d];
    }
    const uint64_t S = (uint64_t) s * G.st[p.a];
    auto R = [&](uint32_t i) { return rec[at + i * S]; };
    auto put = [&](uint32_t i, float pred) {
        const float x = in[at + i * S];
        float v;
        const int code = quantize_select(x, pred, Q.eb, Q.recip, Q.ebf, Q.radius, &v);
        if ((uint32_t) code > 65535u) atomicOr(err, 1u);
        const uint64_t c = off + (i - 1) / 2;
        codes[c] = (uint16_t) code;
        if (code == 0) outval[c] = x;
        rec[at + i * S] = v;
    };
    if (linear || n < 5) {
        for (uint32_t i = 1; i + 1 < n; i += 2) put(i, i_linear(R(i - 1), R(i + 1)));
        if (n % 2 == 0) {
            if (n < 4) put(n - 1, R(n - 2));
            else put(n - 1, i_linear1(R(n - 4), R(n - 2)));
        }
    } else {
        put(1, i_quad1(R(0), R(2), R(4)));
        for (uint32_t i = 3; i + 3 < n; i += 2) put(i, i_cubic(R(i - 3), R(i - 1), R(i + 1), R(i + 3)));
        if (n % 2 == 0) {
            put(n - 3, i_quad2(R(n - 6), R(n - 4), R(n - 2)));
            put(n - 1, i_quad3(R(n - 6), R(n - 4), R(n - 2)));
        } else {
            put(n - 2, i_quad2(R(n - 5), R(n - 3), R(n - 1)));
        }
    }
}

// sampling<T,N>: sample index -> field index per axis.
__global__ void sz3_sample_kernel(const float *__restrict__ in, int N, uint64_t sn, uint64_t s0, uint64_t s1,
                                  uint64_t sd1, uint64_t sd2, uint64_t st0, uint64_t st1, uint64_t dmin, uint64_t sb,
                                  float *__restrict__ out) {
    for (uint64_t q = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; q < sn; q += (uint64_t) gridDim.x * blockDim.x) {
        uint64_t c[3] = {0, 0, 0};
        if (N == 1) {
            c[0] = q;
        } else if (N == 2) {
            c[0] = q / sd1, c[1] = q % sd1;
        } else {
            c[2] = q % sd2;
            c[1] = q / sd2 % sd1;
            c[0] = q / (sd1 * sd2);
        }
        uint64_t src = 0;
        const uint64_t stv[3] = {s0, s1, 1};
        for (int d = 0; d < N; d++) {
            const uint64_t bd = c[d] / (2 * sb), i = c[d] % (2 * sb);
            const uint64_t di = i < sb ? i + sb : dmin - 3 * sb + i;
            src += (bd * dmin + di) * (N == 1 ? 1 : (N == 2 ? (d == 0 ? st0 : 1) : stv[d]));
        }
        out[q] = in[src];
    }
}

// 1-layer 3-D Lorenzo error histogram at the sample positions
// (QO:109-117, MetaLorenzoPredictor.hpp:42-46).
__global__ void sz3_qintv_kernel(const float *__restrict__ in, const uint64_t *__restrict__ pos, uint64_t n,
                                 uint64_t r23, uint64_t r3, double precision, uint32_t *__restrict__ hist) {
    for (uint64_t k = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t) gridDim.x * blockDim.x) {
        const float *p = in + pos[k];
        const int64_t a = (int64_t) r3, c = (int64_t) r23;
        const float pred = p[-1] + p[-a] + p[-c] - p[-a - 1] - p[-c - 1] - p[-c - a] + p[-c - a - 1];
        const float e = fabsf(pred - *p);
        uint64_t idx = (uint64_t) (((double) e / precision + 1) / 2);
        if (idx > 32767) idx = 32767;
        atomicAdd(&hist[idx], 1u);
    }
}

}  // namespace
}  // namespace faz

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
namespace faz {
namespace {

template<class V>
void sput(std::vector<uint8_t> &out, const V &v) {
    const uint8_t *b = reinterpret_cast<const uint8_t *>(&v);
    out.insert(out.end(), b, b + sizeof(V));
}

float sz3_sampling_ratio(size_t block, size_t n, size_t dmin, const size_t *dims, int N) {
    size_t sample_n = 1;
    for (int d = 0; d < N; d++) sample_n *= dims[d] / dmin * 2 * block;
    return (float) (sample_n * 1.0 / n);
}

// sampling<T,N> geometry (Extraction.hpp:17-212): false when sb < 9.
bool sz3_geometry(int N, const size_t *dims, size_t *sdims, size_t &sb, size_t &dmin) {
    size_t num = 1;
    dmin = dims[0];
    for (int d = 0; d < N; d++) num *= dims[d], dmin = std::min(dmin, dims[d]);
    sb = dmin;
    while (sz3_sampling_ratio(sb, num, dmin, dims, N) > 0.035) sb--;
    if (sb * 2 > dmin) sb = dmin / 2;
    if (sb < 9) {
        for (int d = 0; d < N; d++) sdims[d] = dims[d];
        return false;
    }
    for (int d = 0; d < N; d++) sdims[d] = dims[d] / dmin * 2 * sb;
    return true;
}

int sz3_sample(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims, float *d_out, size_t *sdims,
               size_t &sb, bool &sampled) {
    size_t dmin;
    sampled = sz3_geometry(N, dims, sdims, sb, dmin);
    if (!sampled) return FAZ_OK;
    uint64_t sn = 1;
    for (int d = 0; d < N; d++) sn *= sdims[d];
    const uint64_t st0 = N == 3 ? (uint64_t) dims[1] * dims[2] : (N == 2 ? dims[1] : 1);
    const uint64_t st1 = N == 3 ? dims[2] : 1;
    hipLaunchKernelGGL(sz3_sample_kernel, dim3((unsigned) std::min<uint64_t>((sn + 255) / 256, 8192)), dim3(256), 0,
                       ctx->stream, d_in, N, sn, st0, st1, N >= 2 ? (uint64_t) sdims[1] : 1,
                       N == 3 ? (uint64_t) sdims[2] : 1, st0, st1, (uint64_t) dmin, (uint64_t) sb, d_out);
    HIPCHECK(hipGetLastError());
    return FAZ_OK;
}

// do_not_use_this_interp_compress_block_test (SZI:558-576) on the device
// array d_in (N <= 3): the engine buffer in `raw` and its zstd-framed size.
int block_interp_test(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims, double eb, int algo, int dir,
                      size_t bs, std::vector<uint8_t> &raw, size_t &bytes) {
    if (N < 1 || N > 3) CTX_FAIL(FAZ_ERR_INVALID, "block interpolation test: N must be 1..3");
    int nf = N == 1 ? 1 : (N == 2 ? 2 : 6);
    if (dir < 0 || dir >= nf || algo < 0 || algo > 1 || bs < 1 || !(eb > 0))
        CTX_FAIL(FAZ_ERR_INVALID, "block interpolation test: bad interpolator, direction, block or bound");
    BIGeom G;
    memset(&G, 0, sizeof(G));
    G.N = N;
    G.bs = (uint32_t) bs;
    uint64_t s = 1;
    G.nblocks = 1;
    for (int d = N - 1; d >= 0; d--) {
        G.dims[d] = (uint32_t) dims[d];
        G.st[d] = s;
        s *= dims[d];
        G.nb[d] = (uint32_t) ((dims[d] - 1) / bs + 1);
        G.nblocks *= G.nb[d];
    }
    const uint64_t n = s;
    if (n >= (1ull << 32)) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 2^32-1 points");
    if (G.nblocks > 65535) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 65535 blocks");
    BIQuant Q{eb, 1.0 / eb, float_floor(eb), 32768};
    hipStream_t st = ctx->stream;
    HIPCHECK(ctx->gf_recon.ensure(n));
    HIPCHECK(ctx->codes.ensure(n));
    HIPCHECK(ctx->outval.ensure(n));
    HIPCHECK(ctx->err_flag.ensure(1));
    HIPCHECK(hipMemcpyAsync(ctx->gf_recon.p, d_in, n * 4, hipMemcpyDeviceToDevice, st));
    launch_fill_u32(ctx->err_flag.p, 1, 0u, st);
    hipLaunchKernelGGL(bi_first_kernel, dim3((unsigned) ((G.nblocks + 63) / 64)), dim3(64), 0, st, G, Q, d_in,
                       ctx->gf_recon.p, ctx->codes.p, ctx->outval.p);
    // the largest block fixes the level count and line counts of the launches
    uint32_t full[3] = {1, 1, 1};
    for (int d = 0; d < N; d++) full[d] = std::min<uint32_t>(G.bs, G.dims[d]);
    const uint32_t top = bi_levels(N, full);
    BIPass P[3];
    const int np = bi_passes(N, dir, P);
    for (uint32_t l = top; l >= 1; l--) {
        const uint32_t sl = 1u << (l - 1);
        for (int pi = 0; pi < np; pi++) {
            const BIPass &p = P[pi];
            const uint64_t lines = (uint64_t) (p.u >= 0 ? bi_count(full[p.u], sl * p.su) : 1) *
                                   (p.w >= 0 ? bi_count(full[p.w], sl * p.sw) : 1);
            hipLaunchKernelGGL(bi_pass_kernel, dim3((unsigned) ((lines + 255) / 256), (unsigned) G.nblocks), dim3(256),
                               0, st, G, Q, dir, pi, l, algo == 0 ? 1 : 0, d_in, ctx->gf_recon.p, ctx->codes.p,
                               ctx->outval.p, ctx->err_flag.p);
        }
    }
    HIPCHECK(hipGetLastError());
    // outliers in code order, histogram, tree, encode
    const uint32_t nchunks = (uint32_t) ((n + kHistChunk - 1) / kHistChunk);
    HIPCHECK(ctx->scalars.ensure(4));
    HIPCHECK(ctx->tiles.ensure(scan_tiles(nchunks)));
    HIPCHECK(ctx->chunk_zeros.ensure(nchunks));
    HIPCHECK(ctx->chunk_base.ensure(nchunks));
    HIPCHECK(ctx->freq.ensure(kHistSlots));
    HIPCHECK(ctx->first.ensure(65536));
    launch_hist(ctx->codes.p, n, ctx->freq.p, ctx->first.p, ctx->chunk_zeros.p, 32768, st);
    launch_zero_scan(ctx->chunk_zeros.p, ctx->chunk_base.p, nchunks, ctx->tiles.p, ctx->scalars.p, st);
    HIPCHECK(hipGetLastError());
    uint64_t n_unpred = 0;
    uint32_t errflag = 0;
    HIPCHECK(hipMemcpyAsync(&n_unpred, ctx->scalars.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&errflag, ctx->err_flag.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (errflag) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "a quantization code exceeded 16 bits");
    HIPCHECK(ctx->unpred.ensure(n_unpred));
    launch_unpred_gather(ctx->codes.p, n, ctx->outval.p, ctx->chunk_zeros.p, ctx->chunk_base.p, ctx->unpred.p, st);
    HIPCHECK(hipMemcpyAsync(ctx->h_freq.data(), ctx->freq.p, 65536 * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(ctx->h_first.data(), ctx->first.p, 65536 * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    ctx->hist_codes = ctx->codes.p;
    ctx->hist_n = n;
    HuffTable tab;
    if (build_huffman(ctx->h_freq.data(), ctx->h_first.data(), 65536, tab))
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "Huffman tree could not be built");
    faz_huff_table ht{tab.offset, tab.n_states, tab.code.data(), tab.len.data()};
    uint64_t maxbits = 0;
    for (auto l: tab.len) maxbits = std::max<uint64_t>(maxbits, l);
    HIPCHECK(ctx->words.ensure((size_t) ((n * maxbits + 31) / 32 + 2)));
    size_t nbytes = 0;
    int rc = faz_gpu_huffman_encode(ctx, ctx->codes.p, n, &ht, reinterpret_cast<uint8_t *>(ctx->words.p),
                                    ctx->words.cap * 4, &nbytes);
    if (rc) return rc;
    // compress's buffer (SBI:150-165)
    raw.clear();
    for (int d = 0; d < N; d++) sput(raw, (uint64_t) dims[d]);
    sput(raw, (uint32_t) bs);
    sput(raw, (int32_t) algo);
    sput(raw, (int32_t) dir);
    raw.push_back(2);
    sput(raw, eb);
    sput(raw, (int32_t) 32768);
    sput(raw, (uint64_t) n_unpred);
    size_t at = raw.size();
    raw.resize(at + n_unpred * 4 + tab.tree.size() + 8 + nbytes);
    if (n_unpred) HIPCHECK(hipMemcpyAsync(raw.data() + at, ctx->unpred.p, n_unpred * 4, hipMemcpyDeviceToHost, st));
    at += n_unpred * 4;
    memcpy(raw.data() + at, tab.tree.data(), tab.tree.size());
    at += tab.tree.size();
    const uint64_t nb64 = nbytes;
    memcpy(raw.data() + at, &nb64, 8);
    at += 8;
    HIPCHECK(hipMemcpyAsync(raw.data() + at, ctx->words.p, nbytes, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    std::vector<uint8_t> z(ZSTD_compressBound(raw.size()));
    const size_t zs = ZSTD_compress(z.data(), z.size(), raw.data(), raw.size(), 3);
    if (ZSTD_isError(zs)) CTX_FAIL(FAZ_ERR_INVALID, "zstd failed");
    bytes = 8 + zs;
    return FAZ_OK;
}

// optimize_quant_invl_3d (QO:89-147): sample positions from the
// reference's pointer walk, error histogram on the device, interval count.
int quant_intervals(faz_gpu_ctx *ctx, const float *d_in, const size_t *dims, double precision, int &out,
                    size_t *samples) {
    const size_t r1 = dims[0], r2 = dims[1], r3 = dims[2], dist = 100, r23 = r2 * r3, len = r1 * r23;
    std::vector<uint64_t> pos;
    pos.reserve(len / dist + 16);
    size_t offset_count = dist - 2, offset_count_2 = 0, n1 = 1, n2 = 1;
    size_t p = r23 + r3 + offset_count;
    while (p < len) {
        pos.push_back(p);
        offset_count += dist;
        if (offset_count >= r3) {
            n2++;
            if (n2 == r2) {
                n1++;
                n2 = 1;
                p += r3;
            }
            offset_count_2 = (n1 + n2) % dist;
            p += (r3 + dist - offset_count) + (dist - offset_count_2);
            offset_count = dist - offset_count_2;
            if (offset_count == 0) offset_count++;
        } else {
            p += dist;
        }
    }
    const uint64_t n = pos.size();
    if (samples) *samples = n;
    hipStream_t st = ctx->stream;
    HIPCHECK(ctx->chunk_base.ensure(std::max<uint64_t>(n, 1)));
    HIPCHECK(ctx->freq.ensure(kHistSlots));
    launch_fill_u32(ctx->freq.p, 32768, 0u, st);
    if (n) {
        HIPCHECK(hipMemcpyAsync(ctx->chunk_base.p, pos.data(), n * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(sz3_qintv_kernel, dim3((unsigned) std::min<uint64_t>((n + 255) / 256, 4096)), dim3(256), 0,
                           st, d_in, ctx->chunk_base.p, n, (uint64_t) r23, (uint64_t) r3, precision, ctx->freq.p);
    }
    HIPCHECK(hipGetLastError());
    std::vector<uint32_t> hist(32768);
    HIPCHECK(hipMemcpyAsync(hist.data(), ctx->freq.p, 32768 * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    // estimate_quantization_intervals (QO:25-39)
    const size_t target = (size_t) (n * 0.999);
    size_t sum = 0, i = 0;
    for (i = 0; i < hist.size(); i++) {
        sum += hist[i];
        if (sum > target) break;
    }
    if (i == hist.size()) i = hist.size() - 1;
    unsigned acc = 2 * (unsigned) (i + 1), b = acc - 1;
    b |= b >> 1, b |= b >> 2, b |= b >> 4, b |= b >> 8, b |= b >> 16;
    const unsigned num = 2 * (b + 1);
    out = num > 32 ? (int) num : 32;
    return FAZ_OK;
}

// SZ_compress_LorenzoReg's zstd-framed size (LR:95-114) on a device array.
int lorenzo_size(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims, double eb, int bs, int pred_dim,
                 int radius, size_t &bytes) {
    std::vector<uint8_t> raw;
    int rc;
    if (N == 3) {
        faz_lorenzo_params p;
        memset(&p, 0, sizeof(p));
        for (int d = 0; d < 3; d++) p.dims[d] = dims[d];
        p.eb = eb;
        p.block_size = bs;
        p.pred_dim = pred_dim;
        p.lorenzo = p.lorenzo2 = 1;
        p.regression = 0;
        p.radius = radius;
        rc = faz_lorenzo_raw_vec(ctx, d_in, nullptr, &p, raw);
    } else {
        faz_general_params p;
        memset(&p, 0, sizeof(p));
        p.N = N;
        for (int d = 0; d < N; d++) p.dims[d] = dims[d];
        p.eb = eb;
        p.block_size = bs;
        p.lorenzo = p.lorenzo2 = 1;
        p.radius = radius;
        rc = faz_general_raw_vec(ctx, d_in, nullptr, &p, raw);
    }
    if (rc) return rc;
    std::vector<uint8_t> z(ZSTD_compressBound(raw.size()));
    const size_t zs = ZSTD_compress(z.data(), z.size(), raw.data(), raw.size(), 3);
    if (ZSTD_isError(zs)) CTX_FAIL(FAZ_ERR_INVALID, "zstd failed");
    bytes = 8 + zs;
    return FAZ_OK;
}

int default_block_size(int N) { return N == 1 ? 128 : (N == 2 ? 16 : 6); }

// Tuning with predictorTuningRate = 0 (SZI:2096-2180) and the Lorenzo
// branch's tests (SZI:2783-2850).
int sz3_tune(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims, double eb, double rel, faz_sz3_out *o) {
    memset(o, 0, sizeof(*o));
    if (N < 1 || N > 3) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "the default tuning mode runs on 1- to 3-D fields");
    HIPCHECK(hipSetDevice(ctx->device));
    size_t sdims[4] = {1, 1, 1, 1}, sb = 0, num = 1, snum = 1;
    for (int d = 0; d < N; d++) num *= dims[d];
    bool sampled = false;
    {
        size_t dm;
        sampled = sz3_geometry(N, dims, sdims, sb, dm);
    }
    for (int d = 0; d < N; d++) snum *= sdims[d], o->sample_dims[d] = sdims[d];
    o->sampling_block = sb;
    o->sampled = sampled;
    o->interp_algo = 1;  // Config defaults (CFG:417-419)
    o->direction = 0;
    o->use_interp = 1;
    if (!sampled) return FAZ_OK;
    HIPCHECK(ctx->gf_fields.ensure(snum));
    float *d_s = ctx->gf_fields.p;
    int rc = sz3_sample(ctx, d_in, N, dims, d_s, sdims, sb, sampled);
    if (rc) return rc;
    size_t bytes = 0;
    if ((rc = lorenzo_size(ctx, d_s, N, sdims, eb, 5, N, 32768, bytes))) return rc;
    o->lorenzo_bytes = bytes;
    double lorenzo_ratio = snum * 1.0 * sizeof(float) / bytes;
    double best = 0;
    std::vector<uint8_t> raw;
    for (int op = 0; op <= 1; op++) {
        if ((rc = block_interp_test(ctx, d_s, N, sdims, eb, op, o->direction, sb, raw, bytes))) return rc;
        o->interp_bytes[op] = bytes;
        const double r = snum * 1.0 * sizeof(float) / bytes;
        if (r > best) best = r, o->interp_algo = op;
    }
    const int fact = N == 1 ? 1 : (N == 2 ? 2 : 6);
    if ((rc = block_interp_test(ctx, d_s, N, sdims, eb, o->interp_algo, fact - 1, sb, raw, bytes))) return rc;
    o->interp_bytes[2] = bytes;
    {
        const double r = snum * 1.0 * sizeof(float) / bytes;
        if (r > best * 1.02) best = r, o->direction = fact - 1;
    }
    o->lorenzo_ratio = lorenzo_ratio;
    o->interp_ratio = best;
    o->use_interp = !(lorenzo_ratio > best && lorenzo_ratio < 80 && best < 80);
    if (o->use_interp) return FAZ_OK;
    int quantbin = 65536, pred_dim = N;
    const int bsd = default_block_size(N);
    if (N == 3) {
        if ((rc = quant_intervals(ctx, d_in, dims, eb, quantbin, nullptr))) return rc;
        o->quantbin_estimate = quantbin;
        if (quantbin / 2 > 32768)
            CTX_FAIL(FAZ_ERR_UNSUPPORTED, "optimize_quant_invl_3d chose %d bins (above 65536, 16-bit codes)",
                     quantbin);
        pred_dim = 2;
        if ((rc = lorenzo_size(ctx, d_s, N, sdims, eb, bsd, 2, quantbin / 2, bytes))) return rc;
        o->lorenzo_bytes_pd2 = bytes;
        const double r = snum * 1.0 * sizeof(float) / bytes;
        if (r > lorenzo_ratio * 1.02) lorenzo_ratio = r;
        else pred_dim = 3;
    }
    if (rel < 1.01e-6 && lorenzo_ratio > 5 && quantbin != 16384) {
        if ((rc = lorenzo_size(ctx, d_s, N, sdims, eb, bsd, pred_dim, 8192, bytes))) return rc;
        o->lorenzo_bytes_16384 = bytes;
        const double r = snum * 1.0 * sizeof(float) / bytes;
        if (r > lorenzo_ratio * 1.02) lorenzo_ratio = r, quantbin = 16384;
    }
    o->quantbin = quantbin;
    o->final_lorenzo_ratio = lorenzo_ratio;
    return FAZ_OK;
}

}  // namespace
}  // namespace faz

extern "C" {

faz_status faz_gpu_sampling_f32(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims, float *d_out,
                                size_t cap, size_t *sample_dims, size_t *sampling_block, int *sampled) {
    if (!ctx || !d_in || !dims || !sample_dims || !sampling_block || !sampled || N < 1 || N > 3)
        return FAZ_ERR_INVALID;
    HIPCHECK(hipSetDevice(ctx->device));
    size_t dm, sd[4] = {1, 1, 1, 1}, sb = 0;
    bool s = faz::sz3_geometry(N, dims, sd, sb, dm);
    size_t sn = 1;
    for (int d = 0; d < N; d++) sn *= sd[d], sample_dims[d] = sd[d];
    *sampling_block = sb;
    *sampled = s ? 1 : 0;
    if (!s) return FAZ_OK;
    if (!d_out || cap < sn) CTX_FAIL(FAZ_ERR_CAPACITY, "sample needs %zu floats", sn);
    bool again;
    int rc = faz::sz3_sample(ctx, d_in, N, dims, d_out, sd, sb, again);
    if (rc) return (faz_status) rc;
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return FAZ_OK;
}

faz_status faz_gpu_block_interp_test_f32(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims, double eb,
                                         int algo, int direction, size_t block, uint8_t *h_raw, size_t cap,
                                         size_t *raw_size, size_t *bytes) {
    if (!ctx || !d_in || !dims || !bytes) return FAZ_ERR_INVALID;
    HIPCHECK(hipSetDevice(ctx->device));
    std::vector<uint8_t> raw;
    int rc = faz::block_interp_test(ctx, d_in, N, dims, eb, algo, direction, block, raw, *bytes);
    if (rc) return (faz_status) rc;
    if (raw_size) *raw_size = raw.size();
    if (h_raw) {
        if (cap < raw.size()) CTX_FAIL(FAZ_ERR_CAPACITY, "raw buffer needs %zu bytes", raw.size());
        memcpy(h_raw, raw.data(), raw.size());
    }
    return FAZ_OK;
}

faz_status faz_gpu_quant_intervals_f32(faz_gpu_ctx *ctx, const float *d_in, const size_t *dims, double eb,
                                       int *quantbin, size_t *samples) {
    if (!ctx || !d_in || !dims || !quantbin || !(eb > 0)) return FAZ_ERR_INVALID;
    if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2) CTX_FAIL(FAZ_ERR_INVALID, "optimize_quant_invl_3d needs every dimension >= 2");
    HIPCHECK(hipSetDevice(ctx->device));
    return (faz_status) faz::quant_intervals(ctx, d_in, dims, eb, *quantbin, samples);
}

faz_status faz_gpu_sz3_tune_f32(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims, double abs_eb,
                                double rel_eb, faz_sz3_out *out) {
    if (!ctx || !d_in || !dims || !out || !(abs_eb > 0)) return FAZ_ERR_INVALID;
    return (faz_status) faz::sz3_tune(ctx, d_in, N, dims, abs_eb, rel_eb, out);
}

}  // extern "C"
