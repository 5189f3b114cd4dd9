// General-frontend engine on gfx950: SZGeneralFrontend (include/QoZ/
// frontend/SZGeneralFrontend.hpp:36-116) over a list of up to three
// predictors -- 1-layer Lorenzo, 2-layer Lorenzo (LorenzoPredictor.hpp:
// 25-247), RegressionPredictor (RegressionPredictor.hpp:17-246) -- composed by
// ComposedPredictor (ComposedPredictor.hpp:45-93, 203-249) or one alone, as
// make_lorenzo_regression_compressor builds them (SZLorenzoReg.hpp:27-91), for
// N = 1..4.
//
// The reference walks blocks of conf.blockSize in raster order and overwrites
// the field with its reconstruction point by point: a Lorenzo prediction reads
// reconstructed values behind the point and originals elsewhere, and the
// per-block predictor choice sums |x - pred| + noise along the block
// diagonals on that mixed array.  Every read lies at offsets in {0,1,2}^N
// below the point, so a block depends only on blocks whose indices are not
// larger on any axis: all blocks of one anti-diagonal sum(b) = T are
// independent.  Here:
//   * gf_block_kernel, one launch per anti-diagonal, one wave per block: the
//     block (originals) and its two low halo layers (reconstruction, 0
//     outside the field) are staged in LDS, the diagonal estimates are
//     computed in parallel and summed in the reference's order by one lane,
//     then the in-block hyperplanes sum(l) = s are swept (their points are
//     independent of each other);
//   * N = 1 is one dependency chain: gf_line_kernel walks it with one lane.
// The RegressionPredictor fits a plane to the block's originals (its error
// estimate uses that plane) and predicts with the plane's coefficients after
// they are quantized against the last block that chose it, in raster order.
// The fit needs no other block (gf_reg_fit_kernel fits all blocks at once);
// the quantized coefficients depend on which raster-earlier blocks chose the
// regression, a chain the anti-diagonal order does not follow.  The host
// therefore guesses the choices, runs the coefficient chain for them, runs
// the wavefront, and repeats with the choices it produced until they no
// longer change: a block's outcome depends only on the choices of blocks
// before it in raster order, so every pass fixes at least one more block of
// the raster prefix, and a pass whose choices equal its input is the
// reference's sequential result.
// Several arrays of one shape ("fields": the tuner's sample blocks) run in the
// same launches, one per grid z index.  Codes are 32-bit (the tuner's
// CompressTest uses radius 65536) in the reference's block-major order.
#include "faz_gpu.h"
#include "faz_ctx.h"
#include "faz_device.h"
#include "faz_internal.h"

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

constexpr int kGFMaxBlock[5] = {0, 1 << 20, 64, 16, 8};  // largest blockSize staged per N

struct GFGeom {
    int N;
    uint32_t dims[4];
    uint64_t st[4];      // element strides
    uint32_t bs;
    uint32_t nb[4];      // blocks per axis
    uint64_t num;        // points per field
    uint64_t nblocks;    // blocks per field
    uint32_t E;          // staged edge bs + 2
    uint32_t nfields;
};

// The predictor list in the order of LR:36-91.
enum { GF_L1 = 1, GF_L2 = 2, GF_REG = 3 };

struct GFPred {
    int npred;           // 1..3
    int kind[3];         // GF_L1, GF_L2 or GF_REG per list entry
    int reg;             // the list ends with the RegressionPredictor
    float noise[3];      // LorenzoPredictor noise by layer count (LP:25-47)
};

__host__ __device__ __forceinline__ int gf_kind(const GFPred &P, int sid) {
    return sid == 0 ? P.kind[0] : (sid == 1 ? P.kind[1] : P.kind[2]);
}

// RegressionPredictor state, 5 floats per block (N + 1 used): raw = the
// plane fitted to the block's originals, deq = its coefficients after the
// coefficient quantizers.  The 1-D kernel runs the coefficient chain itself
// and writes each block's raw coefficients and their codes.
struct GFReg {
    const float *raw;
    const float *deq;
    float *out_raw;
    uint32_t *out_code;
    double eb_lin, recip_lin, eb_ind, recip_ind;  // RP:22-25
};

// Selection value of a block whose chosen predictor declined it (a
// regression block of extent 1): the reference records no selection there.
constexpr uint8_t kGFNoSelection = 255;

struct GFQuant {
    double eb, recip;
    float ebf;
    int radius;
};

__host__ __device__ __forceinline__ uint32_t gf_len(uint32_t b, uint32_t bs, uint32_t d) {
    return (b + 1) * bs < d ? bs : d - b * bs;
}

// First code of block b in the block-major stream of one field: for each
// axis d, the earlier blocks along d inside the slab fixed by the block's
// coordinates on axes < d (full extent on axes > d).
template<int N>
__device__ __forceinline__ uint64_t gf_code_off(const GFGeom &G, const uint32_t *b, const uint32_t *sz) {
    uint64_t off = 0, pre = 1;
    for (int d = 0; d < N; d++) {
        uint64_t post = 1;
        for (int e = d + 1; e < N; e++) post *= G.dims[e];
        off += pre * ((uint64_t) b[d] * G.bs) * post;
        pre *= sz[d];
    }
    return off;
}

// LorenzoPredictor::do_predict (LP:117-247) on f(a_0, .., a_{N-1}) = the
// value a_d points below along each axis d; float, left to right.  N = 4 uses
// the 1-layer formula for both layer counts (LP:170-177).
template<int N, class F>
__device__ __forceinline__ float lorenzo1(F f) {
    if constexpr (N == 1) return f(1, 0, 0, 0);
    else if constexpr (N == 2) return f(0, 1, 0, 0) + f(1, 0, 0, 0) - f(1, 1, 0, 0);
    else if constexpr (N == 3)
        return f(0, 0, 1, 0) + f(0, 1, 0, 0) + f(1, 0, 0, 0) - f(0, 1, 1, 0) - f(1, 0, 1, 0) - f(1, 1, 0, 0) +
               f(1, 1, 1, 0);
    else
        return f(0, 0, 0, 1) + f(0, 0, 1, 0) - f(0, 0, 1, 1) + f(0, 1, 0, 0) - f(0, 1, 0, 1) - f(0, 1, 1, 0) +
               f(0, 1, 1, 1) + f(1, 0, 0, 0) - f(1, 0, 0, 1) - f(1, 0, 1, 0) + f(1, 0, 1, 1) - f(1, 1, 0, 0) +
               f(1, 1, 0, 1) + f(1, 1, 1, 0) - f(1, 1, 1, 1);
}

template<int N, class F>
__device__ __forceinline__ float lorenzo2(F f) {
    if constexpr (N == 1) return 2 * f(1, 0, 0, 0) - f(2, 0, 0, 0);
    else if constexpr (N == 2)
        return 2 * f(0, 1, 0, 0) - f(0, 2, 0, 0) + 2 * f(1, 0, 0, 0) - 4 * f(1, 1, 0, 0) + 2 * f(1, 2, 0, 0) -
               f(2, 0, 0, 0) + 2 * f(2, 1, 0, 0) - f(2, 2, 0, 0);
    else if constexpr (N == 3)
        return 2 * f(0, 0, 1, 0) - f(0, 0, 2, 0) + 2 * f(0, 1, 0, 0) - 4 * f(0, 1, 1, 0) + 2 * f(0, 1, 2, 0) -
               f(0, 2, 0, 0) + 2 * f(0, 2, 1, 0) - f(0, 2, 2, 0) + 2 * f(1, 0, 0, 0) - 4 * f(1, 0, 1, 0) +
               2 * f(1, 0, 2, 0) - 4 * f(1, 1, 0, 0) + 8 * f(1, 1, 1, 0) - 4 * f(1, 1, 2, 0) + 2 * f(1, 2, 0, 0) -
               4 * f(1, 2, 1, 0) + 2 * f(1, 2, 2, 0) - f(2, 0, 0, 0) + 2 * f(2, 0, 1, 0) - f(2, 0, 2, 0) +
               2 * f(2, 1, 0, 0) - 4 * f(2, 1, 1, 0) + 2 * f(2, 1, 2, 0) - f(2, 2, 0, 0) + 2 * f(2, 2, 1, 0) -
               f(2, 2, 2, 0);
    else
        return lorenzo1<4>(f);
}

template<int N, class F>
__device__ __forceinline__ float lorenzo(int layers, F f) {
    return layers == 1 ? lorenzo1<N>(f) : lorenzo2<N>(f);
}

// RegressionPredictor::predict (RP:91-98) at block-local index l: float,
// from 0, the slopes in axis order, then the constant term.
template<int N>
__host__ __device__ __forceinline__ float reg_predict(const uint32_t *l, const float *c) {
    float pred = 0;
    for (int i = 0; i < N; i++) pred += (float) l[i] * c[i];
    pred += c[N];
    return pred;
}

// RegressionPredictor::precompress_block (RP:48-85) for every block of every
// field in one launch, one lane per block: a block's plane reads only its
// own originals.  Sums in double in the reference's order (rows along the
// last axis, the index-weighted sum of that axis running over all rows); a
// block of extent 1 on an axis declines and is left unwritten.
template<int N>
__global__ void __launch_bounds__(256) gf_reg_fit_kernel(GFGeom G, const float *__restrict__ in,
                                                         float *__restrict__ coef) {
    const uint64_t t = (uint64_t) blockIdx.x * 256 + threadIdx.x;
    if (t >= G.nblocks * G.nfields) return;
    const uint64_t f = t / G.nblocks;
    uint64_t r = t % G.nblocks;
    uint32_t b[4] = {0, 0, 0, 0}, sz[4] = {1, 1, 1, 1};
    for (int d = N - 1; d >= 0; d--) {
        b[d] = (uint32_t) (r % G.nb[d]);
        r /= G.nb[d];
    }
    uint64_t n = 1, at0 = f * G.num;
    for (int d = 0; d < N; d++) {
        sz[d] = gf_len(b[d], G.bs, G.dims[d]);
        if (sz[d] <= 1) return;
        n *= sz[d];
        at0 += (uint64_t) b[d] * G.bs * G.st[d];
    }
    const float recip = (float) (1.0 / (double) n);
    double sum[N + 1];
    for (int i = 0; i <= N; i++) sum[i] = 0;
    uint32_t l[4] = {0, 0, 0, 0};
    const uint64_t rows = n / sz[N - 1];
    for (uint64_t row = 0; row < rows; row++) {
        uint64_t at = at0;
        for (int d = 0; d < N - 1; d++) at += (uint64_t) l[d] * G.st[d];
        double cum = 0;
        for (uint32_t k = 0; k < sz[N - 1]; k++) {
            const float v = in[at + k];
            cum += v;
            sum[N - 1] += (double) k * v;
        }
        for (int i = 0; i < N - 1; i++) sum[i] += cum * (double) l[i];
        sum[N] += cum;
        for (int d = N - 2; d >= 0; d--) {
            if (++l[d] < sz[d]) break;
            l[d] = 0;
        }
    }
    float c[N + 1];
    c[N] = (float) (sum[N] * recip);
    for (int i = 0; i < N; i++) {
        c[i] = (float) ((2 * sum[i] / (double) (sz[i] - 1) - sum[N]) * 6 * recip / (double) (sz[i] + 1));
        c[N] -= (float) (sz[i] - 1) * c[i] / 2;
    }
    for (int i = 0; i <= N; i++) coef[5 * t + i] = c[i];
}

// One wave per block of anti-diagonal T.  Block coordinates: b0 = lo0 +
// blockIdx.x; b1 = lo1 + blockIdx.y / span2 (N >= 3); b2 = lo2 + blockIdx.y %
// span2 (N = 4); the last axis closes the sum.  Field = blockIdx.z.
// MODE 0 compresses in[] into recon/codes/outval/sel; MODE 1 rebuilds recon
// from codes, outval (the unpredictable value of each zero code) and sel.
// tab lists the points of a full bs^N block by hyperplane (8 bits per local
// coordinate), diag[s] where hyperplane s starts.
template<int MODE, int N>
__global__ void __launch_bounds__(64) gf_block_kernel(GFGeom G, GFPred P, GFQuant Q, GFReg R, uint32_t T,
                                                      uint32_t lo0, uint32_t lo1, uint32_t lo2, uint32_t span2,
                                                      const float *__restrict__ in, float *__restrict__ recon,
                                                      uint32_t *__restrict__ codes, float *__restrict__ outval,
                                                      uint8_t *__restrict__ sel, const uint32_t *__restrict__ tab,
                                                      const uint32_t *__restrict__ diag) {
    extern __shared__ float s_gf[];
    const uint32_t E = G.E;
    uint32_t EN = 1;
    for (int d = 0; d < N; d++) EN *= E;
    uint32_t bsN = 1;
    for (int d = 0; d < N; d++) bsN *= G.bs;
    float *s_rec = s_gf;                                          // E^N
    uint32_t *s_code = reinterpret_cast<uint32_t *>(s_gf + EN);   // bs^N
    float *s_e = s_gf + EN + bsN;                                 // 3 * 8 * bs estimates
    __shared__ int s_sid;

    uint32_t b[4] = {0, 0, 0, 0};
    b[0] = lo0 + blockIdx.x;
    if constexpr (N >= 3) b[1] = lo1 + blockIdx.y / span2;
    if constexpr (N == 4) b[2] = lo2 + blockIdx.y % span2;
    uint32_t used = 0;
    for (int d = 0; d < N - 1; d++) {
        if (b[d] >= G.nb[d]) return;
        used += b[d];
    }
    if (used > T || T - used >= G.nb[N - 1]) return;
    b[N - 1] = T - used;
    const uint32_t f = blockIdx.z;
    uint64_t bidx = 0;
    for (int d = 0; d < N; d++) bidx = bidx * G.nb[d] + b[d];
    bidx += (uint64_t) f * G.nblocks;
    const uint64_t base = (uint64_t) f * G.num;
    uint32_t sz[4], o[4];
    uint32_t n = 1, m = 0xffffffffu, nsteps = 1;
    for (int d = 0; d < N; d++) {
        o[d] = b[d] * G.bs;
        sz[d] = gf_len(b[d], G.bs, G.dims[d]);
        n *= sz[d];
        m = min(m, sz[d]);
        nsteps += sz[d] - 1;
    }
    uint32_t Es[4];  // LDS strides
    Es[N - 1] = 1;
    for (int d = N - 2; d >= 0; d--) Es[d] = Es[d + 1] * E;
    const uint64_t coff = base + gf_code_off<N>(G, b, sz);
    const uint32_t lane = threadIdx.x;

    // stage: originals inside the block (compress), reconstruction below it
    for (uint32_t q = lane; q < EN; q += 64) {
        uint32_t r = q, a[4];
        bool skip = false, inside = true, outside = false;
        uint64_t at = base;
        for (int d = N - 1; d >= 0; d--) {
            a[d] = r % E;
            r /= E;
        }
        for (int d = 0; d < N; d++) {
            if (a[d] >= sz[d] + 2) skip = true;
            if (a[d] < 2) inside = false;
            const int64_t x = (int64_t) o[d] + a[d] - 2;
            if (x < 0) outside = true;
            at += (uint64_t) (x < 0 ? 0 : x) * G.st[d];
        }
        if (skip) continue;
        float v = 0.0f;
        if (inside) {
            if constexpr (MODE == 0) v = in[at];
        } else if (!outside) {
            v = recon[at];
        }
        s_rec[q] = v;
    }
    if constexpr (MODE == 1) {
        __syncthreads();
        for (uint32_t p = lane; p < n; p += 64) {
            uint32_t r = p, slot = 0;
            for (int d = N - 1; d >= 0; d--) {
                slot += (r % sz[d] + 2) * Es[d];
                r /= sz[d];
            }
            const uint32_t c = codes[coff + p];
            s_code[p] = c;
            s_rec[slot] = c ? 0.0f : outval[coff + p];
        }
    }
    __syncthreads();

    // predictor choice (CP:45-63, 203-249); a regression block of extent 1
    // declines (RP:48-53) and runs on the 1-layer Lorenzo fallback (GF:47-52)
    const bool degen = m <= 1;
    int sid = 0;
    if (P.npred > 1) {
        if constexpr (MODE == 0) {
            const uint32_t ndiag = 1u << (N - 1);
            const uint32_t npts = m >= 3 ? (m - 2) * ndiag : 0;
            float rc[5] = {0, 0, 0, 0, 0};  // the block's fitted plane (its error estimate)
            if (P.reg && npts)
                for (int i = 0; i <= N; i++) rc[i] = R.raw[5 * bidx + i];
            for (uint32_t idx = lane; idx < npts; idx += 64) {
                const uint32_t k = idx / ndiag, t = idx % ndiag;
                uint32_t l[4] = {k, 0, 0, 0};
                uint32_t slot = (k + 2) * Es[0];
                for (int d = 1; d < N; d++) {
                    l[d] = ((t >> (N - 1 - d)) & 1) ? m - 1 - k : k;
                    slot += (l[d] + 2) * Es[d];
                }
                auto fv = [&](int a0, int a1, int a2, int a3) {
                    const int a[4] = {a0, a1, a2, a3};
                    int off = 0;
                    for (int d = 0; d < N; d++) off += a[d] * (int) Es[d];
                    return s_rec[(int) slot - off];
                };
                const float x = s_rec[slot];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    if (q >= P.npred) break;
                    const int kd = P.kind[q];
                    float e;
                    if (kd == GF_REG) e = fabsf(x - reg_predict<N>(l, rc));
                    else if (kd == GF_L1) e = fabsf(x - lorenzo1<N>(fv)) + P.noise[1];
                    else e = fabsf(x - lorenzo2<N>(fv)) + P.noise[2];
                    s_e[q * 8 * G.bs + idx] = e;
                }
            }
            __syncthreads();
            if (lane == 0) {
                double pe[3] = {0, 0, 0};
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    if (q >= P.npred) break;
                    for (uint32_t idx = 0; idx < npts; idx++) pe[q] += (double) s_e[q * 8 * G.bs + idx];
                }
                int best = 0;  // std::min_element: the first smallest
                if (P.npred > 1 && pe[1] < pe[best]) best = 1;
                if (P.npred > 2 && pe[2] < pe[best]) best = 2;
                s_sid = best;
                sel[bidx] = (gf_kind(P, best) == GF_REG && degen) ? kGFNoSelection : (uint8_t) best;
            }
            __syncthreads();
            sid = s_sid;
        } else {
            sid = sel[bidx];
        }
    }
    int kind = gf_kind(P, sid);
    if (kind == GF_REG && degen) kind = GF_L1;

    if (kind == GF_REG) {
        // the plane after the coefficient chain: no point reads another
        float cq[5];
        for (int i = 0; i <= N; i++) cq[i] = R.deq[5 * bidx + i];
        for (uint32_t p = lane; p < n; p += 64) {
            uint32_t r = p, slot = 0, l[4] = {0, 0, 0, 0};
            for (int d = N - 1; d >= 0; d--) {
                l[d] = r % sz[d];
                r /= sz[d];
                slot += (l[d] + 2) * Es[d];
            }
            const float pred = reg_predict<N>(l, cq);
            float v;
            if constexpr (MODE == 0) {
                const int code = quantize_select(s_rec[slot], pred, Q.eb, Q.recip, Q.ebf, Q.radius, &v);
                s_code[p] = (uint32_t) code;
                if (code == 0) outval[coff + p] = s_rec[slot];
            } else {
                const uint32_t code = s_code[p];
                v = code ? recover_point(pred, (int) code, Q.radius, Q.eb) : s_rec[slot];
            }
            s_rec[slot] = v;
        }
        __syncthreads();
    } else {
        // in-block hyperplanes
        const int layers = kind;
        for (uint32_t s = 0; s < nsteps; s++) {
            const uint32_t q1 = diag[s + 1];
            for (uint32_t q = diag[s] + lane; q < q1; q += 64) {
                const uint32_t e = tab[q];
                uint32_t l[4], slot = 0, p = 0;
                bool skip = false;
                for (int d = 0; d < N; d++) {
                    l[d] = (e >> (8 * d)) & 255u;
                    if (l[d] >= sz[d]) skip = true;
                    slot += (l[d] + 2) * Es[d];
                    p = p * sz[d] + l[d];
                }
                if (skip) continue;
                auto fv = [&](int a0, int a1, int a2, int a3) {
                    const int a[4] = {a0, a1, a2, a3};
                    int off = 0;
                    for (int d = 0; d < N; d++) off += a[d] * (int) Es[d];
                    return s_rec[(int) slot - off];
                };
                const float pred = lorenzo<N>(layers, fv);
                float v;
                if constexpr (MODE == 0) {
                    const int code = quantize_select(s_rec[slot], pred, Q.eb, Q.recip, Q.ebf, Q.radius, &v);
                    s_code[p] = (uint32_t) code;
                    if (code == 0) outval[coff + p] = s_rec[slot];
                } else {
                    const uint32_t code = s_code[p];
                    v = code ? recover_point(pred, (int) code, Q.radius, Q.eb) : s_rec[slot];
                }
                s_rec[slot] = v;
            }
            __syncthreads();
        }
    }
    for (uint32_t p = lane; p < n; p += 64) {
        uint32_t r = p, slot = 0;
        uint64_t at = base;
        for (int d = N - 1; d >= 0; d--) {
            const uint32_t l = r % sz[d];
            r /= sz[d];
            slot += (l + 2) * Es[d];
            at += (uint64_t) (o[d] + l) * G.st[d];
        }
        recon[at] = s_rec[slot];
        if constexpr (MODE == 0) codes[coff + p] = s_code[p];
    }
}

// N = 1: the whole line is one dependency chain; lane 0 of each wave walks
// one field, keeping the two previous reconstructed values in registers, and
// with the RegressionPredictor its current and last committed coefficients
// (a block of one point keeps the previous block's coefficients, which its
// error estimate then reads, as the reference's does).
template<int MODE>
__global__ void __launch_bounds__(64) gf_line_kernel(GFGeom G, GFPred P, GFQuant Q, GFReg R,
                                                     const float *__restrict__ in, float *__restrict__ recon,
                                                     uint32_t *__restrict__ codes, float *__restrict__ outval,
                                                     uint8_t *__restrict__ sel) {
    if (threadIdx.x != 0) return;
    const uint32_t f = blockIdx.x;
    if (f >= G.nfields) return;
    const uint64_t base = (uint64_t) f * G.num;
    const uint32_t d0 = G.dims[0], bs = G.bs;
    float r1 = 0.0f, r2 = 0.0f;  // reconstruction at i-1, i-2 (0 below the field)
    float cur[2] = {0.0f, 0.0f}, prevc[2] = {0.0f, 0.0f};
    for (uint32_t bb = 0; bb < G.nb[0]; bb++) {
        const uint32_t o = bb * bs, m = gf_len(bb, bs, d0);
        const uint64_t bidx = (uint64_t) f * G.nblocks + bb;
        if constexpr (MODE == 0) {
            if (P.reg && m > 1) {
                // precompress_block (RP:48-85) on the block's originals
                double s0 = 0, s1 = 0;
                for (uint32_t t = 0; t < m; t++) {
                    const float v = in[base + o + t];
                    s1 += v;
                    s0 += (double) t * v;
                }
                const float recip = (float) (1.0 / (double) m);
                cur[1] = (float) (s1 * recip);
                cur[0] = (float) ((2 * s0 / (double) (m - 1) - s1) * 6 * recip / (double) (m + 1));
                cur[1] -= (float) (m - 1) * cur[0] / 2;
            }
        }
        int sid = 0;
        if (P.npred > 1) {
            if constexpr (MODE == 0) {
                // estimates at the block's first and last points (CP:192-201)
                auto val = [&](int64_t j) -> float {
                    if (j < 0) return 0.0f;
                    if (j >= (int64_t) o) return in[base + (uint64_t) j];
                    return j == (int64_t) o - 1 ? r1 : r2;
                };
                const int64_t at0 = o, at1 = (int64_t) o + m - 1;
                const float x0 = in[base + o], x1 = in[base + (uint64_t) at1];
                double pe[3] = {0, 0, 0};
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    if (q >= P.npred) break;
                    const int kd = P.kind[q];
                    float e0, e1;
                    if (kd == GF_REG) {
                        const uint32_t l0 = 0, l1 = m - 1;
                        e0 = fabsf(x0 - reg_predict<1>(&l0, cur));
                        e1 = fabsf(x1 - reg_predict<1>(&l1, cur));
                    } else if (kd == GF_L1) {
                        e0 = fabsf(x0 - val(at0 - 1)) + P.noise[1];
                        e1 = fabsf(x1 - val(at1 - 1)) + P.noise[1];
                    } else {
                        e0 = fabsf(x0 - (2 * val(at0 - 1) - val(at0 - 2))) + P.noise[2];
                        e1 = fabsf(x1 - (2 * val(at1 - 1) - val(at1 - 2))) + P.noise[2];
                    }
                    pe[q] += (double) e0;
                    pe[q] += (double) e1;
                }
                if (P.npred > 1 && pe[1] < pe[sid]) sid = 1;
                if (P.npred > 2 && pe[2] < pe[sid]) sid = 2;
            } else {
                sid = sel[bidx];
            }
        }
        int kind = gf_kind(P, sid);
        const bool declined = kind == GF_REG && m <= 1;
        if (declined) kind = GF_L1;
        if constexpr (MODE == 0)
            if (P.npred > 1) sel[bidx] = declined ? kGFNoSelection : (uint8_t) sid;
        float c[2] = {0.0f, 0.0f};
        if (kind == GF_REG) {
            if constexpr (MODE == 0) {
                // precompress_block_commit (RP:87-90, 224-230): slope, then
                // constant term, each against the last committed value
                float v;
                R.out_raw[5 * bidx] = cur[0];
                R.out_raw[5 * bidx + 1] = cur[1];
                R.out_code[5 * bidx] = (uint32_t) quantize_point(cur[0], prevc[0], R.eb_lin, R.recip_lin, 32768, &v);
                cur[0] = v;
                R.out_code[5 * bidx + 1] =
                    (uint32_t) quantize_point(cur[1], prevc[1], R.eb_ind, R.recip_ind, 32768, &v);
                cur[1] = v;
                prevc[0] = cur[0];
                prevc[1] = cur[1];
                c[0] = cur[0];
                c[1] = cur[1];
            } else {
                c[0] = R.deq[5 * bidx];
                c[1] = R.deq[5 * bidx + 1];
            }
        }
        for (uint32_t i = 0; i < m; i++) {
            const uint64_t at = base + o + i;
            float pred;
            if (kind == GF_REG) pred = reg_predict<1>(&i, c);
            else pred = kind == GF_L1 ? r1 : 2 * r1 - r2;
            float v;
            if constexpr (MODE == 0) {
                const float x = in[at];
                const int code = quantize_select(x, pred, Q.eb, Q.recip, Q.ebf, Q.radius, &v);
                codes[at] = (uint32_t) code;
                if (code == 0) outval[at] = x;
            } else {
                const uint32_t code = codes[at];
                v = code ? recover_point(pred, (int) code, Q.radius, Q.eb) : outval[at];
            }
            recon[at] = v;
            r2 = r1;
            r1 = v;
        }
    }
}

__global__ void gf_narrow_kernel(const uint32_t *__restrict__ c32, uint16_t *__restrict__ c16, uint64_t n) {
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x)
        c16[i] = (uint16_t) c32[i];
}

// Copies nfields sample blocks of (sbs+1)^N at `starts` out of the field.
__global__ void gf_gather_kernel(const float *__restrict__ field, int N, uint64_t s0, uint64_t s1, uint32_t e,
                                 const uint64_t *__restrict__ starts, uint32_t nfields, float *__restrict__ out) {
    const uint64_t per = N == 2 ? (uint64_t) e * e : (uint64_t) e * e * e;
    const uint64_t total = per * nfields;
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t) gridDim.x * blockDim.x) {
        const uint64_t f = i / per, r = i % per;
        const uint64_t *st = starts + 3 * f;
        uint64_t at;
        if (N == 2) {
            at = (st[0] + r / e) * s0 + st[1] + r % e;
        } else {
            const uint64_t k = r % e, t = r / e;
            at = (st[0] + t / e) * s0 + (st[1] + t % e) * s1 + st[2] + k;
        }
        out[i] = field[at];
    }
}

size_t gf_lds_bytes(int N, uint32_t bs) {
    size_t en = 1, bn = 1;
    for (int d = 0; d < N; d++) en *= bs + 2, bn *= bs;
    return (en + bn + 24 * (size_t) bs) * 4;
}

template<int MODE, int N>
void gf_launch_diagonals(const GFGeom &G, const GFPred &P, const GFQuant &Q, const GFReg &R, const float *in,
                         float *recon, uint32_t *codes, float *outval, uint8_t *sel, const uint32_t *tab,
                         const uint32_t *diag, hipStream_t st) {
    uint32_t Tmax = 0;
    for (int d = 0; d < N; d++) Tmax += G.nb[d] - 1;
    const size_t lds = gf_lds_bytes(N, G.bs);
    for (uint32_t T = 0; T <= Tmax; T++) {
        uint32_t lo[4], hi[4];
        for (int d = 0; d < N; d++) {
            const uint32_t rest = Tmax - (G.nb[d] - 1);  // most the other axes can add
            lo[d] = T > rest ? T - rest : 0;
            hi[d] = std::min(G.nb[d] - 1, T);
        }
        const uint32_t gx = hi[0] - lo[0] + 1;
        uint32_t gy = 1, span2 = 1;
        if (N >= 3) gy = hi[1] - lo[1] + 1;
        if (N == 4) {
            span2 = hi[2] - lo[2] + 1;
            gy *= span2;
        }
        hipLaunchKernelGGL((gf_block_kernel<MODE, N>), dim3(gx, gy, G.nfields), dim3(64), lds, st, G, P, Q, R, T,
                           lo[0], N >= 3 ? lo[1] : 0, N == 4 ? lo[2] : 0, span2, in, recon, codes, outval, sel, tab,
                           diag);
    }
}

template<int MODE>
void gf_launch(const GFGeom &G, const GFPred &P, const GFQuant &Q, const GFReg &R, const float *in, float *recon,
               uint32_t *codes, float *outval, uint8_t *sel, const uint32_t *tab, const uint32_t *diag,
               hipStream_t st) {
    switch (G.N) {
        case 1:
            hipLaunchKernelGGL(gf_line_kernel<MODE>, dim3(G.nfields), dim3(64), 0, st, G, P, Q, R, in, recon, codes,
                               outval, sel);
            break;
        case 2: gf_launch_diagonals<MODE, 2>(G, P, Q, R, in, recon, codes, outval, sel, tab, diag, st); break;
        case 3: gf_launch_diagonals<MODE, 3>(G, P, Q, R, in, recon, codes, outval, sel, tab, diag, st); break;
        default: gf_launch_diagonals<MODE, 4>(G, P, Q, R, in, recon, codes, outval, sel, tab, diag, st); break;
    }
}

void gf_launch_fit(const GFGeom &G, const float *in, float *coef, hipStream_t st) {
    const uint32_t grid = (uint32_t) ((G.nblocks * G.nfields + 255) / 256);
    switch (G.N) {
        case 2: hipLaunchKernelGGL(gf_reg_fit_kernel<2>, dim3(grid), dim3(256), 0, st, G, in, coef); break;
        case 3: hipLaunchKernelGGL(gf_reg_fit_kernel<3>, dim3(grid), dim3(256), 0, st, G, in, coef); break;
        case 4: hipLaunchKernelGGL(gf_reg_fit_kernel<4>, dim3(grid), dim3(256), 0, st, G, in, coef); break;
        default: break;
    }
}

}  // namespace
}  // namespace faz

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
namespace faz {
namespace {

double gf_elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

template<class V>
void gput(std::vector<uint8_t> &out, const V &v) {
    const uint8_t *b = reinterpret_cast<const uint8_t *>(&v);
    out.insert(out.end(), b, b + sizeof(V));
}

int gf_check(faz_gpu_ctx *ctx, const faz_general_params *p) {
    if (!p) return FAZ_ERR_INVALID;
    if (p->N < 1 || p->N > 4) CTX_FAIL(FAZ_ERR_INVALID, "N must be 1..4");
    uint64_t n = 1;
    for (int d = 0; d < p->N; d++) {
        if (p->dims[d] == 0) CTX_FAIL(FAZ_ERR_INVALID, "empty dimension");
        n *= p->dims[d];
    }
    if (n >= (1ull << 32)) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 2^32-1 points");
    if (!p->lorenzo && !p->lorenzo2 && !p->regression)
        CTX_FAIL(FAZ_ERR_INVALID, "all Lorenzo and regression predictors are disabled");
    if (p->block_size < 1) CTX_FAIL(FAZ_ERR_INVALID, "blockSize must be positive");
    if (p->lorenzo2 && p->block_size < 2)
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "the 2-layer predictor needs blockSize >= 2 (its reads would leave the array)");
    if (p->block_size > kGFMaxBlock[p->N])
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "blockSize %d above the %d the %d-D kernel stages", p->block_size,
                 kGFMaxBlock[p->N], p->N);
    if (p->radius < 1 || p->radius > 65536) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "radius outside 1..65536");
    if (!(p->eb > 0)) CTX_FAIL(FAZ_ERR_INVALID, "error bound must be positive");
    uint64_t nbk[4] = {1, 1, 1, 1};
    for (int d = 0; d < p->N; d++) nbk[d] = (p->dims[d] - 1) / p->block_size + 1;
    for (int d = 0; d < p->N - 1; d++)
        if (nbk[d] > 65535) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 65535 blocks along axis %d", d);
    if (p->N == 4 && nbk[1] * nbk[2] > 65535)
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 65535 blocks on axes 1 and 2 of a 4-D field");
    return FAZ_OK;
}

GFGeom gf_geom(const faz_general_params *p, uint32_t nfields) {
    GFGeom G;
    memset(&G, 0, sizeof(G));
    G.N = p->N;
    G.bs = (uint32_t) p->block_size;
    uint64_t s = 1;
    G.nblocks = 1;
    for (int d = p->N - 1; d >= 0; d--) {
        G.dims[d] = (uint32_t) p->dims[d];
        G.st[d] = s;
        s *= G.dims[d];
        G.nb[d] = (G.dims[d] - 1) / G.bs + 1;
        G.nblocks *= G.nb[d];
    }
    G.num = s;
    G.E = G.bs + 2;
    G.nfields = nfields;
    return G;
}

GFPred gf_pred(const faz_general_params *p) {
    static const double n1[4] = {0.5, 0.81, 1.22, 1.79};
    static const double n2[4] = {1.08, 2.76, 6.8, 0.0};
    GFPred P;
    P.npred = 0;
    if (p->lorenzo) P.kind[P.npred++] = GF_L1;
    if (p->lorenzo2) P.kind[P.npred++] = GF_L2;
    if (p->regression) P.kind[P.npred++] = GF_REG;
    for (int i = P.npred; i < 3; i++) P.kind[i] = GF_L1;
    P.reg = p->regression != 0;
    P.noise[0] = 0.0f;
    P.noise[1] = (float) (n1[p->N - 1] * p->eb);
    P.noise[2] = (float) (n2[p->N - 1] * p->eb);
    return P;
}

GFQuant gf_quant(double eb, int radius) {
    GFQuant Q;
    Q.eb = eb;
    Q.recip = 1.0 / eb;
    Q.ebf = float_floor(eb);
    Q.radius = radius;
    return Q;
}

// The RegressionPredictor's coefficient quantizers (RP:22-25): the constant
// term at eb / (N+1), the slopes at eb / (N+1) / blockSize, radius 32768.
GFReg gf_reg_args(int N, double eb, int bs) {
    GFReg R;
    memset(&R, 0, sizeof(R));
    R.eb_ind = eb / (N + 1);
    R.eb_lin = eb / (N + 1) / bs;
    R.recip_ind = 1.0 / R.eb_ind;
    R.recip_lin = 1.0 / R.eb_lin;
    return R;
}

constexpr int kGFCoefRadius = 32768;
// Wavefront passes after which a compression with the RegressionPredictor
// gives up (each pass settles at least one more block of the raster prefix;
// fields seen in practice settle in a few).
constexpr int kGFMaxPasses = 64;

// The RegressionPredictor's part of one compressed field (RP:100-117).
struct GFRegOut {
    std::vector<int32_t> codes;              // per committed block: slopes, then the constant term
    std::vector<float> unpred_ind, unpred_lin;
    size_t nblocks = 0;                      // blocks predicted by the regression
    int passes = 1;                          // wavefront passes until the choices settled
};

// Block b (raster index within a field) has extent 1 on some axis.
bool gf_block_degen(const GFGeom &G, uint64_t b) {
    for (int d = G.N - 1; d >= 0; d--) {
        const uint32_t i = (uint32_t) (b % G.nb[d]);
        b /= G.nb[d];
        if (gf_len(i, G.bs, G.dims[d]) <= 1) return true;
    }
    return false;
}

// The coefficient chain (precompress_block_commit, RP:87-90, 224-230) in
// raster order for the choices S: each block's plane quantized against the
// last committed coefficients gives deq (what the block predicts with if it
// chose the regression); blocks with S == reg_sid commit.  out (may be NULL)
// collects the committed codes and unpredictable coefficients.
void gf_chain(const GFGeom &G, const GFReg &R, const std::vector<float> &raw, const uint8_t *S, int reg_sid,
              const std::vector<uint8_t> &degen, std::vector<float> &deq, GFRegOut *out) {
    const int N = G.N;
    float prev[5] = {0, 0, 0, 0, 0};
    for (uint64_t b = 0; b < G.nblocks; b++) {
        if (degen[b]) continue;
        const bool commit = S[b] == reg_sid;
        float c[5];
        for (int i = 0; i <= N; i++) {
            const bool lin = i < N;
            float v;
            const int code = quantize_point(raw[5 * b + i], prev[i], lin ? R.eb_lin : R.eb_ind,
                                            lin ? R.recip_lin : R.recip_ind, kGFCoefRadius, &v);
            c[i] = v;
            if (commit && out) {
                out->codes.push_back(code);
                if (!code) (lin ? out->unpred_lin : out->unpred_ind).push_back(raw[5 * b + i]);
            }
        }
        for (int i = 0; i <= N; i++) deq[5 * b + i] = c[i];
        if (commit) {
            for (int i = 0; i <= N; i++) prev[i] = c[i];
            if (out) out->nblocks++;
        }
    }
}

// Points of a full bs^N block ordered by hyperplane sum(l), 8 bits per
// coordinate, and where each hyperplane starts.
int gf_ensure_order(faz_gpu_ctx *ctx, int N, uint32_t bs) {
    if (N == 1) return FAZ_OK;
    const uint32_t key = (uint32_t) N << 16 | bs;
    if (ctx->gf_tab_key == key) return FAZ_OK;
    const uint32_t nplanes = (uint32_t) N * (bs - 1) + 1;
    std::vector<std::vector<uint32_t>> planes(nplanes);
    uint32_t total = 1;
    for (int d = 0; d < N; d++) total *= bs;
    for (uint32_t q = 0; q < total; q++) {
        uint32_t r = q, e = 0, s = 0;
        for (int d = N - 1; d >= 0; d--) {
            const uint32_t l = r % bs;
            r /= bs;
            e |= l << (8 * d);
            s += l;
        }
        planes[s].push_back(e);
    }
    std::vector<uint32_t> tab, diag;
    for (auto &pl: planes) {
        diag.push_back((uint32_t) tab.size());
        tab.insert(tab.end(), pl.begin(), pl.end());
    }
    diag.push_back((uint32_t) tab.size());
    HIPCHECK(ctx->gf_tab.ensure(tab.size()));
    HIPCHECK(ctx->gf_diag.ensure(diag.size()));
    HIPCHECK(hipMemcpy(ctx->gf_tab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(ctx->gf_diag.p, diag.data(), diag.size() * 4, hipMemcpyHostToDevice));
    ctx->gf_tab_key = key;
    return FAZ_OK;
}

// SZGeneralFrontend::compress over nfields arrays of p's shape laid end to
// end at d_in: 32-bit codes (ctx->gf_codes), outlier values at their code
// slots (ctx->outval), selections (ctx->gf_sel), reconstruction (recon).
// With the RegressionPredictor (one field): its coefficients into ro.
int gf_compress(faz_gpu_ctx *ctx, const float *d_in, float *recon, const faz_general_params *p, uint32_t nfields,
                GFRegOut *ro = nullptr) {
    int rc = gf_check(ctx, p);
    if (rc) return rc;
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const GFGeom G = gf_geom(p, nfields);
    const GFPred P = gf_pred(p);
    const GFQuant Q = gf_quant(p->eb, p->radius);
    const uint64_t n = G.num * nfields;
    if ((rc = gf_ensure_order(ctx, G.N, G.bs))) return rc;
    HIPCHECK(ctx->gf_codes.ensure(n));
    HIPCHECK(ctx->outval.ensure(n));
    HIPCHECK(ctx->gf_sel.ensure(G.nblocks * nfields));
    GFRegOut local;
    GFRegOut &o = ro ? *ro : local;
    o = GFRegOut();
    if (!P.reg) {
        GFReg R;
        memset(&R, 0, sizeof(R));
        gf_launch<0>(G, P, Q, R, d_in, recon, ctx->gf_codes.p, ctx->outval.p, ctx->gf_sel.p, ctx->gf_tab.p,
                     ctx->gf_diag.p, st);
        HIPCHECK(hipGetLastError());
        return FAZ_OK;
    }
    if (nfields != 1) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "the RegressionPredictor compresses one field at a time");
    GFReg R = gf_reg_args(G.N, p->eb, p->block_size);
    HIPCHECK(ctx->gf_reg_raw.ensure(5 * G.nblocks));
    HIPCHECK(ctx->gf_reg_deq.ensure(5 * G.nblocks));
    HIPCHECK(ctx->gf_reg_code.ensure(5 * G.nblocks));
    R.raw = ctx->gf_reg_raw.p;
    R.deq = ctx->gf_reg_deq.p;
    R.out_raw = ctx->gf_reg_raw.p;
    R.out_code = ctx->gf_reg_code.p;
    const int reg_sid = P.npred - 1;
    if (G.N == 1) {
        // one lane walks the line and runs the coefficient chain itself
        gf_launch<0>(G, P, Q, R, d_in, recon, ctx->gf_codes.p, ctx->outval.p, ctx->gf_sel.p, ctx->gf_tab.p,
                     ctx->gf_diag.p, st);
        HIPCHECK(hipGetLastError());
        std::vector<uint8_t> sel(P.npred > 1 ? G.nblocks : 0);
        std::vector<uint32_t> code(5 * G.nblocks);
        std::vector<float> raw(5 * G.nblocks);
        if (!sel.empty()) HIPCHECK(hipMemcpyAsync(sel.data(), ctx->gf_sel.p, sel.size(), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(code.data(), ctx->gf_reg_code.p, code.size() * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(raw.data(), ctx->gf_reg_raw.p, raw.size() * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        for (uint64_t b = 0; b < G.nblocks; b++) {
            const bool committed = P.npred > 1 ? sel[b] == reg_sid : gf_len((uint32_t) b, G.bs, G.dims[0]) > 1;
            if (!committed) continue;
            for (int i = 0; i < 2; i++) {
                o.codes.push_back((int32_t) code[5 * b + i]);
                if (!code[5 * b + i]) (i == 0 ? o.unpred_lin : o.unpred_ind).push_back(raw[5 * b + i]);
            }
            o.nblocks++;
        }
        return FAZ_OK;
    }
    // every block's plane, then passes until the choices settle
    gf_launch_fit(G, d_in, ctx->gf_reg_raw.p, st);
    HIPCHECK(hipGetLastError());
    std::vector<float> h_raw(5 * G.nblocks), deq(5 * G.nblocks, 0.0f);
    HIPCHECK(hipMemcpyAsync(h_raw.data(), ctx->gf_reg_raw.p, h_raw.size() * 4, hipMemcpyDeviceToHost, st));
    std::vector<uint8_t> degen(G.nblocks), S(G.nblocks), S2(G.nblocks);
    for (uint64_t b = 0; b < G.nblocks; b++) {
        degen[b] = gf_block_degen(G, b);
        S[b] = degen[b] ? 0 : (uint8_t) reg_sid;  // first guess: the regression everywhere
    }
    HIPCHECK(hipStreamSynchronize(st));
    int pass = 0;
    for (;;) {
        gf_chain(G, R, h_raw, S.data(), reg_sid, degen, deq, nullptr);
        HIPCHECK(hipMemcpyAsync(ctx->gf_reg_deq.p, deq.data(), deq.size() * 4, hipMemcpyHostToDevice, st));
        gf_launch<0>(G, P, Q, R, d_in, recon, ctx->gf_codes.p, ctx->outval.p, ctx->gf_sel.p, ctx->gf_tab.p,
                     ctx->gf_diag.p, st);
        HIPCHECK(hipGetLastError());
        pass++;
        if (P.npred == 1) {
            HIPCHECK(hipStreamSynchronize(st));
            break;
        }
        HIPCHECK(hipMemcpyAsync(S2.data(), ctx->gf_sel.p, G.nblocks, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (S2 == S) break;
        if (pass >= kGFMaxPasses)
            CTX_FAIL(FAZ_ERR_UNSUPPORTED, "predictor choices did not settle in %d wavefront passes", kGFMaxPasses);
        S.swap(S2);
    }
    gf_chain(G, R, h_raw, S.data(), reg_sid, degen, deq, &o);
    o.passes = pass;
    return FAZ_OK;
}

// Huffman of a host-side int stream over nsym symbols: HuffmanEncoder<int>
// preprocess_encode + save + encode (HUF:103-227).
bool gf_host_huffman(const int32_t *s, size_t n, int nsym, std::vector<uint8_t> &out) {
    // counts over the window of the nonzero symbols, symbol 0 apart
    uint32_t lo = UINT32_MAX, hi = 0;
    for (size_t i = 0; i < n; i++) {
        const uint32_t v = (uint32_t) s[i];
        if (v >= (uint32_t) nsym) return false;
        if (v) {
            lo = std::min(lo, v);
            hi = std::max(hi, v);
        }
    }
    if (lo > hi) lo = hi = 1;
    const uint32_t nwin = hi - lo + 1;
    std::vector<uint32_t> freq(nwin, 0), first(nwin, UINT32_MAX);
    uint32_t freq0 = 0, first0 = UINT32_MAX;
    for (size_t i = 0; i < n; i++) {
        const uint32_t v = (uint32_t) s[i];
        if (!v) {
            if (!freq0++) first0 = (uint32_t) i;
        } else if (!freq[v - lo]++) {
            first[v - lo] = (uint32_t) i;
        }
    }
    HuffTable t;
    if (build_huffman_window(freq.data(), first.data(), (int) lo, (int) nwin, freq0, first0, t)) return false;
    out.insert(out.end(), t.tree.begin(), t.tree.end());
    encode_huffman(t, s, n, out);
    return true;
}

// LinearQuantizer::save (Q:185-201).
void gf_save_quantizer(std::vector<uint8_t> &out, double eb, int radius, const std::vector<float> &unpred) {
    out.push_back(2);
    gput(out, eb);
    gput(out, (int32_t) radius);
    gput(out, (uint64_t) unpred.size());
    const uint8_t *u = reinterpret_cast<const uint8_t *>(unpred.data());
    out.insert(out.end(), u, u + unpred.size() * 4);
}

// The frontend's save (GF:98-103): dims, blockSize, each predictor's save in
// list order (LorenzoPredictor: its id, LP:57-59; RegressionPredictor: id,
// coefficient count, both coefficient quantizers, the codes' Huffman,
// RP:100-117), the ComposedPredictor's selections (CP:70-81) without the
// blocks that recorded none.
bool gf_save_front(const faz_general_params *p, const uint8_t *sel, size_t nsel, const GFRegOut *reg,
                   std::vector<uint8_t> &out) {
    for (int d = 0; d < p->N; d++) gput(out, (uint64_t) p->dims[d]);
    gput(out, (uint32_t) p->block_size);
    if (p->lorenzo) out.push_back(1);
    if (p->lorenzo2) out.push_back(1);
    if (p->regression) {
        static const GFRegOut none;
        if (!reg) reg = &none;
        const GFReg R = gf_reg_args(p->N, p->eb, p->block_size);
        out.push_back(2);
        gput(out, (uint64_t) reg->codes.size());
        if (!reg->codes.empty()) {
            gf_save_quantizer(out, R.eb_ind, kGFCoefRadius, reg->unpred_ind);
            gf_save_quantizer(out, R.eb_lin, kGFCoefRadius, reg->unpred_lin);
            if (!gf_host_huffman(reg->codes.data(), reg->codes.size(), 2 * kGFCoefRadius, out)) return false;
        }
    }
    const int npred = (p->lorenzo ? 1 : 0) + (p->lorenzo2 ? 1 : 0) + (p->regression ? 1 : 0);
    if (npred > 1) {
        std::vector<int32_t> s;
        s.reserve(nsel);
        for (size_t i = 0; i < nsel; i++)
            if (sel[i] != kGFNoSelection) s.push_back(sel[i]);
        gput(out, (uint64_t) s.size());
        if (!s.empty() && !gf_host_huffman(s.data(), s.size(), npred, out)) return false;
    }
    return true;
}

// SZGeneralCompressor<SZGeneralFrontend>::compress up to zstd (SGC:30-71) on
// one field: frontend, LinearQuantizer::save, codes' tree and bits (the codes
// narrowed to 16 bits for the device encoder: radius <= 32768 here).
int gf_raw(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_general_params *p,
           std::vector<uint8_t> &raw) {
    int rc = gf_check(ctx, p);
    if (rc) return rc;
    if (p->radius > 32768) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "radius above 32768 (16-bit codes)");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const GFGeom G = gf_geom(p, 1);
    const uint64_t n = G.num;
    float *recon = d_recon;
    if (!recon) {
        HIPCHECK(ctx->gf_recon.ensure(n));
        recon = ctx->gf_recon.p;
    }
    HIPCHECK(hipEventRecord(ctx->ev[0], st));
    GFRegOut ro;
    if ((rc = gf_compress(ctx, d_in, recon, p, 1, &ro))) return rc;
    HIPCHECK(hipEventRecord(ctx->ev[1], st));
    const uint32_t nchunks = (uint32_t) ((n + kHistChunk - 1) / kHistChunk);
    HIPCHECK(ctx->codes.ensure(n));
    HIPCHECK(ctx->scalars.ensure(4));
    HIPCHECK(ctx->tiles.ensure(scan_tiles(nchunks)));
    HIPCHECK(ctx->chunk_zeros.ensure(nchunks));
    HIPCHECK(ctx->chunk_base.ensure(nchunks));
    HIPCHECK(ctx->freq.ensure(kHistSlots));
    HIPCHECK(ctx->first.ensure(65536));
    hipLaunchKernelGGL(gf_narrow_kernel, dim3((uint32_t) std::min<uint64_t>((n + 255) / 256, 8192)), dim3(256), 0,
                       st, ctx->gf_codes.p, ctx->codes.p, n);
    launch_hist(ctx->codes.p, n, ctx->freq.p, ctx->first.p, ctx->chunk_zeros.p, p->radius, st);
    launch_zero_scan(ctx->chunk_zeros.p, ctx->chunk_base.p, nchunks, ctx->tiles.p, ctx->scalars.p, st);
    HIPCHECK(hipGetLastError());
    uint64_t n_unpred = 0;
    HIPCHECK(hipMemcpyAsync(&n_unpred, ctx->scalars.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    HIPCHECK(ctx->unpred.ensure(n_unpred));
    launch_unpred_gather(ctx->codes.p, n, ctx->outval.p, ctx->chunk_zeros.p, ctx->chunk_base.p, ctx->unpred.p, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(ctx->h_freq.data(), ctx->freq.p, 65536 * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(ctx->h_first.data(), ctx->first.p, 65536 * 4, hipMemcpyDeviceToHost, st));
    std::vector<uint8_t> sel(G.nblocks);
    HIPCHECK(hipMemcpyAsync(sel.data(), ctx->gf_sel.p, G.nblocks, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    ctx->hist_codes = ctx->codes.p;
    ctx->hist_n = n;
    HuffTable tab;
    if (build_huffman(ctx->h_freq.data(), ctx->h_first.data(), 65536, tab))
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "Huffman tree could not be built (codes longer than 64 bits)");
    faz_huff_table ht{tab.offset, tab.n_states, tab.code.data(), tab.len.data()};
    uint64_t maxbits = 0;
    for (auto l: tab.len) maxbits = std::max<uint64_t>(maxbits, l);
    HIPCHECK(ctx->words.ensure((size_t) ((n * maxbits + 31) / 32 + 2)));
    size_t nbytes = 0;
    rc = faz_gpu_huffman_encode(ctx, ctx->codes.p, n, &ht, reinterpret_cast<uint8_t *>(ctx->words.p),
                                ctx->words.cap * 4, &nbytes);
    if (rc) return rc;
    raw.clear();
    const size_t nsel = gf_pred(p).npred > 1 ? G.nblocks : 0;
    if (!gf_save_front(p, sel.data(), nsel, &ro, raw)) CTX_FAIL(FAZ_ERR_INVALID, "selection Huffman failed");
    raw.push_back(2);  // LinearQuantizer::save (Q:185-201)
    gput(raw, p->eb);
    gput(raw, (int32_t) p->radius);
    gput(raw, (uint64_t) n_unpred);
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
    HIPCHECK(hipEventRecord(ctx->ev[2], st));
    HIPCHECK(hipStreamSynchronize(st));
    ctx->timings[1] = gf_elapsed_ms(ctx->ev[0], ctx->ev[1]);
    ctx->timings[5] = gf_elapsed_ms(ctx->ev[0], ctx->ev[2]);
    return FAZ_OK;
}

// LinearQuantizer::load (Q:203-226) of a coefficient quantizer.
bool gf_load_quantizer(const uint8_t *&p, const uint8_t *end, double &eb, int32_t &radius,
                       std::vector<float> &unpred) {
    if ((size_t) (end - p) < 1 + 8 + 4 + 8) return false;
    uint64_t U;
    memcpy(&eb, p + 1, 8);
    memcpy(&radius, p + 9, 4);
    memcpy(&U, p + 13, 8);
    p += 21;
    if (U > (size_t) (end - p) / 4) return false;
    unpred.resize(U);
    if (U) memcpy(unpred.data(), p, U * 4);
    p += U * 4;
    return true;
}

// SGC:115-140 + GF:105-116 (CP:82-93, RP:134-156, 233-240) on a pre-zstd
// buffer; the predictor list comes from the stream's Config.
int gf_decompress(faz_gpu_ctx *ctx, const uint8_t *raw, size_t n, int N, int lo, int lo2, int regr, float *d_out,
                  size_t num) {
    const uint8_t *p = raw, *end = raw + n;
    auto need = [&](size_t k) { return (size_t) (end - p) >= k; };
    faz_general_params prm;
    memset(&prm, 0, sizeof(prm));
    prm.N = N;
    prm.lorenzo = lo, prm.lorenzo2 = lo2, prm.regression = regr;
    prm.eb = 1.0;
    prm.radius = 1;
    if (N < 1 || N > 4) CTX_FAIL(FAZ_ERR_INVALID, "N must be 1..4");
    if (!need((size_t) N * 8 + 4)) CTX_FAIL(FAZ_ERR_FORMAT, "general frontend header truncated");
    for (int d = 0; d < N; d++) {
        uint64_t v;
        memcpy(&v, p, 8);
        p += 8;
        prm.dims[d] = (size_t) v;
    }
    uint32_t bs;
    memcpy(&bs, p, 4);
    p += 4;
    prm.block_size = (int) bs;
    int rc = gf_check(ctx, &prm);
    if (rc) return rc;
    const GFGeom G = gf_geom(&prm, 1);
    if (G.num != num) CTX_FAIL(FAZ_ERR_INVALID, "buffer holds %llu points, %zu expected", (unsigned long long) G.num,
                               num);
    const GFPred P = gf_pred(&prm);
    // each predictor's load in list order; the regression's coefficients
    std::vector<int32_t> rcodes;
    std::vector<float> run_ind, run_lin;
    double reb_ind = 1, reb_lin = 1;
    int32_t rrad_ind = 1, rrad_lin = 1;
    for (int k = 0; k < P.npred; k++) {
        if (!need(1)) CTX_FAIL(FAZ_ERR_FORMAT, "predictor ids truncated");
        p += 1;
        if (P.kind[k] != GF_REG) continue;
        if (!need(8)) CTX_FAIL(FAZ_ERR_FORMAT, "regression coefficient count truncated");
        uint64_t nc;
        memcpy(&nc, p, 8);
        p += 8;
        if (nc > (uint64_t) (N + 1) * G.nblocks) CTX_FAIL(FAZ_ERR_FORMAT, "more regression coefficients than blocks");
        if (!nc) continue;
        if (!gf_load_quantizer(p, end, reb_ind, rrad_ind, run_ind) ||
            !gf_load_quantizer(p, end, reb_lin, rrad_lin, run_lin))
            CTX_FAIL(FAZ_ERR_FORMAT, "regression coefficient quantizers truncated");
        if (rrad_ind < 1 || rrad_ind > 65536 || rrad_lin < 1 || rrad_lin > 65536)
            CTX_FAIL(FAZ_ERR_FORMAT, "regression coefficient quantizer radius outside 1..65536");
        const size_t used = decode_huffman(p, end - p, rcodes, nc);
        if (!used) CTX_FAIL(FAZ_ERR_FORMAT, "regression coefficient Huffman section malformed");
        p += used;
    }
    std::vector<uint8_t> sel;
    if (P.npred > 1) {
        if (!need(8)) CTX_FAIL(FAZ_ERR_FORMAT, "selection count truncated");
        uint64_t ns;
        memcpy(&ns, p, 8);
        p += 8;
        if (ns != G.nblocks) CTX_FAIL(FAZ_ERR_FORMAT, "selection count does not match the blocks");
        std::vector<int32_t> s;
        size_t used = decode_huffman(p, end - p, s, ns);
        if (!used) CTX_FAIL(FAZ_ERR_FORMAT, "selection Huffman section malformed");
        p += used;
        sel.resize(ns);
        for (size_t i = 0; i < ns; i++) {
            if (s[i] < 0 || s[i] >= P.npred) CTX_FAIL(FAZ_ERR_FORMAT, "selection out of range");
            sel[i] = (uint8_t) s[i];
        }
    }
    // the coefficient chain (pred_and_recover_coefficients, RP:233-240):
    // each regression block's coefficients recovered against the previous
    // regression block's
    std::vector<float> deq;
    if (P.reg) {
        deq.assign(5 * G.nblocks, 0.0f);
        float cur[5] = {0, 0, 0, 0, 0};
        size_t ci = 0, ui_ind = 0, ui_lin = 0;
        for (uint64_t b = 0; b < G.nblocks; b++) {
            const int kind = P.npred > 1 ? gf_kind(P, sel[b]) : P.kind[0];
            if (kind != GF_REG || gf_block_degen(G, b)) continue;
            if (ci + N + 1 > rcodes.size()) CTX_FAIL(FAZ_ERR_FORMAT, "regression coefficient codes run out");
            for (int i = 0; i <= N; i++) {
                const bool lin = i < N;
                const int32_t code = rcodes[ci++];
                const int32_t rad = lin ? rrad_lin : rrad_ind;
                if (code < 0 || code >= 2 * rad) CTX_FAIL(FAZ_ERR_FORMAT, "regression coefficient code out of range");
                if (code) {
                    cur[i] = recover_point(cur[i], code, rad, lin ? reb_lin : reb_ind);
                } else {
                    std::vector<float> &u = lin ? run_lin : run_ind;
                    size_t &ui = lin ? ui_lin : ui_ind;
                    if (ui >= u.size()) CTX_FAIL(FAZ_ERR_FORMAT, "unpredictable regression coefficients run out");
                    cur[i] = u[ui++];
                }
            }
            for (int i = 0; i <= N; i++) deq[5 * b + i] = cur[i];
        }
    }
    if (!need(1 + 8 + 4 + 8)) CTX_FAIL(FAZ_ERR_FORMAT, "quantizer truncated");
    p += 1;
    double eb;
    int32_t radius;
    uint64_t U;
    memcpy(&eb, p, 8);
    memcpy(&radius, p + 8, 4);
    memcpy(&U, p + 12, 8);
    p += 20;
    if (radius < 1 || radius > 65536) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "radius outside 1..65536");
    if (!(eb > 0)) CTX_FAIL(FAZ_ERR_FORMAT, "error bound not positive");
    if (!need(U * 4)) CTX_FAIL(FAZ_ERR_FORMAT, "unpredictable list truncated");
    const float *unpred_h = reinterpret_cast<const float *>(p);
    p += U * 4;
    prm.eb = eb;
    prm.radius = radius;
    const GFQuant Q = gf_quant(eb, radius);
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if ((rc = gf_ensure_order(ctx, N, G.bs))) return rc;
    HIPCHECK(ctx->gf_codes.ensure(num));
    HIPCHECK(ctx->outval.ensure(num));
    HIPCHECK(ctx->gf_sel.ensure(G.nblocks));
    GFReg R;
    memset(&R, 0, sizeof(R));
    if (P.reg) {
        HIPCHECK(ctx->gf_reg_deq.ensure(deq.size()));
        HIPCHECK(hipMemcpyAsync(ctx->gf_reg_deq.p, deq.data(), deq.size() * 4, hipMemcpyHostToDevice, st));
        R.deq = ctx->gf_reg_deq.p;
    }
    // codes decoded on the device (HUF:236-268), each below 2 * radius, then
    // the unpredictable values at the zero codes
    if ((rc = gpu_huffman_decode<uint32_t>(ctx, p, end - p, num, 2u * (uint32_t) radius, ctx->gf_codes.p, nullptr)))
        return rc;
    if ((rc = gpu_scatter_unpred_u32(ctx, ctx->gf_codes.p, num, unpred_h, U, ctx->outval.p))) return rc;
    if (!sel.empty()) HIPCHECK(hipMemcpyAsync(ctx->gf_sel.p, sel.data(), sel.size(), hipMemcpyHostToDevice, st));
    ctx->hist_codes = nullptr;
    gf_launch<1>(G, P, Q, R, nullptr, d_out, ctx->gf_codes.p, ctx->outval.p, ctx->gf_sel.p, ctx->gf_tab.p,
                 ctx->gf_diag.p, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    return FAZ_OK;
}

}  // namespace

// CompressTest's Lorenzo variant (SZI:1136-1170, 1254-1256, 1283-1294,
// 1353-1392): one general compressor with lorenzo + lorenzo2, blockSize 5,
// radius 65536 over the sample blocks, the squared error of the overwritten
// blocks, encoding_lossless over the accumulated codes, zstd.
int gf_compress_test(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *fdims, const size_t *starts,
                     size_t nblocks, size_t sbs, double eb, std::vector<uint8_t> &raw, size_t *bytes,
                     double *square_error) {
    if (N != 2 && N != 3) CTX_FAIL(FAZ_ERR_INVALID, "CompressTest runs on 2-D and 3-D fields");
    if (!nblocks) CTX_FAIL(FAZ_ERR_INVALID, "no sample blocks");
    for (size_t b = 0; b < nblocks; b++)
        for (int d = 0; d < N; d++)
            if (starts[3 * b + d] + sbs + 1 > fdims[d]) CTX_FAIL(FAZ_ERR_INVALID, "sample block %zu leaves the field", b);
    faz_general_params p;
    memset(&p, 0, sizeof(p));
    p.N = N;
    for (int d = 0; d < N; d++) p.dims[d] = sbs + 1;
    p.eb = eb;
    p.block_size = 5;
    p.lorenzo = p.lorenzo2 = 1;
    p.radius = 65536;
    int rc = gf_check(ctx, &p);
    if (rc) return rc;
    if (nblocks > 65535) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 65535 sample blocks");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const GFGeom G = gf_geom(&p, (uint32_t) nblocks);
    const uint64_t total = G.num * nblocks;
    HIPCHECK(ctx->gf_fields.ensure(2 * total));
    HIPCHECK(ctx->scalars.ensure(3 * nblocks));
    std::vector<uint64_t> st64(starts, starts + 3 * nblocks);
    uint64_t *d_starts = reinterpret_cast<uint64_t *>(ctx->scalars.p);
    HIPCHECK(hipMemcpyAsync(d_starts, st64.data(), st64.size() * 8, hipMemcpyHostToDevice, st));
    const uint64_t s0 = N == 2 ? fdims[1] : (uint64_t) fdims[1] * fdims[2], s1 = N == 3 ? fdims[2] : 1;
    float *blocks = ctx->gf_fields.p, *recon = ctx->gf_fields.p + total;
    hipLaunchKernelGGL(gf_gather_kernel, dim3((uint32_t) std::min<uint64_t>((total + 255) / 256, 8192)), dim3(256),
                       0, st, d_in, N, s0, s1, (uint32_t) (sbs + 1), d_starts, (uint32_t) nblocks, blocks);
    HIPCHECK(hipGetLastError());
    if ((rc = gf_compress(ctx, blocks, recon, &p, (uint32_t) nblocks))) return rc;
    std::vector<float> h_in(total), h_rec(total), h_ov(total);
    std::vector<uint32_t> h_codes(total);
    std::vector<uint8_t> h_sel(G.nblocks * nblocks);
    HIPCHECK(hipMemcpyAsync(h_in.data(), blocks, total * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(h_rec.data(), recon, total * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(h_ov.data(), ctx->outval.p, total * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(h_codes.data(), ctx->gf_codes.p, total * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(h_sel.data(), ctx->gf_sel.p, h_sel.size(), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    double sq = 0;  // SZI:1287-1292: float difference, float square, double sum in order
    for (uint64_t i = 0; i < total; i++) {
        const float v = h_in[i] - h_rec[i];
        sq += v * v;
    }
    raw.clear();
    if (!gf_save_front(&p, h_sel.data(), h_sel.size(), nullptr, raw))
        CTX_FAIL(FAZ_ERR_INVALID, "selection Huffman failed");
    std::vector<float> unpred;
    std::vector<int32_t> c(total);
    for (uint64_t i = 0; i < total; i++) {
        c[i] = (int32_t) h_codes[i];
        if (!h_codes[i]) unpred.push_back(h_ov[i]);
    }
    raw.push_back(2);
    gput(raw, eb);
    gput(raw, (int32_t) p.radius);
    gput(raw, (uint64_t) unpred.size());
    const uint8_t *u = reinterpret_cast<const uint8_t *>(unpred.data());
    raw.insert(raw.end(), u, u + unpred.size() * 4);
    if (!gf_host_huffman(c.data(), c.size(), 2 * p.radius, raw)) CTX_FAIL(FAZ_ERR_INVALID, "code Huffman failed");
    std::vector<uint8_t> z(ZSTD_compressBound(raw.size()));
    const size_t zs = ZSTD_compress(z.data(), z.size(), raw.data(), raw.size(), 3);
    if (ZSTD_isError(zs)) CTX_FAIL(FAZ_ERR_INVALID, "zstd failed");
    *bytes = 8 + zs;
    *square_error = sq;
    return FAZ_OK;
}

}  // namespace faz

int faz_general_raw_vec(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_general_params *p,
                        std::vector<uint8_t> &raw) {
    if (!ctx || !d_in || !p) return FAZ_ERR_INVALID;
    return faz::gf_raw(ctx, d_in, d_recon, p, raw);
}

int faz_general_decompress_raw_vec(faz_gpu_ctx *ctx, const uint8_t *raw, size_t n, int N, int lorenzo, int lorenzo2,
                                   int regression, float *d_out, size_t num) {
    if (!ctx || !raw || !d_out) return FAZ_ERR_INVALID;
    return faz::gf_decompress(ctx, raw, n, N, lorenzo, lorenzo2, regression, d_out, num);
}

extern "C" {

faz_status faz_gpu_general_compress_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon,
                                        const faz_general_params *p, faz_general_out *out) {
    if (!ctx || !d_in || !d_recon || !p || !out) return FAZ_ERR_INVALID;
    faz::GFRegOut ro;
    int rc = faz::gf_compress(ctx, d_in, d_recon, p, 1, &ro);
    if (rc) return (faz_status) rc;
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    const faz::GFGeom G = faz::gf_geom(p, 1);
    out->d_codes = ctx->gf_codes.p;
    out->n_codes = G.num;
    out->d_outval = ctx->outval.p;
    out->d_selection = ctx->gf_sel.p;
    out->n_selection = faz::gf_pred(p).npred > 1 ? G.nblocks : 0;
    out->n_regression = ro.nblocks;
    out->passes = ro.passes;
    return FAZ_OK;
}

faz_status faz_gpu_general_raw_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_general_params *p,
                                   uint8_t *h_raw, size_t cap, size_t *raw_size) {
    if (!ctx || !d_in || !p || !raw_size) return FAZ_ERR_INVALID;
    std::vector<uint8_t> raw;
    int rc = faz::gf_raw(ctx, d_in, d_recon, p, raw);
    if (rc) return (faz_status) rc;
    *raw_size = raw.size();
    if (!h_raw || cap < raw.size()) CTX_FAIL(FAZ_ERR_CAPACITY, "output holds %zu bytes, %zu needed", cap, raw.size());
    memcpy(h_raw, raw.data(), raw.size());
    return FAZ_OK;
}

faz_status faz_gpu_general_decompress_raw(faz_gpu_ctx *ctx, const uint8_t *h_raw, size_t n, int N, int lorenzo,
                                          int lorenzo2, int regression, float *d_out, size_t num) {
    if (!ctx || !h_raw || !d_out) return FAZ_ERR_INVALID;
    return (faz_status) faz::gf_decompress(ctx, h_raw, n, N, lorenzo, lorenzo2, regression, d_out, num);
}

faz_status faz_gpu_compress_test_lorenzo_f32(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *fdims,
                                             const size_t *starts, size_t n_blocks, size_t sbs, double eb,
                                             uint8_t *h_raw, size_t cap, size_t *raw_size, size_t *bytes,
                                             double *square_error) {
    if (!ctx || !d_in || !fdims || !starts || !bytes || !square_error) return FAZ_ERR_INVALID;
    std::vector<uint8_t> raw;
    int rc = faz::gf_compress_test(ctx, d_in, N, fdims, starts, n_blocks, sbs, eb, raw, bytes, square_error);
    if (rc) return (faz_status) rc;
    if (raw_size) *raw_size = raw.size();
    if (h_raw) {
        if (cap < raw.size()) CTX_FAIL(FAZ_ERR_CAPACITY, "output holds %zu bytes, %zu needed", cap, raw.size());
        memcpy(h_raw, raw.data(), raw.size());
    }
    return FAZ_OK;
}

}  // extern "C"
