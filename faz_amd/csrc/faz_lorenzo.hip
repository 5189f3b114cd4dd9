// 3-D Lorenzo/regression engine on gfx950: SZFastFrontend (include/QoZ/
// frontend/SZFastFrontend.hpp:176-345, 429-530) with the predictors of
// MetaLorenzoPredictor.hpp / MetaRegressionPredictor.hpp and the framing of
// SZGeneralCompressor (SZGeneralCompressor.hpp:30-71, 115-140).
//
// The reference sweeps blocks in raster order, one point at a time.  Here:
//   1. lr_stats_kernel, one thread per block: regression coefficients (float
//      sums in the reference's loop order) and the predictor selection, which
//      reads only original values inside the block, so every block is
//      independent;
//   2. the regression coefficients are quantized on the host against the
//      previous regression block's (the one sequential chain of the engine,
//      four floats per regression block);
//   3. lr_regression_kernel: regression blocks predict from their own
//      coefficients only, all at once;
//   4. lr_lorenzo_kernel, one launch per block anti-diagonal bi+bj+bk = T:
//      a Lorenzo point reads reconstructed values at offsets in {0,-1,-2}^3,
//      which lie in its own block or in blocks of smaller diagonals, zero
//      outside the field (the reference's zero-padded pred_buffer, FF:228-243).
//      One wave per block stages the block and its two low halo layers in LDS
//      and walks the in-block hyperplanes i+j+k = s, whose points are
//      independent of each other.
// Codes are block-major as the reference pushes them (FF:191, 212).
#include "faz_gpu.h"
#include "faz_ctx.h"
#include "faz_device.h"
#include "faz_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#pragma clang fp contract(off)

using namespace faz;

namespace faz {
namespace {

constexpr int kSelLorenzo = 0, kSelRegression = 1, kSelLorenzo2 = 2;  // MetaDef.hpp:18-20
constexpr uint32_t kLRMaxBlock = 16;  // largest blockSize the Lorenzo kernel stages (5-bit table coordinates)

struct LRGeom {
    uint32_t d1, d2, d3;  // x (slowest), y, z (contiguous)
    uint64_t s0, s1;      // element strides of x and y
    uint32_t bs;
    uint32_t nx, ny, nz;  // blocks per axis
    uint64_t nblocks;
};

__host__ __device__ __forceinline__ uint32_t blk_len(uint32_t b, uint32_t bs, uint32_t d) {
    return (b + 1) * bs < d ? bs : d - b * bs;
}

// First code of block (bi, bj, bk) in the block-major code stream: whole
// x-slabs, then whole y-rows of this slab, then earlier blocks of the row.
__host__ __device__ __forceinline__ uint64_t blk_code_off(const LRGeom &G, uint32_t bi, uint32_t bj, uint32_t bk,
                                                          uint32_t sx, uint32_t sy) {
    return (uint64_t) bi * G.bs * G.s0 + (uint64_t) sx * ((uint64_t) bj * G.bs * G.d3 + (uint64_t) sy * bk * G.bs);
}

// MetaLorenzoPredictor.hpp:9-77 on f(dx, dy, dz), offsets <= 0 along x, y, z;
// float arithmetic left to right as the reference evaluates it.
template<int PD, bool TWO, class F>
__device__ __forceinline__ float lorenzo_pred(F f) {
    if constexpr (!TWO) {
        if constexpr (PD == 3)
            return f(0, 0, -1) + f(0, -1, 0) + f(-1, 0, 0) - f(0, -1, -1) - f(-1, 0, -1) - f(-1, -1, 0) +
                   f(-1, -1, -1);
        else if constexpr (PD == 2)
            return f(0, 0, -1) + f(-1, 0, 0) - f(-1, 0, -1);
        else
            return f(0, 0, -1);
    } else {
        if constexpr (PD == 3)
            return 2 * f(0, 0, -1) - f(0, 0, -2) + 2 * f(0, -1, 0) - 4 * f(0, -1, -1) + 2 * f(0, -1, -2) -
                   f(0, -2, 0) + 2 * f(0, -2, -1) - f(0, -2, -2) + 2 * f(-1, 0, 0) - 4 * f(-1, 0, -1) +
                   2 * f(-1, 0, -2) - 4 * f(-1, -1, 0) + 8 * f(-1, -1, -1) - 4 * f(-1, -1, -2) +
                   2 * f(-1, -2, 0) - 4 * f(-1, -2, -1) + 2 * f(-1, -2, -2) - f(-2, 0, 0) + 2 * f(-2, 0, -1) -
                   f(-2, 0, -2) + 2 * f(-2, -1, 0) - 4 * f(-2, -1, -1) + 2 * f(-2, -1, -2) - f(-2, -2, 0) +
                   2 * f(-2, -2, -1) - f(-2, -2, -2);
        else if constexpr (PD == 2)
            return 2 * f(-1, 0, 0) - f(-2, 0, 0) + 2 * f(0, 0, -1) - 4 * f(-1, 0, -1) + 2 * f(-2, 0, -1) -
                   f(0, 0, -2) + 2 * f(-1, 0, -2) - f(-2, 0, -2);
        else
            return 2 * f(0, 0, -1) - f(0, 0, -2);
    }
}

// regression_predict_3d (MRP:148-152, 169-171)
__device__ __forceinline__ float reg_pred(const float4 &c, int x, int y, int z) {
    return c.x * (float) x + c.y * (float) y + c.z * (float) z + c.w;
}

struct LRSel {
    int pred_dim;
    int lo, lo2, reg;
    float prec;  // the bound as the frontend's float precision argument
};

// Original-data Lorenzo predictions of the selection samples (FF:429-473).
template<int PD>
__device__ __forceinline__ void sample_lorenzo(const float *__restrict__ p, uint64_t s0, uint64_t s1, bool lo,
                                               bool lo2, double &lp, double &lp2) {
    auto f = [&](int dx, int dy, int dz) { return p[(int64_t) dx * (int64_t) s0 + (int64_t) dy * (int64_t) s1 + dz]; };
    if (lo2) lp2 = lorenzo_pred<PD, true>(f);
    if (lo) lp = lorenzo_pred<PD, false>(f);
}

// One thread per block: compute_regression_coeffcients_3d (MRP:117-145) and
// meta_blockwise_selection_3d (FF:476-530).
__global__ void __launch_bounds__(256) lr_stats_kernel(LRGeom G, LRSel S, const float *__restrict__ in,
                                                       uint8_t *__restrict__ ind, float4 *__restrict__ coef,
                                                       uint32_t *__restrict__ regflag) {
    const uint64_t b = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= G.nblocks) return;
    const uint32_t bk = (uint32_t) (b % G.nz);
    const uint64_t t = b / G.nz;
    const uint32_t bj = (uint32_t) (t % G.ny), bi = (uint32_t) (t / G.ny);
    const int sx = (int) blk_len(bi, G.bs, G.d1), sy = (int) blk_len(bj, G.bs, G.d2), sz = (int) blk_len(bk, G.bs, G.d3);
    const int min_size = min(min(sx, sy), sz);
    const bool use_reg = S.reg && min_size >= 2;
    const float *base = in + (uint64_t) bi * G.bs * G.s0 + (uint64_t) bj * G.bs * G.s1 + (uint64_t) bk * G.bs;
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
    if (use_reg) {
        float fx = 0, fy = 0, fz = 0, f = 0, sum_x, sum_y;
        for (int i = 0; i < sx; i++) {
            sum_x = 0;
            for (int j = 0; j < sy; j++) {
                sum_y = 0;
                const float *row = base + (uint64_t) i * G.s0 + (uint64_t) j * G.s1;
                for (int k = 0; k < sz; k++) {
                    const float v = row[k];
                    sum_y += v;
                    fz += v * (float) k;
                }
                fy += sum_y * (float) j;
                sum_x += sum_y;
            }
            fx += sum_x * (float) i;
            f += sum_x;
        }
        const float coeff = (float) (1.0 / (double) (sx * sy * sz));
        c.x = (2 * fx / (float) (sx - 1) - f) * 6 * coeff / (float) (sx + 1);
        c.y = (2 * fy / (float) (sy - 1) - f) * 6 * coeff / (float) (sy + 1);
        c.z = (2 * fz / (float) (sz - 1) - f) * 6 * coeff / (float) (sz + 1);
        c.w = f * coeff - ((float) (sx - 1) * c.x / 2 + (float) (sy - 1) * c.y / 2 + (float) (sz - 1) * c.z / 2);
    }
    double e_lo = 0, e_lo2 = 0, e_reg = 0;
    const double n1 = S.pred_dim == 3 ? 1.22 : (S.pred_dim == 2 ? 0.81 : 0.5);   // MetaDef.hpp:26-31
    const double n2 = S.pred_dim == 3 ? 6.8 : (S.pred_dim == 2 ? 2.76 : 1.08);
    auto est = [&](int x, int y, int z) {
        const float *p = base + (uint64_t) x * G.s0 + (uint64_t) y * G.s1 + z;
        const float cur = *p;
        float noise = 0, noise2 = 0;
        if (use_reg) e_reg += (double) fabsf(cur - reg_pred(c, x, y, z));
        double lp = 0, lp2 = 0;
        if (S.pred_dim == 3)
            sample_lorenzo<3>(p, G.s0, G.s1, S.lo, S.lo2, lp, lp2);
        else if (S.pred_dim == 2)
            sample_lorenzo<2>(p, G.s0, G.s1, S.lo, S.lo2, lp, lp2);
        else
            sample_lorenzo<1>(p, G.s0, G.s1, S.lo, S.lo2, lp, lp2);
        if (S.lo2) noise2 = (float) (n2 * (double) S.prec);
        if (S.lo) noise = (float) (n1 * (double) S.prec);
        e_lo += fabs((double) cur - lp) + (double) noise;
        e_lo2 += fabs((double) cur - lp2) + (double) noise2;
    };
    for (int i = 2; i < min_size - 1; i++) {
        const int m = min_size - i;
        est(i, i, i);
        est(i, i, m);
        est(i, m, i);
        est(i, m, m);
    }
    if (min_size > 3) est(min_size - 1, min_size - 1, min_size - 1);
    int sel;
    if (use_reg && (!S.lo || e_reg <= e_lo) && (!S.lo2 || e_reg < e_lo2))
        sel = kSelRegression;
    else if (S.lo2 && (!S.lo || e_lo2 <= e_lo) && (!use_reg || e_lo2 <= e_reg))
        sel = kSelLorenzo2;
    else
        sel = kSelLorenzo;
    ind[b] = (uint8_t) sel;
    coef[b] = c;
    regflag[b] = sel == kSelRegression;
}

// Unquantized coefficients of the regression blocks, in block order.
__global__ void lr_gather_coef_kernel(uint64_t nblocks, const uint32_t *__restrict__ regflag,
                                      const uint64_t *__restrict__ rank, const float4 *__restrict__ coef,
                                      float4 *__restrict__ out) {
    const uint64_t b = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nblocks && regflag[b]) out[rank[b]] = coef[b];
}

struct LRQuant {
    double eb, recip;
    float ebf;
    int radius;
};

// Regression blocks (MRP:155-183 compress, 213-247 decompress), one wave per
// block, all at once.  MODE 0: quantize, write codes / outliers / recon;
// MODE 1: recover from codes and the unpredictable value of each code slot.
template<int MODE>
__global__ void __launch_bounds__(256) lr_regression_kernel(LRGeom G, LRQuant Q, const float *__restrict__ in,
                                                            float *__restrict__ recon, uint16_t *__restrict__ codes,
                                                            float *__restrict__ outval,
                                                            const uint8_t *__restrict__ ind,
                                                            const uint64_t *__restrict__ rank,
                                                            const float4 *__restrict__ regq) {
    const uint32_t lane = threadIdx.x % 64;
    const uint64_t nwaves = (uint64_t) gridDim.x * (blockDim.x / 64);
    for (uint64_t b = (uint64_t) blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; b < G.nblocks; b += nwaves) {
        if (ind[b] != kSelRegression) continue;
        const uint32_t bk = (uint32_t) (b % G.nz);
        const uint64_t t = b / G.nz;
        const uint32_t bj = (uint32_t) (t % G.ny), bi = (uint32_t) (t / G.ny);
        const uint32_t sx = blk_len(bi, G.bs, G.d1), sy = blk_len(bj, G.bs, G.d2), sz = blk_len(bk, G.bs, G.d3);
        const float4 c = regq[rank[b]];
        const uint64_t off = blk_code_off(G, bi, bj, bk, sx, sy);
        const uint64_t base = (uint64_t) bi * G.bs * G.s0 + (uint64_t) bj * G.bs * G.s1 + (uint64_t) bk * G.bs;
        const uint32_t n = sx * sy * sz;
        for (uint32_t p = lane; p < n; p += 64) {
            const uint32_t k = p % sz, r = p / sz, j = r % sy, i = r / sy;
            const uint64_t at = base + (uint64_t) i * G.s0 + (uint64_t) j * G.s1 + k;
            const float pred = reg_pred(c, (int) i, (int) j, (int) k);
            float v;
            if constexpr (MODE == 0) {
                const float o = in[at];
                const int code = quantize_select(o, pred, Q.eb, Q.recip, Q.ebf, Q.radius, &v);
                codes[off + p] = (uint16_t) code;
                if (code == 0) outval[off + p] = o;
            } else {
                const int code = codes[off + p];
                v = code ? recover_point(pred, code, Q.radius, Q.eb) : outval[off + p];
            }
            recon[at] = v;
        }
    }
}

// Lorenzo blocks of anti-diagonal T (MLP_:80-215), one wave per block; the
// block's selector picks the 1-layer or 2-layer predictor.
// tab lists the points of a full bs^3 block ordered by i+j+k (i | j << 5 |
// k << 10), diag[s] where hyperplane s starts in it.
template<int MODE, int PD>
__global__ void __launch_bounds__(64) lr_lorenzo_kernel(LRGeom G, LRQuant Q, uint32_t T, uint32_t bi_lo,
                                                        uint32_t bj_lo, const float *__restrict__ in,
                                                        float *__restrict__ recon, uint16_t *__restrict__ codes,
                                                        float *__restrict__ outval, const uint8_t *__restrict__ ind,
                                                        const uint16_t *__restrict__ tab,
                                                        const uint32_t *__restrict__ diag) {
    // LDS sized for the block edge at launch (lr_lds_bytes)
    extern __shared__ float s_lr[];
    float *s_rec = s_lr;                                           // (bs+2)^3
    float *s_val = s_lr + (G.bs + 2) * (G.bs + 2) * (G.bs + 2);    // bs^3
    uint16_t *s_code = reinterpret_cast<uint16_t *>(s_val + G.bs * G.bs * G.bs);
    const uint32_t bi = bi_lo + blockIdx.x, bj = bj_lo + blockIdx.y;
    if (bi >= G.nx || bj >= G.ny || bi + bj > T || T - bi - bj >= G.nz) return;
    const uint32_t bk = T - bi - bj;
    const uint64_t b = ((uint64_t) bi * G.ny + bj) * G.nz + bk;
    const uint8_t sel = ind[b];
    if (sel == kSelRegression) return;
    const bool two = sel == kSelLorenzo2;
    const uint32_t lane = threadIdx.x;
    const uint32_t bs = G.bs, E = bs + 2;
    const uint32_t sx = blk_len(bi, bs, G.d1), sy = blk_len(bj, bs, G.d2), sz = blk_len(bk, bs, G.d3);
    const uint32_t n = sx * sy * sz;
    const uint64_t off = blk_code_off(G, bi, bj, bk, sx, sy);
    const int64_t x0 = (int64_t) bi * bs, y0 = (int64_t) bj * bs, z0 = (int64_t) bk * bs;
    // halo: reconstructed values of the two layers below each face, 0 outside the field
    for (uint32_t q = lane; q < E * E * E; q += 64) {
        const uint32_t c = q % E, r = q / E, bb = r % E, a = r / E;
        if (a >= 2 && bb >= 2 && c >= 2) continue;
        if (a >= sx + 2 || bb >= sy + 2 || c >= sz + 2) continue;
        const int64_t x = x0 + a - 2, y = y0 + bb - 2, z = z0 + c - 2;
        float v = 0.0f;
        if (x >= 0 && y >= 0 && z >= 0) v = recon[(uint64_t) x * G.s0 + (uint64_t) y * G.s1 + (uint64_t) z];
        s_rec[q] = v;
    }
    // the block's originals (compress) or codes and outlier values (decompress), block-major
    for (uint32_t p = lane; p < n; p += 64) {
        if constexpr (MODE == 0) {
            const uint32_t k = p % sz, r = p / sz, j = r % sy, i = r / sy;
            s_val[p] = in[(uint64_t) (x0 + i) * G.s0 + (uint64_t) (y0 + j) * G.s1 + (uint64_t) (z0 + k)];
        } else {
            const uint16_t code = codes[off + p];
            s_code[p] = code;
            s_val[p] = code ? 0.0f : outval[off + p];
        }
    }
    __syncthreads();
    const uint32_t nsteps = sx + sy + sz - 2;
    for (uint32_t s = 0; s < nsteps; s++) {
        const uint32_t q1 = diag[s + 1];
        for (uint32_t q = diag[s] + lane; q < q1; q += 64) {
            const uint32_t e = tab[q];
            const uint32_t i = e & 31, j = (e >> 5) & 31, k = e >> 10;
            if (i >= sx || j >= sy || k >= sz) continue;
            const uint32_t slot = ((i + 2) * E + (j + 2)) * E + (k + 2);
            auto f = [&](int dx, int dy, int dz) {
                return s_rec[(int) slot + dx * (int) (E * E) + dy * (int) E + dz];
            };
            const float pred = two ? lorenzo_pred<PD, true>(f) : lorenzo_pred<PD, false>(f);
            const uint32_t p = (i * sy + j) * sz + k;
            float v;
            if constexpr (MODE == 0) {
                const int code = quantize_select(s_val[p], pred, Q.eb, Q.recip, Q.ebf, Q.radius, &v);
                s_code[p] = (uint16_t) code;
            } else {
                const int code = s_code[p];
                v = code ? recover_point(pred, code, Q.radius, Q.eb) : s_val[p];
            }
            s_rec[slot] = v;
        }
        __syncthreads();
    }
    for (uint32_t p = lane; p < n; p += 64) {
        const uint32_t k = p % sz, r = p / sz, j = r % sy, i = r / sy;
        recon[(uint64_t) (x0 + i) * G.s0 + (uint64_t) (y0 + j) * G.s1 + (uint64_t) (z0 + k)] =
                s_rec[((i + 2) * E + (j + 2)) * E + (k + 2)];
        if constexpr (MODE == 0) {
            const uint16_t code = s_code[p];
            codes[off + p] = code;
            if (code == 0) outval[off + p] = s_val[p];
        }
    }
}

size_t lr_lds_bytes(uint32_t bs) {
    return (size_t) (bs + 2) * (bs + 2) * (bs + 2) * 4 + (size_t) bs * bs * bs * 6;
}

template<int MODE, int PD>
void launch_lorenzo_diagonals(const LRGeom &G, const LRQuant &Q, const float *in, float *recon, uint16_t *codes,
                              float *outval, const uint8_t *ind, const uint16_t *tab, const uint32_t *diag,
                              hipStream_t st) {
    const uint32_t Tmax = G.nx + G.ny + G.nz - 3;
    for (uint32_t T = 0; T <= Tmax; T++) {
        const uint32_t bi_lo = T > (G.ny - 1) + (G.nz - 1) ? T - (G.ny - 1) - (G.nz - 1) : 0;
        const uint32_t bi_hi = std::min(G.nx - 1, T);
        const uint32_t bj_lo = T > (G.nx - 1) + (G.nz - 1) ? T - (G.nx - 1) - (G.nz - 1) : 0;
        const uint32_t bj_hi = std::min(G.ny - 1, T);
        hipLaunchKernelGGL((lr_lorenzo_kernel<MODE, PD>), dim3(bi_hi - bi_lo + 1, bj_hi - bj_lo + 1), dim3(64),
                           lr_lds_bytes(G.bs), st, G, Q, T, bi_lo, bj_lo, in, recon, codes, outval, ind, tab, diag);
    }
}

template<int MODE>
void launch_lorenzo_all(int pd, const LRGeom &G, const LRQuant &Q, const float *in, float *recon, uint16_t *codes,
                        float *outval, const uint8_t *ind, const uint16_t *tab, const uint32_t *diag,
                        hipStream_t st) {
    if (pd == 3)
        launch_lorenzo_diagonals<MODE, 3>(G, Q, in, recon, codes, outval, ind, tab, diag, st);
    else if (pd == 2)
        launch_lorenzo_diagonals<MODE, 2>(G, Q, in, recon, codes, outval, ind, tab, diag, st);
    else
        launch_lorenzo_diagonals<MODE, 1>(G, Q, in, recon, codes, outval, ind, tab, diag, st);
}

}  // namespace
}  // namespace faz

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
namespace faz {
namespace {

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

// meta_params (MetaDef.hpp:33-69) as the frontend builds it (FF:25-31): the
// bound enters as a float, the coefficient bounds are float expressions.
struct MetaParams {
    int block_size, pred_dim;
    bool lo, lo2, reg;
    float eb_indep, eb_linear;
};

MetaParams meta_params(const faz_lorenzo_params *p) {
    MetaParams m;
    m.block_size = p->block_size;
    m.pred_dim = p->pred_dim;
    m.lo = p->lorenzo != 0;
    m.lo2 = p->lorenzo2 != 0;
    m.reg = p->regression != 0;
    const float prec = (float) p->eb;
    m.eb_linear = 1.0f * prec / 4 / (float) p->block_size;  // reg_eb_base * precision / RegCoeffNum3d / bs
    m.eb_indep = (float) p->block_size * m.eb_linear;      // reg_eb_1 = bs
    return m;
}

// Raw bytes of meta_params (44: int, int, 3 bool, pad, int, int, 5 float, bool, pad 3).
void put_meta_params(const MetaParams &m, std::vector<uint8_t> &out) {
    uint8_t b[44];
    memset(b, 0, sizeof(b));
    auto put = [&](size_t at, const void *v, size_t n) { memcpy(b + at, v, n); };
    int lp = 2, cap = 0;
    float one = 1.0f, bsf = (float) m.block_size;
    put(0, &m.block_size, 4);
    put(4, &m.pred_dim, 4);
    b[8] = m.lo, b[9] = m.lo2, b[10] = m.reg;
    put(12, &lp, 4);             // lorenzo_padding_layer
    put(16, &cap, 4);            // capacity
    put(20, &m.eb_indep, 4);     // regression_param_eb_independent
    put(24, &m.eb_linear, 4);    // regression_param_eb_linear
    put(28, &one, 4);            // reg_eb_base
    put(32, &bsf, 4);            // reg_eb_1
    put(36, &one, 4);            // sample_ratio
    b[40] = 1;                   // lossless
    out.insert(out.end(), b, b + sizeof(b));
}

template<class V>
void put(std::vector<uint8_t> &out, const V &v) {
    const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
    out.insert(out.end(), p, p + sizeof(V));
}

// quantize_reg_coeff (MRP:10-27): capacity 65536, radius 32768.
inline int quantize_coeff(float pred, float cur, double precision, double recip, float *dec,
                          std::vector<float> &unpred) {
    double diff = cur - pred;
    double qd = std::fabs(diff) * recip + 1;
    if (qd < 65536) {
        qd = (diff > 0) ? qd : -qd;
        const int qi = (int) (qd / 2) + 32768;
        const float d = (float) (pred + 2 * (qi - 32768) * precision);
        *dec = d;
        if (std::fabs(d - cur) <= precision) return qi;
    }
    *dec = cur;
    unpred.push_back(cur);
    return 0;
}

// compress_regression_coefficient_3d over the regression blocks in order
// (MRP:40-51, FF:209-216): each coefficient is predicted by the same
// coefficient of the previous regression block (0 for the first).
void coefficient_chain(const MetaParams &m, const float4 *raw, size_t nreg, std::vector<int32_t> &types,
                       std::vector<float> &unpred, std::vector<float4> &deq) {
    const float prc[4] = {m.eb_linear, m.eb_linear, m.eb_linear, m.eb_indep};
    float rcp[4];
    for (int i = 0; i < 4; i++) rcp[i] = (float) (1.0 / prc[i]);
    types.resize(4 * nreg);
    deq.resize(nreg);
    unpred.clear();
    float prev[4] = {0, 0, 0, 0};
    for (size_t r = 0; r < nreg; r++) {
        const float c[4] = {raw[r].x, raw[r].y, raw[r].z, raw[r].w};
        float d[4];
        for (int t = 0; t < 4; t++) types[4 * r + t] = quantize_coeff(prev[t], c[t], prc[t], rcp[t], &d[t], unpred);
        deq[r] = make_float4(d[0], d[1], d[2], d[3]);
        memcpy(prev, d, sizeof(d));
    }
}

// recover_reg_coeff over the chain (MRP:29-37, 55-94).
bool coefficient_recover(const MetaParams &m, const int32_t *types, size_t nreg, const float *unpred, size_t n_unpred,
                         std::vector<float4> &deq) {
    const double prc[4] = {m.eb_linear, m.eb_linear, m.eb_linear, m.eb_indep};
    float prev[4] = {0, 0, 0, 0};
    size_t u = 0;
    deq.resize(nreg);
    for (size_t r = 0; r < nreg; r++) {
        float d[4];
        for (int t = 0; t < 4; t++) {
            const int tv = types[4 * r + t];
            if (tv == 0) {
                if (u >= n_unpred) return false;
                d[t] = unpred[u++];
            } else {
                d[t] = (float) (prev[t] + 2 * (tv - 32768) * prc[t]);
            }
        }
        deq[r] = make_float4(d[0], d[1], d[2], d[3]);
        memcpy(prev, d, sizeof(d));
    }
    return true;
}

// Huffman of a small host-side int stream (indicator, coefficient codes):
// HuffmanEncoder<int> preprocess_encode + save + encode (HUF:103-227).
bool host_huffman(const int32_t *s, size_t n, std::vector<uint8_t> &out) {
    std::vector<uint32_t> freq(65536, 0), first(65536, UINT32_MAX);
    for (size_t i = 0; i < n; i++) {
        const uint32_t v = (uint32_t) s[i];
        if (v > 65535) return false;
        if (!freq[v]++) first[v] = (uint32_t) i;
    }
    HuffTable t;
    if (build_huffman(freq.data(), first.data(), 65536, t)) return false;
    out.insert(out.end(), t.tree.begin(), t.tree.end());
    encode_huffman(t, s, n, out);
    return true;
}

// The in-block hyperplane table for block edge bs: points of a bs^3 block
// ordered by i+j+k, and where each hyperplane starts.
void build_order(uint32_t bs, std::vector<uint16_t> &tab, std::vector<uint32_t> &diag) {
    tab.clear();
    diag.assign(3 * bs - 1, 0);
    for (uint32_t s = 0; s + 2 < 3 * bs; s++) {
        diag[s] = (uint32_t) tab.size();
        for (uint32_t i = 0; i < bs; i++)
            for (uint32_t j = 0; j < bs; j++)
                if (s >= i + j && s - i - j < bs) tab.push_back((uint16_t) (i | j << 5 | (s - i - j) << 10));
    }
    diag[3 * bs - 2] = (uint32_t) tab.size();
}

int lr_check(faz_gpu_ctx *ctx, const faz_lorenzo_params *p) {
    if (!p) return FAZ_ERR_INVALID;
    for (int i = 0; i < 3; i++)
        if (p->dims[i] == 0) CTX_FAIL(FAZ_ERR_INVALID, "empty dimension");
    if (p->block_size < 1) CTX_FAIL(FAZ_ERR_INVALID, "blockSize must be positive");
    if (p->block_size > (int) kLRMaxBlock)
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "blockSize %d above the %u the Lorenzo kernels stage", p->block_size,
                 kLRMaxBlock);
    if (p->pred_dim < 1 || p->pred_dim > 3) CTX_FAIL(FAZ_ERR_INVALID, "pred_dim must be 1, 2 or 3");
    if (p->radius < 1 || p->radius > 32768) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "radius outside 1..32768");
    if (!(p->eb > 0)) CTX_FAIL(FAZ_ERR_INVALID, "error bound must be positive");
    const uint64_t n = (uint64_t) p->dims[0] * p->dims[1] * p->dims[2];
    if (n >= (1ull << 32)) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 2^32-1 points");
    if ((p->dims[1] - 1) / p->block_size + 1 > 65535)
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 65535 blocks along the second dimension");
    return FAZ_OK;
}

LRGeom lr_geom(const faz_lorenzo_params *p) {
    LRGeom G;
    G.d1 = (uint32_t) p->dims[0], G.d2 = (uint32_t) p->dims[1], G.d3 = (uint32_t) p->dims[2];
    G.s1 = G.d3;
    G.s0 = (uint64_t) G.d2 * G.d3;
    G.bs = (uint32_t) p->block_size;
    G.nx = (G.d1 - 1) / G.bs + 1, G.ny = (G.d2 - 1) / G.bs + 1, G.nz = (G.d3 - 1) / G.bs + 1;
    G.nblocks = (uint64_t) G.nx * G.ny * G.nz;
    return G;
}

LRQuant lr_quant(double eb, int radius) {
    LRQuant Q;
    Q.eb = eb;
    Q.recip = 1.0 / eb;
    Q.ebf = float_floor(eb);
    Q.radius = radius;
    return Q;
}

int ensure_order(faz_gpu_ctx *ctx, uint32_t bs) {
    if (ctx->lr_tab_bs == bs) return FAZ_OK;
    std::vector<uint16_t> tab;
    std::vector<uint32_t> diag;
    build_order(bs, tab, diag);
    HIPCHECK(ctx->lr_tab.ensure(tab.size()));
    HIPCHECK(ctx->lr_diag.ensure(diag.size()));
    HIPCHECK(hipMemcpy(ctx->lr_tab.p, tab.data(), tab.size() * 2, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(ctx->lr_diag.p, diag.data(), diag.size() * 4, hipMemcpyHostToDevice));
    ctx->lr_tab_bs = bs;
    return FAZ_OK;
}

// Host results of one Lorenzo/regression compression besides the device arrays.
struct LRHost {
    MetaParams m;
    uint64_t n = 0, nblocks = 0, nreg = 0, n_unpred = 0;
    std::vector<int32_t> reg_types;
    std::vector<float> reg_unpred;
};

// SZFastFrontend::compress_3d on the GPU: codes (ctx->codes), outliers
// (ctx->unpred), reconstruction (recon), selectors (ctx->lr_ind); histogram
// and first positions of the codes in ctx->freq / ctx->first.
int lr_compress(faz_gpu_ctx *ctx, const float *d_in, float *recon, const faz_lorenzo_params *p, LRHost &h) {
    int rc = lr_check(ctx, p);
    if (rc) return rc;
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const LRGeom G = lr_geom(p);
    const LRQuant Q = lr_quant(p->eb, p->radius);
    h.m = meta_params(p);
    h.n = (uint64_t) G.d1 * G.d2 * G.d3;
    h.nblocks = G.nblocks;
    if ((rc = ensure_order(ctx, G.bs))) return rc;
    const uint32_t nchunks = (uint32_t) ((h.n + kHistChunk - 1) / kHistChunk);
    HIPCHECK(ctx->codes.ensure(h.n));
    HIPCHECK(ctx->outval.ensure(h.n));
    HIPCHECK(ctx->lr_ind.ensure(G.nblocks));
    HIPCHECK(ctx->lr_coef.ensure(G.nblocks));
    HIPCHECK(ctx->lr_flag.ensure(G.nblocks));
    HIPCHECK(ctx->lr_rank.ensure(G.nblocks));
    HIPCHECK(ctx->scalars.ensure(4));
    HIPCHECK(ctx->tiles.ensure(std::max(scan_tiles((uint32_t) G.nblocks), scan_tiles(nchunks))));
    HIPCHECK(ctx->chunk_zeros.ensure(nchunks));
    HIPCHECK(ctx->chunk_base.ensure(nchunks));
    HIPCHECK(ctx->freq.ensure(kHistSlots));
    HIPCHECK(ctx->first.ensure(65536));
    HIPCHECK(hipEventRecord(ctx->ev[0], st));
    LRSel S{p->pred_dim, p->lorenzo != 0, p->lorenzo2 != 0, p->regression != 0, (float) p->eb};
    const uint32_t sgrid = (uint32_t) ((G.nblocks + 255) / 256);
    hipLaunchKernelGGL(lr_stats_kernel, dim3(sgrid), dim3(256), 0, st, G, S, d_in, ctx->lr_ind.p, ctx->lr_coef.p,
                       ctx->lr_flag.p);
    launch_zero_scan(ctx->lr_flag.p, ctx->lr_rank.p, (uint32_t) G.nblocks, ctx->tiles.p, ctx->scalars.p, st);
    HIPCHECK(hipGetLastError());
    uint64_t nreg = 0;
    HIPCHECK(hipMemcpyAsync(&nreg, ctx->scalars.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    h.nreg = nreg;
    std::vector<float4> raw(nreg), deq;
    if (nreg) {
        HIPCHECK(ctx->lr_regc.ensure(nreg));
        hipLaunchKernelGGL(lr_gather_coef_kernel, dim3(sgrid), dim3(256), 0, st, G.nblocks, ctx->lr_flag.p,
                           ctx->lr_rank.p, ctx->lr_coef.p, ctx->lr_regc.p);
        HIPCHECK(hipMemcpyAsync(raw.data(), ctx->lr_regc.p, nreg * sizeof(float4), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        coefficient_chain(h.m, raw.data(), nreg, h.reg_types, h.reg_unpred, deq);
        HIPCHECK(hipMemcpyAsync(ctx->lr_regc.p, deq.data(), nreg * sizeof(float4), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(lr_regression_kernel<0>, dim3((uint32_t) std::min<uint64_t>((nreg + 3) / 4, 16384)),
                           dim3(256), 0, st, G, Q, d_in, recon, ctx->codes.p, ctx->outval.p, ctx->lr_ind.p,
                           ctx->lr_rank.p, ctx->lr_regc.p);
    } else {
        h.reg_types.clear();
        h.reg_unpred.clear();
    }
    launch_lorenzo_all<0>(p->pred_dim, G, Q, d_in, recon, ctx->codes.p, ctx->outval.p, ctx->lr_ind.p,
                          ctx->lr_tab.p, ctx->lr_diag.p, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(ctx->ev[1], st));
    launch_hist(ctx->codes.p, h.n, ctx->freq.p, ctx->first.p, ctx->chunk_zeros.p, p->radius, st);
    launch_zero_scan(ctx->chunk_zeros.p, ctx->chunk_base.p, nchunks, ctx->tiles.p, ctx->scalars.p, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(&h.n_unpred, ctx->scalars.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    HIPCHECK(ctx->unpred.ensure(h.n_unpred));
    launch_unpred_gather(ctx->codes.p, h.n, ctx->outval.p, ctx->chunk_zeros.p, ctx->chunk_base.p, ctx->unpred.p, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(ctx->ev[2], st));
    HIPCHECK(hipStreamSynchronize(st));
    ctx->hist_codes = ctx->codes.p;
    ctx->hist_n = h.n;
    ctx->timings[1] = elapsed_ms(ctx->ev[0], ctx->ev[1]);
    ctx->timings[2] = elapsed_ms(ctx->ev[1], ctx->ev[2]);
    return FAZ_OK;
}

}  // namespace
}  // namespace faz

namespace faz {
namespace {

// SZGeneralCompressor<SZFastFrontend>::compress up to the zstd stage
// (SGC:30-71): frontend save (FF:50-84), then the codes' Huffman tree and bits.
int lr_raw(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_lorenzo_params *p,
           std::vector<uint8_t> &raw) {
    float *recon = d_recon;
    if (!recon) {
        int rc = lr_check(ctx, p);
        if (rc) return rc;
        HIPCHECK(hipSetDevice(ctx->device));
        HIPCHECK(ctx->work_out.ensure((size_t) p->dims[0] * p->dims[1] * p->dims[2]));
        recon = ctx->work_out.p;
    }
    LRHost h;
    int rc = lr_compress(ctx, d_in, recon, p, h);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    HIPCHECK(hipEventRecord(ctx->ev[3], st));
    HIPCHECK(hipMemcpyAsync(ctx->h_freq.data(), ctx->freq.p, 65536 * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(ctx->h_first.data(), ctx->first.p, 65536 * 4, hipMemcpyDeviceToHost, st));
    std::vector<uint8_t> ind8(h.nblocks);
    HIPCHECK(hipMemcpyAsync(ind8.data(), ctx->lr_ind.p, h.nblocks, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    HuffTable tab;
    if (build_huffman(ctx->h_freq.data(), ctx->h_first.data(), 65536, tab))
        CTX_FAIL(FAZ_ERR_UNSUPPORTED, "Huffman tree could not be built (codes longer than 64 bits)");
    HIPCHECK(hipEventRecord(ctx->ev[4], st));
    faz_huff_table ht{tab.offset, tab.n_states, tab.code.data(), tab.len.data()};
    uint64_t maxbits = 0;
    for (auto l: tab.len) maxbits = std::max<uint64_t>(maxbits, l);
    HIPCHECK(ctx->words.ensure((size_t) ((h.n * maxbits + 31) / 32 + 2)));
    size_t nbytes = 0;
    rc = faz_gpu_huffman_encode(ctx, ctx->codes.p, h.n, &ht, reinterpret_cast<uint8_t *>(ctx->words.p),
                                ctx->words.cap * 4, &nbytes);
    if (rc) return rc;
    HIPCHECK(hipEventRecord(ctx->ev[5], st));
    // frontend section
    raw.clear();
    put_meta_params(h.m, raw);
    put(raw, p->eb);                  // precision
    put(raw, (uint64_t) 0);           // meanInfo: use_mean = false, mean = 0
    put(raw, (uint64_t) h.nreg);      // reg_count
    std::vector<int32_t> ind(ind8.begin(), ind8.end());
    if (!host_huffman(ind.data(), ind.size(), raw)) CTX_FAIL(FAZ_ERR_INVALID, "selector Huffman failed");
    if (h.nreg) {  // encode_regression_coefficients (MRP:96-108)
        put(raw, (uint64_t) h.reg_unpred.size());
        const uint8_t *u = reinterpret_cast<const uint8_t *>(h.reg_unpred.data());
        raw.insert(raw.end(), u, u + h.reg_unpred.size() * 4);
        if (!host_huffman(h.reg_types.data(), h.reg_types.size(), raw))
            CTX_FAIL(FAZ_ERR_INVALID, "coefficient Huffman failed");
    }
    // LinearQuantizer::save (Q:185-201)
    raw.push_back(2);
    put(raw, p->eb);
    put(raw, (int32_t) p->radius);
    put(raw, (uint64_t) h.n_unpred);
    size_t at = raw.size();
    raw.resize(at + h.n_unpred * 4 + tab.tree.size() + 8 + nbytes);
    HIPCHECK(hipMemcpyAsync(raw.data() + at, ctx->unpred.p, h.n_unpred * 4, hipMemcpyDeviceToHost, st));
    at += h.n_unpred * 4;
    memcpy(raw.data() + at, tab.tree.data(), tab.tree.size());
    at += tab.tree.size();
    const uint64_t nb64 = nbytes;
    memcpy(raw.data() + at, &nb64, 8);
    at += 8;
    HIPCHECK(hipMemcpyAsync(raw.data() + at, ctx->words.p, nbytes, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    ctx->timings[3] = elapsed_ms(ctx->ev[3], ctx->ev[4]);
    ctx->timings[4] = elapsed_ms(ctx->ev[4], ctx->ev[5]);
    ctx->timings[5] = elapsed_ms(ctx->ev[0], ctx->ev[5]);
    return FAZ_OK;
}

__global__ void lr_flag_kernel(uint64_t nblocks, const uint8_t *__restrict__ ind, uint32_t *__restrict__ flag) {
    const uint64_t b = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nblocks) flag[b] = ind[b] == kSelRegression;
}

// SZGeneralCompressor::decompress after zstd (SGC:115-140): frontend load
// (FF:86-131, MRP:55-94), codes, then decompress_3d (FF:251-318).
int lr_decompress(faz_gpu_ctx *ctx, const uint8_t *raw, size_t n, const size_t *dims, float *d_out) {
    const uint8_t *p = raw, *end = raw + n;
    auto need = [&](size_t k) { return (size_t) (end - p) >= k; };
    if (!need(44 + 8 + 8 + 8)) CTX_FAIL(FAZ_ERR_FORMAT, "Lorenzo frontend header truncated");
    faz_lorenzo_params prm;
    memset(&prm, 0, sizeof(prm));
    for (int i = 0; i < 3; i++) prm.dims[i] = dims[i];
    MetaParams m;
    memcpy(&m.block_size, p, 4);
    memcpy(&m.pred_dim, p + 4, 4);
    m.lo = p[8] != 0, m.lo2 = p[9] != 0, m.reg = p[10] != 0;
    memcpy(&m.eb_indep, p + 20, 4);
    memcpy(&m.eb_linear, p + 24, 4);
    p += 44;
    double precision;
    memcpy(&precision, p, 8);
    p += 16;  // precision, meanInfo
    uint64_t nreg;
    memcpy(&nreg, p, 8);
    p += 8;
    prm.block_size = m.block_size;
    prm.pred_dim = m.pred_dim;
    prm.lorenzo = m.lo, prm.lorenzo2 = m.lo2, prm.regression = m.reg;
    prm.eb = precision;
    prm.radius = 1;
    int rc = lr_check(ctx, &prm);
    if (rc) return rc;
    const LRGeom G = lr_geom(&prm);
    std::vector<int32_t> ind;
    size_t used = decode_huffman(p, end - p, ind, G.nblocks);
    if (!used) CTX_FAIL(FAZ_ERR_FORMAT, "selector Huffman section malformed");
    p += used;
    uint64_t nregc = 0;
    for (auto v: ind) {
        if (v < 0 || v > 2) CTX_FAIL(FAZ_ERR_FORMAT, "selector out of range");
        nregc += v == kSelRegression;
    }
    if (nregc != nreg) CTX_FAIL(FAZ_ERR_FORMAT, "regression block count mismatch");
    std::vector<float4> deq;
    if (nreg) {
        if (!need(8)) CTX_FAIL(FAZ_ERR_FORMAT, "coefficient section truncated");
        uint64_t ru;
        memcpy(&ru, p, 8);
        p += 8;
        if (!need(ru * 4)) CTX_FAIL(FAZ_ERR_FORMAT, "coefficient outliers truncated");
        std::vector<float> reg_unpred(ru);
        memcpy(reg_unpred.data(), p, ru * 4);
        p += ru * 4;
        std::vector<int32_t> types;
        used = decode_huffman(p, end - p, types, 4 * nreg);
        if (!used) CTX_FAIL(FAZ_ERR_FORMAT, "coefficient Huffman section malformed");
        p += used;
        if (!coefficient_recover(m, types.data(), nreg, reg_unpred.data(), ru, deq))
            CTX_FAIL(FAZ_ERR_FORMAT, "coefficient outliers fewer than their codes");
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
    if (radius < 1 || radius > 32768) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "radius outside 1..32768");
    if (!need(U * 4)) CTX_FAIL(FAZ_ERR_FORMAT, "unpredictable list truncated");
    const float *unpred_h = reinterpret_cast<const float *>(p);
    p += U * 4;
    const uint64_t num = (uint64_t) G.d1 * G.d2 * G.d3;
    std::vector<uint8_t> ind8(ind.begin(), ind.end());
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if ((rc = ensure_order(ctx, G.bs))) return rc;
    const uint32_t nchunks = (uint32_t) ((num + kHistChunk - 1) / kHistChunk);
    HIPCHECK(ctx->codes.ensure(num));
    HIPCHECK(ctx->unpred.ensure(U));
    HIPCHECK(ctx->outval.ensure(num));
    HIPCHECK(ctx->chunk_zeros.ensure(nchunks));
    HIPCHECK(ctx->chunk_base.ensure(nchunks));
    HIPCHECK(ctx->freq.ensure(kHistSlots));
    HIPCHECK(ctx->first.ensure(65536));
    HIPCHECK(ctx->scalars.ensure(4));
    HIPCHECK(ctx->lr_ind.ensure(G.nblocks));
    HIPCHECK(ctx->lr_flag.ensure(G.nblocks));
    HIPCHECK(ctx->lr_rank.ensure(G.nblocks));
    // codes decoded on the device (HUF:236-268), then the outliers' ranks
    if ((rc = gpu_huffman_decode<uint16_t>(ctx, p, end - p, num, 65536u, ctx->codes.p, nullptr))) return rc;
    HIPCHECK(ctx->tiles.ensure(std::max(scan_tiles((uint32_t) G.nblocks), scan_tiles(nchunks))));
    if (U) HIPCHECK(hipMemcpyAsync(ctx->unpred.p, unpred_h, U * 4, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(ctx->lr_ind.p, ind8.data(), G.nblocks, hipMemcpyHostToDevice, st));
    launch_hist(ctx->codes.p, num, ctx->freq.p, ctx->first.p, ctx->chunk_zeros.p, radius, st);
    launch_zero_scan(ctx->chunk_zeros.p, ctx->chunk_base.p, nchunks, ctx->tiles.p, ctx->scalars.p, st);
    uint32_t *hs = ctx->hp_hist.data() + 2 * 65536;  // pinned: zero count
    HIPCHECK(hipMemcpyAsync(hs, ctx->scalars.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(spin_wait(st));
    uint64_t zeros;
    memcpy(&zeros, hs, 8);
    if (zeros != U) CTX_FAIL(FAZ_ERR_FORMAT, "unpredictable count does not match the zero codes");
    launch_unpred_scatter(ctx->codes.p, num, ctx->unpred.p, ctx->chunk_zeros.p, ctx->chunk_base.p, ctx->outval.p, st);
    ctx->hist_codes = nullptr;
    const LRQuant Q = lr_quant(eb, radius);
    const uint32_t sgrid = (uint32_t) ((G.nblocks + 255) / 256);
    if (nreg) {
        HIPCHECK(ctx->lr_regc.ensure(nreg));
        HIPCHECK(hipMemcpyAsync(ctx->lr_regc.p, deq.data(), nreg * sizeof(float4), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(lr_flag_kernel, dim3(sgrid), dim3(256), 0, st, G.nblocks, ctx->lr_ind.p, ctx->lr_flag.p);
        launch_zero_scan(ctx->lr_flag.p, ctx->lr_rank.p, (uint32_t) G.nblocks, ctx->tiles.p, ctx->scalars.p, st);
        hipLaunchKernelGGL(lr_regression_kernel<1>, dim3((uint32_t) std::min<uint64_t>((nreg + 3) / 4, 16384)),
                           dim3(256), 0, st, G, Q, nullptr, d_out, ctx->codes.p, ctx->outval.p, ctx->lr_ind.p,
                           ctx->lr_rank.p, ctx->lr_regc.p);
    }
    launch_lorenzo_all<1>(m.pred_dim, G, Q, nullptr, d_out, ctx->codes.p, ctx->outval.p, ctx->lr_ind.p,
                          ctx->lr_tab.p, ctx->lr_diag.p, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    return FAZ_OK;
}

}  // namespace
}  // namespace faz

int faz_lorenzo_raw_vec(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_lorenzo_params *p,
                        std::vector<uint8_t> &raw) {
    if (!ctx || !d_in || !p) return FAZ_ERR_INVALID;
    return lr_raw(ctx, d_in, d_recon, p, raw);
}

extern "C" {

faz_status faz_gpu_lorenzo_compress_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon,
                                        const faz_lorenzo_params *p, faz_lorenzo_out *out) {
    if (!ctx || !d_in || !d_recon || !p || !out) return FAZ_ERR_INVALID;
    LRHost h;
    int rc = lr_compress(ctx, d_in, d_recon, p, h);
    if (rc) return rc;
    out->d_codes = ctx->codes.p;
    out->n_codes = h.n;
    out->d_unpred = ctx->unpred.p;
    out->n_unpred = h.n_unpred;
    out->d_selector = ctx->lr_ind.p;
    out->n_blocks = h.nblocks;
    out->n_regression = h.nreg;
    return FAZ_OK;
}

faz_status faz_gpu_lorenzo_raw_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_lorenzo_params *p,
                                   uint8_t *h_raw, size_t cap, size_t *raw_size) {
    if (!ctx || !d_in || !p || !raw_size) return FAZ_ERR_INVALID;
    std::vector<uint8_t> raw;
    int rc = lr_raw(ctx, d_in, d_recon, p, raw);
    if (rc) return rc;
    *raw_size = raw.size();
    if (!h_raw || cap < raw.size()) CTX_FAIL(FAZ_ERR_CAPACITY, "raw buffer needs %zu bytes", raw.size());
    memcpy(h_raw, raw.data(), raw.size());
    return FAZ_OK;
}

faz_status faz_gpu_lorenzo_decompress_raw(faz_gpu_ctx *ctx, const uint8_t *h_raw, size_t n, const size_t *dims,
                                          float *d_out) {
    if (!ctx || !h_raw || !dims || !d_out) return FAZ_ERR_INVALID;
    return lr_decompress(ctx, h_raw, n, dims, d_out);
}

}  // extern "C"
