// gfx950 kernels of the FAZ compression hot path.
//
// Bit-exactness with the reference x86-64 build: no FMA contraction (this
// file is compiled with -ffp-contract=off and the pragma below), FP32
// denormals preserved, every float/double operation in the reference's order.
#include "faz_internal.h"
#include "faz_device.h"

#include <hip/hip_runtime.h>
#include <climits>

#pragma clang fp contract(off)

namespace faz {

// Where one predicted point sits in the reference's push order and which
// stencil it uses (block_interpolation_1d, SIC:1306-1417): the line holds n
// points, the point is at odd local position i.
struct PointSlot {
    uint64_t seq;   // index in the code stream
    uint32_t i, n;  // local position / line length along the pass axis
    uint32_t exit;  // cubic: position predicted with quad_2
};

template<int N>
__device__ __forceinline__ bool locate(const PassDesc &d, uint32_t t, uint32_t *x, PointSlot &ps) {
    // lattice coordinates, last natural axis fastest
    // (divisions by the stride are shifts, by lattice extents and the block
    // edge multiplications: FastDiv)
    const uint32_t ls = d.ls, ls2 = d.ls + 1;
    uint32_t r = t;
#pragma unroll
    for (int a = N - 1; a >= 0; --a) {
        uint32_t l = r;
        if (a) {
            const uint32_t q = d.dlat[a].div(r);
            l = r - q * d.lat[a];
            r = q;
        }
        uint32_t k = d.kind[a];
        x[a] = (k == kPassAxis) ? d.s * (2 * l + 1) : (k == kG1 ? d.s * l : 2 * d.s * l);
    }
    // owning block along each axis (multi_dimensional_range raster, IT:361-453;
    // a G1/G2 coordinate equal to a block start belongs to the previous block)
    uint32_t kb[N], b[N], nn[N], g1[N], g2[N];
    uint32_t kl = 0;
#pragma unroll
    for (int a = 0; a < N; ++a) {
        uint32_t xa = x[a];
        uint32_t k = (d.kind[a] == kPassAxis) ? d.dbs.div(xa) : (xa ? d.dbs.div(xa - 1) : 0);
        kb[a] = k;
        b[a] = k * d.bs;
        uint32_t e = min(b[a] + d.bs, d.dims[a] - 1);
        nn[a] = ((e - b[a]) >> ls) + 1;
        g1[a] = b[a] ? (e - b[a]) >> ls : (e >> ls) + 1;
        g2[a] = b[a] ? (e - b[a]) >> ls2 : (e >> ls2) + 1;
        kl = kl * d.nb[a] + k;
    }
    // codes of the earlier passes of this block and the line index; every
    // per-axis array is indexed by an unrolled natural axis (the permutation
    // enters through comparisons of d.invp), so the arrays stay in registers.
    // Pass Pp interpolates axis perm[Pp]: its count is n/2 along that axis
    // times g1 (axes earlier in the permutation) or g2 (later) elsewhere.
    uint64_t off = 0;
#pragma unroll
    for (int a = 0; a < N; ++a) {
        if (d.invp[a] >= d.P) continue;
        uint64_t c = nn[a] / 2;
#pragma unroll
        for (int a2 = 0; a2 < N; ++a2) {
            if (a2 == a) continue;
            c *= d.invp[a2] < d.invp[a] ? g1[a2] : g2[a2];
        }
        off += c;
    }
    // line index: the axes other than the pass axis in permutation order,
    // earlier ones outer
    uint32_t lidx[N], lsize[N];
#pragma unroll
    for (int a = 0; a < N; ++a) {
        if (d.invp[a] < d.P) {
            lidx[a] = b[a] ? ((x[a] - b[a]) >> ls) - 1 : x[a] >> ls;
            lsize[a] = g1[a];
        } else {
            lidx[a] = b[a] ? ((x[a] - b[a]) >> ls2) - 1 : x[a] >> ls2;
            lsize[a] = g2[a];
        }
    }
    uint64_t line = 0;
#pragma unroll
    for (int a = 0; a < N; ++a) {
        if (a == d.axis) continue;
        uint64_t mult = 1;
#pragma unroll
        for (int a2 = 0; a2 < N; ++a2)
            if (a2 != d.axis && d.invp[a2] > d.invp[a]) mult *= lsize[a2];
        line += lidx[a] * mult;
    }
    uint32_t i = 0, n = 0;
#pragma unroll
    for (int a = 0; a < N; ++a)
        if (a == d.axis) {
            i = (x[a] - b[a]) >> ls;
            n = nn[a];
        }
    uint32_t rank, exit = 0;
    if (d.linear || n < 5) {
        rank = (i - 1) / 2;
    } else {
        uint32_t m = (n - 3) / 2 - 1;  // interior cubic points
        exit = 3 + 2 * m;
        if (i == 1)
            rank = m;
        else if (i == exit)
            rank = m + 1;
        else if (i == n - 1)
            rank = m + 2;
        else
            rank = (i - 3) / 2;
    }
    ps.seq = d.level_base + d.block_prefix[kl] + off + line * (uint64_t) (n / 2) + rank;
    ps.i = i;
    ps.n = n;
    ps.exit = exit;
    return true;
}

__device__ __forceinline__ float predict(const float *v, uint64_t f, uint64_t S, const PointSlot &ps,
                                         int linear) {
    const uint32_t i = ps.i, n = ps.n;
    if (linear || n < 5) {
        if (i + 1 < n) return i_linear(v[f - S], v[f + S]);
        if (n < 4) return v[f - S];
        return i_linear1(v[f - 3 * S], v[f - S]);
    }
    if (i == 1) return i_quad1(v[f - S], v[f + S], v[f + 3 * S]);
    if (i == ps.exit) return i_quad2(v[f - 3 * S], v[f - S], v[f + S]);
    if (i == n - 1) return i_quad3(v[f - 5 * S], v[f - 3 * S], v[f - S]);
    return i_cubic(v[f - 3 * S], v[f - S], v[f + S], v[f + 3 * S]);
}

// ---------------------------------------------------------------------------
// One (level, pass) of the interpolation predictor: every lattice point of
// the pass is independent, so the launch covers all blocks at once; each point
// lands at its closed-form position in the code stream.
// ---------------------------------------------------------------------------
template<int N>
__global__ void __launch_bounds__(256) interp_pass_kernel(PassDesc d, const float *in, float *recon,
                                                          uint16_t *__restrict__ codes,
                                                          float *__restrict__ outval, uint32_t *__restrict__ err) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.total) return;
    uint32_t x[N];
    PointSlot ps;
    locate<N>(d, t, x, ps);
    uint64_t f = 0;
#pragma unroll
    for (int a = 0; a < N; ++a) f += (uint64_t) x[a] * d.offs[a];
    const uint64_t S = (uint64_t) d.s * d.offs[d.axis];
    float pred = predict(recon, f, S, ps, d.linear);
    float orig = in[f];
    float rec;
    const double eb = d.ebp ? d.ebp->eb : d.eb, recip = d.ebp ? d.ebp->recip : d.recip;
    int code = quantize_point(orig, pred, eb, recip, d.radius, &rec);
    recon[f] = rec;
    if ((unsigned) code > 65535u) atomicOr(err, 1u);
    codes[ps.seq] = (uint16_t) code;
    if (code == 0) outval[ps.seq] = orig;
}

// Same walk in recover mode (PB_recover, SIC:110-196 with Q:160-180).
template<int N>
__global__ void __launch_bounds__(256) recover_pass_kernel(PassDesc d, float *__restrict__ data,
                                                           const uint16_t *__restrict__ codes,
                                                           const float *__restrict__ unpred_at) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.total) return;
    uint32_t x[N];
    PointSlot ps;
    locate<N>(d, t, x, ps);
    uint64_t f = 0;
#pragma unroll
    for (int a = 0; a < N; ++a) f += (uint64_t) x[a] * d.offs[a];
    const uint64_t S = (uint64_t) d.s * d.offs[d.axis];
    float pred = predict(data, f, S, ps, d.linear);
    int code = codes[ps.seq];
    const float v = code ? recover_point(pred, code, d.radius, d.eb) : unpred_at[ps.seq];
    if (N > 1 && d.P == N - 1 && S == 1) {
        // A last pass along the contiguous axis at stride 1 covers whole rows:
        // each point also stores its left neighbour (the value it read for
        // its stencil, so concurrent readers see the same bits) and every line
        // of the row is written in full instead of every other word.
        const float left = data[f - 1];
        __builtin_nontemporal_store(left, data + f - 1);
        __builtin_nontemporal_store(v, data + f);
    } else {
        data[f] = v;
    }
}

void launch_interp_pass(const PassDesc &d, const float *d_in, float *d_recon, uint16_t *d_codes, float *d_outval,
                        double *, uint32_t *d_err, hipStream_t st) {
    if (d.total == 0) return;
    dim3 grid((unsigned) ((d.total + 255) / 256)), block(256);
    switch (d.N) {
        case 1: hipLaunchKernelGGL(interp_pass_kernel<1>, grid, block, 0, st, d, d_in, d_recon, d_codes, d_outval, d_err); break;
        case 2: hipLaunchKernelGGL(interp_pass_kernel<2>, grid, block, 0, st, d, d_in, d_recon, d_codes, d_outval, d_err); break;
        case 3: hipLaunchKernelGGL(interp_pass_kernel<3>, grid, block, 0, st, d, d_in, d_recon, d_codes, d_outval, d_err); break;
        default: hipLaunchKernelGGL(interp_pass_kernel<4>, grid, block, 0, st, d, d_in, d_recon, d_codes, d_outval, d_err); break;
    }
}

void launch_recover_pass(const PassDesc &d, float *d_data, const uint16_t *d_codes, const float *d_unpred_at,
                         hipStream_t st) {
    if (d.total == 0) return;
    dim3 grid((unsigned) ((d.total + 255) / 256)), block(256);
    switch (d.N) {
        case 1: hipLaunchKernelGGL(recover_pass_kernel<1>, grid, block, 0, st, d, d_data, d_codes, d_unpred_at); break;
        case 2: hipLaunchKernelGGL(recover_pass_kernel<2>, grid, block, 0, st, d, d_data, d_codes, d_unpred_at); break;
        case 3: hipLaunchKernelGGL(recover_pass_kernel<3>, grid, block, 0, st, d, d_data, d_codes, d_unpred_at); break;
        default: hipLaunchKernelGGL(recover_pass_kernel<4>, grid, block, 0, st, d, d_data, d_codes, d_unpred_at); break;
    }
}

// ---------------------------------------------------------------------------
// Anchors (build_grid, SIC:854-900) and the un-anchored first point (SIC:372-374)
// ---------------------------------------------------------------------------
__global__ void anchor_kernel(const float *__restrict__ in, float *__restrict__ recon, uint16_t *__restrict__ codes,
                              float *__restrict__ outval, int N, uint32_t d0, uint32_t d1, uint32_t d2, uint64_t o0,
                              uint64_t o1, uint32_t step, uint32_t count) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    uint64_t f;
    if (N == 2) {
        uint32_t ny = (d1 - 1) / step + 1;
        f = (uint64_t) (t / ny) * step * o0 + (uint64_t) (t % ny) * step;
    } else {
        uint32_t ny = (d1 - 1) / step + 1, nz = (d2 - 1) / step + 1;
        uint32_t z = t % nz, y = (t / nz) % ny, xx = t / (nz * ny);
        f = (uint64_t) xx * step * o0 + (uint64_t) y * step * o1 + (uint64_t) z * step;
    }
    float v = in[f];
    recon[f] = v;
    codes[t] = 0;
    outval[t] = v;
}

void launch_anchor_grid(const float *d_in, float *d_recon, uint16_t *d_codes, float *d_outval, int N,
                        const uint32_t *dims, const uint64_t *offs, uint32_t step, uint64_t count, hipStream_t st) {
    if (!count) return;
    hipLaunchKernelGGL(anchor_kernel, dim3((unsigned) ((count + 255) / 256)), dim3(256), 0, st, d_in, d_recon, d_codes,
                       d_outval, N, dims[0], dims[1], N > 2 ? dims[2] : 1, offs[0], N > 2 ? offs[1] : 1, step,
                       (uint32_t) count);
}

__global__ void anchor_scatter_kernel(const float *__restrict__ vals, float *__restrict__ out, int N, uint32_t d1,
                                      uint32_t d2, uint64_t o0, uint64_t o1, uint32_t step, uint32_t count) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    uint64_t f;
    if (N == 2) {
        uint32_t ny = (d1 - 1) / step + 1;
        f = (uint64_t) (t / ny) * step * o0 + (uint64_t) (t % ny) * step;
    } else {
        uint32_t ny = (d1 - 1) / step + 1, nz = (d2 - 1) / step + 1;
        uint32_t z = t % nz, y = (t / nz) % ny, xx = t / (nz * ny);
        f = (uint64_t) xx * step * o0 + (uint64_t) y * step * o1 + (uint64_t) z * step;
    }
    out[f] = vals[t];
}

// recover_grid (SIC:902-923): anchors take the leading unpredictable values
void launch_anchor_scatter(const float *d_vals, float *d_out, int N, const uint32_t *dims, const uint64_t *offs,
                           uint32_t step, uint64_t count, hipStream_t st) {
    if (!count) return;
    hipLaunchKernelGGL(anchor_scatter_kernel, dim3((unsigned) ((count + 255) / 256)), dim3(256), 0, st, d_vals, d_out,
                       N, dims[1], N > 2 ? dims[2] : 1, offs[0], N > 2 ? offs[1] : 1, step, (uint32_t) count);
}

// The decoder's first point without anchors (SIC:372-374 in recover mode):
// predicted as 0, or the first unpredictable value.
__global__ void first_value_kernel(const uint16_t *codes, const float *unpred, uint64_t n_unpred, int radius,
                                   double eb, float *out) {
    const int code = codes[0];
    out[0] = code ? recover_point(0.0f, code, radius, eb) : (n_unpred ? unpred[0] : 0.0f);
}

void launch_first_value(const uint16_t *d_codes, const float *d_unpred, uint64_t n_unpred, int radius, double eb,
                        float *d_out, hipStream_t st) {
    hipLaunchKernelGGL(first_value_kernel, dim3(1), dim3(1), 0, st, d_codes, d_unpred, n_unpred, radius, eb, d_out);
}

__global__ void first_point_kernel(const float *in, float *recon, uint16_t *codes, float *outval, double eb,
                                   double recip, int radius, uint32_t *err, const LevelEB *ebp) {
    if (ebp) {
        eb = ebp->eb;
        recip = ebp->recip;
    }
    float rec;
    int code = quantize_point(in[0], 0.0f, eb, recip, radius, &rec);
    recon[0] = rec;
    if ((unsigned) code > 65535u) atomicOr(err, 1u);
    codes[0] = (uint16_t) code;
    if (code == 0) outval[0] = in[0];
}

void launch_first_point(const float *d_in, float *d_recon, uint16_t *d_codes, float *d_outval, double eb, double recip,
                        int radius, uint32_t *d_err, hipStream_t st, const LevelEB *d_eb) {
    hipLaunchKernelGGL(first_point_kernel, dim3(1), dim3(1), 0, st, d_in, d_recon, d_codes, d_outval, eb, recip, radius,
                       d_err, d_eb);
}

// ---------------------------------------------------------------------------
// data_range: min/max as order-preserving uint32 keys
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t f2key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Value range: each workgroup writes its (min, max) keys to the partial slots
// after the result pair, and one small workgroup folds them.  (Same-address
// atomics from every workgroup serialised in the L2 and cost a quarter of
// the pass; kRangeGrid workgroups of four float4 loads per lane in flight
// read 512 MiB at about 6 TB/s.)
// When clr.freq is set, the launch also clears the histogram arrays and
// the error flag for the compression that follows (a launch of its own cost
// more than the stores).
struct RangeClear {
    uint32_t *freq, *first, *err;
};

__global__ void __launch_bounds__(256) range_kernel(const float *__restrict__ in, size_t n,
                                                   uint32_t *__restrict__ part, RangeClear clr) {
    if (clr.freq) {
        const uint32_t nt = gridDim.x * blockDim.x;
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t) kHistSlots; i += nt) {
            clr.freq[i] = 0u;
            if (i < 65536u) clr.first[i] = 0xFFFFFFFFu;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) *clr.err = 0u;
    }
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    const size_t stride = (size_t) gridDim.x * blockDim.x;
    const size_t n4 = ((uintptr_t) in & 15) ? 0 : n / 4;
    const float4 *in4 = reinterpret_cast<const float4 *>(in);
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int U = 4;  // 8 loads per lane, or 1024-2048 workgroups: slower on 512 MiB
    for (; i + (U - 1) * stride < n4; i += U * stride) {
        float4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = in4[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; u++) {
            uint32_t a = f2key(v[u].x), b = f2key(v[u].y), c = f2key(v[u].z), e = f2key(v[u].w);
            lo = min(lo, min(min(a, b), min(c, e)));
            hi = max(hi, max(max(a, b), max(c, e)));
        }
    }
    for (; i < n4; i += stride) {
        float4 v = in4[i];
        uint32_t a = f2key(v.x), b = f2key(v.y), c = f2key(v.z), e = f2key(v.w);
        lo = min(lo, min(min(a, b), min(c, e)));
        hi = max(hi, max(max(a, b), max(c, e)));
    }
    for (size_t j = n4 * 4 + (size_t) blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        uint32_t a = f2key(in[j]);
        lo = min(lo, a);
        hi = max(hi, a);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, (uint32_t) __shfl_xor((int) lo, o));
        hi = max(hi, (uint32_t) __shfl_xor((int) hi, o));
    }
    __shared__ uint32_t slo[4], shi[4];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) {
        slo[w] = lo;
        shi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            lo = min(lo, slo[k]);
            hi = max(hi, shi[k]);
        }
        part[2 * blockIdx.x] = lo;
        part[2 * blockIdx.x + 1] = hi;
    }
}

__device__ void level_bounds(const uint32_t *minmax, double rel, const LevelScale &sc, LevelEB *out, int i);

// With lvl set, the same workgroup goes on to the level bounds
// (level_bounds_kernel's arithmetic, one thread per bound).
__global__ void __launch_bounds__(256) range_fold_kernel(const uint32_t *__restrict__ part, uint32_t np,
                                                        uint32_t *__restrict__ minmax, double rel, LevelScale sc,
                                                        LevelEB *__restrict__ lvl) {
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        lo = min(lo, part[2 * i]);
        hi = max(hi, part[2 * i + 1]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, (uint32_t) __shfl_xor((int) lo, o));
        hi = max(hi, (uint32_t) __shfl_xor((int) hi, o));
    }
    __shared__ uint32_t slo[4], shi[4];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) {
        slo[w] = lo;
        shi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            lo = min(lo, slo[k]);
            hi = max(hi, shi[k]);
        }
        minmax[0] = lo;
        minmax[1] = hi;
        slo[0] = lo;
        shi[0] = hi;
    }
    if (!lvl) return;
    __syncthreads();
    const uint32_t folded[2] = {slo[0], shi[0]};
    if ((int) threadIdx.x <= sc.n) level_bounds(folded, rel, sc, lvl, (int) threadIdx.x - 1);
}

// The range's keys back to floats (the inverse of f2key).
__device__ __forceinline__ float key2f(uint32_t key) {
    return __uint_as_float((key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key);
}

// Level bounds from the folded range, in the host's arithmetic: the base
// bound rel * (double) (float) (max - min) (QoZ::data_range and the REL
// conversion of SZ_compress), each level's bound from it (SIC:384-422), the
// reciprocal, the largest float not above the bound, and the bounded
// quantizer's test 4.5 max|x| / eb + 8 < 2^30.
// i = -1: the base bound.
__device__ void level_bounds(const uint32_t *minmax, double rel, const LevelScale &sc, LevelEB *out, int i) {
    const float mn = key2f(minmax[0]), mx = key2f(minmax[1]);
    const double base = rel * (double) (float) (mx - mn);
    const double amn = fabs((double) mn), amx = fabs((double) mx);
    const double absmax = amn < amx ? amx : amn;  // std::max, NaN included
    const double eb = i < 0 ? base : (sc.div[i] ? base / sc.f[i] : base * sc.f[i]);
    const double recip = 1.0 / eb;
    float ebf = (float) eb;
    if ((double) ebf > eb) ebf = __uint_as_float(__float_as_uint(ebf) - 1);  // next float toward 0 (ebf > 0)
    const double v = 4.5 * absmax * recip + 8.0;
    LevelEB r;
    r.eb = eb;
    r.recip = recip;
    r.ebf = ebf;
    r.bounded = isfinite(absmax) && absmax < 1e37 && v < 1073741824.0 ? 1 : 0;
    out[i + 1] = r;
}

__global__ void level_bounds_kernel(const uint32_t *__restrict__ minmax, double rel, LevelScale sc,
                                    LevelEB *__restrict__ out) {
    const int i = (int) threadIdx.x - 1;
    if (i < sc.n) level_bounds(minmax, rel, sc, out, i);
}

void launch_level_bounds(const uint32_t *d_minmax, double rel, const LevelScale &sc, LevelEB *d_out,
                         hipStream_t st) {
    hipLaunchKernelGGL(level_bounds_kernel, dim3(1), dim3(kMaxLevels + 1), 0, st, d_minmax, rel, sc, d_out);
}

void launch_range(const float *d_in, size_t n, uint32_t *d_minmax, hipStream_t st) {
    const size_t want = (n / 4 + 255) / 256;
    const unsigned grid = (unsigned) std::min<size_t>(std::max<size_t>(want, 1), kRangeGrid);
    hipLaunchKernelGGL(range_kernel, dim3(grid), dim3(256), 0, st, d_in, n, d_minmax + 2, RangeClear{});
    hipLaunchKernelGGL(range_fold_kernel, dim3(1), dim3(256), 0, st, d_minmax + 2, grid, d_minmax, 0.0, LevelScale{},
                       (LevelEB *) nullptr);
}

void launch_range_rel(const float *d_in, size_t n, uint32_t *d_minmax, uint32_t *d_freq, uint32_t *d_first,
                      uint32_t *d_err, double rel, const LevelScale &sc, LevelEB *d_lvl, hipStream_t st) {
    const size_t want = (n / 4 + 255) / 256;
    const unsigned grid = (unsigned) std::min<size_t>(std::max<size_t>(want, 1), kRangeGrid);
    hipLaunchKernelGGL(range_kernel, dim3(grid), dim3(256), 0, st, d_in, n, d_minmax + 2,
                       RangeClear{d_freq, d_first, d_err});
    hipLaunchKernelGGL(range_fold_kernel, dim3(1), dim3(256), 0, st, d_minmax + 2, grid, d_minmax, rel, sc, d_lvl);
}

// ---------------------------------------------------------------------------
// Histogram + first occurrence + per-chunk outlier counts
//
// Codes cluster tightly around the quantizer radius, so the 64 symbols
// nearest it are counted in lane-private LDS counters (no two lanes of a
// wave add to one word); the rest of a 4096-symbol window goes to LDS
// atomics, anything beyond it to global atomics.  First positions: LDS
// atomicMin guarded by a read, so each symbol costs at most a few atomics per
// workgroup.
// ---------------------------------------------------------------------------
constexpr int kWin = kHistWin;  // LDS bins around the quantizer radius
constexpr int kLaneWin = 64;  // symbols [hot0, hot0 + 64) counted per lane

// The thread's 8 groups of 8 consecutive codes of chunk c (16-byte loads
// for a full chunk of aligned codes, element loads past the end otherwise).
constexpr int kHistGroups = kHistPerThread / 8;
__device__ __forceinline__ void hist_load(const uint16_t *__restrict__ codes, uint32_t c, uint32_t n, int aligned,
                                          uint4 *q) {
    const uint32_t base = c * kHistChunk;
    const uint32_t end = min(base + kHistChunk, n);
    const bool full = aligned && end - base == kHistChunk;
#pragma unroll
    for (int g = 0; g < kHistGroups; g++) {
        const uint32_t j = base + (threadIdx.x + g * kHistThreads) * 8;
        if (full) {
            q[g] = *reinterpret_cast<const uint4 *>(codes + j);
        } else {
            uint32_t t[8];
#pragma unroll
            for (int k = 0; k < 8; k++) t[k] = j + k < end ? codes[j + k] : 0;
            q[g] = make_uint4(t[0] | t[1] << 16, t[2] | t[3] << 16, t[4] | t[5] << 16, t[6] | t[7] << 16);
        }
    }
}

// Zero codes per chunk only (the decoder's outlier ranks need no histogram):
// one workgroup per chunk, the histogram's coalesced loads.
__global__ void __launch_bounds__(kHistThreads) zero_count_u16_kernel(const uint16_t *__restrict__ codes, uint32_t n,
                                                                     uint32_t *__restrict__ chunk_zeros, int aligned) {
    __shared__ uint32_t s_wave[kHistThreads / 64];
    const uint32_t c = blockIdx.x;
    const uint32_t base = c * kHistChunk;
    const uint32_t end = min(base + kHistChunk, n);
    uint4 q[kHistGroups];
    hist_load(codes, c, n, aligned, q);
    uint32_t z = 0;
#pragma unroll
    for (int g = 0; g < kHistGroups; g++) {
        const uint32_t j = base + (threadIdx.x + g * kHistThreads) * 8;
        const uint32_t w[4] = {q[g].x, q[g].y, q[g].z, q[g].w};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t v = k & 1 ? w[k >> 1] >> 16 : w[k >> 1] & 0xFFFFu;
            z += (v == 0 && j + k < end) ? 1u : 0u;  // past the end: padding, not a code
        }
    }
    for (int o = 32; o > 0; o >>= 1) z += __shfl_xor(z, o);
    if (threadIdx.x % 64 == 0) s_wave[threadIdx.x / 64] = z;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (uint32_t w = 0; w < kHistThreads / 64; w++) t += s_wave[w];
        chunk_zeros[c] = t;
    }
}

// Persistent: a workgroup keeps its LDS window and lane counters across the
// chunks it takes (in increasing order), so they are cleared and flushed to
// the global histogram once per workgroup, not once per chunk.  The next
// chunk's codes are in flight while the current one is counted, and the
// per-chunk outlier counter alternates between two slots, so a chunk costs
// one barrier.
//
// A group of eight codes that all lie in [hot0, hot0 + kLaneWin) takes the
// fast path: one LDS add per code into the lane's own counter of the symbol
// (counters laid out [symbol][lane], so a wave's adds hit 64 distinct banks
// whatever the symbols) and a 64-bit mask of the symbols seen; only symbols
// new to the thread go on to the first-position update.  Any other group
// goes code by code.
__global__ void __launch_bounds__(kHistThreads) hist_kernel(const uint16_t *__restrict__ codes, uint32_t n,
                                                           uint32_t *__restrict__ freq, uint32_t *__restrict__ first,
                                                           uint32_t *__restrict__ chunk_zeros, int lo_,
                                                           int hot0_, int aligned, uint32_t *__restrict__ outside) {
    __shared__ uint32_t s_cnt[kWin];
    __shared__ uint32_t s_first[kWin];  // first global position per window symbol
    __shared__ uint32_t s_lane[kLaneWin * 64];
    __shared__ uint32_t s_zero[2], s_zero_first;
    const uint32_t lo = (uint32_t) lo_, hot0 = (uint32_t) hot0_;
    for (int i = threadIdx.x; i < kWin; i += kHistThreads) {
        s_cnt[i] = 0;
        s_first[i] = 0xFFFFFFFFu;
    }
    for (int i = threadIdx.x; i < kLaneWin * 64; i += kHistThreads) s_lane[i] = 0;
    if (threadIdx.x == 0) {
        s_zero_first = 0xFFFFFFFFu;
        s_zero[0] = s_zero[1] = 0;
    }
    __syncthreads();
    uint64_t seen = 0;  // lane-window symbols whose first position this thread already offered
    uint32_t zeros_total = 0, slot = 0;
    const uint32_t lane = threadIdx.x % 64;
    const uint32_t nchunks = (n + kHistChunk - 1) / kHistChunk;
    constexpr int kGroups = kHistGroups;
    uint4 q[kGroups], qn[kGroups];
    // one code outside the fast path
    auto count_one = [&](uint32_t v, uint32_t gi, uint32_t *sz) {
        const uint32_t d = v - hot0;
        if (d < (uint32_t) kLaneWin) {
            atomicAdd(&s_lane[d * 64 + lane], 1u);
            if (!((seen >> d) & 1u)) {
                seen |= 1ull << d;
                if (s_first[v - lo] > gi) atomicMin(&s_first[v - lo], gi);
            }
        } else if (v == 0) {
            atomicAdd(sz, 1u);
            if (s_zero_first > gi) atomicMin(&s_zero_first, gi);
        } else {
            const uint32_t w = v - lo;
            if (w < (uint32_t) kWin) {
                atomicAdd(&s_cnt[w], 1u);
                if (s_first[w] > gi) atomicMin(&s_first[w], gi);
            } else {
                atomicAdd(&freq[v], 1u);
                atomicMin(&first[v], gi);
                *outside = 1u;
            }
        }
    };
    if (blockIdx.x < nchunks) hist_load(codes, blockIdx.x, n, aligned, q);
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x, slot ^= 1u) {
        const uint32_t base = c * kHistChunk;
        const uint32_t end = min(base + kHistChunk, n);
        uint32_t *sz = &s_zero[slot];
        if (c + gridDim.x < nchunks) hist_load(codes, c + gridDim.x, n, aligned, qn);
#pragma unroll
        for (int g = 0; g < kGroups; g++) {
            const uint32_t j = base + (threadIdx.x + g * kHistThreads) * 8;
            const uint32_t v[8] = {q[g].x & 0xFFFFu, q[g].x >> 16, q[g].y & 0xFFFFu, q[g].y >> 16,
                                   q[g].z & 0xFFFFu, q[g].z >> 16, q[g].w & 0xFFFFu, q[g].w >> 16};
            uint32_t d[8], any = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                d[k] = v[k] - hot0;
                any |= d[k];
            }
            if (j + 8 <= end && any < (uint32_t) kLaneWin) {
                uint64_t m = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    atomicAdd(&s_lane[d[k] * 64 + lane], 1u);
                    m |= 1ull << d[k];
                }
                if (m & ~seen) {
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        if (!((seen >> d[k]) & 1u)) {
                            seen |= 1ull << d[k];
                            const uint32_t gi = j + k;
                            if (s_first[v[k] - lo] > gi) atomicMin(&s_first[v[k] - lo], gi);
                        }
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < 8; k++)
                    if (j + k < end) count_one(v[k], j + k, sz);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // this slot is next used two chunks on, after one more barrier
            chunk_zeros[c] = *sz;
            zeros_total += *sz;
            *sz = 0;
        }
#pragma unroll
        for (int g = 0; g < kGroups; g++) q[g] = qn[g];
    }
    __syncthreads();
    // lane counters into the window: four threads per symbol, 16 lanes each
    {
        const uint32_t d = threadIdx.x / 4, part = threadIdx.x % 4;
        static_assert(kHistThreads == 4 * kLaneWin, "four threads per lane-window symbol");
        uint32_t t = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) t += s_lane[d * 64 + part * 16 + i];
        if (t) atomicAdd(&s_cnt[hot0 + d - lo], t);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kWin; i += kHistThreads) {
        const uint32_t cnt = s_cnt[i];
        if (cnt) {
            atomicAdd(&freq[lo + i], cnt);
            atomicMin(&first[lo + i], s_first[i]);
        }
    }
    if (threadIdx.x == 0 && zeros_total) {
        atomicAdd(&freq[0], zeros_total);
        atomicMin(&first[0], s_zero_first);
    }
}

__global__ void fill_u32_kernel(uint32_t *d, size_t count, uint32_t value) {
    size_t stride = (size_t) gridDim.x * blockDim.x;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) d[i] = value;
}

void launch_fill_u32(uint32_t *d, size_t count, uint32_t value, hipStream_t st) {
    if (!count) return;
    size_t want = (count + 255) / 256;
    unsigned grid = (unsigned) (want < 8192 ? want : 8192);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid), dim3(256), 0, st, d, count, value);
}

void launch_hist_clear(uint32_t *d_freq, uint32_t *d_first, hipStream_t st) {
    launch_fill_u32(d_freq, kHistSlots, 0u, st);
    launch_fill_u32(d_first, 65536, 0xFFFFFFFFu, st);
}

void launch_hist(const uint16_t *d_codes, size_t n, uint32_t *d_freq, uint32_t *d_first, uint32_t *d_chunk_zeros,
                 int radius, hipStream_t st, bool cleared) {
    if (!cleared) launch_hist_clear(d_freq, d_first, st);
    uint32_t nchunks = (uint32_t) ((n + kHistChunk - 1) / kHistChunk);
    // window [lo, lo + kWin) inside 1..65535, hot symbols inside the window
    const int lo = hist_window_lo(radius);
    int hot0 = radius - kLaneWin / 2;
    if (hot0 < lo) hot0 = lo;
    if (hot0 + kLaneWin > lo + kWin) hot0 = lo + kWin - kLaneWin;
    // persistent workgroups: as many as are resident at once (one round);
    // measured once per process (thread-safe static initialisation)
    static const uint32_t resident = [] {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, hist_kernel, kHistThreads, 0) == hipSuccess &&
            cus > 0 && per > 0)
            return (uint32_t) (cus * per);
        return 768u;
    }();
    const uint32_t grid = nchunks < resident ? nchunks : resident;
    if (nchunks)
        hipLaunchKernelGGL(hist_kernel, dim3(grid), dim3(kHistThreads), 0, st, d_codes, (uint32_t) n, d_freq,
                           d_first, d_chunk_zeros, lo, hot0, (int) (((uintptr_t) d_codes & 15) == 0), d_freq + 65536);
}

int hist_window_lo(int radius) {
    int lo = radius - kWin / 2;
    if (lo < 1) lo = 1;
    if (lo + kWin > 65536) lo = 65536 - kWin;
    return lo;
}

// pack may be host memory (coherent pinned): the block goes straight to the
// host, and with eb set the base bound follows it at word kHistPackWords.
__global__ void __launch_bounds__(256) hist_pack_kernel(const uint32_t *__restrict__ freq,
                                                        const uint32_t *__restrict__ first, int lo,
                                                        const uint32_t *__restrict__ scalar0,
                                                        const uint32_t *__restrict__ err, uint32_t *__restrict__ pack,
                                                        const uint32_t *__restrict__ eb) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (eb && i - (uint32_t) kHistPackWords < 2u) {
        pack[i] = eb[i - kHistPackWords];
        return;
    }
    uint32_t v;
    if (i < (uint32_t) kWin) v = freq[lo + i];
    else if (i < 2u * kWin) v = first[lo + i - kWin];
    else if (i == kHistPackFreq0) v = freq[0];
    else if (i == kHistPackFreq0 + 1) v = first[0];
    else if (i == kHistPackFreq0 + 2) v = freq[65536];
    else if (i == kHistPackFreq0 + 3) v = err[0];
    else if (i < (uint32_t) kHistPackWords) v = scalar0[i - (kHistPackFreq0 + 4)];
    else return;
    pack[i] = v;
}

void launch_zero_count_u16(const uint16_t *d_codes, size_t n, uint32_t *d_chunk_zeros, hipStream_t st) {
    const uint32_t nchunks = (uint32_t) ((n + kHistChunk - 1) / kHistChunk);
    if (nchunks)
        hipLaunchKernelGGL(zero_count_u16_kernel, dim3(nchunks), dim3(kHistThreads), 0, st, d_codes, (uint32_t) n,
                           d_chunk_zeros, (int) (((uintptr_t) d_codes & 15) == 0));
}

void launch_hist_pack(const uint32_t *d_freq, const uint32_t *d_first, int radius, const uint64_t *d_scalar0,
                      const uint32_t *d_err, uint32_t *d_pack, hipStream_t st, const double *d_eb) {
    hipLaunchKernelGGL(hist_pack_kernel, dim3((kHistPackWords + 2 + 255) / 256), dim3(256), 0, st, d_freq, d_first,
                       hist_window_lo(radius), reinterpret_cast<const uint32_t *>(d_scalar0), d_err, d_pack,
                       reinterpret_cast<const uint32_t *>(d_eb));
}

// The encode's bit total and the error flag into (host) memory.
__global__ void report_kernel(const uint64_t *__restrict__ bits, const uint32_t *__restrict__ err,
                              uint32_t *__restrict__ out) {
    if (threadIdx.x == 0) {
        const uint64_t b = *bits;
        out[0] = (uint32_t) b;
        out[1] = (uint32_t) (b >> 32);
        out[2] = *err;
    }
}

void launch_report(const uint64_t *d_bits, const uint32_t *d_err, uint32_t *out, hipStream_t st) {
    hipLaunchKernelGGL(report_kernel, dim3(1), dim3(64), 0, st, d_bits, d_err, out);
}

// ---------------------------------------------------------------------------
// Exclusive scans over chunk totals: tiles of kScanTile entries are reduced,
// then every tile scans itself and adds the sum of the tiles before it.
// With words set (the encoder's bit offsets) the apply pass also zeroes the
// output word each offset falls in and the word of the last bit: the only
// words the packer ORs into (every other word is stored whole by one chunk).
// ---------------------------------------------------------------------------
constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanPer = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanPer;

template<typename TIn>
__global__ void __launch_bounds__(kScanThreads) tile_reduce_kernel(const TIn *__restrict__ in, uint32_t count,
                                                                  uint64_t *__restrict__ tile_sum) {
    const uint32_t i0 = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    uint64_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; k++)
        if (i0 + k < count) sum += (uint64_t) in[i0 + k];
    for (int o = 32; o > 0; o >>= 1) sum += (uint64_t) __shfl_xor((long long) sum, o);
    __shared__ uint64_t s[kScanThreads / 64];
    if (threadIdx.x % 64 == 0) s[threadIdx.x / 64] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (uint32_t w = 0; w < kScanThreads / 64; w++) t += s[w];
        tile_sum[blockIdx.x] = t;
    }
}

template<typename TIn>
__global__ void __launch_bounds__(kScanThreads) scan_apply_kernel(const TIn *__restrict__ in, uint32_t count,
                                                                 const uint64_t *__restrict__ tile_sum,
                                                                 uint32_t ntiles, uint64_t *__restrict__ out,
                                                                 uint64_t *total, uint32_t *__restrict__ words) {
    __shared__ uint64_t s_wave[kScanThreads / 64];
    __shared__ uint64_t s_tile_base;
    // sum of the tiles before this one
    uint64_t tb = 0;
    for (uint32_t k = threadIdx.x; k < blockIdx.x; k += kScanThreads) tb += tile_sum[k];
    for (int o = 32; o > 0; o >>= 1) tb += (uint64_t) __shfl_xor((long long) tb, o);
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    if (lane == 0) s_wave[wave] = tb;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (uint32_t w = 0; w < kScanThreads / 64; w++) t += s_wave[w];
        s_tile_base = t;
    }
    __syncthreads();
    const uint32_t i0 = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    uint64_t v[kScanPer];
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; k++) {
        v[k] = i0 + k < count ? (uint64_t) in[i0 + k] : 0;
        mine += v[k];
    }
    uint64_t incl = mine;
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t t = (uint64_t) __shfl_up((long long) incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint64_t run = s_tile_base + incl - mine;
    for (int w = 0; w < wave; w++) run += s_wave[w];
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; k++) {
        if (i0 + k < count) {
            out[i0 + k] = run;
            if (words) words[run >> 5] = 0u;
        }
        run += v[k];
    }
    if (blockIdx.x == ntiles - 1 && threadIdx.x == kScanThreads - 1) {
        if (total) *total = run;
        if (words && run) words[(run - 1) >> 5] = 0u;
    }
}

template<typename TIn>
static void scan_two_level(const TIn *d_in, uint64_t *d_out, uint32_t count, uint64_t *d_tiles, uint64_t *d_total,
                           hipStream_t st, uint32_t *d_words = nullptr) {
    uint32_t ntiles = (count + kScanTile - 1) / kScanTile;
    if (ntiles == 0) return;
    hipLaunchKernelGGL(tile_reduce_kernel<TIn>, dim3(ntiles), dim3(kScanThreads), 0, st, d_in, count, d_tiles);
    hipLaunchKernelGGL(scan_apply_kernel<TIn>, dim3(ntiles), dim3(kScanThreads), 0, st, d_in, count, d_tiles, ntiles,
                       d_out, d_total, d_words);
}

void launch_zero_scan(const uint32_t *d_chunk_zeros, uint64_t *d_chunk_base, uint32_t nchunks, uint64_t *d_tiles,
                      uint64_t *d_total, hipStream_t st) {
    scan_two_level<uint32_t>(d_chunk_zeros, d_chunk_base, nchunks, d_tiles, d_total, st);
}

void launch_encode_offsets(const uint64_t *d_bits, uint64_t *d_off, uint32_t nchunks, uint64_t *d_tiles,
                           uint64_t *d_total, uint32_t *d_words, hipStream_t st) {
    scan_two_level<uint64_t>(d_bits, d_off, nchunks, d_tiles, d_total, st, d_words);
}

void launch_scan_u64(const uint64_t *d_in, uint64_t *d_out, uint32_t count, uint64_t *d_tiles, uint64_t *d_total,
                     hipStream_t st) {
    scan_two_level<uint64_t>(d_in, d_out, count, d_tiles, d_total, st);
}

uint32_t scan_tiles(uint32_t count) { return (count + kScanTile - 1) / kScanTile; }

// ---------------------------------------------------------------------------
// Outlier compaction: the unpredictable list is the values at code-0
// positions in code order.  Only chunks that hold outliers do any work;
// ranks come from a workgroup prefix of per-thread counts.
// ---------------------------------------------------------------------------
// 64 consecutive codes from start as eight 16-byte words (past the end: 1,
// which is never an outlier).
__device__ __forceinline__ void load64(const uint16_t *__restrict__ codes, uint32_t start, uint32_t end, bool aligned,
                                       uint4 *q) {
    if (aligned && start + 64 <= end) {
        const uint4 *p = reinterpret_cast<const uint4 *>(codes + start);
#pragma unroll
        for (int g = 0; g < 8; g++) q[g] = p[g];
    } else {
#pragma unroll
        for (int g = 0; g < 8; g++) {
            uint32_t t[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint32_t j = start + 8 * g + k;
                t[k] = j < end ? codes[j] : 1u;  // past the end: never an outlier
            }
            q[g] = make_uint4(t[0] | t[1] << 16, t[2] | t[3] << 16, t[4] | t[5] << 16, t[6] | t[7] << 16);
        }
    }
}

__device__ __forceinline__ uint16_t code_at(const uint4 *q, int k) {
    const uint4 &w = q[k >> 3];
    const uint32_t word = (k & 6) == 0 ? w.x : (k & 6) == 2 ? w.y : (k & 6) == 4 ? w.z : w.w;
    return (uint16_t) (k & 1 ? word >> 16 : word);
}

// Workgroup exclusive prefix (kHistThreads threads).
__device__ __forceinline__ uint32_t chunk_exclusive(uint32_t x, uint32_t *s_wave) {
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    uint32_t incl = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = (uint32_t) __shfl_up((int) incl, o);
        if (lane >= (uint32_t) o) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (uint32_t w = 0; w < kHistThreads / 64; w++) before += w < wave ? s_wave[w] : 0;
    return before + incl - x;
}

// The outliers of a chunk in code order: a thread takes 64 consecutive codes
// (kHistPerThread), its rank comes from a workgroup prefix of per-thread
// outlier counts.  Only chunks that hold outliers do any work.
template<bool GATHER>
__global__ void __launch_bounds__(kHistThreads) unpred_move_kernel(const uint16_t *__restrict__ codes, uint32_t n,
                                                                  const float *__restrict__ src,
                                                                  const uint32_t *__restrict__ chunk_zeros,
                                                                  const uint64_t *__restrict__ chunk_base,
                                                                  float *__restrict__ dst, int aligned,
                                                                  uint64_t limit) {
    static_assert(kHistPerThread == 64, "one load64 per thread and chunk");
    __shared__ uint32_t s_wave[kHistThreads / 64];
    const uint32_t nchunks = (n + kHistChunk - 1) / kHistChunk;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        if (chunk_zeros[c] == 0) continue;
        const uint32_t base = c * kHistChunk;
        const uint32_t end = min(base + kHistChunk, n);
        const uint32_t my0 = base + threadIdx.x * kHistPerThread;
        uint4 q[8];
        load64(codes, my0, end, aligned != 0, q);
        uint32_t z = 0;
#pragma unroll
        for (int k = 0; k < 64; k++) z += code_at(q, k) == 0;
        uint64_t run = chunk_base[c] + chunk_exclusive(z, s_wave);
        if (z) {
#pragma unroll
            for (int k = 0; k < 64; k++) {
                if (code_at(q, k) == 0) {
                    // gather: the list in code order; scatter: back to code positions
                    if (GATHER)
                        dst[run] = src[my0 + k];
                    else if (run < limit)  // a count mismatch is reported by the caller
                        dst[my0 + k] = src[run];
                    run++;
                }
            }
        }
        __syncthreads();
    }
}

void launch_unpred_gather(const uint16_t *d_codes, size_t n, const float *d_outval, const uint32_t *d_chunk_zeros,
                          const uint64_t *d_chunk_base, float *d_unpred, hipStream_t st) {
    uint32_t nchunks = (uint32_t) ((n + kHistChunk - 1) / kHistChunk);
    if (nchunks > 2048) nchunks = 2048;
    if (nchunks)
        hipLaunchKernelGGL(unpred_move_kernel<true>, dim3(nchunks), dim3(kHistThreads), 0, st, d_codes, (uint32_t) n,
                           d_outval, d_chunk_zeros, d_chunk_base, d_unpred, (int) (((uintptr_t) d_codes & 15) == 0),
                           ~0ull);
}

// The inverse for decompression: unpredictable value of every code-0
// position, indexed by code position (unpred_move_kernel<false>); zeros
// ranked at or beyond count (a stream whose list is too short) take nothing.
void launch_unpred_scatter(const uint16_t *d_codes, size_t n, const float *d_unpred, const uint32_t *d_chunk_zeros,
                           const uint64_t *d_chunk_base, float *d_unpred_at, hipStream_t st, uint64_t count) {
    uint32_t nchunks = (uint32_t) ((n + kHistChunk - 1) / kHistChunk);
    if (nchunks > 2048) nchunks = 2048;
    if (nchunks)
        hipLaunchKernelGGL(unpred_move_kernel<false>, dim3(nchunks), dim3(kHistThreads), 0, st, d_codes,
                           (uint32_t) n, d_unpred, d_chunk_zeros, d_chunk_base, d_unpred_at,
                           (int) (((uintptr_t) d_codes & 15) == 0), (uint64_t) count);
}

// The decompression's checks in one block of (host) memory: the decoded
// symbol count, the decoder's error flag, the zero-code count and the
// recovery's error flag.
__global__ void decode_report_kernel(const uint64_t *__restrict__ symbols, const uint32_t *__restrict__ dec_err,
                                     const uint64_t *__restrict__ zeros, const uint32_t *__restrict__ rec_err,
                                     uint32_t *__restrict__ out) {
    if (threadIdx.x == 0) {
        const uint64_t a = symbols ? *symbols : ~0ull, b = *zeros;  // null: checked by the decoder
        out[0] = (uint32_t) a;
        out[1] = (uint32_t) (a >> 32);
        out[2] = dec_err ? *dec_err : 0u;
        out[3] = *rec_err;
        out[4] = (uint32_t) b;
        out[5] = (uint32_t) (b >> 32);
    }
}

void launch_decode_report(const uint64_t *d_symbols, const uint32_t *d_dec_err, const uint64_t *d_zeros,
                          const uint32_t *d_rec_err, uint32_t *out, hipStream_t st) {
    hipLaunchKernelGGL(decode_report_kernel, dim3(1), dim3(64), 0, st, d_symbols, d_dec_err, d_zeros, d_rec_err, out);
}


}  // namespace faz
