// One level of the 3-D interpolation engine as a single launch.
//
// One 1024-thread workgroup takes one block of the level (block_interpolation
// <3>, SIC:3143-3185): with the axis order (p0, p1, p2) of the direction, the
// points whose p2 coordinate is even ("E", at most 33 x 33 x 17 floats) are
// staged in LDS; passes A (along p0) and B (along p1) run entirely in LDS;
// pass C (along p2) streams the odd-p2 originals from HBM and reads its
// stencil neighbours from E.  The workgroup also recomputes the block's faces
// that the reference assigns to the previous blocks (identical inputs give
// identical bits), so blocks never exchange data within a level.
//
// Codes of the block are staged in LDS at their offsets inside the block's
// segment of the code stream and written out contiguously.  Reconstructed
// values are written for the half-open block [b, b + bs) (through the last
// point for the last block), so every row of a level is written exactly once.
#include "faz_device.h"
#include "faz_internal.h"

namespace faz {

namespace {

constexpr int kThreads = 1024;
constexpr int kTile = kFusedMaxBlock + 1;            // 33 points per axis
constexpr int kHalf = (kTile + 1) / 2;               // 17 even positions
constexpr int kEMax = kTile * kTile * kHalf;         // 18513 floats
constexpr int kCodeMax = kTile * kTile * kTile - kHalf * kHalf * kHalf;  // 31024 codes

struct Blk {
    uint32_t b[3], n[3], whi[3];
    uint32_t first[3];   // 1 when the block starts at 0 (owns its low faces)
    uint32_t g1[3], g2[3];
    uint32_t m[3];       // E extents per natural axis
    uint32_t es[3];      // E strides per natural axis
    uint64_t seg;        // first code of the block
    uint32_t cntA, cntB, total;
};

// x / d for x < 2^20 through a float reciprocal, corrected by one step
struct FastDiv {
    uint32_t d;
    float r;
    __device__ __forceinline__ explicit FastDiv(uint32_t d_) : d(d_), r(1.0f / (float) d_) {}
    __device__ __forceinline__ uint32_t div(uint32_t x) const {
        uint32_t q = (uint32_t) ((float) x * r);
        if (q * d > x)
            q--;
        else if ((q + 1) * d <= x)
            q++;
        return q;
    }
};

__device__ __forceinline__ uint32_t idx_g1(uint32_t l, uint32_t first) { return first ? l : l - 1; }
__device__ __forceinline__ uint32_t idx_g2(uint32_t l, uint32_t first) { return first ? l / 2 : l / 2 - 1; }

__device__ __forceinline__ uint64_t gaddr(const Level3D &L, const Blk &B, uint32_t l0, uint32_t l1, uint32_t l2) {
    return (uint64_t) (B.b[0] + l0 * L.s) * L.offs[0] + (uint64_t) (B.b[1] + l1 * L.s) * L.offs[1] +
           (uint64_t) (B.b[2] + l2 * L.s);
}

// E index of a point whose p2 coordinate is even
__device__ __forceinline__ uint32_t eidx(const Blk &B, int p2, uint32_t l0, uint32_t l1, uint32_t l2) {
    uint32_t e0 = p2 == 0 ? l0 / 2 : l0, e1 = p2 == 1 ? l1 / 2 : l1, e2 = p2 == 2 ? l2 / 2 : l2;
    return e0 * B.es[0] + e1 * B.es[1] + e2 * B.es[2];
}

__device__ __forceinline__ void store_code(uint16_t *s_codes, uint32_t slot, int code, float orig, const Blk &B,
                                           float *__restrict__ outval, uint32_t *__restrict__ err) {
    if ((unsigned) code > 65535u) atomicOr(err, 1u);
    s_codes[slot] = (uint16_t) code;
    if (code == 0) outval[B.seg + slot] = orig;
}

}  // namespace

// axis order of direction DIR: std::next_permutation order of (0, 1, 2)
constexpr int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

template<int DIR>
__global__ void __launch_bounds__(kThreads) interp_level3d_kernel(Level3D L, const float *in, float *recon,
                                                                  uint16_t *__restrict__ codes,
                                                                  float *__restrict__ outval,
                                                                  uint32_t *__restrict__ err) {
    __shared__ float E[kEMax];
    __shared__ uint16_t s_codes[kCodeMax];
    constexpr int p0 = kPerm[DIR][0], p1 = kPerm[DIR][1], P2 = kPerm[DIR][2];
    const int tid = threadIdx.x;

    // ---- block geometry ----------------------------------------------------
    Blk B;
    {
        uint32_t r = blockIdx.x;  // once per workgroup
        uint32_t k2 = r % L.nbw[2];
        r /= L.nbw[2];
        uint32_t k1 = r % L.nbw[1];
        uint32_t k0 = r / L.nbw[1];
        uint32_t k[3] = {k0, k1, k2};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            B.b[a] = k[a] * L.bs;
            uint32_t e = min(B.b[a] + L.bs, L.dims[a] - 1);
            B.n[a] = (e - B.b[a]) / L.s + 1;
            B.whi[a] = e == L.dims[a] - 1 ? B.n[a] - 1 : B.n[a] - 2;
            B.first[a] = B.b[a] == 0;
            B.g1[a] = B.n[a] - 1 + B.first[a];
            B.g2[a] = (B.n[a] - 1) / 2 + B.first[a];
            B.m[a] = a == P2 ? (B.n[a] + 1) / 2 : B.n[a];
        }
        B.es[2] = 1;
        B.es[1] = (P2 == 2 ? kHalf : kTile);
        B.es[0] = B.es[1] * (P2 == 1 ? kHalf : kTile);
        uint64_t kl = ((uint64_t) k0 * L.nb[1] + k1) * L.nb[2] + k2;
        B.seg = L.level_base + L.block_prefix[kl];
        B.cntA = (B.n[p0] / 2) * B.g2[p1] * B.g2[P2];
        B.cntB = B.g1[p0] * (B.n[p1] / 2) * B.g2[P2];
        B.total = B.cntA + B.cntB + B.g1[p0] * B.g1[p1] * (B.n[P2] / 2);
    }
    // the host plan bounds every block segment; refuse to write outside it
    if (B.seg + B.total > L.code_limit || B.total > (uint32_t) kCodeMax) {
        if (tid == 0) atomicOr(err, 2u);
        return;
    }

    // ---- load E: coarse points from recon, the rest from the input ---------
    // (flat index over E with axis 2 fastest; R loads in flight per thread)
    {
        const uint32_t m2 = B.m[2], m1 = B.m[1];
        const uint32_t total = B.m[0] * m1 * m2;
        const FastDiv d2(m2), d1(m1);
        constexpr int R = 8;
        for (uint32_t base = 0; base < total; base += kThreads * R) {
            float v[R];
            uint32_t dst[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint32_t idx = base + r * kThreads + tid;
                dst[r] = 0xFFFFFFFFu;
                if (idx < total) {
                    uint32_t row = d2.div(idx), e2 = idx - row * m2;
                    uint32_t e0 = d1.div(row), e1 = row - e0 * m1;
                    uint32_t l0 = P2 == 0 ? 2 * e0 : e0, l1 = P2 == 1 ? 2 * e1 : e1, l2 = P2 == 2 ? 2 * e2 : e2;
                    const float *src = ((l0 | l1 | l2) & 1) ? in : recon;
                    v[r] = src[gaddr(L, B, l0, l1, l2)];
                    dst[r] = e0 * B.es[0] + e1 * B.es[1] + e2;
                }
            }
#pragma unroll
            for (int r = 0; r < R; r++)
                if (dst[r] != 0xFFFFFFFFu) E[dst[r]] = v[r];
        }
    }
    __syncthreads();

    // ---- pass A: along p0 at even (p1, p2), odd p0 -------------------------
    {
        const uint32_t cn = B.n[p0] / 2, hj = (B.n[p1] + 1) / 2, hk = (B.n[P2] + 1) / 2;
        const uint32_t cnt = cn * hj * hk;
        const uint32_t st = B.es[p0];
        const FastDiv dcn(cn ? cn : 1), dhj(hj);
        for (uint32_t t = tid; t < cnt; t += kThreads) {
            uint32_t r = dcn.div(t), q = t - r * cn;
            uint32_t r2 = dhj.div(r);
            uint32_t j = 2 * (r - r2 * hj), kk = 2 * r2;
            uint32_t i = 2 * q + 1;
            uint32_t l[3];
            l[p0] = i;
            l[p1] = j;
            l[P2] = kk;
            uint32_t ei = eidx(B, P2, l[0], l[1], l[2]);
            float orig = E[ei];
            float pred = stencil_pred(i, B.n[p0], L.linear, [&](int d) { return E[(int) ei + d * (int) st]; });
            float rec;
            int code = quantize_point(orig, pred, L.eb, L.recip, L.radius, &rec);
            E[ei] = rec;
            if ((B.first[p1] || j) && (B.first[P2] || kk)) {
                uint32_t line = idx_g2(j, B.first[p1]) * B.g2[P2] + idx_g2(kk, B.first[P2]);
                store_code(s_codes, line * cn + line_rank(i, B.n[p0], L.linear), code, orig, B, outval, err);
            }
        }
    }
    __syncthreads();

    // ---- pass B: along p1 at any p0, even p2, odd p1 -----------------------
    {
        const uint32_t cn = B.n[p1] / 2, ni = B.n[p0], hk = (B.n[P2] + 1) / 2;
        const uint32_t cnt = cn * ni * hk;
        const uint32_t st = B.es[p1];
        const FastDiv dcn(cn ? cn : 1), dhk(hk);
        for (uint32_t t = tid; t < cnt; t += kThreads) {
            uint32_t r = dcn.div(t), q = t - r * cn;
            uint32_t i0 = dhk.div(r);
            uint32_t kk = 2 * (r - i0 * hk);
            uint32_t i = 2 * q + 1;
            uint32_t l[3];
            l[p0] = i0;
            l[p1] = i;
            l[P2] = kk;
            uint32_t ei = eidx(B, P2, l[0], l[1], l[2]);
            float orig = E[ei];
            float pred = stencil_pred(i, B.n[p1], L.linear, [&](int d) { return E[(int) ei + d * (int) st]; });
            float rec;
            int code = quantize_point(orig, pred, L.eb, L.recip, L.radius, &rec);
            E[ei] = rec;
            if ((B.first[p0] || i0) && (B.first[P2] || kk)) {
                uint32_t line = idx_g1(i0, B.first[p0]) * B.g2[P2] + idx_g2(kk, B.first[P2]);
                store_code(s_codes, B.cntA + line * cn + line_rank(i, B.n[p1], L.linear), code, orig, B, outval,
                           err);
            }
        }
    }
    __syncthreads();

    // ---- pass C: along p2 at any (p0, p1), odd p2 --------------------------
    const uint32_t cnC = B.n[P2] / 2;
    const uint32_t stC = B.es[P2];
    const uint32_t baseC = B.cntA + B.cntB;
    if (P2 == 2) {
        // rows along the contiguous axis, 16 lanes per row: lane jj writes
        // the even value 2jj (lane 15 also 32) and predicts 2jj + 1; RB rows
        // per lane are loaded before any is computed
        const uint32_t lane = tid & 63, wave = tid >> 6;
        const uint32_t sub = lane >> 4, jj = lane & 15;
        const uint32_t rows = B.n[0] * B.n[1];
        const FastDiv dn1(B.n[1]);
        const uint32_t ie = 2 * jj, i = ie + 1;
        const bool odd_ok = i < B.n[2];
        const bool tail = jj == 15 && B.n[2] > 32;  // even position 32
        constexpr int RB = 4;
        constexpr uint32_t kRowsPerStep = (kThreads / 64) * 4;
        for (uint32_t r0 = wave * 4 + sub; r0 < rows; r0 += kRowsPerStep * RB) {
            float orig[RB];
#pragma unroll
            for (int b = 0; b < RB; b++) {
                const uint32_t row = r0 + b * kRowsPerStep;
                if (row < rows && odd_ok) {
                    uint32_t l0 = dn1.div(row), l1 = row - l0 * B.n[1];
                    orig[b] = in[gaddr(L, B, l0, l1, i)];
                }
            }
#pragma unroll
            for (int b = 0; b < RB; b++) {
                const uint32_t row = r0 + b * kRowsPerStep;
                if (row >= rows) break;
                uint32_t l0 = dn1.div(row), l1 = row - l0 * B.n[1];
                const uint32_t eb0 = l0 * B.es[0] + l1 * B.es[1];
                const bool wrow = l0 <= B.whi[0] && l1 <= B.whi[1];
                const uint64_t f = gaddr(L, B, l0, l1, ie);
                if (wrow && ie <= B.whi[2]) recon[f] = E[eb0 + jj];
                if (wrow && tail && 32 <= B.whi[2]) recon[gaddr(L, B, l0, l1, 32)] = E[eb0 + 16];
                if (odd_ok) {
                    const uint32_t ei0 = eb0 + jj;  // E index of position i - 1
                    float pred = stencil_pred(i, B.n[2], L.linear, [&](int d) {
                        return E[(int) ei0 + ((d + 1) / 2) * (int) stC];
                    });
                    float rec;
                    int code = quantize_point(orig[b], pred, L.eb, L.recip, L.radius, &rec);
                    if (wrow && i <= B.whi[2]) recon[gaddr(L, B, l0, l1, i)] = rec;
                    const uint32_t lq0 = p0 == 0 ? l0 : l1, lq1 = p0 == 0 ? l1 : l0;  // (p0, p1) coordinates
                    if ((B.first[p0] || lq0) && (B.first[p1] || lq1)) {
                        uint32_t line = idx_g1(lq0, B.first[p0]) * B.g1[p1] + idx_g1(lq1, B.first[p1]);
                        store_code(s_codes, baseC + line * cnC + line_rank(i, B.n[2], L.linear), code, orig[b], B,
                                   outval, err);
                    }
                }
            }
        }
    } else {
        // p2 is axis 0 or 1: write the E rows, then walk columns along p2
        // with lanes along the contiguous axis 2
        const int q = P2 == 0 ? 1 : 0;  // the other non-contiguous axis
        {
            const uint32_t m2 = B.m[2], m1 = B.m[1];
            const uint32_t total = B.m[0] * m1 * m2;
            const FastDiv d2(m2), d1(m1);
            for (uint32_t idx = tid; idx < total; idx += kThreads) {
                uint32_t row = d2.div(idx), e2 = idx - row * m2;
                uint32_t e0 = d1.div(row), e1 = row - e0 * m1;
                uint32_t l0 = P2 == 0 ? 2 * e0 : e0, l1 = P2 == 1 ? 2 * e1 : e1;
                if (l0 <= B.whi[0] && l1 <= B.whi[1] && e2 <= B.whi[2])
                    recon[gaddr(L, B, l0, l1, e2)] = E[e0 * B.es[0] + e1 * B.es[1] + e2];
            }
        }
        const uint32_t cols = B.n[q] * B.n[2];
        const FastDiv dn2(B.n[2]);
        for (uint32_t c = tid; c < cols; c += kThreads) {
            uint32_t lq = dn2.div(c), l2 = c - lq * B.n[2];
            float orig[kHalf];
#pragma unroll
            for (uint32_t jj = 0; jj < (uint32_t) kHalf - 1; jj++) {
                if (jj < cnC) {
                    uint32_t l[3];
                    l[q] = lq;
                    l[2] = l2;
                    l[P2] = 2 * jj + 1;
                    orig[jj] = in[gaddr(L, B, l[0], l[1], l[2])];
                }
            }
            uint32_t l[3];
            l[q] = lq;
            l[2] = l2;
            l[P2] = 0;
            const uint32_t ecol = eidx(B, P2, l[0], l[1], l[2]);
            const bool wcol = lq <= B.whi[q] && l2 <= B.whi[2];
            const uint32_t lq0 = p0 == 2 ? l2 : lq, lq1 = p1 == 2 ? l2 : lq;  // (p0, p1) coordinates
            const bool owned = (B.first[p0] || lq0) && (B.first[p1] || lq1);
            const uint32_t line = owned ? idx_g1(lq0, B.first[p0]) * B.g1[p1] + idx_g1(lq1, B.first[p1]) : 0;
#pragma unroll
            for (uint32_t jj = 0; jj < (uint32_t) kHalf - 1; jj++) {
                if (jj < cnC) {
                    uint32_t i = 2 * jj + 1;
                    uint32_t ei0 = ecol + jj * stC;  // E index of position i - 1
                    float pred = stencil_pred(i, B.n[P2], L.linear, [&](int d) {
                        return E[(int) ei0 + ((d + 1) / 2) * (int) stC];
                    });
                    float rec;
                    int code = quantize_point(orig[jj], pred, L.eb, L.recip, L.radius, &rec);
                    l[P2] = i;
                    if (wcol && i <= B.whi[P2]) recon[gaddr(L, B, l[0], l[1], l[2])] = rec;
                    if (owned)
                        store_code(s_codes, baseC + line * cnC + line_rank(i, B.n[P2], L.linear), code, orig[jj], B,
                                   outval, err);
                }
            }
        }
    }
    __syncthreads();

    // ---- the block's codes, contiguous in the stream -----------------------
    for (uint32_t t = tid; t < B.total; t += kThreads) codes[B.seg + t] = s_codes[t];
}

void launch_interp_level3d(const Level3D &L, const float *d_in, float *d_recon, uint16_t *d_codes, float *d_outval,
                           uint32_t *d_err, hipStream_t st) {
    const uint32_t nblk = L.nbw[0] * L.nbw[1] * L.nbw[2];
    if (!nblk) return;
    dim3 grid(nblk), block(kThreads);
    int dir = 0;
    for (int d = 0; d < 6; d++)
        if (kPerm[d][0] == L.p[0] && kPerm[d][1] == L.p[1] && kPerm[d][2] == L.p[2]) dir = d;
#define FAZ_L3D(D) hipLaunchKernelGGL(interp_level3d_kernel<D>, grid, block, 0, st, L, d_in, d_recon, d_codes, d_outval, d_err)
    switch (dir) {
        case 0: FAZ_L3D(0); break;
        case 1: FAZ_L3D(1); break;
        case 2: FAZ_L3D(2); break;
        case 3: FAZ_L3D(3); break;
        case 4: FAZ_L3D(4); break;
        default: FAZ_L3D(5); break;
    }
#undef FAZ_L3D
}

}  // namespace faz
