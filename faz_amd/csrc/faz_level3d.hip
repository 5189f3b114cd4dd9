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
// Work mapping: a lane keeps one position of a line (so its stencil, its
// emission rank and its neighbour offsets are fixed) while the lines it
// visits advance by a constant stride; the cubic-family stencils are
// evaluated without branches and the lane's own one selected.
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
constexpr int kWaves = kThreads / 64;
constexpr int kTile = kFusedMaxBlock + 1;            // 33 points per axis
constexpr int kHalf = (kTile + 1) / 2;               // 17 even positions
constexpr int kEMax = kTile * kTile * kHalf;         // 18513 floats
constexpr int kCodeMax = kTile * kTile * kTile - kHalf * kHalf * kHalf;  // 31024 codes

constexpr auto &kPerm = kPerm3;

// stencil kinds of block_interpolation_1d (SIC:1306-1417)
enum Kind : uint32_t { kCubic = 0, kQuad1, kQuad2, kQuad3, kLin, kCopy, kLin1 };

// Kind, emission rank and neighbour displacements (in line positions) of the
// odd position i of a line of n points; a..d are the stencil operands in the
// reference's argument order (unused ones repeat a valid neighbour).
struct Stencil {
    uint32_t kind, rank;
    int da, db, dc, dd;
};

__device__ __forceinline__ Stencil make_stencil(uint32_t i, uint32_t n, bool linear) {
    Stencil s;
    s.rank = line_rank(i, n, linear);
    if (linear || n < 5) {
        if (i + 1 < n) {
            s.kind = kLin, s.da = -1, s.db = 1;
        } else if (n < 4) {
            s.kind = kCopy, s.da = -1, s.db = -1;
        } else {
            s.kind = kLin1, s.da = -3, s.db = -1;
        }
        s.dc = s.dd = s.db;
        return s;
    }
    const uint32_t ex = 2 * ((n - 3) / 2) + 1;
    if (i == 1) {
        s.kind = kQuad1, s.da = -1, s.db = 1, s.dc = 3, s.dd = 3;
    } else if (i == ex) {
        s.kind = kQuad2, s.da = -3, s.db = -1, s.dc = 1, s.dd = 1;
    } else if (i == n - 1) {
        s.kind = kQuad3, s.da = -5, s.db = -3, s.dc = -1, s.dd = -1;
    } else {
        s.kind = kCubic, s.da = -3, s.db = -1, s.dc = 1, s.dd = 3;
    }
    return s;
}

// The stencil of `kind` on its operands (Interpolators.hpp:9-36, each formula
// in the reference's operation order); all candidates are computed, the
// lane's own is selected.
__device__ __forceinline__ float eval_stencil(uint32_t kind, float a, float b, float c, float d) {
    if (kind >= kLin) {
        const float lin = i_linear(a, b);
        const float lin1 = i_linear1(a, b);
        return kind == kLin ? lin : (kind == kCopy ? a : lin1);
    }
    const float cub = i_cubic(a, b, c, d);
    const float q1 = i_quad1(a, b, c);
    const float q2 = i_quad2(a, b, c);
    const float q3 = i_quad3(a, b, c);
    return kind == kCubic ? cub : (kind == kQuad1 ? q1 : (kind == kQuad2 ? q2 : q3));
}

// x / d for x < 2^20 through the hardware float reciprocal (about 1 ulp),
// corrected by one step either way
struct FastDiv {
    uint32_t d;
    float r;
    __device__ __forceinline__ explicit FastDiv(uint32_t d_) : d(d_), r(__builtin_amdgcn_rcpf((float) d_)) {}
    __device__ __forceinline__ uint32_t div(uint32_t x) const {
        uint32_t q = (uint32_t) ((float) x * r);
        if (q * d > x)
            q--;
        else if ((q + 1) * d <= x)
            q++;
        return q;
    }
};

// The (hi, lo) digits of an index over a row length nl, advanced by a
// constant stride without division.
struct Walk {
    uint32_t hi, lo, dhi, dlo, nl;
    __device__ __forceinline__ Walk(uint32_t start, uint32_t step, uint32_t nl_) : nl(nl_) {
        FastDiv f(nl_);
        hi = f.div(start);
        lo = start - hi * nl_;
        dhi = f.div(step);
        dlo = step - dhi * nl_;
    }
    __device__ __forceinline__ void next() {
        lo += dlo;
        hi += dhi;
        if (lo >= nl) {
            lo -= nl;
            hi++;
        }
    }
};

struct Blk {
    uint32_t b[3], n[3], whi[3];
    uint32_t first[3];   // 1 when the block starts at 0 (owns its low faces)
    uint32_t g1[3], g2[3];
    uint32_t es[3];      // E strides per natural axis
    uint64_t seg;        // first code of the block
    uint32_t cntA, cntB, total;
};

__device__ __forceinline__ uint64_t gaddr(const Level3D &L, const Blk &B, uint32_t l0, uint32_t l1, uint32_t l2) {
    return (uint64_t) (B.b[0] + l0 * L.s) * L.offs[0] + (uint64_t) (B.b[1] + l1 * L.s) * L.offs[1] +
           (uint64_t) (B.b[2] + l2 * L.s);
}

__device__ __forceinline__ void store_code(uint16_t *s_codes, uint32_t slot, int code, float orig, const Blk &B,
                                           float *__restrict__ outval, uint32_t *__restrict__ err) {
    if ((unsigned) code > 65535u) atomicOr(err, 1u);
    s_codes[slot] = (uint16_t) code;
    if (code == 0) outval[B.seg + slot] = orig;
}

}  // namespace

template<int DIR, bool LINEAR>
__global__ void __launch_bounds__(kThreads) interp_level3d_kernel(Level3D L, const float *in, float *recon,
                                                                  uint16_t *__restrict__ codes,
                                                                  float *__restrict__ outval,
                                                                  uint32_t *__restrict__ err) {
    if (L.ebp) {
        L.eb = L.ebp->eb;
        L.recip = L.ebp->recip;
    }
    __shared__ float E[kEMax];
    __shared__ uint32_t s_codes32[(kCodeMax + 1) / 2];
    uint16_t *s_codes = reinterpret_cast<uint16_t *>(s_codes32);
    constexpr int p0 = kPerm[DIR][0], p1 = kPerm[DIR][1], P2 = kPerm[DIR][2];
    const int tid = threadIdx.x;
    const uint32_t lane = tid & 63, wave = tid >> 6;

    // ---- block geometry (uniform) ------------------------------------------
    Blk B;
    {
        uint32_t r = blockIdx.x;
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

    // E coordinates along each natural axis: the p2 axis holds even positions only
    const uint32_t m0 = P2 == 0 ? (B.n[0] + 1) / 2 : B.n[0];
    const uint32_t m1 = P2 == 1 ? (B.n[1] + 1) / 2 : B.n[1];
    const uint32_t m2 = P2 == 2 ? (B.n[2] + 1) / 2 : B.n[2];
    constexpr uint32_t st0 = P2 == 0 ? 2 : 1, st1 = P2 == 1 ? 2 : 1, st2 = P2 == 2 ? 2 : 1;

    // ---- load E: rows along the contiguous axis, 16 (or 32) lanes a row -----
    {
        constexpr uint32_t LW = P2 == 2 ? 16 : 32;  // lanes per row (m2 <= 17 or 33)
        constexpr uint32_t RPW = 64 / LW;
        constexpr uint32_t kStep = kWaves * RPW;
        const uint32_t c = lane % LW, sub = lane / LW;
        const uint32_t rows = m0 * m1;
        constexpr int R = 4;
        Walk w(wave * RPW + sub, kStep, m1);
        for (uint32_t r0 = wave * RPW + sub; r0 < rows; r0 += kStep * R) {
            float v[R], vt[R];
            uint32_t dst[R];
#pragma unroll
            for (int u = 0; u < R; u++) {
                const uint32_t row = r0 + u * kStep;
                dst[u] = 0xFFFFFFFFu;
                if (row < rows) {
                    const uint32_t l0 = w.hi * st0, l1 = w.lo * st1;
                    const uint64_t rb = gaddr(L, B, l0, l1, 0);
                    dst[u] = w.hi * B.es[0] + w.lo * B.es[1];
                    if (c < m2) {
                        const uint32_t l2 = c * st2;
                        v[u] = (((l0 | l1 | l2) & 1) ? in : recon)[rb + l2 * L.s];
                    }
                    if (c + LW < m2) {  // the row's tail column
                        const uint32_t l2 = (c + LW) * st2;
                        vt[u] = (((l0 | l1 | l2) & 1) ? in : recon)[rb + l2 * L.s];
                    }
                }
                w.next();
            }
#pragma unroll
            for (int u = 0; u < R; u++) {
                if (dst[u] == 0xFFFFFFFFu) continue;
                if (c < m2) E[dst[u] + c] = v[u];
                if (c + LW < m2) E[dst[u] + c + LW] = vt[u];
            }
        }
    }
    __syncthreads();

    // ---- passes A and B inside LDS -----------------------------------------
    // pass A: along p0, lines over (even p1, even p2); pass B: along p1,
    // lines over (any p0, even p2).  Lane = (line sub-slot, position q).
    const uint32_t q = lane & 15, sub = lane >> 4;
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
        const int pa = pass == 0 ? p0 : p1;      // interpolated axis
        const int pu = pass == 0 ? p1 : p0;      // slow outer axis of the lines
        const uint32_t n = B.n[pa], cn = n / 2;
        const uint32_t nu = pass == 0 ? (B.n[pu] + 1) / 2 : B.n[pu];  // even p1 / any p0
        const uint32_t nw = (B.n[P2] + 1) / 2;                         // even p2
        const uint32_t esa = B.es[pa];
        const uint32_t esu = B.es[pu] * (pass == 0 ? 2 : 1);
        const uint32_t esw = B.es[P2];
        const uint32_t i = 2 * q + 1;
        const bool active = q < cn;
        const Stencil sc = make_stencil(active ? i : 1, n < 2 ? 2 : n, LINEAR);
        const int oa = sc.da * (int) esa, ob = sc.db * (int) esa, oc = sc.dc * (int) esa, od = sc.dd * (int) esa;
        const uint32_t lines = nu * nw;
        const uint32_t cbase = pass == 0 ? 0 : B.cntA;
        const uint32_t fu = B.first[pu], fw = B.first[P2];
        if (active) {
            Walk w(wave * 4 + sub, kWaves * 4, nw);
            for (uint32_t ln = wave * 4 + sub; ln < lines; ln += kWaves * 4, w.next()) {
                const uint32_t ei = i * esa + w.hi * esu + w.lo * esw;
                const float orig = E[ei];
                const float pred = eval_stencil(sc.kind, E[(int) ei + oa], E[(int) ei + ob], E[(int) ei + oc],
                                                E[(int) ei + od]);
                float rec;
                const int code = quantize_point(orig, pred, L.eb, L.recip, L.radius, &rec);
                E[ei] = rec;
                // owned lines: outer coordinates past the block start (or block 0)
                if ((fu || w.hi) && (fw || w.lo)) {
                    const uint32_t line = (w.hi - (1 - fu)) * B.g2[P2] + (w.lo - (1 - fw));
                    store_code(s_codes, cbase + line * cn + sc.rank, code, orig, B, outval, err);
                }
            }
        }
        __syncthreads();
    }

    // ---- pass C: along p2 at any (p0, p1), odd p2 --------------------------
    const uint32_t nC = B.n[P2], cnC = nC / 2;
    const uint32_t baseC = B.cntA + B.cntB;
    if (P2 == 2) {
        // rows along the contiguous axis, 16 lanes a row: lane q writes the
        // even value 2q (lane 15 also 32) and predicts 2q + 1
        const uint32_t i = 2 * q + 1;
        const bool odd_ok = i < nC;
        const bool tail = q == 15 && nC > 32;
        const Stencil sc = make_stencil(odd_ok ? i : 1, nC < 2 ? 2 : nC, LINEAR);
        // E index of position i + d within the row is (i + d) / 2 = q + (d + 1) / 2
        const int oa = (int) q + (sc.da + 1) / 2, ob = (int) q + (sc.db + 1) / 2, oc = (int) q + (sc.dc + 1) / 2,
                  od = (int) q + (sc.dd + 1) / 2;
        const uint32_t rows = B.n[0] * B.n[1];
        constexpr int RB = 4;
        constexpr uint32_t kStep = kWaves * 4;
        Walk w(wave * 4 + sub, kStep, B.n[1]);
        for (uint32_t r0 = wave * 4 + sub; r0 < rows; r0 += kStep * RB) {
            float orig[RB];
            uint32_t l0s[RB], l1s[RB];
#pragma unroll
            for (int u = 0; u < RB; u++) {
                l0s[u] = w.hi;
                l1s[u] = w.lo;
                if (r0 + u * kStep < rows && odd_ok) orig[u] = in[gaddr(L, B, w.hi, w.lo, i)];
                w.next();
            }
#pragma unroll
            for (int u = 0; u < RB; u++) {
                if (r0 + u * kStep >= rows) break;
                const uint32_t l0 = l0s[u], l1 = l1s[u];
                const uint32_t eb0 = l0 * B.es[0] + l1 * B.es[1];
                const bool wrow = l0 <= B.whi[0] && l1 <= B.whi[1];
                const uint64_t f = gaddr(L, B, l0, l1, 0);
                if (wrow && 2 * q <= B.whi[2]) recon[f + 2 * q * L.s] = E[eb0 + q];
                if (wrow && tail && 32 <= B.whi[2]) recon[f + 32 * L.s] = E[eb0 + 16];
                if (odd_ok) {
                    const float pred = eval_stencil(sc.kind, E[eb0 + oa], E[eb0 + ob], E[eb0 + oc], E[eb0 + od]);
                    float rec;
                    const int code = quantize_point(orig[u], pred, L.eb, L.recip, L.radius, &rec);
                    if (wrow && i <= B.whi[2]) recon[f + i * L.s] = rec;
                    const uint32_t lq0 = p0 == 0 ? l0 : l1, lq1 = p0 == 0 ? l1 : l0;  // (p0, p1) coordinates
                    if ((B.first[p0] || lq0) && (B.first[p1] || lq1)) {
                        const uint32_t line = (lq0 - (1 - B.first[p0])) * B.g1[p1] + (lq1 - (1 - B.first[p1]));
                        store_code(s_codes, baseC + line * cnC + sc.rank, code, orig[u], B, outval, err);
                    }
                }
            }
        }
    } else {
        // p2 is axis 0 or 1.  First the E rows (the even p2 planes) go out,
        // then lanes along the contiguous axis 2 predict the odd planes:
        // work items are (lq, position qq) pairs, 32 lanes over l2 each.
        constexpr int qa = P2 == 0 ? 1 : 0;  // the other non-contiguous axis
        const uint32_t c = lane & 31, hs = lane >> 5;
        {
            const uint32_t rows = m0 * m1;
            Walk w(wave * 2 + hs, kWaves * 2, m1);
            for (uint32_t row = wave * 2 + hs; row < rows; row += kWaves * 2, w.next()) {
                const uint32_t l0 = w.hi * st0, l1 = w.lo * st1;
                if (l0 > B.whi[0] || l1 > B.whi[1]) continue;
                const uint64_t rb = gaddr(L, B, l0, l1, 0);
                const uint32_t eb = w.hi * B.es[0] + w.lo * B.es[1];
                if (c <= B.whi[2]) recon[rb + c * L.s] = E[eb + c];
                if (c + 32 <= B.whi[2]) recon[rb + (c + 32) * L.s] = E[eb + c + 32];
            }
        }
        const uint32_t items = B.n[qa] * cnC;  // (lq, qq) pairs
        const int esP = (int) B.es[P2];
        for (uint32_t l2base = 0; l2base < B.n[2]; l2base += 32) {
            const uint32_t l2 = l2base + c;
            const bool col_ok = l2 < B.n[2];
            Walk w(wave * 2 + hs, kWaves * 2, cnC ? cnC : 1);
            for (uint32_t it = wave * 2 + hs; it < items; it += kWaves * 2, w.next()) {
                if (!col_ok) continue;
                const uint32_t lq = w.hi, qq = w.lo;
                const uint32_t i = 2 * qq + 1;
                const Stencil sc = make_stencil(i, nC, LINEAR);
                uint32_t l[3];
                l[qa] = lq;
                l[2] = l2;
                l[P2] = i;
                const uint64_t f = gaddr(L, B, l[0], l[1], l[2]);
                const float orig = in[f];
                const uint32_t ei0 = lq * B.es[qa] + l2 + qq * B.es[P2];  // E index of position i - 1
                const float pred = eval_stencil(sc.kind, E[(int) ei0 + ((sc.da + 1) / 2) * esP],
                                                E[(int) ei0 + ((sc.db + 1) / 2) * esP],
                                                E[(int) ei0 + ((sc.dc + 1) / 2) * esP],
                                                E[(int) ei0 + ((sc.dd + 1) / 2) * esP]);
                float rec;
                const int code = quantize_point(orig, pred, L.eb, L.recip, L.radius, &rec);
                if (lq <= B.whi[qa] && l2 <= B.whi[2] && i <= B.whi[P2]) recon[f] = rec;
                const uint32_t lq0 = p0 == 2 ? l2 : lq, lq1 = p1 == 2 ? l2 : lq;  // (p0, p1) coordinates
                if ((B.first[p0] || lq0) && (B.first[p1] || lq1)) {
                    const uint32_t line = (lq0 - (1 - B.first[p0])) * B.g1[p1] + (lq1 - (1 - B.first[p1]));
                    store_code(s_codes, baseC + line * cnC + sc.rank, code, orig, B, outval, err);
                }
            }
        }
    }
    __syncthreads();

    // ---- the block's codes, contiguous in the stream -----------------------
    if (B.total) {
        uint32_t t0 = 0;
        uint16_t *dst = codes + B.seg;
        if (B.seg & 1) {  // align the global side to 4 bytes
            if (tid == 0) dst[0] = s_codes[0];
            t0 = 1;
        }
        const uint32_t pairs = (B.total - t0) / 2;
        uint32_t *dst32 = reinterpret_cast<uint32_t *>(dst + t0);
        if (t0 == 0) {
            for (uint32_t t = tid; t < pairs; t += kThreads) dst32[t] = s_codes32[t];
        } else {
            for (uint32_t t = tid; t < pairs; t += kThreads)
                dst32[t] = (uint32_t) s_codes[1 + 2 * t] | ((uint32_t) s_codes[2 + 2 * t] << 16);
        }
        if (tid == 0 && ((B.total - t0) & 1)) dst[B.total - 1] = s_codes[B.total - 1];
    }
}

void launch_interp_level3d(const Level3D &L, const float *d_in, float *d_recon, uint16_t *d_codes, float *d_outval,
                           uint32_t *d_err, hipStream_t st) {
    const uint32_t nblk = L.nbw[0] * L.nbw[1] * L.nbw[2];
    if (!nblk) return;
    dim3 grid(nblk), block(kThreads);
    int dir = 0;
    for (int d = 0; d < 6; d++)
        if (kPerm[d][0] == L.p[0] && kPerm[d][1] == L.p[1] && kPerm[d][2] == L.p[2]) dir = d;
#define FAZ_L3D(D)                                                                                                \
    do {                                                                                                          \
        if (L.linear)                                                                                             \
            hipLaunchKernelGGL((interp_level3d_kernel<D, true>), grid, block, 0, st, L, d_in, d_recon, d_codes,   \
                               d_outval, d_err);                                                                  \
        else                                                                                                      \
            hipLaunchKernelGGL((interp_level3d_kernel<D, false>), grid, block, 0, st, L, d_in, d_recon, d_codes,  \
                               d_outval, d_err);                                                                  \
    } while (0)
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
