// One level of the 3-D interpolation engine over whole tiles: the fast path
// for interpBlockSize 32.
//
// A block of the level (block_interpolation<3>, SIC:3143-3185) spans 32 or 33
// lattice points per axis.  Its whole tile (at most 33^3 floats, 152 KB with
// the padded strides below; gfx950 gives a workgroup up to 160 KB of LDS)
// sits in LDS, so passes A (along p0), B (along p1) and C (along p2) never
// touch HBM.  Workgroups are persistent: while one tile is predicted the next
// tile's values are already in flight into registers, so the HBM reads of a
// block overlap the arithmetic of the previous one.
//
// Work mapping inside a pass: four lanes per line, lane t (= lane & 3)
// taking the odd positions i = 8t+1, 8t+3, 8t+5, 8t+7, whose seven even
// neighbours are read once and shared by the four stencils.  Lines have
// 32 or 33 points only, so every position has a fixed stencil (cubic inside,
// quad_1 at i = 1, quad_2 at the loop exit, quad_3 at n-1 for even n:
// SIC:1360-1414) and a fixed emission rank.
//
// A block recomputes the faces that the reference assigns to the previous
// blocks (identical inputs give identical bits) and emits codes only for the
// lines it owns, straight into its segment of the code stream.
// Reconstructed values go back for the half-open block [b, b + bs) (through
// the last point for the last block along an axis), so every point of the
// level is written by exactly one workgroup.
#include <algorithm>

#include "faz_device.h"
#include "faz_internal.h"

namespace faz {

namespace {

constexpr int kThreads = 1024;               // 16 waves: four per SIMD, up to 128 VGPRs each
constexpr int kRow = 35;                     // LDS stride of axis 1 (33 + 2 pad)
constexpr int kPlane = 1157;                 // LDS stride of axis 0 (33 * 35 + 2 pad)
constexpr int kTileFloats = 33 * kPlane;     // 38181 floats = 152724 bytes
constexpr int kEdge = 33 * 32 + 33 * 33;     // l0 = 32 plane (l2 < 32) and l2 = 32 column
constexpr int kEdgeSlots = (kEdge + kThreads - 1) / kThreads;
#ifndef FAZ_TILE_LANES
#define FAZ_TILE_LANES 4
#endif
constexpr int kLanesPerLine = FAZ_TILE_LANES;  // lanes sharing one line of 16 odd positions
// rounds of a pass unrolled together (more loads in flight, more registers)
#ifndef FAZ_TILE_UNROLL
#define FAZ_TILE_UNROLL 1
#endif
#define FAZ_PRAGMA(x) _Pragma(#x)
#define FAZ_UNROLL(n) FAZ_PRAGMA(unroll n)
constexpr int kPerLane = 16 / kLanesPerLine;  // odd positions per lane
constexpr int kLinesPerRound = kThreads / kLanesPerLine;

__host__ __device__ constexpr int lds_stride(int axis) { return axis == 0 ? kPlane : (axis == 1 ? kRow : 1); }

// Flag bits of an edge or coarse slot: the slot exists, and which of its
// coordinates sit at tile position 32 (present only when that axis has 33).
enum : uint32_t { kHas = 1, kAt32_0 = 2, kAt32_1 = 4, kAt32_2 = 8, kAllEven = 16 };

struct Tile {
    uint32_t k[3];      // block coordinates in the launched raster
    uint32_t n[3];
    uint32_t first;     // bit a set when the block starts at 0 on axis a
    uint32_t miss;      // kAt32_a set when axis a has no position 32
    uint32_t nown;      // kAt32_a set when this block does not write position 32 of axis a
    uint64_t gbase;     // element offset of the block origin in the input
    uint64_t cbase;     // ... in the coarser lattice
    uint64_t dbase;     // ... in the destination lattice
    uint64_t seg;       // first code of the block
    uint32_t cnt0, cnt1, total;
    uint32_t emit;      // 1 when the block's code segment lies inside the stream
};

__device__ __forceinline__ uint32_t first_of(const Tile &B, int a) { return (B.first >> a) & 1; }
// lines over all positions / over even positions of axis a that the block owns
__device__ __forceinline__ uint32_t g1_of(const Tile &B, int a) { return B.n[a] - 1 + first_of(B, a); }
__device__ __forceinline__ uint32_t g2_of(const Tile &B, int a) { return (B.n[a] - 1) / 2 + first_of(B, a); }

// Origins and missing-position mask of block blk (what the loads need).
__device__ __forceinline__ void tile_origin(const Level3D &L, uint32_t blk, uint64_t &gbase, uint64_t &cbase,
                                            uint32_t &miss) {
    uint32_t r = blk;
    const uint32_t k2 = r % L.nbw[2];
    r /= L.nbw[2];
    const uint32_t k1 = r % L.nbw[1];
    const uint32_t k0 = r / L.nbw[1];
    const uint32_t k[3] = {k0, k1, k2};
    uint64_t g = 0, c = 0;
    miss = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const uint32_t b = k[a] * L.bs;
        const uint32_t e = min(b + L.bs, L.dims[a] - 1);
        if ((e - b) / L.s + 1 != 33) miss |= kAt32_0 << a;
        g += (uint64_t) b * (a == 2 ? 1 : L.offs[a]);
        c += (uint64_t) (b / (2 * L.s)) * L.cstr[a];
    }
    gbase = g;
    cbase = c;
}

// The block at raster coordinates (k0, k1, k2).
template<int DIR>
__device__ __forceinline__ Tile tile_at(const Level3D &L, uint32_t k0, uint32_t k1, uint32_t k2) {
    constexpr int p0 = kPerm3[DIR][0], p1 = kPerm3[DIR][1], p2 = kPerm3[DIR][2];
    Tile B;
    const uint32_t k[3] = {k0, k1, k2};
    B.k[0] = k0;
    B.k[1] = k1;
    B.k[2] = k2;
    uint32_t b[3];
    B.miss = 0;
    B.nown = 0;
    B.first = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        b[a] = k[a] * L.bs;
        const uint32_t e = min(b[a] + L.bs, L.dims[a] - 1);
        B.n[a] = (e - b[a]) / L.s + 1;
        const uint32_t whi = e == L.dims[a] - 1 ? B.n[a] - 1 : B.n[a] - 2;
        if (b[a] == 0) B.first |= 1u << a;
        if (B.n[a] != 33) B.miss |= kAt32_0 << a;
        if (whi != 32) B.nown |= kAt32_0 << a;
    }
    B.gbase = (uint64_t) b[0] * L.offs[0] + (uint64_t) b[1] * L.offs[1] + b[2];
    B.cbase = B.dbase = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        B.cbase += (uint64_t) (b[a] / (2 * L.s)) * L.cstr[a];
        B.dbase += (uint64_t) (b[a] / L.s) * L.dstr[a];
    }
    // read through the constant address space: a scalar load whose wait does
    // not include the vector loads in flight (the next tile's values)
    const auto *prefix = (const __attribute__((address_space(4))) uint64_t *) L.block_prefix;
    B.seg = L.level_base + prefix[((uint64_t) k0 * L.nb[1] + k1) * L.nb[2] + k2];
    B.cnt0 = 16 * g2_of(B, p1) * g2_of(B, p2);
    B.cnt1 = g1_of(B, p0) * 16 * g2_of(B, p2);
    B.total = B.cnt0 + B.cnt1 + g1_of(B, p0) * g1_of(B, p1) * 16;
    B.emit = B.seg + B.total <= L.code_limit ? 1u : 0u;
    return B;
}

template<int DIR>
__device__ __forceinline__ Tile tile_of(const Level3D &L, uint32_t blk) {
    uint32_t r = blk;
    const uint32_t k2 = r % L.nbw[2];
    r /= L.nbw[2];
    const uint32_t k1 = r % L.nbw[1];
    const uint32_t k0 = r / L.nbw[1];
    return tile_at<DIR>(L, k0, k1, k2);
}

// First code of pass PASS in block B's segment.
template<int DIR, int PASS>
__device__ __forceinline__ uint64_t pass_base(const Tile &B) {
    return B.seg + (PASS == 0 ? 0u : (PASS == 1 ? B.cnt0 : B.cnt0 + B.cnt1));
}

// Per-thread placement of the tile loads, fixed for the whole launch.
//  main: l0 = tid >> 5, l2 = tid & 31, l1 = 0..32 over the unrolled loop
//  edge: the l0 = 32 plane and the l2 = 32 column, kEdgeSlots per thread
// Points with every coordinate even come from the reconstruction of the
// coarser levels, all others from the original data.
struct Place {
    uint32_t gl, ll;                    // main: input byte / LDS offset of (tid >> 5, 0, l2)
    uint32_t cgl, dgl;                  // main: coarse-lattice / destination byte offset
    uint32_t lane_even;                 // l0 and l2 both even
    uint32_t eg[kEdgeSlots];            // edge: input (or coarse-lattice, all even) byte offsets
    uint32_t dg[kEdgeSlots];            // edge: destination byte offsets
    uint32_t ef[kEdgeSlots];            // edge: LDS index << 8 | flags
};

// Element strides per tile step: in the input (S), the coarser lattice per
// even step (C) and the destination lattice (D).
struct Strides {
    uint32_t S[3], C[3], D[3];
};

__device__ __forceinline__ Place place_of(uint32_t tid, const Strides &X) {
    Place P;
    const uint32_t l0 = tid >> 5, l2 = tid & 31;
    P.gl = 4 * (l0 * X.S[0] + l2 * X.S[2]);
    P.cgl = 4 * ((l0 >> 1) * X.C[0] + (l2 >> 1) * X.C[2]);
    P.dgl = 4 * (l0 * X.D[0] + l2 * X.D[2]);
    P.ll = l0 * kPlane + l2;
    P.lane_even = ((l0 | l2) & 1) == 0 ? 1u : 0u;
#pragma unroll
    for (int j = 0; j < kEdgeSlots; j++) {
        const uint32_t e = tid + kThreads * j;
        uint32_t c0 = 0, c1 = 0, c2 = 0, f = 0;
        if (e < 33 * 32) {
            c0 = 32, c1 = e >> 5, c2 = e & 31, f = kHas;
        } else if (e < (uint32_t) kEdge) {
            const uint32_t x = e - 33 * 32;
            c0 = x / 33, c1 = x % 33, c2 = 32, f = kHas;
        }
        const bool even = ((c0 | c1 | c2) & 1) == 0;
        f |= (c0 == 32 ? kAt32_0 : 0) | (c1 == 32 ? kAt32_1 : 0) | (c2 == 32 ? kAt32_2 : 0) | (even ? kAllEven : 0);
        P.eg[j] = even ? 4 * ((c0 >> 1) * X.C[0] + (c1 >> 1) * X.C[1] + (c2 >> 1) * X.C[2])
                       : 4 * (c0 * X.S[0] + c1 * X.S[1] + c2 * X.S[2]);
        P.dg[j] = 4 * (c0 * X.D[0] + c1 * X.D[1] + c2 * X.D[2]);
        P.ef[j] = (c0 * kPlane + c1 * kRow + c2) << 8 | f;
    }
    return P;
}

__device__ __forceinline__ bool slot_on(uint32_t f, uint32_t off) { return (f & kHas) && !(f & off); }

// Buffer resources over a block's origin: every access is a 32-bit lane
// offset plus a uniform (scalar) offset, with no 64-bit address arithmetic
// per lane.  The record limit is the whole 32-bit range; the host bounds the
// offsets of a tile below 2^31 bytes.
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc rsrc_at(const void *p, int bytes = -1) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short) 0, bytes, 0x00020000);
}
// A store at this offset falls outside every bounded resource and is dropped
// by the hardware: predicated stores without branches.
constexpr uint32_t kDropped = 0x80000000u;
__device__ __forceinline__ float ld(Rsrc r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
// Loads of the all-even points from the reconstruction.  They carry the
// non-temporal hint, which also keeps the compiler from merging them with the
// original-data load of the other lanes into one load through a per-lane
// selected (non-uniform) resource.
__device__ __forceinline__ float ld_coarse(Rsrc r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 2));
}
__device__ __forceinline__ void st(Rsrc r, float v, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}

// Values of one tile in flight (registers).
struct Pre {
    float m[33];
    float ed[kEdgeSlots];
};

__device__ __forceinline__ void issue_loads(Pre &R, const Place &P, const Strides &X, uint64_t gbase,
                                            uint64_t cbase, uint32_t miss, const float *__restrict__ in,
                                            const float *__restrict__ csrc) {
    const Rsrc ib = rsrc_at(in + gbase);
    const Rsrc rb = rsrc_at(csrc + cbase);
#pragma unroll
    for (int k = 0; k < 33; k++) {
        if (k == 32 && (miss & kAt32_1)) break;
        if (!(k & 1) && P.lane_even)
            R.m[k] = ld_coarse(rb, P.cgl, 4 * (k >> 1) * X.C[1]);
        else
            R.m[k] = ld(ib, P.gl, 4 * k * X.S[1]);
    }
#pragma unroll
    for (int j = 0; j < kEdgeSlots; j++) {
        if (slot_on(P.ef[j], miss)) {
            if (P.ef[j] & kAllEven)
                R.ed[j] = ld_coarse(rb, P.eg[j], 0);
            else
                R.ed[j] = ld(ib, P.eg[j], 0);
        }
    }
}

__device__ __forceinline__ void fill_tile(float *T, const Pre &R, const Place &P, uint32_t miss) {
#pragma unroll
    for (int k = 0; k < 32; k++) T[P.ll + k * kRow] = R.m[k];
    if (!(miss & kAt32_1)) T[P.ll + 32 * kRow] = R.m[32];
#pragma unroll
    for (int j = 0; j < kEdgeSlots; j++)
        if (slot_on(P.ef[j], miss)) T[P.ef[j] >> 8] = R.ed[j];
}

// The block's part of this level's lattice: [b, b + bs) per axis, through the
// last point for the last block along an axis.
__device__ __forceinline__ void store_tile(const float *T, const Place &P, const Strides &X, const Tile &B,
                                           float *__restrict__ dst) {
    const Rsrc db = rsrc_at(dst + B.dbase);
#pragma unroll
    for (int k = 0; k < 32; k++) st(db, T[P.ll + k * kRow], P.dgl, 4 * k * X.D[1]);
    if (!(B.nown & kAt32_1)) st(db, T[P.ll + 32 * kRow], P.dgl, 4 * 32 * X.D[1]);
#pragma unroll
    for (int j = 0; j < kEdgeSlots; j++)
        if (slot_on(P.ef[j], B.nown)) st(db, T[P.ef[j] >> 8], P.dg[j], 0);
}

// Predictions of a lane's kPerLane odd positions of one line (stride ESA in
// LDS, line start at base + lane offset already applied): lane t reads the
// even positions 2 kPerLane t - 2 + 2k, k < kPerLane + 3 (olo / ohi keep the
// first and last reads inside the line for the first and last lane, which
// never use them), and the end-of-line stencils are selected for the first
// and last lane (SIC:1360-1414).
template<bool LINEAR, int ESA>
__device__ __forceinline__ void line_pred(const float *T, int base, int olo, int ohi, bool first_lane, bool last_lane,
                                          uint32_t na, float *pr) {
    constexpr int NE = kPerLane + 3;
    float e[NE];
    e[0] = LINEAR ? 0.f : T[base + olo];
#pragma unroll
    for (int k = 1; k < NE - 1; k++) e[k] = T[base + (2 * k - 2) * ESA];
    e[NE - 1] = LINEAR ? 0.f : T[base + ohi];
    constexpr int h = kPerLane;  // e[h - 2 .. h + 1] hold positions 26 .. 32 for the last lane
    if (LINEAR) {
#pragma unroll
        for (int j = 0; j < kPerLane; j++) pr[j] = i_linear(e[j + 1], e[j + 2]);
        if (na == 32) {
            const float l1 = i_linear1(e[h - 1], e[h]);
            pr[h - 1] = last_lane ? l1 : pr[h - 1];
        }
    } else {
#pragma unroll
        for (int j = 0; j < kPerLane; j++) pr[j] = i_cubic(e[j], e[j + 1], e[j + 2], e[j + 3]);
        const float q1 = i_quad1(e[1], e[2], e[3]);
        pr[0] = first_lane ? q1 : pr[0];
        if (na == 33) {
            const float q2 = i_quad2(e[h - 1], e[h], e[h + 1]);
            pr[h - 1] = last_lane ? q2 : pr[h - 1];
        } else {
            const float q2 = i_quad2(e[h - 2], e[h - 1], e[h]);
            const float q3 = i_quad3(e[h - 2], e[h - 1], e[h]);
            pr[h - 2] = last_lane ? q2 : pr[h - 2];
            pr[h - 1] = last_lane ? q3 : pr[h - 1];
        }
    }
}

// One pass of the block over the tile in LDS.
template<int DIR, bool LINEAR, bool BOUNDED, int PASS>
__device__ __forceinline__ void tile_pass(float *T, const Tile &B, const Level3D &L, Rsrc cb, Rsrc ob,
                                          uint32_t tid, bool &wide) {
    constexpr int p0 = kPerm3[DIR][0], p1 = kPerm3[DIR][1], p2 = kPerm3[DIR][2];
    constexpr int A = PASS == 0 ? p0 : (PASS == 1 ? p1 : p2);  // interpolated axis
    constexpr int U = PASS == 0 ? p1 : p0;                     // slow line axis (code order)
    constexpr int W = PASS == 2 ? p1 : p2;                     // fast line axis
    constexpr bool UE = PASS == 0, WE = PASS != 2;             // even positions only
    constexpr uint32_t NU = UE ? 17 : 33, NW = WE ? 17 : 33;
    constexpr int ESA = lds_stride(A), ESU = lds_stride(U) * (UE ? 2 : 1), ESW = lds_stride(W) * (WE ? 2 : 1);
    constexpr uint32_t LINES = NU * NW;
    constexpr int ROUNDS = (int) ((LINES + kLinesPerRound - 1) / kLinesPerRound);

    const uint32_t t = tid % kLanesPerLine, line0 = tid / kLanesPerLine;
    const uint32_t na = B.n[A];
    const uint32_t umax = UE ? (B.n[U] - 1) / 2 : B.n[U] - 1;
    const uint32_t wmax = WE ? (B.n[W] - 1) / 2 : B.n[W] - 1;
    const uint32_t fu = first_of(B, U), fw = first_of(B, W);
    const uint32_t gw = WE ? g2_of(B, W) : g1_of(B, W);
    const uint32_t pbase = PASS == 0 ? 0 : (PASS == 1 ? B.cnt0 : B.cnt0 + B.cnt1);
    const bool first_lane = t == 0, last_lane = t == kLanesPerLine - 1;
    // emission ranks of the lane's points q = kPerLane t + j (line_rank for
    // n = 32, 33: interior cubic points first, then q = 0, then the rest)
    uint32_t rk[kPerLane];
    {
        const uint32_t ncub = na == 33 ? 14u : 13u;
#pragma unroll
        for (int j = 0; j < kPerLane; j++) {
            const uint32_t q = kPerLane * t + j;
            rk[j] = LINEAR ? q : (q == 0 ? ncub : (q <= ncub ? q - 1 : q));
        }
    }
    // lane t reads the even positions 2 kPerLane t - 2 + 2k, k < NE; the first
    // and last of them fall outside the line for the first and last lane,
    // which read a position inside it instead and never use the value
    const int olo = first_lane ? 0 : -2 * ESA;
    const int ohi = last_lane ? 2 * kPerLane * ESA : (2 * kPerLane + 2) * ESA;
    FAZ_UNROLL(FAZ_TILE_UNROLL)
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t line = line0 + (uint32_t) kLinesPerRound * r;
        const uint32_t u = line / NW, w = line - u * NW;
        if ((LINES % kLinesPerRound == 0 || line < LINES) && u <= umax && w <= wmax) {
            const int base = (int) (u * ESU + w * ESW) + (int) t * 2 * kPerLane * ESA;
            float o[kPerLane], pr[kPerLane];
#pragma unroll
            for (int j = 0; j < kPerLane; j++) o[j] = T[base + (2 * j + 1) * ESA];
            line_pred<LINEAR, ESA>(T, base, olo, ohi, first_lane, last_lane, na, pr);
            const bool own = B.emit && (fu | u) && (fw | w);
            const uint32_t cl = pbase + ((u + fu - 1) * gw + (w + fw - 1)) * 16;
            const uint32_t coff = own ? 2 * cl : kDropped;  // codes of lines owned elsewhere are dropped
#pragma unroll
            for (int j = 0; j < kPerLane; j++) {
                float rec;
                const int code = BOUNDED ? quantize_bounded(o[j], pr[j], L.eb, L.recip, L.ebf, L.radius, &rec)
                                         : quantize_select(o[j], pr[j], L.eb, L.recip, L.ebf, L.radius, &rec);
                if (!BOUNDED) wide |= (uint32_t) code > 65535u;
                T[base + (2 * j + 1) * ESA] = rec;
                __builtin_amdgcn_raw_buffer_store_b16((uint16_t) code, cb, coff + 2 * rk[j], 0, 0);
                if (own && code == 0) st(ob, o[j], 4 * (cl + rk[j]), 0);
            }
        }
    }
}

// Code lines of one pass as the decoder finds them.  A line of the block at
// position u = 0 (w = 0) of a block that does not start the axis lies on the
// face that the previous block along U (W) owns, at that block's last
// position; its codes sit in that block's segment.  owner[c] for c = mu | 2 mw
// (mu, mw: the line moves to the previous block along U, W) gives the first
// code of the pass in the owning block and its line count along W.
struct LineOwners {
    uint64_t base[4];
    uint32_t gw[4];
    uint32_t ulast, wlast;  // own index of the last U / W position in the previous block
    uint32_t ok;            // every owning block's segment lies inside the stream
};

template<int DIR, int PASS>
__device__ __forceinline__ LineOwners line_owners(const Level3D &L, const Tile &B) {
    constexpr int p0 = kPerm3[DIR][0], p1 = kPerm3[DIR][1], p2 = kPerm3[DIR][2];
    constexpr int U = PASS == 0 ? p1 : p0;
    constexpr int W = PASS == 2 ? p1 : p2;
    constexpr bool UE = PASS == 0, WE = PASS != 2;
    LineOwners O;
    const uint32_t gw = WE ? g2_of(B, W) : g1_of(B, W);
    O.base[0] = pass_base<DIR, PASS>(B);
    O.gw[0] = gw;
    O.base[1] = O.base[2] = O.base[3] = O.base[0];
    O.gw[1] = O.gw[2] = O.gw[3] = gw;
    O.ulast = O.wlast = 0;
    O.ok = B.emit;
    const uint32_t fu = first_of(B, U), fw = first_of(B, W);
    // a previous block along an axis is never the last one: 33 positions,
    // the last of them (16 of the even ones, or 32) owned by it
    if (!fu) {
        uint32_t k[3] = {B.k[0], B.k[1], B.k[2]};
        k[U]--;
        const Tile Q = tile_at<DIR>(L, k[0], k[1], k[2]);
        O.base[1] = pass_base<DIR, PASS>(Q);
        O.ulast = (UE ? 16u : 32u) + first_of(Q, U) - 1;
        O.ok &= Q.emit;
    }
    if (!fw) {
        uint32_t k[3] = {B.k[0], B.k[1], B.k[2]};
        k[W]--;
        const Tile Q = tile_at<DIR>(L, k[0], k[1], k[2]);
        O.base[2] = pass_base<DIR, PASS>(Q);
        O.gw[2] = WE ? g2_of(Q, W) : g1_of(Q, W);
        O.wlast = (WE ? 16u : 32u) + first_of(Q, W) - 1;
        O.ok &= Q.emit;
        if (!fu) {
            k[U]--;
            const Tile R = tile_at<DIR>(L, k[0], k[1], k[2]);
            O.base[3] = pass_base<DIR, PASS>(R);
            O.gw[3] = O.gw[2];
            O.ok &= R.emit;
        }
    }
    return O;
}

// One pass of the decoder over the tile in LDS (PB_recover, SIC:110-196 with
// Q:160-180): the stencils of the compressor, each odd position recovered
// from its code, or the unpredictable value at the code's position.
template<int DIR, bool LINEAR, int PASS>
__device__ __forceinline__ void tile_recover_pass(float *T, const Tile &B, const Level3D &L,
                                                  const uint16_t *__restrict__ codes,
                                                  const float *__restrict__ outval, uint32_t tid, uint32_t *err) {
    constexpr int p0 = kPerm3[DIR][0], p1 = kPerm3[DIR][1], p2 = kPerm3[DIR][2];
    constexpr int A = PASS == 0 ? p0 : (PASS == 1 ? p1 : p2);
    constexpr int U = PASS == 0 ? p1 : p0;
    constexpr int W = PASS == 2 ? p1 : p2;
    constexpr bool UE = PASS == 0, WE = PASS != 2;
    constexpr uint32_t NU = UE ? 17 : 33, NW = WE ? 17 : 33;
    constexpr int ESA = lds_stride(A), ESU = lds_stride(U) * (UE ? 2 : 1), ESW = lds_stride(W) * (WE ? 2 : 1);
    constexpr uint32_t LINES = NU * NW;
    constexpr int ROUNDS = (int) ((LINES + kLinesPerRound - 1) / kLinesPerRound);

    const LineOwners O = line_owners<DIR, PASS>(L, B);
    if (!O.ok) {  // uniform: a neighbour's segment falls outside the stream
        if (tid == 0) atomicOr(err, 2u);
        return;
    }
    const uint32_t t = tid % kLanesPerLine, line0 = tid / kLanesPerLine;
    const uint32_t na = B.n[A];
    const uint32_t umax = UE ? (B.n[U] - 1) / 2 : B.n[U] - 1;
    const uint32_t wmax = WE ? (B.n[W] - 1) / 2 : B.n[W] - 1;
    const uint32_t fu = first_of(B, U), fw = first_of(B, W);
    const bool first_lane = t == 0, last_lane = t == kLanesPerLine - 1;
    uint32_t rk[kPerLane];
    {
        const uint32_t ncub = na == 33 ? 14u : 13u;
#pragma unroll
        for (int j = 0; j < kPerLane; j++) {
            const uint32_t q = kPerLane * t + j;
            rk[j] = LINEAR ? q : (q == 0 ? ncub : (q <= ncub ? q - 1 : q));
        }
    }
    const int olo = first_lane ? 0 : -2 * ESA;
    const int ohi = last_lane ? 2 * kPerLane * ESA : (2 * kPerLane + 2) * ESA;
    // first code of the lane's line in round r (through its owner), or ~0
    // when the lane has no line in that round
    auto line_code = [&](int r) -> uint64_t {
        const uint32_t line = line0 + (uint32_t) kLinesPerRound * r;
        const uint32_t u = line / NW, w = line - u * NW;
        if (!((LINES % kLinesPerRound == 0 || line < LINES) && u <= umax && w <= wmax)) return ~0ull;
        const uint32_t mu = (u | fu) == 0, mw = (w | fw) == 0;
        const uint32_t c = mu | mw << 1;
        const uint64_t lb = c == 0 ? O.base[0] : (c == 1 ? O.base[1] : (c == 2 ? O.base[2] : O.base[3]));
        const uint32_t gw = c < 2 ? O.gw[0] : O.gw[2];
        const uint32_t uu = mu ? O.ulast : u + fu - 1;
        const uint32_t ww = mw ? O.wlast : w + fw - 1;
        return lb + (uint64_t) (uu * gw + ww) * 16;
    };
    // the codes of round r + 1 are read while round r is decoded
    uint64_t cl_next = line_code(0);
    int code_next[kPerLane];
    if (cl_next != ~0ull) {
#pragma unroll
        for (int j = 0; j < kPerLane; j++) code_next[j] = codes[cl_next + rk[j]];
    }
    for (int r = 0; r < ROUNDS; r++) {
        const uint64_t cl = cl_next;
        int code[kPerLane];
#pragma unroll
        for (int j = 0; j < kPerLane; j++) code[j] = code_next[j];
        if (r + 1 < ROUNDS) {
            cl_next = line_code(r + 1);
            if (cl_next != ~0ull) {
#pragma unroll
                for (int j = 0; j < kPerLane; j++) code_next[j] = codes[cl_next + rk[j]];
            }
        }
        if (cl != ~0ull) {
            const uint32_t line = line0 + (uint32_t) kLinesPerRound * r;
            const uint32_t u = line / NW, w = line - u * NW;
            const int base = (int) (u * ESU + w * ESW) + (int) t * 2 * kPerLane * ESA;
            float pr[kPerLane];
            line_pred<LINEAR, ESA>(T, base, olo, ohi, first_lane, last_lane, na, pr);
#pragma unroll
            for (int j = 0; j < kPerLane; j++) {
                const float rec = code[j] ? recover_point(pr[j], code[j], L.radius, L.eb) : outval[cl + rk[j]];
                T[base + (2 * j + 1) * ESA] = rec;
            }
        }
    }
}

// Loads of the coarser lattice only (the decoder predicts every other
// point of the tile before reading it).
__device__ __forceinline__ void issue_lattice_loads(Pre &R, const Place &P, const Strides &X, uint64_t cbase,
                                                    uint32_t miss, const float *__restrict__ csrc) {
    const Rsrc rb = rsrc_at(csrc + cbase);
    if (P.lane_even) {
#pragma unroll
        for (int k = 0; k < 33; k += 2) {
            if (k == 32 && (miss & kAt32_1)) break;
            R.m[k] = ld_coarse(rb, P.cgl, 4 * (k >> 1) * X.C[1]);
        }
    }
#pragma unroll
    for (int j = 0; j < kEdgeSlots; j++)
        if (slot_on(P.ef[j], miss) && (P.ef[j] & kAllEven)) R.ed[j] = ld_coarse(rb, P.eg[j], 0);
}

__device__ __forceinline__ void fill_lattice(float *T, const Pre &R, const Place &P, uint32_t miss) {
    if (P.lane_even) {
#pragma unroll
        for (int k = 0; k < 32; k += 2) T[P.ll + k * kRow] = R.m[k];
        if (!(miss & kAt32_1)) T[P.ll + 32 * kRow] = R.m[32];
    }
#pragma unroll
    for (int j = 0; j < kEdgeSlots; j++)
        if (slot_on(P.ef[j], miss) && (P.ef[j] & kAllEven)) T[P.ef[j] >> 8] = R.ed[j];
}

}  // namespace

template<int DIR, bool LINEAR, bool BOUNDED>
__device__ __forceinline__ void tile3d_compress(float *T, const Level3D &L, const float *__restrict__ in,
                                                uint16_t *__restrict__ codes, float *__restrict__ outval,
                                                uint32_t *__restrict__ err) {
    const uint32_t tid = threadIdx.x;
    const uint32_t nblk = L.nbw[0] * L.nbw[1] * L.nbw[2];
    // blocks of one XCD (workgroup id mod 8) form a contiguous raster range,
    // so neighbouring tiles share that XCD's L2 for their common faces
    uint32_t blk, end, step;
    if (gridDim.x % 8 == 0) {
        const uint32_t per = gridDim.x / 8, xcd = blockIdx.x % 8, chunk = (nblk + 7) / 8;
        blk = xcd * chunk + blockIdx.x / 8;
        end = min(nblk, (xcd + 1) * chunk);
        step = per;
    } else {
        blk = blockIdx.x;
        end = nblk;
        step = gridDim.x;
    }
    if (blk >= end) return;
    Strides X;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        X.S[a] = L.s * (uint32_t) (a == 2 ? 1 : L.offs[a]);
        X.C[a] = (uint32_t) L.cstr[a];
        X.D[a] = (uint32_t) L.dstr[a];
    }
    const Place P = place_of(tid, X);

    bool wide = false;  // a code beyond 16 bits (possible only without the bound)
    Pre R;
    {
        uint64_t g, c;
        uint32_t miss;
        tile_origin(L, blk, g, c, miss);
        issue_loads(R, P, X, g, c, miss, in, L.csrc);
    }
    for (;;) {
        const Tile B = tile_of<DIR>(L, blk);
        if (!B.emit && tid == 0) atomicOr(err, 2u);
        fill_tile(T, R, P, B.miss);
        __syncthreads();
        const uint32_t nxt = blk + step;
        if (nxt < end) {
            uint64_t g, c;
            uint32_t miss;
            tile_origin(L, nxt, g, c, miss);
            issue_loads(R, P, X, g, c, miss, in, L.csrc);
        }
        // bounded to the block's segment of the code stream; the segment is
        // read through a vector load, so its uniformity is stated explicitly
        const uint64_t seg = ((uint64_t) __builtin_amdgcn_readfirstlane((uint32_t) (B.seg >> 32)) << 32) |
                             __builtin_amdgcn_readfirstlane((uint32_t) B.seg);
        const int total = __builtin_amdgcn_readfirstlane((int) B.total);
        const Rsrc cb = rsrc_at(codes + seg, 2 * total);
        const Rsrc ob = rsrc_at(outval + seg, 4 * total);
        tile_pass<DIR, LINEAR, BOUNDED, 0>(T, B, L, cb, ob, tid, wide);
        __syncthreads();
        tile_pass<DIR, LINEAR, BOUNDED, 1>(T, B, L, cb, ob, tid, wide);
        __syncthreads();
        tile_pass<DIR, LINEAR, BOUNDED, 2>(T, B, L, cb, ob, tid, wide);
        __syncthreads();
        if (L.dst) store_tile(T, P, X, B, L.dst);
        if (nxt >= end) break;
        __syncthreads();
        blk = nxt;
    }
    if (!BOUNDED && wide) atomicOr(err, 1u);
}

// MODE 0: the exact quantizer, 1: the bounded one, 2: the bounded one with
// the level's bounds read from L.ebp (computed on the device); a level whose
// bound does not admit it flags bit 3 of err and writes nothing, and the
// caller compresses again with bounds from the host (which then picks the
// exact quantizer): a kernel that holds both forms keeps more uniform values
// spilled than either.
template<int DIR, bool LINEAR, int MODE>
__global__ void __launch_bounds__(kThreads) interp_tile3d_kernel(Level3D L, const float *__restrict__ in,
                                                                 uint16_t *__restrict__ codes,
                                                                 float *__restrict__ outval,
                                                                 uint32_t *__restrict__ err) {
    __shared__ float T[kTileFloats];
    if (MODE == 2) {
        const LevelEB e = *L.ebp;
        if (!e.bounded) {
            if (threadIdx.x == 0) atomicOr(err, 8u);
            return;
        }
        L.eb = e.eb;
        L.recip = e.recip;
        L.ebf = e.ebf;
        tile3d_compress<DIR, LINEAR, true>(T, L, in, codes, outval, err);
    } else {
        tile3d_compress<DIR, LINEAR, MODE == 1>(T, L, in, codes, outval, err);
    }
}

// The decoder of one whole-tile level (SZInterpolationCompressor::decompress
// with block_interpolation in recover mode, SIC:42-199): the same persistent
// tiles, the coarser lattice the only HBM read besides the codes, L.dst the
// level's lattice (at full resolution for the finest level).
template<int DIR, bool LINEAR>
__global__ void __launch_bounds__(kThreads) recover_tile3d_kernel(Level3D L, const uint16_t *__restrict__ codes,
                                                                  const float *__restrict__ outval,
                                                                  uint32_t *__restrict__ err) {
    __shared__ float T[kTileFloats];
    const uint32_t tid = threadIdx.x;
    const uint32_t nblk = L.nbw[0] * L.nbw[1] * L.nbw[2];
    uint32_t blk, end, step;
    if (gridDim.x % 8 == 0) {
        const uint32_t per = gridDim.x / 8, xcd = blockIdx.x % 8, chunk = (nblk + 7) / 8;
        blk = xcd * chunk + blockIdx.x / 8;
        end = min(nblk, (xcd + 1) * chunk);
        step = per;
    } else {
        blk = blockIdx.x;
        end = nblk;
        step = gridDim.x;
    }
    if (blk >= end) return;
    Strides X;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        X.S[a] = L.s * (uint32_t) (a == 2 ? 1 : L.offs[a]);
        X.C[a] = (uint32_t) L.cstr[a];
        X.D[a] = (uint32_t) L.dstr[a];
    }
    const Place P = place_of(tid, X);
    Pre R;
    {
        uint64_t g, c;
        uint32_t miss;
        tile_origin(L, blk, g, c, miss);
        issue_lattice_loads(R, P, X, c, miss, L.csrc);
    }
    for (;;) {
        const Tile B = tile_of<DIR>(L, blk);
        fill_lattice(T, R, P, B.miss);
        __syncthreads();
        const uint32_t nxt = blk + step;
        if (nxt < end) {
            uint64_t g, c;
            uint32_t miss;
            tile_origin(L, nxt, g, c, miss);
            issue_lattice_loads(R, P, X, c, miss, L.csrc);
        }
        if (B.emit) {
            tile_recover_pass<DIR, LINEAR, 0>(T, B, L, codes, outval, tid, err);
            __syncthreads();
            tile_recover_pass<DIR, LINEAR, 1>(T, B, L, codes, outval, tid, err);
            __syncthreads();
            tile_recover_pass<DIR, LINEAR, 2>(T, B, L, codes, outval, tid, err);
            __syncthreads();
            store_tile(T, P, X, B, L.dst);
        } else if (tid == 0) {
            atomicOr(err, 2u);
        }
        if (nxt >= end) break;
        __syncthreads();
        blk = nxt;
    }
}

void launch_recover_tile3d(const Level3D &L, const uint16_t *d_codes, const float *d_outval, uint32_t *d_err,
                           uint32_t grid, hipStream_t st) {
    const uint32_t nblk = L.nbw[0] * L.nbw[1] * L.nbw[2];
    if (!nblk || !grid) return;
    dim3 g(std::min(grid, nblk)), block(kThreads);
    int dir = 0;
    for (int d = 0; d < 6; d++)
        if (kPerm3[d][0] == L.p[0] && kPerm3[d][1] == L.p[1] && kPerm3[d][2] == L.p[2]) dir = d;
#define FAZ_R3D(D)                                                                                                 \
    do {                                                                                                           \
        if (L.linear)                                                                                              \
            hipLaunchKernelGGL((recover_tile3d_kernel<D, true>), g, block, 0, st, L, d_codes, d_outval, d_err);  \
        else                                                                                                       \
            hipLaunchKernelGGL((recover_tile3d_kernel<D, false>), g, block, 0, st, L, d_codes, d_outval, d_err); \
    } while (0)
    switch (dir) {
        case 0: FAZ_R3D(0); break;
        case 1: FAZ_R3D(1); break;
        case 2: FAZ_R3D(2); break;
        case 3: FAZ_R3D(3); break;
        case 4: FAZ_R3D(4); break;
        default: FAZ_R3D(5); break;
    }
#undef FAZ_R3D
}

bool tile3d_eligible(const Level3D &L) {
    if (L.s == 0 || L.bs != 32 * L.s) return false;
    for (int a = 0; a < 3; a++) {
        if (L.dims[a] < 33 || L.nbw[a] == 0) return false;
        const uint64_t b = (uint64_t) (L.nbw[a] - 1) * L.bs;
        const uint64_t e = std::min<uint64_t>(b + L.bs, L.dims[a] - 1);
        const uint64_t n = (e - b) / L.s + 1;
        if (n != 32 && n != 33) return false;
    }
    // every in-tile byte offset stays below 2^31
    const uint64_t span = 32ull * L.s * (L.offs[0] + L.offs[1] + 1);
    return span < (1ull << 29);
}

void launch_interp_tile3d(const Level3D &L, const float *d_in, uint16_t *d_codes, float *d_outval, uint32_t *d_err,
                          uint32_t grid, hipStream_t st) {
    const uint32_t nblk = L.nbw[0] * L.nbw[1] * L.nbw[2];
    if (!nblk || !grid) return;
    dim3 g(std::min(grid, nblk)), block(kThreads);
    int dir = 0;
    for (int d = 0; d < 6; d++)
        if (kPerm3[d][0] == L.p[0] && kPerm3[d][1] == L.p[1] && kPerm3[d][2] == L.p[2]) dir = d;
#define FAZ_T3D_B(D, LIN)                                                                                        \
    do {                                                                                                         \
        if (L.ebp)                                                                                               \
            hipLaunchKernelGGL((interp_tile3d_kernel<D, LIN, 2>), g, block, 0, st, L, d_in, d_codes, d_outval,    \
                               d_err);                                                                           \
        else if (L.bounded)                                                                                      \
            hipLaunchKernelGGL((interp_tile3d_kernel<D, LIN, 1>), g, block, 0, st, L, d_in, d_codes, d_outval,    \
                               d_err);                                                                           \
        else                                                                                                     \
            hipLaunchKernelGGL((interp_tile3d_kernel<D, LIN, 0>), g, block, 0, st, L, d_in, d_codes, d_outval,    \
                               d_err);                                                                           \
    } while (0)
#define FAZ_T3D(D)                 \
    do {                           \
        if (L.linear)              \
            FAZ_T3D_B(D, true);    \
        else                       \
            FAZ_T3D_B(D, false);   \
    } while (0)
    switch (dir) {
        case 0: FAZ_T3D(0); break;
        case 1: FAZ_T3D(1); break;
        case 2: FAZ_T3D(2); break;
        case 3: FAZ_T3D(3); break;
        case 4: FAZ_T3D(4); break;
        default: FAZ_T3D(5); break;
    }
#undef FAZ_T3D
#undef FAZ_T3D_B
}

}  // namespace faz
