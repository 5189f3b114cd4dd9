// Internal interfaces shared by the HIP kernels (faz_kernels.hip) and the
// host orchestration (faz_gpu.hip, host/*.cpp).  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <stddef.h>
#include <stdint.h>

struct faz_gpu_ctx;

namespace faz {

// Largest supported rank and level count.
constexpr int kMaxN = 4;
constexpr int kMaxLevels = 64;

// Axis roles inside one interpolation pass (SIC:3143-3185 generalised to N):
// the pass axis carries the odd multiples of the stride, axes earlier in the
// permutation run over all multiples of the stride (G1), later ones over
// multiples of twice the stride (G2).
enum AxisKind : uint8_t { kPassAxis = 0, kG1 = 1, kG2 = 2 };

// Unsigned 32-bit division by a divisor fixed at launch: one high multiply,
// a subtract and two shifts (the round-up multiplier method of Granlund and
// Montgomery, "Division by Invariant Integers using Multiplication", 1994,
// figure 4.1), exact for every x < 2^32.
struct FastDiv {
    uint32_t m = 1, sh1 = 0, sh2 = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t d) {
        if (d == 0) d = 1;  // an empty lattice: never divided by
        uint32_t l = 0;
        while (l < 32 && ((uint64_t) 1 << l) < d) l++;
        m = (uint32_t) ((((uint64_t) 1 << 32) * (((uint64_t) 1 << l) - d)) / d + 1);
        sh1 = l ? 1 : 0;
        sh2 = l ? l - 1 : 0;
    }
    __host__ __device__ uint32_t div(uint32_t x) const {
#ifdef __HIP_DEVICE_COMPILE__
        const uint32_t t = __umulhi(m, x);
#else
        const uint32_t t = (uint32_t) (((uint64_t) m * x) >> 32);
#endif
        return (t + ((x - t) >> sh1)) >> sh2;
    }
};

// Error bound of one level computed on the device from the value range (the
// REL path without a host round trip after the range pass).
struct LevelEB {
    double eb, recip;  // level error bound and 1.0 / eb
    float ebf;         // largest float <= eb
    int bounded;       // |x - pred| / eb provably below 2^30 (see interp_compress)
};
// How a plan's levels scale the base bound: level i (coarse to fine) has
// base * f[i] (div[i] = 0) or base / f[i] (div[i] = 1), the arithmetic of
// the level schedule (SIC:384-422) with its factors computed on the host.
struct LevelScale {
    int n;
    uint8_t div[kMaxLevels];
    double f[kMaxLevels];
};

// One (level, pass) launch of the interpolation predictor.
struct PassDesc {
    int N;
    int P;                    // pass number within the block (0 = first)
    int axis;                 // natural axis interpolated along
    int perm[kMaxN];          // dimension_sequences[direction]
    int invp[kMaxN];          // invp[perm[Q]] = Q
    uint32_t dims[kMaxN];
    uint64_t offs[kMaxN];     // row-major element strides
    uint32_t lat[kMaxN];      // lattice extent per natural axis
    uint8_t kind[kMaxN];      // AxisKind per natural axis
    uint32_t nb[kMaxN];       // blocks per natural axis at this level
    uint32_t s;               // stride 2^(level-1)
    uint32_t ls;              // log2(s)
    uint32_t bs;              // block edge interpBlockSize * stride
    FastDiv dlat[kMaxN];      // division by lat[a]
    FastDiv dbs;              // division by bs
    uint64_t total;           // lattice points (= codes of this pass)
    uint64_t level_base;      // codes pushed before this level
    const uint64_t *block_prefix;  // codes of earlier blocks in raster order
    double eb, recip;         // level error bound and 1.0 / eb
    int radius;
    int linear;               // 1: linear interpolator, 0: cubic
    int tuning;               // 0 / 1 / 2
    const LevelEB *ebp;       // when set, eb and recip come from here (device)
};

// One level of the 3-D engine as a single launch: one workgroup per block.
struct Level3D {
    uint32_t dims[3];
    uint64_t offs[3];
    uint32_t s, bs;           // stride and block edge
    uint32_t nb[3];           // blocks per axis in the reference raster
    uint32_t nbw[3];          // launched blocks per axis (trailing one-plane blocks skipped)
    int p[3];                 // dimension_sequences[direction]
    int linear;
    double eb, recip;
    float ebf;                // largest float <= eb
    int bounded;              // |x - pred| / eb provably below 2^30 on this level
    int radius;
    uint64_t level_base;
    const uint64_t *block_prefix;
    uint64_t code_limit;      // codes in the whole stream (bounds every write)
    // Whole-tile kernel only: where the points of the coarser lattice (every
    // coordinate a multiple of 2s) are read, and where this level's lattice
    // (multiples of s) is written.  Lattice point c (per axis) lives at
    // csrc[sum c_a * cstr_a] / dst[sum c_a * dstr_a]; dst may be null.
    const float *csrc;
    uint64_t cstr[3];
    float *dst;
    uint64_t dstr[3];
    const LevelEB *ebp;       // when set, eb, recip, ebf and bounded come from here (device)
};

// Largest float not above a positive double.
inline float float_floor(double x) {
    float f = (float) x;
    if ((double) f > x) f = nextafterf(f, 0.0f);
    return f;
}

// Largest interpBlockSize the fused 3-D kernel stages in LDS.
constexpr uint32_t kFusedMaxBlock = 32;

// Axis order of 3-D direction d: std::next_permutation order of (0, 1, 2)
// (dimension_sequences, SIC:820-852).
constexpr int kPerm3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// Whole-tile level kernel: a level qualifies when every block it launches
// spans 32 or 33 lattice points on each axis (interpBlockSize 32) and every
// tile offset fits 32 bits.  `grid` persistent workgroups share the blocks.
bool tile3d_eligible(const Level3D &L);
void launch_interp_tile3d(const Level3D &L, const float *d_in, uint16_t *d_codes, float *d_outval, uint32_t *d_err,
                          uint32_t grid, hipStream_t st);
// The decoder of a whole-tile level: codes and the unpredictable value of
// every code-0 position (indexed by code position) in, L.csrc / L.dst as in
// the compressor; bit 1 of *d_err when a block segment leaves the stream.
void launch_recover_tile3d(const Level3D &L, const uint16_t *d_codes, const float *d_outval, uint32_t *d_err,
                           uint32_t grid, hipStream_t st);
// Zero codes per kHistChunk chunk of n 16-bit codes (the outlier ranks of
// the decoder, without the histogram).
void launch_zero_count_u16(const uint16_t *d_codes, size_t n, uint32_t *d_chunk_zeros, hipStream_t st);

// Kernel launchers (faz_kernels.hip).  All are asynchronous on `st`.
void launch_interp_level3d(const Level3D &L, const float *d_in, float *d_recon, uint16_t *d_codes, float *d_outval,
                           uint32_t *d_err, hipStream_t st);
// Value range of n floats into d_minmax[0..1] (order-preserving keys); the
// buffer holds kRangeScratch words: the pair, then per-workgroup partials.
constexpr unsigned kRangeGrid = 512;
constexpr size_t kRangeScratch = 2 + 2 * (size_t) kRangeGrid;
void launch_range(const float *d_in, size_t n, uint32_t *d_minmax, hipStream_t st);
// After launch_range on the same stream: d_out[0] = the base bound
// rel * (float) (max - min), d_out[1 + i] = level i of the plan (sc), with
// the bounded-quantizer test of each level.
void launch_level_bounds(const uint32_t *d_minmax, double rel, const LevelScale &sc, LevelEB *d_out,
                         hipStream_t st);
// The value range with the REL compression's set-up folded in: the histogram
// arrays and the error flag cleared by the range pass, the level bounds
// computed by the fold (as launch_level_bounds).
void launch_range_rel(const float *d_in, size_t n, uint32_t *d_minmax, uint32_t *d_freq, uint32_t *d_first,
                      uint32_t *d_err, double rel, const LevelScale &sc, LevelEB *d_lvl, hipStream_t st);
void launch_anchor_grid(const float *d_in, float *d_recon, uint16_t *d_codes, float *d_outval, int N,
                        const uint32_t *dims, const uint64_t *offs, uint32_t step, uint64_t count, hipStream_t st);
void launch_anchor_scatter(const float *d_vals, float *d_out, int N, const uint32_t *dims, const uint64_t *offs,
                           uint32_t step, uint64_t count, hipStream_t st);
void launch_first_point(const float *d_in, float *d_recon, uint16_t *d_codes, float *d_outval, double eb,
                        double recip, int radius, uint32_t *d_err, hipStream_t st, const LevelEB *d_eb = nullptr);
// out[0] of a decoder without anchors: code 0 recovered from prediction 0, or
// the first unpredictable value.
void launch_first_value(const uint16_t *d_codes, const float *d_unpred, uint64_t n_unpred, int radius, double eb,
                        float *d_out, hipStream_t st);
void launch_interp_pass(const PassDesc &d, const float *d_in, float *d_recon, uint16_t *d_codes, float *d_outval,
                        double *d_loss, uint32_t *d_err, hipStream_t st);
void launch_recover_pass(const PassDesc &d, float *d_data, const uint16_t *d_codes, const float *d_unpred_at,
                         hipStream_t st);
// d_freq holds kHistSlots words: the 65536 counts, then a flag set when a
// symbol other than 0 falls outside the window [hist_window_lo(radius),
// + kHistWin).  cleared: launch_hist_clear already ran on st for the arrays.
constexpr int kHistWin = 4096;
constexpr int kHistSlots = 65537;
int hist_window_lo(int radius);
void launch_hist(const uint16_t *d_codes, size_t n, uint32_t *d_freq, uint32_t *d_first, uint32_t *d_chunk_zeros,
                 int radius, hipStream_t st, bool cleared = false);
// One readback block after the histogram: counts and first positions of the
// window, then freq[0], first[0], the outside flag, *d_err and the u64 at
// d_scalar0 (two words).
constexpr int kHistPackFreq0 = 2 * kHistWin;
constexpr int kHistPackWords = kHistPackFreq0 + 6;
void launch_hist_pack(const uint32_t *d_freq, const uint32_t *d_first, int radius, const uint64_t *d_scalar0,
                      const uint32_t *d_err, uint32_t *d_pack, hipStream_t st, const double *d_eb = nullptr);
// words 0-1: *d_bits, word 2: *d_err (out may be coherent pinned host memory)
void launch_report(const uint64_t *d_bits, const uint32_t *d_err, uint32_t *out, hipStream_t st);
void launch_hist_clear(uint32_t *d_freq, uint32_t *d_first, hipStream_t st);
void launch_zero_scan(const uint32_t *d_chunk_zeros, uint64_t *d_chunk_base, uint32_t nchunks, uint64_t *d_tiles,
                      uint64_t *d_total, hipStream_t st);
void launch_unpred_gather(const uint16_t *d_codes, size_t n, const float *d_outval, const uint32_t *d_chunk_zeros,
                          const uint64_t *d_chunk_base, float *d_unpred, hipStream_t st);
// Huffman encode (faz_encode.hip).  tab holds one entry per symbol - offset
// (encode_entry).  When every present symbol but 0 lies in table indices
// [win, win + nwin) with nwin <= encode_lds_entries(), that window is staged
// in LDS and e0 is symbol 0's entry; nwin = 0 reads the whole table from
// global memory.
// Entry of a codeword of L bits (MSB-first in the top L bits of code): when
// every codeword of the table has at most 32 bits (narrow), the codeword
// left-aligned in the low word and L in the high word, so the packer ORs it
// in with one shift; otherwise the codeword right-aligned in bits 8..63 and L
// (<= 56) in bits 0..7.
__host__ __device__ inline uint64_t encode_entry(uint64_t code, uint32_t L, bool narrow) {
    if (!L) return 0;
    return narrow ? ((uint64_t) L << 32) | (uint32_t) (code >> 32) : ((code >> (64 - L)) << 8) | L;
}
struct EncodeWindow {
    uint32_t win, nwin;
    uint64_t e0;
    uint32_t max_len = 56;  // longest codeword (runs of codes assemble in registers up to 32)
};
uint32_t encode_lds_entries();
void launch_encode_sizes(const uint16_t *d_codes, size_t n, const uint64_t *d_tab, const EncodeWindow &w,
                         int32_t offset, uint64_t *d_chunk_bits, hipStream_t st);
// The encoder's chunk bit offsets and total (launch_scan_u64), zeroing the
// output words its packer ORs into when every codeword has at most 32 bits.
void launch_encode_offsets(const uint64_t *d_bits, uint64_t *d_off, uint32_t nchunks, uint64_t *d_tiles,
                           uint64_t *d_total, uint32_t *d_words, hipStream_t st);
void launch_scan_u64(const uint64_t *d_in, uint64_t *d_out, uint32_t count, uint64_t *d_tiles, uint64_t *d_total,
                     hipStream_t st);
// tiles of the two-level scans (scratch of this many uint64 per scan)
uint32_t scan_tiles(uint32_t count);
void launch_encode_pack(const uint16_t *d_codes, size_t n, const uint64_t *d_tab, const EncodeWindow &w,
                        int32_t offset, const uint64_t *d_chunk_off, uint32_t *d_words, hipStream_t st);
void launch_fill_u32(uint32_t *d, size_t count, uint32_t value, hipStream_t st);
// One-pass encode (codewords of at most 32 bits): d_status (one u64 per
// chunk) must be zero; writes the bitstream words (no zero fill needed),
// per-chunk bit offsets, the bit total, and sets bit 2 of *d_err if a
// look-back wait ran out.
void launch_encode_onepass(const uint16_t *d_codes, size_t n, const uint64_t *d_tab, const EncodeWindow &w,
                           int32_t offset, uint32_t *d_words, uint64_t *d_status, uint64_t *d_chunk_off,
                           uint32_t *d_head_bits, uint64_t *d_total_bits, uint32_t *d_err, hipStream_t st);
void launch_unpred_scatter(const uint16_t *d_codes, size_t n, const float *d_unpred, const uint32_t *d_chunk_zeros,
                           const uint64_t *d_chunk_base, float *d_unpred_at, hipStream_t st,
                           uint64_t count = ~0ull);
// words 0-1: *d_symbols, 2: *d_dec_err, 3: *d_rec_err, 4-5: *d_zeros
void launch_decode_report(const uint64_t *d_symbols, const uint32_t *d_dec_err, const uint64_t *d_zeros,
                          const uint32_t *d_rec_err, uint32_t *out, hipStream_t st);

// HuffmanEncoder::decode on the device (faz_decode.hip): num symbols into
// d_out on ctx->stream, each symbol + the tree's offset checked to lie in
// [0, limit); returns a faz_status with the stream drained.  With d_buf (the
// section in device memory) the host buf needs only the tree and byte count.
// run_count > 0: the symbols are expected to open with run_count copies of
// run_value (checked on the device, used to skip their decoding).
// With defer set the checks are left to the caller (no final wait): the
// decoded symbol count (>= num expected) at defer->total and the error flag
// (0 expected) at defer->err, both on the device, or null when already
// checked.  The output is written within [0, num) whatever the stream holds.
struct DecodeCheck {
    const uint64_t *total = nullptr;
    const uint32_t *err = nullptr;
};
// zero_counts (ceil(num / kHistChunk) words): the zero symbols of every
// kHistChunk outputs, counted by the write pass (the outlier ranks).
template<class OUT>
int gpu_huffman_decode(faz_gpu_ctx *ctx, const uint8_t *buf, size_t n, size_t num, uint32_t limit, OUT *d_out,
                       size_t *used, const uint8_t *d_buf = nullptr, uint32_t run_value = 0,
                       uint64_t run_count = 0, DecodeCheck *defer = nullptr, uint32_t *zero_counts = nullptr);

// outval[i] = the unpredictable value of zero code i, for n 32-bit codes on
// the device (faz_decode.hip); uploads the U host values first.
int gpu_scatter_unpred_u32(faz_gpu_ctx *ctx, const uint32_t *d_codes, uint64_t n, const float *h_unpred, uint64_t U,
                           float *d_outval);

// Device address of the allocation-free encoder's scratch array on `device`
// (hipGetSymbolAddress, once per device; faz_noalloc.hip).  Call with the
// device current.
hipError_t noalloc_init(int device);
size_t noalloc_scratch_bytes();

// Chunk geometry of the histogram / outlier / encode kernels.
constexpr uint32_t kHistThreads = 256;
constexpr uint32_t kHistPerThread = 64;
constexpr uint32_t kHistChunk = kHistThreads * kHistPerThread;  // codes per workgroup
constexpr uint32_t kEncThreads = 256;
constexpr uint32_t kEncPerThread = 16;
constexpr uint32_t kEncChunk = kEncThreads * kEncPerThread;

}  // namespace faz
