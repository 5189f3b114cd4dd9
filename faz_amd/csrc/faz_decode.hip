// Huffman decode on gfx950: HuffmanEncoder::load + decode (include/QoZ/
// encoder/HuffmanEncoder.hpp:236-296) for the bitstreams the reference
// writes (one MSB-first stream, no chunk index).
//
// The stream is cut into chunks of kChunkBits bits, one thread each.  The
// first codeword of chunk c starts at an entry offset e_c < L (L = the longest
// codeword) past the chunk's boundary, and e_{c+1} is a function x_c of e_c.
// Pass 1 tabulates x_c and the chunk's symbol count for every entry offset:
// it decodes from offset 0, and every other offset decodes only until it
// lands on a codeword start of an earlier trajectory (a Huffman decoder falls
// into step within a few codewords, except inside long runs of one codeword,
// where two or three phases persist and are decoded in full).  Pass 2 composes
// the maps (a parallel prefix inside each group of kDecThreads chunks, then
// one walk over the groups) to get every e_c from e_0 = 0; the counts at the
// true entries are scanned into output offsets, and pass 3 decodes each chunk
// again, writing its symbols.  Passes 1 and 3 stage their workgroup's slice
// of the stream in LDS.
//
// Codewords of up to kTabBits bits come from one table lookup (LDS), longer
// ones through second-level tables of kSubBits bits; multi-codeword tables
// count (pass 1) and write (pass 3) every whole codeword of a window at once.
#include "faz_gpu.h"
#include "faz_ctx.h"
#include "faz_internal.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <functional>
#include <type_traits>
#include <vector>

namespace faz {
namespace {

constexpr int kTabBits = 12;
constexpr uint32_t kChunkBits = 1024;
// parts of the stream uploaded one after another, each followed by its map launch
constexpr int kUploadParts = 4;
constexpr uint32_t kDecThreads = 256;
// stream words a workgroup stages: its chunks, then a margin for the last
// codeword (at most 64 bits) and the 64-bit window read past it
constexpr uint32_t kGroupWords = kDecThreads * kChunkBits / 32;
constexpr uint32_t kStageWords = kGroupWords + 5;
// one padding word per 32 keeps lanes that read the same word of their own
// chunk on different banks
constexpr uint32_t kStageSlots = kStageWords + kStageWords / 32 + 1;
// codeword starts of a trajectory are kept for its first kMaskBits bits
constexpr uint32_t kMaskBits = 128;
// The write pass decodes quarter chunks: the map pass records, for every
// entry offset, where the trajectory's first codeword at or after each of
// the chunk's three inner quarter boundaries starts and how many codewords
// lie before it.
constexpr uint32_t kQuarters = 4;
constexpr uint32_t kQuarterBits = kChunkBits / kQuarters;
// the write pass stages a quarter of the map pass's words per workgroup
constexpr uint32_t kWGroupWords = kGroupWords / kQuarters;
constexpr uint32_t kWStageWords = kWGroupWords + 5;
constexpr uint32_t kWStageSlots = kWStageWords + kWStageWords / 32 + 1;
// Lookup entries.  Primary table (kTabBits bits, LDS) and second-level
// tables (kSubBits bits each, global memory, one per tree node reached after
// kTabBits + j * kSubBits bits):
//   leaf node (bits 0..23) | codeword bits still to consume << 24  (bit 31 clear)
//   kLongFlag | index of the next second-level table
//   kLongFlag | kWalkFlag | tree node: continue bit by bit (tables capped)
//   kOffTree: the bits leave the tree (malformed tree)
constexpr int kSubBits = 8;
constexpr uint32_t kLongFlag = 0x80000000u;
constexpr uint32_t kWalkFlag = 0x40000000u;
constexpr uint32_t kOffTree = 0xFFFFFFFFu;
// second-level entries allowed before deeper nodes fall back to the walk
constexpr size_t kMaxSubEntries = (size_t) 1 << 22;

struct Tree {
    const int32_t *left, *right, *sym;
    const uint8_t *leaf;
    const uint32_t *sub;  // second-level tables
};

__device__ __forceinline__ uint32_t slot(uint32_t i) { return i + (i >> 5); }

// Multi-codeword entries for counting (ctab, one per kTabBits-bit window):
// bits 0..11 the window offsets where a codeword starts, 12..15 how many
// whole codewords the window holds (0: the first one is longer than the
// window or leaves the tree), 16..20 the bits they take.
__device__ __forceinline__ uint32_t cnt_of(uint32_t e) { return (e >> 12) & 15u; }
__device__ __forceinline__ uint32_t used_of(uint32_t e) { return (e >> 16) & 31u; }

struct BitReader {
    const uint32_t *lw;     // staged words (byte order swapped), see slot()
    uint64_t base;          // stream word index of lw's first word
    uint32_t staged = kStageWords;  // words in lw
    const uint32_t *words;  // the stream as 32-bit words (bytes in stream order)
    uint64_t nwords;
    uint64_t buf;           // next bits, MSB first
    int have;               // valid bits in buf
    uint64_t next;          // next word to load

    __device__ uint32_t word(uint64_t i) const {
        const uint64_t li = i - base;
        if (li < staged) return lw[slot((uint32_t) li)];
        return i < nwords ? __builtin_bswap32(words[i]) : 0u;
    }
    __device__ void start(uint64_t pos) {
        next = pos / 32;
        const uint32_t skip = (uint32_t) (pos % 32);
        buf = (uint64_t) word(next) << 32 | (uint64_t) word(next + 1);
        next += 2;
        buf <<= skip;
        have = 64 - (int) skip;
    }
    __device__ void refill() {
        if (have <= 32) {
            buf |= (uint64_t) word(next++) << (32 - have);
            have += 32;
        }
    }
    __device__ uint32_t peek(int k) const { return (uint32_t) (buf >> (64 - k)); }
    __device__ void drop(int k) {
        buf <<= k;
        have -= k;
    }
};

// One codeword at the reader: the leaf node, or -1 when the bits leave the
// tree.  Always advances pos.
__device__ __forceinline__ int32_t decode_one(BitReader &r, const uint32_t *tab, const Tree &T, uint64_t &pos) {
    r.refill();
    uint32_t e = tab[r.peek(kTabBits)];
    int bits = kTabBits;
    int depth = 0;
    for (;;) {
        if (e == kOffTree) {
            r.drop(1);
            pos++;
            return -1;
        }
        if (!(e & kLongFlag)) {
            const int len = (int) (e >> 24);
            r.drop(len);
            pos += (uint64_t) len;
            return (int32_t) (e & 0xFFFFFFu);
        }
        r.drop(bits);
        pos += (uint64_t) bits;
        depth += bits;
        if (e & kWalkFlag) break;
        // codewords are at most 64 bits (the reference keeps them in 64-bit words)
        if (depth >= 64) return -1;
        r.refill();
        e = T.sub[(e & 0x3FFFFFFFu) + r.peek(kSubBits)];
        bits = kSubBits;
    }
    int32_t nd = (int32_t) (e & 0x3FFFFFFFu);
    for (; depth < 64; depth++) {
        r.refill();
        const uint32_t bit = r.peek(1);
        r.drop(1);
        pos++;
        nd = bit ? T.right[nd] : T.left[nd];
        if (nd < 0) return -1;
        if (T.leaf[nd]) return nd;
    }
    return -1;
}

// Loads the lookup table and the workgroup's slice of the stream into LDS.
__device__ __forceinline__ void stage(uint32_t *tab, uint32_t *lw, const uint32_t *__restrict__ gtab,
                                      const uint32_t *__restrict__ words, uint64_t nwords, uint64_t base) {
    for (uint32_t i = threadIdx.x; i < (1u << kTabBits); i += blockDim.x) tab[i] = gtab[i];
    for (uint32_t i = threadIdx.x; i < kStageWords; i += blockDim.x) {
        const uint64_t g = base + i;
        lw[slot(i)] = g < nwords ? __builtin_bswap32(words[g]) : 0u;
    }
    __syncthreads();
}

// A decoding position with the symbols counted so far: those that end inside
// the stream (a codeword cut by the stream's end is not a symbol).
struct Cursor {
    BitReader r;
    uint64_t pos;
    uint32_t k;
    // the first codeword start at or after each inner quarter boundary of
    // the chunk: byte q of qx its distance from boundary q, 16 bits q of qn
    // the codewords before it (boundaries past the trajectory's end: 0 and
    // the final count)
    uint64_t qnext;
    uint32_t qi, qx;
    uint64_t qn;
    __device__ void init(const BitReader &proto, uint64_t p, uint64_t chunk) {
        r = proto;
        r.start(p);
        pos = p;
        k = 0;
        qnext = chunk + kQuarterBits;
        qi = 0;
        qx = 0;
        qn = 0;
    }
    __device__ void qrecord(uint64_t start, uint32_t count) {
        qx |= (uint32_t) (start - qnext) << (8 * qi);
        qn |= (uint64_t) count << (16 * qi);
        qi++;
        qnext = qi < kQuarters - 1 ? qnext + kQuarterBits : ~0ull;
    }
    __device__ void qfinish() {
        for (; qi < kQuarters - 1; qi++) qn |= (uint64_t) k << (16 * qi);
    }
    __device__ void step(const uint32_t *tab, const Tree &T, uint64_t nbits) {
        const int32_t nd = decode_one(r, tab, T, pos);
        if (nd >= 0 && pos <= nbits) k++;
        if (pos >= qnext) qrecord(pos, k);
    }
    // The ctab entry of the next window when every codeword it holds starts
    // before lim (0 otherwise); consume() then takes them.
    __device__ uint32_t peek_multi(const uint32_t *ctab, uint64_t lim) {
        if (pos + kTabBits > lim) return 0;
        r.refill();
        const uint32_t e = ctab[r.peek(kTabBits)];
        return cnt_of(e) ? e : 0u;
    }
    __device__ void consume(uint32_t e) {
        const int used = (int) used_of(e);
        const uint64_t before = pos;
        const uint32_t kb = k;
        r.drop(used);
        pos += (uint64_t) used;
        k += cnt_of(e);
        if (pos >= qnext) {
            // the boundary falls inside the window: its first start there, or
            // the start after the window
            const uint32_t S = e & 0xFFFu, j0 = (uint32_t) (qnext - before);
            const uint32_t hi = S & ~((1u << j0) - 1u);
            if (hi) {
                const uint32_t j = (uint32_t) __builtin_ctz(hi);
                qrecord(before + j, kb + (uint32_t) __popc(S & ((1u << j) - 1u)));
            } else {
                qrecord(pos, k);
            }
        }
    }
};

// Quarter data of a trajectory that joins another one: the same starts,
// counts shifted by the codewords the two have counted differently.
__device__ __forceinline__ uint64_t qshift(uint64_t qn, int32_t delta) {
    uint64_t out = 0;
#pragma unroll
    for (uint32_t q = 0; q < kQuarters - 1; q++)
        out |= (uint64_t) (uint16_t) ((int32_t) ((qn >> (16 * q)) & 0xFFFFu) + delta) << (16 * q);
    return out;
}

// ORs the 12-bit start mask S, placed at offset q, into the 128-bit mask m.
__device__ __forceinline__ void or_mask(uint64_t *m, uint32_t S, uint64_t q) {
    if (q < 64) {
        m[0] |= (uint64_t) S << q;
        if (q > 52) m[1] |= (uint64_t) S >> (64 - q);
    } else if (q < 128) {
        m[1] |= (uint64_t) S << (q - 64);
    }
}
// The 12 bits of m from offset q on (zero past bit 127).
__device__ __forceinline__ uint32_t mask_window(const uint64_t *m, uint64_t q) {
    uint64_t w = 0;
    if (q < 64) {
        w = m[0] >> q;
        if (q > 52) w |= m[1] << (64 - q);
    } else if (q < 128) {
        w = m[1] >> (q - 64);
    }
    return (uint32_t) w & 0xFFFu;
}

__device__ __forceinline__ bool mask_has(const uint64_t *m, uint32_t q) { return (m[q >> 6] >> (q & 63)) & 1u; }
__device__ __forceinline__ uint32_t mask_below(const uint64_t *m, uint32_t q) {
    if (q < 64) return (uint32_t) __popcll(m[0] & ((1ull << q) - 1));
    if (q >= 128) return (uint32_t) (__popcll(m[0]) + __popcll(m[1]));
    return (uint32_t) (__popcll(m[0]) + __popcll(m[1] & ((1ull << (q - 64)) - 1)));
}

// A stream that opens with `count` copies of one codeword (the anchors of an
// interpolation engine buffer are its first codes, all 0): bits [0, bits)
// repeat the len-bit codeword.  *ok is set by huff_run_check_kernel when the
// stream really holds that run; chunks wholly inside it then need no
// trajectory search (the codeword starts are the multiples of len) and no
// decoding.
struct RunPrefix {
    uint64_t bits;       // 0: no run
    uint32_t len;        // codeword length
    uint32_t value;      // output value of the codeword
    uint64_t code;       // codeword, right-aligned
    const uint32_t *ok;  // device flag: 1 when the bits match
};

__device__ __forceinline__ bool run_chunk(const RunPrefix &run, uint64_t b) {
    return run.bits && b + kChunkBits <= run.bits && *run.ok;
}
// codeword starts of the run below position p (p <= run.bits)
__device__ __forceinline__ uint64_t run_starts_below(const RunPrefix &run, uint64_t p) {
    return (p + run.len - 1) / run.len;
}
// first codeword start at or after p (p <= run.bits)
__device__ __forceinline__ uint64_t run_next_start(const RunPrefix &run, uint64_t p) {
    return run_starts_below(run, p) * run.len;
}

// Clears *ok when some bit of [0, run.bits) differs from the repeated
// codeword; one thread per 32-bit stream word.
__global__ void huff_run_check_kernel(const uint32_t *__restrict__ words, RunPrefix run, uint32_t *__restrict__ ok) {
    const uint64_t w = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t p0 = w * 32;
    if (p0 >= run.bits) return;
    const uint32_t got = __builtin_bswap32(words[w]);
    uint32_t want = 0, care = 0;
    uint32_t ph = (uint32_t) (p0 % run.len);
    for (uint32_t j = 0; j < 32; j++) {
        if (p0 + j < run.bits) {
            want |= (uint32_t) ((run.code >> (run.len - 1 - ph)) & 1u) << (31 - j);
            care |= 1u << (31 - j);
        }
        ph = ph + 1 == run.len ? 0 : ph + 1;
    }
    if ((got ^ want) & care) atomicAnd(ok, 0u);
}

// Pass 1: for chunk c and every entry offset o < K, the next chunk's entry
// offset (xmap, clamped to K - 1) and the symbols that start in the chunk
// (nmap).  A trajectory merges with an earlier one when it reaches one of its
// codeword starts; the starts of the first two trajectories are kept as bit
// masks over the chunk's first kMaskBits bits.
__global__ void __launch_bounds__(kDecThreads) huff_map_kernel(const uint32_t *__restrict__ words, uint64_t nwords,
                                                               uint64_t nbits, uint32_t nchunks,
                                                               const uint32_t *__restrict__ gtab,
                                                               const uint32_t *__restrict__ gctab, Tree T, uint32_t K,
                                                               uint8_t *__restrict__ xmap,
                                                               uint16_t *__restrict__ nmap, uint32_t wg0,
                                                               RunPrefix run, uint32_t *__restrict__ qxmap,
                                                               uint64_t *__restrict__ qnmap) {
    __shared__ uint32_t tab[1 << kTabBits];
    __shared__ uint32_t ctab[1 << kTabBits];
    __shared__ uint32_t lw[kStageSlots];
    const uint32_t wg = blockIdx.x + wg0;  // launches cover consecutive ranges of workgroups
    const uint64_t base = (uint64_t) wg * kGroupWords;
    for (uint32_t i = threadIdx.x; i < (1u << kTabBits); i += blockDim.x) ctab[i] = gctab[i];
    stage(tab, lw, gtab, words, nwords, base);
    const uint32_t c = wg * kDecThreads + threadIdx.x;
    if (c >= nchunks) return;
    BitReader proto;
    proto.lw = lw, proto.base = base, proto.words = words, proto.nwords = nwords;
    const uint64_t b = (uint64_t) c * kChunkBits;
    const uint64_t end = min(b + kChunkBits, nbits);
    const uint64_t nxt = b + kChunkBits;
    // masks are exact only where every symbol before kMaskBits ends inside the stream
    const bool masks = b + kMaskBits <= nbits;
    // the maps are offset-major (entry o of chunk c at o * nchunks + c), so
    // a wave's stores for one offset are contiguous
    auto put = [&](uint32_t o, uint64_t x, uint32_t n, uint32_t qx, uint64_t qn) {
        const uint64_t e = x > nxt ? x - nxt : 0;
        const uint64_t at = (uint64_t) o * nchunks + c;
        xmap[at] = (uint8_t) min<uint64_t>(e, K - 1);
        nmap[at] = (uint16_t) n;
        qxmap[at] = qx;
        qnmap[at] = qn;
    };
    if (run_chunk(run, b)) {
        // inside the verified run every entry is mapped to the run's own
        // exit: only the true entry is ever composed
        const uint64_t ex = run_next_start(run, b + kChunkBits);
        const uint64_t k0 = run_starts_below(run, b);
        const uint32_t n = (uint32_t) (run_starts_below(run, b + kChunkBits) - k0);
        uint32_t qx = 0;
        uint64_t qn = 0;
        for (uint32_t q = 0; q < kQuarters - 1; q++) {
            const uint64_t bq = b + (uint64_t) (q + 1) * kQuarterBits;
            qx |= (uint32_t) (run_next_start(run, bq) - bq) << (8 * q);
            qn |= (uint64_t) (run_starts_below(run, bq) - k0) << (16 * q);
        }
        for (uint32_t o = 0; o < K; o++) put(o, ex, n, qx, qn);
        return;
    }
    uint64_t mA[2] = {0, 0}, mB[2] = {0, 0};
    uint64_t xA, xB = 0, qnA, qnB = 0;
    uint32_t nA, nB = 0, qxA, qxB = 0;
    bool hasB = false;
    {
        Cursor A;
        A.init(proto, b, b);
        while (A.pos < end) {
            const uint64_t q = A.pos - b;
            const uint32_t e = A.peek_multi(ctab, end);
            if (e) {
                or_mask(mA, e & 0xFFFu, q);
                A.consume(e);
                continue;
            }
            if (q < kMaskBits) mA[q >> 6] |= 1ull << (q & 63);
            A.step(tab, T, nbits);
        }
        A.qfinish();
        xA = A.pos, nA = A.k, qxA = A.qx, qnA = A.qn;
    }
    put(0, xA, nA, qxA, qnA);
    // a trajectory that joins A (or B) at p after `at` codewords of its own
    auto join_A = [&](uint32_t o, uint32_t at, uint32_t p) {
        const uint32_t below = mask_below(mA, p);
        put(o, xA, at + nA - below, qxA, qshift(qnA, (int32_t) at - (int32_t) below));
    };
    auto join_B = [&](uint32_t o, uint32_t at, uint32_t p) {
        const uint32_t below = mask_below(mB, p);
        put(o, xB, at + nB - below, qxB, qshift(qnB, (int32_t) at - (int32_t) below));
    };
    for (uint32_t o = 1; o < K; o++) {
        if (masks && mask_has(mA, o)) {
            join_A(o, 0, o);
            continue;
        }
        if (masks && hasB && mask_has(mB, o)) {
            join_B(o, 0, o);
            continue;
        }
        Cursor O;
        O.init(proto, b + o, b);
        uint64_t mO[2] = {0, 0};
        bool merged = false;
        while (O.pos < end) {
            const uint64_t q = O.pos - b;
            const bool watch = masks && q < kMaskBits;
            const uint32_t e = O.peek_multi(ctab, end);
            if (e) {
                if (watch) {
                    // the first codeword start of the window on a start of A or B
                    const uint32_t S = e & 0xFFFu;
                    const uint32_t hitA = mask_window(mA, q) & S;
                    const uint32_t hitB = hasB ? mask_window(mB, q) & S : 0u;
                    const uint32_t hit = hitA | hitB;
                    if (hit) {
                        const uint32_t j = (uint32_t) __builtin_ctz(hit);
                        const uint32_t at = O.k + (uint32_t) __popc(S & ((1u << j) - 1u));
                        const uint32_t p = (uint32_t) q + j;
                        if ((hitA >> j) & 1u)
                            join_A(o, at, p);
                        else
                            join_B(o, at, p);
                        merged = true;
                        break;
                    }
                    or_mask(mO, S, q);
                }
                O.consume(e);
                continue;
            }
            if (watch) {
                if (mask_has(mA, (uint32_t) q)) {
                    join_A(o, O.k, (uint32_t) q);
                    merged = true;
                    break;
                }
                if (hasB && mask_has(mB, (uint32_t) q)) {
                    join_B(o, O.k, (uint32_t) q);
                    merged = true;
                    break;
                }
                mO[q >> 6] |= 1ull << (q & 63);
            }
            O.step(tab, T, nbits);
        }
        if (merged) continue;
        O.qfinish();
        put(o, O.pos, O.k, O.qx, O.qn);
        if (!hasB) {
            hasB = true;
            xB = O.pos, nB = O.k, qxB = O.qx, qnB = O.qn;
            mB[0] = mO[0], mB[1] = mO[1];
        }
    }
}

// Pass 2a: inside each group of kDecThreads chunks, the inclusive
// compositions M_i = x_i o ... o x_0 of the group's maps (Hillis-Steele in
// LDS, one thread per chunk); gmap receives them and grp the group's whole
// composition.
__global__ void __launch_bounds__(kDecThreads) huff_group_compose_kernel(const uint8_t *__restrict__ xmap,
                                                                         uint32_t nchunks, uint32_t K,
                                                                         uint8_t *__restrict__ gmap,
                                                                         uint8_t *__restrict__ grp) {
    __shared__ uint8_t buf[2][kDecThreads * 64];
    const uint32_t i = threadIdx.x;
    const uint64_t c0 = (uint64_t) blockIdx.x * kDecThreads;
    const uint32_t n = (uint32_t) min<uint64_t>(kDecThreads, nchunks - c0);
    for (uint32_t j = i; j < n * K; j += kDecThreads) {  // offset-major source, chunk-major in LDS
        const uint32_t o = j / n, ci = j - o * n;
        buf[0][ci * K + o] = xmap[(uint64_t) o * nchunks + c0 + ci];
    }
    __syncthreads();
    int cur = 0;
    for (uint32_t d = 1; d < n; d *= 2) {
        if (i < n) {
            for (uint32_t o = 0; o < K; o++) {
                uint8_t v = buf[cur][i * K + o];
                if (i >= d) v = buf[cur][i * K + buf[cur][(i - d) * K + o]];
                buf[cur ^ 1][i * K + o] = v;
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    for (uint32_t j = i; j < n * K; j += kDecThreads) gmap[c0 * K + j] = buf[cur][j];
    for (uint32_t o = i; o < K; o += kDecThreads) grp[(uint64_t) blockIdx.x * K + o] = buf[cur][(n - 1) * K + o];
}

// Pass 2b (one workgroup): the entry offset of every group, walking the
// group compositions from entry 0, staged through LDS in batches.
__global__ void __launch_bounds__(1024) huff_group_entry_kernel(const uint8_t *__restrict__ grp, uint32_t ngroups,
                                                                uint32_t K, uint8_t *__restrict__ gentry) {
    constexpr uint32_t kBuf = 48 * 1024;
    __shared__ uint8_t buf[kBuf];
    const uint32_t per = kBuf / K;
    uint32_t e = 0;
    for (uint32_t g0 = 0; g0 < ngroups; g0 += per) {
        const uint32_t n = min(per, ngroups - g0);
        for (uint32_t j = threadIdx.x; j < n * K; j += blockDim.x) buf[j] = grp[(uint64_t) g0 * K + j];
        __syncthreads();
        if (threadIdx.x == 0) {
            for (uint32_t j = 0; j < n; j++) {
                gentry[g0 + j] = (uint8_t) e;
                e = buf[j * K + e];
            }
        }
        __syncthreads();
    }
}

// Entry of every chunk (the group's entry, carried through the group's
// earlier chunks), and at that entry the first codeword start of each
// quarter of the chunk and the symbols of each quarter.
__global__ void huff_entry_kernel(const uint8_t *__restrict__ gmap, const uint8_t *__restrict__ gentry,
                                  const uint16_t *__restrict__ nmap, const uint32_t *__restrict__ qxmap,
                                  const uint64_t *__restrict__ qnmap, uint32_t nchunks, uint32_t K,
                                  uint64_t *__restrict__ start, uint32_t *__restrict__ count) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    const uint32_t ge = gentry[c / kDecThreads];
    const uint32_t e = (c % kDecThreads) ? gmap[(uint64_t) (c - 1) * K + ge] : ge;
    const uint64_t at = (uint64_t) e * nchunks + c;
    const uint64_t b = (uint64_t) c * kChunkBits;
    const uint32_t qx = qxmap[at];
    const uint64_t qn = qnmap[at];
    uint32_t before = 0;
#pragma unroll
    for (uint32_t q = 0; q < kQuarters; q++) {
        const uint32_t upto = q + 1 < kQuarters ? (uint32_t) ((qn >> (16 * q)) & 0xFFFFu) : (uint32_t) nmap[at];
        start[(uint64_t) kQuarters * c + q] = q ? b + (uint64_t) q * kQuarterBits + ((qx >> (8 * (q - 1))) & 0xFFu)
                                                : b + e;
        count[(uint64_t) kQuarters * c + q] = upto - before;
        before = upto;
    }
}

// Pass 3: one thread per quarter chunk (chunk t / 4, quarter t % 4) writes
// the symbols of the codewords that start in its quarter from start[t] on,
// at offset[t] (those below num), each sym + sym_offset, which must lie in
// [0, limit) unless limit is 0;
// err[0] = 1 otherwise or when the bits leave the tree.  MULTI: wtab (LDS)
// gives up to three output values per kTabBits-bit window (bits 0..47, in
// order), their count (48..49, 0: decode one codeword through the primary
// table, read from global memory) and their bits (52..56).
template<class OUT, bool MULTI>
__global__ void __launch_bounds__(kDecThreads) huff_write_kernel(const uint32_t *__restrict__ words, uint64_t nwords,
                                                                 uint64_t nbits, uint32_t nchunks,
                                                                 const uint32_t *__restrict__ gtab,
                                                                 const uint64_t *__restrict__ gwtab, Tree T,
                                                                 const uint64_t *__restrict__ start,
                                                                 const uint64_t *__restrict__ offset,
                                                                 int32_t sym_offset, uint32_t limit, uint64_t num,
                                                                 OUT *__restrict__ out, uint32_t *__restrict__ err,
                                                                 RunPrefix run, uint32_t *__restrict__ zc) {
    __shared__ uint32_t tab[MULTI ? 1 : (1 << kTabBits)];
    __shared__ uint64_t wtab[MULTI ? (1 << kTabBits) : 1];
    __shared__ uint32_t lw[kWStageSlots];
    const uint64_t base = (uint64_t) blockIdx.x * kWGroupWords;
    for (uint32_t i = threadIdx.x; i < (1u << kTabBits); i += blockDim.x) {
        if (MULTI)
            wtab[i] = gwtab[i];
        else
            tab[i] = gtab[i];
    }
    for (uint32_t i = threadIdx.x; i < kWStageWords; i += blockDim.x) {
        const uint64_t g = base + i;
        lw[slot(i)] = g < nwords ? __builtin_bswap32(words[g]) : 0u;
    }
    __syncthreads();
    const uint32_t *dtab = MULTI ? gtab : tab;
    const uint64_t t = (uint64_t) blockIdx.x * kDecThreads + threadIdx.x;
    if (t >= (uint64_t) kQuarters * nchunks) return;
    const uint32_t c = (uint32_t) (t / kQuarters), q = (uint32_t) (t % kQuarters);
    const uint64_t end = min((uint64_t) c * kChunkBits + (uint64_t) (q + 1) * kQuarterBits, nbits);
    uint64_t pos = start[t], o = offset[t];
    BitReader r;
    r.lw = lw, r.base = base, r.words = words, r.nwords = nwords;
    r.staged = kWStageWords;
    r.start(pos);
    // symbols leave in 16-byte stores once out + o is 16-byte aligned (each
    // lane writes its own range: wide stores keep the partial-line writes
    // that reach the L2 few)
    constexpr int P = 16 / sizeof(OUT), H = P / 2;
    constexpr int W = 8 * sizeof(OUT);
    uint64_t acc0 = 0, acc1 = 0;
    int na = 0;
    // zc: zero symbols per kHistChunk outputs (the outlier ranks), counted in
    // a register and added once per chunk the thread's range touches
    uint32_t zn = 0, zh = 0;
    auto push = [&](uint32_t v) {
        if (zc && v == 0) {
            const uint32_t h = (uint32_t) (o / kHistChunk);
            if (h != zh && zn) {
                atomicAdd(&zc[zh], zn);
                zn = 0;
            }
            zh = h;
            zn++;
        }
        if (na == 0 && (reinterpret_cast<uintptr_t>(out + o) & 15u) != 0) {
            out[o++] = (OUT) v;
            return;
        }
        const uint64_t x = (uint64_t) (std::make_unsigned_t<OUT>) (OUT) v;
        if (na < H)
            acc0 |= x << (W * na);
        else
            acc1 |= x << (W * (na - H));
        na++;
        o++;
        if (na == P) {
            *reinterpret_cast<uint4 *>(out + (o - P)) =
                make_uint4((uint32_t) acc0, (uint32_t) (acc0 >> 32), (uint32_t) acc1, (uint32_t) (acc1 >> 32));
            acc0 = acc1 = 0;
            na = 0;
        }
    };
    bool bad = false;
    if (run_chunk(run, (uint64_t) c * kChunkBits)) {
        // the run's codewords: one value, no decoding
        for (uint64_t k = run_starts_below(run, end) - run_starts_below(run, pos); k > 0 && o < num; k--)
            push(run.value);
        pos = end;
    }
    while (pos < end && o < num) {
        if (MULTI && pos + kTabBits <= end) {
            r.refill();
            const uint64_t e = wtab[r.peek(kTabBits)];
            const uint32_t cnt = (uint32_t) (e >> 48) & 3u;
            if (cnt && o + cnt <= num) {
                const int used = (int) ((e >> 52) & 31u);
                r.drop(used);
                pos += (uint64_t) used;
                push((uint32_t) e & 0xFFFFu);
                if (cnt > 1) push((uint32_t) (e >> 16) & 0xFFFFu);
                if (cnt > 2) push((uint32_t) (e >> 32) & 0xFFFFu);
                continue;
            }
        }
        const int32_t k = decode_one(r, dtab, T, pos);
        if (k < 0) {
            bad = true;
            break;
        }
        const uint32_t v = (uint32_t) (T.sym[k] + sym_offset);
        if (limit && v >= limit) {
            bad = true;
            break;
        }
        push(v);
    }
    for (int j = 0; j < na; j++) out[o - na + j] = (OUT) (j < H ? acc0 >> (W * j) : acc1 >> (W * (j - H)));
    if (zn) atomicAdd(&zc[zh], zn);
    if (bad) atomicOr(err, 1u);
}

template<class OUT>
__global__ void fill_kernel(OUT *__restrict__ out, uint64_t n, OUT v) {
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t) gridDim.x * blockDim.x)
        out[i] = v;
}

// The saved tree (HuffmanEncoder.hpp:119-137, 569-633) as child and symbol
// arrays indexed by serialised slot (slot 0 is the root; child slot 0 means
// none).  Returns false on a malformed tree; sets the bitstream's byte count
// and its position in buf.
bool parse_tree(const uint8_t *buf, size_t n, int32_t &offset, std::vector<int32_t> &left,
                std::vector<int32_t> &right, std::vector<int32_t> &sym, std::vector<uint8_t> &leaf,
                uint64_t &enc_bytes, size_t &bits_at) {
    if (n < 12) return false;
    memcpy(&offset, buf, 4);
    const unsigned nc = ((unsigned) buf[4] << 24) | ((unsigned) buf[5] << 16) | ((unsigned) buf[6] << 8) | buf[7];
    if (nc == 0 || nc > (1u << 23)) return false;
    const size_t w = nc <= 256 ? 1 : (nc <= 65536 ? 2 : 4);
    const size_t treeBytes = 1 + 2 * nc * w + nc * 4 + nc;
    if (12 + treeBytes + 8 > n) return false;
    const uint8_t *tb = buf + 12;
    auto idx = [&](const uint8_t *arr, unsigned i) -> unsigned {
        if (w == 1) return arr[i];
        if (w == 2) {
            uint16_t v;
            memcpy(&v, arr + 2 * i, 2);
            return v;
        }
        uint32_t v;
        memcpy(&v, arr + 4 * i, 4);
        return v;
    };
    const uint8_t *L = tb + 1, *R = tb + 1 + nc * w, *C = tb + 1 + 2 * nc * w, *F = C + nc * 4;
    left.assign(nc, -1), right.assign(nc, -1), sym.assign(nc, 0), leaf.assign(nc, 0);
    for (unsigned i = 0; i < nc; i++) {
        memcpy(&sym[i], C + 4 * i, 4);
        leaf[i] = F[i] ? 1 : 0;
    }
    if (nc > 256) leaf[0] = 0;  // the reference rebuilds the root as an inner node
    for (unsigned i = 0; i < nc; i++) {
        if (leaf[i]) continue;
        const unsigned l = idx(L, i), r = idx(R, i);
        if (l >= nc || r >= nc) return false;
        left[i] = l ? (int32_t) l : -1;
        right[i] = r ? (int32_t) r : -1;
    }
    // a serialised Huffman tree is a tree: every slot reachable from the
    // root is reached once (no shared children, no cycles) and no path is
    // deeper than 64 bits
    if (!leaf[0]) {
        std::vector<uint8_t> seen(nc, 0);
        std::vector<std::pair<int32_t, uint32_t>> stack{{0, 0u}};
        seen[0] = 1;
        while (!stack.empty()) {
            const auto [nd, d] = stack.back();
            stack.pop_back();
            if (leaf[nd]) continue;
            if (d >= 64) return false;
            for (const int32_t ch: {left[nd], right[nd]}) {
                if (ch < 0) continue;
                if (seen[ch]) return false;
                seen[ch] = 1;
                stack.push_back({ch, d + 1});
            }
        }
    }
    memcpy(&enc_bytes, tb + treeBytes, 8);
    bits_at = 12 + treeBytes + 8;
    return bits_at + enc_bytes <= n;
}

}  // namespace

// Host check of a saved Huffman section (no GPU needed): the tree parses,
// is a tree, and its bitstream fits in n bytes.
extern "C" faz_status faz_huffman_check_section(const uint8_t *buf, size_t n, size_t *nodes, size_t *section_bytes) {
    if (!buf) return FAZ_ERR_INVALID;
    int32_t offset;
    std::vector<int32_t> left, right, sym;
    std::vector<uint8_t> leaf;
    uint64_t enc = 0;
    size_t at = 0;
    if (!parse_tree(buf, n, offset, left, right, sym, leaf, enc, at)) return FAZ_ERR_FORMAT;
    if (nodes) *nodes = sym.size();
    if (section_bytes) *section_bytes = at + enc;
    return FAZ_OK;
}

// HuffmanEncoder::decode on the device: num symbols (each plus the tree's
// offset, which must land in [0, limit) unless limit is 0) into d_out on
// ctx->stream.  Returns with the stream drained; *used = bytes of the Huffman
// section.  d_buf, when given, is the same section in device memory: the
// bitstream is then copied on the device (buf needs only the tree and the
// byte count), otherwise it is uploaded from buf.  run_count > 0 announces
// that the symbols open with run_count copies of run_value (checked on the
// device; a stream without the run decodes the general way).
template<class OUT>
int gpu_huffman_decode(faz_gpu_ctx *ctx, const uint8_t *buf, size_t n, size_t num, uint32_t limit, OUT *d_out,
                       size_t *used, const uint8_t *d_buf, uint32_t run_value, uint64_t run_count,
                       DecodeCheck *defer, uint32_t *zero_counts) {
    if (defer) *defer = DecodeCheck{};
    int32_t offset;
    std::vector<int32_t> left, right, sym;
    std::vector<uint8_t> leaf;
    uint64_t enc = 0;
    size_t at = 0;
    if (!parse_tree(buf, n, offset, left, right, sym, leaf, enc, at)) CTX_FAIL(FAZ_ERR_FORMAT, "Huffman section malformed");
    if (used) *used = at + enc;
    hipStream_t st = ctx->stream;
    if (!num) return FAZ_OK;
    if (leaf[0]) {  // one symbol, no bits (HuffmanEncoder.hpp:245-250)
        const uint32_t v = (uint32_t) (sym[0] + offset);
        if (limit && v >= limit) CTX_FAIL(FAZ_ERR_FORMAT, "Huffman symbol out of range");
        const uint64_t want = (num + 255) / 256;
        hipLaunchKernelGGL(fill_kernel<OUT>, dim3((unsigned) std::min<uint64_t>(want, 8192)), dim3(256), 0, st, d_out,
                           (uint64_t) num, (OUT) v);
        HIPCHECK(hipGetLastError());
        HIPCHECK(spin_wait(st));
        return FAZ_OK;
    }
    // primary table: every kTabBits-bit prefix walked down from the root;
    // second-level tables for the nodes below it
    const size_t nc = sym.size();
    std::vector<uint32_t> sub;
    // entry for the next `bits` bits (value v, MSB first) from node `from`
    auto entry = [&](int32_t from, uint32_t v, int bits, int32_t &stop) -> uint32_t {
        int32_t nd = from;
        for (int d = 0; d < bits; d++) {
            nd = ((v >> (bits - 1 - d)) & 1u) ? right[nd] : left[nd];
            if (nd < 0) return kOffTree;
            if (leaf[nd]) return (uint32_t) nd | (uint32_t) (d + 1) << 24;
        }
        stop = nd;  // an inner node after all `bits` bits
        return kLongFlag;
    };
    // builds the table of node nd (depth `depth`) and those below it
    std::function<uint32_t(int32_t, int)> build_sub = [&](int32_t nd, int depth) -> uint32_t {
        if (depth >= 64 || sub.size() + ((size_t) 1 << kSubBits) > kMaxSubEntries)
            return kLongFlag | kWalkFlag | (uint32_t) nd;
        const size_t at = sub.size();
        sub.resize(at + ((size_t) 1 << kSubBits));
        for (uint32_t v = 0; v < (1u << kSubBits); v++) {
            int32_t stop = -1;
            uint32_t e = entry(nd, v, kSubBits, stop);
            if (e == kLongFlag) e = build_sub(stop, depth + kSubBits);
            sub[at + v] = e;
        }
        return kLongFlag | (uint32_t) at;
    };
    // the primary table by a walk over the tree's first kTabBits levels in
    // code order (the order the second-level tables are built in): a leaf at
    // depth d fills its 2^(kTabBits - d) windows, an inner node at depth
    // kTabBits gets its second-level table, missing children leave kOffTree
    std::vector<uint32_t> tab((size_t) 1 << kTabBits, kOffTree);
    {
        struct Visit {
            int32_t nd;
            uint32_t code;
            int depth;
        };
        std::vector<Visit> st;
        st.reserve(2 * kTabBits + 2);
        st.push_back({0, 0u, 0});
        while (!st.empty()) {
            const Visit f = st.back();
            st.pop_back();
            if (f.depth > 0 && leaf[f.nd]) {
                const uint32_t lo = f.code << (kTabBits - f.depth), cnt = 1u << (kTabBits - f.depth);
                for (uint32_t v = lo; v < lo + cnt; v++) tab[v] = (uint32_t) f.nd | (uint32_t) f.depth << 24;
                continue;
            }
            if (f.depth == kTabBits) {
                tab[f.code] = build_sub(f.nd, kTabBits);
                continue;
            }
            if (right[f.nd] >= 0) st.push_back({right[f.nd], f.code << 1 | 1u, f.depth + 1});
            if (left[f.nd] >= 0) st.push_back({left[f.nd], f.code << 1, f.depth + 1});
        }
    }
    // multi-codeword windows: whole codewords from the window start on
    // (ctab for counting; wtab: up to three output values in [0, limit),
    // for 16-bit outputs), each codeword read through the primary table with
    // the window's remaining bits (a codeword that needs bits past the window
    // ends the run)
    constexpr bool kMulti = sizeof(OUT) == 2;
    std::vector<uint32_t> ctab(tab.size());
    std::vector<uint64_t> wtab(kMulti ? tab.size() : 0);
    for (uint32_t v = 0; v < tab.size(); v++) {
        uint32_t at = 0, cnt = 0, S = 0, wcnt = 0, wused = 0;
        uint64_t wv = 0;
        bool wopen = kMulti && limit && limit <= 65536u;
        while (at < (uint32_t) kTabBits) {
            const uint32_t e = tab[(v << at) & ((1u << kTabBits) - 1)];
            if (e & kLongFlag) break;  // longer than the window, or off the tree
            const uint32_t len = e >> 24;
            if (len > (uint32_t) kTabBits - at) break;
            const int32_t nd = (int32_t) (e & 0xFFFFFFu);
            S |= 1u << at;
            cnt++;
            if (wopen) {
                const uint32_t val = (uint32_t) (sym[nd] + offset);
                if (val < limit && wcnt < 3) {
                    wv |= (uint64_t) val << (16 * wcnt);
                    wcnt++;
                    wused += len;
                } else {
                    wopen = false;
                }
            }
            at += len;
        }
        ctab[v] = S | cnt << 12 | at << 16;
        if (kMulti) wtab[v] = wv | (uint64_t) wcnt << 48 | (uint64_t) wused << 52;
    }
    // longest codeword: the entry offsets of a chunk lie below it
    uint32_t maxlen = 0;
    {
        std::vector<std::pair<int32_t, uint32_t>> stack{{0, 0u}};
        while (!stack.empty()) {
            const auto [nd, d] = stack.back();
            stack.pop_back();
            if (d > 64) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "Huffman codeword longer than 64 bits");
            if (leaf[nd]) {
                maxlen = std::max(maxlen, d);
                continue;
            }
            if (left[nd] >= 0) stack.push_back({left[nd], d + 1});
            if (right[nd] >= 0) stack.push_back({right[nd], d + 1});
        }
    }
    const uint32_t K = std::max<uint32_t>(maxlen, 1);
    if (enc > (uint64_t) UINT32_MAX * (kChunkBits / 8)) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "Huffman stream too long");
    const uint64_t nbits = enc * 8;
    const uint64_t nwords = (enc + 3) / 4;
    const uint32_t nchunks = (uint32_t) std::max<uint64_t>(1, (nbits + kChunkBits - 1) / kChunkBits);
    const uint64_t nmaps = (uint64_t) nchunks * K;
    HIPCHECK(ctx->dec_words.ensure(nwords + 2));
    HIPCHECK(ctx->dec_tab.ensure(2 * tab.size() + 3 * nc + sub.size()));
    HIPCHECK(ctx->dec_wtab.ensure(tab.size()));
    HIPCHECK(ctx->dec_leaf.ensure(nc));
    HIPCHECK(ctx->dec_pos.ensure(2 * (size_t) kQuarters * nchunks));
    HIPCHECK(ctx->dec_cnt.ensure((size_t) kQuarters * nchunks + 3));
    HIPCHECK(ctx->dec_qx.ensure(nmaps));
    HIPCHECK(ctx->dec_qn.ensure(nmaps));
    HIPCHECK(ctx->dec_x.ensure(2 * nmaps));
    HIPCHECK(ctx->dec_g.ensure((size_t) ((nchunks + kDecThreads - 1) / kDecThreads) * (K + 1)));
    HIPCHECK(ctx->dec_n.ensure(nmaps));
    HIPCHECK(ctx->tiles.ensure(scan_tiles(kQuarters * nchunks)));
    HIPCHECK(ctx->scalars.ensure(4));
    // the opening run: its codeword from the tree (a leaf whose value is
    // run_value), checked against the stream before the map launch
    RunPrefix run{0, 0, 0, 0, ctx->dec_cnt.p + (size_t) kQuarters * nchunks + 2};
    if (run_count && K <= 64) {
        std::vector<int32_t> parent(nc, -1);
        for (size_t i = 0; i < nc; i++) {
            if (leaf[i]) continue;
            if (left[i] >= 0) parent[left[i]] = (int32_t) i;
            if (right[i] >= 0) parent[right[i]] = (int32_t) i;
        }
        for (size_t i = 0; i < nc; i++) {
            if (!leaf[i] || (uint32_t) (sym[i] + offset) != run_value) continue;
            // the codeword read from the leaf up to the root (its last bit first)
            uint64_t code = 0;
            uint32_t len = 0;
            bool path = true;
            for (int32_t nd = (int32_t) i; nd != 0;) {
                const int32_t pa = parent[nd];
                if (pa < 0 || len >= 64) {
                    path = false;
                    break;
                }
                code |= (uint64_t) (right[pa] == nd) << len;
                len++;
                nd = pa;
            }
            if (path && len >= 1 && run_count <= nbits / len && (!limit || run_value < limit)) {
                run.bits = run_count * len;
                run.len = len;
                run.value = run_value;
                run.code = code;
            }
            break;
        }
    }
    // the stream (its last word zero-padded) goes up on the upload stream in
    // parts, below, so that the map launches overlap the later parts' copies
    hipStream_t us = ctx->ustream;
    launch_fill_u32(ctx->dec_words.p + (nwords ? nwords - 1 : 0), 2, 0u, d_buf ? st : us);
    uint32_t *d_tab = ctx->dec_tab.p;
    int32_t *d_tree = reinterpret_cast<int32_t *>(ctx->dec_tab.p + tab.size());
    // tables, tree arrays, leaf flags and the multi-codeword table staged in
    // pinned memory in their device layout: three asynchronous copies
    uint32_t *d_sub = ctx->dec_tab.p + tab.size() + 3 * nc;
    uint32_t *d_ctab = d_sub + sub.size();
    {
        const size_t nt = tab.size() + 3 * nc + sub.size() + ctab.size();
        const size_t nl = (nc + 3) / 4, nw = kMulti ? 2 * wtab.size() : 0;
        uint32_t *h = ctx->h_dec.ensure(nt + nl + nw);
        uint32_t *w = h;
        auto put = [&w](const void *src, size_t words) {
            memcpy(w, src, words * 4);
            w += words;
        };
        put(tab.data(), tab.size());
        put(left.data(), nc);
        put(right.data(), nc);
        put(sym.data(), nc);
        if (!sub.empty()) put(sub.data(), sub.size());
        put(ctab.data(), ctab.size());
        memcpy(h + nt, leaf.data(), nc);
        if (kMulti) memcpy(h + nt + nl, wtab.data(), wtab.size() * 8);
        HIPCHECK(hipMemcpyAsync(d_tab, h, nt * 4, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(ctx->dec_leaf.p, h + nt, nc, hipMemcpyHostToDevice, st));
        if (kMulti) HIPCHECK(hipMemcpyAsync(ctx->dec_wtab.p, h + nt + nl, wtab.size() * 8, hipMemcpyHostToDevice, st));
    }
    const Tree T{d_tree, d_tree + nc, d_tree + 2 * nc, ctx->dec_leaf.p, d_sub};
    // per quarter chunk: first codeword start, output offset, symbol count
    const uint32_t nquart = kQuarters * nchunks;
    uint64_t *d_start = ctx->dec_pos.p, *d_off = d_start + nquart;
    uint32_t *d_cnt = ctx->dec_cnt.p, *d_err = ctx->dec_cnt.p + nquart;
    const dim3 grid((nchunks + kDecThreads - 1) / kDecThreads), block(kDecThreads);
    // pass 1: entry -> exit maps and counts, part by part as the stream
    // arrives (the last workgroup of a part stages a few words past its
    // slice, which that part's copy includes)
    {
        const uint32_t nwg = grid.x;
        const int parts = nwg >= 64 && !d_buf ? kUploadParts : 1;
        uint64_t sent = 0;
        uint8_t *dst = reinterpret_cast<uint8_t *>(ctx->dec_words.p);
        if (d_buf) {  // resident: one copy on the stream, one map launch
            if (enc) HIPCHECK(hipMemcpyAsync(dst, d_buf + at, enc, hipMemcpyDeviceToDevice, st));
            sent = enc;
        }
        for (int k = 0; k < parts; k++) {
            const uint32_t w0 = (uint32_t) ((uint64_t) nwg * k / parts);
            const uint32_t w1 = (uint32_t) ((uint64_t) nwg * (k + 1) / parts);
            uint64_t need = k + 1 == parts ? enc : std::min<uint64_t>(enc, ((uint64_t) w1 * kGroupWords + kStageWords) * 4);
            if (need > sent) {
                HIPCHECK(hipMemcpyAsync(dst + sent, buf + at + sent, need - sent, hipMemcpyHostToDevice, us));
                sent = need;
            }
            if (!d_buf) {
                HIPCHECK(hipEventRecord(ctx->dec_ev[k], us));
                HIPCHECK(hipStreamWaitEvent(st, ctx->dec_ev[k], 0));
            }
            if (k == 0 && run.bits) {
                launch_fill_u32(const_cast<uint32_t *>(run.ok), 1, 1u, st);
                const uint64_t rw = (run.bits + 31) / 32;
                hipLaunchKernelGGL(huff_run_check_kernel, dim3((unsigned) ((rw + 255) / 256)), dim3(256), 0, st,
                                   ctx->dec_words.p, run, const_cast<uint32_t *>(run.ok));
            }
            if (w1 > w0)
                hipLaunchKernelGGL(huff_map_kernel, dim3(w1 - w0), block, 0, st, ctx->dec_words.p, nwords, nbits,
                                   nchunks, d_tab, d_ctab, T, K, ctx->dec_x.p, ctx->dec_n.p, w0, run, ctx->dec_qx.p,
                                   ctx->dec_qn.p);
            HIPCHECK(hipGetLastError());
        }
    }
    // pass 2: compositions inside groups of chunks, the groups' entries,
    // then every chunk's entry and count
    const uint32_t ngroups = (nchunks + kDecThreads - 1) / kDecThreads;
    uint8_t *d_gmap = ctx->dec_x.p + nmaps, *d_grp = ctx->dec_g.p, *d_gentry = ctx->dec_g.p + (size_t) ngroups * K;
    hipLaunchKernelGGL(huff_group_compose_kernel, dim3(ngroups), dim3(kDecThreads), 0, st, ctx->dec_x.p, nchunks, K,
                       d_gmap, d_grp);
    hipLaunchKernelGGL(huff_group_entry_kernel, dim3(1), dim3(1024), 0, st, d_grp, ngroups, K, d_gentry);
    ctx->dec_rounds = ngroups;
    hipLaunchKernelGGL(huff_entry_kernel, dim3((nchunks + 255) / 256), dim3(256), 0, st, d_gmap, d_gentry,
                       ctx->dec_n.p, ctx->dec_qx.p, ctx->dec_qn.p, nchunks, K, d_start, d_cnt);
    HIPCHECK(hipGetLastError());
    // output offsets of the quarter chunks (exclusive scan) and the total,
    // then pass 3
    launch_zero_scan(d_cnt, d_off, nquart, ctx->tiles.p, ctx->scalars.p, st);
    launch_fill_u32(d_err, 1, 0u, st);
    if (zero_counts) launch_fill_u32(zero_counts, (num + kHistChunk - 1) / kHistChunk, 0u, st);
    hipLaunchKernelGGL((huff_write_kernel<OUT, kMulti>), dim3((nquart + kDecThreads - 1) / kDecThreads), block, 0, st,
                       ctx->dec_words.p, nwords, nbits, nchunks,
                       d_tab, ctx->dec_wtab.p, T, d_start, d_off, offset, limit, (uint64_t) num, d_out, d_err,
                       run, zero_counts);
    HIPCHECK(hipGetLastError());
    if (defer) {
        defer->total = ctx->scalars.p;
        defer->err = d_err;
        return FAZ_OK;
    }
    uint32_t *hf = ctx->hp_hist.data() + 2 * 65536;  // pinned scalars
    HIPCHECK(hipMemcpyAsync(hf, ctx->scalars.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(hf + 2, d_err, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(spin_wait(st));
    uint64_t total;
    memcpy(&total, hf, 8);
    if (total < num)
        CTX_FAIL(FAZ_ERR_FORMAT, "Huffman bitstream holds %llu symbols, %zu expected", (unsigned long long) total, num);
    if (hf[2]) CTX_FAIL(FAZ_ERR_FORMAT, "Huffman bitstream leaves the tree or a symbol is out of range");
    return FAZ_OK;
}

template int gpu_huffman_decode<uint16_t>(faz_gpu_ctx *, const uint8_t *, size_t, size_t, uint32_t, uint16_t *,
                                          size_t *, const uint8_t *, uint32_t, uint64_t, DecodeCheck *, uint32_t *);
template int gpu_huffman_decode<int32_t>(faz_gpu_ctx *, const uint8_t *, size_t, size_t, uint32_t, int32_t *,
                                          size_t *, const uint8_t *, uint32_t, uint64_t, DecodeCheck *, uint32_t *);
template int gpu_huffman_decode<uint32_t>(faz_gpu_ctx *, const uint8_t *, size_t, size_t, uint32_t, uint32_t *,
                                          size_t *, const uint8_t *, uint32_t, uint64_t, DecodeCheck *, uint32_t *);

// ---------------------------------------------------------------------------
// Unpredictable values of 32-bit code streams (the general frontend): the
// k-th zero code takes the k-th value of the list.
// ---------------------------------------------------------------------------
namespace {

constexpr uint32_t kZeroThreads = 256;
constexpr uint32_t kZeroPer = 16;
constexpr uint32_t kZeroChunk = kZeroThreads * kZeroPer;

__device__ __forceinline__ uint32_t zeros16(const uint32_t *__restrict__ codes, uint64_t i0, uint64_t n,
                                            uint32_t &mask) {
    mask = 0;
#pragma unroll
    for (uint32_t j = 0; j < kZeroPer; j++) {
        const uint64_t i = i0 + j;
        if (i < n && codes[i] == 0) mask |= 1u << j;
    }
    return (uint32_t) __popc(mask);
}

__global__ void __launch_bounds__(kZeroThreads) zero_count_u32_kernel(const uint32_t *__restrict__ codes, uint64_t n,
                                                                      uint32_t *__restrict__ chunk_zeros) {
    __shared__ uint32_t s[kZeroThreads / 64];
    uint32_t mask;
    uint32_t z = zeros16(codes, (uint64_t) blockIdx.x * kZeroChunk + threadIdx.x * kZeroPer, n, mask);
    for (int o = 32; o > 0; o >>= 1) z += __shfl_xor(z, o);
    if (threadIdx.x % 64 == 0) s[threadIdx.x / 64] = z;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kZeroThreads / 64; w++) t += s[w];
        chunk_zeros[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kZeroThreads) zero_scatter_u32_kernel(const uint32_t *__restrict__ codes, uint64_t n,
                                                                        const uint64_t *__restrict__ chunk_base,
                                                                        const float *__restrict__ unpred, uint64_t U,
                                                                        float *__restrict__ outval) {
    __shared__ uint32_t s[kZeroThreads / 64];
    const uint64_t i0 = (uint64_t) blockIdx.x * kZeroChunk + threadIdx.x * kZeroPer;
    uint32_t mask;
    const uint32_t z = zeros16(codes, i0, n, mask);
    // exclusive prefix of the per-thread counts inside the workgroup
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    uint32_t incl = z;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= (uint32_t) o) incl += t;
    }
    if (lane == 63) s[wave] = incl;
    __syncthreads();
    uint64_t rank = chunk_base[blockIdx.x] + incl - z;
    for (uint32_t w = 0; w < wave; w++) rank += s[w];
    while (mask) {
        const uint32_t j = (uint32_t) __builtin_ctz(mask);
        mask &= mask - 1;
        if (rank < U) outval[i0 + j] = unpred[rank];
        rank++;
    }
}

}  // namespace

// Device zero ranks of n 32-bit codes: outval[i] = unpred[rank of zero i]
// for every zero code; fails when there are more zero codes than values.
int gpu_scatter_unpred_u32(faz_gpu_ctx *ctx, const uint32_t *d_codes, uint64_t n, const float *h_unpred, uint64_t U,
                           float *d_outval) {
    hipStream_t st = ctx->stream;
    const uint64_t nz64 = (n + kZeroChunk - 1) / kZeroChunk;
    if (nz64 >= (1ull << 31)) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "too many codes");
    const uint32_t nz = (uint32_t) std::max<uint64_t>(nz64, 1);
    HIPCHECK(ctx->unpred.ensure(U));
    HIPCHECK(ctx->chunk_zeros.ensure(nz));
    HIPCHECK(ctx->chunk_base.ensure(nz));
    HIPCHECK(ctx->tiles.ensure(scan_tiles(nz)));
    HIPCHECK(ctx->scalars.ensure(4));
    if (U) HIPCHECK(hipMemcpyAsync(ctx->unpred.p, h_unpred, U * 4, hipMemcpyHostToDevice, st));
    if (!n) return FAZ_OK;
    hipLaunchKernelGGL(zero_count_u32_kernel, dim3(nz), dim3(kZeroThreads), 0, st, d_codes, n, ctx->chunk_zeros.p);
    launch_zero_scan(ctx->chunk_zeros.p, ctx->chunk_base.p, nz, ctx->tiles.p, ctx->scalars.p, st);
    uint32_t *hs = ctx->hp_hist.data() + 2 * 65536;  // pinned
    HIPCHECK(hipMemcpyAsync(hs, ctx->scalars.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(spin_wait(st));
    uint64_t zeros;
    memcpy(&zeros, hs, 8);
    if (zeros > U) CTX_FAIL(FAZ_ERR_FORMAT, "more zero codes than unpredictable values");
    hipLaunchKernelGGL(zero_scatter_u32_kernel, dim3(nz), dim3(kZeroThreads), 0, st, d_codes, n, ctx->chunk_base.p,
                       ctx->unpred.p, U, d_outval);
    HIPCHECK(hipGetLastError());
    return FAZ_OK;
}

}  // namespace faz

extern "C" {

faz_status faz_gpu_huffman_decode(faz_gpu_ctx *ctx, const uint8_t *h_buf, size_t n, size_t num, uint32_t limit,
                                  int32_t *d_out, size_t *used) {
    if (!ctx || !h_buf || (!d_out && num)) return FAZ_ERR_INVALID;
    HIPCHECK(hipSetDevice(ctx->device));
    return (faz_status) faz::gpu_huffman_decode<int32_t>(ctx, h_buf, n, num, limit, d_out, used);
}

}  // extern "C"
