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
// the maps with a parallel prefix (Hillis-Steele, log2(chunks) steps) to get
// every e_c from e_0 = 0; the counts at the true entries are scanned into
// output offsets, and pass 3 decodes each chunk again, writing its symbols.
// Every pass stages its workgroup's slice of the stream in LDS.
//
// Codewords of up to kTabBits bits come from one table lookup (LDS); longer
// ones continue bit by bit down the tree from the node the table names.
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

struct BitReader {
    const uint32_t *lw;     // staged words (byte order swapped), see slot()
    uint64_t base;          // stream word index of lw's first word
    const uint32_t *words;  // the stream as 32-bit words (bytes in stream order)
    uint64_t nwords;
    uint64_t buf;           // next bits, MSB first
    int have;               // valid bits in buf
    uint64_t next;          // next word to load

    __device__ uint32_t word(uint64_t i) const {
        const uint64_t li = i - base;
        if (li < kStageWords) return lw[slot((uint32_t) li)];
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
    __device__ void init(const BitReader &proto, uint64_t p) {
        r = proto;
        r.start(p);
        pos = p;
        k = 0;
    }
    __device__ void step(const uint32_t *tab, const Tree &T, uint64_t nbits) {
        const int32_t nd = decode_one(r, tab, T, pos);
        if (nd >= 0 && pos <= nbits) k++;
    }
};

__device__ __forceinline__ bool mask_has(const uint64_t *m, uint32_t q) { return (m[q >> 6] >> (q & 63)) & 1u; }
__device__ __forceinline__ uint32_t mask_below(const uint64_t *m, uint32_t q) {
    if (q < 64) return (uint32_t) __popcll(m[0] & ((1ull << q) - 1));
    if (q >= 128) return (uint32_t) (__popcll(m[0]) + __popcll(m[1]));
    return (uint32_t) (__popcll(m[0]) + __popcll(m[1] & ((1ull << (q - 64)) - 1)));
}

// Pass 1: for chunk c and every entry offset o < K, the next chunk's entry
// offset (xmap, clamped to K - 1) and the symbols that start in the chunk
// (nmap).  A trajectory merges with an earlier one when it reaches one of its
// codeword starts; the starts of the first two trajectories are kept as bit
// masks over the chunk's first kMaskBits bits.
__global__ void __launch_bounds__(kDecThreads) huff_map_kernel(const uint32_t *__restrict__ words, uint64_t nwords,
                                                               uint64_t nbits, uint32_t nchunks,
                                                               const uint32_t *__restrict__ gtab, Tree T, uint32_t K,
                                                               uint8_t *__restrict__ xmap,
                                                               uint16_t *__restrict__ nmap) {
    __shared__ uint32_t tab[1 << kTabBits];
    __shared__ uint32_t lw[kStageSlots];
    const uint64_t base = (uint64_t) blockIdx.x * kGroupWords;
    stage(tab, lw, gtab, words, nwords, base);
    const uint32_t c = blockIdx.x * kDecThreads + threadIdx.x;
    if (c >= nchunks) return;
    BitReader proto;
    proto.lw = lw, proto.base = base, proto.words = words, proto.nwords = nwords;
    const uint64_t b = (uint64_t) c * kChunkBits;
    const uint64_t end = min(b + kChunkBits, nbits);
    const uint64_t nxt = b + kChunkBits;
    // masks are exact only where every symbol before kMaskBits ends inside the stream
    const bool masks = b + kMaskBits <= nbits;
    auto put = [&](uint32_t o, uint64_t x, uint32_t n) {
        const uint64_t e = x > nxt ? x - nxt : 0;
        xmap[(uint64_t) c * K + o] = (uint8_t) min<uint64_t>(e, K - 1);
        nmap[(uint64_t) c * K + o] = (uint16_t) n;
    };
    uint64_t mA[2] = {0, 0}, mB[2] = {0, 0};
    uint64_t xA, xB = 0;
    uint32_t nA, nB = 0;
    bool hasB = false;
    {
        Cursor A;
        A.init(proto, b);
        while (A.pos < end) {
            const uint64_t q = A.pos - b;
            if (q < kMaskBits) mA[q >> 6] |= 1ull << (q & 63);
            A.step(tab, T, nbits);
        }
        xA = A.pos, nA = A.k;
    }
    put(0, xA, nA);
    for (uint32_t o = 1; o < K; o++) {
        if (masks && mask_has(mA, o)) {
            put(o, xA, nA - mask_below(mA, o));
            continue;
        }
        if (masks && hasB && mask_has(mB, o)) {
            put(o, xB, nB - mask_below(mB, o));
            continue;
        }
        Cursor O;
        O.init(proto, b + o);
        uint64_t mO[2] = {0, 0};
        bool merged = false;
        while (O.pos < end) {
            const uint64_t q = O.pos - b;
            if (masks && q < kMaskBits) {
                if (mask_has(mA, (uint32_t) q)) {
                    put(o, xA, O.k + nA - mask_below(mA, (uint32_t) q));
                    merged = true;
                    break;
                }
                if (hasB && mask_has(mB, (uint32_t) q)) {
                    put(o, xB, O.k + nB - mask_below(mB, (uint32_t) q));
                    merged = true;
                    break;
                }
                mO[q >> 6] |= 1ull << (q & 63);
            }
            O.step(tab, T, nbits);
        }
        if (merged) continue;
        put(o, O.pos, O.k);
        if (!hasB) {
            hasB = true;
            xB = O.pos, nB = O.k;
            mB[0] = mO[0], mB[1] = mO[1];
        }
    }
}

// Pass 2, one Hillis-Steele step: out_c = in_c o in_{c-d} for c >= d.
__global__ void huff_compose_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ out, uint64_t total,
                                    uint32_t K, uint32_t d) {
    const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint64_t c = i / K, o = i % K;
    uint8_t v = in[i];
    if (c >= d) v = in[c * K + in[(c - d) * K + o]];
    out[i] = v;
}

// Entry of every chunk (chunk c's is the composed map of chunk c - 1 at 0)
// and its symbol count at that entry.
__global__ void huff_entry_kernel(const uint8_t *__restrict__ comp, const uint16_t *__restrict__ nmap,
                                  uint32_t nchunks, uint32_t K, uint64_t *__restrict__ start,
                                  uint32_t *__restrict__ count) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    const uint32_t e = c ? comp[(uint64_t) (c - 1) * K] : 0u;
    start[c] = (uint64_t) c * kChunkBits + e;
    count[c] = nmap[(uint64_t) c * K + e];
}

// Pass 3: chunk c writes its symbols from offset[c] on (those below num),
// each sym + sym_offset, which must lie in [0, limit) unless limit is 0;
// err[0] = 1 otherwise or when the bits leave the tree.
template<class OUT>
__global__ void __launch_bounds__(kDecThreads) huff_write_kernel(const uint32_t *__restrict__ words, uint64_t nwords,
                                                                 uint64_t nbits, uint32_t nchunks,
                                                                 const uint32_t *__restrict__ gtab, Tree T,
                                                                 const uint64_t *__restrict__ start,
                                                                 const uint64_t *__restrict__ offset,
                                                                 int32_t sym_offset, uint32_t limit, uint64_t num,
                                                                 OUT *__restrict__ out, uint32_t *__restrict__ err) {
    __shared__ uint32_t tab[1 << kTabBits];
    __shared__ uint32_t lw[kStageSlots];
    const uint64_t base = (uint64_t) blockIdx.x * kGroupWords;
    stage(tab, lw, gtab, words, nwords, base);
    const uint32_t c = blockIdx.x * kDecThreads + threadIdx.x;
    if (c >= nchunks) return;
    const uint64_t end = min((uint64_t) (c + 1) * kChunkBits, nbits);
    uint64_t pos = start[c], o = offset[c];
    BitReader r;
    r.lw = lw, r.base = base, r.words = words, r.nwords = nwords;
    r.start(pos);
    // symbols leave in 8-byte stores once out + o is 8-byte aligned
    constexpr int P = 8 / sizeof(OUT);
    constexpr int W = 8 * sizeof(OUT);
    uint64_t acc = 0;
    int na = 0;
    bool bad = false;
    while (pos < end && o < num) {
        const int32_t k = decode_one(r, tab, T, pos);
        if (k < 0) {
            bad = true;
            break;
        }
        const uint32_t v = (uint32_t) (T.sym[k] + sym_offset);
        if (limit && v >= limit) {
            bad = true;
            break;
        }
        if (na == 0 && (reinterpret_cast<uintptr_t>(out + o) & 7u) != 0) {
            out[o++] = (OUT) v;
            continue;
        }
        acc |= (uint64_t) (std::make_unsigned_t<OUT>) (OUT) v << (W * na);
        na++;
        o++;
        if (na == P) {
            *reinterpret_cast<uint64_t *>(out + (o - P)) = acc;
            acc = 0;
            na = 0;
        }
    }
    for (int j = 0; j < na; j++) out[o - na + j] = (OUT) (acc >> (W * j));
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
    memcpy(&enc_bytes, tb + treeBytes, 8);
    bits_at = 12 + treeBytes + 8;
    return bits_at + enc_bytes <= n;
}

}  // namespace

// HuffmanEncoder::decode on the device: num symbols (each plus the tree's
// offset, which must land in [0, limit) unless limit is 0) into d_out on
// ctx->stream.  Returns with the stream drained; *used = bytes of the Huffman
// section.
template<class OUT>
int gpu_huffman_decode(faz_gpu_ctx *ctx, const uint8_t *buf, size_t n, size_t num, uint32_t limit, OUT *d_out,
                       size_t *used) {
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
    std::vector<uint32_t> tab((size_t) 1 << kTabBits);
    for (uint32_t v = 0; v < tab.size(); v++) {
        int32_t stop = -1;
        uint32_t e = entry(0, v, kTabBits, stop);
        if (e == kLongFlag) e = build_sub(stop, kTabBits);
        tab[v] = e;
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
    HIPCHECK(ctx->dec_tab.ensure(tab.size() + 3 * nc + sub.size()));
    HIPCHECK(ctx->dec_leaf.ensure(nc));
    HIPCHECK(ctx->dec_pos.ensure(2 * (size_t) nchunks));
    HIPCHECK(ctx->dec_cnt.ensure((size_t) nchunks + 2));
    HIPCHECK(ctx->dec_x.ensure(2 * nmaps));
    HIPCHECK(ctx->dec_n.ensure(nmaps));
    HIPCHECK(ctx->tiles.ensure(scan_tiles(nchunks)));
    HIPCHECK(ctx->scalars.ensure(4));
    // the stream, its last word zero-padded
    launch_fill_u32(ctx->dec_words.p + (nwords ? nwords - 1 : 0), 2, 0u, st);
    if (enc) HIPCHECK(hipMemcpyAsync(ctx->dec_words.p, buf + at, enc, hipMemcpyHostToDevice, st));
    uint32_t *d_tab = ctx->dec_tab.p;
    int32_t *d_tree = reinterpret_cast<int32_t *>(ctx->dec_tab.p + tab.size());
    HIPCHECK(hipMemcpyAsync(d_tab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(d_tree, left.data(), nc * 4, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(d_tree + nc, right.data(), nc * 4, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(d_tree + 2 * nc, sym.data(), nc * 4, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(ctx->dec_leaf.p, leaf.data(), nc, hipMemcpyHostToDevice, st));
    uint32_t *d_sub = ctx->dec_tab.p + tab.size() + 3 * nc;
    if (!sub.empty()) HIPCHECK(hipMemcpyAsync(d_sub, sub.data(), sub.size() * 4, hipMemcpyHostToDevice, st));
    const Tree T{d_tree, d_tree + nc, d_tree + 2 * nc, ctx->dec_leaf.p, d_sub};
    uint64_t *d_start = ctx->dec_pos.p, *d_off = d_start + nchunks;
    uint32_t *d_cnt = ctx->dec_cnt.p, *d_err = ctx->dec_cnt.p + nchunks;
    const dim3 grid((nchunks + kDecThreads - 1) / kDecThreads), block(kDecThreads);
    // pass 1: entry -> exit maps and counts
    hipLaunchKernelGGL(huff_map_kernel, grid, block, 0, st, ctx->dec_words.p, nwords, nbits, nchunks, d_tab, T, K,
                       ctx->dec_x.p, ctx->dec_n.p);
    HIPCHECK(hipGetLastError());
    // pass 2: prefix composition, then every chunk's entry and count
    uint8_t *cur = ctx->dec_x.p, *alt = ctx->dec_x.p + nmaps;
    const dim3 cgrid((unsigned) ((nmaps + 255) / 256));
    uint32_t steps = 0;
    for (uint32_t d = 1; d < nchunks; d *= 2, steps++) {
        hipLaunchKernelGGL(huff_compose_kernel, cgrid, dim3(256), 0, st, cur, alt, nmaps, K, d);
        std::swap(cur, alt);
        if (d > nchunks / 2) break;
    }
    ctx->dec_rounds = steps;
    hipLaunchKernelGGL(huff_entry_kernel, dim3((nchunks + 255) / 256), dim3(256), 0, st, cur, ctx->dec_n.p, nchunks,
                       K, d_start, d_cnt);
    HIPCHECK(hipGetLastError());
    // output offsets of the chunks (exclusive scan) and the total, then pass 3
    launch_zero_scan(d_cnt, d_off, nchunks, ctx->tiles.p, ctx->scalars.p, st);
    launch_fill_u32(d_err, 1, 0u, st);
    hipLaunchKernelGGL((huff_write_kernel<OUT>), grid, block, 0, st, ctx->dec_words.p, nwords, nbits, nchunks, d_tab,
                       T, d_start, d_off, offset, limit, (uint64_t) num, d_out, d_err);
    HIPCHECK(hipGetLastError());
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
                                          size_t *);
template int gpu_huffman_decode<int32_t>(faz_gpu_ctx *, const uint8_t *, size_t, size_t, uint32_t, int32_t *,
                                         size_t *);

}  // namespace faz

extern "C" {

faz_status faz_gpu_huffman_decode(faz_gpu_ctx *ctx, const uint8_t *h_buf, size_t n, size_t num, uint32_t limit,
                                  int32_t *d_out, size_t *used) {
    if (!ctx || !h_buf || (!d_out && num)) return FAZ_ERR_INVALID;
    HIPCHECK(hipSetDevice(ctx->device));
    return (faz_status) faz::gpu_huffman_decode<int32_t>(ctx, h_buf, n, num, limit, d_out, used);
}

}  // extern "C"
