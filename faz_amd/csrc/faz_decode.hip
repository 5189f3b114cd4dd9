// Huffman decode on gfx950: HuffmanEncoder::load + decode (include/QoZ/
// encoder/HuffmanEncoder.hpp:236-296) for the bitstreams the reference
// writes (one MSB-first stream, no chunk index).
//
// The stream is cut into chunks of kChunkBits bits.  A thread decodes one
// chunk from a start position until it passes the chunk's end; its exit
// position is where the next chunk's first codeword begins if its own start
// was right.  Chunk 0 starts at bit 0; every other chunk first starts at its
// boundary, which is usually inside a codeword, but a Huffman decoder
// resynchronises within a few codewords, so after one correction round the
// exits almost always agree.  Rounds repeat until no chunk's start moves
// (chunk i is exact once chunk i - 1 is, so at most one round per chunk).
// Then the symbol counts are scanned and every chunk decodes again, writing
// its symbols at its offset.
//
// Codewords of up to kTabBits bits come from one table lookup (LDS); longer
// ones continue bit by bit down the tree from the node the table names.
#include "faz_gpu.h"
#include "faz_ctx.h"
#include "faz_internal.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

namespace faz {
namespace {

constexpr int kTabBits = 12;
constexpr uint32_t kChunkBits = 4096;
// Table entry: a codeword of at most kTabBits bits -> leaf node (bits 0..23)
// | length << 24; a longer one -> kLongFlag | node reached after kTabBits
// bits; kOffTree: the prefix leaves the tree (malformed tree).
constexpr uint32_t kLongFlag = 0x80000000u;
constexpr uint32_t kOffTree = 0xFFFFFFFFu;

struct Tree {
    const int32_t *left, *right, *sym;
    const uint8_t *leaf;
};

struct BitReader {
    const uint32_t *words;  // the stream as 32-bit words (bytes in stream order)
    uint64_t nwords;
    uint64_t buf;           // next bits, MSB first
    int have;               // valid bits in buf
    uint64_t next;          // next word to load

    __device__ uint32_t word(uint64_t i) const { return i < nwords ? __builtin_bswap32(words[i]) : 0u; }
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
    const uint32_t e = tab[r.peek(kTabBits)];
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
    r.drop(kTabBits);
    pos += kTabBits;
    int32_t nd = (int32_t) (e & 0x7FFFFFFFu);
    // codewords are at most 64 bits (the reference keeps them in 64-bit words)
    for (int depth = kTabBits; depth < 64; depth++) {
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

__device__ __forceinline__ void load_table(uint32_t *tab, const uint32_t *__restrict__ gtab) {
    for (int i = threadIdx.x; i < (1 << kTabBits); i += blockDim.x) tab[i] = gtab[i];
    __syncthreads();
}

// Round 0: every chunk from its boundary.  Later rounds: chunk c restarts
// from chunk c - 1's exit where that differs from its start (changed[0] = 1).
// Records exit[c] and the symbols that end inside the stream.  Bits that
// leave the tree (only a mis-synchronised start can meet them in a well-formed
// stream) restart at the root; the write pass reports them.
__global__ void __launch_bounds__(256) huff_count_kernel(const uint32_t *__restrict__ words, uint64_t nwords,
                                                         uint64_t nbits, uint32_t nchunks,
                                                         const uint32_t *__restrict__ gtab, Tree T,
                                                         uint64_t *__restrict__ start, uint64_t *__restrict__ exitp,
                                                         uint32_t *__restrict__ count, uint32_t *__restrict__ changed,
                                                         int round) {
    __shared__ uint32_t tab[1 << kTabBits];
    load_table(tab, gtab);
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    uint64_t s;
    if (round == 0) {
        s = (uint64_t) c * kChunkBits;
    } else {
        if (c == 0) return;
        s = exitp[c - 1];
        if (s == start[c]) return;
        atomicOr(changed, 1u);
    }
    const uint64_t end = min((uint64_t) (c + 1) * kChunkBits, nbits);
    uint64_t pos = s;
    uint32_t n = 0;
    BitReader r;
    r.words = words, r.nwords = nwords;
    r.start(pos);
    while (pos < end) {
        const int32_t k = decode_one(r, tab, T, pos);
        if (k >= 0 && pos <= nbits) n++;
    }
    start[c] = s;
    exitp[c] = pos;
    count[c] = n;
}

// Chunk c writes its symbols from offset[c] on (those below num), each
// sym + sym_offset, which must lie in [0, limit) unless limit is 0; err[0] = 1
// otherwise or when the bits leave the tree.
template<class OUT>
__global__ void __launch_bounds__(256) huff_write_kernel(const uint32_t *__restrict__ words, uint64_t nwords,
                                                         uint64_t nbits, uint32_t nchunks,
                                                         const uint32_t *__restrict__ gtab, Tree T,
                                                         const uint64_t *__restrict__ start,
                                                         const uint64_t *__restrict__ offset, int32_t sym_offset,
                                                         uint32_t limit, uint64_t num, OUT *__restrict__ out,
                                                         uint32_t *__restrict__ err) {
    __shared__ uint32_t tab[1 << kTabBits];
    load_table(tab, gtab);
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    const uint64_t end = min((uint64_t) (c + 1) * kChunkBits, nbits);
    uint64_t pos = start[c], o = offset[c];
    BitReader r;
    r.words = words, r.nwords = nwords;
    r.start(pos);
    while (pos < end && o < num) {
        const int32_t k = decode_one(r, tab, T, pos);
        if (k < 0) {
            atomicOr(err, 1u);
            return;
        }
        const uint32_t v = (uint32_t) (T.sym[k] + sym_offset);
        if (limit && v >= limit) {
            atomicOr(err, 1u);
            return;
        }
        out[o++] = (OUT) v;
    }
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
    // primary table: every kTabBits-bit prefix walked down from the root
    const size_t nc = sym.size();
    std::vector<uint32_t> tab((size_t) 1 << kTabBits);
    for (uint32_t v = 0; v < tab.size(); v++) {
        int32_t nd = 0;
        uint32_t e = 0;
        bool done = false;
        for (int d = 0; d < kTabBits && !done; d++) {
            const uint32_t bit = (v >> (kTabBits - 1 - d)) & 1u;
            nd = bit ? right[nd] : left[nd];
            if (nd < 0) {
                e = kOffTree;
                done = true;
            } else if (leaf[nd]) {
                e = (uint32_t) nd | (uint32_t) (d + 1) << 24;
                done = true;
            }
        }
        tab[v] = done ? e : (kLongFlag | (uint32_t) nd);
    }
    if (enc > (uint64_t) UINT32_MAX * (kChunkBits / 8)) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "Huffman stream too long");
    const uint64_t nbits = enc * 8;
    const uint64_t nwords = (enc + 3) / 4;
    const uint32_t nchunks = (uint32_t) std::max<uint64_t>(1, (nbits + kChunkBits - 1) / kChunkBits);
    HIPCHECK(ctx->dec_words.ensure(nwords + 2));
    HIPCHECK(ctx->dec_tab.ensure(tab.size() + 3 * nc));
    HIPCHECK(ctx->dec_leaf.ensure(nc));
    HIPCHECK(ctx->dec_pos.ensure(3 * (size_t) nchunks));
    HIPCHECK(ctx->dec_cnt.ensure((size_t) nchunks + 2));
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
    const Tree T{d_tree, d_tree + nc, d_tree + 2 * nc, ctx->dec_leaf.p};
    uint64_t *d_start = ctx->dec_pos.p, *d_exit = d_start + nchunks, *d_off = d_start + 2 * (size_t) nchunks;
    uint32_t *d_cnt = ctx->dec_cnt.p, *d_flags = ctx->dec_cnt.p + nchunks;  // [0] changed, [1] error
    const dim3 grid((nchunks + 255) / 256), block(256);
    uint32_t *hf = ctx->hp_hist.data() + 2 * 65536;  // pinned scalars
    uint32_t round = 0;
    for (;; round++) {
        launch_fill_u32(d_flags, 2, 0u, st);
        hipLaunchKernelGGL(huff_count_kernel, grid, block, 0, st, ctx->dec_words.p, nwords, nbits, nchunks, d_tab, T,
                           d_start, d_exit, d_cnt, d_flags, (int) round);
        HIPCHECK(hipGetLastError());
        if (round == 0) continue;
        HIPCHECK(hipMemcpyAsync(hf, d_flags, 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(spin_wait(st));
        if (!hf[0]) break;
        if (round > nchunks) CTX_FAIL(FAZ_ERR_FORMAT, "Huffman chunk starts did not settle");
    }
    ctx->dec_rounds = round;
    // symbol offsets of the chunks (exclusive scan), the total, then the write
    launch_zero_scan(d_cnt, d_off, nchunks, ctx->tiles.p, ctx->scalars.p, st);
    hipLaunchKernelGGL((huff_write_kernel<OUT>), grid, block, 0, st, ctx->dec_words.p, nwords, nbits, nchunks, d_tab,
                       T, d_start, d_off, offset, limit, (uint64_t) num, d_out, d_flags + 1);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(hf + 2, ctx->scalars.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(hf + 4, d_flags + 1, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(spin_wait(st));
    uint64_t total;
    memcpy(&total, hf + 2, 8);
    if (total < num)
        CTX_FAIL(FAZ_ERR_FORMAT, "Huffman bitstream holds %llu symbols, %zu expected", (unsigned long long) total, num);
    if (hf[4]) CTX_FAIL(FAZ_ERR_FORMAT, "Huffman bitstream leaves the tree or a symbol is out of range");
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
