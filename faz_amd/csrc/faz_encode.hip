// Huffman encode (HuffmanEncoder::encode, HuffmanEncoder.hpp:150-227) on the
// GPU: chunk bit totals, an exclusive scan of them (faz_kernels.hip), then
// per-chunk packing of the MSB-first bitstream into 32-bit words stored with
// big-endian byte order (the reference writes 64-bit big-endian stores of an
// MSB-first stream, which is the same byte sequence).
//
// Workgroups are persistent over chunks of kEncChunk codes, so the code table
// is staged in LDS once per workgroup: the window of at most kTabLds entries
// that holds every symbol but 0, followed by symbol 0's entry (the outlier
// code, far from the others), which every index outside the window clamps
// to.  Tables whose symbols spread wider read global memory.  A thread's 16
// consecutive codes form one run of bits starting at the chunk offset plus a
// workgroup prefix of lengths; the thread assembles the run in a 64-bit
// register (codewords of at most 32 bits are stored left-aligned, so each is
// one shift and an OR) and writes it to an LDS bit buffer, words wholly
// inside the run with plain stores and its first and last word, which it may
// share with the neighbouring threads, with atomic ORs.  The chunk's words then go
// out with plain stores; the first and last word of a chunk, shared with the
// neighbouring chunks, with atomic ORs into the zeroed output.
#include "faz_internal.h"

namespace faz {

namespace {

constexpr uint32_t kTabLds = 4096;                    // table entries staged in LDS
constexpr uint32_t kPackWords = kEncChunk + 4;        // bit buffer: 32 bits per code on average
constexpr uint32_t kEncGridPerCu = 8;

__device__ __forceinline__ uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

// 16 consecutive codes from position j as two 16-byte words (16-byte loads
// when the codes are 16-byte aligned; zero past end).  Loaded one chunk
// ahead of their use, so a workgroup's code reads overlap its packing.
struct Codes16 {
    uint4 a, b;
    __device__ __forceinline__ uint16_t operator[](int k) const {
        const uint4 &w = k < 8 ? a : b;
        const uint32_t word = (k & 6) == 0 ? w.x : (k & 6) == 2 ? w.y : (k & 6) == 4 ? w.z : w.w;
        return (uint16_t) (k & 1 ? word >> 16 : word);
    }
};

__device__ __forceinline__ Codes16 fetch16(const uint16_t *__restrict__ codes, uint32_t j, uint32_t end,
                                           bool aligned) {
    Codes16 r;
    if (aligned && j + 16 <= end) {
        const uint4 *q = reinterpret_cast<const uint4 *>(codes + j);
        r.a = q[0];
        r.b = q[1];
    } else {
        uint32_t t[16];
#pragma unroll
        for (int k = 0; k < 16; k++) t[k] = j + k < end ? codes[j + k] : 0;
        r.a = make_uint4(t[0] | t[1] << 16, t[2] | t[3] << 16, t[4] | t[5] << 16, t[6] | t[7] << 16);
        r.b = make_uint4(t[8] | t[9] << 16, t[10] | t[11] << 16, t[12] | t[13] << 16, t[14] | t[15] << 16);
    }
    return r;
}

__device__ __forceinline__ Codes16 fetch_chunk(const uint16_t *__restrict__ codes, uint32_t c, uint32_t n,
                                               bool aligned) {
    const uint32_t base = c * kEncChunk;
    return fetch16(codes, base + threadIdx.x * kEncPerThread, min(base + kEncChunk, n), aligned);
}

// Table entries [win, win + nwin) live in LDS, then e0, the entry of symbol
// 0.  narrow: every codeword has at most 32 bits (encode_entry's layouts).
struct TabWindow {
    uint32_t win, nwin;
    uint64_t e0;
    uint32_t max_len;  // longest codeword of the table
    __device__ __forceinline__ bool narrow() const { return max_len <= 32; }
};

// Symbol 0 lies below the window (every other symbol is positive), so its
// index wraps above nwin and clamps onto e0's slot.
template<bool LDS_TAB>
__device__ __forceinline__ uint64_t entry(const uint64_t *s_tab, const uint64_t *__restrict__ tab, const TabWindow &w,
                                          uint16_t sym, int32_t offset) {
    if (!LDS_TAB) return tab[(int32_t) sym - offset];
    return s_tab[min((uint32_t) ((int32_t) sym - offset) - w.win, w.nwin)];
}

__device__ __forceinline__ uint32_t entry_len(uint64_t e, bool narrow) {
    return narrow ? (uint32_t) (e >> 32) : (uint32_t) e & 0xFF;
}

// LDS bytes of a staged window
__host__ __device__ constexpr size_t tab_lds_bytes(uint32_t nwin) { return 8 * ((size_t) nwin + 1); }

template<bool LDS_TAB>
__device__ __forceinline__ void stage_table(uint64_t *s_tab, const uint64_t *__restrict__ tab, const TabWindow &w) {
    if (!LDS_TAB) return;
    for (uint32_t i = threadIdx.x; i < w.nwin; i += kEncThreads) s_tab[i] = tab[w.win + i];
    if (threadIdx.x == 0) s_tab[w.nwin] = w.e0;
    __syncthreads();
}

// Workgroup exclusive prefix of x; *total gets the sum.
__device__ __forceinline__ uint32_t wg_exclusive(uint32_t x, uint32_t *s_wave, uint32_t *total) {
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    uint32_t incl = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = (uint32_t) __shfl_up((int) incl, o);
        if (lane >= (uint32_t) o) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (uint32_t w = 0; w < kEncThreads / 64; w++) {
        const uint32_t v = s_wave[w];
        before += w < wave ? v : 0;
        sum += v;
    }
    *total = sum;
    return before + incl - x;
}

// Bit total of chunk c (codes v of this thread; unless FULL, valid below end).
template<bool LDS_TAB, bool FULL>
__device__ __forceinline__ void sizes_chunk(uint32_t c, const Codes16 &v, uint32_t end, const uint64_t *s_tab,
                                            const uint64_t *__restrict__ tab, const TabWindow &tw, int32_t offset,
                                            uint32_t *s_red, uint64_t *__restrict__ chunk_bits) {
    const uint32_t my0 = c * kEncChunk + threadIdx.x * kEncPerThread;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < 16; k++)
        if (FULL || my0 + k < end) sum += entry_len(entry<LDS_TAB>(s_tab, tab, tw, v[k], offset), tw.narrow());
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sum += (uint32_t) __shfl_xor((int) sum, o);
    if (threadIdx.x % 64 == 0) s_red[threadIdx.x / 64] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (uint32_t w = 0; w < kEncThreads / 64; w++) t += s_red[w];
        chunk_bits[c] = t;
    }
    __syncthreads();
}

// OR a code of L bits (right-aligned in c, 1 <= L <= 56) into the MSB-first
// bit buffer at bit position pos: it touches word pos / 32 and, when it
// crosses word boundaries, the next one or two.  W is a word sink.
template<class W>
__device__ __forceinline__ void put_code(uint64_t c, uint32_t L, uint32_t pos, W &&word_or) {
    const uint64_t hi = c << (64 - L);     // codeword in the top L bits
    const uint32_t sh = pos & 31, w = pos >> 5;
    word_or(w, (uint32_t) (hi >> (32 + sh)));
    if (sh + L > 32) word_or(w + 1, (uint32_t) (hi >> sh));
    if (sh + L > 64) word_or(w + 2, (uint32_t) (hi << (32 - sh)));
}

// The thread's run of codes e[0..15] (narrow entries: length << 32 | the
// codeword left-aligned in 32 bits) at bit pos of the zeroed LDS bit buffer,
// MSB-first: each codeword shifted into a 64-bit pair of words at its bit
// offset and OR-ed into the two words it can touch (the second is OR-ed with
// zero when the codeword ends inside the first).  No branch per code, and
// words shared with the neighbouring runs need no special case.
__device__ __forceinline__ void put_run(const uint64_t *e, uint32_t pos, uint32_t *s_words) {
    uint32_t p = pos;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint64_t v = (uint64_t) (uint32_t) e[k] << (32 - (p & 31));
        uint32_t *w = s_words + (p >> 5);
        atomicOr(w, (uint32_t) (v >> 32));
        atomicOr(w + 1, (uint32_t) v);
        p += (uint32_t) (e[k] >> 32);
    }
}

// Bits of chunk c, which starts at bit g0 of the stream.
template<bool LDS_TAB, bool FULL>
__device__ __forceinline__ void pack_chunk(uint32_t c, const Codes16 &v, uint32_t end, uint64_t g0,
                                           const uint64_t *s_tab, const uint64_t *__restrict__ tab,
                                           const TabWindow &tw, int32_t offset, uint32_t *s_words, uint32_t *s_wave,
                                           uint32_t *__restrict__ words) {
    const uint32_t my0 = c * kEncChunk + threadIdx.x * kEncPerThread;
    uint64_t e[16];
    uint32_t mybits = 0;
    const bool narrow = tw.narrow();
#pragma unroll
    for (int k = 0; k < 16; k++) {
        e[k] = FULL || my0 + k < end ? entry<LDS_TAB>(s_tab, tab, tw, v[k], offset) : 0;
        mybits += entry_len(e[k], narrow);
    }
    uint32_t total;
    const uint32_t before = wg_exclusive(mybits, s_wave, &total);
    const uint32_t head = (uint32_t) (g0 & 31);
    const uint32_t nw = (head + total + 31) / 32;
    const uint64_t gw = g0 >> 5;
    uint32_t pos = head + before;
    if (nw < kPackWords && narrow) {  // a run's last OR may touch word nw
        for (uint32_t i = threadIdx.x; i < nw; i += kEncThreads) s_words[i] = 0;
        __syncthreads();
        put_run(e, pos, s_words);
        __syncthreads();
        const bool open_tail = ((head + total) & 31) != 0;
        for (uint32_t i = threadIdx.x; i < nw; i += kEncThreads) {
            const uint32_t val = bswap32(s_words[i]);
            if ((i == 0 && head) || (i == nw - 1 && open_tail))
                atomicOr(&words[gw + i], val);
            else
                words[gw + i] = val;
        }
    } else {
        // long codewords: straight into the zeroed output
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint32_t L = entry_len(e[k], narrow);
            const uint64_t c = narrow ? (uint64_t) ((uint32_t) e[k] >> (32 - L)) : e[k] >> 8;
            if (L) put_code(c, L, pos, [&](uint32_t w, uint32_t bits) { atomicOr(&words[gw + w], bswap32(bits)); });
            pos += L;
        }
    }
    __syncthreads();
}

// Full chunks with 16-byte aligned codes: the next chunk's codes (and, for
// the packer, its bit offset) are in flight while the current one is
// processed, with a fixed number of loads per step so the wait before use
// covers the current chunk only.  The last, partial chunk and unaligned
// codes take element loads.
__device__ __forceinline__ Codes16 fetch_full(const uint16_t *__restrict__ codes, uint32_t c) {
    const uint4 *q = reinterpret_cast<const uint4 *>(codes + (size_t) c * kEncChunk + threadIdx.x * kEncPerThread);
    Codes16 r;
    r.a = q[0];
    r.b = q[1];
    return r;
}

template<bool LDS_TAB>
__global__ void __launch_bounds__(kEncThreads) encode_sizes_kernel(const uint16_t *__restrict__ codes, uint32_t n,
                                                                  const uint64_t *__restrict__ tab, TabWindow tw,
                                                                  int32_t offset, uint64_t *__restrict__ chunk_bits,
                                                                  int aligned) {
    extern __shared__ uint64_t s_tab[];
    __shared__ uint32_t s_red[kEncThreads / 64];
    stage_table<LDS_TAB>(s_tab, tab, tw);
    const uint32_t nchunks = (n + kEncChunk - 1) / kEncChunk;
    const uint32_t nfull = aligned ? n / kEncChunk : 0;
    uint32_t c = blockIdx.x;
    if (c < nfull) {
        Codes16 v = fetch_full(codes, c);
        for (; c < nfull; c += gridDim.x) {
            // the next full chunk (the last one again at the end: unused)
            const Codes16 next = fetch_full(codes, min(c + gridDim.x, nfull - 1));
            sizes_chunk<LDS_TAB, true>(c, v, 0, s_tab, tab, tw, offset, s_red, chunk_bits);
            v = next;
        }
    }
    for (; c < nchunks; c += gridDim.x)
        sizes_chunk<LDS_TAB, false>(c, fetch_chunk(codes, c, n, aligned != 0), min(c * kEncChunk + kEncChunk, n),
                                    s_tab, tab, tw, offset, s_red, chunk_bits);
}

template<bool LDS_TAB>
__global__ void __launch_bounds__(kEncThreads) encode_pack_kernel(const uint16_t *__restrict__ codes, uint32_t n,
                                                                 const uint64_t *__restrict__ tab, TabWindow tw,
                                                                 int32_t offset, const uint64_t *__restrict__ chunk_off,
                                                                 uint32_t *__restrict__ words, int aligned) {
    extern __shared__ uint64_t s_tab[];
    __shared__ uint32_t s_words[kPackWords];
    __shared__ uint32_t s_wave[kEncThreads / 64];
    stage_table<LDS_TAB>(s_tab, tab, tw);
    const uint32_t nchunks = (n + kEncChunk - 1) / kEncChunk;
    const uint32_t nfull = aligned ? n / kEncChunk : 0;
    uint32_t c = blockIdx.x;
    if (c < nfull) {
        Codes16 v = fetch_full(codes, c);
        uint64_t g0 = chunk_off[c];
        for (; c < nfull; c += gridDim.x) {
            // the next full chunk (the last one again at the end: unused)
            const uint32_t cn = min(c + gridDim.x, nfull - 1);
            const Codes16 next = fetch_full(codes, cn);
            const uint64_t g0n = chunk_off[cn];
            pack_chunk<LDS_TAB, true>(c, v, 0, g0, s_tab, tab, tw, offset, s_words, s_wave, words);
            v = next;
            g0 = g0n;
        }
    }
    for (; c < nchunks; c += gridDim.x)
        pack_chunk<LDS_TAB, false>(c, fetch_chunk(codes, c, n, aligned != 0), min(c * kEncChunk + kEncChunk, n),
                                   chunk_off[c], s_tab, tab, tw, offset, s_words, s_wave, words);
}

// One-pass encode: workgroups all resident at once (the grid is one round of
// the device's occupancy), workgroup b taking chunks b, b + G, b + 2G, ... with
// the next chunk's codes in flight.  Each chunk publishes its bit total as
// soon as it is known and finds its bit offset by a decoupled look-back over
// the published totals and inclusive prefixes (Merrill and Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016), 64
// predecessors per step across wave 0; every predecessor of a chunk belongs
// to an earlier round or to a resident workgroup of the same round, so every
// wait ends.  The chunk's bits are assembled in LDS from bit 0 while the
// look-back runs and shifted into place as they are stored.  Every output
// word is stored by exactly one chunk: the one whose range holds the word's
// first bit.  A chunk starting inside a word hands its bits of that word to
// head_bits[c] instead, and a fix-up kernel ORs them in after the pack, so
// the output needs no zero fill.  Needs codewords of at most 32 bits.
constexpr uint64_t kFlagAgg = 1, kFlagPrefix = 2;
constexpr uint32_t kLookbackSpins = 1u << 24;

// Status words are read and written with relaxed device-scope atomics: they
// stay coherent across the XCDs (each has its own L2) without ordering any
// other access, which is all the look-back needs, since a word carries its
// flag and its value together and the bit words and offsets are read only by
// later launches.  Acquire / release would add an L2 invalidate / write-back
// to every access.
__device__ __forceinline__ uint64_t status_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive bit offset of chunk c > 0 (wave 0; every lane returns it).
__device__ __forceinline__ uint64_t look_back(const uint64_t *status, uint32_t c, uint32_t lane, uint32_t *err) {
    uint64_t prefix = 0;
    int64_t j = (int64_t) c - 1;
    uint32_t spins = 0;
    for (;;) {
        const int64_t idx = j - (int64_t) lane;
        // before chunk 0: an inclusive prefix of 0
        const uint64_t sv = idx >= 0 ? status_load(&status[idx]) : kFlagPrefix;
        const uint32_t flag = (uint32_t) (sv & 3);
        const uint64_t pmask = __ballot(flag == kFlagPrefix);
        const uint64_t zmask = __ballot(flag == 0);
        const uint32_t firstp = pmask ? (uint32_t) __ffsll((unsigned long long) pmask) - 1 : 64u;
        // every predecessor nearer than the first prefix must have its total out
        const uint64_t nearer = firstp < 64 ? ((1ull << firstp) - 1) : ~0ull;
        if (zmask & nearer) {
            if (++spins > kLookbackSpins) {
                if (lane == 0) atomicOr(err, 4u);
                return prefix;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        uint64_t add = lane <= firstp ? (sv >> 2) : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) add += (uint64_t) __shfl_xor((long long) add, o);
        prefix += add;
        if (firstp < 64) return prefix;
        j -= 64;
    }
}

template<bool LDS_TAB>
__global__ void __launch_bounds__(kEncThreads) encode_onepass_kernel(
        const uint16_t *__restrict__ codes, uint32_t n, const uint64_t *__restrict__ tab, TabWindow tw, int32_t offset,
        uint32_t *__restrict__ words, uint64_t *__restrict__ status, uint64_t *__restrict__ chunk_off,
        uint32_t *__restrict__ head_bits, uint64_t *__restrict__ total_bits, uint32_t *__restrict__ err,
        int aligned) {
    extern __shared__ uint64_t s_tab[];
    __shared__ uint32_t s_words[kPackWords + 1];
    __shared__ uint32_t s_wave[kEncThreads / 64];
    __shared__ uint64_t s_prefix;
    stage_table<LDS_TAB>(s_tab, tab, tw);
    const uint32_t nchunks = (n + kEncChunk - 1) / kEncChunk;
    const uint32_t nfull = aligned ? n / kEncChunk : 0;
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    uint32_t c = blockIdx.x;
    if (c >= nchunks) return;
    Codes16 v = c < nfull ? fetch_full(codes, c) : fetch_chunk(codes, c, n, aligned != 0);
    for (; c < nchunks; c += gridDim.x) {
        const uint32_t cn = c + gridDim.x;
        Codes16 next{};
        if (cn < nchunks) next = cn < nfull ? fetch_full(codes, cn) : fetch_chunk(codes, cn, n, aligned != 0);
        const uint32_t end = min(c * kEncChunk + kEncChunk, n);
        const uint32_t my0 = c * kEncChunk + threadIdx.x * kEncPerThread;
        uint64_t e[16];
        uint32_t mybits = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            e[k] = my0 + k < end ? entry<LDS_TAB>(s_tab, tab, tw, v[k], offset) : 0;
            mybits += (uint32_t) (e[k] >> 32);  // narrow entries only
        }
        uint32_t total;
        const uint32_t before = wg_exclusive(mybits, s_wave, &total);
        // the total goes out at once; the chunk's bits are assembled from bit 0
        if (threadIdx.x == 0)
            status_store(&status[c], ((uint64_t) total << 2) | (c == 0 ? kFlagPrefix : kFlagAgg));
        const uint32_t nwz = (total + 31) / 32 + 1;
        for (uint32_t i = threadIdx.x; i < nwz; i += kEncThreads) s_words[i] = 0;
        __syncthreads();
        put_run(e, before, s_words);
        if (wave == 0) {
            const uint64_t prefix = c == 0 ? 0 : look_back(status, c, lane, err);
            if (lane == 0) {
                if (c) status_store(&status[c], ((prefix + total) << 2) | kFlagPrefix);
                s_prefix = prefix;
                chunk_off[c] = prefix;
                if (c == nchunks - 1) *total_bits = prefix + total;
            }
        }
        __syncthreads();
        // out: word i of the output range takes the chunk's bits shifted right
        // by head (the bits of the word before come from s_words[i - 1])
        const uint64_t g0 = s_prefix;
        const uint32_t head = (uint32_t) (g0 & 31);
        const uint32_t nw = (head + total + 31) / 32;
        const uint64_t gw = g0 >> 5;
        for (uint32_t i = threadIdx.x; i < nw; i += kEncThreads) {
            const uint32_t hi = i ? s_words[i - 1] : 0u, lo = s_words[i];
            const uint32_t val = bswap32(head ? (hi << (32 - head)) | (lo >> head) : lo);
            if (i == 0 && head) head_bits[c] = val;  // the word's first bit belongs to an earlier chunk
            else words[gw + i] = val;
        }
        if (threadIdx.x == 0 && (nw == 0 || !head)) head_bits[c] = 0;
        __syncthreads();
        v = next;
    }
}

// ORs every chunk's bits of the word it starts inside into the output.
__global__ void encode_head_fix_kernel(const uint64_t *__restrict__ chunk_off, const uint32_t *__restrict__ head_bits,
                                       uint32_t nchunks, uint32_t *__restrict__ words) {
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        const uint32_t hb = head_bits[c];
        if (hb) atomicOr(&words[chunk_off[c] >> 5], hb);
    }
}

uint32_t encode_grid(uint32_t nchunks) {
    // the batch's worker threads enqueue encodes concurrently: thread-safe
    // static initialisation
    static const uint32_t ncu = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            return (uint32_t) v;
        return 256u;
    }();
    return std::min<uint32_t>(nchunks, ncu * kEncGridPerCu);
}

}  // namespace

void launch_encode_sizes(const uint16_t *d_codes, size_t n, const uint64_t *d_tab, const EncodeWindow &w,
                         int32_t offset, uint64_t *d_chunk_bits, hipStream_t st) {
    const uint32_t nchunks = (uint32_t) ((n + kEncChunk - 1) / kEncChunk);
    if (!nchunks) return;
    const int aligned = ((uintptr_t) d_codes & 15) == 0;
    const TabWindow tw{w.win, w.nwin, w.e0, w.max_len};
    if (w.nwin && w.nwin <= kTabLds)
        hipLaunchKernelGGL(encode_sizes_kernel<true>, dim3(encode_grid(nchunks)), dim3(kEncThreads), tab_lds_bytes(w.nwin), st,
                           d_codes, (uint32_t) n, d_tab, tw, offset, d_chunk_bits, aligned);
    else
        hipLaunchKernelGGL(encode_sizes_kernel<false>, dim3(encode_grid(nchunks)), dim3(kEncThreads), 0, st, d_codes,
                           (uint32_t) n, d_tab, tw, offset, d_chunk_bits, aligned);
}

void launch_encode_pack(const uint16_t *d_codes, size_t n, const uint64_t *d_tab, const EncodeWindow &w,
                        int32_t offset, const uint64_t *d_chunk_off, uint32_t *d_words, hipStream_t st) {
    const uint32_t nchunks = (uint32_t) ((n + kEncChunk - 1) / kEncChunk);
    if (!nchunks) return;
    const int aligned = ((uintptr_t) d_codes & 15) == 0;
    const TabWindow tw{w.win, w.nwin, w.e0, w.max_len};
    if (w.nwin && w.nwin <= kTabLds)
        hipLaunchKernelGGL(encode_pack_kernel<true>, dim3(encode_grid(nchunks)), dim3(kEncThreads), tab_lds_bytes(w.nwin), st,
                           d_codes, (uint32_t) n, d_tab, tw, offset, d_chunk_off, d_words, aligned);
    else
        hipLaunchKernelGGL(encode_pack_kernel<false>, dim3(encode_grid(nchunks)), dim3(kEncThreads), 0, st, d_codes,
                           (uint32_t) n, d_tab, tw, offset, d_chunk_off, d_words, aligned);
}

uint32_t encode_lds_entries() { return kTabLds; }

void launch_encode_onepass(const uint16_t *d_codes, size_t n, const uint64_t *d_tab, const EncodeWindow &w,
                           int32_t offset, uint32_t *d_words, uint64_t *d_status, uint64_t *d_chunk_off,
                           uint32_t *d_head_bits, uint64_t *d_total_bits, uint32_t *d_err, hipStream_t st) {
    const uint32_t nchunks = (uint32_t) ((n + kEncChunk - 1) / kEncChunk);
    if (!nchunks) return;
    const int aligned = ((uintptr_t) d_codes & 15) == 0;
    const TabWindow tw{w.win, w.nwin, w.e0, w.max_len};
    const bool lds = w.nwin && w.nwin <= kTabLds;
    const size_t shmem = lds ? tab_lds_bytes(w.nwin) : 0;
    // one round of resident workgroups (the look-back's waits need every
    // predecessor's workgroup running)
    int dev = 0, cus = 0, per = 0;
    uint32_t resident = 256;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, lds ? (const void *) encode_onepass_kernel<true> : (const void *) encode_onepass_kernel<false>,
            kEncThreads, shmem) == hipSuccess &&
        cus > 0 && per > 0)
        resident = (uint32_t) (cus * per);
    const dim3 grid(std::min<uint32_t>(nchunks, resident));
    if (lds)
        hipLaunchKernelGGL(encode_onepass_kernel<true>, grid, dim3(kEncThreads), shmem, st, d_codes, (uint32_t) n,
                           d_tab, tw, offset, d_words, d_status, d_chunk_off, d_head_bits, d_total_bits, d_err,
                           aligned);
    else
        hipLaunchKernelGGL(encode_onepass_kernel<false>, grid, dim3(kEncThreads), 0, st, d_codes, (uint32_t) n, d_tab,
                           tw, offset, d_words, d_status, d_chunk_off, d_head_bits, d_total_bits, d_err, aligned);
    hipLaunchKernelGGL(encode_head_fix_kernel, dim3(std::min<uint32_t>((nchunks + 255) / 256, 1024)), dim3(256), 0,
                       st, d_chunk_off, d_head_bits, nchunks, d_words);
}

}  // namespace faz
