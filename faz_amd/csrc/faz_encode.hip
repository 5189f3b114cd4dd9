// Huffman encode (HuffmanEncoder::encode, HuffmanEncoder.hpp:150-227) on the
// GPU: chunk bit totals, an exclusive scan of them (faz_kernels.hip), then
// per-chunk packing of the MSB-first bitstream into 32-bit words stored with
// big-endian byte order (the reference writes 64-bit big-endian stores of an
// MSB-first stream, which is the same byte sequence).
//
// Workgroups are persistent over chunks of kEncChunk codes, so the code table
// is staged in LDS (sized to the table, at most kTabLds entries) once per
// workgroup.  Every code is placed on its own: its
// bit position is the chunk offset plus a workgroup prefix of lengths, and its
// bits are OR-ed into at most three words of an LDS bit buffer, with no
// sequential accumulator per thread.  Interior words then go out with plain
// stores; the first and last word of a chunk, shared with its neighbours, with
// atomic ORs into the zeroed output.
#include "faz_internal.h"

namespace faz {

namespace {

constexpr uint32_t kTabLds = 4096;                    // table entries staged in LDS
constexpr uint32_t kPackWords = kEncChunk + 4;        // bit buffer: 32 bits per code on average
constexpr uint32_t kEncGridPerCu = 8;

__device__ __forceinline__ uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

// 16 consecutive codes from position j (zero past end), 16-byte loads when
// the codes are 16-byte aligned.
__device__ __forceinline__ void load16(const uint16_t *__restrict__ codes, uint32_t j, uint32_t end, bool aligned,
                                       uint16_t *v) {
    if (aligned && j + 16 <= end) {
        const uint4 *q = reinterpret_cast<const uint4 *>(codes + j);
        const uint4 a = q[0], b = q[1];
        const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            v[2 * k] = (uint16_t) w[k];
            v[2 * k + 1] = (uint16_t) (w[k] >> 16);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = j + k < end ? codes[j + k] : 0;
    }
}

template<bool LDS_TAB>
__device__ __forceinline__ uint64_t entry(const uint64_t *s_tab, const uint64_t *__restrict__ tab, int32_t i) {
    return LDS_TAB ? s_tab[i] : tab[i];
}

template<bool LDS_TAB>
__device__ __forceinline__ void stage_table(uint64_t *s_tab, const uint64_t *__restrict__ tab, uint32_t n_states) {
    if (LDS_TAB) {
        for (uint32_t i = threadIdx.x; i < n_states; i += kEncThreads) s_tab[i] = tab[i];
        __syncthreads();
    }
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

template<bool LDS_TAB>
__global__ void __launch_bounds__(kEncThreads) encode_sizes_kernel(const uint16_t *__restrict__ codes, uint32_t n,
                                                                  const uint64_t *__restrict__ tab,
                                                                  uint32_t n_states, int32_t offset,
                                                                  uint64_t *__restrict__ chunk_bits, int aligned) {
    extern __shared__ uint64_t s_tab[];
    __shared__ uint32_t s_red[kEncThreads / 64];
    stage_table<LDS_TAB>(s_tab, tab, n_states);
    const uint32_t nchunks = (n + kEncChunk - 1) / kEncChunk;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint32_t base = c * kEncChunk, end = min(base + kEncChunk, n);
        const uint32_t my0 = base + threadIdx.x * kEncPerThread;
        uint16_t v[16];
        load16(codes, my0, end, aligned != 0, v);
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (my0 + k < end) sum += (uint32_t) entry<LDS_TAB>(s_tab, tab, (int32_t) v[k] - offset) & 0xFF;
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

template<bool LDS_TAB>
__global__ void __launch_bounds__(kEncThreads) encode_pack_kernel(const uint16_t *__restrict__ codes, uint32_t n,
                                                                 const uint64_t *__restrict__ tab,
                                                                 uint32_t n_states, int32_t offset,
                                                                 const uint64_t *__restrict__ chunk_off,
                                                                 uint32_t *__restrict__ words, int aligned) {
    extern __shared__ uint64_t s_tab[];
    __shared__ uint32_t s_words[kPackWords];
    __shared__ uint32_t s_wave[kEncThreads / 64];
    stage_table<LDS_TAB>(s_tab, tab, n_states);
    const uint32_t nchunks = (n + kEncChunk - 1) / kEncChunk;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint32_t base = c * kEncChunk, end = min(base + kEncChunk, n);
        const uint32_t my0 = base + threadIdx.x * kEncPerThread;
        uint16_t v[16];
        load16(codes, my0, end, aligned != 0, v);
        uint64_t e[16];
        uint32_t mybits = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            e[k] = my0 + k < end ? entry<LDS_TAB>(s_tab, tab, (int32_t) v[k] - offset) : 0;
            mybits += (uint32_t) e[k] & 0xFF;
        }
        uint32_t total;
        const uint32_t before = wg_exclusive(mybits, s_wave, &total);
        const uint64_t g0 = chunk_off[c];
        const uint32_t head = (uint32_t) (g0 & 31);
        const uint32_t nw = (head + total + 31) / 32;
        const uint64_t gw = g0 >> 5;
        uint32_t pos = head + before;
        if (nw <= kPackWords) {
            for (uint32_t i = threadIdx.x; i < nw; i += kEncThreads) s_words[i] = 0;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const uint32_t L = (uint32_t) e[k] & 0xFF;
                if (L) put_code(e[k] >> 8, L, pos, [&](uint32_t w, uint32_t bits) { atomicOr(&s_words[w], bits); });
                pos += L;
            }
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
                const uint32_t L = (uint32_t) e[k] & 0xFF;
                if (L)
                    put_code(e[k] >> 8, L, pos,
                             [&](uint32_t w, uint32_t bits) { atomicOr(&words[gw + w], bswap32(bits)); });
                pos += L;
            }
        }
        __syncthreads();
    }
}

uint32_t encode_grid(uint32_t nchunks) {
    static uint32_t ncu = 0;
    if (!ncu) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            ncu = (uint32_t) v;
        else
            ncu = 256;
    }
    return std::min<uint32_t>(nchunks, ncu * kEncGridPerCu);
}

}  // namespace

void launch_encode_sizes(const uint16_t *d_codes, size_t n, const uint64_t *d_tab, uint32_t n_states, int32_t offset,
                         uint64_t *d_chunk_bits, hipStream_t st) {
    const uint32_t nchunks = (uint32_t) ((n + kEncChunk - 1) / kEncChunk);
    if (!nchunks) return;
    const int aligned = ((uintptr_t) d_codes & 15) == 0;
    const dim3 grid(encode_grid(nchunks)), block(kEncThreads);
    if (n_states <= kTabLds)
        hipLaunchKernelGGL(encode_sizes_kernel<true>, grid, block, n_states * 8, st, d_codes, (uint32_t) n, d_tab,
                           n_states, offset, d_chunk_bits, aligned);
    else
        hipLaunchKernelGGL(encode_sizes_kernel<false>, grid, block, 0, st, d_codes, (uint32_t) n, d_tab, n_states,
                           offset, d_chunk_bits, aligned);
}

void launch_encode_pack(const uint16_t *d_codes, size_t n, const uint64_t *d_tab, uint32_t n_states, int32_t offset,
                        const uint64_t *d_chunk_off, uint32_t *d_words, hipStream_t st) {
    const uint32_t nchunks = (uint32_t) ((n + kEncChunk - 1) / kEncChunk);
    if (!nchunks) return;
    const int aligned = ((uintptr_t) d_codes & 15) == 0;
    const dim3 grid(encode_grid(nchunks)), block(kEncThreads);
    if (n_states <= kTabLds)
        hipLaunchKernelGGL(encode_pack_kernel<true>, grid, block, n_states * 8, st, d_codes, (uint32_t) n, d_tab,
                           n_states, offset, d_chunk_off, d_words, aligned);
    else
        hipLaunchKernelGGL(encode_pack_kernel<false>, grid, block, 0, st, d_codes, (uint32_t) n, d_tab, n_states,
                           offset, d_chunk_off, d_words, aligned);
}

}  // namespace faz
