// Allocation-free Huffman encode: the kernels of faz_gpu_huffman_encode
// (chunk bit totals, exclusive scan, zero fill, pack) with every scratch
// array carved out of one file-scope device array instead of context
// buffers, for callers that may not allocate device memory (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// The array's name is a host-side stand-in, not a device address: host code
// uses only the address hipGetSymbolAddress returns for the current device,
// fetched once per device when the first context on that device is created
// (faz_noalloc_init) and kept per device.
#include "faz_internal.h"
#include "faz_ctx.h"

#include <mutex>

namespace faz {

namespace {

constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Largest input: 2^30 codes (a 4 GiB float field), 65537 table states
// (stateNum = max - min + 2 for 16-bit codes).
constexpr size_t kMaxCodes = (size_t) 1 << 30;
constexpr uint32_t kMaxChunks = (uint32_t) (kMaxCodes / kEncChunk);
constexpr uint32_t kMaxStates = 65537;
constexpr uint32_t kMaxScanTiles = (kMaxChunks + 4095) / 4096 + 1;

// Layout of the scratch array (byte offsets, each region 256-byte aligned).
constexpr size_t kTabOff = 0;                                           // packed table, u64[states]
constexpr size_t kCodeOff = align_up(kTabOff + 8 * (size_t) kMaxStates);  // uploaded codewords, u64[states]
constexpr size_t kLenOff = align_up(kCodeOff + 8 * (size_t) kMaxStates);  // uploaded lengths, u8[states]
constexpr size_t kBitsOff = align_up(kLenOff + kMaxStates);               // chunk bit totals, u64[chunks]
constexpr size_t kOffOff = align_up(kBitsOff + 8 * (size_t) kMaxChunks);  // chunk bit offsets, u64[chunks]
constexpr size_t kTilesOff = align_up(kOffOff + 8 * (size_t) kMaxChunks); // scan tile sums
constexpr size_t kTotalOff = align_up(kTilesOff + 8 * (size_t) kMaxScanTiles);  // bit total, u64
constexpr size_t kScratchBytes = align_up(kTotalOff + 8);

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

// Packed entry per table state (encode_entry).
__global__ void pack_table_kernel(const uint64_t *__restrict__ code, const uint8_t *__restrict__ len, uint32_t ns,
                                  int narrow, uint64_t *__restrict__ tab) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x)
        tab[i] = encode_entry(code[i], len[i], narrow != 0);
}

}  // namespace

}  // namespace faz

static __device__ __align__(256) unsigned char scratch[faz::kScratchBytes];

namespace faz {

hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_err[device] = e;
        g_ptr[device] = e == hipSuccess ? static_cast<unsigned char *>(p) : nullptr;
        if (e != hipSuccess) (void) hipGetLastError();
        g_done[device] = true;
    }
    return g_err[device];
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

}  // namespace faz

using namespace faz;

extern "C" {

faz_status faz_gpu_huffman_encode_noalloc(faz_gpu_ctx *ctx, const uint16_t *d_codes, size_t n,
                                          const faz_huff_table *t, uint8_t *d_bits, size_t cap_bytes,
                                          size_t *nbytes) {
    if (!ctx || !d_codes || !n || !t || !nbytes || !t->code || !t->len || t->n_states <= 0) return FAZ_ERR_INVALID;
    if (n > kMaxCodes)
        CTX_FAIL(FAZ_ERR_INVALID, "allocation-free encode takes at most %zu codes (got %zu)", kMaxCodes, n);
    if ((uint32_t) t->n_states > kMaxStates)
        CTX_FAIL(FAZ_ERR_INVALID, "allocation-free encode takes at most %u table states (got %d)", kMaxStates,
                 t->n_states);
    if (((uintptr_t) d_bits & 3) != 0) CTX_FAIL(FAZ_ERR_INVALID, "bit buffer must be 4-byte aligned");
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices && g_done[ctx->device]) base = g_ptr[ctx->device];
    }
    if (!base) CTX_FAIL(FAZ_ERR_HIP, "no scratch address for device %d (hipGetSymbolAddress failed)", ctx->device);
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const uint32_t ns = (uint32_t) t->n_states;
    // window of the present symbols other than 0 (read from the caller's lengths)
    int64_t lo = -1, hi = -1;
    uint32_t maxlen = 0;
    for (uint32_t i = 0; i < ns; i++) {
        const uint32_t L = t->len[i];
        if (L > 56) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "Huffman codeword of %u bits (at most 56 supported)", L);
        maxlen = std::max(maxlen, L);
        if (L && (int64_t) i + t->offset != 0) {
            if (lo < 0) lo = i;
            hi = i;
        }
    }
    EncodeWindow ew{0, 0, 0};
    ew.max_len = maxlen;
    if (lo >= 0 && hi - lo + 1 <= (int64_t) encode_lds_entries()) {
        ew.win = (uint32_t) lo;
        ew.nwin = (uint32_t) (hi - lo + 1);
        const int64_t zi = -(int64_t) t->offset;
        if (zi >= 0 && zi < (int64_t) ns && t->len[zi])
            ew.e0 = encode_entry(t->code[zi], t->len[zi], maxlen <= 32);
    }
    uint64_t *tab = reinterpret_cast<uint64_t *>(base + kTabOff);
    uint64_t *code = reinterpret_cast<uint64_t *>(base + kCodeOff);
    uint8_t *len = base + kLenOff;
    uint64_t *chunk_bits = reinterpret_cast<uint64_t *>(base + kBitsOff);
    uint64_t *chunk_off = reinterpret_cast<uint64_t *>(base + kOffOff);
    uint64_t *tiles = reinterpret_cast<uint64_t *>(base + kTilesOff);
    uint64_t *total = reinterpret_cast<uint64_t *>(base + kTotalOff);
    const uint32_t nchunks = (uint32_t) ((n + kEncChunk - 1) / kEncChunk);
    if (scan_tiles(nchunks) > kMaxScanTiles) CTX_FAIL(FAZ_ERR_INVALID, "internal: scan tiles exceed the scratch");
    HIPCHECK(hipMemcpyAsync(code, t->code, (size_t) ns * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(len, t->len, ns, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(pack_table_kernel, dim3(std::min<uint32_t>((ns + 255) / 256, 1024)), dim3(256), 0, st, code,
                       len, ns, (int) (maxlen <= 32), tab);
    launch_encode_sizes(d_codes, n, tab, ew, t->offset, chunk_bits, st);
    launch_scan_u64(chunk_bits, chunk_off, nchunks, tiles, total, st);
    HIPCHECK(hipGetLastError());
    uint64_t bits = 0;
    HIPCHECK(hipMemcpyAsync(&bits, total, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const size_t bytes = (size_t) ((bits + 7) / 8);
    const size_t words = (size_t) ((bits + 31) / 32);
    *nbytes = bytes;
    if (!d_bits || cap_bytes < words * 4) CTX_FAIL(FAZ_ERR_CAPACITY, "bit buffer needs %zu bytes", words * 4);
    uint32_t *w = reinterpret_cast<uint32_t *>(d_bits);
    launch_fill_u32(w, words, 0u, st);
    launch_encode_pack(d_codes, n, tab, ew, t->offset, chunk_off, w, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    return FAZ_OK;
}

size_t faz_gpu_noalloc_scratch_bytes(void) { return noalloc_scratch_bytes(); }

}  // extern "C"
