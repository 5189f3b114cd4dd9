// Byte layout of the interpolation engine buffer that the reference hands to
// zstd (SZInterpolationCompressor.hpp:489-518 write, 53-92 read) and of the
// LinearQuantizer section (IntegerQuantizer.hpp:185-226).  Raw native
// little-endian fields, as QoZ::write/read copy them.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <cstring>
#include <vector>

namespace faz {

struct EngineHeader {
    int N = 3;
    uint64_t dims[4] = {0, 0, 0, 0};
    unsigned block_size = 32;
    int interp_algo = 1;
    int direction = 0;
    double alpha = -1, beta = -1;
    uint64_t max_step = 0;
    int levelwise = 0;
    std::vector<uint8_t> ops, dirs;
};

template<class V>
inline void put_raw(uint8_t *&w, const V &v) {
    memcpy(w, &v, sizeof(V));
    w += sizeof(V);
}

template<class V>
inline bool get_raw(const uint8_t *&r, const uint8_t *end, V &v) {
    if ((size_t) (end - r) < sizeof(V)) return false;
    memcpy(&v, r, sizeof(V));
    r += sizeof(V);
    return true;
}

inline size_t engine_header_size(const EngineHeader &h) {
    return (size_t) h.N * 8 + 4 + 4 + 4 + 8 + 8 + 8 + 4 + 1 + 4 + 8 + 4 + 4 + 2 * (size_t) h.levelwise;
}

inline void write_engine_header(const EngineHeader &h, uint8_t *&w) {
    for (int i = 0; i < h.N; i++) put_raw(w, (uint64_t) h.dims[i]);
    put_raw(w, h.block_size);
    put_raw(w, h.interp_algo);
    put_raw(w, h.direction);
    put_raw(w, h.alpha);
    put_raw(w, h.beta);
    put_raw(w, (uint64_t) h.max_step);
    put_raw(w, h.levelwise);
    put_raw(w, (uint8_t) 0);   // blockwiseTuning
    put_raw(w, (int) 0);       // fixBlockSize
    put_raw(w, (uint64_t) 0);  // cross_block
    put_raw(w, (int) 0);       // trimToZero
    put_raw(w, (int) 0);       // blockOrder
    for (int i = 0; i < h.levelwise; i++) put_raw(w, h.ops[i]);
    for (int i = 0; i < h.levelwise; i++) put_raw(w, h.dirs[i]);
}

// Returns false on truncation or on engine modes the GPU path does not serve
// (blockwise tuning, fixed block size, cross-block, trimToZero, block order).
inline bool read_engine_header(const uint8_t *&r, size_t n, int N, EngineHeader &h) {
    const uint8_t *end = r + n;
    h.N = N;
    for (int i = 0; i < N; i++)
        if (!get_raw(r, end, h.dims[i])) return false;
    uint8_t blockwise;
    int fixbs, ttz, border;
    uint64_t cross;
    if (!get_raw(r, end, h.block_size) || !get_raw(r, end, h.interp_algo) || !get_raw(r, end, h.direction) ||
        !get_raw(r, end, h.alpha) || !get_raw(r, end, h.beta) || !get_raw(r, end, h.max_step) ||
        !get_raw(r, end, h.levelwise) || !get_raw(r, end, blockwise) || !get_raw(r, end, fixbs) ||
        !get_raw(r, end, cross) || !get_raw(r, end, ttz) || !get_raw(r, end, border))
        return false;
    if (blockwise || fixbs || cross || ttz || border) return false;
    if (h.levelwise < 0 || h.levelwise > 32 || (size_t) (end - r) < 2 * (size_t) h.levelwise) return false;
    h.ops.assign(r, r + h.levelwise);
    r += h.levelwise;
    h.dirs.assign(r, r + h.levelwise);
    r += h.levelwise;
    return true;
}

inline void write_quantizer_head(double eb, int radius, uint64_t n_unpred, uint8_t *&w) {
    put_raw(w, (uint8_t) 2);
    put_raw(w, eb);
    put_raw(w, radius);
    put_raw(w, n_unpred);
}

inline bool read_quantizer_head(const uint8_t *&r, size_t n, double &eb, int &radius, uint64_t &n_unpred) {
    const uint8_t *end = r + n;
    uint8_t tag;
    return get_raw(r, end, tag) && get_raw(r, end, eb) && get_raw(r, end, radius) && get_raw(r, end, n_unpred);
}

inline size_t engine_raw_size(const EngineHeader &h, uint64_t n_unpred, size_t tree_bytes, size_t bit_bytes) {
    return engine_header_size(h) + 1 + 8 + 4 + 8 + n_unpred * 4 + tree_bytes + 8 + bit_bytes;
}

}  // namespace faz
