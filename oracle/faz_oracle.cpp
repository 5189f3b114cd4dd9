// CPU oracle for the FAZ compression hot path — TEST INFRASTRUCTURE ONLY.
//
// A scalar restatement of the reference algorithm (JLiu-1/FAZ, namespace QoZ),
// following its operation order exactly so that results are bit-identical.
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it,
// and only as the checker; the product in faz_amd/ never links or calls it.
//
// Built with -ffp-contract=off: the reference is built for x86-64 without FMA,
// so every float/double operation below rounds individually.
//
// Parity pins (tests/test_oracle_golden.py, tests/test_oracle_math.py):
//   quantizer, interpolators, block/level/pass order, anchors, tuning modes,
//   Huffman (ska::unordered_map order, heap, codes, tree bytes, bitstream),
//   engine buffer, Config trailer and stream: golden vectors written by the
//   reference engine compiled from its own sources (oracle/_ref/ref_harness)
//   plus closed forms and invariants.
#include "faz_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <vector>

extern "C" {
// libzstd (system shared library); prototypes as published in zstd.h
size_t ZSTD_compress(void *dst, size_t dstCapacity, const void *src, size_t srcSize, int compressionLevel);
size_t ZSTD_decompress(void *dst, size_t dstCapacity, const void *src, size_t compressedSize);
size_t ZSTD_compressBound(size_t srcSize);
unsigned ZSTD_isError(size_t code);
}

namespace {

// ---------------------------------------------------------------------------
// LinearQuantizer (include/QoZ/quantizer/IntegerQuantizer.hpp:14-257)
// ---------------------------------------------------------------------------

// (int) of a non-negative double as x86-64 cvttsd2si does it: values outside
// the int range (and NaN) give INT_MIN.
static inline int32_t x86_trunc_i32(double v) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return INT32_MIN;
    return (int32_t) v;
}

struct Quantizer {
    double eb = 1, recip = 1;
    int radius = 32768;
    std::vector<float> unpred;
    size_t index = 0;  // decompression cursor

    void set_eb(double e) {  // Q:28-31
        eb = e;
        recip = 1.0 / e;
    }

    // quantize_and_overwrite(T&, T, bool) with trimToZero == 0 (Q:69-108)
    int quantize_and_overwrite(float &data, float pred, bool save_unpred = true) {
        float diff = data - pred;
        uint32_t qi = (uint32_t) x86_trunc_i32(std::fabs(diff) * recip) + 1u;
        if ((int32_t) qi < radius * 2) {
            int32_t q = ((int32_t) qi) >> 1;  // arithmetic shift
            int32_t half = q;
            uint32_t qs = (uint32_t) q << 1;
            int32_t shifted;
            if (diff < 0) {
                qs = 0u - qs;
                shifted = radius - half;
            } else {
                shifted = radius + half;
            }
            float dec = (float) ((double) pred + (double) (int32_t) qs * eb);
            if (std::fabs(dec - data) > eb) {
                if (save_unpred) unpred.push_back(data);
                return 0;
            }
            data = dec;
            return shifted;
        }
        if (save_unpred) unpred.push_back(data);
        return 0;
    }

    // recover (Q:160-180)
    float recover(float pred, int code) {
        if (code) return (float) ((double) pred + (double) (2 * (code - radius)) * eb);
        return unpred[index++];
    }
};

// ---------------------------------------------------------------------------
// Interpolators (include/QoZ/utils/Interpolators.hpp:9-36), evaluated in T=float
// left to right; linear1 uses double constants.
// ---------------------------------------------------------------------------
static inline float interp_linear(float a, float b) { return (a + b) / 2; }
static inline float interp_linear1(float a, float b) { return (float) (-0.5 * a + 1.5 * b); }
static inline float interp_quad_1(float a, float b, float c) { return (3 * a + 6 * b - c) / 8; }
static inline float interp_quad_2(float a, float b, float c) { return (-a + 6 * b + 3 * c) / 8; }
static inline float interp_quad_3(float a, float b, float c) { return (3 * a - 10 * b + 15 * c) / 8; }
static inline float interp_cubic(float a, float b, float c, float d) { return (-a + 9 * b + 9 * c - d) / 16; }

// ---------------------------------------------------------------------------
// SZInterpolationCompressor (include/QoZ/compressor/SZInterpolationCompressor.hpp)
// ---------------------------------------------------------------------------
struct InterpEngine {
    int N = 3;
    size_t dims[4] = {0, 0, 0, 0};
    size_t offs[4] = {0, 0, 0, 0};
    size_t num = 0;
    int levels = -1;
    bool anchor = false;
    size_t max_step = 0;
    unsigned block_size = 32;
    int interp_algo = 1, direction = 0;
    double alpha = -1, beta = -1;
    std::vector<uint8_t> ops, dirs;
    std::vector<std::array<int, 4>> perms;

    Quantizer quantizer;
    std::vector<int> codes;
    std::vector<size_t> code_pos;  // flat index of each pushed code (seq-index probe)
    size_t code_index = 0;         // decompression cursor
    bool decompressing = false;
    int tuning = 0;

    void set_params(const oracle_interp_params *p) {
        N = p->N;
        for (int i = 0; i < N; i++) dims[i] = p->dims[i];
        max_step = p->max_step;
        block_size = (unsigned) p->block_size;
        interp_algo = p->interp_algo;
        direction = p->direction;
        alpha = p->alpha;
        beta = p->beta;
        ops.assign(p->ops, p->ops + p->levelwise);
        dirs.assign(p->dirs, p->dirs + p->levelwise);
        quantizer.radius = p->radius;
        quantizer.set_eb(p->eb);
    }

    // init (SIC:820-852)
    void init() {
        num = 1;
        levels = -1;
        for (int i = 0; i < N; i++) {
            if (levels < std::ceil(std::log2((double) dims[i]))) levels = (unsigned) std::ceil(std::log2((double) dims[i]));
            num *= dims[i];
        }
        anchor = false;
        if (max_step > 0) {
            anchor = true;
            levels += 1;
            int max_level = (unsigned) std::log2((double) max_step) + 1;
            if (max_level <= levels) levels = max_level;
        }
        offs[N - 1] = 1;
        for (int i = N - 2; i >= 0; i--) offs[i] = offs[i + 1] * dims[i + 1];
        perms.clear();
        std::array<int, 4> seq = {0, 1, 2, 3};
        do {
            perms.push_back(seq);
        } while (std::next_permutation(seq.begin(), seq.begin() + N));
    }

    // level-wise error bound (SIC:386-406, 113-137)
    void set_level_eb(double eb, unsigned level) {
        if (alpha < 0) {
            quantizer.set_eb(level >= 3 ? eb * 0.5 : eb);
        } else if (alpha >= 1) {
            double r = std::pow(alpha, level - 1);
            if (r > beta) r = beta;
            quantizer.set_eb(eb / r);
        } else {
            double r = 1 - (level - 1) * alpha;
            if (r < beta) r = beta;
            quantizer.set_eb(eb * r);
        }
    }

    void level_op(unsigned level, int &interp, int &dir) const {  // SIC:407-422
        int L = (int) ops.size();
        if (L == 0) {
            interp = interp_algo;
            dir = direction;
        } else if ((int) level - 1 < L) {
            interp = ops[level - 1];
            dir = dirs[level - 1];
        } else {
            interp = ops[L - 1];
            dir = dirs[L - 1];
        }
    }

    // one predicted point: quantize (SIC:924-938) / quantize_tuning (SIC:940-974) / recover (SIC:976-978)
    double point(float *data, float *d, float pred) {
        if (decompressing) {
            *d = quantizer.recover(pred, codes[code_index++]);
            return 0;
        }
        if (tuning == 0) {
            codes.push_back(quantizer.quantize_and_overwrite(*d, pred));
            code_pos.push_back((size_t) (d - data));
            return 0;
        }
        if (tuning == 1) {
            float orig = *d;
            codes.push_back(quantizer.quantize_and_overwrite(*d, pred));
            code_pos.push_back((size_t) (d - data));
            return (*d - orig) * (*d - orig);
        }
        double pred_error = std::fabs(*d - pred);
        quantizer.quantize_and_overwrite(*d, pred, false);
        return pred_error;
    }

    // block_interpolation_1d (SIC:1306-1417)
    double line(float *data, size_t begin, size_t end, size_t stride, bool linear) {
        size_t n = (end - begin) / stride + 1;
        if (n <= 1) return 0;
        double err = 0;
        size_t s3 = 3 * stride, s5 = 5 * stride;
        if (linear || n < 5) {
            for (size_t i = 1; i + 1 < n; i += 2) {
                float *d = data + begin + i * stride;
                err += point(data, d, interp_linear(*(d - stride), *(d + stride)));
            }
            if (n % 2 == 0) {
                float *d = data + begin + (n - 1) * stride;
                if (n < 4)
                    err += point(data, d, *(d - stride));
                else
                    err += point(data, d, interp_linear1(*(d - s3), *(d - stride)));
            }
        } else {
            float *d;
            size_t i;
            for (i = 3; i + 3 < n; i += 2) {
                d = data + begin + i * stride;
                err += point(data, d, interp_cubic(*(d - s3), *(d - stride), *(d + stride), *(d + s3)));
            }
            d = data + begin + stride;
            err += point(data, d, interp_quad_1(*(d - stride), *(d + stride), *(d + s3)));
            d = data + begin + i * stride;
            err += point(data, d, interp_quad_2(*(d - s3), *(d - stride), *(d + stride)));
            if (n % 2 == 0) {
                d = data + begin + (n - 1) * stride;
                err += point(data, d, interp_quad_3(*(d - s5), *(d - s3), *(d - stride)));
            }
        }
        return err;
    }

    // block_interpolation<N> (SIC:3076-3442): pass P runs along axis p[P];
    // axes p[Q<P] step by stride from (begin? begin+stride : 0), axes p[Q>P]
    // step by 2*stride from (begin? begin+2*stride : 0); earlier axes outer.
    double block(float *data, const size_t *begin, const size_t *end, bool linear, int dir, size_t stride) {
        const std::array<int, 4> &p = perms[dir];
        double err = 0;
        for (int P = 0; P < N; P++) {
            int other[3];
            int no = 0;
            for (int Q = 0; Q < N; Q++)
                if (Q != P) other[no++] = Q;
            size_t st[3], lo[3], hi[3], c[3];
            for (int k = 0; k < no; k++) {
                int Q = other[k];
                int ax = p[Q];
                st[k] = Q < P ? stride : 2 * stride;
                lo[k] = begin[ax] ? begin[ax] + st[k] : 0;
                hi[k] = end[ax];
                c[k] = lo[k];
            }
            int ax = p[P];
            // odometer over the outer coordinates (earlier permutation index outer)
            bool empty = false;
            for (int k = 0; k < no; k++)
                if (lo[k] > hi[k]) empty = true;
            if (empty) continue;
            while (true) {
                size_t off = begin[ax] * offs[ax];
                for (int k = 0; k < no; k++) off += c[k] * offs[p[other[k]]];
                err += line(data, off, off + (end[ax] - begin[ax]) * offs[ax], stride * offs[ax], linear);
                int k = no - 1;
                while (k >= 0) {
                    c[k] += st[k];
                    if (c[k] <= hi[k]) break;
                    c[k] = lo[k];
                    k--;
                }
                if (k < 0) break;
            }
        }
        return err;
    }

    // one level: blocks in row-major raster of multi_dimensional_range (IT:361-453)
    double level_pass(float *data, unsigned level) {
        int interp, dir;
        level_op(level, interp, dir);
        size_t stride = (size_t) 1U << (level - 1);
        size_t bs = (size_t) block_size * stride;
        size_t nb[4], bi[4] = {0, 0, 0, 0};
        size_t total = 1;
        for (int i = 0; i < N; i++) {
            nb[i] = (dims[i] - 1) / bs + 1;
            total *= nb[i];
        }
        double err = 0;
        for (size_t b = 0; b < total; b++) {
            size_t begin[4], end[4];
            for (int i = 0; i < N; i++) {
                begin[i] = bi[i] * bs;
                end[i] = begin[i] + bs;
                if (end[i] > dims[i] - 1) end[i] = dims[i] - 1;
            }
            err += block(data, begin, end, interp == 0, dir, stride);
            for (int i = N - 1; i >= 0; i--) {
                if (++bi[i] < nb[i]) break;
                bi[i] = 0;
            }
        }
        return err;
    }

    // build_grid (SIC:854-900): only N == 2 and N == 3 have anchors
    void build_grid(float *data) {
        if (tuning > 1) return;
        size_t s0 = max_step * (tuning == 1);
        if (N == 2) {
            for (size_t x = s0; x < dims[0]; x += max_step)
                for (size_t y = s0; y < dims[1]; y += max_step) {
                    quantizer.unpred.push_back(data[x * dims[1] + y]);
                    codes.push_back(0);
                    code_pos.push_back(x * dims[1] + y);
                }
        } else if (N == 3) {
            for (size_t x = s0; x < dims[0]; x += max_step)
                for (size_t y = s0; y < dims[1]; y += max_step)
                    for (size_t z = s0; z < dims[2]; z += max_step) {
                        size_t idx = x * offs[0] + y * offs[1] + z;
                        quantizer.unpred.push_back(data[idx]);
                        codes.push_back(0);
                        code_pos.push_back(idx);
                    }
        }
    }

    // compress (SIC:328-476), returns the summed loss of the tuning modes
    double compress(float *data, int tuning_, int start_level, int end_level, std::vector<size_t> &level_counts) {
        tuning = tuning_;
        init();
        codes.clear();
        code_pos.clear();
        quantizer.unpred.clear();
        double eb = quantizer.eb;
        if (tuning) level_counts.assign(levels, 0);
        if (start_level <= 0 || start_level > levels) start_level = levels;
        if (end_level >= start_level || end_level < 0) end_level = 0;
        if (!anchor) {
            codes.push_back(quantizer.quantize_and_overwrite(*data, 0));
            code_pos.push_back(0);
        } else if (start_level == levels) {
            if (tuning) level_counts[start_level - 1] = codes.size();
            build_grid(data);
            start_level--;
        }
        double err = 0;
        for (unsigned level = start_level; level > (unsigned) end_level && level <= (unsigned) start_level; level--) {
            set_level_eb(eb, level);
            err += level_pass(data, level);
            if (tuning) level_counts[level - 1] = codes.size();
        }
        quantizer.set_eb(eb);
        return err;
    }

    // decompress level loop (SIC:94-196)
    void decompress(float *data) {
        decompressing = true;
        code_index = 0;
        quantizer.index = 0;
        init();
        double eb = quantizer.eb;
        if (!anchor) {
            *data = quantizer.recover(0, codes[code_index++]);
        } else {
            if (N == 2) {
                for (size_t x = 0; x < dims[0]; x += max_step)
                    for (size_t y = 0; y < dims[1]; y += max_step) {
                        data[x * offs[0] + y] = quantizer.unpred[quantizer.index++];
                        code_index++;
                    }
            } else if (N == 3) {
                for (size_t x = 0; x < dims[0]; x += max_step)
                    for (size_t y = 0; y < dims[1]; y += max_step)
                        for (size_t z = 0; z < dims[2]; z += max_step) {
                            data[x * offs[0] + y * offs[1] + z] = quantizer.unpred[quantizer.index++];
                            code_index++;
                        }
            }
            levels--;
        }
        for (unsigned level = levels; level > 0 && level <= (unsigned) levels; level--) {
            set_level_eb(eb, level);
            level_pass(data, level);
        }
        quantizer.set_eb(eb);
    }
};

// ---------------------------------------------------------------------------
// HuffmanEncoder<int> (include/QoZ/encoder/HuffmanEncoder.hpp)
// ---------------------------------------------------------------------------

// ska::unordered_map<int,size_t> iteration order after inserting the keys in
// first-occurrence order (utils/ska_hash/unordered_map.hpp:254-303, 412-446,
// 604-643; flat_hash_map.hpp:1268-1295): power-of-two buckets indexed by
// Fibonacci hashing, growth to max(4, 2B) when count+1 > B, rehash walks old
// buckets upward pushing each chain element to the front of its new chain,
// new keys go to the chain front, iteration runs from the top bucket down.
static std::vector<int> ska_iteration_order(const std::vector<int> &keys_in_first_occurrence) {
    std::vector<std::vector<int>> chains;  // chain[i][0] = head
    size_t B = 1;
    int shift = 63;
    bool allocated = false;
    auto index = [&](int key) -> size_t {
        return (size_t) ((11400714819323198485ull * (uint64_t) (size_t) (long) key) >> shift);
    };
    size_t count = 0;
    for (int key: keys_in_first_occurrence) {
        if (!allocated || (double) (count + 1) > (double) B * 1.0) {
            size_t nb = std::max<size_t>(4, 2 * (allocated ? B : 1));
            size_t p2 = 1;
            while (p2 < nb) p2 <<= 1;
            nb = std::max<size_t>(2, p2);
            int log2nb = 0;
            while (((size_t) 1 << log2nb) < nb) log2nb++;
            std::vector<std::vector<int>> fresh(nb);
            shift = 64 - log2nb;
            if (allocated) {
                for (size_t i = 0; i < B; i++)
                    for (int e: chains[i]) {
                        auto &c = fresh[index(e)];
                        c.insert(c.begin(), e);
                    }
            }
            chains.swap(fresh);
            B = nb;
            allocated = true;
        }
        auto &c = chains[index(key)];
        c.insert(c.begin(), key);
        count++;
    }
    std::vector<int> order;
    for (size_t i = B; i-- > 0;)
        for (int e: chains[i]) order.push_back(e);
    return order;
}

struct Huffman {
    struct Node {
        int left = -1, right = -1;
        uint64_t freq = 0;
        uint8_t t = 0;
        int c = 0;
    };
    int offset = 0;
    int stateNum = 0;
    std::vector<Node> pool;
    std::vector<int> qq;  // 1-based heap of pool indices
    int qend = 1;
    std::vector<uint64_t> code;
    std::vector<uint8_t> len;
    std::vector<bool> has_code;
    int root = -1;
    unsigned nodeCount = 0;

    int new_leaf(uint64_t freq, int c) {
        Node n;
        n.freq = freq;
        n.c = c;
        n.t = 1;
        pool.push_back(n);
        return (int) pool.size() - 1;
    }
    int new_inner(int a, int b) {
        Node n;
        n.left = a;
        n.right = b;
        n.freq = pool[a].freq + pool[b].freq;
        n.t = 0;
        pool.push_back(n);
        return (int) pool.size() - 1;
    }
    void qinsert(int n) {  // HUF:455-463
        int j, i = qend++;
        if ((int) qq.size() <= i) qq.resize(i + 1);
        while ((j = (i >> 1))) {
            if (pool[qq[j]].freq <= pool[n].freq) break;
            qq[i] = qq[j], i = j;
        }
        qq[i] = n;
    }
    int qremove() {  // HUF:465-485
        int i = 1, l;
        int n = qq[i];
        if (qend < 2) return -1;
        qend--;
        qq[i] = qq[qend];
        while ((l = (i << 1)) < qend) {
            if (l + 1 < qend && pool[qq[l + 1]].freq < pool[qq[l]].freq) l++;
            if (pool[qq[i]].freq > pool[qq[l]].freq) {
                std::swap(qq[i], qq[l]);
                i = l;
            } else {
                break;
            }
        }
        return n;
    }
    bool build_code(int n, int length, uint64_t out1) {  // HUF:493-521 (codes <= 64 bits)
        if (pool[n].t) {
            if (length > 64) return false;
            code[pool[n].c] = length == 0 ? 0 : (out1 << (64 - length));
            len[pool[n].c] = (uint8_t) length;
            has_code[pool[n].c] = true;
            return true;
        }
        if (length >= 64) return false;
        out1 <<= 1;
        if (!build_code(pool[n].left, length + 1, out1)) return false;
        return build_code(pool[n].right, length + 1, out1 | 1);
    }

    // init (HUF:529-566) + preprocess_encode (HUF:95-116)
    bool init(const int32_t *s, size_t n) {
        std::vector<int> first_order;
        std::vector<uint64_t> freq;
        int mn = s[0], mx = s[0];
        for (size_t i = 0; i < n; i++) {
            mn = std::min(mn, s[i]);
            mx = std::max(mx, s[i]);
        }
        // counts and first-occurrence order (dense table over [mn, mx])
        std::vector<uint64_t> cnt((size_t) ((int64_t) mx - mn + 1), 0);
        for (size_t i = 0; i < n; i++) {
            uint64_t &c = cnt[(size_t) ((int64_t) s[i] - mn)];
            if (c == 0) first_order.push_back(s[i]);
            c++;
        }
        offset = mn;
        stateNum = mx - offset + 2;
        pool.clear();
        qq.assign(2, 0);
        qend = 1;
        std::vector<int> order = ska_iteration_order(first_order);
        for (int key: order) qinsert(new_leaf(cnt[(size_t) ((int64_t) key - mn)], key - offset));
        // qinsert(new_node(0, 0, qremove(), qremove())): g++ evaluates the two
        // calls right to left, so the first removed node becomes the right child.
        while (qend > 2) {
            int first = qremove();
            int second = qremove();
            qinsert(new_inner(second, first));
        }
        root = qq[1];
        code.assign(stateNum, 0);
        len.assign(stateNum, 0);
        has_code.assign(stateNum, false);
        if (!build_code(root, 0, 0)) return false;
        unsigned k = 0;
        for (int i = 0; i < stateNum; i++)
            if (has_code[i]) k++;
        nodeCount = k * 2 - 1;
        return true;
    }

    static void put_be32(uint8_t *b, int32_t v) {
        b[0] = (uint8_t) (v >> 24);
        b[1] = (uint8_t) (v >> 16);
        b[2] = (uint8_t) (v >> 8);
        b[3] = (uint8_t) v;
    }

    template<class T1>
    void pad_tree(std::vector<T1> &L, std::vector<T1> &R, std::vector<int32_t> &C, std::vector<uint8_t> &t,
                  unsigned i, int node, unsigned &n_inode) {  // HUF:569-584
        C[i] = pool[node].c;
        t[i] = pool[node].t;
        if (pool[node].left >= 0) {
            n_inode++;
            L[i] = (T1) n_inode;
            pad_tree(L, R, C, t, n_inode, pool[node].left, n_inode);
        }
        if (pool[node].right >= 0) {
            n_inode++;
            R[i] = (T1) n_inode;
            pad_tree(L, R, C, t, n_inode, pool[node].right, n_inode);
        }
    }

    template<class T1>
    void tree_bytes(std::vector<uint8_t> &out) {  // HUF:607-633
        std::vector<T1> L(nodeCount, 0), R(nodeCount, 0);
        std::vector<int32_t> C(nodeCount, 0);
        std::vector<uint8_t> t(nodeCount, 0);
        unsigned n_inode = 0;
        pad_tree(L, R, C, t, 0, root, n_inode);
        out.push_back(0);  // little-endian marker
        const uint8_t *p = (const uint8_t *) L.data();
        out.insert(out.end(), p, p + nodeCount * sizeof(T1));
        p = (const uint8_t *) R.data();
        out.insert(out.end(), p, p + nodeCount * sizeof(T1));
        p = (const uint8_t *) C.data();
        out.insert(out.end(), p, p + nodeCount * sizeof(int32_t));
        out.insert(out.end(), t.begin(), t.end());
    }

    // save (HUF:119-137)
    void save(std::vector<uint8_t> &out) {
        uint8_t b[4];
        memcpy(b, &offset, 4);
        out.insert(out.end(), b, b + 4);
        put_be32(b, (int32_t) nodeCount);
        out.insert(out.end(), b, b + 4);
        put_be32(b, (int32_t) (stateNum / 2));
        out.insert(out.end(), b, b + 4);
        if (nodeCount <= 256)
            tree_bytes<uint8_t>(out);
        else if (nodeCount <= 65536)
            tree_bytes<uint16_t>(out);
        else
            tree_bytes<uint32_t>(out);
    }

    // encode (HUF:150-227): size_t byte count, then the codes MSB-first
    void encode(const int32_t *s, size_t n, std::vector<uint8_t> &out) {
        size_t start = out.size();
        out.resize(start + 8, 0);
        std::vector<uint8_t> bits;
        uint64_t acc = 0;
        int nacc = 0;  // bits pending in acc (MSB-aligned at bit 63)
        for (size_t i = 0; i < n; i++) {
            int st = s[i] - offset;
            int l = len[st];
            uint64_t c = code[st];
            for (int k = 0; k < l; k++) {
                uint64_t bit = (c >> (63 - k)) & 1;
                acc |= bit << (63 - nacc);
                if (++nacc == 8) {
                    bits.push_back((uint8_t) (acc >> 56));
                    acc = 0;
                    nacc = 0;
                }
            }
        }
        if (nacc) bits.push_back((uint8_t) (acc >> 56));
        uint64_t outSize = bits.size();
        memcpy(out.data() + start, &outSize, 8);
        out.insert(out.end(), bits.begin(), bits.end());
    }
};

// load + decode (HUF:236-296, 310-454)
struct HuffDecoder {
    struct Node {
        int left = -1, right = -1;
        int c = 0;
        uint8_t t = 0;
    };
    std::vector<Node> nodes;
    int offset = 0;

    static int32_t get_be32(const uint8_t *b) {
        return (int32_t) (((uint32_t) b[0] << 24) | ((uint32_t) b[1] << 16) | ((uint32_t) b[2] << 8) | b[3]);
    }
    template<class T1>
    void unpad(const T1 *L, const T1 *R, const int32_t *C, const uint8_t *t, unsigned i, int node) {
        if (nodes[node].t == 0) {
            T1 l = L[i];
            if (l != 0) {
                Node m;
                m.c = C[l];
                m.t = t[l];
                nodes.push_back(m);
                int id = (int) nodes.size() - 1;
                nodes[node].left = id;
                unpad(L, R, C, t, l, id);
            }
            T1 r = R[i];
            if (r != 0) {
                Node m;
                m.c = C[r];
                m.t = t[r];
                nodes.push_back(m);
                int id = (int) nodes.size() - 1;
                nodes[node].right = id;
                unpad(L, R, C, t, r, id);
            }
        }
    }
    template<class T1>
    size_t load_tree(const uint8_t *b, unsigned nc) {
        std::vector<T1> L(nc), R(nc);
        std::vector<int32_t> C(nc);
        std::vector<uint8_t> t(nc);
        memcpy(L.data(), b + 1, nc * sizeof(T1));
        memcpy(R.data(), b + 1 + nc * sizeof(T1), nc * sizeof(T1));
        memcpy(C.data(), b + 1 + 2 * nc * sizeof(T1), nc * 4);
        memcpy(t.data(), b + 1 + 2 * nc * sizeof(T1) + nc * 4, nc);
        nodes.clear();
        Node root;
        if (nc <= 256) {
            root.c = C[0];
            root.t = t[0];
        }
        nodes.push_back(root);
        unpad(L.data(), R.data(), C.data(), t.data(), 0, 0);
        return 1 + 2 * nc * sizeof(T1) + nc * 4 + nc;
    }
    // returns bytes consumed by load + decode, 0 on error
    size_t decode(const uint8_t *buf, size_t n, int32_t *out, size_t num) {
        if (n < 12) return 0;
        memcpy(&offset, buf, 4);
        unsigned nc = (unsigned) get_be32(buf + 4);
        size_t used = 12;
        if (nc <= 256)
            used += load_tree<uint8_t>(buf + 12, nc);
        else if (nc <= 65536)
            used += load_tree<uint16_t>(buf + 12, nc);
        else
            used += load_tree<uint32_t>(buf + 12, nc);
        uint64_t enc;
        memcpy(&enc, buf + used, 8);
        used += 8;
        const uint8_t *bytes = buf + used;
        if (nodes[0].t) {
            for (size_t i = 0; i < num; i++) out[i] = nodes[0].c + offset;
            return used + enc;
        }
        int nd = 0;
        size_t count = 0;
        for (size_t i = 0; count < num; i++) {
            if ((i >> 3) >= enc) return 0;
            int bit = (bytes[i >> 3] >> (7 - (i & 7))) & 1;
            nd = bit ? nodes[nd].right : nodes[nd].left;
            if (nd < 0) return 0;
            if (nodes[nd].t) {
                out[count++] = nodes[nd].c + offset;
                nd = 0;
            }
        }
        return used + enc;
    }
};

template<class V>
static void put(std::vector<uint8_t> &out, const V &v) {
    const uint8_t *p = (const uint8_t *) &v;
    out.insert(out.end(), p, p + sizeof(V));
}

template<class V>
static V take(const uint8_t *&p) {
    V v;
    memcpy(&v, p, sizeof(V));
    p += sizeof(V);
    return v;
}

// engine header + quantizer + Huffman (SIC:485-518, Q:185-201)
static bool interp_raw(const oracle_interp_params *prm, const float *input, std::vector<uint8_t> &out) {
    InterpEngine e;
    e.set_params(prm);
    std::vector<float> data(input, input + [&] {
        size_t n = 1;
        for (int i = 0; i < prm->N; i++) n *= prm->dims[i];
        return n;
    }());
    std::vector<size_t> lc;
    e.compress(data.data(), 0, 0, 0, lc);
    Huffman h;
    if (!h.init(e.codes.data(), e.codes.size())) return false;
    out.clear();
    for (int i = 0; i < prm->N; i++) put(out, (size_t) prm->dims[i]);
    put(out, (unsigned) prm->block_size);
    put(out, (int) prm->interp_algo);
    put(out, (int) prm->direction);
    put(out, prm->alpha);
    put(out, prm->beta);
    put(out, (size_t) prm->max_step);
    put(out, (int) prm->levelwise);
    put(out, (bool) false);   // blockwiseTuning
    put(out, (int) 0);        // fixBlockSize
    put(out, (size_t) 0);     // cross_block
    put(out, (int) 0);        // trimToZero
    put(out, (int) 0);        // blockOrder
    out.insert(out.end(), prm->ops, prm->ops + prm->levelwise);
    out.insert(out.end(), prm->dirs, prm->dirs + prm->levelwise);
    out.push_back(2);
    put(out, prm->eb);
    put(out, (int) prm->radius);
    put(out, (size_t) e.quantizer.unpred.size());
    const uint8_t *up = (const uint8_t *) e.quantizer.unpred.data();
    out.insert(out.end(), up, up + e.quantizer.unpred.size() * sizeof(float));
    h.save(out);
    h.encode(e.codes.data(), e.codes.size(), out);
    return true;
}

static size_t interp_header_len(int N, int levels) {
    return N * 8 + 4 + 4 + 4 + 8 + 8 + 8 + 4 + 1 + 4 + 8 + 4 + 4 + 2 * (size_t) levels;
}

static bool interp_decompress_raw(const uint8_t *raw, size_t n, int N, float *out, size_t num) {
    const uint8_t *p = raw;
    oracle_interp_params prm;
    memset(&prm, 0, sizeof(prm));
    prm.N = N;
    size_t total = 1;
    for (int i = 0; i < N; i++) {
        prm.dims[i] = take<size_t>(p);
        total *= prm.dims[i];
    }
    if (total != num) return false;
    prm.block_size = (int) take<unsigned>(p);
    prm.interp_algo = take<int>(p);
    prm.direction = take<int>(p);
    prm.alpha = take<double>(p);
    prm.beta = take<double>(p);
    prm.max_step = take<size_t>(p);
    prm.levelwise = take<int>(p);
    bool blockwise = take<bool>(p);
    int fixbs = take<int>(p);
    size_t cross = take<size_t>(p);
    int ttz = take<int>(p);
    int border = take<int>(p);
    if (blockwise || fixbs || cross || ttz || border) return false;  // out-of-scope engine modes
    if (prm.levelwise < 0 || prm.levelwise > 32) return false;
    memcpy(prm.ops, p, prm.levelwise);
    p += prm.levelwise;
    memcpy(prm.dirs, p, prm.levelwise);
    p += prm.levelwise;
    p += 1;
    prm.eb = take<double>(p);
    prm.radius = take<int>(p);
    size_t U = take<size_t>(p);
    InterpEngine e;
    e.set_params(&prm);
    e.quantizer.unpred.assign((const float *) p, (const float *) p + U);
    p += U * sizeof(float);
    e.codes.resize(num);
    HuffDecoder hd;
    if (!hd.decode(p, n - (p - raw), e.codes.data(), num)) return false;
    e.decompress(out);
    return true;
}

// Config fields of the trailer that differ between the two stream kinds.
struct TrailerFields {
    uint8_t cmprAlgo = 2;  // ALGO_INTERP (Config.hpp:24-25: LORENZO_REG=0, INTERP_LORENZO=1, INTERP=2)
    bool lorenzo = true, lorenzo2 = false, regression = true;
    int blockSize = -1;    // -1: the Config constructor default for N
    int pred_dim = -1;     // -1: N
};

// QoZ::Config trailer (Config.hpp:240-309)
static void config_save(std::vector<uint8_t> &out, const oracle_interp_params *p, double absEB, double relEB,
                        const TrailerFields &tf = TrailerFields()) {
    size_t num = 1;
    for (int i = 0; i < p->N; i++) num *= p->dims[i];
    put(out, (char) p->N);
    for (int i = 0; i < p->N; i++) put(out, (size_t) p->dims[i]);
    put(out, num);
    put(out, tf.cmprAlgo);
    put(out, (uint8_t) 0);  // errorBoundMode = EB_ABS (after calAbsErrorBound)
    put(out, (uint8_t) 0);  // tuningTarget = RD
    put(out, absEB);
    put(out, relEB);
    put(out, p->alpha);
    put(out, p->beta);
    put(out, tf.lorenzo);
    put(out, tf.lorenzo2);
    put(out, tf.regression);
    put(out, (bool) false);  // regression2
    put(out, (uint8_t) p->interp_algo);
    put(out, (uint8_t) p->direction);
    put(out, (int) p->block_size);
    put(out, (uint8_t) 1);  // lossless = zstd
    put(out, (uint8_t) 1);  // encoder = Huffman
    put(out, (int) (p->radius * 2));
    int blockSize = tf.blockSize >= 0 ? tf.blockSize : (p->N == 1 ? 128 : (p->N == 2 ? 16 : 6));
    int stride = p->N == 1 ? 128 : (p->N == 2 ? 16 : 6);  // set by setDims only
    put(out, blockSize);
    put(out, (int) p->levelwise);
    put(out, (bool) false);  // blockwiseTuning
    put(out, stride);
    put(out, (size_t) p->max_step);
    put(out, (int) (tf.pred_dim >= 0 ? tf.pred_dim : p->N));
    put(out, (bool) false);  // openmp
    put(out, (int) 0);       // fixBlockSize
    put(out, (int) 0);       // blockwiseSampleBlockSize
    put(out, (int) 0);       // crossBlock
    put(out, (int) 0);       // wavelet
    put(out, (size_t) 0);    // firstSize (indeterminate in the reference)
    put(out, (int) 0);       // offsetPredictor
    put(out, num);           // coeffs_num
    put(out, (size_t) p->N);
    for (int i = 0; i < p->N; i++) put(out, (size_t) p->dims[i]);
    put(out, (int) 0);       // pid
    put(out, (int) 0);       // blockOrder
    put(out, (int) -1);      // sperr
    put(out, (bool) false);  // sperrWithoutWave
    put(out, (bool) true);   // pyBind
    put(out, (size_t) 0);    // metadata size
    put(out, (int) 0);       // conditioning
}

// ---------------------------------------------------------------------------
// 3-D Lorenzo/regression engine: SZFastFrontend (include/QoZ/frontend/
// SZFastFrontend.hpp:176-345, 429-530) with the predictors of
// MetaLorenzoPredictor.hpp and MetaRegressionPredictor.hpp and the constants
// of MetaDef.hpp:16-31, composed by SZGeneralCompressor (SZGeneralCompressor.hpp:30-71).
// ---------------------------------------------------------------------------
enum { SEL_LORENZO = 0, SEL_REGRESSION = 1, SEL_LORENZO_2LAYER = 2 };

struct LREngine {
    size_t d1 = 0, d2 = 0, d3 = 0, s0 = 0, s1 = 0;
    int bs = 6, pred_dim = 3;
    bool lo = true, lo2 = false, reg = true;
    double eb = 0;
    float prec_f = 0;          // meta_params' float copy of the bound (MD:44-66)
    float reg_linear = 0, reg_indep = 0;
    size_t nx = 0, ny = 0, nz = 0, nblocks = 0;
    Quantizer quantizer;
    std::vector<int32_t> codes, indicator, reg_types;
    std::vector<float> reg_unpred, reg_coeffs, raw_coeffs, recon;

    void set(const oracle_lr_params *p) {
        d1 = p->dims[0], d2 = p->dims[1], d3 = p->dims[2];
        s0 = d2 * d3, s1 = d3;
        bs = p->block_size, pred_dim = p->pred_dim;
        lo = p->lorenzo != 0, lo2 = p->lorenzo2 != 0, reg = p->regression != 0;
        eb = p->eb;
        prec_f = (float) eb;
        // meta_params(): reg_eb_base(1) * precision / RegCoeffNum3d / (float) bs, all float (MD:59-65)
        reg_linear = 1.0f * prec_f / 4 / (float) bs;
        reg_indep = (float) bs * reg_linear;
        nx = (d1 - 1) / bs + 1, ny = (d2 - 1) / bs + 1, nz = (d3 - 1) / bs + 1;
        nblocks = nx * ny * nz;
        quantizer.set_eb(eb);
        quantizer.radius = p->radius;
    }

    // reconstructed value, 0 below the low faces (the zero-padded pred_buffer, FF:228-243)
    float rv(long x, long y, long z) const {
        if (x < 0 || y < 0 || z < 0) return 0.0f;
        return recon[(size_t) x * s0 + (size_t) y * s1 + (size_t) z];
    }

    // MetaLorenzoPredictor.hpp:9-77 on an accessor f(dx,dy,dz) of offsets <= 0
    // (dim0 = x, dim1 = y, unit stride = z); float arithmetic left to right.
    template<class F>
    static float lorenzo(F f, int pd, bool two) {
        if (!two) {
            if (pd == 3)
                return f(0, 0, -1) + f(0, -1, 0) + f(-1, 0, 0) - f(0, -1, -1) - f(-1, 0, -1) - f(-1, -1, 0) +
                       f(-1, -1, -1);
            if (pd == 2) return f(0, 0, -1) + f(-1, 0, 0) - f(-1, 0, -1);
            return f(0, 0, -1);
        }
        if (pd == 3)
            return 2 * f(0, 0, -1) - f(0, 0, -2) + 2 * f(0, -1, 0) - 4 * f(0, -1, -1) + 2 * f(0, -1, -2) -
                   f(0, -2, 0) + 2 * f(0, -2, -1) - f(0, -2, -2) + 2 * f(-1, 0, 0) - 4 * f(-1, 0, -1) +
                   2 * f(-1, 0, -2) - 4 * f(-1, -1, 0) + 8 * f(-1, -1, -1) - 4 * f(-1, -1, -2) +
                   2 * f(-1, -2, 0) - 4 * f(-1, -2, -1) + 2 * f(-1, -2, -2) - f(-2, 0, 0) + 2 * f(-2, 0, -1) -
                   f(-2, 0, -2) + 2 * f(-2, -1, 0) - 4 * f(-2, -1, -1) + 2 * f(-2, -1, -2) - f(-2, -2, 0) +
                   2 * f(-2, -2, -1) - f(-2, -2, -2);
        if (pd == 2)
            return 2 * f(-1, 0, 0) - f(-2, 0, 0) + 2 * f(0, 0, -1) - 4 * f(-1, 0, -1) + 2 * f(-2, 0, -1) -
                   f(0, 0, -2) + 2 * f(-1, 0, -2) - f(-2, 0, -2);
        return 2 * f(0, 0, -1) - f(0, 0, -2);
    }

    // compute_regression_coeffcients_3d (MRP:117-145), float sums in loop order
    void regression_coeffs(size_t x0, size_t y0, size_t z0, int sx, int sy, int sz, const float *data, float *c) const {
        float fx = 0, fy = 0, fz = 0, f = 0, sum_x, sum_y;
        for (int i = 0; i < sx; i++) {
            sum_x = 0;
            for (int j = 0; j < sy; j++) {
                sum_y = 0;
                const float *row = data + (x0 + i) * s0 + (y0 + j) * s1 + z0;
                for (int k = 0; k < sz; k++) {
                    float v = row[k];
                    sum_y += v;
                    fz += v * k;
                }
                fy += sum_y * j;
                sum_x += sum_y;
            }
            fx += sum_x * i;
            f += sum_x;
        }
        float coeff = (float) (1.0 / (sx * sy * sz));
        c[0] = (2 * fx / (sx - 1) - f) * 6 * coeff / (sx + 1);
        c[1] = (2 * fy / (sy - 1) - f) * 6 * coeff / (sy + 1);
        c[2] = (2 * fz / (sz - 1) - f) * 6 * coeff / (sz + 1);
        c[3] = f * coeff - ((sx - 1) * c[0] / 2 + (sy - 1) * c[1] / 2 + (sz - 1) * c[2] / 2);
    }

    static float reg_predict(const float *c, int x, int y, int z) {  // MRP:148-152, 169-171
        return c[0] * (float) x + c[1] * (float) y + c[2] * (float) z + c[3];
    }

    // meta_blockwise_selection_3d + meta_block_error_estimation_3d (FF:429-530):
    // samples of the ORIGINAL data inside the block.
    int select(size_t x0, size_t y0, size_t z0, int min_size, bool use_reg, const float *c, const float *data) const {
        double e_lo = 0, e_lo2 = 0, e_reg = 0;
        auto est = [&](int x, int y, int z) {
            const float *pos = data + (x0 + x) * s0 + (y0 + y) * s1 + (z0 + z);
            auto f = [&](int dx, int dy, int dz) { return pos[(long) dx * (long) s0 + (long) dy * (long) s1 + dz]; };
            float cur = *pos;
            float noise = 0, noise2 = 0;
            if (use_reg) e_reg += std::fabs(cur - reg_predict(c, x, y, z));
            double lp = 0, lp2 = 0;
            const double n1 = pred_dim == 3 ? 1.22 : (pred_dim == 2 ? 0.81 : 0.5);
            const double n2 = pred_dim == 3 ? 6.8 : (pred_dim == 2 ? 2.76 : 1.08);
            if (lo2) {
                lp2 = lorenzo(f, pred_dim, true);
                noise2 = (float) (n2 * prec_f);
            }
            if (lo) {
                lp = lorenzo(f, pred_dim, false);
                noise = (float) (n1 * prec_f);
            }
            e_lo += std::fabs((double) cur - lp) + noise;
            e_lo2 += std::fabs((double) cur - lp2) + noise2;
        };
        for (int i = 2; i < min_size - 1; i++) {
            int b = min_size - i;
            est(i, i, i);
            est(i, i, b);
            est(i, b, i);
            est(i, b, b);
        }
        if (min_size > 3) est(min_size - 1, min_size - 1, min_size - 1);
        if (use_reg && (!lo || e_reg <= e_lo) && (!lo2 || e_reg < e_lo2)) return SEL_REGRESSION;
        if (lo2 && (!lo || e_lo2 <= e_lo) && (!use_reg || e_lo2 <= e_reg)) return SEL_LORENZO_2LAYER;
        return SEL_LORENZO;
    }

    // quantize_reg_coeff (MRP:10-27)
    int quantize_coeff(float pred, float cur, double precision, double recip, float *dec) {
        double diff = cur - pred;
        double qd = std::fabs(diff) * recip + 1;
        if (qd < 65536) {
            qd = (diff > 0) ? qd : -qd;
            int qi = (int) (qd / 2) + 32768;
            float d = (float) (pred + 2 * (qi - 32768) * precision);
            *dec = d;
            if (std::fabs(d - cur) <= precision) return qi;
        }
        *dec = cur;
        reg_unpred.push_back(cur);
        return 0;
    }

    // compress_3d (FF:176-345)
    void compress(const float *data) {
        size_t n = d1 * d2 * d3;
        codes.clear(), indicator.clear(), reg_types.clear(), reg_unpred.clear(), reg_coeffs.clear();
        quantizer.unpred.clear();
        raw_coeffs.assign(4 * nblocks, 0.0f);
        recon.assign(n, 0.0f);
        codes.reserve(n);
        float prev[4] = {0, 0, 0, 0};
        const float prc[4] = {reg_linear, reg_linear, reg_linear, reg_indep};
        float rcp[4];
        for (int i = 0; i < 4; i++) rcp[i] = (float) (1.0 / prc[i]);
        size_t b = 0;
        for (size_t bi = 0; bi < nx; bi++)
            for (size_t bj = 0; bj < ny; bj++)
                for (size_t bk = 0; bk < nz; bk++, b++) {
                    size_t x0 = bi * bs, y0 = bj * bs, z0 = bk * bs;
                    int sx = (bi + 1) * bs < d1 ? bs : (int) (d1 - x0);
                    int sy = (bj + 1) * bs < d2 ? bs : (int) (d2 - y0);
                    int sz = (bk + 1) * bs < d3 ? bs : (int) (d3 - z0);
                    int min_size = std::min(std::min(sx, sy), sz);
                    bool use_reg = reg && min_size >= 2;
                    float c[4] = {0, 0, 0, 0};
                    if (use_reg) regression_coeffs(x0, y0, z0, sx, sy, sz, data, c);
                    memcpy(&raw_coeffs[4 * b], c, sizeof(c));
                    int sel = select(x0, y0, z0, min_size, use_reg, c, data);
                    indicator.push_back(sel);
                    if (sel == SEL_REGRESSION) {
                        for (int t = 0; t < 4; t++) {
                            float d;
                            reg_types.push_back(quantize_coeff(prev[t], c[t], prc[t], rcp[t], &d));
                            c[t] = d;
                        }
                        memcpy(prev, c, sizeof(c));
                        reg_coeffs.insert(reg_coeffs.end(), c, c + 4);
                    }
                    for (int i = 0; i < sx; i++)
                        for (int j = 0; j < sy; j++)
                            for (int k = 0; k < sz; k++) {
                                size_t x = x0 + i, y = y0 + j, z = z0 + k;
                                float pred;
                                if (sel == SEL_REGRESSION) {
                                    pred = reg_predict(c, i, j, k);
                                } else {
                                    auto f = [&](int dx, int dy, int dz) {
                                        return rv((long) x + dx, (long) y + dy, (long) z + dz);
                                    };
                                    pred = lorenzo(f, pred_dim, sel == SEL_LORENZO_2LAYER);
                                }
                                float v = data[x * s0 + y * s1 + z];
                                codes.push_back(quantizer.quantize_and_overwrite(v, pred));
                                recon[x * s0 + y * s1 + z] = v;
                            }
                }
    }

    // decompress_3d (FF:251-318) given codes, indicator, decoded coefficients, unpred lists
    void decompress(const int32_t *cds, float *out) {
        recon.assign(d1 * d2 * d3, 0.0f);
        size_t b = 0, ci = 0, ri = 0;
        for (size_t bi = 0; bi < nx; bi++)
            for (size_t bj = 0; bj < ny; bj++)
                for (size_t bk = 0; bk < nz; bk++, b++) {
                    size_t x0 = bi * bs, y0 = bj * bs, z0 = bk * bs;
                    int sx = (bi + 1) * bs < d1 ? bs : (int) (d1 - x0);
                    int sy = (bj + 1) * bs < d2 ? bs : (int) (d2 - y0);
                    int sz = (bk + 1) * bs < d3 ? bs : (int) (d3 - z0);
                    int sel = indicator[b];
                    const float *c = sel == SEL_REGRESSION ? &reg_coeffs[4 * ri++] : nullptr;
                    for (int i = 0; i < sx; i++)
                        for (int j = 0; j < sy; j++)
                            for (int k = 0; k < sz; k++) {
                                size_t x = x0 + i, y = y0 + j, z = z0 + k;
                                int code = cds[ci++];
                                float v;
                                if (code == 0) {
                                    v = quantizer.recover(0.0f, 0);
                                } else {
                                    float pred;
                                    if (c) {
                                        pred = reg_predict(c, i, j, k);
                                    } else {
                                        auto f = [&](int dx, int dy, int dz) {
                                            return rv((long) x + dx, (long) y + dy, (long) z + dz);
                                        };
                                        pred = lorenzo(f, pred_dim, sel == SEL_LORENZO_2LAYER);
                                    }
                                    v = quantizer.recover(pred, code);
                                }
                                recon[x * s0 + y * s1 + z] = v;
                            }
                }
        memcpy(out, recon.data(), recon.size() * sizeof(float));
    }
};

// SZFastFrontend::save (FF:50-84) + SZGeneralCompressor::compress (SGC:30-71)
static bool lr_raw(const oracle_lr_params *p, const float *data, std::vector<uint8_t> &out) {
    LREngine e;
    e.set(p);
    e.compress(data);
    out.clear();
    // meta_params as a raw struct (MD:33-69): padding bytes written as 0
    put(out, (int) e.bs);
    put(out, (int) e.pred_dim);
    put(out, (bool) e.lo);
    put(out, (bool) e.lo2);
    put(out, (bool) e.reg);
    put(out, (uint8_t) 0);
    put(out, (int) 2);       // lorenzo_padding_layer
    put(out, (int) 0);       // capacity
    put(out, e.reg_indep);
    put(out, e.reg_linear);
    put(out, 1.0f);          // reg_eb_base
    put(out, (float) e.bs);  // reg_eb_1
    put(out, 1.0f);          // sample_ratio
    put(out, (bool) true);   // lossless
    put(out, (uint8_t) 0), put(out, (uint8_t) 0), put(out, (uint8_t) 0);
    put(out, e.eb);          // precision
    put(out, (bool) false);  // meanInfo.use_mean
    put(out, (uint8_t) 0), put(out, (uint8_t) 0), put(out, (uint8_t) 0);
    put(out, 0.0f);          // meanInfo.mean
    size_t reg_count = e.reg_coeffs.size() / 4;
    put(out, reg_count);
    Huffman hi;
    if (!hi.init(e.indicator.data(), e.indicator.size())) return false;
    hi.save(out);
    hi.encode(e.indicator.data(), e.indicator.size(), out);
    if (reg_count) {  // encode_regression_coefficients (MRP:96-108)
        put(out, (size_t) e.reg_unpred.size());
        const uint8_t *u = (const uint8_t *) e.reg_unpred.data();
        out.insert(out.end(), u, u + e.reg_unpred.size() * sizeof(float));
        Huffman hr;
        if (!hr.init(e.reg_types.data(), e.reg_types.size())) return false;
        hr.save(out);
        hr.encode(e.reg_types.data(), e.reg_types.size(), out);
    }
    out.push_back(2);  // LinearQuantizer::save (Q:185-201)
    put(out, e.eb);
    put(out, (int) e.quantizer.radius);
    put(out, (size_t) e.quantizer.unpred.size());
    const uint8_t *up = (const uint8_t *) e.quantizer.unpred.data();
    out.insert(out.end(), up, up + e.quantizer.unpred.size() * sizeof(float));
    Huffman h;
    if (!h.init(e.codes.data(), e.codes.size())) return false;
    h.save(out);
    h.encode(e.codes.data(), e.codes.size(), out);
    return true;
}

// SZFastFrontend::load + decompress_3d (FF:86-131, 251-318), MRP:55-94
static bool lr_decompress_raw(const uint8_t *raw, size_t n, const size_t *dims, float *out) {
    const uint8_t *p = raw, *end = raw + n;
    if (n < 44 + 8 + 8 + 8) return false;
    oracle_lr_params prm;
    memset(&prm, 0, sizeof(prm));
    for (int i = 0; i < 3; i++) prm.dims[i] = dims[i];
    prm.block_size = take<int>(p);
    prm.pred_dim = take<int>(p);
    prm.lorenzo = take<bool>(p);
    prm.lorenzo2 = take<bool>(p);
    prm.regression = take<bool>(p);
    p += 1 + 4 + 4 + 4 * 5 + 1 + 3;
    prm.eb = take<double>(p);
    p += 8;  // meanInfo
    size_t reg_count = take<size_t>(p);
    if (prm.block_size <= 0) return false;
    LREngine e;
    e.set(&prm);
    e.indicator.resize(e.nblocks);
    HuffDecoder hi;
    size_t used = hi.decode(p, end - p, e.indicator.data(), e.nblocks);
    if (!used) return false;
    p += used;
    if (reg_count) {
        size_t U = take<size_t>(p);
        const float *ru = (const float *) p;
        p += U * sizeof(float);
        std::vector<int32_t> types(4 * reg_count);
        HuffDecoder hr;
        used = hr.decode(p, end - p, types.data(), types.size());
        if (!used) return false;
        p += used;
        const float prc[4] = {e.reg_linear, e.reg_linear, e.reg_linear, e.reg_indep};
        float prev[4] = {0, 0, 0, 0};
        size_t ui = 0;
        for (size_t r = 0; r < reg_count; r++)
            for (int t = 0; t < 4; t++) {  // recover_reg_coeff (MRP:29-37)
                int tv = types[4 * r + t];
                float v = tv == 0 ? ru[ui++] : (float) (prev[t] + 2 * (tv - 32768) * (double) prc[t]);
                e.reg_coeffs.push_back(v);
                prev[t] = v;
            }
        (void) U;
    }
    p += 1;  // quantizer uid
    e.quantizer.set_eb(take<double>(p));
    e.quantizer.radius = take<int>(p);
    size_t U = take<size_t>(p);
    e.quantizer.unpred.assign((const float *) p, (const float *) p + U);
    p += U * sizeof(float);
    size_t num = dims[0] * dims[1] * dims[2];
    std::vector<int32_t> codes(num);
    HuffDecoder hd;
    if (!hd.decode(p, end - p, codes.data(), num)) return false;
    e.decompress(codes.data(), out);
    return true;
}

// ---------------------------------------------------------------------------
// General frontend: SZGeneralFrontend (include/QoZ/frontend/
// SZGeneralFrontend.hpp:36-71, 110-125) with ComposedPredictor{Lorenzo 1-layer,
// Lorenzo 2-layer} (ComposedPredictor.hpp:45-63, 70-93, 203-249) or one of
// them alone (make_lorenzo_regression_compressor, SZLorenzoReg.hpp:27-91),
// N = 1..4.  Blocks of conf.blockSize in raster order; the data array is
// overwritten with the reconstruction point by point, so predictions read
// reconstructed values behind the current point and originals elsewhere.
// ---------------------------------------------------------------------------
struct GFEngine {
    int N = 3;
    size_t dims[4] = {1, 1, 1, 1}, st[4] = {0, 0, 0, 0};
    int bs = 5;
    int npred = 2;
    // predictor list in the order of LR:36-91: 1 = 1-layer Lorenzo, 2 =
    // 2-layer Lorenzo, 3 = RegressionPredictor
    int layer[3] = {1, 2, 3};
    float noise[3] = {0, 0, 0};
    double eb = 0;
    Quantizer quantizer;
    std::vector<int32_t> codes, selection;
    // RegressionPredictor (RP:17-246): coefficient quantizers (independent
    // term eb / (N+1), linear terms eb / (N+1) / blockSize), the block's
    // coefficients and the last committed ones, the coefficient codes
    bool reg = false;
    Quantizer q_ind, q_lin;
    float cur[5] = {0, 0, 0, 0, 0}, prev_c[5] = {0, 0, 0, 0, 0};
    std::vector<int32_t> reg_codes;
    size_t reg_index = 0;

    bool set(const oracle_gf_params *p) {
        N = p->N;
        if (N < 1 || N > 4 || p->block_size < 1) return false;
        if (!p->lorenzo && !p->lorenzo2 && !p->regression) return false;
        if (p->lorenzo2 && p->block_size < 2) return false;  // a 2-layer read would leave the array
        npred = 0;
        if (p->lorenzo) layer[npred++] = 1;
        if (p->lorenzo2) layer[npred++] = 2;
        if (p->regression) layer[npred++] = 3;
        reg = p->regression != 0;
        bs = p->block_size;
        size_t s = 1;
        for (int d = N - 1; d >= 0; d--) {
            dims[d] = p->dims[d];
            if (!dims[d]) return false;
            st[d] = s;
            s *= dims[d];
        }
        eb = p->eb;
        // LorenzoPredictor(double eb) noise terms (LP:25-47), stored as T
        static const double n1[4] = {0.5, 0.81, 1.22, 1.79};
        static const double n2[4] = {1.08, 2.76, 6.8, 0.0};
        noise[1] = (float) (n1[N - 1] * eb);
        noise[2] = (float) (n2[N - 1] * eb);
        quantizer.set_eb(eb);
        quantizer.radius = p->radius;
        q_ind.set_eb(eb / (N + 1));        // RP:22-25
        q_lin.set_eb(eb / (N + 1) / bs);
        return true;
    }

    // RegressionPredictor::precompress_block (RP:48-85): least-squares plane
    // of the block's values, row sums in double along the last axis; false
    // (coefficients left as they were) for a block of extent 1 on any axis
    bool reg_fit(const float *data, const size_t *o, const size_t *sz) {
        size_t n = 1;
        for (int d = 0; d < N; d++) {
            if (sz[d] <= 1) return false;
            n *= sz[d];
        }
        const float recip = (float) (1.0 / (double) n);
        double sum[5] = {0, 0, 0, 0, 0};
        const size_t rows = n / sz[N - 1];
        size_t l[4] = {0, 0, 0, 0};
        for (size_t r = 0; r < rows; r++) {
            size_t idx = 0;
            for (int d = 0; d < N - 1; d++) idx += (o[d] + l[d]) * st[d];
            double cum = 0;
            for (size_t t = 0; t < sz[N - 1]; t++) {
                const float v = data[idx + (o[N - 1] + t) * st[N - 1]];
                cum += v;
                sum[N - 1] += (double) t * v;
            }
            for (int i = 0; i < N - 1; i++) sum[i] += cum * (double) l[i];
            sum[N] += cum;
            for (int d = N - 2; d >= 0; d--) {
                if (++l[d] < sz[d]) break;
                l[d] = 0;
            }
        }
        for (int i = 0; i <= N; i++) cur[i] = 0;
        cur[N] = (float) (sum[N] * recip);
        for (int i = 0; i < N; i++) {
            cur[i] = (float) ((2 * sum[i] / (double) (sz[i] - 1) - sum[N]) * 6 * recip / (double) (sz[i] + 1));
            cur[N] -= (float) (sz[i] - 1) * cur[i] / 2;
        }
        return true;
    }

    // RegressionPredictor::predict (RP:91-98) at block-local index l, in T
    float reg_predict(const size_t *l) const {
        float pred = 0;
        for (int i = 0; i < N; i++) pred += (float) l[i] * cur[i];
        pred += cur[N];
        return pred;
    }

    // precompress_block_commit: coefficients predicted from the last
    // committed ones and overwritten with their reconstruction (RP:87-90, 224-230)
    void reg_commit() {
        for (int i = 0; i < N; i++) reg_codes.push_back(q_lin.quantize_and_overwrite(cur[i], prev_c[i]));
        reg_codes.push_back(q_ind.quantize_and_overwrite(cur[N], prev_c[N]));
        for (int i = 0; i <= N; i++) prev_c[i] = cur[i];
    }

    // pred_and_recover_coefficients (RP:233-240)
    bool reg_recover() {
        if (reg_index + N + 1 > reg_codes.size()) return false;
        for (int i = 0; i < N; i++) cur[i] = q_lin.recover(cur[i], reg_codes[reg_index++]);
        cur[N] = q_ind.recover(cur[N], reg_codes[reg_index++]);
        return true;
    }

    size_t num() const {
        size_t n = 1;
        for (int d = 0; d < N; d++) n *= dims[d];
        return n;
    }

    // iterator.prev (Iterator.hpp:283-294): 0 when an offset crosses the low
    // face of the field (bs >= 2 for 2-layer reads keeps every other read in it)
    static float prev(const float *v, const size_t *c, const size_t *st, int N, const int *a) {
        size_t off = 0;
        for (int d = 0; d < N; d++) {
            if (c[d] < (size_t) a[d]) return 0.0f;
            off += (size_t) a[d] * st[d];
        }
        return *(v - off);
    }

    // LorenzoPredictor::do_predict (LP:117-247), float, left to right;
    // N = 4 uses the 1-layer formula for both layer counts (LP:170-177)
    float predict(const float *v, const size_t *c, int L) const {
        auto f1 = [&](int a) { int o[1] = {a}; return prev(v, c, st, 1, o); };
        auto f2 = [&](int a, int b) { int o[2] = {a, b}; return prev(v, c, st, 2, o); };
        auto f3 = [&](int a, int b, int e) { int o[3] = {a, b, e}; return prev(v, c, st, 3, o); };
        auto f4 = [&](int a, int b, int e, int g) { int o[4] = {a, b, e, g}; return prev(v, c, st, 4, o); };
        switch (N) {
            case 1:
                if (L == 1) return f1(1);
                return 2 * f1(1) - f1(2);
            case 2:
                if (L == 1) return f2(0, 1) + f2(1, 0) - f2(1, 1);
                return 2 * f2(0, 1) - f2(0, 2) + 2 * f2(1, 0) - 4 * f2(1, 1) + 2 * f2(1, 2) - f2(2, 0) +
                       2 * f2(2, 1) - f2(2, 2);
            case 3:
                if (L == 1)
                    return f3(0, 0, 1) + f3(0, 1, 0) + f3(1, 0, 0) - f3(0, 1, 1) - f3(1, 0, 1) - f3(1, 1, 0) +
                           f3(1, 1, 1);
                return 2 * f3(0, 0, 1) - f3(0, 0, 2) + 2 * f3(0, 1, 0) - 4 * f3(0, 1, 1) + 2 * f3(0, 1, 2) -
                       f3(0, 2, 0) + 2 * f3(0, 2, 1) - f3(0, 2, 2) + 2 * f3(1, 0, 0) - 4 * f3(1, 0, 1) +
                       2 * f3(1, 0, 2) - 4 * f3(1, 1, 0) + 8 * f3(1, 1, 1) - 4 * f3(1, 1, 2) + 2 * f3(1, 2, 0) -
                       4 * f3(1, 2, 1) + 2 * f3(1, 2, 2) - f3(2, 0, 0) + 2 * f3(2, 0, 1) - f3(2, 0, 2) +
                       2 * f3(2, 1, 0) - 4 * f3(2, 1, 1) + 2 * f3(2, 1, 2) - f3(2, 2, 0) + 2 * f3(2, 2, 1) -
                       f3(2, 2, 2);
            default:
                return f4(0, 0, 0, 1) + f4(0, 0, 1, 0) - f4(0, 0, 1, 1) + f4(0, 1, 0, 0) - f4(0, 1, 0, 1) -
                       f4(0, 1, 1, 0) + f4(0, 1, 1, 1) + f4(1, 0, 0, 0) - f4(1, 0, 0, 1) - f4(1, 0, 1, 0) +
                       f4(1, 0, 1, 1) - f4(1, 1, 0, 0) + f4(1, 1, 0, 1) + f4(1, 1, 1, 0) - f4(1, 1, 1, 1);
        }
    }

    // LorenzoPredictor::estimate_error (LP:102-104): |x - pred| + noise in T;
    // RegressionPredictor::estimate_error (RP:44-46): |x - pred| with the
    // block's unquantized coefficients
    float estimate(const float *data, const size_t *c, int L, const size_t *o) const {
        size_t idx = 0;
        for (int d = 0; d < N; d++) idx += c[d] * st[d];
        float x = data[idx];
        if (L == 3) {
            size_t l[4];
            for (int d = 0; d < N; d++) l[d] = c[d] - o[d];
            return std::fabs(x - reg_predict(l));
        }
        return std::fabs(x - predict(data + idx, c, L)) + noise[L];
    }

    // ComposedPredictor::do_estimate_error (CP:203-249): predictor with the
    // smallest summed estimate over the block diagonals, first on ties.
    int select(const float *data, const size_t *o, const size_t *sz) const {
        double pe[3] = {0, 0, 0};
        size_t m = sz[0];
        for (int d = 1; d < N; d++) m = std::min(m, sz[d]);
        size_t c[4];
        if (N == 1) {
            for (int p = 0; p < npred; p++) {
                c[0] = o[0];
                pe[p] += estimate(data, c, layer[p], o);
                c[0] = o[0] + m - 1;
                pe[p] += estimate(data, c, layer[p], o);
            }
        } else {
            const int ndiag = 1 << (N - 1);
            for (size_t k = 0; k + 2 < m; k++)
                for (int p = 0; p < npred; p++)
                    for (int t = 0; t < ndiag; t++) {
                        c[0] = o[0] + k;
                        for (int d = 1; d < N; d++)
                            c[d] = o[d] + (((t >> (N - 1 - d)) & 1) ? m - 1 - k : k);
                        pe[p] += estimate(data, c, layer[p], o);
                    }
        }
        int sid = 0;
        for (int p = 1; p < npred; p++)
            if (pe[p] < pe[sid]) sid = p;
        return sid;
    }

    // SZGeneralFrontend::compress (GF:36-71) on one array, appending codes,
    // selections and outliers to the engine's lists (the tuner reuses one
    // compressor across its sample blocks).
    void compress(float *data) {
        size_t nb[4], b[4] = {0, 0, 0, 0};
        size_t total = 1;
        for (int d = 0; d < N; d++) {
            nb[d] = (dims[d] - 1) / bs + 1;
            total *= nb[d];
        }
        for (size_t t = 0; t < total; t++) {
            size_t o[4] = {0, 0, 0, 0}, sz[4] = {1, 1, 1, 1};
            for (int d = 0; d < N; d++) {
                o[d] = b[d] * bs;
                sz[d] = b[d] == nb[d] - 1 ? dims[d] - o[d] : (size_t) bs;
            }
            // every predictor's precompress_block, then the choice (CP:45-63);
            // a block the chosen predictor declines runs on the 1-layer
            // Lorenzo fallback with no selection recorded (GF:47-52)
            const bool reg_ok = reg ? reg_fit(data, o, sz) : true;
            int L = layer[0];
            int sid = 0;
            if (npred > 1) sid = select(data, o, sz);
            L = layer[sid];
            const bool ok = L != 3 || reg_ok;
            if (ok) {
                if (npred > 1) selection.push_back(sid);
                if (L == 3) reg_commit();
            } else {
                L = 1;
            }
            size_t bn = 1;
            for (int d = 0; d < N; d++) bn *= sz[d];
            size_t l[4] = {0, 0, 0, 0};
            for (size_t e = 0; e < bn; e++) {
                size_t c[4], idx = 0;
                for (int d = 0; d < N; d++) {
                    c[d] = o[d] + l[d];
                    idx += c[d] * st[d];
                }
                float pred = L == 3 ? reg_predict(l) : predict(data + idx, c, L);
                codes.push_back(quantizer.quantize_and_overwrite(data[idx], pred));
                for (int d = N - 1; d >= 0; d--) {
                    if (++l[d] < sz[d]) break;
                    l[d] = 0;
                }
            }
            for (int d = N - 1; d >= 0; d--) {
                if (++b[d] < nb[d]) break;
                b[d] = 0;
            }
        }
    }

    // SZGeneralFrontend::decompress (GF:72-96) given the codes and selections
    // (one selection per block, as ComposedPredictor::predecompress_block
    // reads them, CP:64-68); false when the lists run out
    bool decompress(const int32_t *cds, float *out) {
        size_t nb[4], b[4] = {0, 0, 0, 0};
        size_t total = 1;
        for (int d = 0; d < N; d++) {
            nb[d] = (dims[d] - 1) / bs + 1;
            total *= nb[d];
        }
        size_t ci = 0;
        quantizer.index = 0;
        for (size_t t = 0; t < total; t++) {
            size_t o[4] = {0, 0, 0, 0}, sz[4] = {1, 1, 1, 1};
            for (int d = 0; d < N; d++) {
                o[d] = b[d] * bs;
                sz[d] = b[d] == nb[d] - 1 ? dims[d] - o[d] : (size_t) bs;
            }
            if (npred > 1 && t >= selection.size()) return false;
            int L = npred > 1 ? layer[selection[t]] : layer[0];
            if (L == 3) {
                bool ok = true;
                for (int d = 0; d < N; d++) ok = ok && sz[d] > 1;
                if (!ok) L = 1;
                else if (!reg_recover()) return false;
            }
            size_t bn = 1;
            for (int d = 0; d < N; d++) bn *= sz[d];
            size_t l[4] = {0, 0, 0, 0};
            for (size_t e = 0; e < bn; e++) {
                size_t c[4], idx = 0;
                for (int d = 0; d < N; d++) {
                    c[d] = o[d] + l[d];
                    idx += c[d] * st[d];
                }
                const float pred = L == 3 ? reg_predict(l) : predict(out + idx, c, L);
                out[idx] = quantizer.recover(pred, cds[ci++]);
                for (int d = N - 1; d >= 0; d--) {
                    if (++l[d] < sz[d]) break;
                    l[d] = 0;
                }
            }
            for (int d = N - 1; d >= 0; d--) {
                if (++b[d] < nb[d]) break;
                b[d] = 0;
            }
        }
        return true;
    }

    // LinearQuantizer::save (Q:185-201)
    static void save_quantizer(const Quantizer &q, std::vector<uint8_t> &out) {
        out.push_back(2);
        put(out, q.eb);
        put(out, (int) q.radius);
        put(out, (size_t) q.unpred.size());
        const uint8_t *u = (const uint8_t *) q.unpred.data();
        out.insert(out.end(), u, u + q.unpred.size() * sizeof(float));
    }

    // SZGeneralFrontend::save (GF:98-103) + quantizer, then the codes'
    // Huffman tree and bits (SGC:30-71 / encoding_lossless SGC:72-91)
    bool save(std::vector<uint8_t> &out) const {
        for (int d = 0; d < N; d++) put(out, (uint64_t) dims[d]);
        put(out, (uint32_t) bs);
        for (int p = 0; p < npred; p++) {
            if (layer[p] != 3) {
                out.push_back(1);  // LorenzoPredictor::save (LP:57-59)
                continue;
            }
            // RegressionPredictor::save (RP:100-117)
            out.push_back(2);
            put(out, (size_t) reg_codes.size());
            if (!reg_codes.empty()) {
                save_quantizer(q_ind, out);
                save_quantizer(q_lin, out);
                Huffman hr;
                if (!hr.init(reg_codes.data(), reg_codes.size())) return false;
                hr.save(out);
                hr.encode(reg_codes.data(), reg_codes.size(), out);
            }
        }
        if (npred > 1) {                                    // ComposedPredictor::save (CP:70-81)
            put(out, (uint64_t) selection.size());
            if (!selection.empty()) {
                Huffman hs;
                if (!hs.init(selection.data(), selection.size())) return false;
                hs.save(out);
                hs.encode(selection.data(), selection.size(), out);
            }
        }
        save_quantizer(quantizer, out);
        Huffman h;
        if (!h.init(codes.data(), codes.size())) return false;
        h.save(out);
        h.encode(codes.data(), codes.size(), out);
        return true;
    }
};

// LinearQuantizer::load (Q:203-226)
static bool load_quantizer(const uint8_t *&p, const uint8_t *end, Quantizer &q) {
    if (end - p < 1 + 8 + 4 + 8) return false;
    p += 1;  // uid
    q.set_eb(take<double>(p));
    q.radius = take<int>(p);
    const size_t U = take<size_t>(p);
    if (U > (size_t) (end - p) / sizeof(float)) return false;
    q.unpred.assign((const float *) p, (const float *) p + U);
    q.index = 0;
    p += U * sizeof(float);
    return true;
}

// SGC:115-140 + GF:105-116 + CP:82-93 on a pre-zstd buffer; N and the
// predictor list come from the Config (lorenzo, lorenzo2, regression).
static bool gf_decompress_raw(const uint8_t *raw, size_t n, int N, int lo, int lo2, int regr, float *out,
                              size_t num) {
    const uint8_t *p = raw, *end = raw + n;
    oracle_gf_params prm;
    memset(&prm, 0, sizeof(prm));
    prm.N = N;
    prm.lorenzo = lo, prm.lorenzo2 = lo2, prm.regression = regr;
    if (n < (size_t) N * 8 + 4) return false;
    for (int d = 0; d < N; d++) prm.dims[d] = (size_t) take<uint64_t>(p);
    prm.block_size = (int) take<uint32_t>(p);
    prm.eb = 1.0;
    prm.radius = 32768;
    GFEngine e;
    if (!e.set(&prm) || e.num() != num) return false;
    for (int k = 0; k < e.npred; k++) {  // predictor ids; the regression predictor's coefficients (RP:134-156)
        p += 1;
        if (e.layer[k] != 3) continue;
        const size_t nc = take<size_t>(p);
        if (!nc) continue;
        if (!load_quantizer(p, end, e.q_ind) || !load_quantizer(p, end, e.q_lin)) return false;
        e.reg_codes.resize(nc);
        HuffDecoder hr;
        const size_t used = hr.decode(p, end - p, e.reg_codes.data(), nc);
        if (!used) return false;
        p += used;
    }
    size_t nblk = 1;
    for (int d = 0; d < N; d++) nblk *= (e.dims[d] - 1) / e.bs + 1;
    if (e.npred > 1) {
        size_t ns = take<uint64_t>(p);
        if (ns != nblk) return false;
        e.selection.resize(ns);
        HuffDecoder hs;
        size_t used = hs.decode(p, end - p, e.selection.data(), ns);
        if (!used) return false;
        p += used;
    }
    p += 1;  // quantizer uid
    e.quantizer.set_eb(take<double>(p));
    e.quantizer.radius = take<int>(p);
    size_t U = take<size_t>(p);
    e.quantizer.unpred.assign((const float *) p, (const float *) p + U);
    p += U * sizeof(float);
    std::vector<int32_t> codes(num);
    HuffDecoder hd;
    if (!hd.decode(p, end - p, codes.data(), num)) return false;
    return e.decompress(codes.data(), out);
}

// ---------------------------------------------------------------------------
// The default (SZ3) tuning mode: Tuning with predictorTuningRate = 0
// (SZInterp.hpp:2096-2180) and SZ_compress_Interp_lorenzo's Lorenzo branch
// (SZI:2783-2898): uniform sampling (Extraction.hpp:8-212), the deprecated
// block interpolation compressor as the interpolation test
// (compressor/deprecated/SZBlockInterpolationCompressor.hpp:110-458, driven by
// do_not_use_this_interp_compress_block_test, SZI:558-576), the Lorenzo
// compressor on the sample, optimize_quant_invl_3d (QuantOptimization.hpp:
// 14-147).
// ---------------------------------------------------------------------------

// cal_sampling_ratio (Extraction.hpp:8-15): float result
static float sz3_sampling_ratio(size_t block, size_t n, size_t dmin, const size_t *dims, int N) {
    size_t sample_n = 1;
    for (int d = 0; d < N; d++) sample_n *= dims[d] / dmin * 2 * block;
    return (float) (sample_n * 1.0 / n);
}

// sampling<T,N> (Extraction.hpp:17-212): 2*sb points per dmin along each
// axis, the sb points after the first sb and the sb points before the last
// sb of every dmin-long segment.  Returns false (sample = whole field) when
// sb < 9.
static bool sz3_sampling(const float *data, int N, const size_t *dims, std::vector<float> &out, size_t *sdims,
                         size_t &sb) {
    size_t num = 1, dmin = dims[0];
    for (int d = 0; d < N; d++) num *= dims[d], dmin = std::min(dmin, dims[d]);
    sb = dmin;
    while (sz3_sampling_ratio(sb, num, dmin, dims, N) > 0.035) sb--;
    if (sb * 2 > dmin) sb = dmin / 2;
    if (sb < 9) {
        for (int d = 0; d < N; d++) sdims[d] = dims[d];
        out.clear();
        return false;
    }
    size_t sn = 1, sst[4], st[4];
    for (int d = 0; d < N; d++) sdims[d] = dims[d] / dmin * 2 * sb, sn *= sdims[d];
    sst[N - 1] = st[N - 1] = 1;
    for (int d = N - 2; d >= 0; d--) sst[d] = sst[d + 1] * sdims[d + 1], st[d] = st[d + 1] * dims[d + 1];
    out.assign(sn, 0.0f);
    for (size_t q = 0; q < sn; q++) {
        size_t src = 0;
        for (int d = 0; d < N; d++) {
            const size_t c = q / sst[d] % sdims[d], bd = c / (2 * sb), i = c % (2 * sb);
            const size_t di = i < sb ? i + sb : dmin - 3 * sb + i;
            src += (bd * dmin + di) * st[d];
        }
        out[q] = data[src];
    }
    return true;
}

// SZBlockInterpolationCompressor (deprecated, SBI:110-458) for N = 1..3:
// independent blocks of blockSize, each with its first point (prediction
// 0) and its own levels; the block's lines in the loop order of its
// direction.
struct BIEngine {
    int N = 3;
    size_t dims[4] = {1, 1, 1, 1}, st[4] = {0, 0, 0, 0};
    size_t bs = 6;
    int algo = 1, dir = 0;
    double eb = 0;
    Quantizer quantizer;
    std::vector<int32_t> codes;

    void quantize(float *d, float pred) { codes.push_back(quantizer.quantize_and_overwrite(*d, pred)); }

    // block_interpolation_1d (SBI:205-262), predict-and-overwrite
    void line(float *data, size_t begin, size_t end, size_t stride) {
        const size_t n = (end - begin) / stride + 1;
        if (n <= 1) return;
        const size_t s3 = 3 * stride, s5 = 5 * stride;
        if (algo == 0 || n < 5) {
            for (size_t i = 1; i + 1 < n; i += 2) {
                float *d = data + begin + i * stride;
                quantize(d, interp_linear(*(d - stride), *(d + stride)));
            }
            if (n % 2 == 0) {
                float *d = data + begin + (n - 1) * stride;
                if (n < 4) quantize(d, *(d - stride));
                else quantize(d, interp_linear1(*(d - s3), *(d - stride)));
            }
        } else {
            float *d = data + begin + stride;
            quantize(d, interp_quad_1(*(d - stride), *(d + stride), *(d + s3)));
            for (size_t i = 3; i + 3 < n; i += 2) {
                d = data + begin + i * stride;
                quantize(d, interp_cubic(*(d - s3), *(d - stride), *(d + stride), *(d + s3)));
            }
            if (n % 2 == 0) {
                d = data + begin + (n - 3) * stride;
                quantize(d, interp_quad_2(*(d - s3), *(d - stride), *(d + stride)));
                d += 2 * stride;
                quantize(d, interp_quad_3(*(d - s5), *(d - s3), *(d - stride)));
            } else {
                d = data + begin + (n - 2) * stride;
                quantize(d, interp_quad_2(*(d - s3), *(d - stride), *(d + stride)));
            }
        }
    }

    size_t off(size_t a, size_t b, size_t c) const { return a * st[0] + b * st[1] + c * st[2]; }

    // block_interpolation (SBI:263-458): the three passes of one level in the
    // direction's loop order (axis, outer axis and step, inner axis and step)
    void level(float *data, const size_t *B, const size_t *E, size_t s) {
        const size_t s2 = 2 * s;
        if (N == 1) {
            line(data, B[0], E[0], s);
            return;
        }
        if (N == 2) {
            if (dir == 0) {
                for (size_t j = B[1]; j <= E[1]; j += s2) line(data, B[0] * st[0] + j, E[0] * st[0] + j, s * st[0]);
                for (size_t i = B[0]; i <= E[0]; i += s) line(data, i * st[0] + B[1], i * st[0] + E[1], s);
            } else {
                for (size_t i = B[0]; i <= E[0]; i += s2) line(data, i * st[0] + B[1], i * st[0] + E[1], s);
                for (size_t j = B[1]; j <= E[1]; j += s) line(data, B[0] * st[0] + j, E[0] * st[0] + j, s * st[0]);
            }
            return;
        }
        // N = 3: {axis, outer axis, outer step factor, inner axis, inner step factor} per pass
        static const int P[6][3][5] = {
            {{0, 1, 2, 2, 2}, {1, 0, 1, 2, 2}, {2, 0, 1, 1, 1}},
            {{0, 1, 2, 2, 2}, {2, 0, 1, 1, 2}, {1, 0, 1, 2, 1}},
            {{1, 0, 2, 2, 2}, {0, 1, 1, 2, 2}, {2, 0, 1, 1, 1}},
            {{1, 0, 2, 2, 2}, {2, 0, 2, 1, 1}, {0, 1, 1, 2, 1}},
            {{2, 0, 2, 1, 2}, {0, 1, 2, 2, 1}, {1, 0, 1, 2, 1}},
            {{2, 0, 2, 1, 2}, {1, 0, 2, 2, 1}, {0, 1, 1, 2, 1}},
        };
        for (int p = 0; p < 3; p++) {
            const int a = P[dir][p][0], u = P[dir][p][1], w = P[dir][p][3];
            const size_t su = s * P[dir][p][2], sw = s * P[dir][p][4];
            for (size_t x = B[u]; x <= E[u]; x += su)
                for (size_t y = B[w]; y <= E[w]; y += sw) {
                    size_t c0[3], c1[3];
                    c0[u] = c1[u] = x;
                    c0[w] = c1[w] = y;
                    c0[a] = B[a];
                    c1[a] = E[a];
                    line(data, off(c0[0], c0[1], c0[2]), off(c1[0], c1[1], c1[2]), s * st[a]);
                }
        }
    }

    // compress (SBI:110-146)
    void compress(float *data) {
        size_t nb[4], b[4] = {0, 0, 0, 0}, total = 1;
        for (int d = 0; d < N; d++) nb[d] = (dims[d] - 1) / bs + 1, total *= nb[d];
        codes.clear();
        quantizer.unpred.clear();
        for (size_t t = 0; t < total; t++) {
            size_t B[3] = {0, 0, 0}, E[3] = {0, 0, 0};
            unsigned max_level = 1;
            for (int d = 0; d < N; d++) {
                B[d] = b[d] * bs;
                const size_t bd = B[d] + bs > dims[d] ? dims[d] - B[d] : bs;
                E[d] = B[d] + bd - 1;
                if (max_level < std::ceil(std::log2((double) bd))) max_level = (unsigned) std::ceil(std::log2((double) bd));
            }
            size_t o = 0;
            for (int d = 0; d < N; d++) o += B[d] * st[d];
            quantize(data + o, 0.0f);
            for (unsigned lv = max_level; lv > 0 && lv <= max_level; lv--) level(data, B, E, (size_t) 1 << (lv - 1));
            for (int d = N - 1; d >= 0; d--) {
                if (++b[d] < nb[d]) break;
                b[d] = 0;
            }
        }
    }

    // compress's buffer (SBI:150-165): dims, blockSize, interpolator id,
    // direction id, quantizer, Huffman tree and bits
    bool save(std::vector<uint8_t> &out) const {
        for (int d = 0; d < N; d++) put(out, (uint64_t) dims[d]);
        put(out, (uint32_t) bs);
        put(out, (int32_t) algo);
        put(out, (int32_t) dir);
        out.push_back(2);
        put(out, eb);
        put(out, (int) quantizer.radius);
        put(out, (size_t) quantizer.unpred.size());
        const uint8_t *u = (const uint8_t *) quantizer.unpred.data();
        out.insert(out.end(), u, u + quantizer.unpred.size() * sizeof(float));
        Huffman h;
        if (!h.init(codes.data(), codes.size())) return false;
        h.save(out);
        h.encode(codes.data(), codes.size(), out);
        return true;
    }
};

// do_not_use_this_interp_compress_block_test (SZI:558-576): the zstd-framed
// size of the block interpolation compressor on a copy of the data.
static size_t sz3_block_test(const float *data, int N, const size_t *dims, double eb, int algo, int dir, size_t bs,
                             std::vector<uint8_t> *raw_out = nullptr) {
    BIEngine e;
    e.N = N;
    size_t s = 1, num = 1;
    for (int d = N - 1; d >= 0; d--) e.dims[d] = dims[d], e.st[d] = s, s *= dims[d];
    num = s;
    e.bs = bs;
    e.algo = algo;
    e.dir = dir;
    e.eb = eb;
    e.quantizer.set_eb(eb);
    e.quantizer.radius = 32768;
    std::vector<float> v(data, data + num);
    e.compress(v.data());
    std::vector<uint8_t> raw;
    if (!e.save(raw)) return 0;
    std::vector<uint8_t> z(ZSTD_compressBound(raw.size()));
    size_t zs = ZSTD_compress(z.data(), z.size(), raw.data(), raw.size(), 3);
    if (raw_out) *raw_out = raw;
    return 8 + zs;
}

// optimize_quant_invl_3d (QO:89-147): the quantizer bin count from a
// histogram of 1-layer 3-D Lorenzo errors (MetaLorenzoPredictor.hpp:42-46)
// at about every 100th point of the pointer walk; the mean estimate it also
// makes does not reach the result.
static int sz3_quant_intervals(const float *data, size_t r1, size_t r2, size_t r3, double precision,
                               size_t *samples = nullptr) {
    std::vector<size_t> intervals(32768, 0);
    const size_t dist = 100, r23 = r2 * r3, len = r1 * r23;
    size_t sample_count = 0, offset_count = dist - 2, offset_count_2 = 0, n1 = 1, n2 = 1;
    const float *p = data + r23 + r3 + offset_count;
    while ((size_t) (p - data) < len) {
        const float pred = p[-1] + p[-(ptrdiff_t) r3] + p[-(ptrdiff_t) r23] - p[-(ptrdiff_t) r3 - 1] -
                           p[-(ptrdiff_t) r23 - 1] - p[-(ptrdiff_t) r23 - (ptrdiff_t) r3] +
                           p[-(ptrdiff_t) r23 - (ptrdiff_t) r3 - 1];
        const float err = std::fabs(pred - *p);
        size_t idx = (size_t) ((err / precision + 1) / 2);
        if (idx >= intervals.size()) idx = intervals.size() - 1;
        intervals[idx]++;
        offset_count += dist;
        if (offset_count >= r3) {
            n2++;
            if (n2 == r2) {
                n1++;
                n2 = 1;
                p += r3;
            }
            offset_count_2 = (n1 + n2) % dist;
            p += (r3 + dist - offset_count) + (dist - offset_count_2);
            offset_count = dist - offset_count_2;
            if (offset_count == 0) offset_count++;
        } else {
            p += dist;
        }
        sample_count++;
    }
    if (samples) *samples = sample_count;
    // estimate_quantization_intervals (QO:25-39)
    const size_t target = (size_t) (sample_count * 0.999);
    size_t sum = 0, i = 0;
    for (i = 0; i < intervals.size(); i++) {
        sum += intervals[i];
        if (sum > target) break;
    }
    if (i == intervals.size()) i = intervals.size() - 1;
    unsigned acc = 2 * (unsigned) (i + 1), b = acc - 1;
    b |= b >> 1, b |= b >> 2, b |= b >> 4, b |= b >> 8, b |= b >> 16;
    const unsigned num_intervals = 2 * (b + 1);
    return num_intervals > 32 ? (int) num_intervals : 32;
}

// Pre-zstd buffer of the general frontend on one array.
static bool gf_raw(const oracle_gf_params *p, const float *data, std::vector<uint8_t> &out) {
    GFEngine e;
    if (!e.set(p)) return false;
    std::vector<float> v(data, data + e.num());
    e.compress(v.data());
    out.clear();
    return e.save(out);
}

// SZ_compress_LorenzoReg's zstd-framed size (LR:95-114) on an array: the
// fast frontend for 3-D, the general frontend otherwise.
static size_t sz3_lorenzo_size(const float *data, int N, const size_t *dims, double eb, int bs, int pred_dim,
                               int radius) {
    std::vector<uint8_t> raw;
    if (N == 3) {
        oracle_lr_params p;
        memset(&p, 0, sizeof(p));
        for (int d = 0; d < 3; d++) p.dims[d] = dims[d];
        p.eb = eb;
        p.block_size = bs;
        p.pred_dim = pred_dim;
        p.lorenzo = p.lorenzo2 = 1;
        p.regression = 0;
        p.radius = radius;
        if (!lr_raw(&p, data, raw)) return 0;
    } else {
        oracle_gf_params p;
        memset(&p, 0, sizeof(p));
        p.N = N;
        for (int d = 0; d < N; d++) p.dims[d] = dims[d];
        p.eb = eb;
        p.block_size = bs;
        p.lorenzo = p.lorenzo2 = 1;
        p.radius = radius;
        if (!gf_raw(&p, data, raw)) return 0;
    }
    std::vector<uint8_t> z(ZSTD_compressBound(raw.size()));
    return 8 + ZSTD_compress(z.data(), z.size(), raw.data(), raw.size(), 3);
}

static int default_block_size(int N) { return N == 1 ? 128 : (N == 2 ? 16 : 6); }

// Tuning with predictorTuningRate = 0 (SZI:2096-2180) and, when Lorenzo
// wins, SZ_compress_Interp_lorenzo's Lorenzo branch (SZI:2783-2898).
static int sz3_tune(const float *data, int N, const size_t *dims, double eb, double rel, oracle_sz3_out *o) {
    memset(o, 0, sizeof(*o));
    size_t num = 1;
    for (int d = 0; d < N; d++) num *= dims[d];
    std::vector<float> sample;
    size_t sdims[4] = {1, 1, 1, 1}, sb = 0;
    const bool sampled = sz3_sampling(data, N, dims, sample, sdims, sb);
    size_t snum = 1;
    for (int d = 0; d < N; d++) snum *= sdims[d], o->sample_dims[d] = sdims[d];
    o->sampling_block = sb;
    o->sampled = sampled;
    o->interp_algo = 1;  // Config defaults (CFG:417-419)
    o->direction = 0;
    o->use_interp = 1;
    if (!sampled) return 0;
    size_t bytes = sz3_lorenzo_size(sample.data(), N, sdims, eb, 5, N, 32768);
    if (!bytes) return -2;
    o->lorenzo_bytes = bytes;
    double lorenzo_ratio = snum * 1.0 * sizeof(float) / bytes;
    double best = 0;
    for (int op = 0; op <= 1; op++) {
        size_t b = sz3_block_test(sample.data(), N, sdims, eb, op, o->direction, sb);
        if (!b) return -2;
        o->interp_bytes[op] = b;
        const double r = snum * 1.0 * sizeof(float) / b;
        if (r > best) best = r, o->interp_algo = op;
    }
    int fact = 1;
    for (int d = 2; d <= N; d++) fact *= d;
    {
        size_t b = sz3_block_test(sample.data(), N, sdims, eb, o->interp_algo, fact - 1, sb);
        if (!b) return -2;
        o->interp_bytes[2] = b;
        const double r = snum * 1.0 * sizeof(float) / b;
        if (r > best * 1.02) best = r, o->direction = fact - 1;
    }
    o->lorenzo_ratio = lorenzo_ratio;
    o->interp_ratio = best;
    o->use_interp = !(lorenzo_ratio > best && lorenzo_ratio < 80 && best < 80);
    if (o->use_interp) return 0;
    // the Lorenzo branch (SZI:2783-2850): setDims resets blockSize and
    // pred_dim, so the tests run with the Config defaults for N and only the
    // bin count survives into the final Config
    int quantbin = 65536, pred_dim = N;
    const int bsd = default_block_size(N);
    if (N == 3) {
        quantbin = sz3_quant_intervals(data, dims[0], dims[1], dims[2], eb);
        o->quantbin_estimate = quantbin;
        pred_dim = 2;
        if (quantbin / 2 > 65536) return -3;
        size_t b = sz3_lorenzo_size(sample.data(), N, sdims, eb, bsd, 2, quantbin / 2);
        if (!b) return -2;
        o->lorenzo_bytes_pd2 = b;
        const double r = snum * 1.0 * sizeof(float) / b;
        if (r > lorenzo_ratio * 1.02) lorenzo_ratio = r;
        else pred_dim = 3;
    }
    if (rel < 1.01e-6 && lorenzo_ratio > 5 && quantbin != 16384) {
        size_t b = sz3_lorenzo_size(sample.data(), N, sdims, eb, bsd, pred_dim, 8192);
        if (!b) return -2;
        o->lorenzo_bytes_16384 = b;
        const double r = snum * 1.0 * sizeof(float) / b;
        if (r > lorenzo_ratio * 1.02) lorenzo_ratio = r, quantbin = 16384;
    }
    o->quantbin = quantbin;
    o->final_lorenzo_ratio = lorenzo_ratio;
    return 0;
}

// ---------------------------------------------------------------------------
// QoZ exhaustive sampling tuner (SZInterp.hpp:1589-2557, QoZ path, RD target,
// no wavelets, no Lorenzo test).
// ---------------------------------------------------------------------------
struct TuneState {
    int N;
    std::vector<size_t> bdims;               // sample block shape, (sbs+1)^N
    std::vector<std::vector<float>> blocks;  // sampled blocks
    oracle_interp_params base;               // engine parameters shared by the runs
};

// sampleBlocks with profiling = 0, var_first = 0 (SZI:1056-1086) and
// sample_blocks (Sample.hpp:79-139): blocks of (sbs+1)^N starting on a grid of
// sbs, every (size_t)(1/rate)-th in raster order.
static void sample_blocks(const float *data, int N, const size_t *dims, size_t sbs, double rate, TuneState &ts,
                          oracle_tune_out *out) {
    size_t stride = (size_t) (1.0 / rate);
    if (stride <= 0) stride = 1;
    size_t idx = 0;
    ts.blocks.clear();
    out->n_blocks = 0;
    auto take = [&](size_t x0, size_t y0, size_t z0) {
        std::vector<float> b;
        if (N == 2) {
            for (size_t i = 0; i <= sbs; i++)
                for (size_t j = 0; j <= sbs; j++) b.push_back(data[(x0 + i) * dims[1] + y0 + j]);
        } else {
            for (size_t i = 0; i <= sbs; i++)
                for (size_t j = 0; j <= sbs; j++)
                    for (size_t k = 0; k <= sbs; k++)
                        b.push_back(data[((x0 + i) * dims[1] + y0 + j) * dims[2] + z0 + k]);
        }
        if (out->n_blocks < 4096) {
            out->block_starts[out->n_blocks][0] = x0;
            out->block_starts[out->n_blocks][1] = y0;
            out->block_starts[out->n_blocks][2] = z0;
        }
        out->n_blocks++;
        ts.blocks.push_back(std::move(b));
    };
    if (N == 2) {
        for (size_t x = 0; x < dims[0] - sbs; x += sbs)
            for (size_t y = 0; y < dims[1] - sbs; y += sbs) {
                if (idx % stride == 0) take(x, y, 0);
                idx++;
            }
    } else if (N == 3) {
        for (size_t x = 0; x < dims[0] - sbs; x += sbs)
            for (size_t y = 0; y < dims[1] - sbs; y += sbs)
                for (size_t z = 0; z < dims[2] - sbs; z += sbs) {
                    if (idx % stride == 0) take(x, y, z);
                    idx++;
                }
    }
}

// setABwithRelBound, configuration 0 (SZI:1394-1414)
static void set_ab(double rel, double &a, double &b) {
    if (rel >= 0.01) a = 2, b = 2;
    else if (rel >= 0.007) a = 1.75, b = 2;
    else if (rel >= 0.004) a = 1.5, b = 2;
    else if (rel > 0.001) a = 1.25, b = 1.5;
    else a = 1, b = 1;
}

// Summed mode-2 loss of one candidate over the sampled blocks (SZI:1952-1975).
static double predictor_loss(const TuneState &ts, oracle_interp_params p, int start, int end) {
    double loss = 0;
    for (const auto &blk: ts.blocks) {
        std::vector<float> cur(blk);
        InterpEngine e;
        e.set_params(&p);
        std::vector<size_t> lc;
        loss += e.compress(cur.data(), 2, start, end, lc);
    }
    return loss;
}

// CompressTest for ALGO_INTERP, RD target (SZI:1131-1392): mode-1 runs over
// the blocks with one quantizer (outliers accumulate), codes interleaved
// level by level across blocks, encoding_lossless (SIC:784-809) through
// zstd, bitrate and PSNR.
static void compress_test(const TuneState &ts, const oracle_interp_params &p, double rng, double &bitrate,
                          double &psnr, size_t &bytes, std::vector<uint8_t> *raw_out = nullptr,
                          double *sq_out = nullptr) {
    double square_error = 0;
    std::vector<std::vector<int>> bq;
    std::vector<size_t> counts;
    std::vector<float> unpred;
    for (const auto &blk: ts.blocks) {
        std::vector<float> cur(blk);
        InterpEngine e;
        e.set_params(&p);
        e.quantizer.radius = 32768;  // LinearQuantizer<T>(eb): default radius
        square_error += e.compress(cur.data(), 1, 0, 0, counts);
        bq.push_back(e.codes);
        unpred.insert(unpred.end(), e.quantizer.unpred.begin(), e.quantizer.unpred.end());
    }
    std::vector<int> q;
    size_t last = 0;
    for (int k = (int) counts.size() - 1; k >= 0; k--) {
        for (size_t l = 0; l < bq.size(); l++)
            for (size_t m = last; m < counts[k]; m++) q.push_back(bq[l][m]);
        last = counts[k];
    }
    std::vector<uint8_t> raw;
    raw.push_back(2);
    put(raw, p.eb);
    put(raw, (int) 32768);
    put(raw, (size_t) unpred.size());
    const uint8_t *u = (const uint8_t *) unpred.data();
    raw.insert(raw.end(), u, u + unpred.size() * sizeof(float));
    Huffman h;
    h.init(q.data(), q.size());
    h.save(raw);
    h.encode(q.data(), q.size(), raw);
    std::vector<uint8_t> z(ZSTD_compressBound(raw.size()));
    size_t zs = ZSTD_compress(z.data(), z.size(), raw.data(), raw.size(), 3);
    bytes = 8 + zs;
    if (raw_out) *raw_out = raw;
    if (sq_out) *sq_out = square_error;
    size_t per = 1;
    for (size_t d: ts.bdims) per *= d;
    const size_t ele = ts.blocks.size() * per;
    bitrate = 8 * double(bytes) / ele;
    double mse = square_error / ele;
    psnr = 20 * std::log10(rng) - 10 * std::log10(mse);
}

}  // namespace

extern "C" int oracle_compress_test_lorenzo(int N, double eb, const float *field, const size_t *fdims,
                                            const size_t *starts, size_t nblocks, size_t sbs, uint8_t *raw, size_t cap,
                                            size_t *raw_size, size_t *bytes, double *square_error);

namespace {

static int qoz_tune(const oracle_tune_in *in, const float *data, oracle_tune_out *out) {
    const int N = in->N;
    if (N != 2 && N != 3) return -1;
    memset(out, 0, sizeof(*out));
    size_t num = 1;
    for (int i = 0; i < N; i++) num *= in->dims[i];
    const float rngf = (float) in->rng;  // Tuning keeps rng as T (SZI:1590)
    const double rel = in->rel_eb > 0 ? in->rel_eb : in->abs_eb / rngf;
    // QoZ defaults (SZI:1594-1613)
    const double rate = N == 2 ? 0.01 : 0.005;
    size_t max_step = in->max_step > 0 ? in->max_step : (N == 2 ? 64 : 32);
    int levelwise = in->levelwise > 0 ? in->levelwise : (N == 2 ? 6 : 4);
    long sbs = in->sample_block_size > 0 ? in->sample_block_size : (N == 2 ? 64 : 32);
    int max_interp_level = -1;
    for (int i = 0; i < N; i++)
        if (max_interp_level < std::ceil(std::log2((double) in->dims[i])))
            max_interp_level = (unsigned) std::ceil(std::log2((double) in->dims[i]));
    if (max_step > 0) {
        int t = (unsigned) std::log2((double) max_step);
        if (t <= max_interp_level) max_interp_level = t;
        if (levelwise > max_interp_level) levelwise = max_interp_level;
    }
    size_t shortest = in->dims[0];
    for (int i = 0; i < N; i++) shortest = std::min(shortest, in->dims[i]);
    // sample block size (SZI:1776-1784)
    const long minimum_sbs = 16;
    if (sbs < minimum_sbs) sbs = minimum_sbs;
    while ((size_t) sbs >= shortest) sbs /= 2;
    while (rate > 0 && sbs >= 2 * minimum_sbs && (std::pow((double) sbs, N) / (double) num) > 1.5 * rate) sbs /= 2;
    TuneState ts;
    ts.N = N;
    sample_blocks(data, N, in->dims, (size_t) sbs, rate, ts, out);
    ts.bdims.assign(N, (size_t) sbs + 1);
    oracle_interp_params p;
    memset(&p, 0, sizeof(p));
    p.N = N;
    for (int i = 0; i < N; i++) p.dims[i] = (size_t) sbs + 1;
    p.eb = in->abs_eb;
    p.max_step = max_step;
    p.block_size = in->interp_block_size;
    p.radius = 32768;
    double a0, b0;
    set_ab(rel, a0, b0);
    p.alpha = a0, p.beta = b0;
    const int ndir = N == 2 ? 2 : 6;
    const int dir_cands[2] = {0, ndir - 1};
    std::vector<uint8_t> ops(levelwise), dirs(levelwise);
    // predictor tuning (SZI:1938-2000)
    for (int level = levelwise; level > 0; level--) {
        int start = level == levelwise ? 9999 : level, end = level - 1;
        int best_op = 1, best_dir = 0;
        double best = std::numeric_limits<double>::max();
        for (int op = 0; op < 2; op++)
            for (int di = 0; di < 2; di++) {
                p.interp_algo = op;
                p.direction = dir_cands[di];
                double loss = predictor_loss(ts, p, start, end);
                out->pred_loss[level - 1][op * 2 + di] = loss;
                if (loss < best) {
                    best = loss;
                    best_op = op;
                    best_dir = dir_cands[di];
                }
            }
        ops[level - 1] = (uint8_t) best_op;
        dirs[level - 1] = (uint8_t) best_dir;
    }
    out->interp_algo = p.interp_algo;  // the last candidate stays in the Config
    out->direction = p.direction;
    // alpha / beta search (SZI:2340-2420)
    p.levelwise = levelwise;
    memcpy(p.ops, ops.data(), levelwise);
    memcpy(p.dirs, dirs.data(), levelwise);
    // init_alphalist / init_betalist, RD target, no wavelets (SZI:869-933)
    std::vector<double> alphas = {1, 1.25, 1.5, 1.75, 2};
    if (in->ab_list == 1) alphas.insert(alphas.end(), {2.25, 2.5});
    if (in->ab_list >= 2) alphas.insert(alphas.end(), {2.25, 2.5, 2.75, 3});
    const double betas[4] = {1.5, 2, 3, 4};
    double best_a = 1, best_b = 1, bestb = 9999, bestm = 0;
    out->use_interp = 1;
    auto test = [&](double a, double b, double eb, double &br, double &ps) {
        oracle_interp_params q = p;
        q.alpha = a, q.beta = b, q.eb = eb;
        size_t bytes;
        compress_test(ts, q, in->rng, br, ps, bytes);
        int t = out->n_tests;
        if (t < 128) {
            out->test_alpha[t] = a, out->test_beta[t] = b, out->test_eb[t] = eb;
            out->test_bitrate[t] = br, out->test_psnr[t] = ps, out->test_bytes[t] = bytes;
        }
        out->n_tests++;
    };
    for (size_t i = 0; i < alphas.size(); i++) {
        for (int j = 0; j < 4; j++) {
            const double alpha = alphas[i], beta = betas[j];
            if (alpha >= 1 && alpha > beta) continue;
            double bitrate, metric;
            test(alpha, beta, in->abs_eb, bitrate, metric);
            const bool stop = (alpha >= 1 && std::pow(alpha, max_interp_level - 1) <= beta) ||
                              (alpha < 1 && alpha * (max_interp_level - 1) <= beta);
            if (metric >= bestm && bitrate <= bestb) {
                best_a = alpha, best_b = beta, bestb = bitrate, bestm = metric;
            } else if (metric <= bestm && bitrate >= bestb) {
                if (stop) break;
                continue;
            } else {
                const double fix = bitrate / bestb;
                double br_r, m_r;
                test(alpha, beta, in->abs_eb * fix, br_r, m_r);
                const double a = (metric - m_r) / (bitrate - br_r);
                const double b = metric - a * bitrate;
                const double reg = a * bestb + b;
                if (reg > bestm) best_a = alpha, best_b = beta, bestb = bitrate, bestm = metric;
            }
            if (stop) break;
        }
    }
    // the Lorenzo test (SZI:2431-2493): CompressTest with ALGO_LORENZO_REG,
    // useFast = false, lorenzoBrFix = 1, recorded with alpha = beta = -1
    if (in->test_lorenzo) {
        const size_t nb = std::min<size_t>(out->n_blocks, 4096);
        std::vector<size_t> starts(3 * nb);
        for (size_t b = 0; b < nb; b++)
            for (int a = 0; a < 3; a++) starts[3 * b + a] = out->block_starts[b][a];
        const double ele = std::pow((double) (sbs + 1), N) * (double) nb;
        auto lor = [&](double eb, double &br, double &ps) -> bool {
            size_t raw_size = 0, bytes = 0;
            double sq = 0;
            // no buffer asked for: -3 (capacity) after the size and error are set
            const int rc = oracle_compress_test_lorenzo(N, eb, data, in->dims, starts.data(), nb, (size_t) sbs,
                                                        nullptr, 0, &raw_size, &bytes, &sq);
            if (rc != 0 && rc != -3) return false;
            br = 8 * double(bytes) / ele;
            ps = 20 * std::log10(in->rng) - 10 * std::log10(sq / ele);
            int t = out->n_tests;
            if (t < 128) {
                out->test_alpha[t] = -1, out->test_beta[t] = -1, out->test_eb[t] = eb;
                out->test_bitrate[t] = br, out->test_psnr[t] = ps, out->test_bytes[t] = bytes;
            }
            out->n_tests++;
            return true;
        };
        double bitrate, metric;
        if (!lor(in->abs_eb, bitrate, metric)) return -2;
        if (metric >= bestm && bitrate <= bestb) {
            bestb = bitrate, bestm = metric, best_a = best_b = -1, out->use_interp = 0;
        } else if (metric <= bestm && bitrate >= bestb) {
            out->use_interp = 1;
        } else {
            const double fix = bitrate / bestb;
            double br_r, m_r;
            if (!lor(in->abs_eb * fix, br_r, m_r)) return -2;
            const double a = (metric - m_r) / (bitrate - br_r);
            if (a > 0) {
                const double b = metric - a * bitrate;
                if (a * bestb + b > bestm) bestb = bitrate, bestm = metric, best_a = best_b = -1, out->use_interp = 0;
            }
        }
    }
    out->alpha = best_a;
    out->beta = best_b;
    out->levelwise = levelwise;
    memcpy(out->ops, ops.data(), levelwise);
    memcpy(out->dirs, dirs.data(), levelwise);
    out->max_step = max_step;
    out->sample_block_size = (int) sbs;
    return 0;
}

}  // namespace

extern "C" {

int oracle_interp_compress(const oracle_interp_params *p, float *data, int tuning, int start_level, int end_level,
                           int32_t *codes, size_t *n_codes, float *unpred, size_t *n_unpred, double *loss,
                           size_t *level_counts, int *n_levels) {
    if (!p || p->N < 1 || p->N > 4) return -1;
    InterpEngine e;
    e.set_params(p);
    std::vector<size_t> lc;
    double err = e.compress(data, tuning, start_level, end_level, lc);
    if (codes) memcpy(codes, e.codes.data(), e.codes.size() * sizeof(int32_t));
    if (n_codes) *n_codes = e.codes.size();
    if (unpred) memcpy(unpred, e.quantizer.unpred.data(), e.quantizer.unpred.size() * sizeof(float));
    if (n_unpred) *n_unpred = e.quantizer.unpred.size();
    if (loss) *loss = err;
    if (level_counts) {
        for (size_t i = 0; i < lc.size() && i < 64; i++) level_counts[i] = lc[i];
    }
    if (n_levels) *n_levels = e.levels;
    return 0;
}

int oracle_interp_raw(const oracle_interp_params *p, const float *data, uint8_t *out, size_t cap, size_t *n_out) {
    std::vector<uint8_t> buf;
    if (!interp_raw(p, data, buf)) return -2;
    *n_out = buf.size();
    if (buf.size() > cap) return -3;
    memcpy(out, buf.data(), buf.size());
    return 0;
}

int oracle_interp_decompress_raw(const uint8_t *raw, size_t n, int N, float *out, size_t num) {
    return interp_decompress_raw(raw, n, N, out, num) ? 0 : -2;
}

int oracle_huffman_encode(const int32_t *codes, size_t n, uint8_t *out, size_t cap, size_t *n_out) {
    if (n == 0) return -1;
    Huffman h;
    if (!h.init(codes, n)) return -2;
    std::vector<uint8_t> buf;
    h.save(buf);
    h.encode(codes, n, buf);
    *n_out = buf.size();
    if (buf.size() > cap) return -3;
    memcpy(out, buf.data(), buf.size());
    return 0;
}

int oracle_huffman_decode(const uint8_t *buf, size_t n, int32_t *out, size_t num, size_t *n_used) {
    HuffDecoder d;
    size_t used = d.decode(buf, n, out, num);
    if (!used) return -2;
    if (n_used) *n_used = used;
    return 0;
}

int oracle_compress_stream(const oracle_interp_params *p, int rel_mode, double eb_in, const float *data,
                           uint8_t *out, size_t cap, size_t *n_out) {
    size_t num = 1;
    for (int i = 0; i < p->N; i++) num *= p->dims[i];
    oracle_interp_params q = *p;
    double rel = -1.0;
    if (rel_mode) {
        float mn, mx;
        oracle_range(data, num, &mn, &mx);
        double rng = (double) (float) (mx - mn);
        rel = eb_in;
        q.eb = rel * rng;
    } else {
        q.eb = eb_in;
    }
    std::vector<uint8_t> raw;
    if (!interp_raw(&q, data, raw)) return -2;
    std::vector<uint8_t> s(8 + ZSTD_compressBound(raw.size()));
    uint64_t rawLen = raw.size();
    memcpy(s.data(), &rawLen, 8);
    size_t z = ZSTD_compress(s.data() + 8, s.size() - 8, raw.data(), raw.size(), 3);
    if (ZSTD_isError(z)) return -4;
    s.resize(8 + z);
    size_t before = s.size();
    config_save(s, &q, q.eb, rel);
    put(s, (int) (s.size() - before));
    *n_out = s.size();
    if (s.size() > cap) return -3;
    memcpy(out, s.data(), s.size());
    return 0;
}

int oracle_decompress_stream(const uint8_t *stream, size_t n, float *out, size_t num) {
    int confSize;
    memcpy(&confSize, stream + n - 4, 4);
    const uint8_t *c = stream + n - 4 - confSize;
    int N = (int) (signed char) c[0];
    size_t dims[4] = {0, 0, 0, 0};
    if (N < 1 || N > 4) return -1;
    memcpy(dims, c + 1, N * sizeof(size_t));
    uint8_t algo = c[1 + N * 8 + 8];
    uint64_t rawLen;
    memcpy(&rawLen, stream, 8);
    std::vector<uint8_t> raw(rawLen);
    size_t z = ZSTD_decompress(raw.data(), rawLen, stream + 8, (n - 4 - confSize) - 8);
    if (ZSTD_isError(z) || z != rawLen) return -4;
    if (algo == 0) {
        // SZ_decompress_LorenzoReg (LR:117-135): fast frontend for 3-D without
        // regression2, else the general frontend with the Config's predictors
        const uint8_t *f = c + 44 + 8 * N;  // lorenzo, lorenzo2, regression, regression2
        if (N == 3 && !f[3]) {
            if (dims[0] * dims[1] * dims[2] != num) return -1;
            return lr_decompress_raw(raw.data(), raw.size(), dims, out) ? 0 : -2;
        }
        if (f[3]) return -1;  // the second-order regression predictor: not restated
        return gf_decompress_raw(raw.data(), raw.size(), N, f[0], f[1], f[2], out, num) ? 0 : -2;
    }
    return interp_decompress_raw(raw.data(), raw.size(), N, out, num) ? 0 : -2;
}

int oracle_lr_compress(const oracle_lr_params *p, const float *data, int32_t *codes, float *recon,
                       int32_t *indicator, size_t *n_blocks, float *raw_coeffs,
                       int32_t *reg_codes, float *reg_coeffs, size_t *n_reg,
                       float *reg_unpred, size_t *n_reg_unpred, float *unpred, size_t *n_unpred) {
    if (!p || p->block_size < 1 || p->pred_dim < 1 || p->pred_dim > 3) return -1;
    if (!p->dims[0] || !p->dims[1] || !p->dims[2]) return -1;
    LREngine e;
    e.set(p);
    e.compress(data);
    if (codes) memcpy(codes, e.codes.data(), e.codes.size() * sizeof(int32_t));
    if (recon) memcpy(recon, e.recon.data(), e.recon.size() * sizeof(float));
    if (indicator) memcpy(indicator, e.indicator.data(), e.indicator.size() * sizeof(int32_t));
    if (n_blocks) *n_blocks = e.nblocks;
    if (raw_coeffs) memcpy(raw_coeffs, e.raw_coeffs.data(), e.raw_coeffs.size() * sizeof(float));
    if (reg_codes) memcpy(reg_codes, e.reg_types.data(), e.reg_types.size() * sizeof(int32_t));
    if (reg_coeffs) memcpy(reg_coeffs, e.reg_coeffs.data(), e.reg_coeffs.size() * sizeof(float));
    if (n_reg) *n_reg = e.reg_coeffs.size() / 4;
    if (reg_unpred) memcpy(reg_unpred, e.reg_unpred.data(), e.reg_unpred.size() * sizeof(float));
    if (n_reg_unpred) *n_reg_unpred = e.reg_unpred.size();
    if (unpred) memcpy(unpred, e.quantizer.unpred.data(), e.quantizer.unpred.size() * sizeof(float));
    if (n_unpred) *n_unpred = e.quantizer.unpred.size();
    return 0;
}

int oracle_lr_raw(const oracle_lr_params *p, const float *data, uint8_t *out, size_t cap, size_t *n_out) {
    if (!p || p->block_size < 1 || p->pred_dim < 1 || p->pred_dim > 3) return -1;
    std::vector<uint8_t> buf;
    if (!lr_raw(p, data, buf)) return -2;
    *n_out = buf.size();
    if (buf.size() > cap) return -3;
    memcpy(out, buf.data(), buf.size());
    return 0;
}

int oracle_lr_decompress_raw(const uint8_t *raw, size_t n, const size_t *dims, float *out) {
    return lr_decompress_raw(raw, n, dims, out) ? 0 : -2;
}

int oracle_lr_compress_stream(const oracle_lr_params *p, int rel_mode, double eb_in, const float *data,
                              uint8_t *out, size_t cap, size_t *n_out) {
    if (!p || p->block_size < 1 || p->pred_dim < 1 || p->pred_dim > 3) return -1;
    size_t num = p->dims[0] * p->dims[1] * p->dims[2];
    oracle_lr_params q = *p;
    double rel = -1.0;
    if (rel_mode) {
        float mn, mx;
        oracle_range(data, num, &mn, &mx);
        rel = eb_in;
        q.eb = rel * (double) (float) (mx - mn);
    } else {
        q.eb = eb_in;
    }
    std::vector<uint8_t> raw;
    if (!lr_raw(&q, data, raw)) return -2;
    std::vector<uint8_t> s(8 + ZSTD_compressBound(raw.size()));
    uint64_t rawLen = raw.size();
    memcpy(s.data(), &rawLen, 8);
    size_t z = ZSTD_compress(s.data() + 8, s.size() - 8, raw.data(), raw.size(), 3);
    if (ZSTD_isError(z)) return -4;
    s.resize(8 + z);
    oracle_interp_params ip;
    memset(&ip, 0, sizeof(ip));
    ip.N = 3;
    for (int i = 0; i < 3; i++) ip.dims[i] = q.dims[i];
    ip.interp_algo = 1;       // Config defaults the trailer carries (Config.hpp:394-499)
    ip.alpha = -1, ip.beta = -1;
    ip.block_size = 32;
    ip.radius = q.radius;
    TrailerFields tf;
    tf.cmprAlgo = 0;          // ALGO_LORENZO_REG
    tf.lorenzo = q.lorenzo != 0, tf.lorenzo2 = q.lorenzo2 != 0, tf.regression = q.regression != 0;
    tf.blockSize = q.block_size;
    tf.pred_dim = q.pred_dim;
    size_t before = s.size();
    config_save(s, &ip, q.eb, rel, tf);
    put(s, (int) (s.size() - before));
    *n_out = s.size();
    if (s.size() > cap) return -3;
    memcpy(out, s.data(), s.size());
    return 0;
}

int oracle_compress_test(const oracle_interp_params *p, const float *field, const size_t *fdims, const size_t *starts,
                         size_t nblocks, size_t sbs, double rng, uint8_t *raw, size_t cap, size_t *raw_size,
                         size_t *bytes, double *square_error) {
    if (!p || (p->N != 2 && p->N != 3)) return -1;
    TuneState ts;
    ts.N = p->N;
    ts.bdims.assign(p->N, sbs + 1);
    oracle_tune_out *tmp = new oracle_tune_out;
    memset(tmp, 0, sizeof(*tmp));
    for (size_t b = 0; b < nblocks; b++) {
        TuneState one;
        size_t d[3] = {fdims[0], fdims[1], p->N == 3 ? fdims[2] : 1};
        // one block at the given start: the sampling walk with a stride of 1 over a 1-block grid
        std::vector<float> blk;
        const size_t *s0 = starts + 3 * b;
        if (p->N == 2) {
            for (size_t i = 0; i <= sbs; i++)
                for (size_t j = 0; j <= sbs; j++) blk.push_back(field[(s0[0] + i) * d[1] + s0[1] + j]);
        } else {
            for (size_t i = 0; i <= sbs; i++)
                for (size_t j = 0; j <= sbs; j++)
                    for (size_t k = 0; k <= sbs; k++)
                        blk.push_back(field[((s0[0] + i) * d[1] + s0[1] + j) * d[2] + s0[2] + k]);
        }
        ts.blocks.push_back(std::move(blk));
    }
    delete tmp;
    double br, ps;
    std::vector<uint8_t> r;
    compress_test(ts, *p, rng, br, ps, *bytes, &r, square_error);
    *raw_size = r.size();
    if (r.size() > cap) return -3;
    memcpy(raw, r.data(), r.size());
    return 0;
}

int oracle_gf_compress(const oracle_gf_params *p, const float *data, int32_t *codes, float *recon,
                       int32_t *selection, size_t *n_sel, float *unpred, size_t *n_unpred) {
    GFEngine e;
    if (!p || !e.set(p)) return -1;
    std::vector<float> v(data, data + e.num());
    e.compress(v.data());
    if (codes) memcpy(codes, e.codes.data(), e.codes.size() * sizeof(int32_t));
    if (recon) memcpy(recon, v.data(), v.size() * sizeof(float));
    if (selection) memcpy(selection, e.selection.data(), e.selection.size() * sizeof(int32_t));
    if (n_sel) *n_sel = e.selection.size();
    if (unpred) memcpy(unpred, e.quantizer.unpred.data(), e.quantizer.unpred.size() * sizeof(float));
    if (n_unpred) *n_unpred = e.quantizer.unpred.size();
    return 0;
}

int oracle_gf_raw(const oracle_gf_params *p, const float *data, uint8_t *out, size_t cap, size_t *n_out) {
    if (!p) return -1;
    std::vector<uint8_t> buf;
    if (!gf_raw(p, data, buf)) return -2;
    *n_out = buf.size();
    if (buf.size() > cap) return -3;
    memcpy(out, buf.data(), buf.size());
    return 0;
}

int oracle_gf_decompress_raw(const uint8_t *raw, size_t n, int N, int lorenzo, int lorenzo2, int regression,
                             float *out, size_t num) {
    return gf_decompress_raw(raw, n, N, lorenzo, lorenzo2, regression, out, num) ? 0 : -2;
}

int oracle_gf_compress_stream(const oracle_gf_params *p, int rel_mode, double eb_in, const float *data,
                              uint8_t *out, size_t cap, size_t *n_out) {
    if (!p || p->N < 1 || p->N > 4) return -1;
    size_t num = 1;
    for (int d = 0; d < p->N; d++) num *= p->dims[d];
    oracle_gf_params q = *p;
    double rel = -1.0;
    if (rel_mode) {
        float mn, mx;
        oracle_range(data, num, &mn, &mx);
        rel = eb_in;
        q.eb = rel * (double) (float) (mx - mn);
    } else {
        q.eb = eb_in;
    }
    std::vector<uint8_t> raw;
    if (!gf_raw(&q, data, raw)) return -2;
    std::vector<uint8_t> s(8 + ZSTD_compressBound(raw.size()));
    uint64_t rawLen = raw.size();
    memcpy(s.data(), &rawLen, 8);
    size_t z = ZSTD_compress(s.data() + 8, s.size() - 8, raw.data(), raw.size(), 3);
    if (ZSTD_isError(z)) return -4;
    s.resize(8 + z);
    oracle_interp_params ip;
    memset(&ip, 0, sizeof(ip));
    ip.N = q.N;
    for (int i = 0; i < q.N; i++) ip.dims[i] = q.dims[i];
    ip.interp_algo = 1;
    ip.alpha = -1, ip.beta = -1;
    ip.block_size = 32;
    ip.radius = q.radius;
    TrailerFields tf;
    tf.cmprAlgo = 0;  // ALGO_LORENZO_REG
    tf.lorenzo = q.lorenzo != 0, tf.lorenzo2 = q.lorenzo2 != 0, tf.regression = q.regression != 0;
    tf.blockSize = q.block_size;
    size_t before = s.size();
    config_save(s, &ip, q.eb, rel, tf);
    put(s, (int) (s.size() - before));
    *n_out = s.size();
    if (s.size() > cap) return -3;
    memcpy(out, s.data(), s.size());
    return 0;
}

int oracle_compress_test_lorenzo(int N, double eb, const float *field, const size_t *fdims, const size_t *starts,
                                 size_t nblocks, size_t sbs, uint8_t *raw, size_t cap, size_t *raw_size,
                                 size_t *bytes, double *square_error) {
    if (N != 2 && N != 3) return -1;
    oracle_gf_params p;
    memset(&p, 0, sizeof(p));
    p.N = N;
    for (int d = 0; d < N; d++) p.dims[d] = sbs + 1;
    p.eb = eb;
    p.block_size = 5;
    p.lorenzo = p.lorenzo2 = 1;
    p.radius = 65536;
    GFEngine e;
    if (!e.set(&p)) return -1;
    const size_t per = e.num();
    double sq = 0;
    std::vector<float> blk, cur;
    for (size_t b = 0; b < nblocks; b++) {
        const size_t *s0 = starts + 3 * b;
        blk.clear();
        if (N == 2) {
            for (size_t i = 0; i <= sbs; i++)
                for (size_t j = 0; j <= sbs; j++) blk.push_back(field[(s0[0] + i) * fdims[1] + s0[1] + j]);
        } else {
            for (size_t i = 0; i <= sbs; i++)
                for (size_t j = 0; j <= sbs; j++)
                    for (size_t k = 0; k <= sbs; k++)
                        blk.push_back(field[((s0[0] + i) * fdims[1] + s0[1] + j) * fdims[2] + s0[2] + k]);
        }
        cur = blk;
        e.compress(cur.data());
        for (size_t j = 0; j < per; j++) {  // SZI:1287-1292: float difference, float square
            float v = blk[j] - cur[j];
            sq += v * v;
        }
    }
    std::vector<uint8_t> r;
    if (!e.save(r)) return -2;
    std::vector<uint8_t> z(ZSTD_compressBound(r.size()));
    size_t zs = ZSTD_compress(z.data(), z.size(), r.data(), r.size(), 3);
    if (ZSTD_isError(zs)) return -4;
    *bytes = 8 + zs;
    *square_error = sq;
    *raw_size = r.size();
    if (r.size() > cap) return -3;
    memcpy(raw, r.data(), r.size());
    return 0;
}

int oracle_sampling(const float *data, int N, const size_t *dims, float *out, size_t cap, size_t *sample_dims,
                    size_t *sampling_block) {
    if (N < 1 || N > 4) return -1;
    std::vector<float> s;
    bool sampled = sz3_sampling(data, N, dims, s, sample_dims, *sampling_block);
    if (!sampled) return 1;
    if (s.size() > cap) return -3;
    memcpy(out, s.data(), s.size() * sizeof(float));
    return 0;
}

int oracle_block_interp_test(const float *data, int N, const size_t *dims, double eb, int algo, int dir, size_t bs,
                             uint8_t *raw, size_t cap, size_t *raw_size, size_t *bytes) {
    if (N < 1 || N > 3) return -1;
    std::vector<uint8_t> r;
    *bytes = sz3_block_test(data, N, dims, eb, algo, dir, bs, &r);
    if (!*bytes) return -2;
    *raw_size = r.size();
    if (raw) {
        if (r.size() > cap) return -3;
        memcpy(raw, r.data(), r.size());
    }
    return 0;
}

int oracle_quant_intervals(const float *data, const size_t *dims, double eb, size_t *samples) {
    return sz3_quant_intervals(data, dims[0], dims[1], dims[2], eb, samples);
}

int oracle_sz3_tune(const float *data, int N, const size_t *dims, double eb, double rel, oracle_sz3_out *out) {
    if (N < 1 || N > 3 || !out) return -1;
    return sz3_tune(data, N, dims, eb, rel, out);
}

// SZ_compress for ALGO_INTERP_LORENZO without conf.QoZ / conf.FZ (the
// default `faz` mode): calAbsErrorBound, range, the SZ3 tuning, then
// SZ_compress_Interp or SZ_compress_LorenzoReg with the tuned Config.
int oracle_sz3_compress_stream(const float *data, int N, const size_t *dims, int rel_mode, double eb_in,
                               uint8_t *out, size_t cap, size_t *n_out, oracle_sz3_out *tune) {
    if (N < 1 || N > 3) return -1;
    size_t num = 1;
    for (int d = 0; d < N; d++) num *= dims[d];
    float mn, mx;
    oracle_range(data, num, &mn, &mx);
    const double rng = (double) (float) (mx - mn);
    double eb, rel;
    if (rel_mode) {
        rel = eb_in;
        eb = rel * rng;
    } else {
        eb = eb_in;
        rel = eb / rng;  // SZI:2569-2570
    }
    oracle_sz3_out t;
    int rc = sz3_tune(data, N, dims, eb, rel, &t);
    if (rc) return rc;
    if (tune) *tune = t;
    std::vector<uint8_t> s;
    oracle_interp_params ip;
    memset(&ip, 0, sizeof(ip));
    ip.N = N;
    for (int d = 0; d < N; d++) ip.dims[d] = dims[d];
    ip.interp_algo = t.interp_algo;
    ip.direction = t.direction;
    ip.alpha = -1, ip.beta = -1;
    ip.block_size = 32;
    ip.radius = 32768;
    TrailerFields tf;
    if (t.use_interp) {
        ip.eb = eb;
        std::vector<uint8_t> raw;
        if (!interp_raw(&ip, data, raw)) return -2;
        s.resize(8 + ZSTD_compressBound(raw.size()));
        uint64_t rawLen = raw.size();
        memcpy(s.data(), &rawLen, 8);
        size_t z = ZSTD_compress(s.data() + 8, s.size() - 8, raw.data(), raw.size(), 3);
        s.resize(8 + z);
        tf.cmprAlgo = 2;
    } else {
        std::vector<uint8_t> raw;
        const int bsd = default_block_size(N);
        if (N == 3) {
            oracle_lr_params p;
            memset(&p, 0, sizeof(p));
            for (int d = 0; d < 3; d++) p.dims[d] = dims[d];
            p.eb = eb;
            p.block_size = bsd;
            p.pred_dim = 3;
            p.lorenzo = p.lorenzo2 = 1;
            p.radius = t.quantbin / 2;
            if (!lr_raw(&p, data, raw)) return -2;
        } else {
            oracle_gf_params p;
            memset(&p, 0, sizeof(p));
            p.N = N;
            for (int d = 0; d < N; d++) p.dims[d] = dims[d];
            p.eb = eb;
            p.block_size = bsd;
            p.lorenzo = p.lorenzo2 = 1;
            p.radius = t.quantbin / 2;
            if (!gf_raw(&p, data, raw)) return -2;
        }
        s.resize(8 + ZSTD_compressBound(raw.size()));
        uint64_t rawLen = raw.size();
        memcpy(s.data(), &rawLen, 8);
        size_t z = ZSTD_compress(s.data() + 8, s.size() - 8, raw.data(), raw.size(), 3);
        s.resize(8 + z);
        tf.cmprAlgo = 0;
        tf.lorenzo = true, tf.lorenzo2 = true, tf.regression = false;
        tf.blockSize = bsd;
        ip.radius = t.quantbin / 2;
    }
    size_t before = s.size();
    config_save(s, &ip, eb, rel, tf);
    put(s, (int) (s.size() - before));
    *n_out = s.size();
    if (s.size() > cap) return -3;
    memcpy(out, s.data(), s.size());
    return 0;
}

int oracle_qoz_tune(const oracle_tune_in *in, const float *data, oracle_tune_out *out) {
    if (!in || !data || !out) return -1;
    return qoz_tune(in, data, out);
}

int oracle_range(const float *data, size_t n, float *mn, float *mx) {
    float a = data[0], b = data[0];
    for (size_t i = 1; i < n; i++) {
        if (b < data[i]) b = data[i];
        if (a > data[i]) a = data[i];
    }
    *mn = a;
    *mx = b;
    return 0;
}

int oracle_interp_seq_index(const oracle_interp_params *p, int64_t *seq) {
    InterpEngine e;
    e.set_params(p);
    size_t num = 1;
    for (int i = 0; i < p->N; i++) num *= p->dims[i];
    std::vector<float> data(num, 0.0f);
    std::vector<size_t> lc;
    e.compress(data.data(), 0, 0, 0, lc);
    for (size_t i = 0; i < num; i++) seq[i] = -1;
    for (size_t k = 0; k < e.code_pos.size(); k++) seq[e.code_pos[k]] = (int64_t) k;
    return 0;
}

}  // extern "C"
