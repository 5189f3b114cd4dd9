// Host side of the Huffman stage: tree construction, code table, tree
// serialisation and bit-serial decoding, reproducing HuffmanEncoder<int>
// (include/QoZ/encoder/HuffmanEncoder.hpp) byte for byte.
//
// The GPU supplies per-symbol counts and the first code position of every
// symbol.  The reference counts symbols in a ska::unordered_map and builds its
// heap in that map's iteration order, which depends only on the order in
// which distinct keys were first inserted; the first positions give that
// order.
#include "huffman_host.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace faz {

namespace {

// ska::unordered_map<int, size_t> bucket/chain structure
// (utils/ska_hash/unordered_map.hpp:254-303, 412-446, 604-643;
// flat_hash_map.hpp:1268-1295): Fibonacci hashing into a power-of-two bucket
// array, growth to max(4, 2*buckets) when size+1 would exceed the bucket count
// (max load factor 1), new keys at the head of their chain, rehash walking the
// old buckets upwards and pushing each element to the head of its new chain,
// iteration from the top bucket down and along each chain.
class SkaOrder {
public:
    explicit SkaOrder(size_t expected) { next_.reserve(expected); keys_.reserve(expected); }

    void insert(int key) {
        if (heads_.empty() || keys_.size() + 1 > heads_.size()) grow();
        int e = (int) keys_.size();
        keys_.push_back(key);
        next_.push_back(-1);
        link_front(e);
    }

    std::vector<int> iteration_order() const {
        std::vector<int> out;
        out.reserve(keys_.size());
        for (size_t b = heads_.size(); b-- > 0;)
            for (int e = heads_[b]; e >= 0; e = next_[e]) out.push_back(keys_[e]);
        return out;
    }

private:
    std::vector<int> heads_;
    std::vector<int> next_;
    std::vector<int> keys_;
    int shift_ = 63;

    size_t bucket(int key) const {
        return (size_t) ((0x9E3779B97F4A7C15ull * (uint64_t) (int64_t) key) >> shift_);
    }
    void link_front(int e) {
        size_t b = bucket(keys_[e]);
        next_[e] = heads_[b];
        heads_[b] = e;
    }
    void grow() {
        size_t old = heads_.empty() ? 1 : heads_.size();
        size_t want = std::max<size_t>(4, 2 * old);
        size_t nb = 2;
        while (nb < want) nb <<= 1;
        int lg = 0;
        while (((size_t) 1 << lg) < nb) lg++;
        std::vector<int> old_heads;
        old_heads.swap(heads_);
        heads_.assign(nb, -1);
        shift_ = 64 - lg;
        for (size_t b = 0; b < old_heads.size(); b++) {
            int e = old_heads[b];
            while (e >= 0) {
                int nx = next_[e];
                link_front(e);
                e = nx;
            }
        }
    }
};

// HuffmanEncoder's 1-based binary heap (qinsert/qremove, HUF:455-485).  The
// entries carry their node's frequency, so the sift loops compare without
// reaching into the node arrays.  qremove's sift-down is done bottom-up: the
// hole descends along the smaller child (the left one on ties) to the
// bottom, then the moved element climbs back while its parent's frequency is
// not below its own.  Along that path frequencies never decrease, so this
// moves up exactly the elements smaller than the moved one, as the
// reference's top-down loop does: the same array after every operation,
// with one branch-free compare per level on the way down.
class Heap {
public:
    void reset(size_t cap) {
        if (q_.size() < 4 * cap + 8) q_.resize(4 * cap + 8);  // prefetches reach 4 i
        end_ = 1;
    }
    void insert(int n, uint64_t f) {
        int i = end_++, j;
        if (q_.size() <= 4 * (size_t) i + 8) q_.resize(8 * (size_t) i + 8);
        while ((j = i >> 1)) {
            if (q_[j].f <= f) break;
            q_[i] = q_[j];
            i = j;
        }
        q_[i] = {f, n};
    }
    int remove() {
        const int n = q_[1].n;
        end_--;
        const E x = q_[end_];
        int i = 1, l;
        while ((l = i << 1) < end_) {
            // the four grandchildren of i share a cache line: fetch it while
            // the children are compared
            __builtin_prefetch(&q_[(size_t) 4 * i]);
            // q_[end_] (the moved entry's old slot) stays readable: no bounds branch
            l += (int) ((l + 1 < end_) & (q_[l + 1].f < q_[l].f));
            q_[i] = q_[l];
            i = l;
        }
        while (i > 1) {
            const int p = i >> 1;
            if (q_[p].f < x.f) break;
            q_[i] = q_[p];
            i = p;
        }
        q_[i] = x;
        return n;
    }
    int size() const { return end_ - 1; }
    int top() const { return q_[1].n; }

private:
    struct E {
        uint64_t f;
        int n;
    };
    std::vector<E> q_;
    int end_ = 1;
};

// Pre-order numbering of the tree (pad_tree, HUF:569-584; internal nodes
// carry C = 0) written as HuffmanEncoder::save's arrays at w: the
// endianness byte, L and R (T1 each), C (int32), t (uint8).
template<class T1>
void serialise(const HuffTable &tab, unsigned nodeCount, uint8_t *w) {
    *w++ = 0;  // endianness marker: little-endian
    uint8_t *L = w, *R = w + (size_t) nodeCount * sizeof(T1), *C = w + 2 * (size_t) nodeCount * sizeof(T1);
    uint8_t *t = C + (size_t) nodeCount * 4;
    memset(w, 0, (size_t) nodeCount * (2 * sizeof(T1) + 5));
    auto put1 = [](uint8_t *a, unsigned i, unsigned v) {
        const T1 x = (T1) v;
        memcpy(a + (size_t) i * sizeof(T1), &x, sizeof(T1));
    };
    struct Item {
        int32_t node;
        uint32_t at;    // slot of the node, or of the parent for a pending right child
        uint32_t right; // 1: a right child whose slot is the next number
    };
    std::vector<Item> st;
    st.reserve(160);
    st.push_back({tab.root, 0, 0});
    unsigned counter = 0;
    while (!st.empty()) {
        const Item it = st.back();
        st.pop_back();
        unsigned slot = it.at;
        if (it.right) {
            slot = ++counter;
            put1(R, it.at, slot);
        }
        const int32_t nd = it.node;
        const bool leaf = tab.left[nd] < 0 && tab.right[nd] < 0;
        if (leaf) {
            memcpy(C + (size_t) slot * 4, &tab.sym[nd], 4);
            t[slot] = 1;
            continue;
        }
        // the right subtree is numbered after the whole left one
        if (tab.right[nd] >= 0) st.push_back({tab.right[nd], slot, 1});
        if (tab.left[nd] >= 0) {
            const unsigned ls = ++counter;
            put1(L, slot, ls);
            st.push_back({tab.left[nd], ls, 0});
        }
    }
}


}  // namespace

namespace {

// Ascending sort of the keys (first position << 32 | symbol): LSD radix over
// 11-bit digits, skipping digits that every key shares; std::sort for short
// lists.
void sort_keys(std::vector<uint64_t> &k, std::vector<uint64_t> &tmp) {
    const size_t n = k.size();
    if (n < 256) {
        std::sort(k.begin(), k.end());
        return;
    }
    uint64_t all_or = 0, all_and = ~0ull;
    for (uint64_t v: k) {
        all_or |= v;
        all_and &= v;
    }
    const uint64_t varying = all_or ^ all_and;
    tmp.resize(n);
    uint64_t *src = k.data(), *dst = tmp.data();
    constexpr int kBits = 11;
    constexpr uint32_t kBuckets = 1u << kBits;
    std::vector<uint32_t> cnt(kBuckets);
    for (int sh = 0; sh < 64; sh += kBits) {
        if (!((varying >> sh) & (kBuckets - 1))) continue;
        std::fill(cnt.begin(), cnt.end(), 0u);
        for (size_t i = 0; i < n; i++) cnt[(src[i] >> sh) & (kBuckets - 1)]++;
        uint32_t run = 0;
        for (uint32_t b = 0; b < kBuckets; b++) {
            const uint32_t c = cnt[b];
            cnt[b] = run;
            run += c;
        }
        for (size_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & (kBuckets - 1)]++] = src[i];
        std::swap(src, dst);
    }
    if (src != k.data()) std::copy(src, src + n, k.data());
}

// The tree of the present symbols (tab.keys: first position << 32 | symbol)
// and its code table; the tree bytes too unless defer.
template<class FreqOf>
int build_present(FreqOf freq_of, HuffTable &tab, bool defer) {
    std::vector<uint64_t> &keys = tab.keys;
    if (keys.empty()) return -1;
    sort_keys(keys, tab.scratch);
    const size_t np = keys.size();
    int mn = (int) (uint32_t) keys[0], mx = mn;
    for (uint64_t k: keys) {
        const int sy = (int) (uint32_t) k;
        mn = std::min(mn, sy);
        mx = std::max(mx, sy);
    }
    // a table of the same span keeps its arrays: only last build's codewords
    // are cleared (the outlier code 0 makes the span tens of thousands of
    // states wide for a few hundred present symbols)
    const bool reuse = tab.offset == mn && tab.n_states == mx - mn + 1 && (int) tab.code.size() == tab.n_states &&
                       (int) tab.len.size() == tab.n_states;
    tab.offset = mn;
    tab.n_states = mx - mn + 1;

    SkaOrder ska(np);
    for (uint64_t k: keys) ska.insert((int) (uint32_t) k);
    const std::vector<int> order = ska.iteration_order();

    // nodes: leaves 0..np-1 in the map's iteration order, then the merges
    const size_t nn = 2 * np - 1;
    tab.left.assign(nn, -1);
    tab.right.assign(nn, -1);
    tab.sym.resize(nn);
    std::vector<uint64_t> freq(nn);
    static thread_local Heap heap;
    heap.reset(2 * np);
    for (size_t i = 0; i < np; i++) {
        const int key = order[i];
        freq[i] = freq_of(key);
        tab.sym[i] = key - mn;
        heap.insert((int) i, freq[i]);
    }
    // qinsert(new_node(0, 0, qremove(), qremove())) in the reference's g++
    // build evaluates the arguments right to left: the first node removed is
    // the right child.
    int32_t next = (int32_t) np;
    while (heap.size() > 1) {
        const int right = heap.remove();
        const int left = heap.remove();
        tab.left[next] = left;
        tab.right[next] = right;
        tab.sym[next] = 0;
        freq[next] = freq[left] + freq[right];
        heap.insert(next, freq[next]);
        next++;
    }
    tab.root = heap.top();

    // codes: left = 0, right = 1, MSB-aligned (build_code, HUF:493-521)
    if (reuse) {
        for (int32_t k: tab.used) {
            tab.code[k] = 0;
            tab.len[k] = 0;
        }
    } else {
        tab.code.assign(tab.n_states, 0);
        tab.len.assign(tab.n_states, 0);
    }
    tab.used.clear();
    struct Item {
        int32_t node;
        int32_t depth;
        uint64_t bits;
    };
    Item st[136];
    int sp = 0;
    st[sp++] = {tab.root, 0, 0};
    while (sp) {
        const Item it = st[--sp];
        const int32_t l = tab.left[it.node], r = tab.right[it.node];
        if (l < 0 && r < 0) {
            if (it.depth > 64) return -2;  // codes beyond 64 bits need > 2.7e13 symbols
            const int32_t c = tab.sym[it.node];
            tab.code[c] = it.depth ? it.bits << (64 - it.depth) : 0;
            tab.len[c] = (uint8_t) it.depth;
            tab.used.push_back(c);
            continue;
        }
        if (it.depth >= 64) return -2;
        st[sp++] = {r, it.depth + 1, (it.bits << 1) | 1};
        st[sp++] = {l, it.depth + 1, it.bits << 1};
    }
    tab.tree.clear();
    tab.tree_pending = true;
    if (!defer) finish_huffman_tree(tab);
    return 0;
}

int build_full(const uint32_t *freq, const uint32_t *first, int nsym, HuffTable &tab, bool defer) {
    tab.keys.clear();
    int s0 = 0;
    for (; s0 + 4 <= nsym; s0 += 4) {  // four counts at a time: most are zero
        uint64_t a, b;
        memcpy(&a, freq + s0, 8);
        memcpy(&b, freq + s0 + 2, 8);
        if (!(a | b)) continue;
        for (int s = s0; s < s0 + 4; s++)
            if (freq[s]) tab.keys.push_back((uint64_t) first[s] << 32 | (uint32_t) s);
    }
    for (int s = s0; s < nsym; s++)
        if (freq[s]) tab.keys.push_back((uint64_t) first[s] << 32 | (uint32_t) s);
    return build_present([&](int s) { return freq[s]; }, tab, defer);
}

int build_window(const uint32_t *wfreq, const uint32_t *wfirst, int lo, int nwin, uint32_t freq0, uint32_t first0,
                 HuffTable &tab, bool defer) {
    tab.keys.clear();
    if (freq0) tab.keys.push_back((uint64_t) first0 << 32);
    for (int i = 0; i < nwin; i++)
        if (wfreq[i]) tab.keys.push_back((uint64_t) wfirst[i] << 32 | (uint32_t) (lo + i));
    return build_present([&](int s) { return s == 0 ? freq0 : wfreq[s - lo]; }, tab, defer);
}

}  // namespace

int build_huffman(const uint32_t *freq, const uint32_t *first, int nsym, HuffTable &tab) {
    return build_full(freq, first, nsym, tab, false);
}

int build_huffman_codes(const uint32_t *freq, const uint32_t *first, int nsym, HuffTable &tab) {
    return build_full(freq, first, nsym, tab, true);
}

int build_huffman_window(const uint32_t *wfreq, const uint32_t *wfirst, int lo, int nwin, uint32_t freq0,
                         uint32_t first0, HuffTable &tab) {
    return build_window(wfreq, wfirst, lo, nwin, freq0, first0, tab, false);
}

int build_huffman_window_codes(const uint32_t *wfreq, const uint32_t *wfirst, int lo, int nwin, uint32_t freq0,
                               uint32_t first0, HuffTable &tab) {
    return build_window(wfreq, wfirst, lo, nwin, freq0, first0, tab, true);
}

// save (HUF:119-137)
void finish_huffman_tree(HuffTable &tab) {
    if (!tab.tree_pending) return;
    const unsigned nodeCount = (unsigned) tab.left.size();
    const int stateNum = tab.n_states + 1;  // HUF:551
    const size_t w = nodeCount <= 256 ? 1 : (nodeCount <= 65536 ? 2 : 4);
    tab.tree.resize(12 + 1 + (size_t) nodeCount * (2 * w + 5));
    uint8_t *o = tab.tree.data();
    memcpy(o, &tab.offset, 4);
    auto be32 = [](uint8_t *d, uint32_t v) {
        d[0] = (uint8_t) (v >> 24);
        d[1] = (uint8_t) (v >> 16);
        d[2] = (uint8_t) (v >> 8);
        d[3] = (uint8_t) v;
    };
    be32(o + 4, nodeCount);
    be32(o + 8, (uint32_t) (stateNum / 2));
    if (w == 1)
        serialise<uint8_t>(tab, nodeCount, o + 12);
    else if (w == 2)
        serialise<uint16_t>(tab, nodeCount, o + 12);
    else
        serialise<uint32_t>(tab, nodeCount, o + 12);
    tab.tree_pending = false;
}

// HuffmanEncoder::load + decode (HUF:236-296, 310-454)
void encode_huffman(const HuffTable &tab, const int32_t *s, size_t n, std::vector<uint8_t> &out) {
    uint64_t total = 0;
    for (size_t i = 0; i < n; i++) total += tab.len[s[i] - tab.offset];
    const uint64_t nbytes = (total + 7) / 8;
    const size_t at = out.size();
    out.resize(at + 8 + nbytes, 0);
    memcpy(out.data() + at, &nbytes, 8);
    uint8_t *w = out.data() + at + 8;
    // pending bits MSB-aligned in acc, fewer than 32 of them between codes;
    // codewords (MSB-aligned, zero below their length) go in at most 32
    // bits at a time and whole 32-bit groups leave big-endian
    uint64_t acc = 0;
    uint32_t nacc = 0;
    auto flush32 = [&]() {
        const uint32_t hi = (uint32_t) (acc >> 32);
        w[0] = (uint8_t) (hi >> 24);
        w[1] = (uint8_t) (hi >> 16);
        w[2] = (uint8_t) (hi >> 8);
        w[3] = (uint8_t) hi;
        w += 4;
        acc <<= 32;
        nacc -= 32;
    };
    for (size_t i = 0; i < n; i++) {
        const int st = s[i] - tab.offset;
        const uint32_t l = tab.len[st];
        const uint64_t c = tab.code[st];
        if (l <= 32) {
            acc |= c >> nacc;
            nacc += l;
        } else {
            acc |= (c & 0xFFFFFFFF00000000ull) >> nacc;
            nacc += 32;
            flush32();
            acc |= (c << 32) >> nacc;
            nacc += l - 32;
        }
        if (nacc >= 32) flush32();
    }
    while (nacc > 0) {
        *w++ = (uint8_t) (acc >> 56);
        acc <<= 8;
        nacc = nacc > 8 ? nacc - 8 : 0;
    }
}

size_t decode_huffman(const uint8_t *buf, size_t n, std::vector<int32_t> &out, size_t num) {
    if (n < 12) return 0;
    int32_t offset;
    memcpy(&offset, buf, 4);
    unsigned nc = ((unsigned) buf[4] << 24) | ((unsigned) buf[5] << 16) | ((unsigned) buf[6] << 8) | buf[7];
    if (nc == 0 || nc > (1u << 26)) return 0;
    size_t w = nc <= 256 ? 1 : (nc <= 65536 ? 2 : 4);
    size_t treeBytes = 1 + 2 * nc * w + nc * 4 + nc;
    if (12 + treeBytes + 8 > n) return 0;
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
    const uint8_t *L = tb + 1, *R = tb + 1 + nc * w, *C = tb + 1 + 2 * nc * w, *T = C + nc * 4;
    // flatten into child tables indexed by serialised slot (unpad_tree walks
    // the same links); slot 0 is the root
    std::vector<int32_t> left(nc, -1), right(nc, -1), sym(nc, 0);
    std::vector<uint8_t> leaf(nc, 0);
    for (unsigned i = 0; i < nc; i++) {
        int32_t c;
        memcpy(&c, C + 4 * i, 4);
        sym[i] = c;
        leaf[i] = T[i];
    }
    if (nc > 256) leaf[0] = 0;  // reference root is rebuilt as an inner node
    for (unsigned i = 0; i < nc; i++) {
        if (leaf[i]) continue;
        unsigned l = idx(L, i), r = idx(R, i);
        if (l >= nc || r >= nc) return 0;
        left[i] = l ? (int32_t) l : -1;
        right[i] = r ? (int32_t) r : -1;
    }
    uint64_t enc;
    memcpy(&enc, tb + treeBytes, 8);
    size_t used = 12 + treeBytes + 8;
    if (used + enc > n) return 0;
    const uint8_t *bytes = buf + used;
    out.resize(num);
    if (leaf[0]) {
        for (size_t i = 0; i < num; i++) out[i] = sym[0] + offset;
        return used + enc;
    }
    int nd = 0;
    size_t count = 0;
    uint64_t nbits = enc * 8;
    for (uint64_t i = 0; count < num; i++) {
        if (i >= nbits) return 0;
        int bit = (bytes[i >> 3] >> (7 - (i & 7))) & 1;
        nd = bit ? right[nd] : left[nd];
        if (nd < 0) return 0;
        if (leaf[nd]) {
            out[count++] = sym[nd] + offset;
            nd = 0;
        }
    }
    return used + enc;
}

}  // namespace faz
