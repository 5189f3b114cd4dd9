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

struct Node {
    int left = -1, right = -1;
    uint64_t freq = 0;
    int c = 0;
    bool leaf = false;
};

// HuffmanEncoder's 1-based binary heap (qinsert/qremove, HUF:455-485).  The
// entries carry their node's frequency, so the sift loops compare without
// reaching into the node pool; the operations are the reference's.  The
// sift-down moves a hole instead of swapping (the same final array) and
// picks the smaller child arithmetically: that choice is data-dependent and
// mispredicts as a branch (about half the cost of a build of a few thousand
// symbols).
class Heap {
public:
    explicit Heap(size_t cap) { q_.resize(cap + 3); }
    void insert(int n, uint64_t f) {
        int i = end_++, j;
        if ((int) q_.size() <= i + 1) q_.resize(2 * (size_t) i + 3);
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
            // q_[end_] (the moved entry's old slot) stays readable: no bounds branch
            l += (int) ((l + 1 < end_) & (q_[l + 1].f < q_[l].f));
            if (!(x.f > q_[l].f)) break;
            q_[i] = q_[l];
            i = l;
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

template<class T1>
void serialise(const std::vector<Node> &pool, int root, unsigned nodeCount, std::vector<uint8_t> &out) {
    // pre-order numbering (pad_tree, HUF:569-584); internal nodes carry C = 0
    std::vector<T1> L(nodeCount, 0), R(nodeCount, 0);
    std::vector<int32_t> C(nodeCount, 0);
    std::vector<uint8_t> t(nodeCount, 0);
    unsigned counter = 0;
    struct Frame { int node; unsigned slot; int stage; };
    std::vector<Frame> st;
    st.push_back({root, 0, 0});
    while (!st.empty()) {
        Frame &f = st.back();
        const Node &nd = pool[f.node];
        if (f.stage == 0) {
            C[f.slot] = nd.leaf ? nd.c : 0;
            t[f.slot] = nd.leaf ? 1 : 0;
            f.stage = 1;
            if (nd.left >= 0) {
                counter++;
                L[f.slot] = (T1) counter;
                st.push_back({nd.left, counter, 0});
                continue;
            }
        }
        if (f.stage == 1) {
            f.stage = 2;
            if (nd.right >= 0) {
                counter++;
                R[f.slot] = (T1) counter;
                st.push_back({nd.right, counter, 0});
                continue;
            }
        }
        st.pop_back();
    }
    out.push_back(0);  // endianness marker: little-endian
    auto app = [&out](const void *p, size_t n) {
        const uint8_t *b = (const uint8_t *) p;
        out.insert(out.end(), b, b + n);
    };
    app(L.data(), nodeCount * sizeof(T1));
    app(R.data(), nodeCount * sizeof(T1));
    app(C.data(), nodeCount * sizeof(int32_t));
    app(t.data(), nodeCount);
}

void put_be32(std::vector<uint8_t> &out, uint32_t v) {
    out.push_back((uint8_t) (v >> 24));
    out.push_back((uint8_t) (v >> 16));
    out.push_back((uint8_t) (v >> 8));
    out.push_back((uint8_t) v);
}

}  // namespace

namespace {

// The tree of the present symbols: (first position, symbol) pairs and a
// count lookup by symbol.
template<class FreqOf>
int build_present(std::vector<std::pair<uint32_t, int>> &present, FreqOf freq_of, HuffTable &tab);

}  // namespace

int build_huffman(const uint32_t *freq, const uint32_t *first, int nsym, HuffTable &tab) {
    std::vector<std::pair<uint32_t, int>> present;
    present.reserve(1024);
    int s0 = 0;
    for (; s0 + 4 <= nsym; s0 += 4) {  // four counts at a time: most are zero
        uint64_t a, b;
        memcpy(&a, freq + s0, 8);
        memcpy(&b, freq + s0 + 2, 8);
        if (!(a | b)) continue;
        for (int s = s0; s < s0 + 4; s++)
            if (freq[s]) present.push_back({first[s], s});
    }
    for (int s = s0; s < nsym; s++)
        if (freq[s]) present.push_back({first[s], s});
    return build_present(present, [&](int s) { return freq[s]; }, tab);
}

int build_huffman_window(const uint32_t *wfreq, const uint32_t *wfirst, int lo, int nwin, uint32_t freq0,
                         uint32_t first0, HuffTable &tab) {
    std::vector<std::pair<uint32_t, int>> present;
    present.reserve(1024);
    if (freq0) present.push_back({first0, 0});
    for (int i = 0; i < nwin; i++)
        if (wfreq[i]) present.push_back({wfirst[i], lo + i});
    return build_present(present, [&](int s) { return s == 0 ? freq0 : wfreq[s - lo]; }, tab);
}

namespace {

template<class FreqOf>
int build_present(std::vector<std::pair<uint32_t, int>> &present, FreqOf freq_of, HuffTable &tab) {
    if (present.empty()) return -1;
    std::sort(present.begin(), present.end());
    int mn = present[0].second, mx = present[0].second;
    for (auto &p: present) {
        mn = std::min(mn, p.second);
        mx = std::max(mx, p.second);
    }
    // a table of the same span keeps its arrays: only last build's codewords
    // are cleared (the outlier code 0 makes the span tens of thousands of
    // states wide for a few hundred present symbols)
    const bool reuse = tab.offset == mn && tab.n_states == mx - mn + 1 && (int) tab.code.size() == tab.n_states &&
                       (int) tab.len.size() == tab.n_states;
    tab.offset = mn;
    tab.n_states = mx - mn + 1;
    const int stateNum = mx - mn + 2;  // HUF:551

    SkaOrder ska(present.size());
    for (auto &p: present) ska.insert(p.second);
    std::vector<int> order = ska.iteration_order();

    std::vector<Node> pool;
    pool.reserve(2 * present.size());
    Heap heap(2 * present.size());
    for (int key: order) {
        Node n;
        n.leaf = true;
        n.freq = freq_of(key);
        n.c = key - mn;
        pool.push_back(n);
        heap.insert((int) pool.size() - 1, n.freq);
    }
    // qinsert(new_node(0, 0, qremove(), qremove())) in the reference's g++
    // build evaluates the arguments right to left: the first node removed is
    // the right child.
    while (heap.size() > 1) {
        int right = heap.remove();
        int left = heap.remove();
        Node n;
        n.left = left;
        n.right = right;
        n.freq = pool[left].freq + pool[right].freq;
        pool.push_back(n);
        heap.insert((int) pool.size() - 1, n.freq);
    }
    const int root = heap.top();

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
    struct Item { int node; int depth; uint64_t bits; };
    std::vector<Item> st;
    st.push_back({root, 0, 0});
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        const Node &nd = pool[it.node];
        if (nd.leaf) {
            if (it.depth > 64) return -2;  // codes beyond 64 bits need > 2.7e13 symbols
            tab.code[nd.c] = it.depth ? it.bits << (64 - it.depth) : 0;
            tab.len[nd.c] = (uint8_t) it.depth;
            tab.used.push_back(nd.c);
            continue;
        }
        if (it.depth >= 64) return -2;
        st.push_back({nd.right, it.depth + 1, (it.bits << 1) | 1});
        st.push_back({nd.left, it.depth + 1, it.bits << 1});
    }

    // save (HUF:119-137)
    const unsigned nodeCount = (unsigned) (2 * present.size() - 1);
    tab.tree.clear();
    uint8_t b4[4];
    memcpy(b4, &mn, 4);
    tab.tree.insert(tab.tree.end(), b4, b4 + 4);
    put_be32(tab.tree, nodeCount);
    put_be32(tab.tree, (uint32_t) (stateNum / 2));
    if (nodeCount <= 256)
        serialise<uint8_t>(pool, root, nodeCount, tab.tree);
    else if (nodeCount <= 65536)
        serialise<uint16_t>(pool, root, nodeCount, tab.tree);
    else
        serialise<uint32_t>(pool, root, nodeCount, tab.tree);
    return 0;
}

}  // namespace

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
