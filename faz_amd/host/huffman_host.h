// Host Huffman stage (tree, code table, serialisation, decoding).
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <vector>

namespace faz {

struct HuffTable {
    int32_t offset = 0;            // smallest symbol
    int32_t n_states = 0;          // max - offset + 1
    std::vector<uint64_t> code;    // MSB-aligned codewords by symbol - offset
    std::vector<uint8_t> len;      // codeword lengths
    std::vector<uint8_t> tree;     // HuffmanEncoder::save bytes
    std::vector<int32_t> used;     // indices with a codeword (what a rebuild of the same span clears)
    // nodes of the last build (children -1 for leaves, symbol - offset of a
    // leaf), kept for a deferred serialisation and reused across builds
    std::vector<int32_t> left, right, sym;
    std::vector<uint64_t> keys, scratch;  // sort buffers of the present symbols
    int32_t root = -1;
    bool tree_pending = false;     // nodes built, tree bytes not yet written
};

// Builds the reference Huffman tree from counts and first positions of
// symbols 0..nsym-1.  Returns 0, or < 0 for an empty histogram / codes longer
// than 64 bits.
int build_huffman(const uint32_t *freq, const uint32_t *first, int nsym, HuffTable &tab);
// The same when every present symbol but 0 lies in [lo, lo + nwin): counts
// and first positions of that window, and of symbol 0.
int build_huffman_window(const uint32_t *wfreq, const uint32_t *wfirst, int lo, int nwin, uint32_t freq0,
                         uint32_t first0, HuffTable &tab);
// The same with the tree bytes left for finish_huffman_tree (the code table
// is complete): a caller can enqueue the encode first.
int build_huffman_window_codes(const uint32_t *wfreq, const uint32_t *wfirst, int lo, int nwin, uint32_t freq0,
                               uint32_t first0, HuffTable &tab);
int build_huffman_codes(const uint32_t *freq, const uint32_t *first, int nsym, HuffTable &tab);
// Writes tab.tree (HuffmanEncoder::save) after a *_codes build; no-op when
// the bytes are already there.
void finish_huffman_tree(HuffTable &tab);

// HuffmanEncoder::encode (HuffmanEncoder.hpp:150-227) on the host: appends
// the size_t byte count and the MSB-first bitstream of s[0..n) to out.  Every
// symbol must have a code in tab.
void encode_huffman(const HuffTable &tab, const int32_t *s, size_t n, std::vector<uint8_t> &out);

// Loads a saved tree followed by a size_t-prefixed bitstream and decodes num
// symbols.  Returns the bytes consumed, 0 on malformed input.
size_t decode_huffman(const uint8_t *buf, size_t n, std::vector<int32_t> &out, size_t num);

}  // namespace faz
