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
};

// Builds the reference Huffman tree from counts and first positions of
// symbols 0..nsym-1.  Returns 0, or < 0 for an empty histogram / codes longer
// than 64 bits.
int build_huffman(const uint32_t *freq, const uint32_t *first, int nsym, HuffTable &tab);
// The same when every present symbol but 0 lies in [lo, lo + nwin): counts
// and first positions of that window, and of symbol 0.
int build_huffman_window(const uint32_t *wfreq, const uint32_t *wfirst, int lo, int nwin, uint32_t freq0,
                         uint32_t first0, HuffTable &tab);

// HuffmanEncoder::encode (HuffmanEncoder.hpp:150-227) on the host: appends
// the size_t byte count and the MSB-first bitstream of s[0..n) to out.  Every
// symbol must have a code in tab.
void encode_huffman(const HuffTable &tab, const int32_t *s, size_t n, std::vector<uint8_t> &out);

// Loads a saved tree followed by a size_t-prefixed bitstream and decodes num
// symbols.  Returns the bytes consumed, 0 on malformed input.
size_t decode_huffman(const uint8_t *buf, size_t n, std::vector<int32_t> &out, size_t num);

}  // namespace faz
