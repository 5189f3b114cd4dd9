"""Host side of the product, no GPU needed: the C-ABI library loads and
exports every entry point of include/faz_gpu.h, the host Huffman stage
(faz_huffman_build) reproduces the reference tree bytes and code table, and
the faz command line runs."""
import os
import re
import subprocess

import numpy as np
import pytest

import faz_amd
from golden_cases import HUFFMAN, load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "faz_gpu.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    declared = set(re.findall(r"\b(faz_\w+)\s*\(", header))
    lib = faz_amd.lib()
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(faz_amd.EXPORTS)


def _freq_first(codes):
    freq = np.zeros(65536, np.uint32)
    first = np.full(65536, 0xFFFFFFFF, np.uint32)
    syms, idx, cnt = np.unique(codes, return_index=True, return_counts=True)
    freq[syms] = cnt
    first[syms] = idx
    return freq, first


def _pack(codes, offset, code, ln):
    """MSB-first packing of the code table (independent of the encoder kernels)."""
    st = codes.astype(np.int64) - offset
    L = ln[st].astype(np.int64)
    bits = []
    for s, l in zip(st, L):
        c = int(code[s])
        bits.extend(((c >> (63 - k)) & 1) for k in range(int(l)))
    b = np.array(bits + [0] * (-len(bits) % 8), np.uint8)
    return np.packbits(b).tobytes() if b.size else b""


@pytest.mark.parametrize("name", [h for h in HUFFMAN if h != "h_negative"])
def test_host_huffman_matches_reference(name):
    c = load(name)
    codes = c["codes"]
    ref = c["encoded"].tobytes()
    freq, first = _freq_first(codes)
    offset, code, ln, tree = faz_amd.huffman_build(freq, first)
    assert tree == ref[:len(tree)]
    nbytes = int(np.frombuffer(ref[len(tree):len(tree) + 8], np.uint64)[0])
    if codes.size <= 60000:
        assert _pack(codes, offset, code, ln) == ref[len(tree) + 8:len(tree) + 8 + nbytes]
    total_bits = int(ln[codes.astype(np.int64) - offset].astype(np.int64).sum())
    assert (total_bits + 7) // 8 == nbytes


def test_host_huffman_rejects_empty_histogram():
    with pytest.raises(faz_amd.FazError):
        faz_amd.huffman_build(np.zeros(65536, np.uint32), np.full(65536, 0xFFFFFFFF, np.uint32))


def test_cli_usage_runs_without_gpu():
    r = subprocess.run([faz_amd.CLI_PATH, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Usage: faz" in r.stdout


FASTDIV_CHECK = r'''
#include <cstdio>
#include <cstdint>
#include <random>
#include "faz_internal.h"
int main() {
    std::mt19937_64 g(7);
    long bad = 0;
    auto check = [&](uint32_t d, uint32_t x) {
        if (faz::FastDiv(d).div(x) != x / d) bad++;
    };
    for (uint32_t d = 1; d <= 20000; d++)
        for (uint32_t x : {0u, 1u, d - 1, d, d + 1, 2 * d - 1, 2 * d, 0xFFFFFFFFu, 0x80000000u, 0x7FFFFFFFu,
                           (uint32_t) g(), (uint32_t) g()})
            check(d, x);
    for (int k = 0; k < 200000; k++) {
        uint32_t d = (uint32_t) g();
        check(d ? d : 1, (uint32_t) g());
    }
    for (uint32_t s = 0; s < 32; s++) {
        check(1u << s, 0xFFFFFFFFu);
        check((1u << s) + 1, 0xFFFFFFFFu);
        check(s ? (1u << s) - 1 : 1, 0xFFFFFFFFu);
    }
    printf("%ld\n", bad);
    return bad != 0;
}
'''


def test_fastdiv_is_exact(tmp_path):
    """The multiply-shift division the pass kernels use for lattice extents and
    block edges equals integer division (divisors 1..20000 at their edge
    values, random divisors, powers of two and their neighbours)."""
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "fastdiv.cpp"
    src.write_text(FASTDIV_CHECK)
    exe = tmp_path / "fastdiv"
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(here, "faz_amd", "csrc"), str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "0"


# --- malformed input, host only ------------------------------------------------

def _section(left, right, leafs, syms, enc=b"", nbits_bytes=None):
    """A Huffman section in the saved layout (HuffmanEncoder.hpp:119-137,
    569-633): int offset, big-endian node count and stateNum / 2, then the
    endian byte, L[], R[] (u8 for <= 256 nodes), C[] (int32), t[] and the
    size_t byte count of the bits."""
    import struct
    nc = len(left)
    assert nc <= 256
    b = struct.pack("<i", 0) + struct.pack(">I", nc) + struct.pack(">I", 1)
    b += b"\0" + bytes(left) + bytes(right) + b"".join(struct.pack("<i", c) for c in syms) + bytes(leafs)
    b += struct.pack("<Q", len(enc) if nbits_bytes is None else nbits_bytes) + enc
    return b


def test_huffman_check_accepts_reference_sections():
    for name in HUFFMAN:
        sec = load(name)["encoded"].tobytes()
        nodes, used = faz_amd.huffman_check_section(sec)
        assert used == len(sec) and nodes >= 1, name


def test_huffman_check_rejects_dag_cycle_deep_and_truncated_trees():
    # a valid three-node tree: root with two leaves
    ok = _section([1, 0, 0], [2, 0, 0], [0, 1, 1], [0, 5, 7], enc=b"\x80")
    assert faz_amd.huffman_check_section(ok) == (3, len(ok))
    # DAG: inner nodes i share one child (left = right = i + 1) for 60 levels;
    # a walk over all paths would take 2^60 steps
    n = 62
    L = [i + 1 for i in range(n - 1)] + [0]
    leafs = [0] * (n - 1) + [1]
    dag = _section(L, L, leafs, [0] * n)
    with pytest.raises(faz_amd.FazError) as e:
        faz_amd.huffman_check_section(dag)
    assert e.value.status == -5
    # cycle: 1 -> 2 -> 1
    cyc = _section([1, 2, 1, 0], [3, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 4])
    with pytest.raises(faz_amd.FazError):
        faz_amd.huffman_check_section(cyc)
    # a proper tree 70 levels deep (each inner node: left leaf, right inner)
    depth = 70
    nc = 2 * depth + 1
    Ld, Rd, lf = [0] * nc, [0] * nc, [0] * nc
    for i in range(depth):
        inner, leaf_i, nxt = 2 * i, 2 * i + 1, 2 * i + 2
        Ld[inner], Rd[inner] = leaf_i, nxt
        lf[leaf_i] = 1
    lf[nc - 1] = 1
    with pytest.raises(faz_amd.FazError):
        faz_amd.huffman_check_section(_section(Ld, Rd, lf, list(range(nc))))
    # bitstream longer than the buffer; child index out of range
    with pytest.raises(faz_amd.FazError):
        faz_amd.huffman_check_section(_section([1, 0, 0], [2, 0, 0], [0, 1, 1], [0, 5, 7], nbits_bytes=1000))
    with pytest.raises(faz_amd.FazError):
        faz_amd.huffman_check_section(_section([1, 0, 0], [9, 0, 0], [0, 1, 1], [0, 5, 7]))
    with pytest.raises(faz_amd.FazError):
        faz_amd.huffman_check_section(ok[:10])


def _trailer_offset(stream):
    n = int(np.frombuffer(stream[-4:], np.int32)[0])
    return len(stream) - 4 - n


def test_stream_info_reads_reference_trailers():
    from golden_cases import INTERP
    for name in INTERP:
        c = load(name)
        s = c["stream"].tobytes()
        N, dims, algo = faz_amd.stream_info(s)
        assert dims == tuple(c["params"]["dims"]) and N == len(dims) and algo == 2, name


def test_stream_info_rejects_truncated_and_corrupt_trailers():
    import struct
    from golden_cases import INTERP
    s = bytearray(load(INTERP[0])["stream"].tobytes())
    at = _trailer_offset(s)
    conf_size = len(s) - 4 - at

    def bad(b):
        with pytest.raises(faz_amd.FazError) as e:
            faz_amd.stream_info(bytes(b))
        assert e.value.status == -5

    with pytest.raises(faz_amd.FazError):
        faz_amd.stream_info(b"")
    bad(s[:3])
    # trailer size beyond the stream, or a trailer cut short
    bad(s[:-4] + struct.pack("<i", len(s)))
    for keep in (1, 8, conf_size // 2):
        t = s[at:at + keep]
        bad(s[:at] + t + struct.pack("<i", len(t)))
    # N outside 1..4 (a signed char in the reference)
    for v in (0, 5, 0x80, 0xFF):
        b = bytearray(s)
        b[at] = v
        bad(b)
    # a huge dimension count would ask for an impossible allocation in the reference
    b = bytearray(s)
    b[at] = 0x7F
    bad(b)
