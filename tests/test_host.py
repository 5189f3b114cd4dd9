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
