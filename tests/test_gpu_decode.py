"""GPU Huffman decode (faz_gpu_huffman_decode; needs an MI355X).

Bar: the symbols of every golden Huffman section the reference wrote (single
leaf, negative offsets, a deep Fibonacci tree, a 4-byte-index tree of 40000
symbols) and of large oracle-encoded streams whose codewords cross every
chunk boundary decode bit-exactly, with the section length the reference's
decoder consumes; short or out-of-range streams are errors, not faults.  The
decompressors that run on it are covered by the decompress tests of each
engine (512^3 included).
"""
import numpy as np
import pytest

import oracle
from golden_cases import HUFFMAN, load

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("name", HUFFMAN)
def test_golden_sections(engine, name):
    c = load(name)
    codes = c["codes"].astype(np.int32)
    print(f"[gpu] Huffman decode {name} n={codes.size}", flush=True)
    out, used = engine.huffman_decode(c["encoded"].tobytes(), codes.size)
    assert used == c["encoded"].size
    assert np.array_equal(out.cpu().numpy(), codes)


def _codes(kind, n, seed):
    r = np.random.default_rng(seed)
    if kind == "peaked":  # quantization codes around the radius
        return (32768 + np.round(r.laplace(0, 1.5, n))).astype(np.int32)
    if kind == "wide":  # every 16-bit symbol: codewords longer than the table
        return r.integers(0, 65536, n, dtype=np.int32)
    if kind == "zipf":  # a deep tree: long codewords for rare symbols
        return np.minimum(r.zipf(1.3, n), 200000).astype(np.int32) - 1
    if kind == "two":  # one-bit codewords
        return r.integers(0, 2, n, dtype=np.int32) + 7
    raise ValueError(kind)


CASES = [("peaked", 3_000_001, 1), ("wide", 1_000_003, 2), ("zipf", 2_000_000, 3), ("two", 4096 * 8, 4),
         ("two", 4096 * 8 + 1, 5), ("peaked", 1, 6), ("peaked", 4097, 7)]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_random_streams(engine, i):
    kind, n, seed = CASES[i]
    codes = _codes(kind, n, seed)
    sec = oracle.huffman_encode(codes)
    print(f"[gpu] Huffman decode {kind} n={n} bytes={len(sec)}", flush=True)
    out, used = engine.huffman_decode(sec + b"\0" * 16, n)
    assert used == len(sec)
    assert np.array_equal(out.cpu().numpy(), codes)
    # a prefix of the symbols decodes from the same section
    k = max(1, n // 3)
    out, _ = engine.huffman_decode(sec, k)
    assert np.array_equal(out.cpu().numpy(), codes[:k])


def test_unaligned_output(engine):
    """Output that starts 4 bytes past an 8-byte boundary: the first symbol
    is stored alone, the rest in 8-byte stores; nothing outside is touched."""
    codes = _codes("peaked", 100003, 11)
    sec = oracle.huffman_encode(codes)
    buf = torch.full((codes.size + 3,), -7, dtype=torch.int32, device="cuda")
    print("[gpu] Huffman decode into an unaligned slice", flush=True)
    out, _ = engine.huffman_decode(sec, codes.size, out=buf[1:1 + codes.size])
    h = buf.cpu().numpy()
    assert h[0] == -7 and h[-2] == -7 and h[-1] == -7
    assert np.array_equal(h[1:1 + codes.size], codes)


def test_errors(engine):
    import faz_amd
    codes = _codes("peaked", 100000, 9)
    sec = oracle.huffman_encode(codes)
    print("[gpu] Huffman decode error paths", flush=True)
    with pytest.raises(faz_amd.FazError):  # more symbols than the stream holds
        engine.huffman_decode(sec, 2 * codes.size)
    with pytest.raises(faz_amd.FazError):  # symbols outside [0, limit)
        engine.huffman_decode(sec, codes.size, limit=32768)
    with pytest.raises(faz_amd.FazError):  # section cut short
        engine.huffman_decode(sec[:len(sec) // 2], codes.size)
    out, _ = engine.huffman_decode(sec, codes.size, limit=65536)
    assert np.array_equal(out.cpu().numpy(), codes)
