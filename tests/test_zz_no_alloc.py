"""The allocation-free Huffman encode (faz_gpu_huffman_encode_noalloc) against
the allocating entry point and the oracle, through the C-ABI (needs an
MI355X).  Two calls in a row on the shared static scratch, the second with a
smaller input, so stale scratch contents of the first call must not leak."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _codes(shape, seed, rel):
    import faz_amd
    r = np.random.default_rng(seed)
    x = r.standard_normal(shape)
    for ax in range(len(shape)):
        x = np.cumsum(x, axis=ax)
    x = x.astype(np.float32)
    eb = rel * float(np.float32(x.max() - x.min()))
    p = oracle.make_params(shape, eb, algo=1, max_step=32)
    return oracle.interp_compress(x, p)["codes"].astype(np.uint16), faz_amd


def _table(codes):
    import faz_amd
    freq = np.zeros(65536, np.uint32)
    first = np.full(65536, 0xFFFFFFFF, np.uint32)
    syms, idx, cnt = np.unique(codes, return_index=True, return_counts=True)
    freq[syms] = cnt
    first[syms] = idx
    return faz_amd.huffman_build(freq, first)


def _encode_both(engine, codes):
    offset, code, ln, tree = _table(codes)
    d = torch.from_numpy(codes.view(np.int16).copy()).cuda()
    bits = int(ln[codes.astype(np.int64) - offset].astype(np.int64).sum())
    cap = (bits + 31) // 32 + 16
    a = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    b = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    na = engine.huffman_encode(d.data_ptr(), codes.size, offset, code, ln, a)
    nb = engine.huffman_encode_noalloc(d.data_ptr(), codes.size, offset, code, ln, b)
    assert na == nb == (bits + 7) // 8
    ba = a.cpu().numpy().view(np.uint8)[:na].tobytes()
    bb = b.cpu().numpy().view(np.uint8)[:nb].tobytes()
    assert ba == bb
    # the oracle's section: tree, size_t byte count, MSB-first bits
    sec = oracle.huffman_encode(codes.astype(np.int32))
    assert sec[:len(tree)] == tree
    assert int(np.frombuffer(sec[len(tree):len(tree) + 8], np.uint64)[0]) == nb
    assert sec[len(tree) + 8:len(tree) + 8 + nb] == bb
    return nb


def test_noalloc_encode_matches_allocating_entry_point(engine):
    import faz_amd
    assert faz_amd.lib().faz_gpu_noalloc_scratch_bytes() > 0
    big, _ = _codes((96, 100, 110), 7, 1e-4)
    print(f"[gpu] no-alloc encode, call 1: {big.size} codes", flush=True)
    n1 = _encode_both(engine, big)
    small, _ = _codes((40, 33, 50), 8, 1e-3)
    print(f"[gpu] no-alloc encode, call 2: {small.size} codes", flush=True)
    n2 = _encode_both(engine, small)
    assert n2 < n1
    # a ragged size that ends inside an encode chunk, after the larger calls
    print("[gpu] no-alloc encode, call 3: 4097 codes", flush=True)
    _encode_both(engine, big[:4097].copy())


def test_noalloc_encode_rejects_sizes_beyond_its_scratch(engine):
    import faz_amd
    codes, _ = _codes((20, 20, 20), 3, 1e-3)
    offset, code, ln, _ = _table(codes)
    d = torch.from_numpy(codes.view(np.int16).copy()).cuda()
    out = torch.zeros(1024, dtype=torch.int32, device="cuda")
    with pytest.raises(faz_amd.FazError, match="at most"):
        engine.huffman_encode_noalloc(d.data_ptr(), (1 << 30) + 1, offset, code, ln, out)
