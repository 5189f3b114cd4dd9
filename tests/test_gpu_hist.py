"""GPU histogram of 16-bit code streams (needs an MI355X).

Bar: the counts equal numpy's bincount and the first positions equal the
first index of every symbol (UINT32_MAX for absent ones), exactly — the
frequency loop of HuffmanEncoder::init (HUF:529-538) plus the insertion
order that fixes ska::unordered_map's iteration (HUF:455-566).  The cases
cover the lane-counter window around the radius, the LDS window, codes far
outside it, unpredictable (zero) codes, chunk-boundary sizes and a stream
that starts at an odd element.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _codes(n, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.integers(0, 65536, n).astype(np.uint16)
    if kind == "const":
        return np.full(n, 32768, np.uint16)
    if kind == "zeros":
        return np.zeros(n, np.uint16)
    sigma = float(kind)
    c = np.rint(32768 + sigma * rng.standard_normal(n))
    c = np.clip(c, 1, 65535)
    c[rng.random(n) < 1e-3] = 0  # unpredictable points
    return c.astype(np.uint16)


def _want(c):
    freq = np.bincount(c, minlength=65536).astype(np.uint32)
    first = np.full(65536, 0xFFFFFFFF, np.uint32)
    sym, idx = np.unique(c, return_index=True)
    first[sym] = idx
    return freq, first


CASES = [(n, kind) for n in (1, 7, 16383, 16384, 16385, 1000003) for kind in ("0.5", "3", "40")] + [
    (3 * 16384 + 5, "uniform"), (2_000_000, "const"), (100_000, "zeros"), (4_000_001, "600"),
    (4_000_000, "3000"), (16_777_216 + 3, "2"),
]


@pytest.mark.parametrize("i", range(len(CASES)))
@pytest.mark.parametrize("offset", [0, 1])
def test_hist_matches_numpy(engine, i, offset):
    n, kind = CASES[i]
    print(f"[gpu] hist n={n} kind={kind} offset={offset}", flush=True)
    c = _codes(n, kind, 700 + i)
    buf = torch.zeros(n + offset, dtype=torch.int16, device="cuda")
    buf[offset:] = torch.from_numpy(c.view(np.int16)).cuda()
    freq, first = engine.huffman_hist(buf[offset:].data_ptr(), n)
    want_f, want_p = _want(c)
    assert np.array_equal(freq, want_f)
    assert np.array_equal(first, want_p)


def test_hist_same_buffer_new_contents(engine):
    """A caller's buffer refilled in place is counted again."""
    n = 50_000
    buf = torch.empty(n, dtype=torch.int16, device="cuda")
    for seed, kind in ((1, "3"), (2, "40")):
        c = _codes(n, kind, seed)
        buf.copy_(torch.from_numpy(c.view(np.int16)))
        freq, first = engine.huffman_hist(buf.data_ptr(), n)
        want_f, want_p = _want(c)
        assert np.array_equal(freq, want_f) and np.array_equal(first, want_p)
