"""Pins of the oracle by what the mathematics fixes (CPU only).

These hold for any correct implementation and fail for plausible slips
(a wrong stencil weight or sign, a dropped term in the quantizer, a transposed
axis, a non-Huffman tree):
  * every interpolation stencil is exact on linear data, so a linear integer
    field with eb = 1/2 reconstructs exactly;
  * quantizer semantics over a dense (pred, value) grid;
  * the code stream visits every point exactly once;
  * |x - x'| <= eb in FP64 and decompress(compress) reproduces the recon;
  * the Huffman code satisfies Kraft's equality, has the optimal weighted
    length of a textbook Huffman construction, and round-trips.
"""
import heapq
import itertools

import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("dims,algo,direction,max_step", [
    ((33, 29, 40), 1, 0, 0), ((33, 29, 40), 1, 5, 32), ((17, 70), 1, 1, 16), ((70, 17), 0, 0, 0),
    ((200,), 1, 0, 0), ((9, 11, 13, 7), 1, 13, 0), ((65, 65, 3), 0, 3, 32)])
def test_stencils_exact_on_linear_fields(dims, algo, direction, max_step):
    grids = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    coef = [3, -2, 5, 1][:len(dims)]
    x = (7 + sum(c * g for c, g in zip(coef, grids))).astype(np.float32)
    p = oracle.make_params(dims, 0.5, algo=algo, direction=direction, max_step=max_step)
    r = oracle.interp_compress(x, p)
    assert np.array_equal(r["recon"], x.ravel())
    assert r["unpred"].size == (r["codes"] == 0).sum()


def test_quantizer_semantics_dense_grid():
    # dims (2,): point 0 is quantized against 0, point 1 against point 0's
    # reconstruction (a 2-point line uses the left neighbour)
    eb = 2.0 ** -10
    radius = 64
    p0 = np.float32(2 * 37 * eb)  # exact multiple of 2*eb: reconstructs exactly
    vals = p0 + np.linspace(-300 * eb, 300 * eb, 4001).astype(np.float64)
    vals = np.concatenate([vals, [p0 + 127.4 * eb, p0 - 127.6 * eb, p0 + 5e6 * eb]]).astype(np.float32)
    for v in vals:
        r = oracle.interp_compress(np.array([p0, v], np.float32), oracle.make_params((2,), eb, radius=radius))
        c0, c1 = r["codes"]
        assert r["recon"][0] == p0
        if c1 == 0:
            assert r["recon"][1] == v
            # unpredictable only when out of range or when the bin centre misses
            q = abs(float(v) - float(p0)) / eb
            assert q >= 2 * radius - 1 or abs(np.float32(p0 + 2 * round(q / 2) * eb) - v) > eb
        else:
            dec = np.float32(np.float64(p0) + 2 * (int(c1) - radius) * eb)
            assert r["recon"][1] == dec
            assert abs(np.float64(dec) - np.float64(v)) <= eb
            assert abs((int(c1) - radius) - (float(v) - float(p0)) / (2 * eb)) <= 0.5 + 1e-6


@pytest.mark.parametrize("dims,max_step,bs", [((17, 17, 17), 0, 32), ((33, 20, 9), 8, 32), ((40, 41), 32, 16),
                                              ((100,), 0, 32), ((5, 6, 7, 8), 0, 4), ((65, 1, 33), 32, 32)])
def test_every_point_coded_once(dims, max_step, bs):
    p = oracle.make_params(dims, 1e-3, max_step=max_step, block_size=bs)
    seq = oracle.seq_index(p)
    assert np.array_equal(np.sort(seq), np.arange(int(np.prod(dims))))


@pytest.mark.parametrize("seed", range(4))
def test_error_bound_and_round_trip(seed):
    rng = np.random.default_rng(seed)
    dims = tuple(int(v) for v in rng.integers(5, 40, 3))
    x = np.cumsum(np.cumsum(rng.standard_normal(dims), 0), 2).astype(np.float32)
    eb = float(rng.choice([1e-1, 1e-2, 1e-3])) * float(x.max() - x.min())
    alpha, beta = [(-1, -1), (1.5, 3.0), (0.5, 0.2)][seed % 3]
    p = oracle.make_params(dims, eb, algo=seed % 2, direction=int(rng.integers(0, 6)), alpha=alpha, beta=beta,
                           max_step=[0, 16, 32][seed % 3])
    r = oracle.interp_compress(x, p)
    err = np.abs(r["recon"].astype(np.float64) - x.ravel().astype(np.float64))
    assert err.max() <= eb
    raw = oracle.interp_raw(x, p)
    dec = oracle.interp_decompress_raw(raw, dims)
    assert np.array_equal(dec.view(np.uint32), r["recon"].view(np.uint32))


def _textbook_huffman_cost(freqs):
    if len(freqs) == 1:
        return 0
    h = list(freqs)
    heapq.heapify(h)
    cost = 0
    while len(h) > 1:
        a = heapq.heappop(h)
        b = heapq.heappop(h)
        cost += a + b
        heapq.heappush(h, a + b)
    return cost


def _parse_tree_lengths(encoded):
    """Leaf depths from the serialised tree (pre-order L/R/C/t arrays)."""
    nc = int.from_bytes(encoded[4:8], "big")
    w = 1 if nc <= 256 else (2 if nc <= 65536 else 4)
    base = 12
    L = np.frombuffer(encoded[base + 1:base + 1 + nc * w], {1: np.uint8, 2: np.uint16, 4: np.uint32}[w])
    R = np.frombuffer(encoded[base + 1 + nc * w:base + 1 + 2 * nc * w], L.dtype)
    C = np.frombuffer(encoded[base + 1 + 2 * nc * w:base + 1 + 2 * nc * w + 4 * nc], np.int32)
    T = np.frombuffer(encoded[base + 1 + 2 * nc * w + 4 * nc:base + 1 + 2 * nc * w + 5 * nc], np.uint8)
    depth = {}
    stack = [(0, 0)]
    while stack:
        i, d = stack.pop()
        if T[i]:
            depth[int(C[i])] = d
            continue
        if L[i]:
            stack.append((int(L[i]), d + 1))
        if R[i]:
            stack.append((int(R[i]), d + 1))
    return depth, base + 1 + 2 * nc * w + 5 * nc


@pytest.mark.parametrize("seed,nsym,n", [(0, 2, 1000), (1, 50, 5000), (2, 300, 40000), (3, 3000, 100000)])
def test_huffman_kraft_optimal_round_trip(seed, nsym, n):
    rng = np.random.default_rng(seed)
    w = rng.pareto(1.5, nsym) + 1e-3
    codes = (32768 - nsym // 2 + rng.choice(nsym, n, p=w / w.sum())).astype(np.int32)
    enc = oracle.huffman_encode(codes)
    depth, end = _parse_tree_lengths(enc)
    offset = int(np.frombuffer(enc[:4], np.int32)[0])
    syms, freq = np.unique(codes, return_counts=True)
    assert offset == syms.min()
    lens = np.array([depth[int(s) - offset] for s in syms])
    if syms.size > 1:
        assert sum(2.0 ** -l for l in lens) == 1.0
    assert int((lens * freq).sum()) == _textbook_huffman_cost(freq.tolist())
    nbytes = int(np.frombuffer(enc[end:end + 8], np.uint64)[0])
    assert nbytes == (int((lens * freq).sum()) + 7) // 8
    dec, used = oracle.huffman_decode(enc, codes.size)
    assert used == len(enc) and np.array_equal(dec, codes)
