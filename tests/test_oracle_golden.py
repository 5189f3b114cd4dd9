"""The oracle against the reference engine's golden vectors (CPU only).

Every fixture was produced by the reference classes compiled from their own
sources (oracle/_ref/ref_harness, see tests/golden/make_golden.py); the oracle
must reproduce codes, unpredictable values, reconstruction, the pre-zstd engine
buffer and the Config trailer bit for bit, and must decode the reference's
own zstd stream.
"""
import numpy as np
import pytest

import oracle
from golden_cases import (CTEST, CTEST_LORENZO, GENERAL, HUFFMAN, INTERP, LORENZO, REGRESSION, SZ3, TUNING, ctest_params,
                          gf_kwargs,
                          load, lr_kwargs, lr_raw_equal, param_kwargs, trailer)


def _params(case):
    return oracle.make_params(**param_kwargs(case))


@pytest.mark.parametrize("name", INTERP)
def test_interp_codes_unpred_recon(name):
    c = load(name)
    r = oracle.interp_compress(c["input"], _params(c))
    assert np.array_equal(r["codes"], c["codes"])
    assert np.array_equal(r["unpred"].view(np.uint32), c["unpred"].view(np.uint32))
    assert np.array_equal(r["recon"].view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", INTERP)
def test_interp_raw_buffer(name):
    c = load(name)
    raw = oracle.interp_raw(c["input"], _params(c))
    assert raw == c["raw"].tobytes()


@pytest.mark.parametrize("name", INTERP)
def test_decode_reference_buffers_and_stream(name):
    c = load(name)
    dims = c["params"]["dims"]
    rec = oracle.interp_decompress_raw(c["raw"].tobytes(), dims)
    assert np.array_equal(rec.view(np.uint32), c["recon"].view(np.uint32))
    rec2 = oracle.decompress_stream(c["stream"].tobytes(), int(np.prod(dims)))
    assert np.array_equal(rec2.view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", INTERP)
def test_stream_trailer(name):
    """Config trailer (Config.hpp:240-309) equals the reference's byte for byte."""
    c = load(name)
    p = c["params"]
    kw = param_kwargs(c)
    rel = p.get("mode", "abs") == "rel"
    s = oracle.compress_stream(c["input"], oracle.make_params(**kw), p["eb"], rel=rel)
    ref = c["stream"].tobytes()
    conf_ref = int(np.frombuffer(ref[-4:], np.int32)[0])
    conf_ours = int(np.frombuffer(s[-4:], np.int32)[0])
    assert conf_ref == conf_ours
    assert s[-4 - conf_ours:] == ref[-4 - conf_ref:]
    # the zstd frames may differ (zstd 1.4.5 vs the system's) but decode identically
    out = oracle.decompress_stream(s, int(np.prod(p["dims"])))
    assert np.array_equal(out.view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", HUFFMAN)
def test_huffman_bytes(name):
    c = load(name)
    enc = oracle.huffman_encode(c["codes"])
    assert enc == c["encoded"].tobytes()
    dec, used = oracle.huffman_decode(c["encoded"].tobytes(), c["codes"].size)
    assert used == c["encoded"].size
    assert np.array_equal(dec, c["codes"])


@pytest.mark.parametrize("name", TUNING)
def test_tuning_modes(name):
    c = load(name)
    p = c["params"]
    r = oracle.interp_compress(c["input"], _params(c), tuning=p["tuning"], start=p.get("start", 0),
                               end=p.get("end", 0))
    assert np.array_equal(r["codes"], c["codes"])
    assert np.array_equal(r["recon"].view(np.uint32), c["recon"].view(np.uint32))
    assert r["loss"] == float(c["loss"][0])
    assert np.array_equal(r["level_counts"], c["level_counts"])


@pytest.mark.parametrize("name", LORENZO)
def test_lorenzo_codes_recon(name):
    """SZFastFrontend codes (block-major) and the reconstruction the decoder rebuilds."""
    c = load(name)
    r = oracle.lr_compress(c["input"], oracle.make_lr_params(**lr_kwargs(c)))
    assert np.array_equal(r["codes"], c["codes"])
    assert np.array_equal(r["recon"].view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", LORENZO)
def test_lorenzo_raw_buffer(name):
    """Pre-zstd buffer: meta_params, indicator tree, coefficients, quantizer, codes."""
    c = load(name)
    raw = oracle.lr_raw(c["input"], oracle.make_lr_params(**lr_kwargs(c)))
    assert lr_raw_equal(raw, c["raw"].tobytes())


@pytest.mark.parametrize("name", LORENZO)
def test_lorenzo_decode_reference(name):
    c = load(name)
    dims = c["params"]["dims"]
    rec = oracle.lr_decompress_raw(c["raw"].tobytes(), dims)
    assert np.array_equal(rec.view(np.uint32), c["recon"].view(np.uint32))
    rec2 = oracle.decompress_stream(c["stream"].tobytes(), int(np.prod(dims)))
    assert np.array_equal(rec2.view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", LORENZO)
def test_lorenzo_stream_trailer(name):
    c = load(name)
    p = c["params"]
    kw = lr_kwargs(c)
    s = oracle.lr_compress_stream(c["input"], oracle.make_lr_params(**kw), p["eb"], rel=p.get("mode") == "rel")
    ref = c["stream"].tobytes()
    conf_ref = int(np.frombuffer(ref[-4:], np.int32)[0])
    conf_ours = int(np.frombuffer(s[-4:], np.int32)[0])
    assert conf_ref == conf_ours
    assert s[-4 - conf_ours:] == ref[-4 - conf_ref:]
    out = oracle.decompress_stream(s, int(np.prod(p["dims"])))
    assert np.array_equal(out.view(np.uint32), c["recon"].view(np.uint32))


def test_lorenzo_fixtures_cover_every_branch():
    """The fixtures exercise all three selectors, both unpredictable lists,
    partial blocks and blocks too thin for regression."""
    seen = np.zeros(3, int)
    unpred = reg_unpred = thin = 0
    for name in LORENZO:
        c = load(name)
        kw = lr_kwargs(c)
        r = oracle.lr_compress(c["input"], oracle.make_lr_params(**kw))
        seen += np.bincount(r["indicator"], minlength=3)
        unpred += r["unpred"].size
        reg_unpred += r["reg_unpred"].size
        thin += any(d % kw["block_size"] == 1 or d == 1 for d in kw["dims"])
    assert (seen > 10).all() and unpred > 100 and reg_unpred > 10 and thin >= 2


@pytest.mark.parametrize("name", CTEST)
def test_compress_test_buffer(name):
    """CompressTest (SZInterp.hpp:1131-1392): the encoding_lossless buffer the
    reference hands to zstd and the summed squared error, exactly; the zstd
    frame size of the reference's vendored zstd is within 0.5%."""
    c = load(name)
    p = c["params"]
    raw, nbytes, sq = oracle.compress_test(c["input"], c["starts"], p["sbs"], ctest_params(c, oracle.make_params),
                                           1.0)
    assert raw == c["raw"].tobytes()
    assert sq == float(c["square_error"][0])
    assert abs(nbytes - int(c["ref_zstd_bytes"][0])) <= 0.005 * nbytes + 8


@pytest.mark.parametrize("name", GENERAL + REGRESSION)
def test_general_raw_recon_decode(name):
    """General frontend (SZGeneralFrontend + ComposedPredictor, N = 1..4)
    with the Lorenzo predictors and the RegressionPredictor: the pre-zstd
    buffer bit for bit, the overwritten array equals the reference's decoded
    reconstruction, and the oracle decodes the reference's buffer to it."""
    c = load(name)
    kw = gf_kwargs(c)
    prm = oracle.make_gf_params(**kw)
    assert oracle.gf_raw(c["input"], prm) == c["raw"].tobytes()
    r = oracle.gf_compress(c["input"], prm)
    assert np.array_equal(r["recon"].view(np.uint32), c["recon"].view(np.uint32))
    dec = oracle.gf_decompress_raw(c["raw"].tobytes(), kw["dims"], kw["lorenzo"], kw["lorenzo2"], kw["regression"])
    assert np.array_equal(dec.view(np.uint32), c["recon"].view(np.uint32))
    x = c["input"].astype(np.float64).ravel()
    bad = np.abs(r["recon"].astype(np.float64) - x) > kw["eb"]
    assert not bad.any()
    assert r["unpred"].size == int((r["codes"] == 0).sum())


@pytest.mark.parametrize("name", GENERAL + REGRESSION)
def test_general_stream(name):
    """SZ_compress stream: the Config trailer bit for bit; for N != 3 (where
    SZ_decompress takes the general frontend, LR:117-135) the oracle decodes
    the reference's stream and its own to the reference's reconstruction."""
    c = load(name)
    kw = gf_kwargs(c)
    p = c["params"]
    s = oracle.gf_compress_stream(c["input"], oracle.make_gf_params(**kw), p["eb"], rel=p.get("mode") == "rel")
    ref = c["stream"].tobytes()
    assert trailer(s) == trailer(ref)
    if len(kw["dims"]) != 3:
        for st in (ref, s):
            dec = oracle.decompress_stream(st, c["input"].size)
            assert np.array_equal(dec.view(np.uint32), c["recon"].view(np.uint32))


def test_general_fixtures_cover_every_branch():
    """Both predictors selected, every N, partial and single-point blocks,
    outliers, a blockSize of 1 and of 2."""
    sel = np.zeros(2, int)
    ns, unpred, thin = set(), 0, 0
    for name in GENERAL:
        c = load(name)
        kw = gf_kwargs(c)
        r = oracle.gf_compress(c["input"], oracle.make_gf_params(**kw))
        sel += np.bincount(r["selection"], minlength=2)
        ns.add(len(kw["dims"]))
        unpred += r["unpred"].size
        thin += any(d % kw["block_size"] == 1 for d in kw["dims"])
    assert (sel > 50).all() and ns == {1, 2, 3, 4} and unpred > 100 and thin >= 3
    assert {gf_kwargs(load(n))["block_size"] for n in GENERAL} >= {1, 2}


def test_regression_fixtures_cover_every_branch():
    """The RegressionPredictor chosen in many blocks next to Lorenzo blocks,
    alone, with both Lorenzo predictors, N = 1..4, a last block of extent 1,
    coefficients beyond the coefficient quantizers' range."""
    chosen, mixed, ns = 0, 0, set()
    for name in REGRESSION:
        c = load(name)
        kw = gf_kwargs(c)
        r = oracle.gf_compress(c["input"], oracle.make_gf_params(**kw))
        ns.add(len(kw["dims"]))
        reg_sid = int(kw["lorenzo"]) + int(kw["lorenzo2"])  # the regression predictor's index
        if r["selection"].size:
            k = int((r["selection"] == reg_sid).sum())
            chosen += k
            mixed += 0 < k < r["selection"].size
    assert chosen > 100 and mixed >= 4 and ns == {1, 2, 3, 4}
    assert any(d % gf_kwargs(load(n))["block_size"] == 1 for n in REGRESSION for d in gf_kwargs(load(n))["dims"])


@pytest.mark.parametrize("name", CTEST_LORENZO)
def test_compress_test_lorenzo_buffer(name):
    """CompressTest's Lorenzo variant (SZI:1136-1170, 1283-1294): the
    encoding_lossless buffer and the squared error exactly; the zstd frame
    size within 0.5% of the reference's vendored zstd."""
    c = load(name)
    p = c["params"]
    raw, nbytes, sq = oracle.compress_test_lorenzo(c["input"], c["starts"], p["sbs"], p["eb"])
    assert raw == c["raw"].tobytes()
    assert sq == float(c["square_error"][0])
    assert abs(nbytes - int(c["ref_zstd_bytes"][0])) <= 0.005 * nbytes + 8


def test_compress_test_lorenzo_wide_codes():
    """radius 65536: the wide-codes fixture holds codes above 65535."""
    import struct
    c = load("k2d_65_wide_codes")
    raw = c["raw"].tobytes()
    n = len(c["params"]["dims"])
    off = 8 * n + 4 + 2                       # dims, blockSize, two predictor ids
    nsel = struct.unpack_from("<Q", raw, off)[0]
    assert nsel == len(c["starts"]) * ((c["params"]["sbs"] + 1 + 4) // 5) ** n
    x = c["input"]
    blocks = [x[s[0]:s[0] + 65, s[1]:s[1] + 65] for s in c["starts"]]
    prm = oracle.make_gf_params((65, 65), c["params"]["eb"], block_size=5, radius=65536)
    codes = np.concatenate([oracle.gf_compress(b, prm)["codes"] for b in blocks])
    assert codes.max() > 65535


def sz3_raws(c):
    off = np.concatenate([[0], np.cumsum(c["raw_lens"])])
    return [c["raws"][off[i]:off[i + 1]].tobytes() for i in range(len(c["raw_lens"]))]


@pytest.mark.parametrize("name", SZ3)
def test_sz3_sampling_block_interp_quant_intervals(name):
    """The default (SZ3) tuning mode's pieces against the reference classes:
    sampling<T,N> (Extraction.hpp), the block interpolation compressor's buffer
    for every interpolator and direction (deprecated/
    SZBlockInterpolationCompressor.hpp; zstd size within 0.5%) and
    optimize_quant_invl_3d (QuantOptimization.hpp)."""
    c = load(name)
    p = c["params"]
    x = c["input"]
    s, sdims, sb = oracle.sampling(x)
    assert sb == p["sampling_block"] and (s is not None) == bool(p["sampled"])
    if s is not None:
        assert list(sdims) == p["sample_dims"]
        assert np.array_equal(s.ravel().view(np.uint32), c["sample"].view(np.uint32))
    sample = c["sample"].reshape(p["sample_dims"])
    for (algo, d), raw, ref in zip(c["keys"], sz3_raws(c), c["ref_zstd_bytes"]):
        ours, nbytes = oracle.block_interp_test(sample, p["eb"], int(algo), int(d), p["block"])
        assert ours == raw, (algo, d)
        assert abs(nbytes - int(ref)) <= 0.005 * nbytes + 8
    if x.ndim == 3:
        for f, q in zip((1.0, 0.01), c["quantbin"]):
            assert oracle.quant_intervals(x, p["eb"] * f)[0] == int(q)


def test_sz3_block_interp_round_trip_bound():
    """Block interpolation test: every point of every block gets exactly one
    code, and the buffer's outlier count equals its zero codes."""
    import struct
    x = np.random.default_rng(5).standard_normal((40, 33, 27)).cumsum(0).cumsum(1).astype(np.float32)
    eb = 1e-3 * float(x.max() - x.min())
    raw, _ = oracle.block_interp_test(x, eb, 1, 3, 11)
    o = 3 * 8 + 4 + 4 + 4
    assert raw[o] == 2
    nunpred = struct.unpack_from("<Q", raw, o + 1 + 8 + 4)[0]
    codes, _ = oracle.huffman_decode(raw[o + 1 + 8 + 4 + 8 + 4 * nunpred:], x.size)
    assert codes.size == x.size and int((codes == 0).sum()) == nunpred
