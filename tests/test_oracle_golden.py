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
from golden_cases import (CTEST, HUFFMAN, INTERP, LORENZO, TUNING, ctest_params, load, lr_kwargs, lr_raw_equal,
                          param_kwargs)


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
