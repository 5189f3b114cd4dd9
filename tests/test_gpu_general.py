"""GPU parity of the general-frontend engine (SZGeneralFrontend with
ComposedPredictor over the 1-layer and 2-layer Lorenzo predictors and the
RegressionPredictor, or one of them alone, N = 1..4; needs an MI355X).

Bar: bit-exact block-major codes, selections, unpredictable values,
reconstructions and pre-zstd buffers against the reference's golden vectors
and the oracle; the GPU decoder reads the reference's own streams; Config
trailers match; the tuner's Lorenzo CompressTest gives the reference's buffer
and squared error; a 1800x3600 field matches the oracle in full.
"""
import numpy as np
import pytest

import oracle
from golden_cases import CTEST_LORENZO, GENERAL, REGRESSION, gf_kwargs, load, trailer

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32).ravel()).cuda()


def _params(kw):
    import faz_amd
    return faz_amd.make_general_params(**kw)


def _check_against_oracle(engine, x, kw):
    o = oracle.gf_compress(x, oracle.make_gf_params(**kw))
    r = engine.general_compress(_dev(x), _params(kw))
    sel = r["selection"][r["selection"] != 255]  # 255: the block recorded no selection
    assert np.array_equal(sel.astype(np.int32), o["selection"])
    assert np.array_equal(r["codes"].astype(np.int32), o["codes"])
    assert np.array_equal(r["unpred"].view(np.uint32), o["unpred"].view(np.uint32))
    assert np.array_equal(r["recon"].cpu().numpy().view(np.uint32), o["recon"].view(np.uint32))
    if kw.get("regression"):
        print(f"  {r['n_regression']} regression blocks, {r['passes']} wavefront passes", flush=True)
    return o


@pytest.mark.parametrize("name", GENERAL + REGRESSION)
def test_golden_raw_recon_decode(engine, name):
    c = load(name)
    kw = gf_kwargs(c)
    print(f"[gpu] general golden {name} {kw}", flush=True)
    _check_against_oracle(engine, c["input"], kw)
    recon = torch.empty(c["input"].size, dtype=torch.float32, device="cuda")
    raw = engine.general_raw(_dev(c["input"]), _params(kw), recon=recon)
    assert raw == c["raw"].tobytes()
    assert np.array_equal(recon.cpu().numpy().view(np.uint32), c["recon"].view(np.uint32))
    out = torch.empty_like(recon)
    engine.general_decompress_raw(c["raw"].tobytes(), kw["dims"], out, kw["lorenzo"], kw["lorenzo2"],
                                  kw["regression"])
    assert np.array_equal(out.cpu().numpy().view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", [n for n in GENERAL + REGRESSION if not n[1:].startswith("3d")])
def test_golden_stream(engine, name):
    """SZ_compress stream (N != 3 takes the general frontend): the reference's
    Config trailer; the reference's stream decodes on the GPU and ours on the
    oracle."""
    import faz_amd
    c = load(name)
    p = c["params"]
    kw = gf_kwargs(c)
    num = c["input"].size
    print(f"[gpu] general stream {name}", flush=True)
    pp = faz_amd.make_general_params(**dict(kw, eb=0.0))
    s = engine.sz_compress_general(c["input"], pp, p["eb"], rel=p.get("mode") == "rel")
    ref = c["stream"].tobytes()
    assert trailer(s) == trailer(ref)
    assert np.array_equal(oracle.decompress_stream(s, num).view(np.uint32), c["recon"].view(np.uint32))
    for st in (ref, s):
        dec = engine.sz_decompress(st, num)
        assert np.array_equal(dec.view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", CTEST_LORENZO)
def test_compress_test_lorenzo_golden(engine, name):
    c = load(name)
    p = c["params"]
    x = c["input"]
    print(f"[gpu] Lorenzo CompressTest {name}", flush=True)
    raw, nbytes, sq = engine.compress_test_lorenzo(_dev(x), x.shape, c["starts"], p["sbs"], p["eb"])
    assert raw == c["raw"].tobytes()
    assert sq == float(c["square_error"][0])
    _, obytes, _ = oracle.compress_test_lorenzo(x, c["starts"], p["sbs"], p["eb"])
    assert nbytes == obytes


def _field(shape, seed):
    r = np.random.default_rng(seed)
    x = r.standard_normal(shape)
    for ax in range(len(shape)):
        x = np.cumsum(np.cumsum(x, axis=ax), axis=ax)
    x = (x - x.min()) / (x.max() - x.min())
    x += 1e-3 * r.standard_normal(shape)
    return x.astype(np.float32)


CASES = [
    ((20000,), dict(), 1e-4),
    ((4097,), dict(block_size=5, lorenzo=False), 1e-3),
    ((300, 410), dict(), 1e-3),
    ((257, 130), dict(block_size=5), 1e-4),
    ((99, 101), dict(block_size=64), 1e-3),
    ((200, 150), dict(block_size=7, lorenzo2=False), 1e-5),
    ((60, 50, 70), dict(block_size=5), 1e-3),
    ((33, 34, 35), dict(block_size=16), 1e-4),
    ((17, 19, 21, 23), dict(), 1e-3),
    ((20, 9, 10, 31), dict(block_size=8, lorenzo=False), 1e-3),
    ((150, 160), dict(block_size=16, radius=512), 1e-6),
    # the RegressionPredictor: composed with each Lorenzo predictor, alone,
    # all three, blocks of extent 1 at the high faces, N = 1..4
    ((20001,), dict(regression=True), 1e-4),
    ((3001,), dict(block_size=10, lorenzo=False, lorenzo2=False, regression=True), 1e-3),
    ((2500,), dict(block_size=7, lorenzo2=True, regression=True), 1e-3),
    ((300, 410), dict(lorenzo2=False, regression=True), 1e-3),
    ((257, 193), dict(block_size=8, lorenzo2=False, regression=True), 1e-4),
    ((129, 97), dict(block_size=16, lorenzo=False, lorenzo2=False, regression=True), 1e-3),
    ((201, 150), dict(block_size=5, lorenzo2=True, regression=True), 1e-3),
    ((61, 50, 73), dict(block_size=6, lorenzo2=False, regression=True), 1e-3),
    ((31, 34, 25), dict(block_size=6, lorenzo=False, regression=True), 1e-5),
    ((17, 19, 21, 25), dict(block_size=6, lorenzo2=False, regression=True), 1e-3),
    ((13, 9, 10, 31), dict(block_size=4, lorenzo=False, lorenzo2=False, regression=True), 1e-3),
]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_random_fields_against_oracle(engine, i):
    shape, extra, rel = CASES[i]
    x = _field(shape, 700 + i)
    eb = rel * float(np.float32(x.max() - x.min()))
    kw = dict(dims=shape, eb=eb, **extra)
    if "block_size" not in kw:
        kw["block_size"] = 128 if len(shape) == 1 else (16 if len(shape) == 2 else 6)
    print(f"[gpu] general oracle parity {shape} {extra} eb={eb}", flush=True)
    o = _check_against_oracle(engine, x, kw)
    raw = engine.general_raw(_dev(x), _params(kw))
    assert raw == oracle.gf_raw(x, oracle.make_gf_params(**kw))
    out = torch.empty(x.size, dtype=torch.float32, device="cuda")
    lo, lo2, reg = kw.get("lorenzo", True), kw.get("lorenzo2", True), kw.get("regression", False)
    nblocks = int(np.prod([(d - 1) // kw["block_size"] + 1 for d in shape]))
    if int(lo) + int(lo2) + int(reg) > 1 and o["selection"].size < nblocks:
        # a 1-D block of one point chose the regression, which declined it:
        # no selection was recorded and the reference's decoder would read
        # past its list, so both decoders reject the buffer
        import faz_amd
        with pytest.raises(faz_amd.FazError):
            engine.general_decompress_raw(raw, shape, out, lo, lo2, reg)
        return
    engine.general_decompress_raw(raw, shape, out, lo, lo2, reg)
    o = oracle.gf_decompress_raw(raw, shape, lo, lo2, reg)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), o.view(np.uint32))


def test_1800x3600_full_parity(engine):
    """The 2-D climate-grid shape at blockSize 16: the whole pre-zstd buffer
    equals the oracle's, the bound holds in FP64 and the GPU decoder
    reproduces the reconstruction."""
    x = _field((1800, 3600), 11)
    eb = 1e-3 * float(np.float32(x.max() - x.min()))
    kw = dict(dims=(1800, 3600), eb=eb, block_size=16)
    print("[gpu] general 1800x3600", flush=True)
    recon = torch.empty(x.size, dtype=torch.float32, device="cuda")
    d = _dev(x)
    raw = engine.general_raw(d, _params(kw), recon=recon)
    assert raw == oracle.gf_raw(x, oracle.make_gf_params(**kw))
    assert (recon.double() - d.double()).abs().max().item() <= eb
    out = torch.empty_like(recon)
    engine.general_decompress_raw(raw, (1800, 3600), out)
    assert torch.equal(out.view(torch.int32), recon.view(torch.int32))


def test_1800x3600_regression_full_parity(engine):
    """ALGO_LORENZO_REG's default predictors for a 2-D field (1-layer Lorenzo
    and the RegressionPredictor, blockSize 16) on the climate-grid shape: the
    pre-zstd buffer equals the oracle's, the stream decodes on the GPU and on
    the oracle to the same reconstruction, within the bound in FP64.  The
    field is smooth on its left half (planes win) and carries a ripple of
    period close to the block size on its right half (Lorenzo wins), so the
    choices mix and the passes over them matter."""
    import faz_amd
    x = _field((1800, 3600), 12)
    i, j = np.meshgrid(np.arange(1800), np.arange(1800, 3600), indexing="ij")
    x[:, 1800:] += (0.02 * np.sin(2 * np.pi * i / 23.0) * np.sin(2 * np.pi * j / 17.0)).astype(np.float32)
    eb = 1e-3 * float(np.float32(x.max() - x.min()))
    kw = dict(dims=(1800, 3600), eb=eb, block_size=16, lorenzo=True, lorenzo2=False, regression=True)
    print("[gpu] general 1800x3600 with the RegressionPredictor", flush=True)
    d = _dev(x)
    r = engine.general_compress(d, _params(kw))
    print(f"  {r['n_regression']} of {113 * 225} blocks regression, {r['passes']} passes", flush=True)
    assert 0 < r["n_regression"] < 113 * 225
    raw = engine.general_raw(d, _params(kw))
    assert raw == oracle.gf_raw(x, oracle.make_gf_params(**kw))
    out = torch.empty(x.size, dtype=torch.float32, device="cuda")
    engine.general_decompress_raw(raw, (1800, 3600), out, True, False, True)
    assert (out.double() - d.double()).abs().max().item() <= eb
    o = oracle.gf_decompress_raw(raw, (1800, 3600), True, False, True)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), o.view(np.uint32))
    # the stream path with the same Config
    pp = faz_amd.make_general_params(**dict(kw, eb=0.0))
    s = engine.sz_compress_general(x, pp, eb, rel=False)
    dec = engine.sz_decompress(s, x.size)
    assert np.array_equal(dec.view(np.uint32), o.view(np.uint32))
    assert np.array_equal(oracle.decompress_stream(s, x.size).view(np.uint32), o.view(np.uint32))


def test_tuner_blocks_against_oracle(engine):
    """The tuner's Lorenzo CompressTest on a field of its own: 40 sample blocks
    of 33^3, compared with the oracle (buffer, size, squared error)."""
    x = _field((132, 99, 165), 5)
    r = np.random.default_rng(3)
    starts = np.array([[r.integers(0, 100), r.integers(0, 67), r.integers(0, 133)] for _ in range(40)], np.int64)
    eb = 2e-4
    print("[gpu] Lorenzo CompressTest 40 blocks", flush=True)
    g = engine.compress_test_lorenzo(_dev(x), x.shape, starts, 32, eb)
    o = oracle.compress_test_lorenzo(x, starts, 32, eb)
    assert g == o


def test_rejects_unsupported(engine):
    import faz_amd
    x = _dev(np.zeros(64 * 64, np.float32))
    with pytest.raises(faz_amd.FazError):
        engine.general_compress(x, faz_amd.make_general_params((64, 64), 1e-3, block_size=65))
    with pytest.raises(faz_amd.FazError):
        engine.general_compress(x, faz_amd.make_general_params((64, 64), 1e-3, block_size=1))
    with pytest.raises(faz_amd.FazError):
        engine.general_compress(x, faz_amd.make_general_params((64, 64), 1e-3, lorenzo=False, lorenzo2=False))
    with pytest.raises(faz_amd.FazError):
        engine.general_compress(x, faz_amd.make_general_params((64, 64), 1e-3, radius=65537))
    with pytest.raises(faz_amd.FazError):
        engine.general_raw(x, faz_amd.make_general_params((64, 64), 1e-3, radius=65536))
