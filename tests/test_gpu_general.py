"""GPU parity of the general-frontend Lorenzo engine (SZGeneralFrontend with
ComposedPredictor{1-layer, 2-layer Lorenzo}, N = 1..4; needs an MI355X).

Bar: bit-exact block-major codes, selections, unpredictable values,
reconstructions and pre-zstd buffers against the reference's golden vectors
and the oracle; the GPU decoder reads the reference's own streams; Config
trailers match; the tuner's Lorenzo CompressTest gives the reference's buffer
and squared error; a 1800x3600 field matches the oracle in full.
"""
import numpy as np
import pytest

import oracle
from golden_cases import CTEST_LORENZO, GENERAL, gf_kwargs, load, trailer

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
    assert np.array_equal(r["selection"].astype(np.int32), o["selection"])
    assert np.array_equal(r["codes"].astype(np.int32), o["codes"])
    assert np.array_equal(r["unpred"].view(np.uint32), o["unpred"].view(np.uint32))
    assert np.array_equal(r["recon"].cpu().numpy().view(np.uint32), o["recon"].view(np.uint32))
    return o


@pytest.mark.parametrize("name", GENERAL)
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
    engine.general_decompress_raw(c["raw"].tobytes(), kw["dims"], out, kw["lorenzo"], kw["lorenzo2"])
    assert np.array_equal(out.cpu().numpy().view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", [n for n in GENERAL if not n.startswith("g3d")])
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
    _check_against_oracle(engine, x, kw)
    raw = engine.general_raw(_dev(x), _params(kw))
    assert raw == oracle.gf_raw(x, oracle.make_gf_params(**kw))
    out = torch.empty(x.size, dtype=torch.float32, device="cuda")
    engine.general_decompress_raw(raw, shape, out, kw.get("lorenzo", True), kw.get("lorenzo2", True))
    o = oracle.gf_decompress_raw(raw, shape, kw.get("lorenzo", True), kw.get("lorenzo2", True))
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
