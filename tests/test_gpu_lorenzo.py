"""GPU parity of the 3-D Lorenzo/regression engine (SZFastFrontend; needs an MI355X).

Bar: bit-exact block-major codes, selectors, unpredictable lists,
reconstructions and pre-zstd buffers against the reference's golden vectors
and the oracle; the GPU decoder reads the reference's own streams; Config
trailers match; on a 256^3 field, size-independent properties (error bound in
FP64, decode reproduces the reconstruction, outlier count = zero codes).
"""
import numpy as np
import pytest

import oracle
from golden_cases import LORENZO, load, lr_kwargs, lr_raw_equal

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32).ravel()).cuda()


def _params(kw):
    import faz_amd
    return faz_amd.make_lorenzo_params(**kw)


def _check_against_oracle(engine, x, kw):
    o = oracle.lr_compress(x, oracle.make_lr_params(**kw))
    r = engine.lorenzo_compress(_dev(x), _params(kw))
    assert np.array_equal(r["selector"].astype(np.int32), o["indicator"])
    assert r["n_regression"] == o["reg_coeffs"].shape[0]
    assert np.array_equal(r["codes"].astype(np.int32), o["codes"])
    assert np.array_equal(r["unpred"].view(np.uint32), o["unpred"].view(np.uint32))
    assert np.array_equal(r["recon"].cpu().numpy().view(np.uint32), o["recon"].view(np.uint32))
    return o


@pytest.mark.parametrize("name", LORENZO)
def test_golden_codes_and_reconstruction(engine, name):
    c = load(name)
    kw = lr_kwargs(c)
    print(f"[gpu] lorenzo golden {name} {kw}", flush=True)
    o = _check_against_oracle(engine, c["input"], kw)
    assert np.array_equal(o["codes"], c["codes"])
    r = engine.lorenzo_compress(_dev(c["input"]), _params(kw))
    assert np.array_equal(r["recon"].cpu().numpy().view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", LORENZO)
def test_golden_raw_buffer_and_decode(engine, name):
    c = load(name)
    kw = lr_kwargs(c)
    dims = c["params"]["dims"]
    print(f"[gpu] lorenzo raw {name}", flush=True)
    raw = engine.lorenzo_raw(_dev(c["input"]), _params(kw))
    assert lr_raw_equal(raw, c["raw"].tobytes())
    out = torch.empty(int(np.prod(dims)), dtype=torch.float32, device="cuda")
    engine.lorenzo_decompress_raw(c["raw"].tobytes(), dims, out)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", LORENZO)
def test_golden_stream(engine, name):
    """SZ_compress stream: same Config trailer as the reference, decodes on the
    oracle; the reference's own stream decodes on the GPU."""
    import faz_amd
    c = load(name)
    p = c["params"]
    kw = lr_kwargs(c)
    num = int(np.prod(p["dims"]))
    print(f"[gpu] lorenzo stream {name}", flush=True)
    pp = faz_amd.make_lorenzo_params(**dict(kw, eb=0.0))
    s = engine.sz_compress_lorenzo(c["input"], pp, p["eb"], rel=p.get("mode") == "rel")
    ref = c["stream"].tobytes()
    conf_ref = int(np.frombuffer(ref[-4:], np.int32)[0])
    conf_ours = int(np.frombuffer(s[-4:], np.int32)[0])
    assert conf_ours == conf_ref
    assert s[-4 - conf_ours:] == ref[-4 - conf_ref:]
    assert np.array_equal(oracle.decompress_stream(s, num).view(np.uint32), c["recon"].view(np.uint32))
    dec = engine.sz_decompress(ref, num)
    assert np.array_equal(dec.view(np.uint32), c["recon"].view(np.uint32))
    dec2 = engine.sz_decompress(s, num)
    assert np.array_equal(dec2.view(np.uint32), c["recon"].view(np.uint32))


def _field(shape, seed, fronts=3):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape)
    for ax in range(3):
        x = np.cumsum(x, axis=ax)
    g = np.stack(np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij"))
    for _ in range(fronts):
        nrm = rng.standard_normal(3)
        nrm /= np.linalg.norm(nrm)
        c0 = rng.uniform(0, max(shape))
        x += 0.5 * x.std() * np.tanh((np.tensordot(nrm, g, 1) - c0) / 4.0)
    x += 0.02 * x.std() * rng.standard_normal(shape)
    return x.astype(np.float32)


CASES = [
    ((64, 48, 80), dict(block_size=6), 1e-3),
    ((61, 50, 47), dict(block_size=5, lorenzo2=True), 1e-3),
    ((40, 70, 33), dict(block_size=6, lorenzo2=True, pred_dim=2), 1e-4),
    ((37, 29, 90), dict(block_size=4, lorenzo2=True, pred_dim=1), 1e-3),
    ((50, 50, 50), dict(block_size=8, lorenzo=False, lorenzo2=True), 1e-2),
    ((33, 65, 34), dict(block_size=16, lorenzo2=True), 1e-3),
    ((30, 31, 32), dict(block_size=1, lorenzo2=True), 1e-3),
    ((45, 45, 45), dict(block_size=6, regression=False), 1e-5),
    ((45, 20, 60), dict(block_size=7, lorenzo2=True, radius=512), 1e-6),
]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_random_fields_against_oracle(engine, i):
    shape, extra, rel = CASES[i]
    x = _field(shape, 300 + i)
    eb = rel * float(np.float32(x.max() - x.min()))
    kw = dict(dims=shape, eb=eb, **extra)
    print(f"[gpu] lorenzo oracle parity {shape} {extra} eb={eb}", flush=True)
    _check_against_oracle(engine, x, kw)
    raw = engine.lorenzo_raw(_dev(x), _params(kw))
    assert lr_raw_equal(raw, oracle.lr_raw(x, oracle.make_lr_params(**kw)))
    out = torch.empty(x.size, dtype=torch.float32, device="cuda")
    engine.lorenzo_decompress_raw(raw, shape, out)
    o = oracle.lr_decompress_raw(raw, shape)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), o.view(np.uint32))


def test_rejects_unsupported(engine):
    import faz_amd
    x = _dev(np.zeros(8 * 8 * 8, np.float32))
    with pytest.raises(faz_amd.FazError):
        engine.lorenzo_compress(x, faz_amd.make_lorenzo_params((8, 8, 8), 1e-3, block_size=17))
    with pytest.raises(faz_amd.FazError):
        engine.lorenzo_compress(x, faz_amd.make_lorenzo_params((8, 8, 8), 1e-3, pred_dim=4))
    with pytest.raises(faz_amd.FazError):
        engine.lorenzo_compress(x, faz_amd.make_lorenzo_params((8, 8, 8), 0.0))


def test_256_properties(engine):
    """256^3 GRF with fronts: error bound in FP64, outliers = zero codes,
    the GPU decoder and the oracle decoder reproduce the reconstruction."""
    import faz_amd
    import bench
    x = bench.make_field((256, 256, 256), seed=3, fronts=16)
    mn, mx = engine.range(x)
    eb = 1e-3 * float(np.float32(mx - mn))
    for extra in (dict(block_size=6), dict(block_size=5, lorenzo2=True)):
        print(f"[gpu] lorenzo 256^3 {extra}", flush=True)
        p = faz_amd.make_lorenzo_params((256, 256, 256), eb, **extra)
        recon = torch.empty_like(x)
        r = engine.lorenzo_compress(x, p, recon=recon)
        err = (recon.double() - x.double()).abs().max().item()
        assert err <= eb
        assert r["unpred"].size == int((r["codes"] == 0).sum())
        raw = engine.lorenzo_raw(x, p)
        out = torch.empty_like(x)
        engine.lorenzo_decompress_raw(raw, (256, 256, 256), out)
        assert torch.equal(out.view(torch.int32), recon.view(torch.int32))
        o = oracle.lr_decompress_raw(raw, (256, 256, 256))
        assert np.array_equal(o.view(np.uint32), recon.cpu().numpy().view(np.uint32))
