"""GPU parity of the interpolation + quantization + Huffman path (needs an MI355X).

Bar: bit-exact codes, unpredictable lists, reconstructions, pre-zstd engine
buffers and Config trailers against the reference's golden vectors and the
oracle; the GPU decoder reads the reference's own streams; at the full
512^3 size, size-independent properties (error bound in FP64, decode of the
stream reproduces the reconstruction, code/outlier counts agree).
"""
import numpy as np
import pytest

import oracle
from golden_cases import INTERP, load, param_kwargs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _faz_params(case):
    import faz_amd
    return faz_amd.make_params(**param_kwargs(case))


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32).ravel()).cuda()


@pytest.mark.parametrize("name", INTERP)
def test_golden_codes_unpred_recon(engine, name):
    c = load(name)
    print(f"[gpu] golden case {name} dims={c['params']['dims']}", flush=True)
    r = engine.interp_compress(_dev(c["input"]), _faz_params(c))
    assert np.array_equal(r["codes"].astype(np.int32), c["codes"])
    assert np.array_equal(r["unpred"].view(np.uint32), c["unpred"].view(np.uint32))
    rec = r["recon"].cpu().numpy()
    assert np.array_equal(rec.view(np.uint32), c["recon"].view(np.uint32))


@pytest.mark.parametrize("name", INTERP)
def test_golden_raw_buffer(engine, name):
    c = load(name)
    print(f"[gpu] raw buffer {name}", flush=True)
    raw = engine.interp_raw(_dev(c["input"]), _faz_params(c))
    assert raw == c["raw"].tobytes()


@pytest.mark.parametrize("name", INTERP)
def test_golden_stream_and_reference_stream_decode(engine, name):
    c = load(name)
    p = c["params"]
    print(f"[gpu] stream {name}", flush=True)
    import faz_amd
    kw = param_kwargs(c)
    s = engine.sz_compress(c["input"], faz_amd.make_params(**kw), p["eb"], rel=p.get("mode", "abs") == "rel")
    ref = c["stream"].tobytes()
    conf = int(np.frombuffer(ref[-4:], np.int32)[0])
    assert s[-4 - conf:] == ref[-4 - conf:]
    num = int(np.prod(p["dims"]))
    # our stream decodes on the oracle, the reference's stream on the GPU
    assert np.array_equal(oracle.decompress_stream(s, num).view(np.uint32), c["recon"].view(np.uint32))
    dec = engine.sz_decompress(ref, num)
    assert np.array_equal(dec.view(np.uint32), c["recon"].view(np.uint32))


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape)
    for ax in range(len(shape)):
        x = np.cumsum(x, axis=ax)
    x += 0.05 * rng.standard_normal(shape) * x.std()
    return x.astype(np.float32)


CONFIGS = [
    ((64, 96, 80), dict(algo=1, direction=0, max_step=32), 1e-3),
    ((97, 65, 130), dict(algo=1, direction=5, max_step=32, alpha=1.25, beta=2.0), 1e-4),
    ((129, 129, 129), dict(algo=0, direction=2), 1e-3),
    ((70, 200, 33), dict(algo=1, direction=3, max_step=16, ops=(1, 0, 1, 1), dirs=(0, 5, 1, 2)), 1e-2),
    ((300, 421), dict(algo=1, direction=1, max_step=64), 1e-3),
    ((100000,), dict(algo=1), 1e-4),
    ((12, 20, 24, 30), dict(algo=1, direction=7), 1e-3),
    ((48, 48, 48), dict(algo=1, max_step=32, block_size=8), 1e-3),
    ((40, 50, 60), dict(algo=1, radius=256), 1e-5),
]


@pytest.mark.parametrize("i", range(len(CONFIGS)))
def test_random_fields_against_oracle(engine, i):
    import faz_amd
    shape, kw, rel = CONFIGS[i]
    x = _field(shape, 100 + i)
    eb = rel * float(np.float32(x.max() - x.min()))
    print(f"[gpu] oracle parity {shape} {kw} eb={eb}", flush=True)
    o = oracle.interp_compress(x, oracle.make_params(shape, eb, **kw))
    r = engine.interp_compress(_dev(x), faz_amd.make_params(shape, eb, **kw))
    assert np.array_equal(r["codes"].astype(np.int32), o["codes"])
    assert np.array_equal(r["unpred"].view(np.uint32), o["unpred"].view(np.uint32))
    assert np.array_equal(r["recon"].cpu().numpy().view(np.uint32), o["recon"].view(np.uint32))
    raw = engine.interp_raw(_dev(x), faz_amd.make_params(shape, eb, **kw))
    assert raw == oracle.interp_raw(x, oracle.make_params(shape, eb, **kw))
    out = torch.empty(x.size, dtype=torch.float32, device="cuda")
    engine.decompress_raw(raw, len(shape), out)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), o["recon"].view(np.uint32))


FUSED3D = [
    ((65, 66, 67), dict(algo=1, direction=d, max_step=32), 1e-3) for d in range(6)
] + [
    ((33, 97, 40), dict(algo=0, direction=5, max_step=16), 1e-4),
    ((1, 70, 90), dict(algo=1, direction=0, max_step=32), 1e-3),
    ((70, 1, 90), dict(algo=1, direction=4), 1e-3),
    ((2, 3, 129), dict(algo=1, direction=1), 1e-3),
    ((64, 64, 64), dict(algo=1, direction=0, block_size=16, max_step=16), 1e-3),
    ((50, 45, 40), dict(algo=1, direction=2, alpha=1.5, beta=2.0, ops=(1, 0, 1), dirs=(0, 5, 3)), 1e-3),
    ((96, 96, 96), dict(algo=1, direction=5, block_size=8), 1e-2),
    ((40, 40, 40), dict(algo=1, radius=128), 1e-6),
]


@pytest.mark.parametrize("path", ["fused", "pass"])
@pytest.mark.parametrize("i", range(len(FUSED3D)))
def test_3d_level_kernels_against_oracle(engine_tile, engine_pass, path, i):
    import faz_amd
    eng = engine_tile if path == "fused" else engine_pass
    shape, kw, rel = FUSED3D[i]
    x = _field(shape, 300 + i)
    eb = rel * float(np.float32(x.max() - x.min()))
    print(f"[gpu] 3-D {path} {shape} {kw} eb={eb}", flush=True)
    o = oracle.interp_compress(x, oracle.make_params(shape, eb, **kw))
    r = eng.interp_compress(_dev(x), faz_amd.make_params(shape, eb, **kw))
    assert np.array_equal(r["codes"].astype(np.int32), o["codes"])
    assert np.array_equal(r["unpred"].view(np.uint32), o["unpred"].view(np.uint32))
    assert np.array_equal(r["recon"].cpu().numpy().view(np.uint32), o["recon"].view(np.uint32))


# Fields whose every dimension is 32k or 32k + 1: their fine levels run on the
# whole-tile kernel (blocks of 32 or 33 points per axis), coarser ones on the
# per-block level kernel.
TILE3D = [
    ((64, 65, 96), dict(algo=1, direction=d, max_step=32), 1e-3) for d in range(6)
] + [
    ((97, 64, 33), dict(algo=0, direction=3, max_step=32), 1e-4),
    ((65, 65, 65), dict(algo=1, direction=0), 1e-3),
    ((128, 96, 65), dict(algo=1, direction=5, max_step=32, alpha=1.5, beta=2.0), 1e-3),
    ((33, 128, 64), dict(algo=0, direction=1, max_step=16), 1e-2),
    ((96, 64, 64), dict(algo=1, max_step=32, ops=(1, 0, 1), dirs=(0, 5, 3)), 1e-3),
    ((64, 64, 64), dict(algo=1, max_step=32, radius=128), 1e-6),
    ((64, 97, 64), dict(algo=0, direction=4, max_step=32), 1e-3),
]


@pytest.mark.parametrize("i", range(len(TILE3D)))
def test_tile_kernel_against_oracle(engine_tile, i):
    import faz_amd
    shape, kw, rel = TILE3D[i]
    x = _field(shape, 500 + i)
    eb = rel * float(np.float32(x.max() - x.min()))
    print(f"[gpu] tile {shape} {kw} eb={eb}", flush=True)
    o = oracle.interp_compress(x, oracle.make_params(shape, eb, **kw))
    r = engine_tile.interp_compress(_dev(x), faz_amd.make_params(shape, eb, **kw))
    assert np.array_equal(r["codes"].astype(np.int32), o["codes"])
    assert np.array_equal(r["unpred"].view(np.uint32), o["unpred"].view(np.uint32))
    assert np.array_equal(r["recon"].cpu().numpy().view(np.uint32), o["recon"].view(np.uint32))


@pytest.mark.parametrize("algo", [0, 1])
def test_tile_kernel_exact_quantizer(engine_tile, algo):
    """An error bound so small against max |x| that |x - pred| / eb may pass
    2^31: the whole-tile levels keep the quantizer's x86 out-of-range path."""
    import faz_amd
    shape = (64, 65, 64)
    x = _field(shape, 800 + algo) + np.float32(1e4)
    eb = 1e-9 * float(np.float32(x.max() - x.min()))
    kw = dict(algo=algo, direction=2, max_step=32)
    print(f"[gpu] tile exact quantizer {shape} eb={eb}", flush=True)
    o = oracle.interp_compress(x, oracle.make_params(shape, eb, **kw))
    r = engine_tile.interp_compress(_dev(x), faz_amd.make_params(shape, eb, **kw))
    assert np.array_equal(r["codes"].astype(np.int32), o["codes"])
    assert np.array_equal(r["unpred"].view(np.uint32), o["unpred"].view(np.uint32))
    assert np.array_equal(r["recon"].cpu().numpy().view(np.uint32), o["recon"].view(np.uint32))


@pytest.mark.parametrize("i", [0, 2, 7, 9, 10, 11])
def test_compression_without_reconstruction(engine_tile, i):
    """SZ_compress returns no reconstruction: without one the whole-tile levels
    hand compact lattices to each other and the finest level writes nothing;
    the engine buffer stays bit-exact."""
    import faz_amd
    shape, kw, rel = TILE3D[i]
    x = _field(shape, 700 + i)
    eb = rel * float(np.float32(x.max() - x.min()))
    print(f"[gpu] no-reconstruction path {shape} {kw}", flush=True)
    size = engine_tile.compress_device(_dev(x), faz_amd.make_params(shape, eb, **kw), False, eb)
    assert engine_tile.fetch_raw(size) == oracle.interp_raw(x, oracle.make_params(shape, eb, **kw))


@pytest.mark.parametrize("rel", [3e-5, 3e-4])
def test_codes_outside_histogram_window(engine_tile, rel):
    """White noise at a small bound: quantization codes spread far beyond the
    4096 symbols around the radius that the compression step reads back, so
    the whole histogram is read instead; the buffer stays bit-exact."""
    import faz_amd
    shape = (40, 48, 56)
    x = np.random.default_rng(31).standard_normal(shape).astype(np.float32)
    eb = rel * float(np.float32(x.max() - x.min()))
    params = dict(algo=1, direction=0, max_step=32)
    want = oracle.interp_raw(x, oracle.make_params(shape, eb, **params))
    codes = oracle.interp_compress(x, oracle.make_params(shape, eb, **params))["codes"]
    far = np.count_nonzero((codes != 0) & (np.abs(codes - 32768) > 2048))
    print(f"[gpu] histogram window fallback rel={rel} far codes={far}", flush=True)
    assert far > 0
    size = engine_tile.compress_device(_dev(x), faz_amd.make_params(shape, eb, **params), False, eb)
    assert engine_tile.fetch_raw(size) == want
    # the next compression of the same context reads the window again
    y = _field(shape, 32)
    eb = 1e-3 * float(np.float32(y.max() - y.min()))
    size = engine_tile.compress_device(_dev(y), faz_amd.make_params(shape, eb, **params), False, eb)
    assert engine_tile.fetch_raw(size) == oracle.interp_raw(y, oracle.make_params(shape, eb, **params))


REL_CASES = [
    ((64, 65, 96), dict(algo=1, direction=0, max_step=32), 1e-3),
    ((97, 64, 33), dict(algo=0, direction=3, max_step=32), 1e-4),
    ((128, 96, 65), dict(algo=1, direction=5, max_step=32, alpha=1.5, beta=2.0), 1e-3),
    ((50, 45, 40), dict(algo=1, direction=2, alpha=0.5, beta=0.75, ops=(1, 0, 1), dirs=(0, 5, 3)), 1e-2),
    ((96, 64, 64), dict(algo=1, max_step=32, alpha=1.25, beta=1.5), 1e-5),
    ((300, 421), dict(algo=1, direction=1, max_step=64), 1e-3),
    ((100000,), dict(algo=1), 1e-4),
    ((64, 65, 64), dict(algo=1, direction=2, max_step=32), 1e-9),  # exact quantizer on the tile levels
]


@pytest.mark.parametrize("i", range(len(REL_CASES)))
def test_rel_bounds_on_device(engine, engine_tile, engine_hostbounds, i):
    """REL compression computes the level bounds on the device from the value
    range (no host round trip): the engine buffer equals the oracle's at
    eb = rel * (max - min) and the host-bound path's."""
    import faz_amd
    shape, kw, rel = REL_CASES[i]
    x = _field(shape, 1100 + i) + np.float32(1e4 if rel < 1e-8 else 0.0)
    mn, mx = oracle.data_range(x)
    eb = rel * float(np.float32(mx - mn))
    print(f"[gpu] device level bounds {shape} {kw} rel={rel}", flush=True)
    want = oracle.interp_raw(x, oracle.make_params(shape, eb, **kw))
    d = _dev(x)
    for eng in (engine, engine_tile, engine_hostbounds):
        size = eng.compress_device(d, faz_amd.make_params(shape, 1.0, **kw), True, rel)
        assert eng.fetch_raw(size) == want


def test_rel_bounds_constant_field_is_an_error(engine):
    import faz_amd
    d = _dev(np.full((40, 40, 40), 3.0, np.float32))
    with pytest.raises(faz_amd.FazError):
        engine.compress_device(d, faz_amd.make_params((40, 40, 40), 1.0, algo=1, max_step=32), True, 1e-3)
    # the context compresses again afterwards
    x = _field((40, 40, 40), 1200)
    mn, mx = oracle.data_range(x)
    eb = 1e-3 * float(np.float32(mx - mn))
    p = dict(algo=1, max_step=32)
    size = engine.compress_device(_dev(x), faz_amd.make_params((40, 40, 40), 1.0, **p), True, 1e-3)
    assert engine.fetch_raw(size) == oracle.interp_raw(x, oracle.make_params((40, 40, 40), eb, **p))


def test_range_matches_oracle(engine):
    x = _field((33, 45, 57), 5) - 3.0
    mn, mx = engine.range(_dev(x))
    assert (mn, mx) == oracle.data_range(x)


def test_rejects_unsupported_modes(engine):
    import faz_amd
    x = _dev(np.zeros(64, np.float32))
    with pytest.raises(faz_amd.FazError):
        engine.interp_compress(x, faz_amd.make_params((64,), 1e-3, max_step=16))  # 1-D anchors
    with pytest.raises(faz_amd.FazError):
        engine.interp_compress(x, faz_amd.make_params((64,), 1e-3, radius=65536))
    with pytest.raises(faz_amd.FazError):
        engine.interp_compress(x, faz_amd.make_params((64,), 0.0))


def _grf_512(seed=0):
    """512^3 GRF with P(k) ~ k^-11/3 plus 16 tanh fronts, synthesised on the GPU."""
    import bench
    return bench.make_field((512, 512, 512), seed=seed, fronts=16)


@pytest.mark.parametrize("rel", [1e-3, 1e-5])
def test_full_size_512_properties(engine, rel):
    import faz_amd
    x = _grf_512()
    mn, mx = engine.range(x)
    eb = rel * float(np.float32(mx - mn))
    print(f"[gpu] 512^3 rel={rel} eb={eb}", flush=True)
    p = faz_amd.make_params((512, 512, 512), eb, algo=1, max_step=32)
    recon = torch.empty_like(x)
    raw = engine.interp_raw(x, p, recon=recon)
    err = (recon.double() - x.double()).abs().max().item()
    assert err <= eb
    out = torch.empty_like(x)
    engine.decompress_raw(raw, 3, out)
    assert torch.equal(out.view(torch.int32), recon.view(torch.int32))
    r = engine.interp_compress(x, p)
    assert r["n_codes"] == x.numel()
    assert r["unpred"].size == int((r["codes"] == 0).sum())
    # the bench path: no reconstruction requested, same engine buffer
    size = engine.compress_device(x, p, False, eb)
    assert engine.fetch_raw(size) == raw
    del x, recon, out
    torch.cuda.empty_cache()


def test_full_size_kernels_agree(engine, engine_level, engine_pass):
    """At 512^3 the whole-tile kernel, the per-block level kernel and the
    per-pass launches agree bit for bit."""
    import faz_amd
    x = _grf_512(seed=1)
    mn, mx = engine.range(x)
    eb = 1e-3 * float(np.float32(mx - mn))
    for d, algo in ((0, 1), (5, 1), (3, 0)):
        p = faz_amd.make_params((512, 512, 512), eb, algo=algo, direction=d, max_step=32)
        print(f"[gpu] 512^3 tile vs level vs per-pass dir={d} algo={algo}", flush=True)
        a = engine.interp_compress(x, p)
        for other in (engine_level, engine_pass):
            b = other.interp_compress(x, p)
            assert np.array_equal(a["codes"], b["codes"])
            assert np.array_equal(a["unpred"].view(np.uint32), b["unpred"].view(np.uint32))
            assert torch.equal(a["recon"].view(torch.int32), b["recon"].view(torch.int32))
            del b
    del x
    torch.cuda.empty_cache()


@pytest.mark.parametrize("rel", [1e-3, 1e-5])
def test_one_pass_encoder_agrees(engine, engine_onepass, rel):
    """The one-pass Huffman encoder (decoupled look-back over the chunks)
    writes the same engine buffer as the sizes + scan + pack encoder, at
    512^3 and on a small field of a few chunks."""
    import faz_amd
    for shape, seed in (((512, 512, 512), 2), ((40, 50, 60), 3)):
        x = _grf_512(seed=seed) if shape[0] == 512 else _dev(np.cumsum(np.random.default_rng(seed).standard_normal(shape), 2))
        mn, mx = engine.range(x)
        eb = rel * float(np.float32(mx - mn))
        p = faz_amd.make_params(shape, eb, algo=1, max_step=32 if shape[0] == 512 else 16)
        print(f"[gpu] one-pass vs two-pass encoder {shape} rel={rel}", flush=True)
        assert engine_onepass.interp_raw(x, p) == engine.interp_raw(x, p)
        del x
    torch.cuda.empty_cache()
