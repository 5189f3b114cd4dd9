"""The benchmark configurations at their real sizes on the GPU (needs an MI355X).

C1  one 128^3 k^-11/3 Gaussian random field (min-max normalised, plus
    N(0, 1e-4^2) noise), REL 1e-3, the reference's default tuning mode as
    served here (SZ3 tuning, `faz -F 0`) and the QoZ tuner (`faz -q`):
    streams equal the oracle's, decode within the bound.
C2  the 512^3 field of bench.py (GRF + 16 tanh fronts) at REL 1e-3 and 1e-5:
    ALGO_INTERP cubic maxStep 32 engine buffers byte-equal to the oracle's,
    the exhaustive QoZ tuner's decisions equal the oracle's restatement and
    the tuned engine buffer byte-equal to the oracle's.
C3  bench.py's seeded 48-field batch (1800x3600 lat-lon, 100x500x500,
    256x384x384 with exact-zero boxes), REL 1e-3: every batched engine
    buffer equals the single call's, one field of each shape equals the
    oracle's engine buffer.
C4  held-out 384^3 fields at REL 1e-3 and 1e-4: the learned tuner's choice
    compresses within 2% of the exhaustive tuner's ratio at equal PSNR.
C5  1024x1024x512 fields (2 GiB) with tanh fronts at REL 1e-2, 1e-3, 1e-4
    and 1e-6 with the learned tuner's choice: every point within the bound in
    FP64, GPU decode equal to the compression's reconstruction; at REL 1e-3
    the ALGO_INTERP engine buffer byte-equal to the oracle's.
"""
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _free():
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _abs_eb(engine, x, rel):
    mn, mx = engine.range(x)
    return rel * float(np.float32(mx - mn)), float(np.float32(mx - mn)), mn, mx


# --- C1 ---------------------------------------------------------------------

def _c1_field(n=128, seed=2024):
    """P(k) ~ k^-11/3 GRF synthesised by FFT from a seed, min-max normalised
    to [0, 1], plus N(0, (1e-4)^2) white noise."""
    r = np.random.default_rng(seed)
    k = np.meshgrid(np.fft.fftfreq(n), np.fft.fftfreq(n), np.fft.rfftfreq(n), indexing="ij")
    kk = np.sqrt(k[0] ** 2 + k[1] ** 2 + k[2] ** 2)
    kk[0, 0, 0] = 1.0
    amp = kk ** (-11.0 / 6.0)
    amp[0, 0, 0] = 0.0
    spec = (r.standard_normal(amp.shape) + 1j * r.standard_normal(amp.shape)) * amp
    x = np.fft.irfftn(spec, s=(n, n, n))
    x = (x - x.min()) / (x.max() - x.min())
    return (x + 1e-4 * r.standard_normal(x.shape)).astype(np.float32)


def test_c1_default_mode_round_trip(engine):
    x = _c1_field()
    rel = 1e-3
    eb = rel * float(np.float32(x.max() - x.min()))
    print("[gpu] C1 128^3 default (SZ3) tuning mode, REL 1e-3", flush=True)
    s = engine.sz_compress_sz3(x, x.shape, rel, rel=True)
    ref, _ = oracle.sz3_compress_stream(x, rel, rel=True)
    assert s == ref
    dec = engine.sz_decompress(s, x.size)
    assert np.array_equal(dec.view(np.uint32), oracle.decompress_stream(s, x.size).view(np.uint32))
    assert np.abs(dec.astype(np.float64) - x.ravel().astype(np.float64)).max() <= eb


def test_c1_qoz_round_trip(engine):
    x = _c1_field()
    rel = 1e-3
    rng = float(np.float32(x.max() - x.min()))
    eb = rel * rng
    print("[gpu] C1 128^3 QoZ tuner (faz -q), REL 1e-3", flush=True)
    s = engine.sz_compress_qoz(x, x.shape, rel, rel=True)
    o = oracle.qoz_tune(x, eb, rel, rng)
    p = oracle.make_params(x.shape, eb, algo=o["interp_algo"], direction=o["direction"], alpha=o["alpha"],
                           beta=o["beta"], max_step=o["max_step"], ops=tuple(o["ops"]), dirs=tuple(o["dirs"]))
    ref = oracle.interp_compress(x, p)["recon"]
    dec = engine.sz_decompress(s, x.size)
    assert np.array_equal(dec.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(oracle.decompress_stream(s, x.size).view(np.uint32), ref.view(np.uint32))
    assert np.abs(dec.astype(np.float64) - x.ravel().astype(np.float64)).max() <= eb


# --- C2 ---------------------------------------------------------------------

@pytest.fixture(scope="module")
def field512():
    import bench
    x = bench.make_field((512, 512, 512), seed=0, fronts=16)
    yield x, x.cpu().numpy().reshape(512, 512, 512)
    del x
    _free()


@pytest.mark.parametrize("rel", [1e-3, 1e-5])
def test_c2_512_engine_buffer_equals_oracle(engine, field512, rel):
    import faz_amd
    x, host = field512
    dims = (512, 512, 512)
    eb, rng, mn, mx = _abs_eb(engine, x, rel)
    assert (mn, mx) == oracle.data_range(host)
    print(f"[gpu] C2 512^3 ALGO_INTERP cubic maxStep 32 rel={rel}: GPU vs oracle engine buffer", flush=True)
    size = engine.compress_device(x, faz_amd.make_params(dims, 0.0, algo=1, max_step=32), True, rel)
    raw = engine.fetch_raw(size)
    want = oracle.interp_raw(host, oracle.make_params(dims, eb, algo=1, max_step=32))
    assert raw == want
    # SZ_compress's stream: same engine buffer and a Config trailer the oracle decodes
    s = engine.sz_compress(x, faz_amd.make_params(dims, eb, algo=1, max_step=32), rel, rel=True)
    dec = engine.sz_decompress(s, host.size)
    ref = oracle.decompress_stream(s, host.size)
    assert np.array_equal(dec.view(np.uint32), ref.view(np.uint32))
    assert np.abs(dec.astype(np.float64) - host.ravel().astype(np.float64)).max() <= eb


@pytest.mark.parametrize("rel", [1e-3, 1e-5])
def test_c2_512_qoz_tuner_equals_oracle(engine, field512, rel):
    import faz_amd
    from test_gpu_tune import _compare
    x, host = field512
    dims = (512, 512, 512)
    eb, rng, _, _ = _abs_eb(engine, x, rel)
    print(f"[gpu] C2 512^3 QoZ tuner rel={rel}: GPU vs oracle decisions", flush=True)
    g = engine.tune_qoz(x, dims, eb, rel, rng)
    o = oracle.qoz_tune(host, eb, rel, rng)
    _compare(g, o)
    kw = dict(algo=o["interp_algo"], direction=o["direction"], alpha=o["alpha"], beta=o["beta"],
              max_step=o["max_step"], ops=tuple(o["ops"]), dirs=tuple(o["dirs"]))
    print(f"[gpu] C2 512^3 tuned engine buffer {kw}", flush=True)
    raw = engine.interp_raw(x, faz_amd.make_params(dims, eb, **kw))
    assert raw == oracle.interp_raw(host, oracle.make_params(dims, eb, **kw))


# --- C3 ---------------------------------------------------------------------

def test_c3_48_field_batch(engine):
    import bench
    import faz_amd
    fields = bench.batch_fields(48, 17)
    ps = [faz_amd.make_params(s, 0.0, algo=1, max_step=64 if len(s) == 2 else 32) for s, _ in fields]
    print("[gpu] C3 48-field batch, REL 1e-3", flush=True)
    b = faz_amd.Batch(engine, ps, [f for _, f in fields], autotune=True)
    try:
        raws, aeb, _ = b.compress_raw(True, 1e-3)
        raws2, _, _ = b.compress_raw(True, 1e-3)  # graph replay
        assert raws2 == raws
    finally:
        b.close()
    seen = set()
    for i, ((shape, f), p) in enumerate(zip(fields, ps)):
        size = engine.compress_device(f, p, True, 1e-3)
        assert engine.fetch_raw(size) == raws[i], (i, shape)
        if shape not in seen:
            seen.add(shape)
            print(f"[gpu] C3 field {i} {shape}: batched buffer vs oracle", flush=True)
            host = f.cpu().numpy().reshape(shape)
            kw = dict(algo=1, max_step=64 if len(shape) == 2 else 32)
            assert raws[i] == oracle.interp_raw(host, oracle.make_params(shape, float(aeb[i]), **kw))
    assert seen == {(1800, 3600), (100, 500, 500), (256, 384, 384)}
    del fields
    _free()


# --- C4 ---------------------------------------------------------------------

@pytest.fixture(scope="module")
def model(engine):
    from faz_amd import learned
    net, mu, sigma = learned.load_model(engine)
    yield net, mu, sigma, learned.load_gates()
    net.close()


@pytest.mark.parametrize("rel,slope,fronts", [(1e-3, 2.6, 4), (1e-3, 4.4, 12), (1e-4, 3.3, 8), (1e-4, 4.8, 0)])
def test_c4_learned_tuner_384(engine, model, rel, slope, fronts):
    from faz_amd import learned
    net, mu, sigma, gates = model
    print(f"[gpu] C4 384^3 rel={rel} slope={slope} fronts={fronts}: learned vs exhaustive", flush=True)
    r = learned.evaluate(engine, net, mu, sigma, (384, 384, 384), (1 << 43) + int(slope * 1000) + fronts, rel,
                         slope=slope, fronts=fronts, gates=gates)
    ex, le = r["exhaustive"], r["learned"]
    cr_e = 32.0 / ex["bitrate"]
    cr_l = learned.ratio_at_equal_psnr(le["bitrate"], le["psnr"], ex["psnr"])
    print(f"  exhaustive ({ex['alpha']},{ex['beta']}) {ex['tune_ms']:.1f} ms CR {cr_e:.3f}; learned "
          f"({le['alpha']},{le['beta']}) {le['tune_ms']:.1f} ms CR@PSNR {cr_l:.3f}", flush=True)
    assert cr_l >= 0.98 * cr_e
    assert le["tune_ms"] * 3 < ex["tune_ms"]
    _free()


# --- C5 ---------------------------------------------------------------------

C5_SHAPE = (1024, 1024, 512)


@pytest.fixture(scope="module")
def field_c5():
    from faz_amd import learned
    x = learned.grf_field(C5_SHAPE, 5 << 40, slope=3.5, fronts=8)
    yield x
    del x
    _free()


@pytest.mark.parametrize("rel", [1e-2, 1e-3, 1e-4, 1e-6])
def test_c5_1024x1024x512_learned_choice(engine, model, field_c5, rel):
    """The learned tuner chooses among the interpolation candidates and the
    Lorenzo test (the exhaustive search as fallback when its gates say so);
    the chosen engine keeps every point within the bound in FP64 and the GPU
    decode equals the compression's reconstruction."""
    import faz_amd
    from faz_amd import learned
    net, mu, sigma, gates = model
    x = field_c5
    eb, rng, _, _ = _abs_eb(engine, x, rel)
    print(f"[gpu] C5 1024x1024x512 rel={rel}: learned choice, bound, decode", flush=True)
    r = learned.tune_learned(engine, x, C5_SHAPE, eb, rel, rng, net, mu, sigma, gates)
    recon = torch.empty_like(x)
    out = torch.empty_like(x)
    if r["use_interp"]:
        raw = engine.interp_raw(x, learned.choice_params(C5_SHAPE, eb, r), recon=recon)
        engine.decompress_raw(raw, 3, out)
    else:
        p = faz_amd.make_lorenzo_params(C5_SHAPE, eb, block_size=5, pred_dim=3, lorenzo=True, lorenzo2=True,
                                        regression=False)
        raw = engine.lorenzo_raw(x, p, recon=recon)
        engine.lorenzo_decompress_raw(raw, C5_SHAPE, out)
    err = (recon.double() - x.double()).abs().max().item()
    assert err <= eb
    mse = ((recon.double() - x.double()) ** 2).mean().item()
    assert torch.equal(out.view(torch.int32), recon.view(torch.int32))
    print(f"  use_interp {r['use_interp']} alpha {r['alpha']} beta {r['beta']} ops {r['ops']} path "
          f"{r['learned_path']}: raw ratio {4.0 * x.numel() / len(raw):.2f}, PSNR "
          f"{20 * math.log10(rng) - 10 * math.log10(mse):.2f} dB, max err {err:.3e} <= {eb:.3e}", flush=True)
    del recon, out
    _free()


def test_c5_1024x1024x512_engine_buffer_equals_oracle(engine, field_c5):
    import faz_amd
    x = field_c5
    rel = 1e-3
    eb, _, _, _ = _abs_eb(engine, x, rel)
    print("[gpu] C5 1024x1024x512 rel=1e-3 ALGO_INTERP: GPU vs oracle engine buffer", flush=True)
    size = engine.compress_device(x, faz_amd.make_params(C5_SHAPE, 0.0, algo=1, max_step=32), True, rel)
    raw = engine.fetch_raw(size)
    host = x.cpu().numpy().reshape(C5_SHAPE)
    assert raw == oracle.interp_raw(host, oracle.make_params(C5_SHAPE, eb, algo=1, max_step=32))
