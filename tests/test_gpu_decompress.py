"""GPU decompression of interpolation engine buffers (needs an MI355X).

Bar: the decoded field is bit-identical to the oracle's reconstruction
(SZInterpolationCompressor::decompress restated), for engine buffers written
by the oracle, both from host memory and from device memory, on the
whole-tile decoder (every 3-D direction, linear and cubic, alpha/beta and
level-wise schedules, blocks ending short) and on the per-pass recovery.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape)
    for ax in range(len(shape)):
        x = np.cumsum(x, axis=ax)
    x += 0.05 * rng.standard_normal(shape) * x.std()
    return x.astype(np.float32)


# every dimension 32k or 32k + 1 somewhere, so the fine levels qualify for
# the whole-tile decoder; the rest exercise partial tiles and flat axes
CASES = [
    ((64, 65, 96), dict(algo=1, direction=d, max_step=32), 1e-3) for d in range(6)
] + [
    ((65, 64, 97), dict(algo=0, direction=d, max_step=32), 1e-3) for d in (0, 3, 5)
] + [
    ((97, 64, 33), dict(algo=0, direction=3, max_step=32), 1e-4),
    ((65, 65, 65), dict(algo=1, direction=0), 1e-3),
    ((128, 96, 65), dict(algo=1, direction=5, max_step=32, alpha=1.5, beta=2.0), 1e-3),
    ((33, 128, 64), dict(algo=0, direction=1, max_step=16), 1e-2),
    ((96, 64, 64), dict(algo=1, max_step=32, ops=(1, 0, 1), dirs=(0, 5, 3)), 1e-3),
    ((64, 64, 64), dict(algo=1, max_step=32, radius=128), 1e-6),
    ((64, 97, 64), dict(algo=0, direction=4, max_step=32), 1e-3),
    ((70, 45, 100), dict(algo=1, direction=2, max_step=32), 1e-3),
    ((1, 70, 90), dict(algo=1, direction=0, max_step=32), 1e-3),
    ((300, 421), dict(algo=1, direction=1, max_step=64), 1e-3),
]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_decoders_match_oracle(engine_tile, engine_pass, i):
    shape, kw, rel = CASES[i]
    x = _field(shape, 900 + i)
    eb = rel * float(np.float32(x.max() - x.min()))
    print(f"[gpu] decode {shape} {kw} eb={eb}", flush=True)
    op = oracle.make_params(shape, eb, **kw)
    raw = oracle.interp_raw(x, op)
    want = oracle.interp_compress(x, op)["recon"].view(np.uint32)
    d_raw = torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda()
    for eng in (engine_tile, engine_pass):
        out = torch.empty(x.size, dtype=torch.float32, device="cuda")
        eng.decompress_raw(raw, len(shape), out)
        assert np.array_equal(out.cpu().numpy().view(np.uint32), want)
        out.fill_(np.nan)
        eng.decompress_dev(d_raw, len(shape), out)
        assert np.array_equal(out.cpu().numpy().view(np.uint32), want)


def test_device_buffer_at_odd_offset(engine_tile):
    """The engine buffer may start at any byte of device memory."""
    shape = (64, 65, 64)
    x = _field(shape, 950)
    eb = 1e-3 * float(np.float32(x.max() - x.min()))
    op = oracle.make_params(shape, eb, algo=1, max_step=32)
    raw = oracle.interp_raw(x, op)
    want = oracle.interp_compress(x, op)["recon"].view(np.uint32)
    buf = torch.zeros(len(raw) + 3, dtype=torch.uint8, device="cuda")
    buf[3:] = torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda()
    out = torch.empty(x.size, dtype=torch.float32, device="cuda")
    engine_tile.decompress_dev(buf[3:], 3, out)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want)


def test_truncated_device_buffer_is_an_error(engine_tile):
    import faz_amd
    shape = (40, 33, 50)
    x = _field(shape, 951)
    eb = 1e-3 * float(np.float32(x.max() - x.min()))
    raw = oracle.interp_raw(x, oracle.make_params(shape, eb, algo=1, max_step=32))
    d_raw = torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda()
    out = torch.empty(x.size, dtype=torch.float32, device="cuda")
    for n in (10, len(raw) // 2, len(raw) - 1):
        with pytest.raises(faz_amd.FazError):
            engine_tile.decompress_dev(d_raw, 3, out, nbytes=n)
    with pytest.raises(faz_amd.FazError):  # wrong point count
        engine_tile.decompress_dev(d_raw, 3, torch.empty(x.size + 1, dtype=torch.float32, device="cuda"))
    # the context still decodes afterwards
    engine_tile.decompress_dev(d_raw, 3, out)
    want = oracle.interp_compress(x, oracle.make_params(shape, eb, algo=1, max_step=32))["recon"]
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want.view(np.uint32))


def _bitstream_at(raw, N=3, levelwise=0):
    """Offset of the Huffman bitstream in an engine buffer (header SIC:489-512,
    LinearQuantizer::save Q:185-201, HuffmanEncoder::save HUF:119-137)."""
    import struct
    h = N * 8 + 61 + 2 * levelwise
    U = struct.unpack_from("<Q", raw, h + 13)[0]
    t = h + 21 + 4 * U
    nc = int.from_bytes(raw[t + 4:t + 8], "big")
    w = 1 if nc <= 256 else (2 if nc <= 65536 else 4)
    at = t + 12 + 1 + 2 * nc * w + 5 * nc + 8
    assert struct.unpack_from("<Q", raw, at - 8)[0] == len(raw) - at
    return at


def test_anchor_run_shortcut_agrees_with_full_search(engine_tile, engine_norun):
    """The decoder skips the trajectory search inside the anchors' opening
    run of zero codes once a device check has matched the run's bits; with a
    bit of that run flipped the check fails and the decoder must behave as
    the full search does (same output or the same error)."""
    import faz_amd
    shape = (129, 130, 131)  # 5^3 anchors at maxStep 32: about two chunks of run
    x = _field(shape, 952)
    eb = 1e-4 * float(np.float32(x.max() - x.min()))
    op = oracle.make_params(shape, eb, algo=1, max_step=32)
    raw = bytearray(oracle.interp_raw(x, op))
    want = oracle.interp_compress(x, op)["recon"].view(np.uint32)
    for eng in (engine_tile, engine_norun):
        out = torch.empty(x.size, dtype=torch.float32, device="cuda")
        eng.decompress_dev(torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda(), 3, out)
        assert np.array_equal(out.cpu().numpy().view(np.uint32), want)
    bits_at = _bitstream_at(raw)
    raw[bits_at + 40] ^= 0x10  # inside the anchors' run
    results = []
    for eng in (engine_tile, engine_norun):
        out = torch.zeros(x.size, dtype=torch.float32, device="cuda")
        try:
            eng.decompress_dev(torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda(), 3, out)
            results.append(("ok", out.cpu().numpy().view(np.uint32).copy()))
        except faz_amd.FazError as e:
            results.append(("error", str(e).split(":")[0]))
    assert results[0][0] == results[1][0]
    if results[0][0] == "ok":
        assert np.array_equal(results[0][1], results[1][1])


@pytest.mark.parametrize("rel", [1e-3, 1e-5])
def test_full_size_512_device_decode(engine, engine_norun, rel):
    """512^3: the device-resident decoder reproduces the compression's
    reconstruction bit for bit (whole-tile decoder on levels 3-1)."""
    import bench
    import faz_amd
    x = bench.make_field((512, 512, 512), seed=4, fronts=16)
    mn, mx = engine.range(x)
    eb = rel * float(np.float32(mx - mn))
    print(f"[gpu] 512^3 device decode rel={rel}", flush=True)
    p = faz_amd.make_params((512, 512, 512), eb, algo=1, max_step=32)
    recon = torch.empty_like(x)
    raw = engine.interp_raw(x, p, recon=recon)
    d_raw = torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda()
    out = torch.empty_like(x)
    engine.decompress_dev(d_raw, 3, out)
    assert torch.equal(out.view(torch.int32), recon.view(torch.int32))
    out.fill_(float("nan"))
    engine_norun.decompress_raw(raw, 3, out)
    assert torch.equal(out.view(torch.int32), recon.view(torch.int32))
    del x, recon, out, d_raw
    torch.cuda.empty_cache()
