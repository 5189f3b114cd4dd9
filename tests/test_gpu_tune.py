"""GPU parity of the exhaustive QoZ tuner (`faz -q`, needs an MI355X).

Bar: the same sample blocks, bit-identical summed mode-2 losses for every
(level, interpolator, direction), the same sequence of CompressTest calls with
bit-identical bitrates and PSNRs (same compressed sizes), and the same final
alpha, beta and level-wise lists as the oracle's restatement of Tuning().
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _grf(shape, slope=11 / 3, seed=0, fronts=0):
    r = np.random.default_rng(seed)
    k = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    kk = np.sqrt(sum(a * a for a in k))
    kk[(0,) * len(shape)] = 1
    f = (r.standard_normal(shape) + 1j * r.standard_normal(shape)) * kk ** (-slope / 2)
    f[(0,) * len(shape)] = 0
    x = np.real(np.fft.ifftn(f))
    g = np.stack(np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij"))
    for _ in range(fronts):
        nrm = r.standard_normal(len(shape))
        nrm /= np.linalg.norm(nrm)
        x += 0.5 * x.std() * np.tanh((np.tensordot(nrm, g, 1) - r.uniform(0, max(shape))) / 4.0)
    x = (x - x.min()) / (x.max() - x.min())
    return (x + 1e-4 * r.standard_normal(shape)).astype(np.float32)


CASES = [
    ((96, 96, 96), 1e-3, 0),
    ((200, 200, 200), 1e-3, 4),
    ((130, 140, 150), 1e-4, 2),
    ((300, 400), 1e-3, 0),
    ((64, 80, 300), 1e-2, 1),
    ((160, 160, 160), 1e-5, 3),
]


def _compare(g, o):
    assert g["sample_block_size"] == o["sample_block_size"]
    assert g["n_blocks"] == o["n_blocks"]
    assert np.array_equal(g["block_starts"], o["block_starts"])
    assert np.array_equal(g["pred_loss"], o["pred_loss"])
    assert len(g["tests"]) == len(o["tests"])
    for a, b in zip(g["tests"], o["tests"]):
        assert (a["alpha"], a["beta"], a["eb"]) == (b["alpha"], b["beta"], b["eb"])
        assert a["bytes"] == b["bytes"]
        assert a["bitrate"] == b["bitrate"] and a["psnr"] == b["psnr"]
    for k in ("alpha", "beta", "levelwise", "ops", "dirs", "interp_algo", "direction", "max_step", "use_interp"):
        assert g[k] == o[k], k


@pytest.mark.parametrize("i", range(len(CASES)))
def test_tuner_matches_oracle(engine, i):
    shape, rel, fronts = CASES[i]
    x = _grf(shape, seed=40 + i, fronts=fronts)
    rng = float(np.float32(x.max() - x.min()))
    eb = rel * rng
    print(f"[gpu] tuner {shape} rel={rel}", flush=True)
    o = oracle.qoz_tune(x, eb, rel, rng)
    g = engine.tune_qoz(torch.from_numpy(x.ravel()).cuda(), shape, eb, rel, rng)
    _compare(g, o)


def _walk(shape, seed):
    """Random walk along every axis: rough at the sample-block scale, where
    the Lorenzo test can win."""
    x = np.random.default_rng(seed).standard_normal(shape)
    for ax in range(len(shape)):
        x = np.cumsum(x, axis=ax)
    return x.astype(np.float32)


@pytest.mark.parametrize("shape,rel,kind,ab_list", [((96, 96, 96), 1e-3, "walk", 0), ((300, 400), 1e-2, "walk", 1),
                                                    ((130, 140, 150), 1e-4, "grf", 2), ((200, 200, 200), 1e-2, "grf", 1),
                                                    ((120, 130, 110), 1e-5, "walk", 2)])
def test_tuner_lorenzo_test_and_alpha_lists_match_oracle(engine, shape, rel, kind, ab_list):
    """conf.testLorenzo (the Lorenzo CompressTest after the alpha/beta search,
    SZI:2431-2493) and the wider alpha lists of conf.abList (SZI:883-891)."""
    x = _walk(shape, 5) if kind == "walk" else _grf(shape, seed=6, fronts=2)
    rng = float(np.float32(x.max() - x.min()))
    eb = rel * rng
    print(f"[gpu] tuner {shape} {kind} rel={rel} abList={ab_list} testLorenzo=1", flush=True)
    o = oracle.qoz_tune(x, eb, rel, rng, ab_list=ab_list, test_lorenzo=1)
    g = engine.tune_qoz(torch.from_numpy(x.ravel()).cuda(), shape, eb, rel, rng, ab_list=ab_list, test_lorenzo=1)
    _compare(g, o)
    assert any(t["alpha"] < 0 for t in g["tests"])  # the Lorenzo test ran
    if ab_list:
        assert max(t["alpha"] for t in g["tests"]) > 2.0 or g["tests"][-1]["alpha"] < 0


def test_tuner_lorenzo_wins_on_a_rough_field(engine):
    x = _walk((96, 96, 96), 5)
    rng = float(np.float32(x.max() - x.min()))
    g = engine.tune_qoz(torch.from_numpy(x.ravel()).cuda(), x.shape, 1e-3 * rng, 1e-3, rng, test_lorenzo=1)
    o = oracle.qoz_tune(x, 1e-3 * rng, 1e-3, rng, test_lorenzo=1)
    print(f"[gpu] rough field: use_interp GPU {g['use_interp']} oracle {o['use_interp']}", flush=True)
    assert g["use_interp"] == o["use_interp"]


from golden_cases import CTEST, ctest_params, load  # noqa: E402


@pytest.mark.parametrize("name", CTEST)
def test_compress_test_golden(engine, name):
    """GPU CompressTest against the reference's encoding_lossless buffer and
    squared error (fixtures from the reference classes)."""
    import faz_amd
    c = load(name)
    p = c["params"]
    x = c["input"]
    print(f"[gpu] CompressTest {name}", flush=True)
    raw, nbytes, sq = engine.compress_test(torch.from_numpy(x.ravel()).cuda(), x.shape, c["starts"], p["sbs"],
                                           ctest_params(c, faz_amd.make_params), 1.0)
    assert raw == c["raw"].tobytes()
    assert sq == float(c["square_error"][0])


def _trailer(stream):
    """The leading Config fields of a stream's trailer (Config.hpp:240-270)."""
    import struct
    n = int(np.frombuffer(stream[-4:], np.int32)[0])
    c = stream[-4 - n:-4]
    N = c[0]
    o = 1 + 8 * N + 8
    algo = c[o]
    o += 3
    absb, relb, alpha, beta = struct.unpack_from("<4d", c, o)
    o += 32 + 4
    interp, direction = c[o], c[o + 1]
    o += 2
    ibs, = struct.unpack_from("<i", c, o)
    o += 4 + 2
    qbins, bs, lw = struct.unpack_from("<3i", c, o)
    o += 12 + 1
    stride, = struct.unpack_from("<i", c, o)
    o += 4
    max_step, = struct.unpack_from("<Q", c, o)
    return dict(algo=algo, abs=absb, rel=relb, alpha=alpha, beta=beta, interp=interp, direction=direction, ibs=ibs,
                levelwise=lw, max_step=max_step)


@pytest.mark.parametrize("shape,rel", [((96, 96, 96), 1e-3), ((150, 170, 190), 1e-4), ((300, 400), 1e-3)])
def test_qoz_stream(engine, shape, rel):
    """faz -q: tuned Config in the trailer, engine buffer decodes on the oracle
    to the oracle's reconstruction with the oracle tuner's parameters."""
    x = _grf(shape, seed=77, fronts=2)
    rng = float(np.float32(x.max() - x.min()))
    eb = rel * rng
    print(f"[gpu] qoz stream {shape} rel={rel}", flush=True)
    s = engine.sz_compress_qoz(x, shape, rel, rel=True)
    o = oracle.qoz_tune(x, eb, rel, rng)
    t = _trailer(s)
    assert t["algo"] == 2 and t["abs"] == eb and t["rel"] == rel
    assert (t["alpha"], t["beta"], t["levelwise"], t["max_step"]) == (o["alpha"], o["beta"], o["levelwise"],
                                                                       o["max_step"])
    assert (t["interp"], t["direction"]) == (o["interp_algo"], o["direction"])
    p = oracle.make_params(shape, eb, algo=o["interp_algo"], direction=o["direction"], alpha=o["alpha"],
                           beta=o["beta"], max_step=o["max_step"], ops=tuple(o["ops"]), dirs=tuple(o["dirs"]))
    ref = oracle.interp_compress(x, p)["recon"]
    dec = oracle.decompress_stream(s, x.size)
    assert np.array_equal(dec.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(engine.sz_decompress(s, x.size).view(np.uint32), ref.view(np.uint32))


def test_compress_test_batch_matches_oracle(engine):
    """Batched CompressTest = one CompressTest per candidate, against the oracle."""
    import faz_amd
    x = _grf((130, 120, 110), seed=9, fronts=3)
    rng = float(np.float32(x.max() - x.min()))
    starts = np.array([[0, 0, 0], [64, 32, 16], [96, 80, 70]], np.int64)
    sbs = 32
    cands = [dict(eb=1e-3 * rng * s, alpha=a, beta=b) for a, b, s in
             [(1.0, 1.5, 1.0), (1.5, 3.0, 0.9), (2.0, 4.0, 1.25), (1.25, 2.0, 1.0)]]
    ops, dirs = (1, 0, 1, 1), (0, 5, 0, 5)
    ps = [faz_amd.make_params((33, 33, 33), c["eb"], alpha=c["alpha"], beta=c["beta"], max_step=32, ops=ops,
                              dirs=dirs) for c in cands]
    print("[gpu] batched CompressTest", flush=True)
    nb, sq = engine.compress_test_batch(torch.from_numpy(x.ravel()).cuda(), x.shape, starts, sbs, ps, rng)
    for c, b, s in zip(cands, nb, sq):
        p = oracle.make_params((33, 33, 33), c["eb"], alpha=c["alpha"], beta=c["beta"], max_step=32, ops=ops,
                               dirs=dirs)
        _, ob, osq = oracle.compress_test(x, starts, sbs, p, rng)
        assert (b, s) == (ob, osq)
