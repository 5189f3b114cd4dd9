"""GPU parity of the default (SZ3) tuning mode (`faz` without -q; needs an MI355X).

Bar: the sample, every block interpolation buffer and the quantization
interval count equal the reference's (golden fixtures); the tuning decisions,
every intermediate size and ratio, and the final stream equal the oracle's on
fields where interpolation and where Lorenzo wins (3-D fast frontend, 1-D and
2-D general frontend, optimize_quant_invl_3d, the 16384-bin test).
"""
import numpy as np
import pytest

import oracle
from golden_cases import SZ3, load

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32).ravel()).cuda()


def _raws(c):
    off = np.concatenate([[0], np.cumsum(c["raw_lens"])])
    return [c["raws"][off[i]:off[i + 1]].tobytes() for i in range(len(c["raw_lens"]))]


@pytest.mark.parametrize("name", SZ3)
def test_golden_sampling_block_interp_quant_intervals(engine, name):
    c = load(name)
    p = c["params"]
    x = c["input"]
    print(f"[gpu] sz3 pieces {name}", flush=True)
    s, sdims, sb = engine.sampling(_dev(x), x.shape)
    assert sb == p["sampling_block"] and (s is not None) == bool(p["sampled"])
    if s is not None:
        assert list(sdims) == p["sample_dims"]
        assert np.array_equal(s.cpu().numpy().view(np.uint32), c["sample"].view(np.uint32))
    sample = _dev(c["sample"])
    for (algo, d), raw in zip(c["keys"], _raws(c)):
        ours, nbytes = engine.block_interp_test(sample, p["sample_dims"], p["eb"], int(algo), int(d), p["block"])
        assert ours == raw, (algo, d)
        _, obytes = oracle.block_interp_test(c["sample"].reshape(p["sample_dims"]), p["eb"], int(algo), int(d),
                                             p["block"])
        assert nbytes == obytes
    if x.ndim == 3:
        for f, q in zip((1.0, 0.01), c["quantbin"]):
            assert engine.quant_intervals(_dev(x), x.shape, p["eb"] * f)[0] == int(q)


def _field(shape, kind, seed):
    r = np.random.default_rng(seed)
    x = r.standard_normal(shape)
    if kind == "walk1":
        for ax in range(len(shape)):
            x = np.cumsum(x, axis=ax)
    else:
        for ax in range(len(shape)):
            x = np.cumsum(np.cumsum(x, axis=ax), axis=ax)
        if kind == "smooth":
            x = (x - x.min()) / (x.max() - x.min()) + 0.001 * r.standard_normal(shape)
    return x.astype(np.float32)


CASES = [
    ((64, 80, 72), "walk1", 1e-3),    # Lorenzo, 3-D fast frontend, 32 bins
    ((100, 100, 100), "walk2", 1e-4),  # Lorenzo, cubic direction 5 tested
    ((64, 80, 72), "walk2", 1e-7),    # Lorenzo, 16384-bin test
    ((400, 500), "walk1", 1e-3),      # Lorenzo, 2-D general frontend
    ((20000,), "walk1", 1e-3),        # Lorenzo, 1-D general frontend
    ((400, 500), "walk2", 1e-3),      # interpolation
    ((100, 100, 100), "smooth", 1e-3),  # interpolation
    ((40, 40, 40), "smooth", 1e-3),   # too small to sample
]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_tune_and_stream_against_oracle(engine, i):
    shape, kind, rel = CASES[i]
    x = _field(shape, kind, 60 + i)
    eb = rel * float(np.float32(x.max() - x.min()))
    print(f"[gpu] sz3 tune {shape} {kind} rel={rel}", flush=True)
    g = engine.sz3_tune(_dev(x), shape, eb, rel)
    o = oracle.sz3_tune(x, eb, rel)
    assert g == o
    s = engine.sz_compress_sz3(x, shape, rel, rel=True)
    ref, _ = oracle.sz3_compress_stream(x, rel, rel=True)
    assert s == ref
    dec = engine.sz_decompress(s, x.size)
    assert np.array_equal(dec.view(np.uint32), oracle.decompress_stream(s, x.size).view(np.uint32))
    assert np.abs(dec.astype(np.float64) - x.ravel().astype(np.float64)).max() <= eb


def test_abs_mode_stream(engine):
    x = _field((64, 80, 72), "walk1", 5)
    eb = 0.01
    print("[gpu] sz3 ABS stream", flush=True)
    s = engine.sz_compress_sz3(x, x.shape, eb, rel=False)
    ref, _ = oracle.sz3_compress_stream(x, eb, rel=False)
    assert s == ref


def test_rejects_wide_bins(engine):
    """optimize_quant_invl_3d may ask for 131072 bins (radius above 16-bit
    codes): rejected loudly."""
    import faz_amd
    x = _field((100, 100, 100), "walk1", 60)
    rel = 1e-7
    o = oracle.sz3_tune(x, rel * float(np.float32(x.max() - x.min())), rel)
    if o["use_interp"] or o["quantbin_estimate"] <= 65536:
        pytest.skip("this field does not ask for more than 65536 bins")
    with pytest.raises(faz_amd.FazError):
        engine.sz_compress_sz3(x, x.shape, rel, rel=True)
