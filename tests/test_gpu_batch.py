"""GPU batches of uneven fields (faz_gpu_batch_*; needs an MI355X).

Bar: every field's stream from the batch is byte-identical to a single-field
SZ_compress call on the same data (and so decodes to the oracle's
reconstruction); replays after refilling the bound inputs stay identical;
ABS bounds replay the captured graph, changed REL ranges re-capture and update
it; every launch schedule the autotuner may pick writes the same buffer.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _field(shape, seed, zeros=False):
    r = np.random.default_rng(seed)
    x = r.standard_normal(shape)
    for ax in range(len(shape)):
        x = np.cumsum(np.cumsum(x, axis=ax), axis=ax)
    x = (x - x.min()) / (x.max() - x.min() + 1e-30)
    x += 1e-4 * r.standard_normal(shape)
    if zeros:  # masked regions of exact zeros
        m = np.zeros(shape, bool)
        for _ in range(3):
            sl = tuple(slice(int(r.integers(0, max(1, n // 2))), int(r.integers(n // 2, n + 1))) for n in shape)
            m[sl] = True
        x[m] = 0.0
    return x.astype(np.float32)


SHAPES = [((90, 180), dict(max_step=64)), ((30, 50, 50), dict(max_step=32)), ((64, 48, 48), dict(max_step=32)),
          ((33, 65, 17), dict()), ((5000,), dict()), ((10, 12, 9, 8), dict()), ((1, 300), dict()),
          ((70, 66, 68), dict(max_step=32, algo=0))]


def _batch(engine, seed, autotune=True):
    import faz_amd
    xs = [_field(s, seed + i, zeros=(i == 2)) for i, (s, _) in enumerate(SHAPES)]
    ts = [torch.from_numpy(x.ravel()).cuda() for x in xs]
    ps = [faz_amd.make_params(s, 0.0, **kw) for s, kw in SHAPES]
    return xs, ts, ps, faz_amd.Batch(engine, ps, ts, autotune=autotune)


def test_batch_streams_equal_single_calls(engine):
    xs, ts, ps, b = _batch(engine, 100)
    try:
        print("[gpu] batch of", len(xs), "fields, REL 1e-3", flush=True)
        streams = b.sz_compress(rel=True, eb=1e-3)
        for x, p, s in zip(xs, ps, streams):
            ref = engine.sz_compress(x, p, 1e-3, rel=True)
            assert s == ref
            dec = oracle.decompress_stream(s, x.size)
            assert np.array_equal(engine.sz_decompress(s, x.size).view(np.uint32), dec.view(np.uint32))
            eb = 1e-3 * float(np.float32(x.max() - x.min()))
            assert np.abs(dec.astype(np.float64) - x.ravel().astype(np.float64)).max() <= eb
        st = b.stats()
        assert st["captures"] == 1 and st["tuned_fields"] == sum(len(s) == 3 for s, _ in SHAPES)
        # refill the bound inputs in place: new ranges -> re-capture + in-place update
        print("[gpu] batch refill, REL 1e-3", flush=True)
        xs2 = [_field(x.shape, 900 + i) * np.float32(1.5) for i, x in enumerate(xs)]
        for t, x in zip(ts, xs2):
            t.copy_(torch.from_numpy(x.ravel()).cuda())
        streams2 = b.sz_compress(rel=True, eb=1e-3)
        for x, p, s in zip(xs2, ps, streams2):
            assert s == engine.sz_compress(x, p, 1e-3, rel=True)
        st = b.stats()
        assert st["captures"] == 2 and st["updates"] + st["instantiations"] == 2
        # absolute bounds: the second call replays the graph as captured
        print("[gpu] batch ABS replay", flush=True)
        ebs = [2e-3] * len(xs2)
        a1 = b.sz_compress(rel=False, eb=ebs)
        a2 = b.sz_compress(rel=False, eb=ebs)
        assert a1 == a2
        for x, p, s in zip(xs2, ps, a1):
            assert s == engine.sz_compress(x, p, 2e-3, rel=False)
        assert b.stats()["replays"] >= 1
    finally:
        b.close()


def test_batch_schedules_agree(engine):
    """Whatever schedule the autotuner picks, the engine buffers agree with
    the untuned batch's."""
    xs, ts, ps, b0 = _batch(engine, 300, autotune=False)
    try:
        print("[gpu] untuned batch", flush=True)
        raw0, _, _ = b0.compress_raw(rel=True, eb=1e-4)
    finally:
        b0.close()
    import faz_amd
    b1 = faz_amd.Batch(engine, ps, ts, autotune=True)
    try:
        print("[gpu] tuned batch", flush=True)
        raw1, _, _ = b1.compress_raw(rel=True, eb=1e-4)
        assert raw0 == raw1
        print("schedules", b1.stats()["schedules"], flush=True)
    finally:
        b1.close()
