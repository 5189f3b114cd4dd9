"""The learned tuner on held-out fields (needs an MI355X and the committed
model faz_amd/models/learned_tuner.npz, trained by faz_amd/train_learned.py).

What the trained network gains: its predictions of the exhaustive tuner's
CompressTest bitrates and PSNRs are accurate where an untrained network of
the same shape is not; its alpha/beta choices give the exhaustive tuner's
compression ratio (within 2% at equal PSNR) several times faster.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HELD_OUT = [((256, 256, 256), 1e-3, 2.3, 3), ((256, 256, 256), 1e-4, 4.1, 9), ((192, 256, 320), 5e-4, 3.2, 0),
            ((256, 256, 256), 2e-3, 2.8, 12)]


@pytest.fixture(scope="module")
def model(engine):
    from faz_amd import learned
    net, mu, sigma = learned.load_model(engine)
    yield net, mu, sigma
    net.close()


def _predict(engine, net, mu, sigma, x, shape, rel):
    """Exhaustive tuner tests on the field and the network's predictions of them."""
    import faz_amd
    mn, mx = engine.range(x)
    rng = float(np.float32(mx - mn))
    eb = rel * rng
    ex = engine.tune_qoz(x, shape, eb, rel, rng)
    feats = engine.block_features(x, shape, ex["block_starts"], ex["sample_block_size"], mn, rng, eb / rng)
    rows = np.stack([faz_amd.learned_input(feats, eb / rng, rel, t["alpha"], t["beta"], t["eb"] / eb, ex["ops"],
                                           ex["dirs"]) for t in ex["tests"]])
    y = net.forward(torch.from_numpy((rows - mu) / sigma).float().cuda()).cpu().numpy()
    scale = np.array([t["eb"] / eb for t in ex["tests"]])
    br = 4.0 * y[:, 0]
    ps = 5.0 * y[:, 1] - 20 * np.log10(eb / rng * scale)
    true_br = np.array([t["bitrate"] for t in ex["tests"]])
    true_ps = np.array([t["psnr"] for t in ex["tests"]])
    return np.abs(br - true_br) / true_br, np.abs(ps - true_ps)


def test_trained_network_predicts_compress_tests(engine, model):
    import faz_amd
    from faz_amd import learned
    net, mu, sigma = model
    untrained = faz_amd.MLP(engine, faz_amd.LEARNED_INPUTS, 256, 3, 2, seed=12345)
    errs, errs_u = [], []
    for shape, rel, slope, fronts in HELD_OUT:
        x = learned.grf_field(shape, (1 << 42) + int(slope * 100), slope, fronts)
        print(f"[gpu] learned predictions {shape} rel={rel} slope={slope}", flush=True)
        errs.append(_predict(engine, net, mu, sigma, x, shape, rel))
        errs_u.append(_predict(engine, untrained, mu, sigma, x, shape, rel))
    br = np.concatenate([e[0] for e in errs])
    ps = np.concatenate([e[1] for e in errs])
    br_u = np.concatenate([e[0] for e in errs_u])
    ps_u = np.concatenate([e[1] for e in errs_u])
    print(f"bitrate rel. error trained {br.mean():.4f} untrained {br_u.mean():.4f}; "
          f"PSNR error trained {ps.mean():.3f} dB untrained {ps_u.mean():.3f} dB", flush=True)
    assert br.mean() < 0.05 and ps.mean() < 0.5
    assert br_u.mean() > 5 * br.mean() and ps_u.mean() > 5 * ps.mean()
    untrained.close()


def test_learned_choice_matches_exhaustive_ratio(engine, model):
    from faz_amd import learned
    net, mu, sigma = model
    for shape, rel, slope, fronts in HELD_OUT:
        print(f"[gpu] learned vs exhaustive {shape} rel={rel} slope={slope}", flush=True)
        r = learned.evaluate(engine, net, mu, sigma, shape, (1 << 42) + 7 + int(slope * 100), rel, slope=slope,
                             fronts=fronts)
        ex, le = r["exhaustive"], r["learned"]
        cr_e = 32.0 / ex["bitrate"]
        cr_l = learned.ratio_at_equal_psnr(le["bitrate"], le["psnr"], ex["psnr"])
        print(f"  exhaustive ({ex['alpha']},{ex['beta']}) {ex['tune_ms']:.1f} ms CR {cr_e:.3f}; "
              f"learned ({le['alpha']},{le['beta']}) {le['tune_ms']:.1f} ms CR@PSNR {cr_l:.3f}", flush=True)
        assert cr_l >= 0.98 * cr_e
        assert le["tune_ms"] * 3 < ex["tune_ms"]
