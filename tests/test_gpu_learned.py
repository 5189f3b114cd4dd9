"""The learned tuner on held-out fields (needs an MI355X and the committed
model faz_amd/models/learned_tuner.npz, trained by faz_amd/train_learned.py).

What the trained network gains over the method without it:
* its bitrate and PSNR predictions of CompressTests are accurate where an
  untrained network of the same shape is not;
* on held-out fields where the exhaustive tuner's alpha/beta beat the
  untuned default (setABwithRelBound), the learned choice keeps the
  exhaustive ratio within 2% at equal PSNR and beats the default, several
  times faster than the exhaustive search;
* the confidence gates: an input outside the training range falls back to
  the exhaustive search (same decisions bit for bit), close calls are
  decided by real CompressTests;
* with the Lorenzo test on, the learned choice of predictor (interpolation
  or Lorenzo) compresses as well as the exhaustive one.
The block statistics the network reads are checked against their CPU
restatement (oracle/features_ref.py).
"""
import math

import numpy as np
import pytest

import oracle
from oracle.features_ref import block_features as features_ref

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (shape, rel, slope, fronts, seed): fields where the exhaustive choice beats
# the untuned default
DISCRIMINATING = [((384, 384, 384), 1e-2, 4.12, 11, 2882111006751),
                  ((384, 384, 384), 1e-2, 2.02, 2, 2289755631944),
                  ((384, 384, 384), 1e-3, 4.4, 12, (1 << 43) + 4400 + 12)]
PREDICT = [((256, 256, 256), 1e-3, 2.3, 3), ((256, 256, 256), 1e-4, 4.1, 9), ((192, 256, 320), 5e-4, 3.2, 0),
           ((256, 256, 256), 1e-2, 3.8, 12), ((256, 256, 256), 1e-5, 2.8, 4)]


@pytest.fixture(scope="module")
def model(engine):
    from faz_amd import learned
    net, mu, sigma = learned.load_model(engine)
    yield net, mu, sigma, learned.load_gates()
    net.close()


def test_block_features_match_the_cpu_restatement(engine):
    r = np.random.default_rng(3)
    for shape, sbs in (((70, 80, 90), 32), ((140, 150), 64), ((50, 40, 60), 16)):
        x = r.standard_normal(shape)
        for ax in range(len(shape)):
            x = np.cumsum(x, axis=ax)
        x = x.astype(np.float32)
        vmin, rng = float(x.min()), float(np.float32(x.max() - x.min()))
        starts = np.array([[1, 2, 3][:len(shape)] + [0] * (3 - len(shape)),
                           [n - sbs - 1 for n in shape] + [0] * (3 - len(shape))], np.int64)
        print(f"[gpu] block features {shape} sbs={sbs}", flush=True)
        g = engine.block_features(torch.from_numpy(x.ravel()).cuda(), shape, starts, sbs, vmin, rng, 1e-3)
        for b, s0 in enumerate(starts):
            blk = x[tuple(slice(s0[a], s0[a] + sbs + 1) for a in range(len(shape)))]
            want = features_ref(blk, vmin, rng, 1e-3)
            np.testing.assert_allclose(g[b], want, rtol=2e-5, atol=1e-6)


def _predict(engine, net, mu, sigma, x, shape, rel):
    """Exhaustive tuner tests on the field and the network's predictions of them."""
    import faz_amd
    mn, mx = engine.range(x)
    rng = float(np.float32(mx - mn))
    eb = rel * rng
    ex = engine.tune_qoz(x, shape, eb, rel, rng, test_lorenzo=1)
    feats = engine.block_features(x, shape, ex["block_starts"], ex["sample_block_size"], mn, rng, eb / rng)
    rows = np.stack([faz_amd.learned_input(feats, eb / rng, rel, max(t["alpha"], 0.0), max(t["beta"], 0.0),
                                           t["eb"] / eb, ex["ops"], ex["dirs"], ex["level_res"], t["alpha"] < 0)
                     for t in ex["tests"]])
    y = net.forward(torch.from_numpy((rows - mu) / sigma).float().cuda()).cpu().numpy()
    scale = np.array([t["eb"] / eb for t in ex["tests"]])
    br = np.exp2(y[:, 0])
    ps = 5.0 * y[:, 1] - 20 * np.log10(eb / rng * scale)
    true_br = np.array([t["bitrate"] for t in ex["tests"]])
    true_ps = np.array([t["psnr"] for t in ex["tests"]])
    return np.abs(br - true_br) / true_br, np.abs(ps - true_ps)


def test_trained_network_predicts_compress_tests(engine, model):
    import faz_amd
    from faz_amd import learned
    net, mu, sigma, _ = model
    untrained = faz_amd.MLP(engine, faz_amd.LEARNED_INPUTS, 256, 3, 2, seed=12345)
    errs, errs_u = [], []
    for shape, rel, slope, fronts in PREDICT:
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
    assert br.mean() < 0.12 and ps.mean() < 0.5
    assert br_u.mean() > 4 * br.mean() and ps_u.mean() > 4 * ps.mean()
    untrained.close()


@pytest.mark.parametrize("i", range(len(DISCRIMINATING)))
def test_learned_choice_where_alpha_beta_matter(engine, model, i):
    from faz_amd import learned
    net, mu, sigma, gates = model
    shape, rel, slope, fronts, seed = DISCRIMINATING[i]
    print(f"[gpu] learned vs exhaustive vs default {shape} rel={rel} slope={slope} fronts={fronts}", flush=True)
    r = learned.evaluate(engine, net, mu, sigma, shape, seed, rel, slope=slope, fronts=fronts, gates=gates)
    ex, le, de = r["exhaustive"], r["learned"], r["default"]
    cr_e = 32.0 / ex["bitrate"]
    cr_l = learned.ratio_at_equal_psnr(le["bitrate"], le["psnr"], ex["psnr"])
    cr_d = learned.ratio_at_equal_psnr(de["bitrate"], de["psnr"], ex["psnr"])
    print(f"  exhaustive ({ex['alpha']},{ex['beta']}) {ex['tune_ms']:.1f} ms CR {cr_e:.3f}; learned "
          f"({le['alpha']},{le['beta']}) path {le['learned_path']} verified {le['n_verified']} "
          f"{le['tune_ms']:.1f} ms CR@PSNR {cr_l:.3f}; default ({de['alpha']},{de['beta']}) CR@PSNR {cr_d:.3f}",
          flush=True)
    assert cr_e > 1.02 * cr_d  # the search matters on this field
    assert cr_l >= 0.98 * cr_e
    assert cr_l > cr_d
    assert le["tune_ms"] * 3 < ex["tune_ms"]


def test_gate_out_of_range_falls_back_to_the_exhaustive_search(engine, model):
    from faz_amd import learned
    net, mu, sigma, gates = model
    shape, rel = (256, 256, 256), 1e-3
    x = learned.grf_field(shape, 77, 3.5, 6)
    mn, mx = engine.range(x)
    rng = float(np.float32(mx - mn))
    ex = engine.tune_qoz(x, shape, rel * rng, rel, rng, test_lorenzo=1)
    closed = dict(gates, lo=np.zeros_like(gates["lo"]), hi=np.zeros_like(gates["hi"]))
    print("[gpu] learned tuner with every input outside the gate", flush=True)
    le = learned.tune_learned(engine, x, shape, rel * rng, rel, rng, net, mu, sigma, closed)
    assert le["learned_path"] == 2
    for k in ("alpha", "beta", "ops", "dirs", "use_interp", "levelwise"):
        assert le[k] == ex[k], k
    assert [(t["alpha"], t["beta"], t["eb"], t["bytes"]) for t in le["tests"]] == \
           [(t["alpha"], t["beta"], t["eb"], t["bytes"]) for t in ex["tests"]]
    print("[gpu] learned tuner verifying every reachable candidate", flush=True)
    wide = dict(gates, margin=1e9, verify=64, margin_lorenzo=1e9)
    lv = learned.tune_learned(engine, x, shape, rel * rng, rel, rng, net, mu, sigma, wide)
    assert lv["learned_path"] == 1 and lv["n_verified"] >= 9
    real = [t for t in lv["tests"] if t["bytes"] > 0]
    cost = [t["bitrate"] - t["psnr"] / 6.02 for t in real]
    best = real[int(np.argmin(cost))]
    if lv["use_interp"]:
        assert (lv["alpha"], lv["beta"]) == (best["alpha"], best["beta"])
    else:
        assert best["alpha"] < 0
    print("[gpu] learned tuner without verification", flush=True)
    ln = learned.tune_learned(engine, x, shape, rel * rng, rel, rng, net, mu, sigma, dict(gates, margin=-1.0, margin_lorenzo=-1.0))
    assert ln["learned_path"] == 0 and ln["n_verified"] == 0


def _walk(shape, seed):
    x = np.random.default_rng(seed).standard_normal(shape)
    for ax in range(len(shape)):
        x = np.cumsum(x, axis=ax)
    return (x / np.abs(x).max()).astype(np.float32)


@pytest.mark.parametrize("rel", [1e-3, 1e-2])
def test_learned_predictor_choice_with_the_lorenzo_test(engine, model, rel):
    """A rough field where the Lorenzo test can win: the learned choice of
    predictor compresses within 2% of the exhaustive choice at equal PSNR."""
    from faz_amd import learned
    net, mu, sigma, gates = model
    shape = (256, 256, 256)
    x = torch.from_numpy(_walk(shape, 8).ravel()).cuda()
    mn, mx = engine.range(x)
    rng = float(np.float32(mx - mn))
    eb = rel * rng
    print(f"[gpu] predictor choice on a random walk, rel={rel}", flush=True)
    ex = engine.tune_qoz(x, shape, eb, rel, rng, test_lorenzo=1)
    le = learned.tune_learned(engine, x, shape, eb, rel, rng, net, mu, sigma, gates)
    res = {}
    for name, t in (("exhaustive", ex), ("learned", le)):
        if t["use_interp"]:
            nbytes, mse = learned.final_stats(engine, x, shape, eb, learned.choice_params(shape, eb, t))
        else:
            nbytes, mse = learned.lorenzo_stats(engine, x, shape, eb)
        res[name] = (8.0 * nbytes / x.numel(), 20 * math.log10(rng) - 10 * math.log10(mse))
    cr_e = 32.0 / res["exhaustive"][0]
    cr_l = learned.ratio_at_equal_psnr(res["learned"][0], res["learned"][1], res["exhaustive"][1])
    print(f"  exhaustive use_interp={ex['use_interp']} CR {cr_e:.3f}; learned use_interp={le['use_interp']} "
          f"path {le['learned_path']} CR@PSNR {cr_l:.3f}", flush=True)
    assert cr_l >= 0.98 * cr_e
