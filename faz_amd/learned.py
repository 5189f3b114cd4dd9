"""Learned tuner of the MI355X FAZ engine: data, training and evaluation.

The exhaustive QoZ tuner (faz -q) ranks alpha/beta candidates by running
CompressTest (tuning-mode-1 compression of the sample blocks, Huffman, zstd)
for each.  The learned tuner keeps the tuner's sampling and its per-level
interpolator choice and replaces those CompressTests by an MLP that predicts
the same (bitrate, PSNR) from statistics of the sample blocks; the decision
rule on top is the exhaustive tuner's.  The quantizer still enforces the
error bound exactly: only the configuration choice is learned.

Pipeline (all arithmetic in the engine's own kernels through the C-ABI):
  * fields: seeded k^-slope Gaussian random fields with tanh fronts
    (synthesised with torch.fft on the device),
  * tuples: for each field, the exhaustive tuner's own CompressTests plus a
    grid of CompressTests over every alpha/beta pair and bound scale, with
    the tuner's interpolators, -> (block statistics, candidate) -> (bitrate,
    PSNR),
  * training: faz_mlp_train_step (BF16 matrix-core forward/backward, Adam),
  * model: parameters + input standardisation in one .npz.
"""
import os
import time

import numpy as np

import faz_amd

ALPHAS = (1.0, 1.25, 1.5, 1.75, 2.0)
BETAS = (1.5, 2.0, 3.0, 4.0)
EB_SCALES = (0.8, 0.9, 1.0, 1.12, 1.25)
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "learned_tuner.npz")


def grf_field(shape, seed, slope=11.0 / 3.0, fronts=0, noise=1e-4, width=4.0, amp=0.5, grid=None):
    """k^-slope Gaussian random field in [0, 1] plus tanh fronts, float32 flat
    device tensor (synthetic input data).  With `grid` (a shape at least
    `shape` on every axis) the field is synthesised on that grid and a
    seeded sub-box is cut out, so fields of many shapes share one FFT plan."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(int(seed))
    full = tuple(grid) if grid else tuple(shape)
    ks = [torch.fft.fftfreq(n, device="cuda") for n in full[:-1]] + [torch.fft.rfftfreq(full[-1], device="cuda")]
    grids = torch.meshgrid(*ks, indexing="ij")
    k = torch.sqrt(sum(a * a for a in grids))
    k.view(-1)[0] = 1.0
    ampl = k.pow(-slope / 2.0)
    ampl.view(-1)[0] = 0.0
    spec = torch.complex(torch.randn(ampl.shape, generator=g, device="cuda") * ampl,
                         torch.randn(ampl.shape, generator=g, device="cuda") * ampl)
    f = torch.fft.irfftn(spec, s=full)
    if grid:
        off = [int(torch.randint(0, n - m + 1, (1,), generator=g, device="cuda").item()) for n, m in zip(full, shape)]
        f = f[tuple(slice(o, o + m) for o, m in zip(off, shape))].contiguous()
    f = (f - f.min()) / (f.max() - f.min())
    f += noise * torch.randn(shape, generator=g, device="cuda")
    if fronts:
        coords = torch.meshgrid(*[torch.arange(n, device="cuda", dtype=torch.float32) for n in shape], indexing="ij")
        for _ in range(fronts):
            nv = torch.randn(len(shape), generator=g, device="cuda")
            nv = nv / nv.norm()
            c = (torch.rand(len(shape), generator=g, device="cuda") *
                 torch.tensor(shape, device="cuda", dtype=torch.float32)).dot(nv)
            f += amp * torch.tanh((sum(coords[a] * nv[a] for a in range(len(shape))) - c) / width)
    out = f.to(torch.float32).contiguous().view(-1)
    torch.cuda.synchronize()
    return out


LORENZO_SCALES = (0.6, 0.8, 0.9, 1.0, 1.12, 1.25, 1.6)


def field_tuples(engine, x, shape, rel):
    """(rows, targets, info) of one field: the exhaustive tuner's tests (with
    the Lorenzo test), a CompressTest grid over ALPHAS x BETAS x EB_SCALES and
    the Lorenzo CompressTest at LORENZO_SCALES."""
    mn, mx = engine.range(x)
    rng = float(np.float32(mx - mn))
    eb = rel * rng
    ex = engine.tune_qoz(x, shape, eb, rel, rng, test_lorenzo=1)
    sbs, starts = ex["sample_block_size"], ex["block_starts"]
    eb_norm = eb / rng
    feats = engine.block_features(x, shape, starts, sbs, mn, rng, eb_norm)
    rows, targets, kinds = [], [], []

    def add(alpha, beta, eb_t, bitrate, psnr):
        lor = alpha < 0
        rows.append(faz_amd.learned_input(feats, eb_norm, rel, 0.0 if lor else alpha, 0.0 if lor else beta,
                                          eb_t / eb, ex["ops"], ex["dirs"], ex["level_res"], lor))
        targets.append(faz_amd.learned_target(bitrate, psnr, eb_norm, eb_t / eb))
        kinds.append((alpha, beta, eb_t / eb))

    for t in ex["tests"]:
        add(t["alpha"], t["beta"], t["eb"], t["bitrate"], t["psnr"])
    bd = (sbs + 1,) * len(shape)
    per = (sbs + 1) ** len(shape) * len(starts)
    grid = [(a, b, s) for a in ALPHAS for b in BETAS if a <= b for s in EB_SCALES]
    ps = [faz_amd.make_params(bd, eb * s, algo=ex["interp_algo"], direction=ex["direction"], alpha=a, beta=b,
                              max_step=ex["max_step"], ops=tuple(ex["ops"]), dirs=tuple(ex["dirs"])) for a, b, s in grid]
    nbytes, sq = engine.compress_test_batch(x, shape, starts, sbs, ps, rng)
    for (a, b, s), nbt, sqe in zip(grid, nbytes, sq):
        bitrate = 8.0 * float(nbt) / per
        psnr = 20 * np.log10(rng) - 10 * np.log10(sqe / per) if sqe > 0 else 200.0
        add(a, b, eb * s, bitrate, psnr)
    for s in LORENZO_SCALES:
        _, nbt, sqe = engine.compress_test_lorenzo(x, shape, starts, sbs, eb * s, want_raw=False)
        bitrate = 8.0 * float(nbt) / per
        psnr = 20 * np.log10(rng) - 10 * np.log10(sqe / per) if sqe > 0 else 200.0
        add(-1.0, -1.0, eb * s, bitrate, psnr)
    return np.array(rows, np.float32), np.array(targets, np.float32), dict(ex=ex, rng=rng, eb=eb, kinds=kinds)


TRAIN_GRIDS = (256, 320, 384, 512)  # cubic FFT grids the training fields are synthesised on


def walk_field(shape, seed, noise=0.0):
    """Random walk along every axis (cumulative sums of white noise), min-max
    normalised to [0, 1]: rough at the sample-block scale, where the Lorenzo
    predictor can beat interpolation (synthetic input data)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(int(seed))
    f = torch.randn(shape, generator=g, device="cuda")
    for a in range(len(shape)):
        f = torch.cumsum(f, dim=a)
    f = (f - f.min()) / (f.max() - f.min())
    if noise:
        f += noise * torch.randn(shape, generator=g, device="cuda")
    out = f.to(torch.float32).contiguous().view(-1)
    torch.cuda.synchronize()
    return out


def training_fields(n, seed=0):
    """Seeded draws for the training set: a field synthesised on one of
    TRAIN_GRIDS (so FFT plans are reused) and, for half of the draws, a
    sub-box of edges 0.5..1 of the grid cut out of it; slope k^-1..k^-5.5,
    0..16 fronts of width 2..8, noise 0..1e-2, REL 1e-6..1e-2 (log-uniform);
    15% are random walks instead (walk_field)."""
    r = np.random.default_rng(seed)
    out = []
    for i in range(n):
        g = int(r.choice(TRAIN_GRIDS))
        if r.random() < 0.5:
            shape = (g, g, g)
        else:
            shape = tuple(int(v) for v in r.integers(max(96, g // 2), g + 1, 3))
        out.append(dict(shape=shape, grid=(g, g, g), seed=int(r.integers(1 << 40)), slope=float(r.uniform(1.0, 5.5)),
                        fronts=int(r.integers(0, 17)), width=float(r.choice([2.0, 4.0, 8.0])),
                        noise=float(r.choice([0.0, 1e-5, 1e-4, 1e-4, 1e-3, 1e-2])), rel=float(10 ** r.uniform(-6, -2)),
                        kind="walk" if r.random() < 0.15 else "grf"))
    return out


def build_dataset(engine, specs, log=None, deadline=None):
    """Rows, targets and the field index of every tuple (fields after the
    time.time() deadline, if any, are left out)."""
    import torch
    X, T, G = [], [], []
    t0 = time.time()
    for i, sp in enumerate(specs):
        if deadline is not None and time.time() > deadline:
            if log:
                log(f"deadline: stopping after {i} fields")
            break
        grid = sp.get("grid")
        if sp.get("kind") == "walk":
            x = walk_field(sp["shape"], sp["seed"], noise=sp.get("noise", 0.0))
        else:
            x = grf_field(sp["shape"], sp["seed"], sp["slope"], sp["fronts"], noise=sp.get("noise", 1e-4),
                          width=sp.get("width", 4.0), grid=None if grid == sp["shape"] else grid)
        rows, targets, _ = field_tuples(engine, x, sp["shape"], sp["rel"])
        X.append(rows)
        T.append(targets)
        G.append(np.full(len(rows), i, np.int32))
        del x
        if log and (i + 1) % 100 == 0:
            log(f"fields {i + 1}/{len(specs)}: {sum(len(a) for a in X)} tuples, {time.time() - t0:.0f} s")
            torch.cuda.empty_cache()
    return np.concatenate(X), np.concatenate(T), np.concatenate(G)


def train(engine, X, T, epochs=40, batch=1024, lr=1e-3, seed=0, hidden=256, layers=3, log=None):
    """Trains a fresh MLP on standardised inputs; returns (mlp, mu, sigma,
    history).  The Lorenzo candidate's rows (a few per field) are repeated
    four times so the predictor choice weighs as much as the alpha/beta grid."""
    import torch
    lor = X[:, 53] > 0.5
    X = np.concatenate([X] + [X[lor]] * 3)
    T = np.concatenate([T] + [T[lor]] * 3)
    mu = X.mean(0)
    sigma = X.std(0)
    sigma[sigma < 1e-6] = 1.0
    Xs = ((X - mu) / sigma).astype(np.float32)
    net = faz_amd.MLP(engine, X.shape[1], hidden, layers, T.shape[1], seed)
    xd = torch.from_numpy(Xs).cuda()
    td = torch.from_numpy(np.ascontiguousarray(T, np.float32)).cuda()
    r = np.random.default_rng(seed)
    hist = []
    n = X.shape[0]
    for ep in range(epochs):
        perm = torch.from_numpy(r.permutation(n)).cuda()
        step_lr = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))
        tot = 0.0
        for s in range(0, n - batch + 1, batch):
            idx = perm[s:s + batch]
            tot += net.train_step(xd[idx].contiguous(), td[idx].contiguous(), step_lr)
        hist.append(tot / max(1, n // batch))
        if log:
            log(f"epoch {ep + 1}/{epochs} loss {hist[-1]:.5f}")
    return net, mu.astype(np.float32), sigma.astype(np.float32), hist


def save_model(path, net, mu, sigma, meta=None, gates=None):
    """Parameters, input standardisation and the confidence gates (lo / hi:
    standardised input range accepted without the exhaustive fallback;
    margin: predicted cost gap, in bits/value, below which the closest
    candidates are verified by real CompressTests)."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    g = gates or {}
    np.savez_compressed(path, params=net.get_params(), mu=mu, sigma=sigma, K=np.array(net.K), N=np.array(net.N),
                        in_dim=np.array([net.in_dim]), out_dim=np.array([net.out_dim]),
                        lo=np.asarray(g.get("lo", np.full(net.in_dim, -np.inf)), np.float32),
                        hi=np.asarray(g.get("hi", np.full(net.in_dim, np.inf)), np.float32),
                        margin=np.array([g.get("margin", -1.0)]), verify=np.array([g.get("verify", 4)]),
                        margin_lorenzo=np.array([g.get("margin_lorenzo", -1.0)]),
                        meta=np.array(repr(meta or {})))


def load_model(engine, path=MODEL_PATH, hidden=256, layers=3):
    z = np.load(path, allow_pickle=False)
    net = faz_amd.MLP(engine, int(z["in_dim"][0]), hidden, layers, int(z["out_dim"][0]), 0)
    if list(z["K"]) != net.K or list(z["N"]) != net.N:
        raise faz_amd.FazError(-1, f"model layout {list(z['K'])}x{list(z['N'])} does not match the network")
    net.set_params(z["params"])
    return net, z["mu"], z["sigma"]


def load_gates(path=MODEL_PATH):
    """Confidence gates stored with the model (see save_model)."""
    z = np.load(path, allow_pickle=False)
    ml = float(z["margin_lorenzo"][0]) if "margin_lorenzo" in z.files else -1.0
    return dict(lo=z["lo"], hi=z["hi"], margin=float(z["margin"][0]), verify=int(z["verify"][0]), margin_lorenzo=ml)


def tune_learned(engine, x, shape, eb, rel, rng, net, mu, sigma, gates=None, test_lorenzo=1):
    """The learned tuner with the model's gates (or none)."""
    g = gates or {}
    return engine.tune_learned(x, shape, eb, rel, rng, net, mu, sigma, test_lorenzo=test_lorenzo,
                               margin=g.get("margin", -1.0), verify=g.get("verify", 0), lo=g.get("lo"),
                               hi=g.get("hi"), margin_lorenzo=g.get("margin_lorenzo", -1.0))


def predict(net, mu, sigma, rows):
    """(bitrate, PSNR relative to the uniform-quantisation PSNR) of rows."""
    import torch
    y = net.forward(torch.from_numpy(((rows - mu) / sigma).astype(np.float32)).cuda()).cpu().numpy()
    return y


def reachable(X, max_interp_level=5):
    """Rows whose candidate the exhaustive search can reach (the Lorenzo
    test, and per alpha the betas up to the first one its stop rule ends on;
    3-D fields with maxStep 32 have max_interp_level 5)."""
    lor = X[:, 53] > 0.5
    a = X[:, 40].astype(np.float64) + 1.0
    b = X[:, 41].astype(np.float64) * 4.0
    ok = np.zeros(len(X), bool)
    for i in range(len(X)):
        if lor[i]:
            ok[i] = True
            continue
        first = None
        for bb in BETAS:
            if a[i] > bb:
                continue
            if a[i] ** (max_interp_level - 1) <= bb:
                first = bb
                break
        ok[i] = first is None or b[i] <= first + 1e-9
    return ok


def calibrate(net, mu, sigma, X, T, G, kinds_scale, quantile=0.9, tol=0.01, log=None, quantile_lorenzo=0.8):
    """Gates from held-out tuples.

    lo/hi: the standardised range of the rows widened by one unit.
    margin: per held-out field, the predicted cost gap between the predicted
    best and the truly best reachable base-bound candidate, counted only
    where taking the predicted best would cost more than `tol` of the best
    bitrate (elsewhere the prediction alone is good enough, gap 0); the
    margin is the `quantile` of these gaps.
    margin_lorenzo: the predicted cost difference between the Lorenzo
    candidate and the best interpolation candidate, the `quantile_lorenzo`
    (x 1.1) of its size over the held-out fields where its sign was wrong by
    more than `tol` of the best bitrate.  Cost = bitrate - PSNR / 6.02, bits
    per value at equal PSNR."""
    Z = (X - mu) / sigma
    lo = Z.min(0) - 1.0
    hi = Z.max(0) + 1.0
    y = predict(net, mu, sigma, X)
    reach = reachable(X)
    base = np.abs(kinds_scale - 1.0) < 1e-9
    lorm = X[:, 53] > 0.5
    gaps, regrets = [], []
    wrong, n_lor, n_lor_wrong = [], 0, 0
    for g in np.unique(G):
        m = (G == g) & base & reach & ~lorm
        if m.sum() < 2:
            continue
        # the PSNR offset is the same for every candidate of a field and cancels
        true = np.exp2(T[m, 0]) - 5.0 * T[m, 1] / 6.02
        pred = np.exp2(y[m, 0]) - 5.0 * y[m, 1] / 6.02
        br_best = float(np.exp2(T[m, 0])[np.argmin(true)])
        regret = float(true[np.argmin(pred)] - true.min()) / br_best
        regrets.append(regret)
        gaps.append(float(pred[np.argmin(true)] - pred.min()) if regret > tol else 0.0)
        ml = (G == g) & base & lorm
        if ml.any():
            n_lor += 1
            t_l = float((np.exp2(T[ml, 0]) - 5.0 * T[ml, 1] / 6.02).min())
            p_l = float((np.exp2(y[ml, 0]) - 5.0 * y[ml, 1] / 6.02).min())
            d_true, d_pred = t_l - true.min(), p_l - pred.min()
            if np.sign(d_true) != np.sign(d_pred) and abs(d_true) > tol * br_best:
                n_lor_wrong += 1
                wrong.append(abs(d_pred))
    margin = float(np.quantile(np.array(gaps), quantile)) if gaps else -1.0
    margin_lor = 1.1 * float(np.quantile(np.array(wrong), quantile_lorenzo)) if wrong else 0.0
    if log:
        r = np.array(regrets)
        log(f"calibration over {len(gaps)} fields: regret > {tol:.0%} in {(r > tol).mean():.1%} "
            f"(p50 {np.median(r):.4f}, p90 {np.quantile(r, 0.9):.4f}); predictor choice wrong by more than "
            f"{tol:.0%} in {n_lor_wrong} of {n_lor}")
    return dict(lo=lo.astype(np.float32), hi=hi.astype(np.float32), margin=margin, verify=3,
                margin_lorenzo=float(margin_lor)), np.array(gaps)


def final_stats(engine, x, shape, eb, params):
    """Compressed stream size and PSNR of the engine with `params` on the field."""
    import torch
    recon = torch.empty_like(x)
    raw = engine.interp_raw(x, params, recon=recon)
    mse = torch.mean((recon.double() - x.double()) ** 2).item()
    return len(raw), mse


def lorenzo_stats(engine, x, shape, eb):
    """Size and MSE of the 3-D Lorenzo engine the tuner's Lorenzo outcome
    selects (lorenzo + lorenzo2, blockSize 5, SZI:2800-2806)."""
    import torch
    p = faz_amd.make_lorenzo_params(shape, eb, block_size=5, pred_dim=3, lorenzo=True, lorenzo2=True,
                                    regression=False)
    recon = torch.empty_like(x)
    raw = engine.lorenzo_raw(x, p, recon=recon)
    mse = torch.mean((recon.double() - x.double()) ** 2).item()
    return len(raw), mse


def choice_params(shape, eb, r):
    return faz_amd.make_params(shape, eb, algo=r["interp_algo"], direction=r["direction"], alpha=r["alpha"],
                               beta=r["beta"], max_step=r["max_step"], ops=tuple(r["ops"]), dirs=tuple(r["dirs"]))


def evaluate(engine, net, mu, sigma, shape, seed, rel, slope=11.0 / 3.0, fronts=8, gates=None, test_lorenzo=1,
             noise=1e-4):
    """Exhaustive vs learned vs default (setABwithRelBound, no alpha/beta
    search) on one field: tuning times, engine buffer sizes, PSNRs."""
    import torch
    x = grf_field(shape, seed, slope, fronts, noise=noise)
    mn, mx = engine.range(x)
    rng = float(np.float32(mx - mn))
    eb = rel * rng
    out = {}
    for name in ("exhaustive", "learned"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if name == "exhaustive":
            r = engine.tune_qoz(x, shape, eb, rel, rng, test_lorenzo=test_lorenzo)
        else:
            r = tune_learned(engine, x, shape, eb, rel, rng, net, mu, sigma, gates, test_lorenzo)
        torch.cuda.synchronize()
        out[name] = dict(tune_ms=(time.perf_counter() - t0) * 1e3, alpha=r["alpha"], beta=r["beta"], ops=r["ops"],
                         dirs=r["dirs"], use_interp=r["use_interp"], learned_path=r["learned_path"],
                         n_verified=r["n_verified"], ex=r)
    ex = out["exhaustive"]["ex"]
    a0, b0 = default_ab(rel)
    out["default"] = dict(tune_ms=0.0, alpha=a0, beta=b0, ops=ex["ops"], dirs=ex["dirs"], use_interp=1)
    n = x.numel()
    for name, o in out.items():
        if o["use_interp"]:
            p = faz_amd.make_params(shape, eb, algo=ex["interp_algo"], direction=ex["direction"], alpha=o["alpha"],
                                    beta=o["beta"], max_step=ex["max_step"], ops=tuple(o["ops"]),
                                    dirs=tuple(o["dirs"]))
            nbytes, mse = final_stats(engine, x, shape, eb, p)
        else:
            nbytes, mse = lorenzo_stats(engine, x, shape, eb)
        o["bitrate"] = 8.0 * nbytes / n
        o["psnr"] = 20 * np.log10(rng) - 10 * np.log10(mse)
        o.pop("ex", None)
    del x
    return out


def default_ab(rel):
    """setABwithRelBound, configuration 0 (include/QoZ/api/impl/SZInterp.hpp:1394-1414)."""
    if rel >= 0.01:
        return 2.0, 2.0
    if rel >= 0.007:
        return 1.75, 2.0
    if rel >= 0.004:
        return 1.5, 2.0
    if rel > 0.001:
        return 1.25, 1.5
    return 1.0, 1.0


def ratio_at_equal_psnr(bitrate, psnr, ref_psnr, bits=32.0, slope_db_per_bit=6.02):
    """Compression ratio shifted to the reference PSNR along the high-rate
    rate-distortion slope (6.02 dB per bit)."""
    return bits / (bitrate + (ref_psnr - psnr) / slope_db_per_bit)
