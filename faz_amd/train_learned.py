"""Trains the learned tuner's MLP on tuples from the exhaustive QoZ tuner,
calibrates its confidence gates on held-out fields and evaluates it on
held-out 384^3 fields.

    python faz_amd/train_learned.py --fields 3000 --out faz_amd/models/learned_tuner.npz

Writes the model (.npz, with the gates) and a JSON report of the held-out
evaluation next to it.  Needs the GPU.
"""
import argparse
import ast
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import faz_amd  # noqa: E402
from faz_amd import learned  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fields", type=int, default=3000)
    ap.add_argument("--val-fields", type=int, default=300)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--out", default=learned.MODEL_PATH)
    ap.add_argument("--eval-fields", type=int, default=4)
    ap.add_argument("--eval-shape", type=int, default=384)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--budget-s", type=float, default=0, help="stop generating tuples after this many seconds")
    ap.add_argument("--recalibrate", default="", help="keep this model's network; rebuild only the validation "
                                                      "fields and its gates")
    ap.add_argument("--quantile", type=float, default=0.9)
    ap.add_argument("--tol", type=float, default=0.01,
                    help="bitrate loss (relative) below which a misprediction needs no verification")
    ap.add_argument("--quantile-lorenzo", type=float, default=0.8,
                    help="share of the held-out predictor-choice errors the Lorenzo margin covers")
    ap.add_argument("--dump-val", default="", help="also write the validation tuples (X, T, G) to this .npz")
    a = ap.parse_args()

    def log(m):
        print(m, flush=True)

    e = faz_amd.Engine(0)
    t0 = time.time()
    if a.recalibrate:
        specs = learned.training_fields(a.fields + a.val_fields, a.seed)[a.fields:]
        X, T, G = learned.build_dataset(e, specs, log=log)
        net, mu, sigma = learned.load_model(e, a.recalibrate)
        old = learned.load_gates(a.recalibrate)
        scale = np.exp2(X[:, 48].astype(np.float64))
        if a.dump_val:
            np.savez_compressed(a.dump_val, X=X, T=T, G=G)
        gates, gaps = learned.calibrate(net, mu, sigma, X, T, G, scale, a.quantile, a.tol, log, a.quantile_lorenzo)
        gates["lo"], gates["hi"] = old["lo"], old["hi"]
        z = np.load(a.recalibrate, allow_pickle=False)
        meta = ast.literal_eval(str(z["meta"])) if str(z["meta"]).startswith("{") else {}
        meta["recalibrated"] = dict(val_fields=a.val_fields, quantile=a.quantile, tol=a.tol,
                                    quantile_lorenzo=a.quantile_lorenzo)
        learned.save_model(a.out, net, mu, sigma, meta, gates)
        log(f"recalibrated: margin {gates['margin']:.4f} bits/value (gap p50 {np.median(gaps):.4f}, "
            f"p{int(100 * a.quantile)} {gates['margin']:.4f}); Lorenzo margin {gates['margin_lorenzo']:.4f}")
        report = dict(recalibrated=True, margin=gates["margin"], results=[])
        evaluate_heldout(e, net, mu, sigma, gates, a, report, log)
        return
    specs = learned.training_fields(a.fields + a.val_fields, a.seed)
    # validation fields first, so a time budget cuts training fields only
    order = specs[a.fields:] + specs[:a.fields]
    X, T, G = learned.build_dataset(e, order, log=log, deadline=t0 + a.budget_s if a.budget_s > 0 else None)
    log(f"dataset: {X.shape[0]} tuples from {int(G.max()) + 1} fields in {time.time() - t0:.0f} s")
    train_m = G >= a.val_fields
    t1 = time.time()
    net, mu, sigma, hist = learned.train(e, X[train_m], T[train_m], epochs=a.epochs, seed=a.seed, log=log)
    log(f"training: {time.time() - t1:.1f} s, final loss {hist[-1]:.5f}")
    # kinds: eb scale of every tuple (input 48 = log2 scale)
    scale = np.exp2(X[:, 48].astype(np.float64))
    gates, gaps = learned.calibrate(net, mu, sigma, X[~train_m], T[~train_m], G[~train_m], scale[~train_m],
                                    a.quantile, a.tol, log, a.quantile_lorenzo)
    if a.dump_val:
        np.savez_compressed(a.dump_val, X=X[~train_m], T=T[~train_m], G=G[~train_m])
    Ztrain = (X[train_m] - mu) / sigma
    gates["lo"] = (Ztrain.min(0) - 1.0).astype(np.float32)
    gates["hi"] = (Ztrain.max(0) + 1.0).astype(np.float32)
    y = learned.predict(net, mu, sigma, X[~train_m])
    br_err = np.abs(np.exp2(y[:, 0]) / np.exp2(T[~train_m, 0]) - 1)
    ps_err = 5 * np.abs(y[:, 1] - T[~train_m, 1])
    log(f"validation: bitrate rel. error mean {br_err.mean():.4f} p90 {np.quantile(br_err, 0.9):.4f}; PSNR error "
        f"mean {ps_err.mean():.3f} dB; margin {gates['margin']:.4f} bits/value (gap p50 {np.median(gaps):.4f}); "
        f"Lorenzo margin {gates['margin_lorenzo']:.4f}")
    n_train = int(G.max()) + 1 - a.val_fields
    learned.save_model(a.out, net, mu, sigma, dict(fields=n_train, val_fields=a.val_fields, tuples=int(X.shape[0]),
                                                   epochs=a.epochs, seed=a.seed, final_loss=float(hist[-1])), gates)
    report = dict(fields=n_train, val_fields=a.val_fields, tuples=int(X.shape[0]), epochs=a.epochs, loss=hist,
                  validation=dict(bitrate_rel_err=float(br_err.mean()), psnr_err_db=float(ps_err.mean()),
                                  margin=gates["margin"]), results=[])
    evaluate_heldout(e, net, mu, sigma, gates, a, report, log)


def evaluate_heldout(e, net, mu, sigma, gates, a, report, log):
    r = np.random.default_rng(a.seed + 1)
    shape = (a.eval_shape,) * 3
    for rel in (1e-2, 1e-3, 1e-4):
        for k in range(a.eval_fields):
            slope = float(r.uniform(2.0, 5.0))
            fronts = int(r.integers(0, 17))
            seed = int(r.integers(1 << 40)) | (1 << 41)  # disjoint from the training draws
            res = learned.evaluate(e, net, mu, sigma, shape, seed, rel, slope=slope, fronts=fronts, gates=gates)
            ex, le, de = res["exhaustive"], res["learned"], res["default"]
            cr_e = 32.0 / ex["bitrate"]
            cr_l = learned.ratio_at_equal_psnr(le["bitrate"], le["psnr"], ex["psnr"])
            cr_d = learned.ratio_at_equal_psnr(de["bitrate"], de["psnr"], ex["psnr"])
            row = dict(rel=rel, slope=slope, fronts=fronts, seed=seed, exhaustive=ex, learned=le, default=de,
                       ratio_exhaustive=cr_e, ratio_learned_at_equal_psnr=cr_l, ratio_default_at_equal_psnr=cr_d,
                       speedup=ex["tune_ms"] / max(le["tune_ms"], 1e-9))
            report["results"].append(row)
            log(f"rel={rel} slope={slope:.2f} fronts={fronts}: exhaustive ({ex['alpha']},{ex['beta']}) "
                f"{ex['tune_ms']:.1f} ms CR {cr_e:.3f} | learned ({le['alpha']},{le['beta']}) path "
                f"{le['learned_path']} verified {le['n_verified']} {le['tune_ms']:.1f} ms CR@PSNR {cr_l:.3f} | "
                f"default ({de['alpha']},{de['beta']}) CR@PSNR {cr_d:.3f}")
    with open(os.path.splitext(a.out)[0] + "_eval.json", "w") as f:
        json.dump(report, f, indent=1, default=float)
    log("done")


if __name__ == "__main__":
    main()
