"""Trains the learned tuner's MLP on tuples from the exhaustive QoZ tuner and
evaluates it on held-out 384^3 fields.

    python faz_amd/train_learned.py --fields 2000 --out faz_amd/models/learned_tuner.npz

Writes the model (.npz) and a JSON report of the held-out evaluation next to
it.  Needs the GPU.
"""
import argparse
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
    ap.add_argument("--fields", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--out", default=learned.MODEL_PATH)
    ap.add_argument("--eval-fields", type=int, default=4)
    ap.add_argument("--eval-shape", type=int, default=384)
    ap.add_argument("--seed", type=int, default=2026)
    a = ap.parse_args()

    def log(m):
        print(m, flush=True)

    e = faz_amd.Engine(0)
    t0 = time.time()
    X, T = learned.build_dataset(e, learned.training_fields(a.fields, a.seed), log=log)
    log(f"dataset: {X.shape[0]} tuples from {a.fields} fields in {time.time() - t0:.0f} s")
    t1 = time.time()
    net, mu, sigma, hist = learned.train(e, X, T, epochs=a.epochs, seed=a.seed, log=log)
    log(f"training: {time.time() - t1:.1f} s, final loss {hist[-1]:.5f}")
    learned.save_model(a.out, net, mu, sigma, dict(fields=a.fields, tuples=int(X.shape[0]), epochs=a.epochs,
                                                   seed=a.seed, final_loss=float(hist[-1])))
    report = dict(fields=a.fields, tuples=int(X.shape[0]), epochs=a.epochs, loss=hist, results=[])
    r = np.random.default_rng(a.seed + 1)
    shape = (a.eval_shape,) * 3
    for rel in (1e-3, 1e-4):
        for k in range(a.eval_fields):
            slope = float(r.uniform(2.0, 5.0))
            seed = int(r.integers(1 << 40)) | (1 << 41)  # disjoint from the training draws
            res = learned.evaluate(e, net, mu, sigma, shape, seed, rel, slope=slope, fronts=int(r.integers(0, 17)))
            ex, le, de = res["exhaustive"], res["learned"], res["default"]
            cr_e = 32.0 / ex["bitrate"]
            cr_l = learned.ratio_at_equal_psnr(le["bitrate"], le["psnr"], ex["psnr"])
            cr_d = learned.ratio_at_equal_psnr(de["bitrate"], de["psnr"], ex["psnr"])
            row = dict(rel=rel, slope=slope, seed=seed, exhaustive=ex, learned=le, default=de, ratio_exhaustive=cr_e,
                       ratio_learned_at_equal_psnr=cr_l, ratio_default_at_equal_psnr=cr_d,
                       speedup=ex["tune_ms"] / max(le["tune_ms"], 1e-9))
            report["results"].append(row)
            log(f"rel={rel} slope={slope:.2f}: exhaustive ({ex['alpha']},{ex['beta']}) {ex['tune_ms']:.1f} ms "
                f"CR {cr_e:.3f} PSNR {ex['psnr']:.2f} | learned ({le['alpha']},{le['beta']}) {le['tune_ms']:.1f} ms "
                f"CR@PSNR {cr_l:.3f} | default ({de['alpha']},{de['beta']}) CR@PSNR {cr_d:.3f}")
    with open(os.path.splitext(a.out)[0] + "_eval.json", "w") as f:
        json.dump(report, f, indent=1, default=float)
    log("done")


if __name__ == "__main__":
    main()
