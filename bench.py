#!/usr/bin/env python3
"""Throughput of the FAZ interpolation + quantization + Huffman path on one MI355X.

Workload (plain GPU drop-in, single field): a 512x512x512 float32 Gaussian
random field with power spectrum P(k) ~ k^-11/3, min-max normalised to [0,1]
plus N(0, (1e-4)^2) white noise, plus 16 randomly placed tanh fronts (width 4
cells, amplitude 0.5); REL error bound 1e-3, cubic interpolation, anchor
stride 32 (the reference's QoZ 3-D defaults).  Synthesised on the GPU from a
seed; the field is resident in HBM before the timed region starts.

A step is one device-resident compression of the field, what SZ_compress
does before zstd: value range, interpolation + quantization (one launch per
level; per pass on the coarse levels), histogram and outlier compaction,
host Huffman tree, GPU encode.  Like SZ_compress it returns the engine
buffer and no reconstruction, so the finest level writes no reconstructed
values.  `value` = 4 * n bytes of float32 input per step / step time (GB/s).
The ratio, PSNR and max error are measured once after the timed steps on a
compression that also returns the reconstruction (same engine buffer).

    python bench.py [--steps K] [--warmup W] [--rel 1e-3] [--no-cpu-baseline]
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time
import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "compress GB/s + ratio at REL 1e-3 (bit-exact); learned vs exhaustive tuner"
PEAK_HBM_GBS = 8000.0


def make_field(shape, seed=0, fronts=16, slope=11.0 / 3.0, noise=1e-4, width=4.0, amp=0.5):
    """k^-slope Gaussian random field + tanh fronts on the GPU (float32, contiguous, flat)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n0, n1, n2 = shape
    kx = torch.fft.fftfreq(n0, device="cuda").view(-1, 1, 1)
    ky = torch.fft.fftfreq(n1, device="cuda").view(1, -1, 1)
    kz = torch.fft.rfftfreq(n2, device="cuda").view(1, 1, -1)
    k = torch.sqrt(kx * kx + ky * ky + kz * kz)
    k[0, 0, 0] = 1.0
    ampl = k.pow(-slope / 2.0)
    ampl[0, 0, 0] = 0.0
    del k
    re = torch.randn(ampl.shape, generator=g, device="cuda")
    im = torch.randn(ampl.shape, generator=g, device="cuda")
    spec = torch.complex(re * ampl, im * ampl)
    del re, im, ampl
    f = torch.fft.irfftn(spec, s=shape)
    del spec
    f = (f - f.min()) / (f.max() - f.min())
    f += noise * torch.randn(shape, generator=g, device="cuda")
    if fronts:
        i = torch.arange(n0, device="cuda", dtype=torch.float32).view(-1, 1, 1)
        j = torch.arange(n1, device="cuda", dtype=torch.float32).view(1, -1, 1)
        l = torch.arange(n2, device="cuda", dtype=torch.float32).view(1, 1, -1)
        for _ in range(fronts):
            nv = torch.randn(3, generator=g, device="cuda")
            nv = nv / nv.norm()
            pnt = torch.rand(3, generator=g, device="cuda") * torch.tensor(shape, device="cuda",
                                                                           dtype=torch.float32)
            c = pnt.dot(nv)
            f += amp * torch.tanh((i * nv[0] + j * nv[1] + l * nv[2] - c) / width)
    out = f.to(torch.float32).contiguous().view(-1)
    torch.cuda.synchronize()
    return out


def measured_traffic(kernel, dims):
    """HBM bytes per launch of `kernel` on this workload, as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in profiles/ (None when
    no summary matches)."""
    import re
    path = os.path.join(ROOT, "profiles", "%s_%s_traffic.md" % (kernel, "x".join(map(str, dims))))
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", open(path).read())
    return int(m.group(1)) if m else None


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo, as lscpu prints it)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(x, dims, rel):
    """Reference CPU compressor (oracle/_ref, built from the reference sources)
    on this host, one thread; falls back to the oracle port when absent."""
    import oracle
    sample = "full %s field, REL %g, ALGO_INTERP cubic maxStep 32, SZ_compress incl. zstd, best of 2" % (
        "x".join(map(str, dims)), rel)
    host = x.cpu().numpy()
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "field.f32")
        host.tofile(path)
        n = host.size
        if os.path.exists(oracle.REF_HARNESS):
            out = subprocess.run([oracle.REF_HARNESS, "time", f"in={path}", "dims=" + ",".join(map(str, dims)),
                                  "mode=rel", f"eb={rel!r}", "algo=1", "maxstep=32", "reps=2"],
                                 capture_output=True, text=True, timeout=600)
            if out.returncode == 0:
                r = json.loads(out.stdout.strip().splitlines()[-1])
                return {"value": round(4.0 * n / r["seconds"] / 1e9, 4), "unit": "GB/s", "cores": 1,
                        "kind": "reference", "sample": sample, "cpu_model": cpu_model(),
                        "ratio": round(4.0 * n / r["bytes"], 3)}
    # oracle port: time one interpolation pass over a 256^3 crop (single thread)
    crop = host.reshape(dims)[:256, :256, :256].copy()
    mn, mx = oracle.data_range(crop)
    p = oracle.make_params(crop.shape, rel * float((mx - mn)), algo=1, max_step=32)
    t0 = time.perf_counter()
    oracle.compress_stream(crop, p, rel, rel=True)
    dt = time.perf_counter() - t0
    return {"value": round(4.0 * crop.size / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": "256^3 crop of the field, REL %g, oracle port incl. zstd" % rel, "cpu_model": cpu_model()}


def tuner_report(eng, x, dims, rel, mn, mx):
    """The exhaustive QoZ tuner (faz -q) and the learned tuner on the bench
    field: tuning times (best of 3, device + host), chosen alpha/beta, the
    engine buffer size and PSNR of each choice, the ratio of the learned
    choice moved to the exhaustive choice's PSNR (6.02 dB per bit), and the
    oracle's CPU restatement of the exhaustive tuner on one thread."""
    import torch
    import faz_amd
    from faz_amd import learned
    net, mu, sigma = learned.load_model(eng)
    gates = learned.load_gates()
    rng = float(np.float32(mx - mn))
    eb = rel * rng
    n = x.numel()
    res = {}
    for name in ("exhaustive", "learned"):
        best = None
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            # faz -q: the QoZ preset clears testLorenzo (SZI:1616) for both
            r = eng.tune_qoz(x, dims, eb, rel, rng) if name == "exhaustive" else \
                learned.tune_learned(eng, x, dims, eb, rel, rng, net, mu, sigma, gates, test_lorenzo=0)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            best = ms if best is None else min(best, ms)
        p = faz_amd.make_params(dims, eb, algo=r["interp_algo"], direction=r["direction"], alpha=r["alpha"],
                                beta=r["beta"], max_step=r["max_step"], ops=tuple(r["ops"]), dirs=tuple(r["dirs"]))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        size = eng.compress_device(x, p, False, eb)
        torch.cuda.synchronize()
        cms = (time.perf_counter() - t0) * 1e3
        nbytes, mse = learned.final_stats(eng, x, dims, eb, p)
        res[name] = {"tune_ms": round(best, 3), "compress_ms": round(cms, 3), "alpha": r["alpha"], "beta": r["beta"],
                     "ops": r["ops"], "dirs": r["dirs"], "compress_tests": len(r["tests"]), "raw_bytes": size,
                     "learned_path": r["learned_path"], "n_verified": r["n_verified"],
                     "bitrate": round(8.0 * nbytes / n, 5), "psnr_db": round(20 * math.log10(rng) -
                                                                             10 * math.log10(mse), 4)}
    ex, le = res["exhaustive"], res["learned"]
    out = dict(res)
    out["tuning_speedup"] = round(ex["tune_ms"] / le["tune_ms"], 2)
    out["end_to_end_speedup"] = round((ex["tune_ms"] + ex["compress_ms"]) / (le["tune_ms"] + le["compress_ms"]), 2)
    out["ratio_exhaustive"] = round(32.0 / ex["bitrate"], 4)
    out["ratio_learned_at_equal_psnr"] = round(learned.ratio_at_equal_psnr(le["bitrate"], le["psnr_db"],
                                                                           ex["psnr_db"]), 4)
    net.close()
    try:
        import oracle
        host = x.cpu().numpy().reshape(dims)
        t0 = time.perf_counter()
        o = oracle.qoz_tune(host, eb, rel, rng)
        out["cpu_oracle_exhaustive_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        out["cpu_oracle_same_choice"] = (o["alpha"], o["beta"], o["ops"], o["dirs"]) == (ex["alpha"], ex["beta"],
                                                                                         ex["ops"], ex["dirs"])
    except Exception as e:
        out["cpu_oracle_exhaustive_ms"] = None
        out["cpu_oracle_error"] = str(e)[:200]
    return out


def rel_report(eng, x, dims, rel, steps=5):
    """The bench step at another error bound on the same field: time per
    step, GB/s, engine buffer ratio, PSNR and the bound checked in FP64."""
    import torch
    import faz_amd
    p = faz_amd.make_params(dims, 0.0, algo=1, max_step=32)
    eng.compress_device(x, p, True, rel)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        size = eng.compress_device(x, p, True, rel)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    recon = torch.empty_like(x)
    eng.compress_device(x, p, True, rel, recon=recon)
    mn, mx = eng.range(x)
    rng = float(np.float32(mx - mn))
    err = recon.double() - x.double()
    mse = (err * err).mean().item()
    out = {"rel_eb": rel, "ms_per_step": round(dt * 1e3, 4), "GB/s": round(4.0 * x.numel() / dt / 1e9, 2),
           "raw_ratio": round(4.0 * x.numel() / size, 3), "psnr_db": round(20 * math.log10(rng) - 10 * math.log10(mse), 3),
           "max_abs_err": err.abs().max().item(), "eb": rel * rng}
    out["bound_holds"] = out["max_abs_err"] <= out["eb"]
    del recon, err
    return out


def c5_report(eng, fields, rels=(1e-2, 1e-3, 1e-4, 1e-6), shape=(1024, 1024, 512), seed=500):
    """The combined path at scale: fields of 1024x1024x512 f32 (2 GiB) with
    spectral slopes k^-2..k^-5 plus tanh fronts, generated one at a time; per
    field and REL bound the learned tuner (Lorenzo test on, gates on:
    out-of-range inputs fall back to the exhaustive search) picks the
    interpolation configuration or the Lorenzo engine, and the choice is
    compressed on the device.  Reported per bound: mean engine-buffer ratio
    (pre-zstd) and PSNR, the compression GB/s, the learned tuning time, the
    predictor and gate outcomes, and the same field through the plain GPU
    path (cubic, untuned alpha/beta) for comparison; the first field also
    runs the exhaustive tuner."""
    import torch
    import faz_amd
    from faz_amd import learned
    net, mu, sigma = learned.load_model(eng)
    gates = learned.load_gates()
    n = int(np.prod(shape))
    per = {rel: {"ratio": [], "psnr": [], "gbs": [], "tune_ms": [], "plain_ratio": [], "plain_psnr": [],
                 "lorenzo": 0, "path": [0, 0, 0], "bound_ok": True} for rel in rels}
    exhaustive = {}
    t_start = time.perf_counter()
    for k in range(fields):
        slope = 2.0 + 3.0 * k / max(fields - 1, 1)
        x = make_field(shape, seed=seed + k, fronts=8, slope=slope)
        mn, mx = eng.range(x)
        rng = float(np.float32(mx - mn))
        for rel in rels:
            eb = rel * rng
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            t = learned.tune_learned(eng, x, shape, eb, rel, rng, net, mu, sigma, gates, test_lorenzo=1)
            torch.cuda.synchronize()
            r = per[rel]
            r["tune_ms"].append((time.perf_counter() - t0) * 1e3)
            r["path"][t["learned_path"]] += 1
            recon = torch.empty_like(x)
            if t["use_interp"]:
                p = learned.choice_params(shape, eb, t)
                eng.compress_device(x, p, False, eb)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                size = eng.compress_device(x, p, False, eb)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                eng.compress_device(x, p, False, eb, recon=recon)
            else:
                r["lorenzo"] += 1
                lp = faz_amd.make_lorenzo_params(shape, eb, block_size=5, pred_dim=3, lorenzo=True, lorenzo2=True,
                                                 regression=False)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                size = len(eng.lorenzo_raw(x, lp, recon=recon))
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
            err = recon.double() - x.double()
            r["bound_ok"] &= err.abs().max().item() <= eb
            mse = (err * err).mean().item()
            r["ratio"].append(4.0 * n / size)
            r["psnr"].append(20 * math.log10(rng) - 10 * math.log10(mse))
            r["gbs"].append(4.0 * n / dt / 1e9)
            del err
            plain = faz_amd.make_params(shape, eb, algo=1, max_step=32)
            psize = eng.compress_device(x, plain, False, eb, recon=recon)
            err = recon.double() - x.double()
            r["plain_ratio"].append(4.0 * n / psize)
            r["plain_psnr"].append(20 * math.log10(rng) - 10 * math.log10((err * err).mean().item()))
            del err, recon
            if k == 0:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                ex = eng.tune_qoz(x, shape, eb, rel, rng, test_lorenzo=1)
                torch.cuda.synchronize()
                exhaustive[str(rel)] = {"tune_ms": round((time.perf_counter() - t0) * 1e3, 2),
                                        "same_choice": (ex["use_interp"], ex["alpha"], ex["beta"]) ==
                                                       (t["use_interp"], t["alpha"], t["beta"]),
                                        "learned_tune_ms": round(r["tune_ms"][-1], 2)}
        del x
        torch.cuda.empty_cache()
    net.close()
    out = {"fields": fields, "shape": list(shape), "slopes": "k^-2..k^-5", "ratio": "engine buffer (pre-zstd)",
           "seconds": round(time.perf_counter() - t_start, 1), "exhaustive_first_field": exhaustive, "by_rel": {}}
    for rel, r in per.items():
        out["by_rel"][str(rel)] = {
            "ratio_mean": round(float(np.mean(r["ratio"])), 3), "psnr_mean": round(float(np.mean(r["psnr"])), 3),
            "GB/s_mean": round(float(np.mean(r["gbs"])), 1), "learned_tune_ms_mean": round(float(np.mean(r["tune_ms"])), 2),
            "plain_ratio_mean": round(float(np.mean(r["plain_ratio"])), 3),
            "plain_psnr_mean": round(float(np.mean(r["plain_psnr"])), 3),
            "lorenzo_chosen": r["lorenzo"], "learned_paths_direct_verified_fallback": r["path"],
            "bound_holds": bool(r["bound_ok"])}
    return out


def batch_fields(count, seed):
    """The batched workload: `count` float32 fields drawn from a seed among
    (i) 2-D 1800x3600 lat-lon fields: zonal cos(lat) bands + k^-3 GRF + 1e-3
    noise, (ii) 3-D 100x500x500 k^-11/3 GRFs, (iii) 3-D 256x384x384 GRFs with
    5% of the points in exact-zero masked boxes."""
    import torch
    from faz_amd.learned import grf_field
    r = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = int(r.integers(0, 3))
        s = int(r.integers(1 << 40))
        if kind == 0:
            shape = (1800, 3600)
            f = grf_field(shape, s, slope=3.0, noise=1e-3).view(shape)
            lat = torch.linspace(-math.pi / 2, math.pi / 2, shape[0], device="cuda").view(-1, 1)
            f = 0.5 * f + 0.5 * (torch.cos(lat) * (0.6 + 0.4 * torch.cos(6 * lat))).expand(shape)
        elif kind == 1:
            shape = (100, 500, 500)
            f = grf_field(shape, s, slope=11.0 / 3.0).view(shape)
        else:
            shape = (256, 384, 384)
            f = grf_field(shape, s, slope=11.0 / 3.0).view(shape)
            g = torch.Generator(device="cuda")
            g.manual_seed(s)
            masked = 0
            while masked < 0.05 * f.numel():
                c = [int(torch.randint(0, n, (1,), generator=g, device="cuda").item()) for n in shape]
                h = [int(torch.randint(8, 48, (1,), generator=g, device="cuda").item()) for _ in shape]
                sl = tuple(slice(ci, min(n, ci + hi)) for ci, hi, n in zip(c, h, shape))
                f[sl] = 0.0
                masked = int((f == 0).sum().item())
        out.append((shape, f.contiguous().view(-1)))
    torch.cuda.synchronize()
    return out


def batch_report(eng, count, seed, rel, steps=3):
    """Batched uneven fields: one hipGraph of per-field branches, autotuned
    launch schedules, against the same fields compressed one call at a time."""
    import torch
    import faz_amd
    fields = batch_fields(count, seed)
    ps = [faz_amd.make_params(s, 0.0, algo=1, max_step=64 if len(s) == 2 else 32) for s, _ in fields]
    t0 = time.perf_counter()
    b = faz_amd.Batch(eng, ps, [f for _, f in fields], autotune=True)
    create_s = time.perf_counter() - t0
    try:
        b.compress(True, rel)  # captures the prediction graph
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sizes = b.compress(True, rel)
        torch.cuda.synchronize()
        t_batch = (time.perf_counter() - t0) / steps
        st = b.stats()
        # the same fields one SZ_compress-equivalent call at a time
        for (s, f), p in zip(fields[:2], ps[:2]):
            eng.compress_device(f, p, True, rel)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        single_sizes = []
        for _ in range(steps):
            single_sizes = [eng.compress_device(f, p, True, rel) for (s, f), p in zip(fields, ps)]
        torch.cuda.synchronize()
        t_single = (time.perf_counter() - t0) / steps
        same = all(int(a) == int(c) for a, c in zip(sizes, single_sizes))
        # byte check on the last field compressed singly and on the first one
        last = eng.fetch_raw(single_sizes[-1])
        same = same and last == b.fetch(len(fields) - 1, sizes[-1])
        nbytes = sum(4 * f.numel() for _, f in fields)
        kinds = {}
        for s, _ in fields:
            kinds["x".join(map(str, s))] = kinds.get("x".join(map(str, s)), 0) + 1
        return {"fields": len(fields), "shapes": kinds, "input_gb": round(nbytes / 1e9, 3), "rel_eb": rel,
                "batch_ms": round(t_batch * 1e3, 3), "batch_gbs": round(nbytes / t_batch / 1e9, 2),
                "single_calls_ms": round(t_single * 1e3, 3), "single_calls_gbs": round(nbytes / t_single / 1e9, 2),
                "speedup": round(t_single / t_batch, 3), "ratio": round(nbytes / float(sum(sizes)), 3),
                "engine_buffers_identical": bool(same), "create_s": round(create_s, 3),
                "graph": {k: st[k] for k in ("captures", "updates", "instantiations", "replays")},
                "phases_ms": {k: round(st[k], 3) for k in ("range_ms", "pred_ms", "encode_ms", "total_ms")},
                "schedules": {str(k): st["schedules"].count(k) for k in sorted(set(st["schedules"]))}}
    finally:
        b.close()
        del fields
        torch.cuda.empty_cache()


def decompress_report(eng, raw, recon, steps=5):
    """SZInterpolationCompressor::decompress of the step's engine buffer on
    the GPU (Huffman decode, outlier scatter, level recovery), the output
    checked equal to the compression's reconstruction bit for bit: from
    device memory (the buffer resident in HBM, the headline figure) and from
    host memory as SZ_decompress hands it over (PCIe upload included)."""
    import torch
    out = torch.empty_like(recon)
    n = recon.numel()
    d_raw = torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda()

    def leg(call, label):
        out.fill_(float("nan"))
        call()
        torch.cuda.synchronize()
        if not torch.equal(out.view(torch.int32), recon.view(torch.int32)):
            raise RuntimeError(f"GPU decompression ({label}) differs from the compression's reconstruction")
        decode = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            call()
            decode += eng.timings()[6]
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        return wall, decode / steps

    wd, dd = leg(lambda: eng.decompress_dev(d_raw, 3, out), "device")
    wh, dh = leg(lambda: eng.decompress_raw(raw, 3, out), "host")
    return {"ms": round(wd * 1e3, 4), "GB/s": round(4.0 * n / wd / 1e9, 2),
            "huffman_host_ms": round(dd, 4), "raw_bytes": len(raw), "equal_to_recon": True,
            "input": "pre-zstd engine buffer resident in device memory",
            "host_input": {"ms": round(wh * 1e3, 4), "GB/s": round(4.0 * n / wh / 1e9, 2),
                           "huffman_host_ms": round(dh, 4),
                           "input": "pre-zstd engine buffer in pageable host memory (PCIe upload included)"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rel", type=float, default=1e-3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tuner", action="store_true", help="skip the exhaustive / learned tuner comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven-fields report")
    ap.add_argument("--batch-fields", type=int, default=48)
    ap.add_argument("--no-rel-sweep", action="store_true", help="skip the REL 1e-5 step on the same field")
    ap.add_argument("--no-decompress", action="store_true", help="skip the decompression report")
    ap.add_argument("--c5-fields", type=int, default=4,
                    help="1024x1024x512 fields of the combined-path leg (0 skips it; the full workload is 24)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="one device copy of the field (4n bytes read, 4n written) before the steps, the known-size "
                         "kernel that calibrates FETCH_SIZE / WRITE_SIZE in rocprofv3 --pmc runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the compression of one field on one MI355X; "
              "multi-GPU sharding is design-only (see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import torch
    import faz_amd

    dims = (args.n, args.n, args.n)
    n = args.n ** 3
    x = make_field(dims, seed=args.seed)
    if args.copy_probe:
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del y
    eng = faz_amd.Engine(0)
    p = faz_amd.make_params(dims, 0.0, algo=1, max_step=32)

    for _ in range(args.warmup):
        eng.compress_device(x, p, True, args.rel)
    torch.cuda.synchronize()
    eng.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        raw_size = eng.compress_device(x, p, True, args.rel)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof_ms, launches, points, kname, prof_bytes = eng.profile()
    raw = eng.fetch_raw(raw_size)
    # phase timings from one more compression outside the timed region (their
    # event records idle the stream between phases)
    eng.set_profiling(False, phases=True)
    eng.compress_device(x, p, True, args.rel)
    timings = eng.timings()
    eng.set_profiling(False)

    dt = (t1 - t0) / args.steps
    value = 4.0 * n / dt / 1e9
    avg_launch_ms = prof_ms / max(launches, 1)
    achieved = prof_bytes / (prof_ms / 1e3) / 1e9 if prof_ms > 0 else 0.0

    # quality of the stream (outside the timed region), from a compression that
    # also returns the reconstruction; its engine buffer must be the same
    recon = torch.empty_like(x)
    raw_size2 = eng.compress_device(x, p, True, args.rel, recon=recon)
    if eng.fetch_raw(raw_size2) != raw:
        raise SystemExit("engine buffer differs between the runs with and without the reconstruction")
    mn, mx = eng.range(x)
    eb = args.rel * float(torch.tensor(mx - mn, dtype=torch.float32).item())
    stream = eng.sz_compress(x, p, args.rel, rel=True)
    err = (recon.double() - x.double())
    max_err = err.abs().max().item()
    mse = (err * err).mean().item()
    psnr = 20 * math.log10(float(mx - mn)) - 10 * math.log10(mse) if mse > 0 else float("inf")

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded k^-11/3 Gaussian random field + 16 tanh fronts, generated on the GPU",
        "config": {"workload": "single field 512x512x512 f32, REL %g, cubic interpolation, maxStep 32, "
                               "interp+quantize+Huffman on GPU (ALGO_INTERP drop-in)" % args.rel,
                   "dims": list(dims), "rel_eb": args.rel, "interp": "cubic", "max_step": 32},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": measured_traffic(kname, dims),
                     "kernel": kname, "launches_per_step": launches // max(args.steps, 1),
                     "avg_launch_ms": round(avg_launch_ms, 5),
                     "bytes_per_launch": round(prof_bytes / max(launches, 1)),
                     "bytes_per_point": round(prof_bytes / max(points, 1), 4)},
        "quality": {"ratio": round(4.0 * n / len(stream), 3), "stream_bytes": len(stream), "raw_bytes": raw_size,
                    "max_abs_err": max_err, "eb": eb, "psnr_db": round(psnr, 3)},
        "phases_ms": {"range": round(timings[0], 4), "interp": round(timings[1], 4),
                      "hist_outliers": round(timings[2], 4), "tree": round(timings[3], 4),
                      "encode": round(timings[4], 4), "device_total": round(timings[5], 4),
                      "host_tree": round(timings[8], 4), "host_encode_enqueue": round(timings[9], 4)},
    }
    if not args.no_decompress:
        try:
            line["decompress"] = decompress_report(eng, raw, recon)
        except Exception as e:  # reported beside the measurement
            line["decompress"] = {"error": str(e)[:200]}
    if not args.no_rel_sweep:
        try:
            line["rel_1e-5"] = rel_report(eng, x, dims, 1e-5)
        except Exception as e:  # reported beside the measurement
            line["rel_1e-5"] = {"error": str(e)[:200]}
    if not args.no_tuner:
        try:
            line["tuner"] = tuner_report(eng, x, dims, args.rel, mn, mx)
        except Exception as e:  # reported beside the measurement
            line["tuner"] = {"error": str(e)[:200]}
    if not args.no_batch:
        try:
            line["batch"] = batch_report(eng, args.batch_fields, args.seed + 17, args.rel)
        except Exception as e:  # reported beside the measurement
            line["batch"] = {"error": str(e)[:200]}
    if args.c5_fields > 0:
        try:
            line["c5"] = c5_report(eng, args.c5_fields)
        except Exception as e:  # reported beside the measurement
            line["c5"] = {"error": str(e)[:200]}
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(x, dims, args.rel)
        except Exception as e:  # the baseline is a report, not the measurement
            line["cpu_baseline"] = {"value": None, "error": str(e)[:200]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
