"""Writes the golden fixtures under tests/golden/ from the reference engine.

Runs only where the reference tree is mounted and oracle/_ref/ref_harness has
been built from it (``make -C oracle ref``).  Each fixture stores the input,
the parameters and what the reference engine produced for it: quantization
codes, unpredictable list, reconstruction, the pre-zstd engine buffer and the
full stream (zstd frame from the reference's vendored zstd 1.4.5 + Config
trailer).  Huffman fixtures store a symbol sequence and the reference
HuffmanEncoder's save+encode bytes; tuning fixtures the codes, loss and level
counts of SZInterpolationCompressor::compress in tuning modes 1 and 2;
Lorenzo/regression fixtures (``l*``) the codes, reconstruction, pre-zstd
buffer and stream of SZ_compress with cmprAlgo = ALGO_LORENZO_REG on 3-D
fields (SZFastFrontend).

    python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def smooth(rng, shape, k=3):
    x = rng.standard_normal(shape)
    for ax in range(len(shape)):
        for _ in range(k):
            x = np.cumsum(x, axis=ax)
    x = (x - x.min()) / (x.max() - x.min() + 1e-30)
    return x.astype(np.float32)


def cases():
    r = np.random.default_rng(20261203)
    c = []
    add = lambda name, x, **kw: c.append((name, x, kw))
    add("c3d_17_cubic_rel1e-3", smooth(r, (17, 17, 17)), mode="rel", eb=1e-3)
    add("c3d_17_anchor32", smooth(r, (17, 17, 17)), mode="rel", eb=1e-3, maxstep=32)
    add("c3d_33x20x9_levelwise", smooth(r, (33, 20, 9), 2), mode="rel", eb=1e-2, maxstep=8, alpha=1.5,
        beta=3.0, ops=[1, 0, 1], dirs=[0, 5, 2], dir=2)
    add("c3d_65x33x17_alpha", smooth(r, (65, 33, 17)), mode="rel", eb=1e-4, maxstep=32, alpha=1.25, beta=2.0)
    add("c3d_33cube_degenerate", smooth(r, (33, 33, 33), 2), mode="rel", eb=1e-3, dir=4)
    add("c3d_24x31x20_bs16", smooth(r, (24, 31, 20)), mode="rel", eb=5e-4, bs=16, dir=1, maxstep=16)
    add("c3d_20_qbins1024", smooth(r, (20, 20, 20), 1), mode="rel", eb=1e-5, qbins=1024)
    add("c3d_13_outliers", r.random((13, 13, 13)).astype(np.float32), mode="abs", eb=1e-9)
    spikes = smooth(r, (11, 12, 13))
    spikes.flat[r.choice(spikes.size, 20, replace=False)] = np.float32(3e30)
    add("c3d_spikes_overflow", spikes, mode="abs", eb=1e-3)
    add("c3d_const_9", np.full((9, 9, 9), 3.25, np.float32), mode="abs", eb=1e-2)
    add("c3d_linear_dir3", smooth(r, (19, 23, 29), 2), mode="rel", eb=1e-3, algo=0, dir=3)
    add("c3d_alpha_lt1", smooth(r, (21, 18, 27)), mode="rel", eb=1e-3, alpha=0.5, beta=0.2, maxstep=16)
    add("c2d_45x37_anchor", smooth(r, (45, 37)), mode="rel", eb=1e-2, maxstep=32)
    add("c2d_100x37_linear", smooth(r, (100, 37), 2), mode="rel", eb=1e-4, algo=0, dir=1, alpha=0.5, beta=0.2)
    add("c2d_90x130_cubic", smooth(r, (90, 130), 2), mode="rel", eb=1e-3, maxstep=64, ops=[1] * 6,
        dirs=[0, 1, 0, 1, 0, 1])
    add("c1d_1000", smooth(r, (1000,), 2), mode="rel", eb=1e-3)
    add("c1d_4097_linear", smooth(r, (4097,), 1), mode="rel", eb=1e-4, algo=0)
    add("c4d_9x10x11x12", smooth(r, (9, 10, 11, 12), 1), mode="rel", eb=1e-2, dir=17)
    add("c3d_1x40x40_flat_axis", smooth(r, (1, 40, 40), 2), mode="rel", eb=1e-3, maxstep=32)
    add("c3d_negative_values", (smooth(r, (16, 17, 18)) - 0.7) * 1e3, mode="rel", eb=1e-3, maxstep=32)
    return c


def huffman_cases():
    r = np.random.default_rng(7)
    c = []
    c.append(("h_geometric", (32768 + np.round(r.laplace(0, 6, 20000))).astype(np.int32)))
    c.append(("h_600_symbols", (30000 + r.integers(0, 600, 30000)).astype(np.int32)))
    z = r.integers(0, 40000, 200000).astype(np.int32)
    c.append(("h_40000_symbols_u32_tree", z))
    c.append(("h_single", np.full(1000, 32768, np.int32)))
    c.append(("h_two", np.array([5, 9, 9, 5, 9] * 40, np.int32)))
    c.append(("h_with_zero", np.where(r.random(5000) < 0.1, 0, 32768 + r.integers(-3, 4, 5000)).astype(np.int32)))
    c.append(("h_negative", r.integers(-50, 50, 3000).astype(np.int32)))
    fib = []
    a, b = 1, 1
    for s in range(30):
        fib += [s] * a
        a, b = b, a + b
    c.append(("h_fibonacci_deep", np.array(fib, np.int32)[r.permutation(len(fib))][:200000]))
    return c


def tuning_cases():
    r = np.random.default_rng(11)
    c = []
    x = smooth(r, (33, 33, 33), 2)
    c.append(("t3d_33_mode1", x, dict(mode="abs", eb=1e-3, maxstep=32, tuning=1)))
    c.append(("t3d_33_mode2_L3", x, dict(mode="abs", eb=1e-3, maxstep=32, tuning=2, start=3, end=2, algo=0)))
    c.append(("t3d_33_mode2_L1", x, dict(mode="abs", eb=1e-3, maxstep=32, tuning=2, start=1, end=0, dir=5)))
    c.append(("t3d_17_mode1_lw", smooth(r, (17, 17, 17)), dict(mode="abs", eb=2e-3, maxstep=16, tuning=1,
                                                               alpha=1.5, beta=2, ops=[1, 1, 0, 1],
                                                               dirs=[0, 5, 0, 5])))
    c.append(("t2d_65_mode1", smooth(r, (65, 65)), dict(mode="abs", eb=1e-3, maxstep=64, tuning=1)))
    return c


def lorenzo_cases():
    """3-D Lorenzo/regression fixtures (SZFastFrontend through SZ_compress)."""
    r = np.random.default_rng(424242)
    c = []
    add = lambda name, x, **kw: c.append((name, x, kw))
    add("l3d_20x17x23_default", smooth(r, (20, 17, 23)), mode="rel", eb=1e-3)
    add("l3d_18_lo2_reg", smooth(r, (18, 18, 18), 2), mode="rel", eb=1e-3, lo2=1)
    add("l3d_19x14x22_pd2_lo2", smooth(r, (19, 14, 22), 2), mode="rel", eb=1e-4, pd=2, lo2=1)
    add("l3d_16x21x13_pd1", smooth(r, (16, 21, 13), 2), mode="rel", eb=1e-3, pd=1, lo2=1)
    add("l3d_26x11x31_bs5", smooth(r, (26, 11, 31)), mode="rel", eb=1e-3, block=5)
    add("l3d_21x24x20_bs8_lo2only", smooth(r, (21, 24, 20), 1), mode="rel", eb=1e-3, block=8, lo=0, lo2=1, reg=0)
    add("l3d_17x19x15_regonly", smooth(r, (17, 19, 15)), mode="rel", eb=1e-2, lo=0, reg=1)
    add("l3d_15_noise_tiny_eb", r.random((15, 15, 15)).astype(np.float32), mode="abs", eb=1e-9, lo2=1)
    add("l3d_20_qbins1024", smooth(r, (20, 20, 20), 1), mode="rel", eb=1e-6, qbins=1024, lo2=1)
    spikes = smooth(r, (14, 13, 12))
    spikes.flat[r.choice(spikes.size, 25, replace=False)] = np.float32(3e30)
    add("l3d_spikes_overflow", spikes, mode="abs", eb=1e-3, lo2=1)
    add("l3d_const_9", np.full((9, 9, 9), -2.5, np.float32), mode="abs", eb=1e-2, lo2=1)
    # piecewise field: linear ramps (regression wins), smooth bumps (Lorenzo 2-layer) and noise
    g = np.stack(np.meshgrid(*[np.arange(n, dtype=np.float64) for n in (24, 30, 27)], indexing="ij"))
    mixed = 0.01 * g[0] - 0.02 * g[1] + 0.005 * g[2]
    mixed[:12] = np.sin(g[0][:12] / 3.0) * np.cos(g[2][:12] / 4.0)
    mixed[:, :, 18:] += 0.3 * r.standard_normal(mixed[:, :, 18:].shape)
    add("l3d_24x30x27_mixed", mixed.astype(np.float32), mode="rel", eb=1e-3, lo2=1)
    add("l3d_negative_abs", ((smooth(r, (13, 25, 14)) - 0.6) * 1e4).astype(np.float32), mode="abs", eb=0.5)
    add("l3d_1x40x40_thin", smooth(r, (1, 40, 40), 2), mode="rel", eb=1e-3, lo2=1)
    add("l3d_40x2x33_thin", smooth(r, (40, 2, 33), 2), mode="rel", eb=1e-3)
    add("l3d_37x3x7_bs3", smooth(r, (37, 3, 7), 2), mode="rel", eb=1e-3, block=3, lo2=1)
    add("l3d_12_bs12_single_block", smooth(r, (12, 12, 12)), mode="rel", eb=1e-3, block=12, lo2=1)
    return c


def ctest_cases():
    """CompressTest fixtures: a field, sample-block starts, the block edge and
    the engine parameters of one tuner candidate."""
    r = np.random.default_rng(99)
    c = []
    f3 = smooth(r, (34, 34, 51), 2)
    c.append(("q3d_17_lw", f3, [(0, 0, 0), (16, 16, 16), (0, 16, 32), (16, 0, 34)], 16,
              dict(eb=1e-3, alpha=1.25, beta=2.0, maxstep=16, ops=[1, 0, 1, 1], dirs=[0, 5, 5, 0])))
    c.append(("q3d_17_noanchor", f3, [(0, 0, 0), (17, 17, 17)], 16,
              dict(eb=2e-4, alpha=1.5, beta=3.0, maxstep=0, algo=1, dir=3)))
    c.append(("q3d_17_small_eb", f3, [(5, 6, 7), (16, 16, 30)], 16,
              dict(eb=3e-6, alpha=1.0, beta=1.5, maxstep=16, ops=[1, 1, 1, 1], dirs=[0, 0, 0, 0])))
    f33 = smooth(r, (66, 34, 34), 2)
    c.append(("q3d_33_qoz", f33, [(0, 0, 0), (32, 0, 0), (33, 1, 1)], 32,
              dict(eb=5e-4, alpha=1.75, beta=4.0, maxstep=32, ops=[0, 1, 1, 0], dirs=[5, 0, 5, 0])))
    f2 = smooth(r, (130, 131), 2)
    c.append(("q2d_65_qoz", f2, [(0, 0), (64, 64), (65, 0)], 64,
              dict(eb=1e-3, alpha=1.25, beta=1.5, maxstep=64, ops=[1, 0, 1, 1, 0, 1], dirs=[0, 1, 1, 0, 0, 1])))
    return c


def general_cases():
    """General-frontend fixtures (SZGeneralFrontend with the Lorenzo
    predictors, N = 1..4) through SZ_compress."""
    r = np.random.default_rng(515151)
    c = []
    add = lambda name, x, **kw: c.append((name, x, kw))
    add("g1d_1000_default", smooth(r, (1000,), 2), mode="rel", eb=1e-3)
    add("g1d_777_bs5_lo2only", smooth(r, (777,), 2), mode="rel", eb=1e-4, block=5, lo=0, lo2=1)
    add("g1d_500_bs7_loonly", smooth(r, (500,), 1), mode="rel", eb=1e-3, block=7, lo=1, lo2=0)
    add("g1d_301_bs2", smooth(r, (301,), 2), mode="rel", eb=1e-3, block=2)
    add("g2d_100x130_default", smooth(r, (100, 130), 2), mode="rel", eb=1e-3)
    add("g2d_61x47_bs5", smooth(r, (61, 47), 2), mode="rel", eb=1e-4, block=5)
    add("g2d_33x90_bs6_lo2only", smooth(r, (33, 90), 2), mode="rel", eb=1e-3, block=6, lo=0, lo2=1)
    add("g2d_40x41_loonly_abs", ((smooth(r, (40, 41), 1) - 0.5) * 100).astype(np.float32), mode="abs", eb=0.05,
        lo=1, lo2=0)
    add("g2d_30x30_bs1_loonly", smooth(r, (30, 30), 2), mode="rel", eb=1e-3, block=1, lo=1, lo2=0)
    add("g2d_35x29_bs2", smooth(r, (35, 29), 2), mode="rel", eb=1e-3, block=2)
    add("g2d_25_noise_tiny_eb", r.random((25, 25)).astype(np.float32), mode="abs", eb=1e-9)
    add("g2d_50x60_qbins1024", smooth(r, (50, 60), 1), mode="rel", eb=1e-6, qbins=1024)
    spikes = smooth(r, (40, 37))
    spikes.flat[r.choice(spikes.size, 20, replace=False)] = np.float32(3e30)
    add("g2d_spikes_overflow", spikes, mode="abs", eb=1e-3)
    add("g2d_const_20", np.full((20, 20), 1.75, np.float32), mode="abs", eb=1e-2)
    add("g2d_1x200_thin", smooth(r, (1, 200), 2), mode="rel", eb=1e-3, block=5)
    add("g2d_200x1_thin", smooth(r, (200, 1), 2), mode="rel", eb=1e-3, block=5)
    add("g3d_20x23x17_bs5", smooth(r, (20, 23, 17), 2), mode="rel", eb=1e-3, block=5)
    add("g3d_13x9x11_bs4_lo2only", smooth(r, (13, 9, 11), 2), mode="rel", eb=1e-3, block=4, lo=0, lo2=1)
    add("g4d_9x10x11x12_default", smooth(r, (9, 10, 11, 12), 2), mode="rel", eb=1e-3)
    add("g4d_8x7x9x10_bs3_lo2only", smooth(r, (8, 7, 9, 10), 2), mode="rel", eb=1e-3, block=3, lo=0, lo2=1)
    add("g4d_6x7x5x8_bs2", smooth(r, (6, 7, 5, 8), 2), mode="rel", eb=1e-4, block=2)
    # RegressionPredictor inside the ComposedPredictor: the Config default
    # (lorenzo, regression; CFG:410-413) and the other compositions, blocks
    # where regression wins (planes plus noise) and where Lorenzo does
    plane = lambda shape: (sum(np.indices(shape, dtype=np.float64)[a] * (0.013 * (a + 1))
                               for a in range(len(shape))) + 1e-4 * r.standard_normal(shape)).astype(np.float32)
    mixed = lambda shape: (np.where(np.indices(shape)[0] < shape[0] // 2, plane(shape), smooth(r, shape, 2))
                           ).astype(np.float32)
    add("r1d_1000_default", smooth(r, (1000,), 2), mode="rel", eb=1e-3, lo=1, lo2=0, reg=1)
    add("r1d_900_planes", plane((900,)), mode="rel", eb=1e-4, block=30, lo=1, lo2=0, reg=1)
    add("r2d_100x130_default", smooth(r, (100, 130), 2), mode="rel", eb=1e-3, lo=1, lo2=0, reg=1)
    add("r2d_96x112_mixed", mixed((96, 112)), mode="rel", eb=1e-4, lo=1, lo2=0, reg=1)
    add("r2d_61x47_bs5_all3", mixed((61, 47)), mode="rel", eb=1e-3, block=5, lo=1, lo2=1, reg=1)
    add("r2d_64x80_regonly", plane((64, 80)), mode="rel", eb=1e-3, lo=0, lo2=0, reg=1)
    add("r2d_40x56_lo2reg", mixed((40, 56)), mode="rel", eb=1e-3, block=8, lo=0, lo2=1, reg=1)
    add("r2d_50x60_qbins1024", mixed((50, 60)), mode="rel", eb=1e-6, qbins=1024, lo=1, lo2=0, reg=1)
    spikes = plane((48, 40))
    spikes.flat[r.choice(spikes.size, 12, replace=False)] = np.float32(3e30)
    add("r2d_spikes_overflow", spikes, mode="abs", eb=1e-3, block=8, lo=1, lo2=0, reg=1)
    add("r3d_20x24x18_bs6", mixed((20, 24, 18)), mode="rel", eb=1e-3, block=6, lo=1, lo2=0, reg=1)
    add("r4d_9x10x11x12_default", smooth(r, (9, 10, 11, 12), 2), mode="rel", eb=1e-3, lo=1, lo2=0, reg=1)
    add("r4d_8x9x10x12_mixed", mixed((8, 9, 10, 12)), mode="rel", eb=1e-4, block=3, lo=1, lo2=0, reg=1)
    # a last block of extent 1 along an axis (33 = 2 * 16 + 1): the
    # regression predictor declines it and the 1-layer Lorenzo fallback runs
    add("r2d_33x40_edge1", mixed((33, 40)), mode="rel", eb=1e-3, lo=1, lo2=0, reg=1)
    return c


def ctest_lorenzo_cases():
    """CompressTest fixtures of the Lorenzo variant (general frontend,
    lorenzo + lorenzo2, blockSize 5, radius 65536)."""
    r = np.random.default_rng(77)
    c = []
    f3 = smooth(r, (34, 34, 51), 2)
    c.append(("k3d_17", f3, [(0, 0, 0), (16, 16, 16), (0, 16, 32), (16, 0, 34)], 16, dict(eb=1e-3)))
    c.append(("k3d_17_small_eb", f3, [(5, 6, 7), (16, 16, 30)], 16, dict(eb=3e-6)))
    f33 = smooth(r, (66, 34, 34), 2)
    c.append(("k3d_33", f33, [(0, 0, 0), (32, 0, 0), (33, 1, 1)], 32, dict(eb=5e-4)))
    f2 = smooth(r, (130, 131), 2)
    c.append(("k2d_65", f2, [(0, 0), (64, 64), (65, 0)], 64, dict(eb=1e-3)))
    c.append(("k2d_65_wide_codes", f2, [(0, 0), (65, 66)], 64, dict(eb=2e-7)))
    return c


def sz3_cases():
    """Default-mode (SZ3) tuning fixtures: a field, its sample, block
    interpolation buffers of the sample and optimize_quant_invl_3d."""
    r = np.random.default_rng(31337)
    c = []
    c.append(("z1d_3000", smooth(r, (3000,), 2), 1e-3))
    c.append(("z2d_300x410", smooth(r, (300, 410), 2), 1e-3))
    c.append(("z2d_97x250_noisy", (smooth(r, (97, 250), 1) + 0.05 * r.standard_normal((97, 250))).astype(np.float32),
              1e-2))
    c.append(("z3d_66x64x70", smooth(r, (66, 64, 70), 2), 1e-3))
    c.append(("z3d_48_noisy", (smooth(r, (48, 48, 48), 1) + 0.02 * r.standard_normal((48, 48, 48))).astype(np.float32),
              1e-3))
    c.append(("z3d_60x40x50", smooth(r, (60, 40, 50), 3), 1e-2))
    return c


def harness_args(inp, out, dims, kw):
    a = [HARNESS, "interp", f"in={inp}", f"out={out}", "dims=" + ",".join(map(str, dims)),
         f"mode={kw.get('mode', 'abs')}", f"eb={kw['eb']!r}", f"algo={kw.get('algo', 1)}",
         f"dir={kw.get('dir', 0)}", f"alpha={kw.get('alpha', -1.0)!r}", f"beta={kw.get('beta', -1.0)!r}",
         f"maxstep={kw.get('maxstep', 0)}", f"bs={kw.get('bs', 32)}", f"qbins={kw.get('qbins', 65536)}"]
    if kw.get("ops"):
        a += ["ops=" + ",".join(map(str, kw["ops"])), "dirs=" + ",".join(map(str, kw["dirs"]))]
    if kw.get("tuning"):
        a += [f"tuning={kw['tuning']}", f"start={kw.get('start', 0)}", f"end={kw.get('end', 0)}"]
    return a


def main():
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref/ref_harness missing: build it with `make -C oracle ref`")
    tmp = tempfile.mkdtemp()
    only = sys.argv[1] if len(sys.argv) > 1 else ""  # name prefix: regenerate only those fixtures
    for name, x, kw in (cases() if not only else []):
        inp = os.path.join(tmp, "in.f32")
        out = os.path.join(tmp, "o")
        x.astype(np.float32).tofile(inp)
        subprocess.check_call(harness_args(inp, out, x.shape, kw))
        ld = lambda ext, dt: np.fromfile(out + ext, dt)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), input=x.astype(np.float32),
                            codes=ld(".codes", np.int32), recon=ld(".recon", np.float32),
                            unpred=ld(".unpred", np.float32), raw=ld(".raw", np.uint8),
                            stream=ld(".stream", np.uint8), eb=ld(".eb", np.float64),
                            params=np.array(json.dumps(dict(kw, dims=list(x.shape)))))
        print("wrote", name)
    for name, codes in (huffman_cases() if not only else []):
        inp = os.path.join(tmp, "codes.i32")
        out = os.path.join(tmp, "huff.bin")
        codes.tofile(inp)
        subprocess.check_call([HARNESS, "huffman", f"in={inp}", f"out={out}"])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), codes=codes, encoded=np.fromfile(out, np.uint8))
        print("wrote", name)
    for name, x, kw in (lorenzo_cases() if not only else []):
        inp = os.path.join(tmp, "in.f32")
        out = os.path.join(tmp, "l")
        x.astype(np.float32).tofile(inp)
        a = [HARNESS, "lorenzo", f"in={inp}", f"out={out}", "dims=" + ",".join(map(str, x.shape)),
             f"mode={kw.get('mode', 'abs')}", f"eb={kw['eb']!r}", f"qbins={kw.get('qbins', 65536)}",
             f"block={kw.get('block', 6)}", f"pd={kw.get('pd', 3)}", f"lo={kw.get('lo', 1)}",
             f"lo2={kw.get('lo2', 0)}", f"reg={kw.get('reg', 1)}"]
        subprocess.check_call(a)
        ld = lambda ext, dt: np.fromfile(out + ext, dt)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), input=x.astype(np.float32),
                            codes=ld(".codes", np.int32), recon=ld(".recon", np.float32),
                            raw=ld(".raw", np.uint8), stream=ld(".stream", np.uint8), eb=ld(".eb", np.float64),
                            params=np.array(json.dumps(dict(kw, dims=list(x.shape)))))
        print("wrote", name)
    for name, x, kw in general_cases():
        if only and not name.startswith(only):
            continue
        inp = os.path.join(tmp, "in.f32")
        out = os.path.join(tmp, "g")
        x.astype(np.float32).tofile(inp)
        a = [HARNESS, "general", f"in={inp}", f"out={out}", "dims=" + ",".join(map(str, x.shape)),
             f"mode={kw.get('mode', 'abs')}", f"eb={kw['eb']!r}", f"qbins={kw.get('qbins', 65536)}",
             f"lo={kw.get('lo', 1)}", f"lo2={kw.get('lo2', 1)}", f"reg={kw.get('reg', 0)}"]
        if "block" in kw:
            a.append(f"block={kw['block']}")
        subprocess.check_call(a)
        ld = lambda ext, dt: np.fromfile(out + ext, dt)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), input=x.astype(np.float32),
                            recon=ld(".recon", np.float32), raw=ld(".raw", np.uint8),
                            stream=ld(".stream", np.uint8), eb=ld(".eb", np.float64),
                            params=np.array(json.dumps(dict(kw, dims=list(x.shape)))))
        print("wrote", name)
    for name, x, starts, sbs, kw in ctest_lorenzo_cases():
        if only and not name.startswith(only):
            continue
        inp = os.path.join(tmp, "field.f32")
        st = os.path.join(tmp, "starts.i64")
        raw = os.path.join(tmp, "ct.raw")
        x.astype(np.float32).tofile(inp)
        s3 = np.array([list(t) + [0] * (3 - len(t)) for t in starts], np.int64)
        s3.tofile(st)
        bd = ",".join([str(sbs + 1)] * x.ndim)
        a = [HARNESS, "ctest", "ctalgo=lorenzo", f"in={inp}", "fdims=" + ",".join(map(str, x.shape)),
             f"starts={st}", f"sbs={sbs}", "dims=" + bd, "mode=abs", f"eb={kw['eb']!r}", f"raw={raw}"]
        res = json.loads(subprocess.check_output(a))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), input=x.astype(np.float32), starts=s3,
                            raw=np.fromfile(raw, np.uint8), square_error=np.array([res["square_error"]]),
                            ref_zstd_bytes=np.array([res["bytes"]]),
                            params=np.array(json.dumps(dict(kw, sbs=sbs, dims=list(x.shape)))))
        print("wrote", name)
    for name, x, rel in sz3_cases():
        if only and not name.startswith(only):
            continue
        inp = os.path.join(tmp, "in.f32")
        out = os.path.join(tmp, "sample.f32")
        x.astype(np.float32).tofile(inp)
        dims = ",".join(map(str, x.shape))
        res = json.loads(subprocess.check_output([HARNESS, "sampling", f"in={inp}", f"out={out}", f"dims={dims}"]))
        sample = np.fromfile(out, np.float32) if res["sampled"] else x.ravel()
        sdims = res["sample_dims"] if res["sampled"] else list(x.shape)
        sample.astype(np.float32).tofile(out)
        eb = rel * float(np.float32(x.max() - x.min()))
        block = max(res["sampling_block"], 6)
        nf = {1: 1, 2: 2, 3: 6}[x.ndim]
        raws, sizes, keys = [], [], []
        for algo in (0, 1):
            for d in range(nf):
                raw = os.path.join(tmp, "bi.raw")
                b = json.loads(subprocess.check_output(
                    [HARNESS, "blockinterp", f"in={out}", "dims=" + ",".join(map(str, sdims)), f"eb={eb!r}",
                     f"block={block}", f"algo={algo}", f"dir={d}", f"raw={raw}"]))
                raws.append(np.fromfile(raw, np.uint8))
                sizes.append(b["bytes"])
                keys.append((algo, d))
        q = []
        if x.ndim == 3:
            for f in (1.0, 0.01):
                q.append(json.loads(subprocess.check_output(
                    [HARNESS, "quantinvl", f"in={inp}", f"dims={dims}", f"eb={eb * f!r}"]))["quantbin"])
        lens = np.array([len(v) for v in raws], np.int64)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), input=x.astype(np.float32),
                            sample=sample.astype(np.float32), raws=np.concatenate(raws), raw_lens=lens,
                            ref_zstd_bytes=np.array(sizes, np.int64), keys=np.array(keys, np.int32),
                            quantbin=np.array(q, np.int32),
                            params=np.array(json.dumps(dict(rel=rel, eb=eb, block=block, dims=list(x.shape),
                                                            sampled=res["sampled"], sample_dims=sdims,
                                                            sampling_block=res["sampling_block"]))))
        print("wrote", name)
    if only:
        return
    for name, x, starts, sbs, kw in ctest_cases():
        inp = os.path.join(tmp, "field.f32")
        st = os.path.join(tmp, "starts.i64")
        raw = os.path.join(tmp, "ct.raw")
        x.astype(np.float32).tofile(inp)
        s3 = np.array([list(t) + [0] * (3 - len(t)) for t in starts], np.int64)
        s3.tofile(st)
        bd = ",".join([str(sbs + 1)] * x.ndim)
        a = [HARNESS, "ctest", f"in={inp}", "fdims=" + ",".join(map(str, x.shape)), f"starts={st}", f"sbs={sbs}",
             "dims=" + bd, "mode=abs", f"eb={kw['eb']!r}", f"alpha={kw['alpha']!r}", f"beta={kw['beta']!r}",
             f"maxstep={kw['maxstep']}", f"algo={kw.get('algo', 1)}", f"dir={kw.get('dir', 0)}", f"raw={raw}"]
        if kw.get("ops"):
            a += ["ops=" + ",".join(map(str, kw["ops"])), "dirs=" + ",".join(map(str, kw["dirs"]))]
        res = json.loads(subprocess.check_output(a))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), input=x.astype(np.float32), starts=s3,
                            raw=np.fromfile(raw, np.uint8), square_error=np.array([res["square_error"]]),
                            ref_zstd_bytes=np.array([res["bytes"]]),
                            params=np.array(json.dumps(dict(kw, sbs=sbs, dims=list(x.shape)))))
        print("wrote", name)
    for name, x, kw in tuning_cases():
        inp = os.path.join(tmp, "in.f32")
        out = os.path.join(tmp, "t")
        x.astype(np.float32).tofile(inp)
        subprocess.check_call(harness_args(inp, out, x.shape, kw))
        ld = lambda ext, dt: np.fromfile(out + ext, dt)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), input=x.astype(np.float32),
                            codes=ld(".codes", np.int32), recon=ld(".recon", np.float32),
                            loss=ld(".loss", np.float64), level_counts=ld(".levelcounts", np.uint64),
                            params=np.array(json.dumps(dict(kw, dims=list(x.shape)))))
        print("wrote", name)


if __name__ == "__main__":
    main()
