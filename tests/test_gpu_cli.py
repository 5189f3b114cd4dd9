"""The `faz` command line (test/faz.cpp's flags) and a C++ program written
against include/faz/sz.hpp, run as processes on the GPU (needs an MI355X).

Bar: `-z` in the default mode (the SZ3 tuning mode, which the CLI runs when
-F is not given) writes the oracle's stream byte for byte; `-q` streams and
ALGO_INTERP `-c` config streams decode (on the oracle and through `-x`) to
the oracle's reconstruction with the same parameters; `-F 0` writes the
default mode's stream without the notice; ALGO_LORENZO_REG `-c` config
streams (2-D: the general frontend with the Config's 1-layer Lorenzo +
RegressionPredictor; 3-D: the fast frontend) equal the oracle's byte for
byte; `-a` prints the error statistics; `-F 1` (FZ, wavelets) fails loudly.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FAZ = os.path.join(ROOT, "faz_amd", "faz")
ROUNDTRIP = os.path.join(ROOT, "faz_amd", "sz_roundtrip")


def _field(shape, seed):
    r = np.random.default_rng(seed)
    x = r.standard_normal(shape)
    for ax in range(len(shape)):
        x = np.cumsum(np.cumsum(x, axis=ax), axis=ax)
    x = (x - x.min()) / (x.max() - x.min())
    return (x + 1e-3 * r.standard_normal(shape)).astype(np.float32)


def _dims(shape):
    """-1 .. -4 with the fastest dimension first."""
    return [f"-{len(shape)}"] + [str(d) for d in reversed(shape)]


def _run(exe, args, cwd, ok=True):
    print(f"[gpu] {os.path.basename(exe)} {' '.join(args)}", flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300, cwd=cwd)
    if ok:
        assert r.returncode == 0, r.stdout + r.stderr
    return r


@pytest.mark.parametrize("shape", [(64, 80, 72), (300, 410)])
def test_cli_default_mode_stream_and_decode(tmp_path, shape):
    x = _field(shape, 3)
    f = tmp_path / "f.dat"
    x.tofile(f)
    z, out = tmp_path / "f.faz", tmp_path / "f.out"
    r = _run(FAZ, ["-f", "-i", str(f), "-z", str(z)] + _dims(shape) + ["-M", "REL", "1e-3"], tmp_path)
    assert "FZ mode" in r.stderr and "compression ratio" in r.stdout
    s = z.read_bytes()
    ref, _ = oracle.sz3_compress_stream(x, 1e-3, rel=True)
    assert s == ref
    r = _run(FAZ, ["-f", "-s", str(z), "-x", str(out), "-i", str(f), "-a"] + _dims(shape), tmp_path)
    assert "Max absolute error" in r.stdout and "PSNR" in r.stdout
    dec = np.fromfile(out, np.float32)
    assert np.array_equal(dec.view(np.uint32), oracle.decompress_stream(s, x.size).view(np.uint32))
    eb = 1e-3 * float(np.float32(x.max() - x.min()))
    assert np.abs(dec.astype(np.float64) - x.ravel().astype(np.float64)).max() <= eb


def test_cli_qoz_stream_decodes_to_the_oracle(tmp_path):
    shape = (96, 96, 96)
    x = _field(shape, 4)
    f, z, out = tmp_path / "f.dat", tmp_path / "f.faz", tmp_path / "f.out"
    x.tofile(f)
    _run(FAZ, ["-f", "-q", "-i", str(f), "-z", str(z)] + _dims(shape) + ["-M", "REL", "1e-3"], tmp_path)
    s = z.read_bytes()
    rng = float(np.float32(x.max() - x.min()))
    o = oracle.qoz_tune(x, 1e-3 * rng, 1e-3, rng)
    p = oracle.make_params(shape, 1e-3 * rng, algo=o["interp_algo"], direction=o["direction"], alpha=o["alpha"],
                           beta=o["beta"], max_step=o["max_step"], ops=tuple(o["ops"]), dirs=tuple(o["dirs"]))
    want = oracle.interp_compress(x, p)["recon"]
    assert np.array_equal(oracle.decompress_stream(s, x.size).view(np.uint32), want.view(np.uint32))
    _run(FAZ, ["-f", "-s", str(z), "-x", str(out)] + _dims(shape), tmp_path)
    assert np.array_equal(np.fromfile(out, np.float32).view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("mode,bound", [("REL", "1e-4"), ("ABS", "2e-3")])
def test_cli_config_file_algo_interp(tmp_path, mode, bound):
    shape = (70, 66, 68)
    x = _field(shape, 5)
    f, z, out, cfg = tmp_path / "f.dat", tmp_path / "f.faz", tmp_path / "f.out", tmp_path / "faz.config"
    x.tofile(f)
    cfg.write_text("[GlobalSettings]\nCmprAlgo = ALGO_INTERP\n\n[AlgoSettings]\n"
                   "InterpolationAlgo = INTERP_ALGO_LINEAR\nInterpDirection = 3\nmaxStep = 32\n")
    _run(FAZ, ["-f", "-c", str(cfg), "-i", str(f), "-z", str(z)] + _dims(shape) + ["-M", mode, bound], tmp_path)
    s = z.read_bytes()
    import faz_amd
    assert faz_amd.stream_info(s) == (3, shape, 2)
    eb = float(bound) * (float(np.float32(x.max() - x.min())) if mode == "REL" else 1.0)
    want = oracle.interp_compress(x, oracle.make_params(shape, eb, algo=0, direction=3, max_step=32))["recon"]
    assert np.array_equal(oracle.decompress_stream(s, x.size).view(np.uint32), want.view(np.uint32))
    _run(FAZ, ["-f", "-s", str(z), "-x", str(out), "-t"] + _dims(shape), tmp_path)
    txt = np.loadtxt(out, dtype=np.float64)
    assert txt.size == x.size and np.abs(txt - want.ravel().astype(np.float64)).max() <= 1e-9 * max(1.0, eb)


def test_cli_sz3_mode_flag(tmp_path):
    shape = (64, 80, 72)
    x = _field(shape, 3)
    f, z = tmp_path / "f.dat", tmp_path / "f.faz"
    x.tofile(f)
    r = _run(FAZ, ["-f", "-F", "0", "-i", str(f), "-z", str(z)] + _dims(shape) + ["-M", "REL", "1e-3"], tmp_path)
    assert "FZ mode" not in r.stderr
    ref, _ = oracle.sz3_compress_stream(x, 1e-3, rel=True)
    assert z.read_bytes() == ref


@pytest.mark.parametrize("shape", [(300, 410), (60, 70, 80)])
def test_cli_config_file_algo_lorenzo_reg(tmp_path, shape):
    x = _field(shape, 7)
    i, j = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    ripple = 0.02 * np.sin(2 * np.pi * i / 9.0) * np.sin(2 * np.pi * j / 7.0)
    x = (x + (ripple if len(shape) == 2 else ripple[..., None]) * (np.arange(shape[-1]) > shape[-1] // 2)).astype(np.float32)
    f, z, out, cfg = tmp_path / "f.dat", tmp_path / "f.faz", tmp_path / "f.out", tmp_path / "faz.config"
    x.tofile(f)
    cfg.write_text("[GlobalSettings]\nCmprAlgo = ALGO_LORENZO_REG\n")
    _run(FAZ, ["-f", "-c", str(cfg), "-i", str(f), "-z", str(z)] + _dims(shape) + ["-M", "REL", "1e-3"], tmp_path)
    s = z.read_bytes()
    if len(shape) == 2:  # Config defaults for N = 2: blockSize 16, lorenzo + regression
        p = oracle.make_gf_params(shape, 0.0, block_size=16, lorenzo=True, lorenzo2=False, regression=True)
        ref = oracle.gf_compress_stream(x, p, 1e-3, rel=True)
    else:
        p = oracle.make_lr_params(shape, 0.0, block_size=6, pred_dim=3, lorenzo=True, lorenzo2=False, regression=True)
        ref = oracle.lr_compress_stream(x, p, 1e-3, rel=True)
    assert s == ref
    _run(FAZ, ["-f", "-s", str(z), "-x", str(out)] + _dims(shape), tmp_path)
    assert np.array_equal(np.fromfile(out, np.float32).view(np.uint32),
                          oracle.decompress_stream(s, x.size).view(np.uint32))


def test_cli_fz_mode_fails_loudly(tmp_path):
    shape = (40, 40, 40)
    x = _field(shape, 6)
    f = tmp_path / "f.dat"
    x.tofile(f)
    r = _run(FAZ, ["-f", "-F", "1", "-i", str(f), "-z", str(tmp_path / "f.faz")] + _dims(shape) + ["-M", "REL", "1e-3"],
             tmp_path, ok=False)
    assert r.returncode != 0 and "FZ" in r.stderr


@pytest.mark.parametrize("mode,shape", [("interp", (64, 80, 72)), ("qoz", (96, 96, 96)), ("lorenzo", (40, 48, 56)),
                                        ("interp", (300, 410))])
def test_cpp_program_against_sz_hpp(tmp_path, mode, shape):
    """SZ_compress<float> / SZ_decompress<float> from a program built against
    include/faz/sz.hpp: the stream decodes on the oracle to what the program
    decoded on the GPU, within the bound."""
    x = _field(shape, 7)
    f, z, out = tmp_path / "f.dat", tmp_path / "f.faz", tmp_path / "f.out"
    x.tofile(f)
    r = _run(ROUNDTRIP, [str(f), str(z), str(out), "1e-3", mode] + [str(d) for d in shape], tmp_path)
    s = z.read_bytes()
    assert f"bytes={len(s)}" in r.stdout
    dec = np.fromfile(out, np.float32)
    assert np.array_equal(dec.view(np.uint32), oracle.decompress_stream(s, x.size).view(np.uint32))
    eb = 1e-3 * float(np.float32(x.max() - x.min()))
    assert np.abs(dec.astype(np.float64) - x.ravel().astype(np.float64)).max() <= eb
