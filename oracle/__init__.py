"""ctypes bindings for the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

The oracle is a scalar CPU restatement of the reference FAZ hot path
(interpolation predictor, linear quantizer, Huffman coder, engine buffer and
stream framing).  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
CPU-baseline leg import it, always as the checker, never as the thing shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")


class InterpParams(ctypes.Structure):
    """Mirror of oracle_interp_params (faz_oracle.h)."""

    _fields_ = [
        ("N", ctypes.c_int),
        ("dims", ctypes.c_size_t * 4),
        ("eb", ctypes.c_double),
        ("interp_algo", ctypes.c_int),
        ("direction", ctypes.c_int),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("max_step", ctypes.c_size_t),
        ("block_size", ctypes.c_int),
        ("radius", ctypes.c_int),
        ("levelwise", ctypes.c_int),
        ("ops", ctypes.c_uint8 * 32),
        ("dirs", ctypes.c_uint8 * 32),
    ]


def make_params(dims, eb, algo=1, direction=0, alpha=-1.0, beta=-1.0, max_step=0,
                block_size=32, radius=32768, ops=(), dirs=()):
    p = InterpParams()
    p.N = len(dims)
    for i, d in enumerate(dims):
        p.dims[i] = int(d)
    p.eb = float(eb)
    p.interp_algo = int(algo)
    p.direction = int(direction)
    p.alpha = float(alpha)
    p.beta = float(beta)
    p.max_step = int(max_step)
    p.block_size = int(block_size)
    p.radius = int(radius)
    p.levelwise = len(ops)
    for i, (o, d) in enumerate(zip(ops, dirs)):
        p.ops[i] = int(o)
        p.dirs[i] = int(d)
    return p


class LRParams(ctypes.Structure):
    """Mirror of oracle_lr_params (faz_oracle.h)."""

    _fields_ = [
        ("dims", ctypes.c_size_t * 3),
        ("eb", ctypes.c_double),
        ("block_size", ctypes.c_int),
        ("pred_dim", ctypes.c_int),
        ("lorenzo", ctypes.c_int),
        ("lorenzo2", ctypes.c_int),
        ("regression", ctypes.c_int),
        ("radius", ctypes.c_int),
    ]


def make_lr_params(dims, eb, block_size=6, pred_dim=3, lorenzo=True, lorenzo2=False,
                   regression=True, radius=32768):
    p = LRParams()
    for i, d in enumerate(dims):
        p.dims[i] = int(d)
    p.eb = float(eb)
    p.block_size = int(block_size)
    p.pred_dim = int(pred_dim)
    p.lorenzo = int(bool(lorenzo))
    p.lorenzo2 = int(bool(lorenzo2))
    p.regression = int(bool(regression))
    p.radius = int(radius)
    return p


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.oracle_interp_compress.argtypes = [
            P(InterpParams), ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_void_p, P(ctypes.c_size_t), ctypes.c_void_p, P(ctypes.c_size_t),
            P(ctypes.c_double), ctypes.c_void_p, P(ctypes.c_int)]
        L.oracle_interp_raw.argtypes = [P(InterpParams), ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_size_t, P(ctypes.c_size_t)]
        L.oracle_interp_decompress_raw.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                                   ctypes.c_void_p, ctypes.c_size_t]
        L.oracle_huffman_encode.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                            ctypes.c_size_t, P(ctypes.c_size_t)]
        L.oracle_huffman_decode.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                            ctypes.c_size_t, P(ctypes.c_size_t)]
        L.oracle_compress_stream.argtypes = [P(InterpParams), ctypes.c_int, ctypes.c_double,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                             P(ctypes.c_size_t)]
        L.oracle_decompress_stream.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                               ctypes.c_size_t]
        L.oracle_range.argtypes = [ctypes.c_void_p, ctypes.c_size_t, P(ctypes.c_float),
                                   P(ctypes.c_float)]
        L.oracle_interp_seq_index.argtypes = [P(InterpParams), ctypes.c_void_p]
        vp, sz = ctypes.c_void_p, P(ctypes.c_size_t)
        L.oracle_lr_compress.argtypes = [P(LRParams), vp, vp, vp, vp, sz, vp, vp, vp, sz, vp, sz, vp, sz]
        L.oracle_lr_raw.argtypes = [P(LRParams), vp, vp, ctypes.c_size_t, sz]
        L.oracle_lr_decompress_raw.argtypes = [vp, ctypes.c_size_t, vp, vp]
        L.oracle_lr_compress_stream.argtypes = [P(LRParams), ctypes.c_int, ctypes.c_double, vp, vp,
                                                ctypes.c_size_t, sz]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def interp_compress(data, params, tuning=0, start=0, end=0):
    """Returns dict(codes, unpred, recon, loss, level_counts, levels)."""
    recon = np.ascontiguousarray(data, dtype=np.float32).ravel().copy()
    n = recon.size
    codes = np.zeros(n + 1, np.int32)
    unpred = np.zeros(n + 1, np.float32)
    nc = ctypes.c_size_t()
    nu = ctypes.c_size_t()
    loss = ctypes.c_double()
    lc = np.zeros(64, np.uint64)
    nl = ctypes.c_int()
    rc = lib().oracle_interp_compress(ctypes.byref(params), _ptr(recon), tuning, start, end,
                                      _ptr(codes), ctypes.byref(nc), _ptr(unpred), ctypes.byref(nu),
                                      ctypes.byref(loss), _ptr(lc), ctypes.byref(nl))
    assert rc == 0, rc
    return dict(codes=codes[:nc.value].copy(), unpred=unpred[:nu.value].copy(), recon=recon,
                loss=loss.value, level_counts=lc[:max(nl.value, 0)].copy(), levels=nl.value)


def interp_raw(data, params):
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    cap = 64 * 1024 + data.size * 16
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t()
    rc = lib().oracle_interp_raw(ctypes.byref(params), _ptr(data), _ptr(out), cap, ctypes.byref(n))
    assert rc == 0, rc
    return out[:n.value].tobytes()


def interp_decompress_raw(raw, dims):
    buf = np.frombuffer(raw, np.uint8).copy()
    num = int(np.prod(dims))
    out = np.zeros(num, np.float32)
    rc = lib().oracle_interp_decompress_raw(_ptr(buf), buf.size, len(dims), _ptr(out), num)
    assert rc == 0, rc
    return out


def huffman_encode(codes):
    codes = np.ascontiguousarray(codes, dtype=np.int32)
    cap = 64 * 1024 + codes.size * 16 + 16 * 200000
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t()
    rc = lib().oracle_huffman_encode(_ptr(codes), codes.size, _ptr(out), cap, ctypes.byref(n))
    assert rc == 0, rc
    return out[:n.value].tobytes()


def huffman_decode(buf, num):
    b = np.frombuffer(buf, np.uint8).copy()
    out = np.zeros(num, np.int32)
    used = ctypes.c_size_t()
    rc = lib().oracle_huffman_decode(_ptr(b), b.size, _ptr(out), num, ctypes.byref(used))
    assert rc == 0, rc
    return out, used.value


def compress_stream(data, params, eb, rel=True):
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    cap = 64 * 1024 + data.size * 16
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t()
    rc = lib().oracle_compress_stream(ctypes.byref(params), 1 if rel else 0, float(eb), _ptr(data),
                                      _ptr(out), cap, ctypes.byref(n))
    assert rc == 0, rc
    return out[:n.value].tobytes()


def decompress_stream(stream, num):
    b = np.frombuffer(stream, np.uint8).copy()
    out = np.zeros(num, np.float32)
    rc = lib().oracle_decompress_stream(_ptr(b), b.size, _ptr(out), num)
    assert rc == 0, rc
    return out


def data_range(data):
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    mn = ctypes.c_float()
    mx = ctypes.c_float()
    lib().oracle_range(_ptr(data), data.size, ctypes.byref(mn), ctypes.byref(mx))
    return mn.value, mx.value


def seq_index(params):
    num = 1
    for i in range(params.N):
        num *= params.dims[i]
    seq = np.zeros(num, np.int64)
    lib().oracle_interp_seq_index(ctypes.byref(params), _ptr(seq))
    return seq


def lr_compress(data, params):
    """3-D Lorenzo/regression engine. Returns dict(codes, recon, indicator,
    raw_coeffs, reg_codes, reg_coeffs, reg_unpred, unpred)."""
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    n = data.size
    bs = params.block_size
    nb = 1
    for i in range(3):
        nb *= (params.dims[i] - 1) // bs + 1
    codes = np.zeros(n, np.int32)
    recon = np.zeros(n, np.float32)
    ind = np.zeros(nb, np.int32)
    raw_c = np.zeros(4 * nb, np.float32)
    reg_codes = np.zeros(4 * nb, np.int32)
    reg_coeffs = np.zeros(4 * nb, np.float32)
    reg_unpred = np.zeros(4 * nb, np.float32)
    unpred = np.zeros(n, np.float32)
    nbk, nreg, nru, nu = (ctypes.c_size_t() for _ in range(4))
    rc = lib().oracle_lr_compress(ctypes.byref(params), _ptr(data), _ptr(codes), _ptr(recon), _ptr(ind),
                                  ctypes.byref(nbk), _ptr(raw_c), _ptr(reg_codes), _ptr(reg_coeffs),
                                  ctypes.byref(nreg), _ptr(reg_unpred), ctypes.byref(nru), _ptr(unpred),
                                  ctypes.byref(nu))
    assert rc == 0, rc
    assert nbk.value == nb
    r = nreg.value
    return dict(codes=codes, recon=recon, indicator=ind, raw_coeffs=raw_c.reshape(-1, 4),
                reg_codes=reg_codes[:4 * r].copy(), reg_coeffs=reg_coeffs[:4 * r].reshape(-1, 4).copy(),
                reg_unpred=reg_unpred[:nru.value].copy(), unpred=unpred[:nu.value].copy())


def lr_raw(data, params):
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    cap = 64 * 1024 + data.size * 16
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t()
    rc = lib().oracle_lr_raw(ctypes.byref(params), _ptr(data), _ptr(out), cap, ctypes.byref(n))
    assert rc == 0, rc
    return out[:n.value].tobytes()


def lr_decompress_raw(raw, dims):
    buf = np.frombuffer(raw, np.uint8).copy()
    d = np.array(dims, np.uint64)
    out = np.zeros(int(np.prod(dims)), np.float32)
    rc = lib().oracle_lr_decompress_raw(_ptr(buf), buf.size, _ptr(d), _ptr(out))
    assert rc == 0, rc
    return out


def lr_compress_stream(data, params, eb, rel=True):
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    cap = 64 * 1024 + data.size * 16
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t()
    rc = lib().oracle_lr_compress_stream(ctypes.byref(params), 1 if rel else 0, float(eb), _ptr(data),
                                         _ptr(out), cap, ctypes.byref(n))
    assert rc == 0, rc
    return out[:n.value].tobytes()


class TuneIn(ctypes.Structure):
    """Mirror of oracle_tune_in (faz_oracle.h)."""

    _fields_ = [
        ("N", ctypes.c_int),
        ("dims", ctypes.c_size_t * 4),
        ("abs_eb", ctypes.c_double),
        ("rel_eb", ctypes.c_double),
        ("rng", ctypes.c_double),
        ("interp_block_size", ctypes.c_int),
        ("max_step", ctypes.c_size_t),
        ("levelwise", ctypes.c_int),
        ("sample_block_size", ctypes.c_int),
        ("ab_list", ctypes.c_int),
        ("test_lorenzo", ctypes.c_int),
    ]


class TuneOut(ctypes.Structure):
    _fields_ = [
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("levelwise", ctypes.c_int),
        ("ops", ctypes.c_uint8 * 32),
        ("dirs", ctypes.c_uint8 * 32),
        ("interp_algo", ctypes.c_int),
        ("direction", ctypes.c_int),
        ("max_step", ctypes.c_size_t),
        ("sample_block_size", ctypes.c_int),
        ("n_blocks", ctypes.c_size_t),
        ("block_starts", (ctypes.c_size_t * 3) * 4096),
        ("pred_loss", (ctypes.c_double * 4) * 32),
        ("n_tests", ctypes.c_int),
        ("test_alpha", ctypes.c_double * 128),
        ("test_beta", ctypes.c_double * 128),
        ("test_eb", ctypes.c_double * 128),
        ("test_bitrate", ctypes.c_double * 128),
        ("test_psnr", ctypes.c_double * 128),
        ("test_bytes", ctypes.c_size_t * 128),
        ("use_interp", ctypes.c_int),
    ]


def qoz_tune(data, abs_eb, rel_eb, rng, interp_block_size=32, max_step=0, levelwise=0, sample_block_size=0,
             ab_list=0, test_lorenzo=0):
    """QoZ exhaustive tuner (faz -q).  Returns a dict of the decisions and of
    every intermediate result."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    L = lib()
    L.oracle_qoz_tune.argtypes = [ctypes.POINTER(TuneIn), ctypes.c_void_p, ctypes.POINTER(TuneOut)]
    t = TuneIn()
    t.N = data.ndim
    for i, d in enumerate(data.shape):
        t.dims[i] = d
    t.abs_eb, t.rel_eb, t.rng = float(abs_eb), float(rel_eb), float(rng)
    t.interp_block_size, t.max_step, t.levelwise, t.sample_block_size = interp_block_size, max_step, levelwise, \
        sample_block_size
    t.ab_list, t.test_lorenzo = int(ab_list), int(test_lorenzo)
    o = TuneOut()
    rc = L.oracle_qoz_tune(ctypes.byref(t), _ptr(data), ctypes.byref(o))
    assert rc == 0, rc
    lw = o.levelwise
    nt = min(o.n_tests, 128)
    nb = o.n_blocks
    return dict(alpha=o.alpha, beta=o.beta, levelwise=lw, ops=list(o.ops[:lw]), dirs=list(o.dirs[:lw]),
                interp_algo=o.interp_algo, direction=o.direction, max_step=o.max_step,
                sample_block_size=o.sample_block_size, n_blocks=nb,
                block_starts=np.array([list(o.block_starts[i]) for i in range(min(nb, 4096))], np.int64),
                pred_loss=np.array([list(o.pred_loss[i]) for i in range(32)]),
                tests=[dict(alpha=o.test_alpha[i], beta=o.test_beta[i], eb=o.test_eb[i], bitrate=o.test_bitrate[i],
                            psnr=o.test_psnr[i], bytes=o.test_bytes[i]) for i in range(nt)],
                use_interp=o.use_interp)


def compress_test(field, starts, sbs, params, rng):
    """CompressTest over the sample blocks at `starts`: (raw bytes, compressed size, squared error)."""
    field = np.ascontiguousarray(field, dtype=np.float32)
    L = lib()
    vp = ctypes.c_void_p
    L.oracle_compress_test.argtypes = [ctypes.POINTER(InterpParams), vp, vp, vp, ctypes.c_size_t, ctypes.c_size_t,
                                       ctypes.c_double, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_double)]
    fd = np.array(list(field.shape) + [1] * (3 - field.ndim), np.uint64)
    st = np.ascontiguousarray(np.asarray(starts, np.uint64).reshape(-1, 3))
    cap = 64 * 1024 + st.shape[0] * (sbs + 1) ** field.ndim * 16
    raw = np.zeros(cap, np.uint8)
    rs, by, sq = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_double()
    rc = L.oracle_compress_test(ctypes.byref(params), _ptr(field), _ptr(fd), _ptr(st), st.shape[0], sbs, float(rng),
                                _ptr(raw), cap, ctypes.byref(rs), ctypes.byref(by), ctypes.byref(sq))
    assert rc == 0, rc
    return raw[:rs.value].tobytes(), by.value, sq.value


class GFParams(ctypes.Structure):
    """Mirror of oracle_gf_params (faz_oracle.h)."""

    _fields_ = [
        ("N", ctypes.c_int),
        ("dims", ctypes.c_size_t * 4),
        ("eb", ctypes.c_double),
        ("block_size", ctypes.c_int),
        ("lorenzo", ctypes.c_int),
        ("lorenzo2", ctypes.c_int),
        ("radius", ctypes.c_int),
        ("regression", ctypes.c_int),
    ]


def make_gf_params(dims, eb, block_size=5, lorenzo=True, lorenzo2=True, radius=32768, regression=False):
    p = GFParams()
    p.N = len(dims)
    for i, d in enumerate(dims):
        p.dims[i] = int(d)
    p.eb = float(eb)
    p.block_size = int(block_size)
    p.lorenzo = int(bool(lorenzo))
    p.lorenzo2 = int(bool(lorenzo2))
    p.radius = int(radius)
    p.regression = int(bool(regression))
    return p


def _gf_lib():
    L = lib()
    vp, sz = ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)
    P = ctypes.POINTER(GFParams)
    L.oracle_gf_compress.argtypes = [P, vp, vp, vp, vp, sz, vp, sz]
    L.oracle_gf_raw.argtypes = [P, vp, vp, ctypes.c_size_t, sz]
    L.oracle_gf_decompress_raw.argtypes = [vp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, vp, ctypes.c_size_t]
    L.oracle_gf_compress_stream.argtypes = [P, ctypes.c_int, ctypes.c_double, vp, vp, ctypes.c_size_t, sz]
    L.oracle_compress_test_lorenzo.argtypes = [ctypes.c_int, ctypes.c_double, vp, vp, vp, ctypes.c_size_t,
                                               ctypes.c_size_t, vp, ctypes.c_size_t, sz, sz,
                                               ctypes.POINTER(ctypes.c_double)]
    return L


def gf_compress(data, params):
    """General frontend with the Lorenzo predictors. Returns dict(codes,
    recon, selection, unpred)."""
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    n = data.size
    nb = 1
    for i in range(params.N):
        nb *= (params.dims[i] - 1) // params.block_size + 1
    codes = np.zeros(n, np.int32)
    recon = np.zeros(n, np.float32)
    sel = np.zeros(nb, np.int32)
    unpred = np.zeros(n, np.float32)
    ns, nu = ctypes.c_size_t(), ctypes.c_size_t()
    rc = _gf_lib().oracle_gf_compress(ctypes.byref(params), _ptr(data), _ptr(codes), _ptr(recon), _ptr(sel),
                                      ctypes.byref(ns), _ptr(unpred), ctypes.byref(nu))
    assert rc == 0, rc
    return dict(codes=codes, recon=recon, selection=sel[:ns.value].copy(), unpred=unpred[:nu.value].copy())


def gf_raw(data, params):
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    cap = 64 * 1024 + data.size * 16
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t()
    rc = _gf_lib().oracle_gf_raw(ctypes.byref(params), _ptr(data), _ptr(out), cap, ctypes.byref(n))
    assert rc == 0, rc
    return out[:n.value].tobytes()


def gf_decompress_raw(raw, dims, lorenzo=True, lorenzo2=True, regression=False):
    buf = np.frombuffer(raw, np.uint8).copy()
    num = int(np.prod(dims))
    out = np.zeros(num, np.float32)
    rc = _gf_lib().oracle_gf_decompress_raw(_ptr(buf), buf.size, len(dims), int(bool(lorenzo)),
                                            int(bool(lorenzo2)), int(bool(regression)), _ptr(out), num)
    assert rc == 0, rc
    return out


def gf_compress_stream(data, params, eb, rel=True):
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    cap = 64 * 1024 + data.size * 16
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t()
    rc = _gf_lib().oracle_gf_compress_stream(ctypes.byref(params), 1 if rel else 0, float(eb), _ptr(data),
                                             _ptr(out), cap, ctypes.byref(n))
    assert rc == 0, rc
    return out[:n.value].tobytes()


def compress_test_lorenzo(field, starts, sbs, eb):
    """CompressTest's Lorenzo variant over the sample blocks at `starts`:
    (raw bytes, compressed size, squared error)."""
    field = np.ascontiguousarray(field, dtype=np.float32)
    fd = np.array(list(field.shape) + [1] * (3 - field.ndim), np.uint64)
    st = np.ascontiguousarray(np.asarray(starts, np.uint64).reshape(-1, 3))
    cap = 64 * 1024 + st.shape[0] * (sbs + 1) ** field.ndim * 16
    raw = np.zeros(cap, np.uint8)
    rs, by, sq = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_double()
    rc = _gf_lib().oracle_compress_test_lorenzo(field.ndim, float(eb), _ptr(field), _ptr(fd), _ptr(st), st.shape[0],
                                                sbs, _ptr(raw), cap, ctypes.byref(rs), ctypes.byref(by),
                                                ctypes.byref(sq))
    assert rc == 0, rc
    return raw[:rs.value].tobytes(), by.value, sq.value


class SZ3Out(ctypes.Structure):
    """Mirror of oracle_sz3_out (faz_oracle.h)."""

    _fields_ = [
        ("sampled", ctypes.c_int),
        ("sample_dims", ctypes.c_size_t * 4),
        ("sampling_block", ctypes.c_size_t),
        ("lorenzo_bytes", ctypes.c_size_t),
        ("interp_bytes", ctypes.c_size_t * 3),
        ("lorenzo_ratio", ctypes.c_double),
        ("interp_ratio", ctypes.c_double),
        ("use_interp", ctypes.c_int),
        ("interp_algo", ctypes.c_int),
        ("direction", ctypes.c_int),
        ("quantbin_estimate", ctypes.c_int),
        ("lorenzo_bytes_pd2", ctypes.c_size_t),
        ("lorenzo_bytes_16384", ctypes.c_size_t),
        ("quantbin", ctypes.c_int),
        ("final_lorenzo_ratio", ctypes.c_double),
    ]


def sz3_out_dict(o):
    d = {k: getattr(o, k) for k, _ in SZ3Out._fields_}
    d["sample_dims"] = list(o.sample_dims)
    d["interp_bytes"] = list(o.interp_bytes)
    return d


def _sz3_lib():
    L = lib()
    vp, sz = ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)
    L.oracle_sampling.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_size_t, vp, sz]
    L.oracle_block_interp_test.argtypes = [vp, ctypes.c_int, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_size_t, vp, ctypes.c_size_t, sz, sz]
    L.oracle_quant_intervals.argtypes = [vp, vp, ctypes.c_double, sz]
    L.oracle_sz3_tune.argtypes = [vp, ctypes.c_int, vp, ctypes.c_double, ctypes.c_double, ctypes.POINTER(SZ3Out)]
    L.oracle_sz3_compress_stream.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_double, vp, ctypes.c_size_t,
                                             sz, ctypes.POINTER(SZ3Out)]
    return L


def sampling(x):
    """sampling<T,N>: (sample array or None, sample dims, sampling block)."""
    x = np.ascontiguousarray(x, np.float32)
    d = np.array(x.shape, np.uint64)
    out = np.zeros(x.size, np.float32)
    sd = np.zeros(4, np.uint64)
    sb = ctypes.c_size_t()
    rc = _sz3_lib().oracle_sampling(_ptr(x), x.ndim, _ptr(d), _ptr(out), out.size, _ptr(sd), ctypes.byref(sb))
    assert rc in (0, 1), rc
    shape = tuple(int(v) for v in sd[:x.ndim])
    return (None if rc == 1 else out[:int(np.prod(shape))].reshape(shape).copy()), shape, sb.value


def block_interp_test(x, eb, algo, direction, block):
    """Block interpolation test: (pre-zstd buffer, zstd-framed size)."""
    x = np.ascontiguousarray(x, np.float32)
    d = np.array(x.shape, np.uint64)
    cap = 64 * 1024 + x.size * 16
    raw = np.zeros(cap, np.uint8)
    rs, by = ctypes.c_size_t(), ctypes.c_size_t()
    rc = _sz3_lib().oracle_block_interp_test(_ptr(x), x.ndim, _ptr(d), float(eb), int(algo), int(direction), int(block),
                                             _ptr(raw), cap, ctypes.byref(rs), ctypes.byref(by))
    assert rc == 0, rc
    return raw[:rs.value].tobytes(), by.value


def quant_intervals(x, eb):
    x = np.ascontiguousarray(x, np.float32)
    d = np.array(x.shape, np.uint64)
    n = ctypes.c_size_t()
    return _sz3_lib().oracle_quant_intervals(_ptr(x), _ptr(d), float(eb), ctypes.byref(n)), n.value


def sz3_tune(x, eb, rel):
    x = np.ascontiguousarray(x, np.float32)
    d = np.array(list(x.shape) + [1] * (4 - x.ndim), np.uint64)
    o = SZ3Out()
    rc = _sz3_lib().oracle_sz3_tune(_ptr(x), x.ndim, _ptr(d), float(eb), float(rel), ctypes.byref(o))
    assert rc == 0, rc
    return sz3_out_dict(o)


def sz3_compress_stream(x, eb, rel=True):
    """SZ_compress in the default mode: (stream, tuning decisions)."""
    x = np.ascontiguousarray(x, np.float32)
    d = np.array(list(x.shape) + [1] * (4 - x.ndim), np.uint64)
    cap = 64 * 1024 + x.size * 16
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t()
    o = SZ3Out()
    rc = _sz3_lib().oracle_sz3_compress_stream(_ptr(x), x.ndim, _ptr(d), 1 if rel else 0, float(eb), _ptr(out), cap,
                                               ctypes.byref(n), ctypes.byref(o))
    assert rc == 0, rc
    return out[:n.value].tobytes(), sz3_out_dict(o)
