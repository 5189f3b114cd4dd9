"""Python bindings of the MI355X FAZ engine (ctypes over the C-ABI in include/faz_gpu.h).

The compute path is the HIP library ``faz_amd/libfaz_amd.so``; importing this
package never falls back to anything else.  PyTorch is used only to hold
device memory (``tensor.data_ptr()`` is handed to the C-ABI).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfaz_amd.so")
CLI_PATH = os.path.join(HERE, "faz")

FAZ_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "unsupported", -4: "capacity", -5: "format"}


class FazError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status}: {ERRORS.get(status, '?')})")
        self.status = status


class InterpParams(ctypes.Structure):
    """Mirror of faz_interp_params."""

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
        ("tuning", ctypes.c_int),
        ("start_level", ctypes.c_int),
        ("end_level", ctypes.c_int),
    ]


class InterpOut(ctypes.Structure):
    _fields_ = [
        ("d_codes", ctypes.c_void_p),
        ("n_codes", ctypes.c_size_t),
        ("d_unpred", ctypes.c_void_p),
        ("n_unpred", ctypes.c_size_t),
        ("loss", ctypes.c_double),
        ("level_counts", ctypes.c_size_t * 64),
        ("levels", ctypes.c_int),
    ]


class HuffTable(ctypes.Structure):
    _fields_ = [
        ("offset", ctypes.c_int32),
        ("n_states", ctypes.c_int32),
        ("code", ctypes.c_void_p),
        ("len", ctypes.c_void_p),
    ]


class LorenzoParams(ctypes.Structure):
    """Mirror of faz_lorenzo_params."""

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


class LorenzoOut(ctypes.Structure):
    _fields_ = [
        ("d_codes", ctypes.c_void_p),
        ("n_codes", ctypes.c_size_t),
        ("d_unpred", ctypes.c_void_p),
        ("n_unpred", ctypes.c_size_t),
        ("d_selector", ctypes.c_void_p),
        ("n_blocks", ctypes.c_size_t),
        ("n_regression", ctypes.c_size_t),
    ]


class GeneralParams(ctypes.Structure):
    """Mirror of faz_general_params."""

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


class GeneralOut(ctypes.Structure):
    _fields_ = [
        ("d_codes", ctypes.c_void_p),
        ("n_codes", ctypes.c_size_t),
        ("d_outval", ctypes.c_void_p),
        ("d_selection", ctypes.c_void_p),
        ("n_selection", ctypes.c_size_t),
        ("n_regression", ctypes.c_size_t),
        ("passes", ctypes.c_int),
    ]


def make_general_params(dims, eb, block_size=None, lorenzo=True, lorenzo2=True, radius=32768, regression=False):
    """General-frontend parameters; block_size defaults to the Config's for N
    (128, 16, 6, 6)."""
    p = GeneralParams()
    p.N = len(dims)
    for i, d in enumerate(dims):
        p.dims[i] = int(d)
    p.eb = float(eb)
    if block_size is None:
        block_size = 128 if p.N == 1 else (16 if p.N == 2 else 6)
    p.block_size = int(block_size)
    p.lorenzo = int(bool(lorenzo))
    p.lorenzo2 = int(bool(lorenzo2))
    p.radius = int(radius)
    p.regression = int(bool(regression))
    return p


def make_lorenzo_params(dims, eb, block_size=6, pred_dim=3, lorenzo=True, lorenzo2=False, regression=True,
                        radius=32768):
    p = LorenzoParams()
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


class TuneIn(ctypes.Structure):
    """Mirror of faz_tune_in."""

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
        ("learned_margin", ctypes.c_double),
        ("learned_verify", ctypes.c_int),
        ("learned_lo", ctypes.c_void_p),
        ("learned_hi", ctypes.c_void_p),
        ("learned_margin_lorenzo", ctypes.c_double),
    ]


class TuneOut(ctypes.Structure):
    """Mirror of faz_tune_out."""

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
        ("ms", ctypes.c_double),
        ("use_interp", ctypes.c_int),
        ("learned_path", ctypes.c_int),
        ("n_verified", ctypes.c_int),
        ("level_res", ctypes.c_double * 32),
    ]


def tune_out_dict(o):
    lw, nt, nb = o.levelwise, min(o.n_tests, 128), o.n_blocks
    return dict(alpha=o.alpha, beta=o.beta, levelwise=lw, ops=list(o.ops[:lw]), dirs=list(o.dirs[:lw]),
                interp_algo=o.interp_algo, direction=o.direction, max_step=o.max_step,
                sample_block_size=o.sample_block_size, n_blocks=nb,
                block_starts=np.array([list(o.block_starts[i]) for i in range(min(nb, 4096))], np.int64),
                pred_loss=np.array([list(o.pred_loss[i]) for i in range(32)]),
                tests=[dict(alpha=o.test_alpha[i], beta=o.test_beta[i], eb=o.test_eb[i], bitrate=o.test_bitrate[i],
                            psnr=o.test_psnr[i], bytes=o.test_bytes[i]) for i in range(nt)],
                ms=getattr(o, "ms", 0.0), use_interp=o.use_interp, learned_path=o.learned_path,
                n_verified=o.n_verified, level_res=np.array(o.level_res[:lw]))


def make_params(dims, eb, algo=1, direction=0, alpha=-1.0, beta=-1.0, max_step=0, block_size=32,
                radius=32768, ops=(), dirs=()):
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


# Exported symbols declared in include/faz_gpu.h (checked by the CPU tests).
EXPORTS = [
    "faz_gpu_ctx_create", "faz_gpu_ctx_destroy", "faz_gpu_last_error", "faz_gpu_ctx_device",
    "faz_gpu_range_f32", "faz_gpu_interp_compress_f32", "faz_gpu_huffman_hist", "faz_gpu_huffman_encode",
    "faz_huffman_build", "faz_gpu_interp_decompress_raw", "faz_gpu_interp_decompress_dev", "faz_gpu_interp_raw_f32",
    "faz_sz_compress_f32", "faz_sz_decompress_f32", "faz_free", "faz_gpu_last_timings",
    "faz_gpu_compress_f32", "faz_gpu_fetch_raw", "faz_gpu_set_profiling", "faz_gpu_profile_read",
    "faz_gpu_profile_kernel",
    "faz_gpu_profile_bytes",
    "faz_gpu_lorenzo_compress_f32", "faz_gpu_lorenzo_raw_f32", "faz_gpu_lorenzo_decompress_raw",
    "faz_sz_compress_lorenzo_f32", "faz_gpu_tune_qoz_f32", "faz_gpu_compress_test_f32",
    "faz_sz_compress_qoz_f32",
    "faz_mlp_create", "faz_mlp_destroy", "faz_mlp_layout", "faz_mlp_forward", "faz_mlp_gradients",
    "faz_mlp_train_step", "faz_mlp_get_params", "faz_mlp_set_params",
    "faz_gpu_compress_test_batch_f32", "faz_gpu_block_features_f32", "faz_learned_input", "faz_learned_target",
    "faz_gpu_tune_learned_f32",
    "faz_gpu_general_compress_f32", "faz_gpu_general_raw_f32", "faz_gpu_general_decompress_raw",
    "faz_gpu_compress_test_lorenzo_f32", "faz_sz_compress_general_f32",
    "faz_gpu_batch_create", "faz_gpu_batch_destroy", "faz_gpu_batch_compress", "faz_gpu_batch_fetch_raw",
    "faz_gpu_batch_stats", "faz_sz_compress_batch_f32", "faz_gpu_batch_field",
    "faz_gpu_sampling_f32", "faz_gpu_block_interp_test_f32", "faz_gpu_quant_intervals_f32", "faz_gpu_sz3_tune_f32",
    "faz_sz_compress_sz3_f32", "faz_gpu_huffman_decode", "faz_gpu_huffman_encode_noalloc",
    "faz_gpu_noalloc_scratch_bytes", "faz_huffman_check_section", "faz_stream_info",
]


class SZ3Out(ctypes.Structure):
    """Mirror of faz_sz3_out."""

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


class BatchStats(ctypes.Structure):
    """Mirror of faz_batch_stats."""

    _fields_ = [
        ("calls", ctypes.c_uint64),
        ("captures", ctypes.c_uint64),
        ("updates", ctypes.c_uint64),
        ("instantiations", ctypes.c_uint64),
        ("replays", ctypes.c_uint64),
        ("tuned_fields", ctypes.c_uint64),
        ("range_ms", ctypes.c_double),
        ("pred_ms", ctypes.c_double),
        ("encode_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
    ]

LEARNED_FEATURES = 40
LEARNED_INPUTS = 54

_lib = None
_hip = None


def lib():
    """Loads libfaz_amd.so; raises when the HIP extension has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FazError(-2, f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C faz_amd)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.faz_gpu_ctx_create.argtypes = [ctypes.c_int, P(vp)]
        L.faz_gpu_ctx_destroy.argtypes = [vp]
        L.faz_gpu_ctx_destroy.restype = None
        L.faz_gpu_last_error.argtypes = [vp]
        L.faz_gpu_last_error.restype = ctypes.c_char_p
        L.faz_gpu_ctx_device.argtypes = [vp]
        L.faz_gpu_range_f32.argtypes = [vp, vp, sz, P(ctypes.c_float), P(ctypes.c_float)]
        L.faz_gpu_interp_compress_f32.argtypes = [vp, vp, vp, P(InterpParams), P(InterpOut)]
        L.faz_gpu_huffman_hist.argtypes = [vp, vp, sz, vp, vp]
        L.faz_gpu_huffman_encode.argtypes = [vp, vp, sz, P(HuffTable), vp, sz, P(sz)]
        L.faz_gpu_huffman_encode_noalloc.argtypes = [vp, vp, sz, P(HuffTable), vp, sz, P(sz)]
        L.faz_gpu_noalloc_scratch_bytes.argtypes = []
        L.faz_gpu_noalloc_scratch_bytes.restype = sz
        L.faz_huffman_check_section.argtypes = [vp, sz, P(sz), P(sz)]
        L.faz_stream_info.argtypes = [vp, sz, P(ctypes.c_int), vp, P(ctypes.c_int)]
        L.faz_huffman_build.argtypes = [vp, vp, P(ctypes.c_int32), P(ctypes.c_int32), vp, vp, vp, sz, P(sz)]
        L.faz_gpu_interp_decompress_raw.argtypes = [vp, vp, sz, ctypes.c_int, vp, sz]
        L.faz_gpu_interp_decompress_dev.argtypes = [vp, vp, sz, ctypes.c_int, vp, sz]
        L.faz_gpu_huffman_decode.argtypes = [vp, vp, sz, sz, ctypes.c_uint32, vp, P(sz)]
        L.faz_gpu_interp_raw_f32.argtypes = [vp, vp, vp, P(InterpParams), vp, sz, P(sz)]
        L.faz_sz_compress_f32.argtypes = [vp, P(InterpParams), ctypes.c_int, ctypes.c_double, vp, ctypes.c_int,
                                          P(vp), P(sz)]
        L.faz_sz_decompress_f32.argtypes = [vp, vp, sz, vp, sz, ctypes.c_int]
        L.faz_free.argtypes = [vp]
        L.faz_free.restype = None
        L.faz_gpu_last_timings.argtypes = [vp, P(ctypes.c_double), ctypes.c_int]
        L.faz_gpu_compress_f32.argtypes = [vp, vp, vp, P(InterpParams), ctypes.c_int, ctypes.c_double, P(sz)]
        L.faz_gpu_fetch_raw.argtypes = [vp, vp, sz]
        L.faz_gpu_set_profiling.argtypes = [vp, ctypes.c_int]
        L.faz_gpu_profile_read.argtypes = [vp, P(ctypes.c_double), P(ctypes.c_uint64), P(ctypes.c_uint64)]
        L.faz_gpu_profile_kernel.argtypes = [vp]
        L.faz_gpu_profile_bytes.argtypes = [vp]
        L.faz_gpu_profile_bytes.restype = ctypes.c_double
        L.faz_gpu_lorenzo_compress_f32.argtypes = [vp, vp, vp, P(LorenzoParams), P(LorenzoOut)]
        L.faz_gpu_lorenzo_raw_f32.argtypes = [vp, vp, vp, P(LorenzoParams), vp, sz, P(sz)]
        L.faz_gpu_lorenzo_decompress_raw.argtypes = [vp, vp, sz, vp, vp]
        L.faz_sz_compress_lorenzo_f32.argtypes = [vp, P(LorenzoParams), ctypes.c_int, ctypes.c_double, vp,
                                                  ctypes.c_int, P(vp), P(sz)]
        L.faz_gpu_tune_qoz_f32.argtypes = [vp, vp, P(TuneIn), P(TuneOut)]
        L.faz_gpu_general_compress_f32.argtypes = [vp, vp, vp, P(GeneralParams), P(GeneralOut)]
        L.faz_gpu_general_raw_f32.argtypes = [vp, vp, vp, P(GeneralParams), vp, sz, P(sz)]
        L.faz_gpu_general_decompress_raw.argtypes = [vp, vp, sz, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_int, vp, sz]
        L.faz_gpu_compress_test_lorenzo_f32.argtypes = [vp, vp, ctypes.c_int, vp, vp, sz, sz, ctypes.c_double, vp,
                                                        sz, P(sz), P(sz), P(ctypes.c_double)]
        L.faz_sz_compress_general_f32.argtypes = [vp, P(GeneralParams), ctypes.c_int, ctypes.c_double, vp,
                                                  ctypes.c_int, P(vp), P(sz)]
        L.faz_gpu_sampling_f32.argtypes = [vp, vp, ctypes.c_int, vp, vp, sz, vp, P(sz), P(ctypes.c_int)]
        L.faz_gpu_block_interp_test_f32.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_double, ctypes.c_int,
                                                    ctypes.c_int, sz, vp, sz, P(sz), P(sz)]
        L.faz_gpu_quant_intervals_f32.argtypes = [vp, vp, vp, ctypes.c_double, P(ctypes.c_int), P(sz)]
        L.faz_gpu_sz3_tune_f32.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_double, ctypes.c_double, P(SZ3Out)]
        L.faz_sz_compress_sz3_f32.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_double, vp, ctypes.c_int,
                                              P(vp), P(sz)]
        L.faz_gpu_batch_create.argtypes = [vp, P(InterpParams), P(vp), ctypes.c_int, ctypes.c_int, P(vp)]
        L.faz_gpu_batch_destroy.argtypes = [vp]
        L.faz_gpu_batch_destroy.restype = None
        L.faz_gpu_batch_compress.argtypes = [vp, vp, vp, vp, vp, vp]
        L.faz_gpu_batch_fetch_raw.argtypes = [vp, ctypes.c_int, vp, sz]
        L.faz_gpu_batch_stats.argtypes = [vp, P(BatchStats), vp]
        L.faz_sz_compress_batch_f32.argtypes = [vp, vp, vp, vp, vp]
        L.faz_gpu_batch_field.argtypes = [vp, ctypes.c_int, P(InterpParams)]
        L.faz_mlp_create.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                     P(vp)]
        L.faz_mlp_destroy.argtypes = [vp]
        L.faz_mlp_destroy.restype = None
        L.faz_mlp_layout.argtypes = [vp, P(ctypes.c_int), vp, vp]
        L.faz_mlp_forward.argtypes = [vp, vp, vp, ctypes.c_int, vp]
        L.faz_mlp_gradients.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp, sz, P(ctypes.c_double)]
        L.faz_mlp_train_step.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_float, P(ctypes.c_double)]
        L.faz_mlp_get_params.argtypes = [vp, vp, vp, sz, P(sz)]
        L.faz_mlp_set_params.argtypes = [vp, vp, vp, sz]
        L.faz_gpu_compress_test_batch_f32.argtypes = [vp, vp, vp, vp, sz, vp, sz, sz, ctypes.c_double, vp, vp]
        L.faz_gpu_block_features_f32.argtypes = [vp, vp, ctypes.c_int, vp, vp, sz, sz, ctypes.c_float,
                                                 ctypes.c_float, ctypes.c_double, vp]
        L.faz_learned_input.argtypes = [vp, sz, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, vp, vp, vp, ctypes.c_int, vp]
        L.faz_learned_target.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, vp]
        L.faz_gpu_tune_learned_f32.argtypes = [vp, vp, P(TuneIn), vp, vp, vp, P(TuneOut)]
        L.faz_sz_compress_qoz_f32.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_double, vp, ctypes.c_int,
                                              P(vp), P(sz)]
        L.faz_gpu_compress_test_f32.argtypes = [vp, vp, vp, P(InterpParams), vp, sz, sz, ctypes.c_double, vp, sz,
                                                P(sz), P(sz), P(ctypes.c_double)]
        _lib = L
    return _lib


def hip():
    """HIP runtime, for copying context-owned device arrays to the host."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def d2h(ptr, nbytes, dtype):
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
    if nbytes:
        rc = hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), nbytes, 2)
        if rc != 0:
            raise FazError(-2, f"hipMemcpy D2H failed ({rc})")
    return out


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def huffman_build(freq, first):
    """Host Huffman tree (faz_huffman_build): returns (offset, code, len, tree_bytes)."""
    freq = np.ascontiguousarray(freq, np.uint32)
    first = np.ascontiguousarray(first, np.uint32)
    assert freq.size == 65536 and first.size == 65536
    code = np.zeros(65536, np.uint64)
    ln = np.zeros(65536, np.uint8)
    off = ctypes.c_int32()
    ns = ctypes.c_int32()
    cap = 4 * 1024 * 1024
    tree = np.zeros(cap, np.uint8)
    tsz = ctypes.c_size_t()
    rc = lib().faz_huffman_build(_np_ptr(freq), _np_ptr(first), ctypes.byref(off), ctypes.byref(ns),
                                 _np_ptr(code), _np_ptr(ln), _np_ptr(tree), cap, ctypes.byref(tsz))
    if rc:
        raise FazError(rc, "faz_huffman_build")
    return off.value, code[:ns.value].copy(), ln[:ns.value].copy(), tree[:tsz.value].tobytes()


def huffman_check_section(section):
    """faz_huffman_check_section: (node count, section bytes); raises
    FazError(-5) for a malformed section."""
    b = np.frombuffer(bytes(section), np.uint8)
    nodes, used = ctypes.c_size_t(), ctypes.c_size_t()
    rc = lib().faz_huffman_check_section(_np_ptr(b) if b.size else None, b.size, ctypes.byref(nodes),
                                         ctypes.byref(used))
    if rc:
        raise FazError(rc, "huffman_check_section")
    return nodes.value, used.value


def stream_info(stream):
    """faz_stream_info: (N, dims, cmprAlgo) of a stream's Config trailer."""
    b = np.frombuffer(bytes(stream), np.uint8)
    N, algo = ctypes.c_int(), ctypes.c_int()
    dims = np.zeros(4, np.uint64)
    rc = lib().faz_stream_info(_np_ptr(b) if b.size else None, b.size, ctypes.byref(N), _np_ptr(dims),
                               ctypes.byref(algo))
    if rc:
        raise FazError(rc, "stream_info")
    return N.value, tuple(int(d) for d in dims[:N.value]), algo.value


class Engine:
    """One GPU context (stream + device scratch) of the C-ABI."""

    def __init__(self, device=0):
        # the bindings hand torch device tensors to the library: let torch's
        # HIP runtime claim the device first (it does not attach afterwards)
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except ImportError:
            pass
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.faz_gpu_ctx_create(int(device), ctypes.byref(h))
        if rc:
            raise FazError(rc, f"faz_gpu_ctx_create(device={device})")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self._L.faz_gpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            msg = self._L.faz_gpu_last_error(self.h)
            raise FazError(rc, f"{what}: {msg.decode() if msg else ''}")

    @staticmethod
    def _dev(t):
        import torch
        assert isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()
        return ctypes.c_void_p(t.data_ptr())

    def range(self, x):
        import torch
        torch.cuda.synchronize()
        mn, mx = ctypes.c_float(), ctypes.c_float()
        self._check(self._L.faz_gpu_range_f32(self.h, self._dev(x), x.numel(), ctypes.byref(mn),
                                              ctypes.byref(mx)), "range")
        return mn.value, mx.value

    def interp_compress(self, x, params, recon=None):
        """Returns dict(codes uint16, unpred float32, recon tensor, levels)."""
        import torch
        if recon is None:
            recon = torch.empty_like(x)
        torch.cuda.synchronize()
        out = InterpOut()
        self._check(self._L.faz_gpu_interp_compress_f32(self.h, self._dev(x), self._dev(recon),
                                                        ctypes.byref(params), ctypes.byref(out)),
                    "interp_compress")
        codes = d2h(out.d_codes, out.n_codes * 2, np.uint16)
        unpred = d2h(out.d_unpred, out.n_unpred * 4, np.float32)
        return dict(codes=codes, unpred=unpred, recon=recon, levels=out.levels, d_codes=out.d_codes,
                    n_codes=out.n_codes)

    def huffman_hist(self, d_codes, n):
        freq = np.zeros(65536, np.uint32)
        first = np.zeros(65536, np.uint32)
        self._check(self._L.faz_gpu_huffman_hist(self.h, ctypes.c_void_p(d_codes), n, _np_ptr(freq),
                                                 _np_ptr(first)), "huffman_hist")
        return freq, first

    def huffman_encode(self, d_codes, n, offset, code, ln, out_tensor):
        t = HuffTable()
        t.offset = offset
        t.n_states = code.size
        t.code = code.ctypes.data
        t.len = ln.ctypes.data
        nb = ctypes.c_size_t()
        self._check(self._L.faz_gpu_huffman_encode(self.h, ctypes.c_void_p(d_codes), n, ctypes.byref(t),
                                                   self._dev(out_tensor), out_tensor.numel() *
                                                   out_tensor.element_size(), ctypes.byref(nb)),
                    "huffman_encode")
        return nb.value

    def huffman_encode_noalloc(self, d_codes, n, offset, code, ln, out_tensor):
        """faz_gpu_huffman_encode_noalloc: the encode with the library's static
        scratch array (no device allocation); returns the byte count."""
        t = HuffTable()
        t.offset = offset
        t.n_states = code.size
        t.code = code.ctypes.data
        t.len = ln.ctypes.data
        nb = ctypes.c_size_t()
        self._check(self._L.faz_gpu_huffman_encode_noalloc(self.h, ctypes.c_void_p(d_codes), n, ctypes.byref(t),
                                                           self._dev(out_tensor), out_tensor.numel() *
                                                           out_tensor.element_size(), ctypes.byref(nb)),
                    "huffman_encode_noalloc")
        return nb.value

    def interp_raw(self, x, params, recon=None):
        import torch
        torch.cuda.synchronize()
        need = ctypes.c_size_t()
        rp = self._dev(recon) if recon is not None else None
        rc = self._L.faz_gpu_interp_raw_f32(self.h, self._dev(x), rp, ctypes.byref(params), None, 0,
                                            ctypes.byref(need))
        if rc != -4:
            self._check(rc, "interp_raw")
        buf = np.zeros(need.value, np.uint8)
        self._check(self._L.faz_gpu_interp_raw_f32(self.h, self._dev(x), rp, ctypes.byref(params),
                                                   _np_ptr(buf), buf.size, ctypes.byref(need)), "interp_raw")
        return buf[:need.value].tobytes()

    def huffman_decode(self, section, num, limit=0, out=None):
        """HuffmanEncoder::load + decode (HuffmanEncoder.hpp:236-296) on the
        device: `num` symbols of a saved Huffman section as an int32 tensor
        on the context's GPU (or into `out`), and the bytes the section took.
        limit > 0 rejects symbols outside [0, limit)."""
        import torch
        b = np.frombuffer(section, np.uint8)  # read only by the library
        if out is None:
            out = torch.empty(max(int(num), 1), dtype=torch.int32, device=f"cuda:{self.device}")
        assert out.dtype == torch.int32 and out.is_contiguous() and out.numel() >= int(num)
        used = ctypes.c_size_t()
        torch.cuda.synchronize()
        self._check(self._L.faz_gpu_huffman_decode(self.h, _np_ptr(b), b.size, int(num), int(limit),
                                                   ctypes.c_void_p(out.data_ptr()), ctypes.byref(used)),
                    "huffman_decode")
        return out[:int(num)], used.value

    def decompress_raw(self, raw, N, out):
        import torch
        b = np.frombuffer(raw, np.uint8)  # read only by the library
        torch.cuda.synchronize()
        self._check(self._L.faz_gpu_interp_decompress_raw(self.h, _np_ptr(b), b.size, int(N), self._dev(out),
                                                          out.numel()), "decompress_raw")
        return out

    def decompress_dev(self, d_raw, N, out, nbytes=None):
        """Decompression of an engine buffer held in device memory (a uint8
        tensor on this context's GPU)."""
        import torch
        assert d_raw.is_cuda and d_raw.dtype == torch.uint8 and d_raw.is_contiguous()
        n = d_raw.numel() if nbytes is None else int(nbytes)
        assert 0 <= n <= d_raw.numel()
        torch.cuda.synchronize()
        self._check(self._L.faz_gpu_interp_decompress_dev(self.h, ctypes.c_void_p(d_raw.data_ptr()), n, int(N),
                                                          self._dev(out), out.numel()), "decompress_dev")
        return out

    def sz_compress(self, data, params, eb, rel=True):
        """SZ_compress stream of `data` (numpy on host, or a device tensor)."""
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        if isinstance(data, np.ndarray):
            a = np.ascontiguousarray(data, np.float32)
            ptr, ondev = _np_ptr(a), 0
        else:
            import torch
            torch.cuda.synchronize()
            ptr, ondev = self._dev(data), 1
        self._check(self._L.faz_sz_compress_f32(self.h, ctypes.byref(params), 1 if rel else 0, float(eb), ptr,
                                                ondev, ctypes.byref(p), ctypes.byref(n)), "sz_compress")
        try:
            return ctypes.string_at(p, n.value)
        finally:
            self._L.faz_free(p)

    def sz_decompress(self, stream, num):
        out = np.zeros(num, np.float32)
        b = np.frombuffer(stream, np.uint8).copy()
        self._check(self._L.faz_sz_decompress_f32(self.h, _np_ptr(b), b.size, _np_ptr(out), num, 0),
                    "sz_decompress")
        return out

    def compress_device(self, x, params, rel, eb, recon=None):
        """Device-resident compression; returns the engine buffer size."""
        n = ctypes.c_size_t()
        rp = self._dev(recon) if recon is not None else None
        self._check(self._L.faz_gpu_compress_f32(self.h, self._dev(x), rp, ctypes.byref(params), 1 if rel else 0,
                                                 float(eb), ctypes.byref(n)), "compress_device")
        return n.value

    def fetch_raw(self, size):
        buf = np.zeros(size, np.uint8)
        self._check(self._L.faz_gpu_fetch_raw(self.h, _np_ptr(buf), size), "fetch_raw")
        return buf.tobytes()

    def set_profiling(self, on, phases=False):
        """on: HIP-event timing of the interpolation kernels (profile());
        phases: the compression's phase timings (timings())."""
        flags = (1 if on else 0) | (2 if phases else 0)
        self._check(self._L.faz_gpu_set_profiling(self.h, flags), "set_profiling")

    def profile(self):
        ms, la, pts = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self._L.faz_gpu_profile_read(self.h, ctypes.byref(ms), ctypes.byref(la), ctypes.byref(pts)),
                    "profile_read")
        kind = self._L.faz_gpu_profile_kernel(self.h)
        name = {0: "interp_tile3d_kernel", 1: "interp_pass_kernel", 2: "interp_level3d_kernel"}.get(kind, "?")
        return ms.value, la.value, pts.value, name, self._L.faz_gpu_profile_bytes(self.h)

    def timings(self):
        t = (ctypes.c_double * 10)()
        self._L.faz_gpu_last_timings(self.h, t, 10)
        return list(t)

    # --- 3-D Lorenzo/regression engine -------------------------------------
    def lorenzo_compress(self, x, params, recon=None):
        """Returns dict(codes uint16 block-major, unpred, selector uint8, n_regression, recon tensor)."""
        import torch
        if recon is None:
            recon = torch.empty_like(x)
        torch.cuda.synchronize()
        out = LorenzoOut()
        self._check(self._L.faz_gpu_lorenzo_compress_f32(self.h, self._dev(x), self._dev(recon),
                                                         ctypes.byref(params), ctypes.byref(out)),
                    "lorenzo_compress")
        return dict(codes=d2h(out.d_codes, out.n_codes * 2, np.uint16),
                    unpred=d2h(out.d_unpred, out.n_unpred * 4, np.float32),
                    selector=d2h(out.d_selector, out.n_blocks, np.uint8), n_regression=out.n_regression,
                    recon=recon)

    def lorenzo_raw(self, x, params, recon=None):
        import torch
        torch.cuda.synchronize()
        need = ctypes.c_size_t()
        rp = self._dev(recon) if recon is not None else None
        rc = self._L.faz_gpu_lorenzo_raw_f32(self.h, self._dev(x), rp, ctypes.byref(params), None, 0,
                                             ctypes.byref(need))
        if rc != -4:
            self._check(rc, "lorenzo_raw")
        buf = np.zeros(need.value, np.uint8)
        self._check(self._L.faz_gpu_lorenzo_raw_f32(self.h, self._dev(x), rp, ctypes.byref(params), _np_ptr(buf),
                                                    buf.size, ctypes.byref(need)), "lorenzo_raw")
        return buf[:need.value].tobytes()

    def lorenzo_decompress_raw(self, raw, dims, out):
        import torch
        b = np.frombuffer(raw, np.uint8)  # read only by the library
        d = np.array(dims, np.uint64)
        torch.cuda.synchronize()
        self._check(self._L.faz_gpu_lorenzo_decompress_raw(self.h, _np_ptr(b), b.size, _np_ptr(d), self._dev(out)),
                    "lorenzo_decompress_raw")
        return out

    def sz_compress_lorenzo(self, data, params, eb, rel=True):
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        if isinstance(data, np.ndarray):
            a = np.ascontiguousarray(data, np.float32)
            ptr, ondev = _np_ptr(a), 0
        else:
            import torch
            torch.cuda.synchronize()
            ptr, ondev = self._dev(data), 1
        self._check(self._L.faz_sz_compress_lorenzo_f32(self.h, ctypes.byref(params), 1 if rel else 0, float(eb),
                                                        ptr, ondev, ctypes.byref(p), ctypes.byref(n)),
                    "sz_compress_lorenzo")
        try:
            return ctypes.string_at(p, n.value)
        finally:
            self._L.faz_free(p)

    # --- general frontend (Lorenzo predictors, any N) ------------------------
    def general_compress(self, x, params, recon=None):
        """Returns dict(codes uint32 block-major, outval at zero codes,
        selection uint8 (255: none recorded), recon tensor, n_regression,
        passes)."""
        import torch
        if recon is None:
            recon = torch.empty_like(x)
        torch.cuda.synchronize()
        out = GeneralOut()
        self._check(self._L.faz_gpu_general_compress_f32(self.h, self._dev(x), self._dev(recon),
                                                         ctypes.byref(params), ctypes.byref(out)),
                    "general_compress")
        codes = d2h(out.d_codes, out.n_codes * 4, np.uint32)
        outval = d2h(out.d_outval, out.n_codes * 4, np.float32)
        return dict(codes=codes, unpred=outval[codes == 0].copy(),
                    selection=d2h(out.d_selection, out.n_selection, np.uint8), recon=recon,
                    n_regression=out.n_regression, passes=out.passes)

    def general_raw(self, x, params, recon=None):
        import torch
        torch.cuda.synchronize()
        need = ctypes.c_size_t()
        rp = self._dev(recon) if recon is not None else None
        rc = self._L.faz_gpu_general_raw_f32(self.h, self._dev(x), rp, ctypes.byref(params), None, 0,
                                             ctypes.byref(need))
        if rc != -4:
            self._check(rc, "general_raw")
        buf = np.zeros(need.value, np.uint8)
        self._check(self._L.faz_gpu_general_raw_f32(self.h, self._dev(x), rp, ctypes.byref(params), _np_ptr(buf),
                                                    buf.size, ctypes.byref(need)), "general_raw")
        return buf[:need.value].tobytes()

    def general_decompress_raw(self, raw, dims, out, lorenzo=True, lorenzo2=True, regression=False):
        import torch
        b = np.frombuffer(raw, np.uint8)  # read only by the library
        torch.cuda.synchronize()
        self._check(self._L.faz_gpu_general_decompress_raw(self.h, _np_ptr(b), b.size, len(dims), int(bool(lorenzo)),
                                                           int(bool(lorenzo2)), int(bool(regression)),
                                                           self._dev(out), int(np.prod(dims))),
                    "general_decompress_raw")
        return out

    def sz_compress_general(self, data, params, eb, rel=True):
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        if isinstance(data, np.ndarray):
            a = np.ascontiguousarray(data, np.float32)
            ptr, ondev = _np_ptr(a), 0
        else:
            import torch
            torch.cuda.synchronize()
            ptr, ondev = self._dev(data), 1
        self._check(self._L.faz_sz_compress_general_f32(self.h, ctypes.byref(params), 1 if rel else 0, float(eb),
                                                        ptr, ondev, ctypes.byref(p), ctypes.byref(n)),
                    "sz_compress_general")
        try:
            return ctypes.string_at(p, n.value)
        finally:
            self._L.faz_free(p)

    def compress_test_lorenzo(self, x, shape, starts, sbs, eb, want_raw=True):
        """CompressTest's Lorenzo variant: (pre-zstd buffer or None, compressed size, squared error)."""
        import torch
        torch.cuda.synchronize()
        fd = np.array(list(shape) + [1] * (3 - len(shape)), np.uint64)
        st = np.ascontiguousarray(np.asarray(starts, np.uint64).reshape(-1, 3))
        rs, by, sq = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_double()

        def call(buf, cap):
            return self._L.faz_gpu_compress_test_lorenzo_f32(self.h, self._dev(x), len(shape), _np_ptr(fd),
                                                             _np_ptr(st), st.shape[0], sbs, float(eb), buf, cap,
                                                             ctypes.byref(rs), ctypes.byref(by), ctypes.byref(sq))
        if not want_raw:
            self._check(call(None, 0), "compress_test_lorenzo")
            return None, by.value, sq.value
        cap = 64 * 1024 + st.shape[0] * (sbs + 1) ** len(shape) * 16
        raw = np.zeros(cap, np.uint8)
        rc = call(_np_ptr(raw), cap)
        if rc == -4:  # the buffer's size is known now
            raw = np.zeros(rs.value, np.uint8)
            rc = call(_np_ptr(raw), raw.size)
        self._check(rc, "compress_test_lorenzo")
        return raw[:rs.value].tobytes(), by.value, sq.value

    # --- default (SZ3) tuning mode --------------------------------------------
    def sampling(self, x, shape):
        """sampling<T,N> of a device field: (sample tensor or None, sample dims, sampling block)."""
        import torch
        torch.cuda.synchronize()
        d = np.array(list(shape), np.uint64)
        out = torch.empty(x.numel(), dtype=torch.float32, device=x.device)
        sd = np.zeros(4, np.uint64)
        sb = ctypes.c_size_t()
        sampled = ctypes.c_int()
        self._check(self._L.faz_gpu_sampling_f32(self.h, self._dev(x), len(shape), _np_ptr(d), self._dev(out),
                                                 out.numel(), _np_ptr(sd), ctypes.byref(sb), ctypes.byref(sampled)),
                    "sampling")
        sdims = tuple(int(v) for v in sd[:len(shape)])
        n = int(np.prod(sdims))
        return (out[:n] if sampled.value else None), sdims, sb.value

    def block_interp_test(self, x, shape, eb, algo, direction, block):
        """Block interpolation test on a device array: (pre-zstd buffer, zstd-framed size)."""
        import torch
        torch.cuda.synchronize()
        d = np.array(list(shape), np.uint64)
        cap = 64 * 1024 + int(np.prod(shape)) * 16
        raw = np.zeros(cap, np.uint8)
        rs, by = ctypes.c_size_t(), ctypes.c_size_t()
        self._check(self._L.faz_gpu_block_interp_test_f32(self.h, self._dev(x), len(shape), _np_ptr(d), float(eb),
                                                          int(algo), int(direction), int(block), _np_ptr(raw), cap,
                                                          ctypes.byref(rs), ctypes.byref(by)), "block_interp_test")
        return raw[:rs.value].tobytes(), by.value

    def quant_intervals(self, x, shape, eb):
        import torch
        torch.cuda.synchronize()
        d = np.array(list(shape), np.uint64)
        q, n = ctypes.c_int(), ctypes.c_size_t()
        self._check(self._L.faz_gpu_quant_intervals_f32(self.h, self._dev(x), _np_ptr(d), float(eb), ctypes.byref(q),
                                                        ctypes.byref(n)), "quant_intervals")
        return q.value, n.value

    def sz3_tune(self, x, shape, abs_eb, rel_eb):
        import torch
        torch.cuda.synchronize()
        d = np.array(list(shape) + [1] * (4 - len(shape)), np.uint64)
        o = SZ3Out()
        self._check(self._L.faz_gpu_sz3_tune_f32(self.h, self._dev(x), len(shape), _np_ptr(d), float(abs_eb),
                                                 float(rel_eb), ctypes.byref(o)), "sz3_tune")
        r = {k: getattr(o, k) for k, _ in SZ3Out._fields_}
        r["sample_dims"] = list(o.sample_dims)
        r["interp_bytes"] = list(o.interp_bytes)
        return r

    def sz_compress_sz3(self, data, shape, eb, rel=True):
        """`faz` default-mode stream (SZ3 tuning, then the chosen engine)."""
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        d = np.array(list(shape), np.uint64)
        if isinstance(data, np.ndarray):
            a = np.ascontiguousarray(data, np.float32)
            ptr, ondev = _np_ptr(a), 0
        else:
            import torch
            torch.cuda.synchronize()
            ptr, ondev = self._dev(data), 1
        self._check(self._L.faz_sz_compress_sz3_f32(self.h, len(shape), _np_ptr(d), 1 if rel else 0, float(eb), ptr,
                                                    ondev, ctypes.byref(p), ctypes.byref(n)), "sz_compress_sz3")
        try:
            return ctypes.string_at(p, n.value)
        finally:
            self._L.faz_free(p)

    # --- exhaustive QoZ tuner -----------------------------------------------
    def tune_qoz(self, x, shape, abs_eb, rel_eb, rng, interp_block_size=32, max_step=0, levelwise=0,
                 sample_block_size=0, ab_list=0, test_lorenzo=0):
        """faz -q tuner on the device field x (flat tensor of `shape`)."""
        import torch
        torch.cuda.synchronize()
        t = TuneIn()
        t.N = len(shape)
        for i, d in enumerate(shape):
            t.dims[i] = int(d)
        t.abs_eb, t.rel_eb, t.rng = float(abs_eb), float(rel_eb), float(rng)
        t.interp_block_size, t.max_step, t.levelwise, t.sample_block_size = interp_block_size, max_step, \
            levelwise, sample_block_size
        t.ab_list, t.test_lorenzo = int(ab_list), int(test_lorenzo)
        o = TuneOut()
        self._check(self._L.faz_gpu_tune_qoz_f32(self.h, self._dev(x), ctypes.byref(t), ctypes.byref(o)), "tune_qoz")
        return tune_out_dict(o)

    def compress_test_batch(self, x, shape, starts, sbs, params_list, rng):
        """CompressTest of every candidate in params_list: (bytes, squared errors)."""
        import torch
        torch.cuda.synchronize()
        fd = np.array(list(shape) + [1] * (3 - len(shape)), np.uint64)
        st = np.ascontiguousarray(np.asarray(starts, np.uint64).reshape(-1, 3))
        arr = (InterpParams * len(params_list))(*params_list)
        nb = np.zeros(len(params_list), np.uint64)
        sq = np.zeros(len(params_list), np.float64)
        self._check(self._L.faz_gpu_compress_test_batch_f32(self.h, self._dev(x), _np_ptr(fd), ctypes.cast(arr,
                                                            ctypes.c_void_p), len(params_list), _np_ptr(st),
                                                            st.shape[0], sbs, float(rng), _np_ptr(nb), _np_ptr(sq)),
                    "compress_test_batch")
        return nb.astype(np.int64), sq

    def compress_test(self, x, shape, starts, sbs, params, rng):
        """One CompressTest: (pre-zstd buffer, compressed size, squared error)."""
        import torch
        torch.cuda.synchronize()
        fd = np.array(list(shape) + [1] * (3 - len(shape)), np.uint64)
        st = np.ascontiguousarray(np.asarray(starts, np.uint64).reshape(-1, 3))
        cap = 64 * 1024 + st.shape[0] * (sbs + 1) ** len(shape) * 16
        raw = np.zeros(cap, np.uint8)
        rs, by, sq = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_double()
        self._check(self._L.faz_gpu_compress_test_f32(self.h, self._dev(x), _np_ptr(fd), ctypes.byref(params),
                                                      _np_ptr(st), st.shape[0], sbs, float(rng), _np_ptr(raw), cap,
                                                      ctypes.byref(rs), ctypes.byref(by), ctypes.byref(sq)),
                    "compress_test")
        return raw[:rs.value].tobytes(), by.value, sq.value

    def sz_compress_qoz(self, data, shape, eb, rel=True):
        """`faz -q` stream: QoZ tuner + interpolation engine (numpy host data or a device tensor)."""
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        d = np.array(list(shape), np.uint64)
        if isinstance(data, np.ndarray):
            a = np.ascontiguousarray(data, np.float32)
            ptr, ondev = _np_ptr(a), 0
        else:
            import torch
            torch.cuda.synchronize()
            ptr, ondev = self._dev(data), 1
        self._check(self._L.faz_sz_compress_qoz_f32(self.h, len(shape), _np_ptr(d), 1 if rel else 0, float(eb), ptr,
                                                    ondev, ctypes.byref(p), ctypes.byref(n)), "sz_compress_qoz")
        try:
            return ctypes.string_at(p, n.value)
        finally:
            self._L.faz_free(p)

    # --- learned tuner ---------------------------------------------------------
    def block_features(self, x, shape, starts, sbs, vmin, rng, eb_norm):
        """FAZ_LEARNED_FEATURES statistics of every sample block (host array)."""
        import torch
        torch.cuda.synchronize()
        d = np.array(list(shape), np.uint64)
        st = np.ascontiguousarray(np.asarray(starts, np.uint64).reshape(-1, 3))
        out = np.zeros((st.shape[0], LEARNED_FEATURES), np.float32)
        self._check(self._L.faz_gpu_block_features_f32(self.h, self._dev(x), len(shape), _np_ptr(d), _np_ptr(st),
                                                       st.shape[0], sbs, float(vmin), float(rng), float(eb_norm),
                                                       _np_ptr(out)), "block_features")
        return out

    def tune_learned(self, x, shape, abs_eb, rel_eb, rng, mlp, mu, sigma, interp_block_size=32, test_lorenzo=0,
                     margin=-1.0, verify=0, lo=None, hi=None, ab_list=0, margin_lorenzo=-1.0):
        """The learned tuner: the network's choice among the alpha/beta (and,
        with test_lorenzo, Lorenzo) candidates; candidates predicted within
        `margin` bits/value of the best are verified by real CompressTests (at
        most `verify`), rows outside the standardised range [lo, hi] fall back
        to the exhaustive search."""
        import torch
        torch.cuda.synchronize()
        t = TuneIn()
        t.N = len(shape)
        for i, dd in enumerate(shape):
            t.dims[i] = int(dd)
        t.abs_eb, t.rel_eb, t.rng = float(abs_eb), float(rel_eb), float(rng)
        t.interp_block_size = interp_block_size
        t.ab_list, t.test_lorenzo = int(ab_list), int(test_lorenzo)
        t.learned_margin, t.learned_verify = float(margin), int(verify)
        t.learned_margin_lorenzo = float(margin_lorenzo)
        keep = []
        if lo is not None and hi is not None:
            lo = np.ascontiguousarray(lo, np.float32)
            hi = np.ascontiguousarray(hi, np.float32)
            keep = [lo, hi]
            t.learned_lo, t.learned_hi = lo.ctypes.data, hi.ctypes.data
        mu = np.ascontiguousarray(mu, np.float32)
        sigma = np.ascontiguousarray(sigma, np.float32)
        o = TuneOut()
        self._check(self._L.faz_gpu_tune_learned_f32(self.h, self._dev(x), ctypes.byref(t), mlp.h, _np_ptr(mu),
                                                     _np_ptr(sigma), ctypes.byref(o)), "tune_learned")
        return tune_out_dict(o)


class MLP:
    """Device MLP of the learned tuner (faz_mlp_* C-ABI): BF16 matrix-core
    products with FP32 accumulation, FP32 Adam."""

    def __init__(self, engine, in_dim, hidden=256, hidden_layers=3, out_dim=2, seed=0):
        self.e = engine
        self.L = engine._L
        h = ctypes.c_void_p()
        engine._check(self.L.faz_mlp_create(engine.h, in_dim, hidden, hidden_layers, out_dim, seed, ctypes.byref(h)),
                      "mlp_create")
        self.h = h
        self.in_dim, self.out_dim = in_dim, out_dim
        n = ctypes.c_int()
        K = np.zeros(32, np.int32)
        N = np.zeros(32, np.int32)
        self.L.faz_mlp_layout(self.h, ctypes.byref(n), _np_ptr(K), _np_ptr(N))
        self.K, self.N = [int(v) for v in K[:n.value]], [int(v) for v in N[:n.value]]
        self.n_params = sum(k * m + m for k, m in zip(self.K, self.N))

    def close(self):
        if getattr(self, "h", None):
            self.L.faz_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, x):
        import torch
        x = x.contiguous()
        y = torch.empty((x.shape[0], self.out_dim), dtype=torch.float32, device=x.device)
        torch.cuda.synchronize()
        self.e._check(self.L.faz_mlp_forward(self.e.h, self.h, self.e._dev(x), x.shape[0], self.e._dev(y)),
                      "mlp_forward")
        return y

    def gradients(self, x, t):
        import torch
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_double()
        torch.cuda.synchronize()
        self.e._check(self.L.faz_mlp_gradients(self.e.h, self.h, self.e._dev(x.contiguous()),
                                               self.e._dev(t.contiguous()), x.shape[0], _np_ptr(g), g.size,
                                               ctypes.byref(loss)), "mlp_gradients")
        return loss.value, g

    def train_step(self, x, t, lr):
        loss = ctypes.c_double()
        self.e._check(self.L.faz_mlp_train_step(self.e.h, self.h, self.e._dev(x), self.e._dev(t), x.shape[0],
                                                float(lr), ctypes.byref(loss)), "mlp_train_step")
        return loss.value

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        n = ctypes.c_size_t()
        self.e._check(self.L.faz_mlp_get_params(self.e.h, self.h, _np_ptr(p), p.size, ctypes.byref(n)),
                      "mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self.e._check(self.L.faz_mlp_set_params(self.e.h, self.h, _np_ptr(p), p.size), "mlp_set_params")


def learned_input(block_feats, eb_norm, rel_eb, alpha, beta, eb_scale, ops, dirs, level_res, lorenzo=False):
    """MLP input row of one candidate (faz_learned_input); level_res: the
    tuner's per-level mean prediction errors in units of the bound."""
    f = np.ascontiguousarray(block_feats, np.float32)
    o = np.ascontiguousarray(ops, np.uint8)
    d = np.ascontiguousarray(dirs, np.uint8)
    lr = np.ascontiguousarray(level_res, np.float64)
    assert lr.size >= len(o)
    row = np.zeros(LEARNED_INPUTS, np.float32)
    rc = lib().faz_learned_input(_np_ptr(f), f.shape[0], float(eb_norm), float(rel_eb), float(alpha), float(beta),
                                 float(eb_scale), len(o), _np_ptr(o), _np_ptr(d), _np_ptr(lr), 1 if lorenzo else 0,
                                 _np_ptr(row))
    if rc:
        raise FazError(rc, "learned_input")
    return row


def learned_target(bitrate, psnr, eb_norm, eb_scale):
    t = np.zeros(2, np.float32)
    rc = lib().faz_learned_target(float(bitrate), float(psnr), float(eb_norm), float(eb_scale), _np_ptr(t))
    if rc:
        raise FazError(rc, "learned_target")
    return t


class Batch:
    """Many fields of uneven shapes compressed together (faz_gpu_batch_*):
    inputs bound at creation (device tensors the caller refills in place),
    one graph of range launches, one graph of per-field prediction branches
    replayed while the bounds stay the same, per-field launch schedules picked
    by the autotuner."""

    def __init__(self, engine, params, tensors, autotune=True):
        import torch
        self.engine = engine
        self.n = len(params)
        self._params = (InterpParams * self.n)(*params)
        self._tensors = list(tensors)  # keeps the bound inputs alive
        ptrs = (ctypes.c_void_p * self.n)(*[t.data_ptr() for t in self._tensors])
        torch.cuda.synchronize()
        h = ctypes.c_void_p()
        engine._check(engine._L.faz_gpu_batch_create(engine.h, self._params, ptrs, self.n, 1 if autotune else 0,
                                                     ctypes.byref(h)), "batch_create")
        self.h = h

    def _modes(self, rel, eb):
        rel = [rel] * self.n if isinstance(rel, (bool, int)) else list(rel)
        eb = [eb] * self.n if isinstance(eb, (float, int)) else list(eb)
        return np.array([1 if r else 0 for r in rel], np.int32), np.array(eb, np.float64)

    def compress(self, rel=True, eb=1e-3):
        """Compresses every field on the device; returns the engine buffer sizes
        (the buffers stay on the device until compress_raw / fetch)."""
        import torch
        m, e = self._modes(rel, eb)
        sizes = np.zeros(self.n, np.uint64)
        torch.cuda.synchronize()
        self.engine._check(self.engine._L.faz_gpu_batch_compress(self.h, _np_ptr(m), _np_ptr(e), _np_ptr(sizes), None,
                                                                 None), "batch_compress")
        return sizes

    def fetch(self, i, size):
        buf = np.zeros(int(size), np.uint8)
        self.engine._check(self.engine._L.faz_gpu_batch_fetch_raw(self.h, i, _np_ptr(buf), buf.size),
                           "batch_fetch_raw")
        return buf.tobytes()

    def compress_raw(self, rel=True, eb=1e-3):
        """Engine buffers of every field: (list of bytes, absolute bounds, ranges)."""
        import torch
        m, e = self._modes(rel, eb)
        sizes = np.zeros(self.n, np.uint64)
        aeb = np.zeros(self.n, np.float64)
        rng = np.zeros(self.n, np.float64)
        torch.cuda.synchronize()
        L = self.engine._L
        self.engine._check(L.faz_gpu_batch_compress(self.h, _np_ptr(m), _np_ptr(e), _np_ptr(sizes), _np_ptr(aeb),
                                                    _np_ptr(rng)), "batch_compress")
        out = []
        for i in range(self.n):
            buf = np.zeros(int(sizes[i]), np.uint8)
            self.engine._check(L.faz_gpu_batch_fetch_raw(self.h, i, _np_ptr(buf), buf.size), "batch_fetch_raw")
            out.append(buf.tobytes())
        return out, aeb, rng

    def sz_compress(self, rel=True, eb=1e-3):
        """SZ_compress stream of every field."""
        import torch
        m, e = self._modes(rel, eb)
        ptrs = (ctypes.c_void_p * self.n)()
        sizes = (ctypes.c_size_t * self.n)()
        torch.cuda.synchronize()
        self.engine._check(self.engine._L.faz_sz_compress_batch_f32(self.h, _np_ptr(m), _np_ptr(e), ptrs, sizes),
                           "sz_compress_batch")
        out = []
        for i in range(self.n):
            try:
                out.append(ctypes.string_at(ptrs[i], sizes[i]))
            finally:
                self.engine._L.faz_free(ptrs[i])
        return out

    def stats(self):
        s = BatchStats()
        sched = np.zeros(self.n, np.int32)
        self.engine._check(self.engine._L.faz_gpu_batch_stats(self.h, ctypes.byref(s), _np_ptr(sched)), "batch_stats")
        d = {k: getattr(s, k) for k, _ in BatchStats._fields_}
        d["schedules"] = sched.tolist()
        return d

    def close(self):
        if getattr(self, "h", None):
            self.engine._L.faz_gpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
