"""Loading helpers for the golden fixtures written by tests/golden/make_golden.py."""
import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names(prefix):
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))


INTERP = names("c")
HUFFMAN = names("h_")
TUNING = names("t")
LORENZO = names("l")
CTEST = names("q")
GENERAL = names("g")
SZ3 = names("z")
CTEST_LORENZO = names("k")
REGRESSION = names("r")


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    if "params" in d:
        d["params"] = json.loads(str(d["params"]))
    return d


def param_kwargs(case):
    """Engine parameters of a fixture, with its absolute error bound."""
    p = case["params"]
    if "eb" in case and np.asarray(case["eb"]).size:
        eb = float(np.asarray(case["eb"]).ravel()[0])
    elif p.get("mode", "abs") == "abs":
        eb = float(p["eb"])
    else:
        x = case["input"].ravel()
        eb = float(p["eb"]) * float(np.float32(x.max() - x.min()))
    return dict(dims=p["dims"], eb=eb, algo=p.get("algo", 1), direction=p.get("dir", 0),
                alpha=p.get("alpha", -1.0), beta=p.get("beta", -1.0), max_step=p.get("maxstep", 0),
                block_size=p.get("bs", 32), radius=p.get("qbins", 65536) // 2,
                ops=tuple(p.get("ops", ())), dirs=tuple(p.get("dirs", ())))


def lr_kwargs(case):
    """Lorenzo/regression engine parameters of an ``l*`` fixture, absolute bound."""
    p = case["params"]
    return dict(dims=p["dims"], eb=float(np.asarray(case["eb"]).ravel()[0]), block_size=p.get("block", 6),
                pred_dim=p.get("pd", 3), lorenzo=p.get("lo", 1), lorenzo2=p.get("lo2", 0),
                regression=p.get("reg", 1), radius=p.get("qbins", 65536) // 2)


def gf_kwargs(case):
    """General-frontend parameters of a ``g*`` fixture, absolute bound; the
    block size defaults to the Config's for N (Config.hpp:71)."""
    p = case["params"]
    n = len(p["dims"])
    return dict(dims=p["dims"], eb=float(np.asarray(case["eb"]).ravel()[0]),
                block_size=p.get("block", 128 if n == 1 else (16 if n == 2 else 6)),
                lorenzo=p.get("lo", 1), lorenzo2=p.get("lo2", 1), radius=p.get("qbins", 65536) // 2,
                regression=p.get("reg", 0))


def trailer(stream):
    """The Config trailer of an SZ_compress stream (sz.hpp:54-91)."""
    n = int(np.frombuffer(stream[-4:], np.int32)[0])
    return stream[-4 - n:]


# Struct padding the reference writes with whatever its memory held
# (meta_params bytes 11 and 41-43, meanInfo bytes 1-3 at offset 52).
LR_RAW_PADDING = (11, 41, 42, 43, 53, 54, 55)


def lr_raw_equal(a, b):
    a = np.frombuffer(a, np.uint8).copy()
    b = np.frombuffer(b, np.uint8).copy()
    if a.size != b.size:
        return False
    a[list(LR_RAW_PADDING)] = 0
    b[list(LR_RAW_PADDING)] = 0
    return bool(np.array_equal(a, b))


def ctest_params(case, make):
    """Engine parameters (for `make`: oracle.make_params or faz_amd.make_params)
    of a CompressTest fixture: the sample-block shape and the candidate."""
    p = case["params"]
    n = len(p["dims"])
    return make((p["sbs"] + 1,) * n, p["eb"], algo=p.get("algo", 1), direction=p.get("dir", 0), alpha=p["alpha"],
                beta=p["beta"], max_step=p["maxstep"], block_size=32, radius=32768, ops=tuple(p.get("ops", ())),
                dirs=tuple(p.get("dirs", ())))
