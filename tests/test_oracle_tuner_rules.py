"""What the reference's tuning rules fix, checked on the oracle's QoZ tuner
(oracle.qoz_tune) independently of its restated loop, from the intermediate
results it reports (the GPU tuner gets the same checks in
tests/test_gpu_tune.py):

* sampleBlocks with profiling and var_first off (SZI:1056-1086): block
  origins on the sbs-grid in raster order, every (1/rate)-th one kept, sbs
  from the QoZ defaults (SZI:1597-1613, 1776-1784);
* per level, the interpolator and direction are the argmin of that level's
  four mode-2 losses, first minimum in (op, direction) order (SZI:1952-2012);
* the alpha/beta search (SZI:2340-2420): candidates in alpha-major order
  without alpha > beta, the first candidate always taken, a re-test exactly
  when neither dominates, at the bound scaled by bitrate / best bitrate, and
  the final choice equal to the rule replayed on the reported tests.
"""
import math

import numpy as np
import pytest

import oracle

ALPHAS = {0: [1, 1.25, 1.5, 1.75, 2], 1: [1, 1.25, 1.5, 1.75, 2, 2.25, 2.5],
          2: [1, 1.25, 1.5, 1.75, 2, 2.25, 2.5, 2.75, 3]}
BETAS = [1.5, 2, 3, 4]


def _field(shape, seed, kind):
    r = np.random.default_rng(seed)
    x = r.standard_normal(shape)
    for ax in range(len(shape)):
        x = np.cumsum(x, axis=ax)
        if kind == "smooth":
            x = np.cumsum(x, axis=ax)
    x = (x - x.min()) / (x.max() - x.min())
    return (x + 1e-4 * r.standard_normal(shape)).astype(np.float32)


def expected_starts(shape):
    N = len(shape)
    rate = 0.01 if N == 2 else 0.005
    sbs = 64 if N == 2 else 32
    num = int(np.prod(shape))
    while sbs >= min(shape):
        sbs //= 2
    while sbs >= 32 and sbs ** N / num > 1.5 * rate:
        sbs //= 2
    stride = int(1.0 / rate)
    grid = [range(0, n - sbs, sbs) for n in shape]
    idx, out = 0, []
    for p in np.ndindex(*[len(g) for g in grid]):
        if idx % stride == 0:
            out.append([grid[a][p[a]] for a in range(N)] + [0] * (3 - N))
        idx += 1
    return sbs, np.array(out, np.int64).reshape(-1, 3)


def check_levels(t, N):
    dir_cands = (0, 1 if N == 2 else 5)  # {0, N! - 1}
    for level in range(1, t["levelwise"] + 1):
        losses = t["pred_loss"][level - 1]
        k = int(np.argmin(losses))  # first minimum in (op, direction) order
        assert t["ops"][level - 1] == k // 2, level
        assert t["dirs"][level - 1] == dir_cands[k % 2], level


def replay_alpha_beta(t, abs_eb, max_interp_level, ab_list=0):
    """SZI:2340-2420 on the reported tests; returns (alpha, beta) and checks
    the order and the re-test rule along the way."""
    tests = [x for x in t["tests"] if x["alpha"] >= 0]
    k = 0
    best_a, best_b, bestb, bestm = 1.0, 1.0, 9999.0, 0.0
    for a in ALPHAS[ab_list]:
        for b in BETAS:
            if a > b:
                continue
            cur = tests[k]
            k += 1
            assert (cur["alpha"], cur["beta"], cur["eb"]) == (a, b, abs_eb)
            br, m = cur["bitrate"], cur["psnr"]
            stop = a ** (max_interp_level - 1) <= b
            if m >= bestm and br <= bestb:
                best_a, best_b, bestb, bestm = a, b, br, m
            elif m <= bestm and br >= bestb:
                if stop:
                    break
                continue
            else:
                re = tests[k]
                k += 1
                assert (re["alpha"], re["beta"]) == (a, b)
                assert re["eb"] == abs_eb * (br / bestb)
                with np.errstate(divide="ignore", invalid="ignore"):  # IEEE division, as in C
                    slope = np.float64(m - re["psnr"]) / np.float64(br - re["bitrate"])
                    reg = slope * bestb + (m - slope * br)
                if reg > bestm:
                    best_a, best_b, bestb, bestm = a, b, br, m
            if stop:
                break
    assert k == len(tests)
    return best_a, best_b


CASES = [((96, 96, 96), 1e-3, "smooth", 0), ((130, 140, 150), 1e-4, "walk", 0), ((300, 400), 1e-3, "smooth", 0),
         ((200, 210, 90), 1e-2, "smooth", 1), ((120, 160, 100), 1e-5, "walk", 2), ((500, 450), 1e-2, "walk", 2)]


@pytest.mark.parametrize("shape,rel,kind,ab_list", CASES)
def test_oracle_tuner_follows_the_rules(shape, rel, kind, ab_list):
    x = _field(shape, 11, kind)
    rng = float(np.float32(x.max() - x.min()))
    eb = rel * rng
    t = oracle.qoz_tune(x, eb, rel, rng, ab_list=ab_list)
    sbs, starts = expected_starts(shape)
    assert t["sample_block_size"] == sbs
    assert np.array_equal(t["block_starts"], starts)
    check_levels(t, len(shape))
    mil = min(int(math.ceil(math.log2(max(shape)))), int(math.log2(t["max_step"])))
    assert t["tests"][0]["alpha"] == 1 and t["tests"][0]["beta"] == 1.5 and t["tests"][0]["eb"] == eb
    assert (t["alpha"], t["beta"]) == replay_alpha_beta(t, eb, mil, ab_list)
