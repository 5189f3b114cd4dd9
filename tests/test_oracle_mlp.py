"""The MLP oracle's backward pass against central finite differences (CPU)."""
import numpy as np

from oracle import mlp_ref


def _net(seed=0, K=(32, 32, 32), N=(32, 32, 32)):
    r = np.random.default_rng(seed)
    p = []
    for k, n in zip(K, N):
        p.append(r.normal(0, 1 / np.sqrt(k), (k, n)).ravel())
        p.append(r.normal(0, 0.1, n))
    return np.concatenate(p), list(K), list(N)


def test_exact_gradients_match_finite_differences():
    params, K, N = _net()
    r = np.random.default_rng(1)
    x = r.normal(size=(16, 20))
    t = r.normal(size=(16, 3))
    loss, g = mlp_ref.gradients(params, K, N, x, t, round=False)
    idx = r.choice(params.size, 60, replace=False)
    for i in idx:
        h = 1e-6
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        l1, _ = mlp_ref.gradients(p1, K, N, x, t, round=False)
        l2, _ = mlp_ref.gradients(p2, K, N, x, t, round=False)
        fd = (l1 - l2) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-40], np.float32)
    b = mlp_ref.bf16(x)
    # bf16 keeps 8 significant bits; ties go to even
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == np.float32(1.0078125)
    assert abs(b[3] - x[3]) <= abs(x[3]) * 2 ** -8
    assert (b.view(np.uint32) & 0xFFFF == 0).all()


def test_rounded_gradients_close_to_exact():
    params, K, N = _net(2)
    r = np.random.default_rng(3)
    x = r.normal(size=(32, 30))
    t = r.normal(size=(32, 2))
    _, g_exact = mlp_ref.gradients(params, K, N, x, t, round=False)
    _, g_round = mlp_ref.gradients(params, K, N, x, t, round=True)
    assert np.linalg.norm(g_round - g_exact) <= 0.02 * np.linalg.norm(g_exact)
