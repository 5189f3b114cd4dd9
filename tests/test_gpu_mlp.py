"""GPU MLP kernels of the learned tuner against the CPU oracle (needs an MI355X).

Forward outputs and every parameter gradient of the BF16 matrix-core kernels
are compared with oracle/mlp_ref.py, which restates the same BF16 roundings
(FP32 accumulation order is the only difference, so the bar is a relative
tolerance: 1e-2 on norms, 0.999 cosine per layer); Adam training on a smooth
regression problem must cut the loss by 20x.
"""
import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _mlp(engine, din, hidden, layers, dout, seed):
    import faz_amd
    return faz_amd.MLP(engine, din, hidden, layers, dout, seed)


@pytest.mark.parametrize("din,hidden,layers,batch", [(47, 256, 3, 200), (40, 64, 1, 32), (70, 96, 2, 65)])
def test_forward_matches_oracle(engine, din, hidden, layers, batch):
    net = _mlp(engine, din, hidden, layers, 2, seed=din)
    r = np.random.default_rng(din)
    x = r.normal(size=(batch, din)).astype(np.float32)
    print(f"[gpu] mlp forward {din}->{hidden}x{layers} batch {batch}", flush=True)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    yr, _ = mlp_ref.forward(net.get_params().astype(np.float64), net.K, net.N, x.astype(np.float64))
    yr = yr[:batch, :2]
    assert np.linalg.norm(y - yr) <= 1e-2 * np.linalg.norm(yr) + 1e-6
    net.close()


@pytest.mark.parametrize("din,hidden,layers,batch", [(47, 256, 3, 256), (33, 64, 2, 96)])
def test_gradients_match_oracle(engine, din, hidden, layers, batch):
    net = _mlp(engine, din, hidden, layers, 2, seed=7)
    r = np.random.default_rng(8)
    x = r.normal(size=(batch, din)).astype(np.float32)
    t = r.normal(size=(batch, 2)).astype(np.float32)
    print(f"[gpu] mlp gradients {din}->{hidden}x{layers} batch {batch}", flush=True)
    loss, g = net.gradients(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    lr_, gr = mlp_ref.gradients(net.get_params().astype(np.float64), net.K, net.N, x.astype(np.float64),
                                t.astype(np.float64))
    assert abs(loss - lr_) <= 1e-3 * abs(lr_)
    assert np.linalg.norm(g - gr) <= 1e-2 * np.linalg.norm(gr)
    at = 0
    for k, n in zip(net.K, net.N):
        for size in (k * n, n):
            a, b = g[at:at + size], gr[at:at + size]
            at += size
            if np.linalg.norm(b) > 0:
                assert a @ b / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.999
    net.close()


def test_training_reduces_loss(engine):
    net = _mlp(engine, 8, 64, 2, 2, seed=3)
    r = np.random.default_rng(4)
    x = r.uniform(-1, 1, size=(512, 8)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], np.cos(x[:, 3]) - 0.5 * x[:, 4] ** 2], 1).astype(np.float32)
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    first = net.train_step(xd, td, 3e-3)
    for _ in range(400):
        last = net.train_step(xd, td, 3e-3)
    print(f"[gpu] mlp training loss {first:.4g} -> {last:.4g}", flush=True)
    assert last < first / 20
    net.close()
