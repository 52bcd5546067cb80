"""The surrogate MLP's GPU forward and backward kernels (BF16 MFMA, FP32
accumulation) against the FP64 CPU oracle, and GPU training."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from phi_eng_amd.world import Surrogate3

pytestmark = pytest.mark.gpu


@pytest.fixture
def net(gpu):
    s = Surrogate3(seed=5)
    rng = np.random.default_rng(5)
    p = rng.normal(0, 0.25, size=M.NPARAM).astype(np.float32)
    s.set_params(p)
    yield s, p.astype(np.float64), rng
    s.close()


@pytest.mark.parametrize("rows", [1, 37, 256, 1000])
def test_forward_matches_oracle(net, rows):
    s, p, rng = net
    x = rng.normal(0, 1, size=(rows, 16)).astype(np.float32)
    y = s.forward(x)
    ref = M.forward(p, x.astype(np.float64), bf=True)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"forward rows={rows}: max err / max |y| = {err:.2e}", flush=True)
    assert err < 2e-3


@pytest.mark.parametrize("rows", [16, 300, 1024])
def test_backward_matches_oracle(net, rows):
    s, p, rng = net
    x = rng.normal(0, 1, size=(rows, 16)).astype(np.float32)
    dy = rng.normal(0, 1, size=(rows, 16)).astype(np.float32)
    g = s.backward(x, dy)
    ref = M.backward(p, x.astype(np.float64), dy.astype(np.float64), bf=True)
    for name, a, b in [("W1", M.OFF_W1, M.OFF_B1), ("b1", M.OFF_B1, M.OFF_W2), ("W2", M.OFF_W2, M.OFF_B2),
                       ("b2", M.OFF_B2, M.NPARAM)]:
        err = np.abs(g[a:b] - ref[a:b]).max() / np.abs(ref[a:b]).max()
        print(f"backward rows={rows} {name}: {err:.2e}", flush=True)
        assert err < 1e-2, name


def test_training_reduces_loss_and_beats_no_network(net):
    """Synthetic samples whose target needs the network: a_t = a0 + f(x) b0."""
    s, _, rng = net
    s.set_params(np.zeros(M.NPARAM, dtype=np.float32) + rng.normal(0, 0.05, M.NPARAM).astype(np.float32))
    n = 8192
    smp = np.zeros((n, 32), dtype=np.float32)
    x = rng.normal(0, 1, size=(n, 16)).astype(np.float32)
    x[:, 10:] = 0
    smp[:, :16] = x
    b = rng.normal(0, 1, size=(n, 3, 3)).astype(np.float32)
    smp[:, 16:25] = b.reshape(n, 9)
    a0 = rng.normal(0, 1, size=(n, 3)).astype(np.float32)
    smp[:, 25:28] = a0
    coef = np.tanh(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]
    smp[:, 28:31] = a0 + coef[:, None] * b[:, 0, :]
    smp[:, 31] = 1.0
    s.add_samples(smp)
    first, last, base = s.train(epochs=30, batch=1024, lr=3e-3, seed=1)
    cpu = M.sample_loss(s.params().astype(np.float64), smp.astype(np.float64))
    print(f"loss first {first:.4f} last {last:.4f} no-network {base:.4f} oracle-eval {cpu:.4f}", flush=True)
    assert last < 0.2 * base
    ev, ev0 = s.evaluate()
    assert abs(ev - cpu) <= 0.05 * cpu + 1e-4
