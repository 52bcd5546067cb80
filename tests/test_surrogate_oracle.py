"""Pins of the MLP oracle (CPU): the analytic gradient equals central
finite differences of the real-valued network, and the bf16 emulation
rounds exactly like the device conversion (round to nearest even)."""
import numpy as np

from oracle import mlp_oracle as M


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, size=M.NPARAM)
    x = rng.normal(0, 1, size=(7, M.IN))
    dy = rng.normal(0, 1, size=(7, M.OUT))
    g = M.backward(p, x, dy, bf=False)
    eps = 1e-6
    for k in rng.choice(M.NPARAM, size=60, replace=False):
        e = np.zeros(M.NPARAM)
        e[k] = eps
        fd = ((M.forward(p + e, x, bf=False) - M.forward(p - e, x, bf=False)) * dy).sum() / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_gelu_closed_form():
    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    h, dg = M.gelu(z)
    exact = 0.5 * z * (1 + np.tanh(np.sqrt(2 / np.pi) * (z + 0.044715 * z ** 3)))
    assert np.allclose(h, exact, rtol=1e-15, atol=0)
    fd = (M.gelu(z + 1e-7)[0] - M.gelu(z - 1e-7)[0]) / 2e-7
    assert np.allclose(dg, fd, rtol=1e-6, atol=1e-9)


def test_bf16_rounding_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5e-3, 3.0e5], dtype=np.float32)
    r = M.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even
    assert r[2] == 1.0 + 4 * 2 ** -8  # tie -> even (up)
    assert r[3] == 1.0
    assert abs(r[4] + 2.5e-3) <= 2.5e-3 * 2 ** -8
    assert abs(r[5] - 3.0e5) <= 3.0e5 * 2 ** -8


def test_sample_loss_zero_when_exact():
    rng = np.random.default_rng(1)
    p = np.zeros(M.NPARAM)
    s = rng.normal(size=(10, 32))
    s[:, 28:31] = s[:, 25:28]  # target = a0, network output 0
    s[:, 31] = 2.0
    assert M.sample_loss(p, s) == 0.0
