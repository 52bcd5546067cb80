"""CPU oracle of the surrogate MLP (16 -> 128 tanh-GELU -> 16) and its
gradient — TEST INFRASTRUCTURE ONLY (the checker of
phi_eng_amd/csrc/mlp.h; no reference counterpart exists for this part).

FP64 arithmetic.  With ``bf16=True`` the values the GPU feeds to its BF16
matrix cores (inputs, weights, hidden activations, output and hidden
gradients) are rounded to bfloat16 (round to nearest even) at the same
points, so the GPU differs only by its FP32 accumulation.  With
``bf16=False`` it is the plain real-valued network; its gradient is pinned
by central finite differences (tests/test_surrogate_oracle.py).

Parameter vector layout: W1 [16][128], b1 [128], W2 [128][16], b2 [16].
"""
import numpy as np

IN, HID, OUT = 16, 128, 16
OFF_W1 = 0
OFF_B1 = IN * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * OUT
NPARAM = OFF_B2 + OUT
K0, K1 = 0.7978845608028654, 0.044715


def bf16(a):
    """Round float64/float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(IN, HID), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, OUT),
            p[OFF_B2:NPARAM])


def gelu(z):
    th = np.tanh(K0 * (z + K1 * z ** 3))
    dg = 0.5 * (1 + th) + 0.5 * z * (1 - th * th) * K0 * (1 + 3 * K1 * z * z)
    return 0.5 * z * (1 + th), dg


def forward(p, x, bf=True):
    w1, b1, w2, b2 = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    z = r(x) @ r(w1) + b1
    h, _ = gelu(z)
    return r(h) @ r(w2) + b2


def backward(p, x, dy, bf=True):
    """Gradient of sum(dy * forward(p, x)) with respect to p."""
    w1, b1, w2, b2 = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    xb, dyb = r(x), r(dy)
    z = xb @ r(w1) + b1
    h, dg = gelu(z)
    dh = dyb @ r(w2).T
    dz = dh * dg
    g = np.zeros(NPARAM)
    g[OFF_W1:OFF_B1] = (xb.T @ r(dz)).ravel()
    g[OFF_B1:OFF_W2] = dz.sum(axis=0)
    g[OFF_W2:OFF_B2] = (r(h).T @ dyb).ravel()
    g[OFF_B2:NPARAM] = dyb.sum(axis=0)
    return g


def sample_loss(p, samples, ncoef=3, bf=True):
    """Mean over samples of |a0 + sum_k c_k b_k - a_t|^2 / s^2 (layout of
    phi_eng_amd/csrc/surrogate.h: x[0:16], b[16:25], a0[25:28], at[28:31], s[31])."""
    s = np.asarray(samples, dtype=np.float64)
    c = forward(p, s[:, 0:16], bf)[:, :ncoef]
    b = s[:, 16:16 + 3 * ncoef].reshape(-1, ncoef, 3)
    pred = s[:, 25:28] + np.einsum("nk,nka->na", c, b)
    d = pred - s[:, 28:31]
    return float(np.mean((d * d).sum(axis=1) / s[:, 31] ** 2))
