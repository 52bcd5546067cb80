"""The allocation-free MLP forward (phieng_surrogate3_mlp_forward_noalloc,
no reference counterpart) through the C ABI: two calls in a row, the second
smaller, each equal bit for bit to phieng_surrogate3_mlp_forward on the same
rows; a call above the scratch maximum is refused with an error."""
import ctypes

import numpy as np
import pytest

from phi_eng_amd import ffi
from phi_eng_amd.world import Surrogate3

pytestmark = pytest.mark.gpu

MAX_ROWS = 65536


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def test_noalloc_forward_matches_forward(gpu):
    l = ffi.lib()
    net = Surrogate3(seed=9)
    rng = np.random.default_rng(17)
    for rows in (MAX_ROWS, 3001):
        x = rng.normal(0.0, 1.5, size=(rows, 16)).astype(np.float32)
        y_ref = np.zeros((rows, 16), np.float32)
        assert l.phieng_surrogate3_mlp_forward(net.ptr, _fp(x), rows, _fp(y_ref)) == 1
        y = np.full((rows, 16), np.nan, np.float32)
        assert l.phieng_surrogate3_mlp_forward_noalloc(net.ptr, _fp(x), rows, _fp(y)) == 1, \
            l.phieng_surrogate3_last_error().decode()
        print(f"rows {rows}: max |y| {np.abs(y_ref).max():.3e}", flush=True)
        assert np.array_equal(y.view(np.uint32), y_ref.view(np.uint32)), rows
    big = np.zeros((MAX_ROWS + 1, 16), np.float32)
    out = np.zeros_like(big)
    assert l.phieng_surrogate3_mlp_forward_noalloc(net.ptr, _fp(big), MAX_ROWS + 1, _fp(out)) == 0
    assert "exceeds the scratch maximum" in l.phieng_surrogate3_last_error().decode()
    net.close()
