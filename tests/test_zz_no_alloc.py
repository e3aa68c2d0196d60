"""clay_hip_mlp_gradient_noalloc (the gradient of the tail predictor with every
device buffer in the library's fixed scratch array) against
clay_hip_mlp_gradient, through the C ABI: bit-identical gradient and loss on
two calls in a row, the second with fewer samples (the array keeps the first
call's contents), and a clear error beyond the array's capacity."""

import ctypes as C

import numpy as np
import pytest

import claysampler_amd as ca
from claysampler_amd import abi

pytestmark = pytest.mark.gpu

F = C.POINTER(C.c_float)


def samples(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, abi.MLP_INPUTS)).astype(np.float32) * 3 + 1
    y = (x[:, 0] * 0.5 - x[:, 3] + rng.normal(size=n)).astype(np.float32)
    return np.ascontiguousarray(x), np.ascontiguousarray(y)


def gradient(r, fn, model, x, y):
    g = np.zeros(abi.MLP_PARAMS, np.float32)
    loss = C.c_float()
    status = fn(r.handle, C.byref(model), x.ctypes.data_as(F), y.ctypes.data_as(F), len(y), g.ctypes.data_as(F),
                C.byref(loss))
    r._check(status)
    return g, loss.value


def test_noalloc_gradient_matches_the_allocating_entry_point():
    r = ca.Renderer(taps=16)
    lib = r.lib
    lib.clay_hip_mlp_gradient_noalloc.argtypes = [C.c_void_p, C.POINTER(abi.TailModel), F, F, C.c_size_t, F, F]
    lib.clay_hip_mlp_gradient_noalloc.restype = C.c_int
    x, y = samples(5000, 1)
    model = abi.TailModel()
    assert lib.clay_mlp_init(C.byref(model), x.ctypes.data_as(F), y.ctypes.data_as(F), len(y), 7) == abi.OK
    for n, seed in ((5000, 1), (1200, 2)):  # the second call is smaller
        x, y = samples(n, seed)
        print(f"no-alloc gradient: n={n}", flush=True)
        want_g, want_l = gradient(r, lib.clay_hip_mlp_gradient, model, x, y)
        got_g, got_l = gradient(r, lib.clay_hip_mlp_gradient_noalloc, model, x, y)
        assert np.array_equal(got_g, want_g) and got_l == want_l, n
        assert np.all(np.isfinite(got_g)) and np.abs(got_g).max() > 0
    x, y = samples(65537, 3)
    g = np.zeros(abi.MLP_PARAMS, np.float32)
    status = lib.clay_hip_mlp_gradient_noalloc(r.handle, C.byref(model), x.ctypes.data_as(F), y.ctypes.data_as(F),
                                               len(y), g.ctypes.data_as(F), None)
    assert status == abi.E_INVALID
    assert b"65536" in lib.clay_hip_last_error(r.handle)
    r.close()
