"""The DPP wavefront primitives (hipr_common.hpp) against numpy: inclusive
sum and max scans over 64 lanes, and the one-lane shifts with their fill
values.  The shuffle forms are checked the same way."""
import ctypes

import numpy as np
import pytest

from skity_amd import device

pytestmark = pytest.mark.gpu


def test_dpp_scans_and_shifts():
    L = device.lib()
    L.hipr_debug_wave_ops.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_int32]
    L.hipr_debug_wave_ops.restype = ctypes.c_int
    ctx = device.Context(0)
    rng = np.random.default_rng(9)
    n = 32
    x = rng.integers(-1000, 1000, (n, 64)).astype(np.int32)
    x[1] = np.arange(64)
    x[2] = -np.arange(64)
    x[3] = np.iinfo(np.int32).min + 5
    out = np.zeros((n, 6, 64), np.int32)
    device.check(L.hipr_debug_wave_ops(ctx.handle, x.ctypes.data, out.ctypes.data, n),
                 "hipr_debug_wave_ops")
    ctx.close()
    csum = np.cumsum(x.astype(np.int64), 1).astype(np.int32)
    cmax = np.maximum.accumulate(x, 1)
    up = np.concatenate([np.full((n, 1), -7, np.int32), x[:, :-1]], 1)
    dn = np.concatenate([x[:, 1:], np.full((n, 1), -9, np.int32)], 1)
    assert np.array_equal(out[:, 1], csum), "shuffle sum scan"
    assert np.array_equal(out[:, 3], cmax), "shuffle max scan"
    assert np.array_equal(out[:, 0], csum), "DPP sum scan"
    assert np.array_equal(out[:, 2], cmax), "DPP max scan"
    assert np.array_equal(out[:, 4], up), "DPP shift up"
    assert np.array_equal(out[:, 5], dn), "DPP shift down"
