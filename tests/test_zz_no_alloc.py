"""The allocation-free StackBlur entry point against the allocating one.

hipr_stackblur_device_noalloc takes device pointers and keeps its
intermediate image in a static device buffer; its output must equal
hipr_stackblur_ex (MFMA) byte for byte, on two calls in a row with the
second smaller, so that the buffer's leftover contents cannot leak in.
"""
import ctypes

import numpy as np
import pytest

from skity_amd import abi, device

pytestmark = pytest.mark.gpu


def _noalloc(ctx, img, radius, torch):
    lib = device.lib()
    lib.hipr_stackblur_device_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p, ctypes.c_int32,
                                                  ctypes.c_int32, ctypes.c_int32]
    lib.hipr_stackblur_device_noalloc.restype = ctypes.c_int
    h, w = img.shape[:2]
    src = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    dst = torch.zeros_like(src)
    torch.cuda.synchronize()
    device.check(lib.hipr_stackblur_device_noalloc(ctx.handle, src.data_ptr(), dst.data_ptr(),
                                                   w, h, radius),
                 "hipr_stackblur_device_noalloc")
    return dst.cpu().numpy()


def test_noalloc_stackblur_equals_allocating_entry_point():
    import torch
    ctx = device.Context(0)
    rng = np.random.default_rng(21)
    for (h, w, r) in [(700, 900, 37), (61, 83, 5), (128, 40, 1)]:
        img = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
        img[..., :3] = np.minimum(img[..., :3], img[..., 3:4])
        want = ctx.stackblur(img, r, impl=abi.BLUR_IMPL_MFMA)
        got = _noalloc(ctx, img, r, torch)
        print(f"{h}x{w} r={r}: {int((got != want).any(-1).sum())} pixels differ", flush=True)
        assert np.array_equal(got, want)
    # larger than the static buffer: refused
    lib = device.lib()
    big = torch.zeros((2049, 2048, 4), dtype=torch.uint8, device="cuda")
    rc = lib.hipr_stackblur_device_noalloc(ctx.handle, big.data_ptr(), big.data_ptr(), 2048,
                                           2049, 5)
    assert rc == abi.ERROR_CAPACITY
    ctx.close()
