"""rc_post_image_noalloc (the no-allocation variant of rc_post_image) through
the C ABI: on two calls in a row, the second with a smaller image, its HDR
output equals rc_post_image's bit for bit and its RGBA8 display is identical;
images beyond its scratch are refused with RC_ERR_LIMIT. Device buffers come
from torch (the test's own allocations)."""
import ctypes as C

import numpy as np
import pytest
import torch

from retrocam_amd import _native as N
from retrocam_amd.post import VHS_PASSES, PostProcessStack, TemporalInputs, make_params

pytestmark = pytest.mark.gpu


def _image(seed, w, h):
    rng = np.random.default_rng(seed)
    a = np.ones((h * w, 4), np.float32)
    a[:, :3] = rng.gamma(1.0, 0.6, (h * w, 3))
    a.reshape(h, w, 4)[h // 3:h // 2, :, :3] = 1.4  # a bright band: smear and edges
    return a


def _noalloc(ctx, img, w, h, p, t):
    lib = N.hip_lib()
    src = torch.from_numpy(img).to("cuda:0")
    hdr = torch.full((h * w, 4), -7.0, dtype=torch.float32, device="cuda:0")
    rgba8 = torch.zeros((h * w, 4), dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    tc = t.to_c()
    rc = lib.rc_post_image_noalloc(ctx.h, C.c_void_p(src.data_ptr()), w, h, C.byref(p), C.byref(tc),
                                   C.c_void_p(hdr.data_ptr()), C.c_void_p(rgba8.data_ptr()))
    ctx.check(rc)
    ctx.sync()
    return hdr.cpu().numpy(), rgba8.cpu().numpy()


def test_noalloc_equals_post_image_on_two_calls(gpu_ctx):
    ps = PostProcessStack(gpu_ctx)
    p = make_params(enable=VHS_PASSES, trackingSeverity=0.8, dropoutBurstRate=1.0, tapeAge=0.9)
    for seed, (w, h), f in ((3, (640, 360), 31), (4, (200, 96), 902)):
        print(f"no-alloc post chain {w}x{h} frame {f}", flush=True)
        img = _image(seed, w, h)
        t = TemporalInputs(time=f / 29.97, agcGain=1.7, frameIndex=f)
        ref = ps.run_on_image(img, w, h, p, t)
        hdr, rgba8 = _noalloc(gpu_ctx, img, w, h, p, t)
        assert np.array_equal(hdr.view(np.uint32), ref.hdr.view(np.uint32))
        assert np.array_equal(rgba8, ref.rgba8)


def test_noalloc_rejects_images_beyond_its_scratch(gpu_ctx):
    lib = N.hip_lib()
    p = make_params(enable=VHS_PASSES)
    t = TemporalInputs().to_c()
    buf = torch.zeros(16, dtype=torch.float32, device="cuda:0")
    ptr = C.c_void_p(buf.data_ptr())
    rc = lib.rc_post_image_noalloc(gpu_ctx.h, ptr, 1281, 720, C.byref(p), C.byref(t), ptr, ptr)
    assert rc == N.RC_ERR_LIMIT
    assert b"scratch" in lib.rc_last_error(gpu_ctx.h)
    rc = lib.rc_post_image_noalloc(gpu_ctx.h, None, 64, 64, C.byref(p), C.byref(t), ptr, ptr)
    assert rc == N.RC_ERR_INVALID_ARG
