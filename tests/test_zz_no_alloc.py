"""The allocation-free SSIM batch (iqm_ssim_batch_dev_noalloc) against the
allocating one it varies (iqm_ssim_batch_dev): two calls in a row, the second
with a smaller input, each bit-equal in every MSSIM and map value (the same
kernels and the same fixed-order reduction run; only where the workgroup
partials live differs).  Batches beyond the static scratch are refused before
anything is launched."""
import ctypes

import numpy as np
import pytest

import iqm_amd

pytestmark = pytest.mark.gpu


def _pairs(ctx, w, h, ch, n, seed):
    stride = w * h * ch
    d_img = ctx.alloc(stride * n)
    d_ref = ctx.alloc(stride * n)
    if ch == 3:
        for i in range(n):
            ctx.synth_pair(w, h, seed + i, d_img + i * stride, d_ref + i * stride)
    else:
        rng = np.random.default_rng(seed)
        ctx.upload(d_img, rng.integers(0, 256, (n, h, w, ch), dtype=np.uint8))
        ctx.upload(d_ref, rng.integers(0, 256, (n, h, w, ch), dtype=np.uint8))
    return d_img, d_ref, stride


@pytest.mark.parametrize("calls", [[(1920, 1080, 3, 3), (640, 360, 3, 2)],
                                   [(512, 512, 1, 4), (101, 77, 4, 1)]], ids=["rgb", "gray_rgba"])
def test_noalloc_matches_batch_dev(ctx, calls):
    ssim = iqm_amd.SSIM()
    d_ms = ctx.alloc(8 * 16)
    for (w, h, ch, n) in calls:
        print(f"[noalloc] {w}x{h}x{ch}, {n} pairs", flush=True)
        d_img, d_ref, stride = _pairs(ctx, w, h, ch, n, seed=77 + w)
        d_map_a = ctx.alloc(4 * w * h * n)
        d_map_b = ctx.alloc(4 * w * h * n)
        want = ssim.compute_batch_dev(ctx, d_img, d_ref, w, h, ch, stride, n, d_map_a)
        ssim.compute_batch_dev_noalloc(ctx, d_img, d_ref, w, h, ch, stride, n, d_ms, d_map_b)
        ctx.sync()
        got = ctx.download(d_ms, np.empty(n, np.float64))
        ma = ctx.download(d_map_a, np.empty(w * h * n, np.float32))
        mb = ctx.download(d_map_b, np.empty(w * h * n, np.float32))
        assert np.all(np.isfinite(got)) and np.all(got <= 1.0)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(mb, ma)
        for p in (d_img, d_ref, d_map_a, d_map_b):
            ctx.free(p)
    ctx.free(d_ms)


def test_noalloc_rejects_oversized_batch(ctx):
    # 1920 x 1080 takes 8 x 17 = 136 workgroups per pair; 1000 pairs need
    # 136,000 partials, more than the static scratch's 131,072.  The check is
    # on the host, before any launch; the pointers are never dereferenced.
    L = iqm_amd.lib()
    p = iqm_amd.SsimParams(11, 0.01, 0.03, 1.5)
    d = ctx.alloc(64)
    rc = L.iqm_ssim_batch_dev_noalloc(ctx.handle, d, d, 1920, 1080, 3, 1920 * 1080 * 3, 1000,
                                      ctypes.byref(p), None, d)
    assert rc == iqm_amd.IQM_ERR_INVALID
    assert b"at most 131072" in L.iqm_last_error()
    rc = L.iqm_ssim_batch_dev_noalloc(ctx.handle, d, d, 1920, 1080, 3, 1920 * 1080 * 3, 0,
                                      ctypes.byref(p), None, d)
    assert rc == iqm_amd.IQM_ERR_INVALID  # n = 0, also refused on the host
    ctx.free(d)
