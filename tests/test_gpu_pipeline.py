"""The host-fed input pipeline (iqm_fsim_batch_host: pinned ring, H2D on a
copy stream overlapped with FSIM; the reference stages each call instead,
src/gpu/fsim/fsim.cpp:170-252) and the stage marks (src/timestamps.h,
src/gpu/fsim/fsim.cpp:85-155, src/gpu/ssim/ssim.cpp:78-235)."""
import numpy as np
import pytest

import iqm_amd
from iqm_amd import synth

pytestmark = pytest.mark.gpu


def _frames(ctx, w, h, n, seed):
    pb = w * h * 3
    di, dr = ctx.alloc(n * pb), ctx.alloc(n * pb)
    for i in range(n):
        ctx.synth_pair(w, h, seed + i, di + i * pb, dr + i * pb)
    img = ctx.download(di, np.empty((n, h, w, 3), np.uint8))
    ref = ctx.download(dr, np.empty((n, h, w, 3), np.uint8))
    return di, dr, img, ref


@pytest.mark.parametrize("w,h,n,chunk", [(640, 360, 5, 2), (1920, 1080, 7, 3), (333, 211, 1, 4)])
def test_host_fed_equals_device_batch(ctx, w, h, n, chunk):
    print(f"[host-fed] {w}x{h} n={n} chunk={chunk}", flush=True)
    fsim = iqm_amd.FSIM()
    di, dr, img, ref = _frames(ctx, w, h, n, 8100 + w)
    want_fs, want_fc = fsim.compute_batch_dev(ctx, di, dr, w, h, 3, w * h * 3, n)
    # pageable numpy arrays: staged through the library's pinned slots
    fs, fc = fsim.compute_batch_host(ctx, img, ref, chunk=chunk)
    assert np.array_equal(fs, want_fs) and np.array_equal(fc, want_fc)
    # pinned arrays: copied by DMA straight from them; twice, so the second
    # call reuses the slots (and graphs) of the first
    pi = ctx.host_alloc(img.nbytes).reshape(img.shape)
    pr = ctx.host_alloc(ref.nbytes).reshape(ref.shape)
    pi[...] = img
    pr[...] = ref
    for _ in range(2):
        fs2, fc2 = fsim.compute_batch_host(ctx, pi, pr, chunk=chunk)
        assert np.array_equal(fs2, want_fs) and np.array_equal(fc2, want_fc)
    ctx.host_free(pi)
    ctx.host_free(pr)
    ctx.free(di)
    ctx.free(dr)


def test_stage_marks(ctx):
    img, ref = synth.rgb_pair(320, 240, seed=35)
    ctx.set_marks(True)
    try:
        r = iqm_amd.FSIM().compute_metric(ctx, img, ref)
        m = ctx.marks()
        assert [n for n, _ in m] == ["downscale factor computed", "images sent to gpu", "FFT library initialized",
                                     "pre median work recorded", "noise powers computed", "FSIM, FSIMc computed"]
        assert all(b[1] >= a[1] for a, b in zip(m, m[1:]))
        s = iqm_amd.SSIM().compute_metric(ctx, img, ref)
        m = ctx.marks()
        assert [n for n, _ in m] == ["start GPU pipeline", "end GPU pipeline", "end copy from GPU",
                                     "end MSSIM compute"]
        assert all(b[1] >= a[1] for a, b in zip(m, m[1:]))
    finally:
        ctx.set_marks(False)
    # marks off: same scores, no marks
    r2 = iqm_amd.FSIM().compute_metric(ctx, img, ref)
    s2 = iqm_amd.SSIM().compute_metric(ctx, img, ref)
    assert (r2.fsim, r2.fsimc, s2.mssim) == (r.fsim, r.fsimc, s.mssim)
    assert ctx.marks() == []
