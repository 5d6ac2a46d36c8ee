"""The device generator of the benchmark inputs against the host recipe
(iqm_amd/synth.py).  On a 2/3/5/7-smooth size the device reference uses the
host's samples and differs only where the FP32 FFT moves a value across a
rounding boundary (<= 1 level on < 1% of the bytes); the device distortion of
a given reference differs from the host's only across quantisation
boundaries (<= one 17-level step on < 1% of the bytes)."""
import numpy as np
import pytest

import iqm_amd
from iqm_amd import synth

pytestmark = pytest.mark.gpu


def _pair(ctx, w, h, seed):
    di, dr = ctx.alloc(w * h * 3), ctx.alloc(w * h * 3)
    ctx.synth_pair(w, h, seed, di, dr)
    img = ctx.download(di, np.empty((h, w, 3), np.uint8))
    ref = ctx.download(dr, np.empty((h, w, 3), np.uint8))
    ctx.free(di)
    ctx.free(dr)
    return img, ref


def test_reference_and_distortion_match_host(ctx):
    w, h, seed = 240, 180, 17
    print(f"[synth] {w}x{h} seed={seed}", flush=True)
    img, ref = _pair(ctx, w, h, seed)
    host_ref = synth.rgb_reference(w, h, seed)
    d = np.abs(ref.astype(int) - host_ref.astype(int))
    print(f"[synth] reference: max={d.max()} differing={np.mean(d > 0):.2e}", flush=True)
    assert d.max() <= 1 and np.mean(d > 0) < 1e-2
    host_img, _ = synth.distort(ref, seed + 1000003)
    d = np.abs(img.astype(int) - host_img.astype(int))
    print(f"[synth] distorted: max={d.max()} differing={np.mean(d > 0):.2e}", flush=True)
    assert d.max() <= 17 and np.mean(d > 0) < 1e-2


def test_non_smooth_size_and_strided_distort(ctx):
    w, h = 203, 151
    img, ref = _pair(ctx, w, h, 5)
    assert ref.min() == 0 and ref.max() == 255
    img2, ref2 = _pair(ctx, w, h, 5)
    assert np.array_equal(img, img2) and np.array_equal(ref, ref2)
    # distortion of a crop read through a row stride equals that of the packed crop
    cw, ch_, x0, y0 = 120, 90, 31, 17
    dsrc = ctx.alloc(ref.nbytes)
    ctx.upload(dsrc, ref)
    dout, dpk, dout2 = ctx.alloc(cw * ch_ * 3), ctx.alloc(cw * ch_ * 3), ctx.alloc(cw * ch_ * 3)
    off = (y0 * w + x0) * 3
    ctx.synth_distort(dsrc + off, w * 3, cw, ch_, 99, dout)
    ctx.copy2d(dpk, cw * 3, dsrc + off, w * 3, cw * 3, ch_)
    ctx.synth_distort(dpk, cw * 3, cw, ch_, 99, dout2)
    a = ctx.download(dout, np.empty((ch_, cw, 3), np.uint8))
    b = ctx.download(dout2, np.empty((ch_, cw, 3), np.uint8))
    crop = ctx.download(dpk, np.empty((ch_, cw, 3), np.uint8))
    for p in (dsrc, dout, dpk, dout2):
        ctx.free(p)
    assert np.array_equal(crop, ref[y0:y0 + ch_, x0:x0 + cw])
    assert np.array_equal(a, b)
