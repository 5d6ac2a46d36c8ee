"""GPU parity of the HIP path against the CPU oracle, through the C-ABI.

Tolerances (FP32 path): scalars within 1e-5 absolute, maps within 1e-4
max-abs, against both the FP32 mirror and the FP64 oracle; the errors
against FP64 are printed for every case.
"""
import os

import numpy as np
import pytest

import iqm_amd
from iqm_amd import synth
from oracle import oracle

pytestmark = pytest.mark.gpu

SCALAR_TOL = 1e-5
MAP_TOL = 1e-4


def _maxabs(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)))


def _report(tag, **errs):
    print(f"[{tag}] " + " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


# line plans: two-level splits (480, 270, 64, 48, 45, 36; 221 = 13*17 and 527 = 17*31 by the prime
# register DFTs), Rader (97, 61, 17, 23, 31, 929 and 74 = 2*37, 202 = 2*101, 611 = 13*47,
# 629 = 17*37, 1184 = 32*37), Bluestein (1009, 659), trivial (1, 3, 7, 8)
@pytest.mark.parametrize("w,h", [(480, 270), (270, 480), (64, 48), (45, 36), (1, 8), (7, 1), (97, 61), (1009, 3), (659, 17),
                                 (221, 527), (611, 202), (629, 1184), (929, 74), (23, 31)])
def test_fft2d_matches_numpy(ctx, w, h):
    rng = np.random.default_rng(w * 1000 + h)
    x = (rng.standard_normal((3, h, w)) + 1j * rng.standard_normal((3, h, w))).astype(np.complex64)
    for inverse in (False, True):
        y = ctx.fft2d(x, inverse=inverse, normalize=inverse)
        ref = np.fft.ifft2(x.astype(np.complex128)) if inverse else np.fft.fft2(x.astype(np.complex128))
        rel = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
        _report(f"fft {w}x{h} inv={inverse}", rel=rel)
        assert rel < 2e-6


def test_ssim_gray_512(ctx):
    img, ref = synth.ssim_gray_pair(512, 512, seed=1)
    r = iqm_amd.SSIM().compute_metric(ctx, img, ref)
    m64, map64 = oracle.ssim(img, ref, prec=64)
    m32, map32 = oracle.ssim(img, ref, prec=32)
    e = dict(d_mssim64=abs(r.mssim - m64), d_mssim32=abs(r.mssim - m32),
             d_map64=_maxabs(r.image_data, map64), d_map32=_maxabs(r.image_data, map32))
    _report("ssim 512 gray", **e)
    assert e["d_mssim64"] <= SCALAR_TOL and e["d_mssim32"] <= SCALAR_TOL
    assert e["d_map64"] <= MAP_TOL and e["d_map32"] <= MAP_TOL


@pytest.mark.parametrize("w,h,ch", [(1920, 1080, 3), (333, 211, 4), (100, 37, 1), (11, 11, 3)])
def test_ssim_shapes(ctx, w, h, ch):
    img, ref = synth.rgb_pair(w, h, seed=w + h)
    if ch == 1:
        img, ref = img[:, :, 1].copy(), ref[:, :, 1].copy()
    elif ch == 4:
        img, ref = synth.rgba(img), synth.rgba(ref)
    r = iqm_amd.SSIM().compute_metric(ctx, img, ref)
    m64, map64 = oracle.ssim(img, ref, prec=64)
    e = dict(d_mssim64=abs(r.mssim - m64), d_map64=_maxabs(r.image_data, map64))
    _report(f"ssim {w}x{h}x{ch}", **e)
    assert e["d_mssim64"] <= SCALAR_TOL and e["d_map64"] <= MAP_TOL


def test_ssim_identity_and_params(ctx):
    img, ref = synth.rgb_pair(200, 150, seed=3)
    r = iqm_amd.SSIM().compute_metric(ctx, ref, ref)
    assert abs(r.mssim - 1.0) < 1e-6 and np.max(np.abs(r.image_data - 1.0)) < 1e-5
    s = iqm_amd.SSIM()
    s.kernel_size, s.sigma, s.k_1, s.k_2 = 7, 1.0, 0.02, 0.05
    r = s.compute_metric(ctx, img, ref)
    m64, map64 = oracle.ssim(img, ref, prec=64, kernel_size=7, k1=0.02, k2=0.05, sigma=1.0)
    assert abs(r.mssim - m64) <= SCALAR_TOL and _maxabs(r.image_data, map64) <= MAP_TOL


def _fsim_case(ctx, img, ref, tag, stages=False):
    r = iqm_amd.FSIM().compute_metric(ctx, img, ref, maps=iqm_amd.FSIM.MAPS, detail=stages)
    names = ("simc", "sim", "pcm", "pc_in", "pc_ref")
    o64 = oracle.fsim(img, ref, prec=64, maps=names + (("yiq_in", "yiq_ref", "grad_in", "grad_ref",
                                                         "an_e", "keys") if stages else ()))
    o32 = oracle.fsim(img, ref, prec=32, maps=("simc",))
    e = dict(d_fsim64=abs(r.fsim - o64.fsim), d_fsimc64=abs(r.fsimc - o64.fsimc),
             d_fsim32=abs(r.fsim - o32.fsim), d_fsimc32=abs(r.fsimc - o32.fsimc),
             d_simc64=_maxabs(r.maps["simc"], o64.maps["simc"]),
             d_simc32=_maxabs(r.maps["simc"], o32.maps["simc"]))
    for n in ("sim", "pcm", "pc_in", "pc_ref"):
        e[f"d_{n}64"] = _maxabs(r.maps[n], o64.maps[n])
    _report(tag, **e)
    assert e["d_fsim64"] <= SCALAR_TOL and e["d_fsimc64"] <= SCALAR_TOL
    assert e["d_fsim32"] <= SCALAR_TOL and e["d_fsimc32"] <= SCALAR_TOL
    for k, v in e.items():
        if k.startswith("d_") and not k.startswith("d_fsim"):
            assert v <= MAP_TOL, (k, v)
    return r, o64


def test_fsim_stages_small(ctx):
    img, ref = synth.rgb_pair(320, 240, seed=11)
    r, o = _fsim_case(ctx, img, ref, "fsim 320x240 stages", stages=True)
    d = r.detail
    np.testing.assert_allclose(d["em_n"], o.em_n, rtol=1e-6)
    np.testing.assert_allclose(d["est"], o.est, rtol=1e-5)
    np.testing.assert_allclose(d["median"], o.median, rtol=1e-4)
    np.testing.assert_allclose(d["T"], o.T, rtol=1e-4)
    np.testing.assert_allclose(d["sums"], o.sums, rtol=1e-5)
    for n in ("yiq_in", "yiq_ref", "grad_in", "grad_ref"):
        scale = max(1.0, float(np.max(np.abs(o.maps[n]))))
        assert _maxabs(d[n], o.maps[n]) / scale <= 1e-5, n
    # the oracle keeps (An, E) per [img*4 + o]; the library stores per image
    # sum_o An, then E_0..E_3
    ae = o.maps["an_e"].reshape(2, 4, 2, *o.maps["an_e"].shape[1:])
    expect = {"an_e": np.concatenate([np.concatenate([ae[i, :, 0].sum(axis=0)[None], ae[i, :, 1]])
                                      for i in range(2)]),
              "keys": o.maps["keys"]}
    for n in ("an_e", "keys"):
        scale = float(np.max(np.abs(expect[n])))
        err = _maxabs(d[n], expect[n]) / scale
        _report(f"stage {n}", rel=err)
        assert err <= 1e-4, n


def test_fsim_1080p(ctx):
    img, ref = synth.rgb_pair(1920, 1080, seed=2)
    _fsim_case(ctx, img, ref, "fsim 1920x1080 rgb")


@pytest.mark.parametrize("w,h,ch", [(512, 512, 1), (640, 480, 4), (999, 601, 3)])
def test_fsim_shapes(ctx, w, h, ch):
    img, ref = synth.rgb_pair(w, h, seed=w)
    if ch == 1:
        img, ref = img[:, :, 0].copy(), ref[:, :, 0].copy()
    elif ch == 4:
        img, ref = synth.rgba(img), synth.rgba(ref)
    _fsim_case(ctx, img, ref, f"fsim {w}x{h}x{ch}")


@pytest.mark.parametrize("w,h,ch", [(3840, 2160, 3), (2560, 1536, 3), (3136, 1792, 4), (1280, 1280, 4)])
def test_fsim_wide_boxes(ctx, w, h, ch):
    """Downscale factors 8, 6, 7 and 5 take the register column-sum prep
    kernel (rows 16-byte aligned); its planes must equal the LDS-staged
    kernel's bit for bit (same integer box sums), and the result must match
    the oracle."""
    img, ref = synth.rgb_pair(w, h, seed=h)
    if ch == 4:
        img, ref = synth.rgba(img), synth.rgba(ref)
    r, _ = _fsim_case(ctx, img, ref, f"fsim {w}x{h}x{ch}", stages=True)
    os.environ["IQM_FSIM_PREP2"] = "0"
    try:
        r0 = iqm_amd.FSIM().compute_metric(ctx, img, ref, maps=iqm_amd.FSIM.MAPS, detail=True)
    finally:
        del os.environ["IQM_FSIM_PREP2"]
    for name in ("yiq_in", "yiq_ref"):
        assert np.array_equal(r.detail[name], r0.detail[name]), name
    assert r.fsim == r0.fsim and r.fsimc == r0.fsimc


def test_fsim_1080p_prep_kernels_agree(ctx):
    """At 1920x1080 (F = 4) the register column-sum prep and the LDS-staged
    prep give the same planes bit for bit."""
    img, ref = synth.rgb_pair(1920, 1080, seed=2)
    r = iqm_amd.FSIM().compute_metric(ctx, img, ref, detail=True)
    os.environ["IQM_FSIM_PREP2"] = "0"
    try:
        r0 = iqm_amd.FSIM().compute_metric(ctx, img, ref, detail=True)
    finally:
        del os.environ["IQM_FSIM_PREP2"]
    for name in ("yiq_in", "yiq_ref"):
        assert np.array_equal(r.detail[name], r0.detail[name]), name
    assert r.fsim == r0.fsim and r.fsimc == r0.fsimc


def test_fsim_select_crowded_bucket(ctx):
    """A vertical grating with the scale-0 period (6 px) on a width holding
    whole periods gives orientation 0 the same response at every pixel, so
    nearly every median key of that orientation falls in one level-1 bucket:
    more keys than the select keeps in LDS, and its last level reads the keys
    from memory again.  The medians (where they are not rounding noise) must
    still be the oracle's, and so must the scores and maps."""
    w, h = 258, 256
    rng = np.random.default_rng(7)
    row = np.round(128 + 60 * np.sin(2 * np.pi * np.arange(w) / 6.0)).astype(np.uint8)
    img = np.repeat(row[None, :], h, axis=0)
    ref = np.clip(img.astype(np.int32) + rng.integers(-3, 4, img.shape), 0, 255).astype(np.uint8)
    r, o = _fsim_case(ctx, img, ref, "fsim grating 258x256", stages=True)
    keys = np.asarray(r.detail["keys"], np.float32).reshape(8, -1)
    med = np.asarray(r.detail["median"], np.float64)
    crowded = 0
    for s in range(8):
        b = (np.float32(med[s]).view(np.uint32) >> 21) & 1023
        crowded = max(crowded, int(np.sum(((keys[s].view(np.uint32) >> 21) & 1023) == b)))
    _report("grating crowded bucket", keys=float(crowded))
    assert crowded > 16384, crowded
    big = o.median > 1e-6 * np.max(o.median)
    np.testing.assert_allclose(med[big], o.median[big], rtol=1e-4)


def test_fsim_flat_images(ctx):
    """Flat 1920x1080 images: the box downscale's zero padding darkens the
    border pixels (F = 4), so the only structure is the image frame.  The
    scores must be the oracle's to the scalar tolerance.  Far from the frame
    sum_o An is small and PC is a ratio of small numbers, so the maps there
    are ill-conditioned in FP32: the oracle's own FP32 mirror departs from the
    FP64 maps by more than the usual map tolerance, and the bar for the GPU
    maps is that distance (with 50% margin).  (At F = 1 a flat pair has no
    structure at all: every response is FFT rounding noise and FSIM is 0/0 in
    exact arithmetic, so that case has no reference value to compare with.)"""
    img = np.full((1080, 1920, 3), 100, np.uint8)
    ref = np.full((1080, 1920, 3), 120, np.uint8)
    names = ("simc", "sim", "pcm", "pc_in", "pc_ref")
    r = iqm_amd.FSIM().compute_metric(ctx, img, ref, maps=iqm_amd.FSIM.MAPS)
    o64 = oracle.fsim(img, ref, prec=64, maps=names)
    o32 = oracle.fsim(img, ref, prec=32, maps=names)
    e = dict(d_fsim64=abs(r.fsim - o64.fsim), d_fsimc64=abs(r.fsimc - o64.fsimc))
    assert e["d_fsim64"] <= SCALAR_TOL and e["d_fsimc64"] <= SCALAR_TOL
    for n in names:
        gpu = _maxabs(r.maps[n], o64.maps[n])
        mirror = _maxabs(o32.maps[n], o64.maps[n])
        e[f"d_{n}64"], e[f"mirror_{n}64"] = gpu, mirror
        assert gpu <= max(MAP_TOL, 1.5 * mirror), (n, gpu, mirror)
    _report("fsim flat 1920x1080", **e)


def test_fsim_identity(ctx):
    img, ref = synth.rgb_pair(400, 300, seed=5)
    r = iqm_amd.FSIM().compute_metric(ctx, ref, ref)
    assert abs(r.fsim - 1.0) < 1e-6 and abs(r.fsimc - 1.0) < 1e-6


def test_fsim_batch_dev_matches_single(ctx):
    w, h, n = 640, 360, 3
    pairs = [synth.rgb_pair(w, h, seed=100 + i) for i in range(n)]
    imgs = np.stack([p[0] for p in pairs])
    refs = np.stack([p[1] for p in pairs])
    di, dr = ctx.alloc(imgs.nbytes), ctx.alloc(refs.nbytes)
    try:
        ctx.upload(di, imgs)
        ctx.upload(dr, refs)
        fs, fc = iqm_amd.FSIM().compute_batch_dev(ctx, di, dr, w, h, 3, w * h * 3, n)
        ss = iqm_amd.SSIM().compute_batch_dev(ctx, di, dr, w, h, 3, w * h * 3, n)
    finally:
        ctx.free(di)
        ctx.free(dr)
    for i in range(n):
        r = iqm_amd.FSIM().compute_metric(ctx, pairs[i][0], pairs[i][1])
        s = iqm_amd.SSIM().compute_metric(ctx, pairs[i][0], pairs[i][1], want_map=False)
        assert fs[i] == r.fsim and fc[i] == r.fsimc
        assert ss[i] == s.mssim


def test_fsim_batch_dev_async_matches_blocking(ctx):
    """The enqueue-only batch call leaves the blocking call's scores in its
    device array; three calls in flight at once (different batches and
    sizes, one with maps) each keep their own results."""
    f = iqm_amd.FSIM()
    cases = [(640, 360, 3), (512, 512, 2), (640, 360, 1)]
    bufs, want = [], []
    try:
        for k, (w, h, n) in enumerate(cases):
            pairs = [synth.rgb_pair(w, h, seed=300 + 10 * k + i) for i in range(n)]
            imgs = np.stack([p[0] for p in pairs])
            refs = np.stack([p[1] for p in pairs])
            di, dr, ds = ctx.alloc(imgs.nbytes), ctx.alloc(refs.nbytes), ctx.alloc(n * 16)
            bufs += [di, dr, ds]
            ctx.upload(di, imgs)
            ctx.upload(dr, refs)
            want.append(f.compute_batch_dev(ctx, di, dr, w, h, 3, w * h * 3, n))
        _, wd, hd = iqm_amd.fsim_dims(*cases[0][:2])
        dm = ctx.alloc(cases[0][2] * wd * hd * 4)
        bufs.append(dm)
        for k, (w, h, n) in enumerate(cases):
            di, dr, ds = bufs[3 * k: 3 * k + 3]
            f.compute_batch_dev_async(ctx, di, dr, w, h, 3, w * h * 3, n, ds, dm if k == 0 else None)
        ctx.sync()
        for k, (w, h, n) in enumerate(cases):
            got = ctx.download(bufs[3 * k + 2], np.empty((n, 2), np.float64))
            assert np.array_equal(got[:, 0], want[k][0]), k
            assert np.array_equal(got[:, 1], want[k][1]), k
        with pytest.raises(iqm_amd.IqmError):
            f.compute_batch_dev_async(ctx, bufs[0], bufs[1], 640, 360, 3, 640 * 360 * 3, 1, 0)
    finally:
        for b in bufs:
            ctx.free(b)


def test_errors(ctx):
    a = np.zeros((10, 12, 3), np.uint8)
    b = np.zeros((10, 13, 3), np.uint8)
    with pytest.raises(iqm_amd.IqmError, match="Compared images must have the same size"):
        iqm_amd.SSIM().compute_metric(ctx, a, b)
    with pytest.raises(iqm_amd.IqmError, match="Compared images must have the same size"):
        iqm_amd.FSIM().compute_metric(ctx, a, b)
    tiny = np.zeros((6, 6, 3), np.uint8)
    with pytest.raises(iqm_amd.IqmError) as ei:
        iqm_amd.SSIM().compute_metric(ctx, tiny, tiny)
    assert ei.value.code == iqm_amd.IQM_ERR_TOO_SMALL


def test_graph_replay_matches_eager(ctx):
    """A repeated batch call replays a captured hipGraph; every replay, with
    eager work and synchronisations in between, equals the eager result."""
    w, h, n = 960, 540, 4
    _, wd, hd = iqm_amd.fsim_dims(w, h)  # F = 2 here: 480 x 270
    pairs = [synth.rgb_pair(w, h, seed=200 + i) for i in range(2)]
    imgs = np.stack([pairs[i % 2][0] for i in range(n)])
    refs = np.stack([pairs[i % 2][1] for i in range(n)])
    maps = np.empty((n, hd, wd), np.float32)
    di, dr, dm = ctx.alloc(imgs.nbytes), ctx.alloc(refs.nbytes), ctx.alloc(maps.nbytes)
    try:
        ctx.upload(di, imgs)
        ctx.upload(dr, refs)
        f = iqm_amd.FSIM()
        ctx.set_graphs(False)
        fs0, fc0 = f.compute_batch_dev(ctx, di, dr, w, h, 3, w * h * 3, n, dm)
        m0 = ctx.download(dm, maps).copy()
        ctx.set_graphs(True)
        for rep in range(4):
            fs, fc = f.compute_batch_dev(ctx, di, dr, w, h, 3, w * h * 3, n, dm)
            m = ctx.download(dm, maps)
            assert np.array_equal(fs, fs0) and np.array_equal(fc, fc0), rep
            assert np.array_equal(m, m0), rep
            iqm_amd.SSIM().compute_metric(ctx, pairs[0][0], pairs[0][1], want_map=False)
            ctx.sync()
    finally:
        ctx.set_graphs(True)
        ctx.free(di)
        ctx.free(dr)
        ctx.free(dm)


def test_autotune_and_plan_file(ctx):
    """The autotuner picks tiles for a size, the plan file round-trips, and
    the tuned launch sequence gives the same scores."""
    import json

    img, ref = synth.rgb_pair(1920, 1080, seed=77)
    before = iqm_amd.FSIM().compute_metric(ctx, img, ref)
    us = ctx.autotune_fsim(1920, 1080, 3, 2)
    assert us > 0
    plan = json.loads(ctx.export_tunes())
    entry = [e for e in plan["fsim"] if (e["wd"], e["hd"]) == (480, 270)][0]
    assert entry["col_bx"] in (8, 16) and entry["row_rb"] in (1, 2, 4)
    after = iqm_amd.FSIM().compute_metric(ctx, img, ref)
    assert abs(after.fsim - before.fsim) < 1e-7 and abs(after.fsimc - before.fsimc) < 1e-7
    ctx.import_tunes(json.dumps({"fsim": [dict(entry, col_bx=8, row_rb=1, fft_rb=2)]}))
    again = json.loads(ctx.export_tunes())["fsim"]
    assert [e for e in again if (e["wd"], e["hd"]) == (480, 270)][0]["col_bx"] == 8
    r = iqm_amd.FSIM().compute_metric(ctx, img, ref)
    assert abs(r.fsim - before.fsim) < 1e-7
    ctx.import_tunes(json.dumps({"fsim": [entry]}))
