"""The uneven-size batch at its benchmark workload: 256 RGB pairs with sides
in [360, 2160] and aspect ratio in [1, 2] (iqm_amd.synth.mixed_sizes, seeds
3000+i, made on the device), SSIM + FSIM/FSIMc per pair through
iqm_batch_dev (the batched form of the reference's per-pair drivers,
src/main.cpp:69-113,180-214).

  * the first 10 pairs against the FP64 oracle: MSSIM, FSIM and FSIMc within
    1e-5 absolute, the SSIM map and the FSIMc map within 1e-4 max-abs (most
    of these pairs have downscaled lines with a prime factor above 13, i.e.
    the Bluestein transforms);
  * the full 256-pair batch by properties that hold at any size: every score
    finite and in [0, 1] (FSIM, FSIMc) or [-1, 1] (MSSIM), equal to the
    per-pair calls within 1e-6 (they run the same stage bodies, except for the
    compile-time two-level FFT kernels a single call takes for 480/270/256-long
    lines), and the captured-graph replay bit-equal to the eager first call;
  * a batch holding the single-pair path's two-level sizes (1920x1080 RGB,
    512x512 gray, 256x256 RGB) against the single calls and the oracle;
  * the graph caches stay bounded when every call brings new buffers.
"""
import numpy as np
import pytest

import iqm_amd
from iqm_amd import synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL_SCALAR = 1e-5
TOL_MAP = 1e-4


def _make_jobs(ctx, sizes, seed0, ch=3):
    jobs = []
    for i, (w, h) in enumerate(sizes):
        _, wd, hd = iqm_amd.fsim_dims(w, h)
        di, dr = ctx.alloc(w * h * 3), ctx.alloc(w * h * 3)
        ctx.synth_pair(w, h, seed0 + i, di, dr)
        jobs.append({"d_img": di, "d_ref": dr, "w": w, "h": h, "ch": ch, "methods": 3,
                     "d_ssim_map": ctx.alloc(4 * w * h), "d_simc_map": ctx.alloc(4 * wd * hd),
                     "wd": wd, "hd": hd})
    ctx.sync()
    return jobs


def _free(ctx, jobs):
    for j in jobs:
        for k in ("d_img", "d_ref", "d_ssim_map", "d_simc_map"):
            if j.get(k):
                ctx.free(j[k])


def _host_pair(ctx, j):
    w, h, ch = j["w"], j["h"], j["ch"]
    shape = (h, w) if ch == 1 else (h, w, ch)
    img = ctx.download(j["d_img"], np.empty(shape, np.uint8))
    ref = ctx.download(j["d_ref"], np.empty(shape, np.uint8))
    return img, ref


def _check_vs_oracle(ctx, jobs, res, tag):
    worst = np.zeros(5)
    for k, j in enumerate(jobs):
        img, ref = _host_pair(ctx, j)
        ms, smap = oracle.ssim(img, ref, prec=64)
        o = oracle.fsim(img, ref, prec=64, maps=("simc",))
        gs = ctx.download(j["d_ssim_map"], np.empty((j["h"], j["w"]), np.float32))
        gc = ctx.download(j["d_simc_map"], np.empty((j["hd"], j["wd"]), np.float32))
        e = np.array([abs(res[k, 0] - ms), abs(res[k, 1] - o.fsim), abs(res[k, 2] - o.fsimc),
                      float(np.max(np.abs(gs - smap))), float(np.max(np.abs(gc - o.maps["simc"])))])
        worst = np.maximum(worst, e)
        print(f"[{tag}] pair {k} {j['w']}x{j['h']} -> {j['wd']}x{j['hd']}: |dMSSIM| {e[0]:.2e} |dFSIM| {e[1]:.2e} "
              f"|dFSIMc| {e[2]:.2e} ssim map {e[3]:.2e} simc map {e[4]:.2e}", flush=True)
        assert e[0] <= TOL_SCALAR and e[1] <= TOL_SCALAR and e[2] <= TOL_SCALAR, (k, e)
        assert e[3] <= TOL_MAP and e[4] <= TOL_MAP, (k, e)
    return worst


def test_mixed_workload_pairs_vs_oracle(ctx):
    sizes = synth.mixed_sizes(256, 3000)[:10]
    print(f"[mixed-oracle] sizes {sizes}", flush=True)
    jobs = _make_jobs(ctx, sizes, 3000)
    res = ctx.batch_dev(jobs)
    worst = _check_vs_oracle(ctx, jobs, res, "mixed-oracle")
    print(f"[mixed-oracle] worst vs FP64: {worst}", flush=True)
    _free(ctx, jobs)


def test_mixed_workload_full_batch_properties(ctx):
    sizes = synth.mixed_sizes(256, 3000)
    jobs = _make_jobs(ctx, sizes, 3000)
    print(f"[mixed-256] {len(jobs)} pairs made", flush=True)
    first = ctx.batch_dev(jobs)            # eager run + capture
    ssim_first = [ctx.download(j["d_ssim_map"], np.empty(j["w"] * j["h"], np.float32)) for j in jobs[:8]]
    again = ctx.batch_dev(jobs)            # graph replay
    ctx.sync()
    assert np.array_equal(first, again)
    for j, m in zip(jobs[:8], ssim_first):
        assert np.array_equal(ctx.download(j["d_ssim_map"], np.empty(j["w"] * j["h"], np.float32)), m)
    assert np.all(np.isfinite(first))
    assert np.all((first[:, 1:] >= 0) & (first[:, 1:] <= 1))
    assert np.all((first[:, 0] >= -1) & (first[:, 0] <= 1))
    ssim, fsim = iqm_amd.SSIM(), iqm_amd.FSIM()
    ctx.set_graphs(False)
    sep = np.empty_like(first)
    for k, j in enumerate(jobs):
        sep[k, 0] = ssim.compute_batch_dev(ctx, j["d_img"], j["d_ref"], j["w"], j["h"], 3, 0, 1)[0]
        fs, fc = fsim.compute_batch_dev(ctx, j["d_img"], j["d_ref"], j["w"], j["h"], 3, 0, 1)
        sep[k, 1], sep[k, 2] = fs[0], fc[0]
    ctx.set_graphs(True)
    d = np.abs(sep - first)
    two_level = [k for k, j in enumerate(jobs) if {j["wd"], j["hd"]} & {480, 270, 256}]
    print(f"[mixed-256] max |batch - per-pair| = {d.max():.3e} (pairs with two-level lines: {two_level}); "
          f"FSIMc range [{first[:, 2].min():.4f}, {first[:, 2].max():.4f}]", flush=True)
    assert d.max() <= 1e-6
    if not two_level:
        assert d.max() == 0.0
    _free(ctx, jobs)


def test_batch_with_two_level_sizes(ctx):
    """1080p RGB, 512x512 gray and 256x256 RGB in one mixed batch: the batch runs
    the run-time transforms, a single call the compile-time two-level kernels."""
    specs = [((1920, 1080), 3), ((512, 512), 1), ((256, 256), 3)]
    jobs = []
    for i, ((w, h), ch) in enumerate(specs):
        img, ref = synth.rgb_pair(w, h, seed=60 + i)
        if ch == 1:
            img, ref = img[:, :, 1].copy(), ref[:, :, 1].copy()
        _, wd, hd = iqm_amd.fsim_dims(w, h)
        di, dr = ctx.alloc(img.nbytes), ctx.alloc(ref.nbytes)
        ctx.upload(di, img)
        ctx.upload(dr, ref)
        jobs.append({"d_img": di, "d_ref": dr, "w": w, "h": h, "ch": ch, "methods": 3,
                     "d_ssim_map": ctx.alloc(4 * w * h), "d_simc_map": ctx.alloc(4 * wd * hd), "wd": wd, "hd": hd})
    res = ctx.batch_dev(jobs)
    fsim, ssim = iqm_amd.FSIM(), iqm_amd.SSIM()
    for k, j in enumerate(jobs):
        img, ref = _host_pair(ctx, j)
        single = fsim.compute_metric(ctx, img, ref, maps=("simc",))
        s = ssim.compute_metric(ctx, img, ref)
        gc = ctx.download(j["d_simc_map"], np.empty((j["hd"], j["wd"]), np.float32))
        gs = ctx.download(j["d_ssim_map"], np.empty((j["h"], j["w"]), np.float32))
        e = (abs(res[k, 1] - single.fsim), abs(res[k, 2] - single.fsimc), float(np.max(np.abs(gc - single.maps["simc"]))))
        print(f"[two-level] {j['w']}x{j['h']}x{j['ch']}: batch vs single |dFSIM| {e[0]:.2e} |dFSIMc| {e[1]:.2e} "
              f"simc map {e[2]:.2e}", flush=True)
        assert e[0] <= 1e-6 and e[1] <= 1e-6 and e[2] <= 1e-5
        assert res[k, 0] == s.mssim and np.array_equal(gs, s.image_data)
    _check_vs_oracle(ctx, jobs, res, "two-level")
    _free(ctx, jobs)


def test_graph_caches_stay_bounded(ctx):
    w, h, n = 320, 240, 1
    fsim = iqm_amd.FSIM()
    bufs = []
    for i in range(20):
        di, dr = ctx.alloc(w * h * 3), ctx.alloc(w * h * 3)
        ctx.synth_pair(w, h, 500 + i, di, dr)
        bufs.append((di, dr))
    want = None
    for di, dr in bufs:
        fs, fc = fsim.compute_batch_dev(ctx, di, dr, w, h, 3, 0, n)
        ctx.batch_dev([{"d_img": di, "d_ref": dr, "w": w, "h": h, "ch": 3, "methods": 2}])
        a, b = ctx.graph_cache_sizes()
        assert a <= 16 and b <= 16, (a, b)
        want = (fs, fc) if want is None else want
    a, b = ctx.graph_cache_sizes()
    print(f"[graphs] after 20 distinct buffers: {a} same-size, {b} mixed graphs", flush=True)
    assert a == 16 and b == 16
    # the least recently used entries were evicted; the first buffers still score the same
    fs, fc = fsim.compute_batch_dev(ctx, bufs[0][0], bufs[0][1], w, h, 3, 0, n)
    assert np.array_equal(fs, want[0]) and np.array_equal(fc, want[1])
    for di, dr in bufs:
        ctx.free(di)
        ctx.free(dr)
