"""Mixed batches (pairs of uneven sizes, SSIM and/or FSIM per pair) through
iqm_batch_dev: every scalar and map equals the single-pair calls bit for bit
(the same kernels run), on the eager first call and on graph replays with
other work on the stream in between."""
import numpy as np
import pytest

import iqm_amd
from iqm_amd import synth

pytestmark = pytest.mark.gpu

CASES = [((320, 240, 3), 3), ((240, 320, 3), 3), ((401, 299, 3), 2), ((97, 131, 3), 1),
         ((513, 257, 3), 3), ((360, 360, 1), 3), ((288, 200, 4), 3)]
# the batched SSIM sizes its staged rows for the widest input of the batch:
# 4 channels for CASES, 3 and 1 for these
CASES_RGB_GRAY = [c for c in CASES if c[0][2] in (1, 3)]
CASES_GRAY = [((320, 240, 1), 3), ((97, 131, 1), 1), ((401, 299, 1), 2)]


def _upload(ctx, a):
    d = ctx.alloc(a.nbytes)
    ctx.upload(d, a)
    return d


def _make(ctx, cases=CASES):
    jobs = []
    for i, ((w, h, ch), methods) in enumerate(cases):
        img, ref = synth.rgb_pair(w, h, seed=900 + i)
        if ch == 1:
            img, ref = img[:, :, 1].copy(), ref[:, :, 1].copy()
        elif ch == 4:
            img, ref = synth.rgba(img), synth.rgba(ref)
        _, wd, hd = iqm_amd.fsim_dims(w, h)
        j = {"d_img": _upload(ctx, img), "d_ref": _upload(ctx, ref), "w": w, "h": h, "ch": ch, "methods": methods,
             "d_ssim_map": ctx.alloc(4 * w * h) if methods & 1 else None,
             "d_simc_map": ctx.alloc(4 * wd * hd) if methods & 2 else None, "wd": wd, "hd": hd}
        jobs.append(j)
    return jobs


def _maps(ctx, j):
    out = {}
    if j["d_ssim_map"]:
        out["ssim"] = ctx.download(j["d_ssim_map"], np.empty(j["w"] * j["h"], np.float32))
    if j["d_simc_map"]:
        out["simc"] = ctx.download(j["d_simc_map"], np.empty(j["wd"] * j["hd"], np.float32))
    return out


@pytest.mark.parametrize("cases", [CASES, CASES_RGB_GRAY, CASES_GRAY], ids=["widest4", "widest3", "gray"])
def test_mixed_batch_matches_single_calls(ctx, cases):
    print(f"[batch] {len(cases)} pairs, channels {sorted({c[0][2] for c in cases})}", flush=True)
    jobs = _make(ctx, cases)
    ssim, fsim = iqm_amd.SSIM(), iqm_amd.FSIM()
    # single-pair reference values and maps
    want, want_maps = [], []
    for j in jobs:
        w, h, ch = j["w"], j["h"], j["ch"]
        m = [np.nan, np.nan, np.nan]
        if j["methods"] & 1:
            m[0] = ssim.compute_batch_dev(ctx, j["d_img"], j["d_ref"], w, h, ch, 0, 1, j["d_ssim_map"])[0]
        if j["methods"] & 2:
            fs, fc = fsim.compute_batch_dev(ctx, j["d_img"], j["d_ref"], w, h, ch, 0, 1, j["d_simc_map"])
            m[1], m[2] = fs[0], fc[0]
        want.append(m)
        want_maps.append(_maps(ctx, j))
    want = np.array(want)
    for j in jobs:  # clear the maps so the batch has to write them
        for k in ("d_ssim_map", "d_simc_map"):
            if j[k]:
                ctx.upload(j[k], np.zeros(j["w"] * j["h"] if k == "d_ssim_map" else j["wd"] * j["hd"], np.float32))
    for rep in range(3):
        got = ctx.batch_dev(jobs)
        assert np.array_equal(np.isnan(got), np.isnan(want)), rep
        assert np.array_equal(np.nan_to_num(got), np.nan_to_num(want)), rep
        for j, wm in zip(jobs, want_maps):
            gm = _maps(ctx, j)
            for k in wm:
                assert np.array_equal(gm[k], wm[k]), (rep, k)
        # other work on the stream between replays
        j0 = jobs[0]
        ssim.compute_batch_dev(ctx, j0["d_img"], j0["d_ref"], j0["w"], j0["h"], j0["ch"], 0, 1)
    for j in jobs:
        for k in ("d_img", "d_ref", "d_ssim_map", "d_simc_map"):
            if j[k]:
                ctx.free(j[k])


def test_batch_rejects_bad_entries(ctx):
    d = ctx.alloc(64 * 64 * 3)
    with pytest.raises(iqm_amd.IqmError, match=r"pair 1"):
        ctx.batch_dev([{"d_img": d, "d_ref": d, "w": 64, "h": 64, "ch": 3, "methods": 1},
                       {"d_img": d, "d_ref": d, "w": 8, "h": 8, "ch": 3, "methods": 1}])
    with pytest.raises(iqm_amd.IqmError, match=r"pair 0"):
        ctx.batch_dev([{"d_img": d, "d_ref": d, "w": 64, "h": 64, "ch": 2, "methods": 3}])
    ctx.free(d)
