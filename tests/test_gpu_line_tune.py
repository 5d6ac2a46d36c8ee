"""FFT plan choice of the generic line transforms (iqm_autotune_lines and the
"lines" entries of the plan file), against the FP64 oracle.

The reference takes VkFFT's plan for every length (src/gpu/fsim/fsim.cpp:94,
160, configured at :370-437); here the two-level split of a line's plan is a
tunable choice per (length, role).  Any split computes the same DFT, so:

  * forced non-default splits of the three plan kinds -- a direct two-level
    length (480 = 32 x 15 instead of 20 x 24), Rader (281: p - 1 = 280 as
    28 x 10 for the columns and 10 x 28 for the rows instead of 14 x 20) and
    Bluestein (269: m = 540 as 30 x 18 instead of 20 x 27) -- give FSIM,
    FSIMc and the FSIMc map within 1e-5 / 1e-4 of the FP64 oracle and within
    1e-6 of the default plans;
  * a split that does not factor the plan is refused;
  * iqm_autotune_lines over config-3 pairs times every (length, role) of the
    batch, and a second context that imports the exported plan file gives
    bit-equal scores.
"""
import json

import numpy as np
import pytest

import iqm_amd
from iqm_amd import synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL_SCALAR = 1e-5
TOL_MAP = 1e-4


def _jobs(ctx, sizes, seed0):
    jobs = []
    for i, (w, h) in enumerate(sizes):
        _, wd, hd = iqm_amd.fsim_dims(w, h)
        di, dr = ctx.alloc(w * h * 3), ctx.alloc(w * h * 3)
        ctx.synth_pair(w, h, seed0 + i, di, dr)
        jobs.append({"d_img": di, "d_ref": dr, "w": w, "h": h, "ch": 3, "methods": iqm_amd.JOB_FSIM,
                     "d_simc_map": ctx.alloc(4 * wd * hd), "wd": wd, "hd": hd})
    ctx.sync()
    return jobs


def _free(ctx, jobs):
    for j in jobs:
        for k in ("d_img", "d_ref", "d_simc_map"):
            ctx.free(j[k])


def _maps(ctx, jobs):
    return [ctx.download(j["d_simc_map"], np.empty((j["hd"], j["wd"]), np.float32)) for j in jobs]


def _pairs_from(ctx, jobs, other):
    """The same pairs as `jobs` (made on ctx) copied to the context `other`."""
    out = []
    for j in jobs:
        w, h = j["w"], j["h"]
        img = ctx.download(j["d_img"], np.empty((h, w, 3), np.uint8))
        ref = ctx.download(j["d_ref"], np.empty((h, w, 3), np.uint8))
        di, dr = other.alloc(img.nbytes), other.alloc(ref.nbytes)
        other.upload(di, img)
        other.upload(dr, ref)
        out.append(dict(j, d_img=di, d_ref=dr, d_simc_map=other.alloc(4 * j["wd"] * j["hd"])))
    other.sync()
    return out


FORCED = {"lines": [{"line": 480, "role": 1, "split1": 32}, {"line": 269, "role": 0, "split1": 30},
                    {"line": 281, "role": 0, "split1": 28}, {"line": 281, "role": 1, "split1": 10}]}


def test_forced_splits_match_oracle_and_default():
    # 1920x1076 -> 480 x 269 (direct rows, Bluestein columns); 1124x1124 -> 281 x 281 (Rader)
    sizes = [(1920, 1076), (1124, 1124)]
    print(f"[line_tune] forced splits {FORCED['lines']} on {sizes}", flush=True)
    ref_ctx = iqm_amd.Context(0)
    ctx = iqm_amd.Context(0)
    try:
        jd = _jobs(ref_ctx, sizes, 7100)
        r_def = ref_ctx.batch_dev(jd)
        m_def = _maps(ref_ctx, jd)
        ctx.import_tunes(json.dumps(FORCED))
        plan = json.loads(ctx.export_tunes())
        assert sorted((e["line"], e["role"], e["split1"]) for e in plan["lines"]) == \
            sorted((e["line"], e["role"], e["split1"]) for e in FORCED["lines"])
        jf = _pairs_from(ref_ctx, jd, ctx)
        r = ctx.batch_dev(jf)
        m = _maps(ctx, jf)
        for k, j in enumerate(jd):
            img = ref_ctx.download(j["d_img"], np.empty((j["h"], j["w"], 3), np.uint8))
            ref = ref_ctx.download(j["d_ref"], np.empty((j["h"], j["w"], 3), np.uint8))
            o = oracle.fsim(img, ref, prec=64, maps=("simc",))
            e = (abs(r[k, 1] - o.fsim), abs(r[k, 2] - o.fsimc), float(np.max(np.abs(m[k] - o.maps["simc"]))))
            d = (float(np.max(np.abs(r[k, 1:] - r_def[k, 1:]))), float(np.max(np.abs(m[k] - m_def[k]))))
            print(f"[line_tune] pair {k} {j['wd']}x{j['hd']}: |dFSIM| {e[0]:.2e} |dFSIMc| {e[1]:.2e} "
                  f"map {e[2]:.2e}; vs default plans: scores {d[0]:.2e} map {d[1]:.2e}", flush=True)
            assert e[0] <= TOL_SCALAR and e[1] <= TOL_SCALAR and e[2] <= TOL_MAP, (k, e)
            assert d[0] <= 1e-6 and d[1] <= 1e-5, (k, d)
        _free(ctx, jf)
        _free(ref_ctx, jd)
    finally:
        ctx.close()
        ref_ctx.close()


def test_split_that_does_not_factor_is_refused(ctx):
    bad = {"lines": [{"line": 480, "role": 0, "split1": 7}]}
    with pytest.raises(iqm_amd.IqmError):
        ctx.import_tunes(json.dumps(bad))


def test_autotune_lines_batch_and_plan_file():
    sizes = synth.mixed_sizes(256, 3000)[:6]
    ctx = iqm_amd.Context(0)
    other = iqm_amd.Context(0)
    try:
        jobs = _jobs(ctx, sizes, 3000)
        r0 = ctx.batch_dev(jobs)
        n_tuned = ctx.autotune_lines(jobs)
        roles = {(j["hd"], 0) for j in jobs} | {(j["wd"], 1) for j in jobs}
        print(f"[line_tune] {n_tuned} (length, role) entries tuned for {sizes}", flush=True)
        assert n_tuned == len(roles)
        plan = ctx.export_tunes()
        lines = json.loads(plan)["lines"]
        assert {(e["line"], e["role"]) for e in lines} == roles
        r1 = ctx.batch_dev(jobs)
        chosen = [(e["line"], e["role"], e["split1"]) for e in lines if e["split1"]]
        print(f"[line_tune] non-default splits chosen: {chosen}; max |tuned - default| "
              f"{float(np.max(np.abs(r1[:, 1:] - r0[:, 1:]))):.2e}", flush=True)
        assert np.all(np.isfinite(r1[:, 1:]))
        assert float(np.max(np.abs(r1[:, 1:] - r0[:, 1:]))) <= 1e-6
        other.import_tunes(plan)
        jo = _pairs_from(ctx, jobs, other)
        r2 = other.batch_dev(jo)
        assert np.array_equal(r2[:, 1:], r1[:, 1:])
        _free(other, jo)
        _free(ctx, jobs)
    finally:
        other.close()
        ctx.close()
