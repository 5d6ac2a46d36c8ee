"""The 4K stream workload on the GPU: consecutive 3840x2160 RGB frame pairs
cut from a seeded 6240x4560 1/f^2 canvas (seed 5) shifted by (dx, dy) in
{1..4}^2 per frame (seeds 5000+t), test frame = distortion of the crop
(seed 6000+t), all made on the device; FSIMc with the network's PC and an
exact spot check on every 30th frame (tolerance 0.02), in batches of 30.

  * the spot-check error of each batch equals a host recomputation over the
    same seeded tiles (network PC from iqm_sur_infer, exact PC from iqm_fsim);
  * a batch that falls back gives exactly iqm_fsim_batch_dev's scores;
  * one stream frame's exact FSIM/FSIMc against the FP64 oracle (1e-5), and
    its simc map (1e-4).
Two batches (60 frames) keep the test short; the benchmark runs 600.
"""
import os

import numpy as np
import pytest

import iqm_amd
from iqm_amd import synth
from oracle import oracle

from test_gpu_surrogate import _spot_err

pytestmark = pytest.mark.gpu

FW, FH, CW, CHH, B = 3840, 2160, 6240, 4560, 30


def test_stream_4k_surrogate_spot_checks(ctx):
    fb = FW * FH * 3
    d_canvas = ctx.alloc(CW * CHH * 3)
    ctx.synth_reference(CW, CHH, 5, d_canvas)
    offs, ox, oy = [], 0, 0
    for t in range(2 * B):
        if t:
            u = synth.uniform(5000 + t, 2)
            ox += 1 + min(3, int(4 * u[0]))
            oy += 1 + min(3, int(4 * u[1]))
        offs.append((ox, oy))
    d_img, d_ref = ctx.alloc(B * fb), ctx.alloc(B * fb)
    sur = iqm_amd.Surrogate(ctx, seed=1)
    with open(os.path.join(os.path.dirname(iqm_amd.__file__), "weights", "pc_surrogate.iqmsur"), "rb") as f:
        sur.load(f.read())
    sur.set_failure_memory(False)  # every batch takes its own spot check here
    fsim = iqm_amd.FSIM()
    _, wd, hd = iqm_amd.fsim_dims(FW, FH)
    for b0 in (0, B):
        for i in range(B):
            t = b0 + i
            off = (offs[t][1] * CW + offs[t][0]) * 3
            ctx.copy2d(d_ref + i * fb, FW * 3, d_canvas + off, CW * 3, FW * 3, FH)
            ctx.synth_distort(d_canvas + off, CW * 3, FW, FH, 6000 + t, d_img + i * fb)
        fs, fc, err, back = sur.fsim_batch_dev(d_img, d_ref, FW, FH, 3, fb, B, spot_every=30, spot_tol=0.02)
        ex_fs, ex_fc = fsim.compute_batch_dev(ctx, d_img, d_ref, FW, FH, 3, fb, B)
        # host recomputation of the spot check on frame b0 (the batch's pair 0)
        img = ctx.download(d_img, np.empty((FH, FW, 3), np.uint8))
        ref = ctx.download(d_ref, np.empty((FH, FW, 3), np.uint8))
        r = fsim.compute_metric(ctx, img, ref, maps=("pc_in", "pc_ref", "simc"), detail=True)
        x = (np.stack([r.detail["yiq_in"][2], r.detail["yiq_ref"][2]]) * np.float32(1.0 / 255.0)).astype(np.float32)
        dx, dy = ctx.alloc(x.nbytes), ctx.alloc(x.nbytes)
        ctx.upload(dx, x)
        sur.infer(dx, dy, 2, hd, wd, fused=True)
        pc_net = ctx.download(dy, np.empty((2, hd, wd), np.float32))
        ctx.free(dx)
        ctx.free(dy)
        pc_ex = [r.maps["pc_in"].reshape(hd, wd), r.maps["pc_ref"].reshape(hd, wd)]
        host_err = _spot_err(pc_net, pc_ex, wd, hd, 0)
        print(f"[stream] frames {b0}..{b0 + B - 1}: spot_err gpu {err:.6f} host {host_err:.6f} "
              f"{'fallback' if back else 'network'}; exact FSIMc[0] {ex_fc[0]:.6f}", flush=True)
        assert abs(err - host_err) <= 1e-6
        assert back == (not host_err <= 0.02)
        if back:
            assert np.array_equal(fs, ex_fs) and np.array_equal(fc, ex_fc)
        assert r.fsim == ex_fs[0] and r.fsimc == ex_fc[0]
        if b0 == B:
            o = oracle.fsim(img, ref, prec=64, maps=("simc",))
            e = (abs(r.fsim - o.fsim), abs(r.fsimc - o.fsimc), float(np.max(np.abs(r.maps["simc"] - o.maps["simc"]))))
            print(f"[stream] frame {b0} vs FP64 oracle: |dFSIM| {e[0]:.2e} |dFSIMc| {e[1]:.2e} simc map {e[2]:.2e}",
                  flush=True)
            assert e[0] <= 1e-5 and e[1] <= 1e-5 and e[2] <= 1e-4
    sur.close()
    for p in (d_img, d_ref, d_canvas):
        ctx.free(p)
