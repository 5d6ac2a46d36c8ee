"""GPU tests of the learned phase-congruency surrogate through the C-ABI.

The network (6 conv3x3 layers, BF16 on the matrix cores, FP32 accumulation)
is checked against tests/sur_ref.py, PyTorch's CPU conv2d in float64 with the
same BF16 roundings.  Bounds:
  forward   max |out_gpu - out_ref| <= 2e-2 * max |out_ref|, mean <= 2e-3 * max |out_ref|
            (FP32 against FP64 sums before each BF16 rounding can move single
            activations by one BF16 step; a wrong matrix-core lane layout gives
            errors of the order of the output itself);
  gradient  per layer: ||g_gpu - g_ref|| <= 3e-2 ||g_ref|| (the HIP backward
            keeps the activation gradients in BF16);
  PC path   FSIM/FSIMc with PC from the network within 1e-5 of the same
            formulas evaluated in FP64 on the host from the network's own maps.
"""
import numpy as np
import pytest

import iqm_amd
from iqm_amd import synth

import sur_ref

pytestmark = pytest.mark.gpu


def _dev(ctx, a):
    a = np.ascontiguousarray(a)
    d = ctx.alloc(a.nbytes)
    ctx.upload(d, a)
    return d


def _layer_slices():
    s = [("w1+b1", slice(sur_ref.OFF_W1, sur_ref.OFF_MID))]
    for l in range(4):
        base = sur_ref.OFF_MID + l * sur_ref.MID
        s.append((f"layer{l + 2}", slice(base, base + sur_ref.MID)))
    s.append(("w6+b6", slice(sur_ref.OFF_W6, sur_ref.TOTAL)))
    return s


@pytest.fixture(scope="module")
def sur(ctx):
    s = iqm_amd.Surrogate(ctx, seed=11)
    yield s
    s.close()


def test_param_count():
    assert iqm_amd.Surrogate.param_count() == sur_ref.TOTAL == 37601


@pytest.mark.parametrize("n,h,w", [(3, 37, 29), (2, 16, 16), (1, 5, 70)])
def test_forward_matches_torch(ctx, sur, n, h, w):
    print(f"[sur forward] n={n} {w}x{h}", flush=True)
    p = sur_ref.he_uniform(100 + h)
    sur.set_params(p)
    assert np.array_equal(sur.params(), p)
    x = np.random.default_rng(h * w).uniform(0, 1, (n, h, w)).astype(np.float32)
    dx, dy = _dev(ctx, x), ctx.alloc(x.nbytes)
    sur.infer(dx, dy, n, h, w)
    y = ctx.download(dy, np.empty_like(x))
    ctx.free(dx)
    ctx.free(dy)
    ref = sur_ref.predict(p, x)
    scale = float(np.max(np.abs(ref)))
    err = np.abs(y - ref)
    print(f"[sur forward] max={err.max():.3e} mean={err.mean():.3e} scale={scale:.3e}", flush=True)
    assert scale > 1e-3
    assert err.max() <= 2e-2 * scale and err.mean() <= 2e-3 * scale


@pytest.mark.parametrize("n,h,w", [(3, 37, 29), (1, 5, 70), (2, 270, 480), (1, 1, 7), (1, 131, 200)])
def test_fused_forward_matches_layers(ctx, sur, n, h, w):
    """The single all-layer kernel equals the layer-by-layer kernels: layers
    1-5 bit for bit, layer 6 up to FP32 summation order."""
    print(f"[sur fused] n={n} {w}x{h}", flush=True)
    p = sur_ref.he_uniform(300 + w)
    sur.set_params(p)
    x = np.random.default_rng(h + w).uniform(0, 1, (n, h, w)).astype(np.float32)
    dx, d1, d2 = _dev(ctx, x), ctx.alloc(x.nbytes), ctx.alloc(x.nbytes)
    sur.infer(dx, d1, n, h, w, fused=False)
    sur.infer(dx, d2, n, h, w, fused=True)
    a = ctx.download(d1, np.empty_like(x))
    b = ctx.download(d2, np.empty_like(x))
    for d_ in (dx, d1, d2):
        ctx.free(d_)
    scale = float(np.max(np.abs(a)))
    err = float(np.max(np.abs(a - b)))
    print(f"[sur fused] max |fused - layers| = {err:.3e} (scale {scale:.3e})", flush=True)
    assert scale > 1e-3 and err <= 1e-5 * scale
    if n * h * w <= 20000:
        ref = sur_ref.predict(p, x)
        assert np.max(np.abs(b - ref)) <= 2e-2 * float(np.max(np.abs(ref)))


def test_gradient_matches_torch(ctx, sur):
    n, h, w = 4, 40, 36
    print(f"[sur gradient] n={n} {w}x{h}", flush=True)
    p = sur_ref.he_uniform(21)
    sur.set_params(p)
    rng = np.random.default_rng(22)
    x = rng.uniform(0, 1, (n, h, w)).astype(np.float32)
    t = rng.uniform(0, 1, (n, h, w)).astype(np.float32)
    dx, dt = _dev(ctx, x), _dev(ctx, t)
    loss = sur.loss_grad(dx, dt, n, h, w)
    g = sur.grad()
    ctx.free(dx)
    ctx.free(dt)
    loss_ref, g_ref, _ = sur_ref.loss_and_grad(p, x, t)
    print(f"[sur gradient] loss={loss:.6e} ref={loss_ref:.6e}", flush=True)
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref
    for name, sl in _layer_slices():
        a, b = g[sl].astype(np.float64), g_ref[sl]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"[sur gradient] {name}: rel={rel:.3e} |g|={np.linalg.norm(b):.3e}", flush=True)
        assert rel <= 3e-2, name


def test_weighted_gradient_matches_torch(ctx, sur):
    """The FSIM-weighted training loss (each pixel's squared error times
    max(t of its pair's two images) + eps) against the same loss in torch."""
    n, h, w, eps = 4, 24, 20, 0.05
    print(f"[sur weighted gradient] n={n} {w}x{h} eps={eps}", flush=True)
    p = sur_ref.he_uniform(23)
    sur.set_params(p)
    rng = np.random.default_rng(24)
    x = rng.uniform(0, 1, (n, h, w)).astype(np.float32)
    t = rng.uniform(0, 1, (n, h, w)).astype(np.float32)
    dx, dt = _dev(ctx, x), _dev(ctx, t)
    sur.set_loss_weight(eps)
    try:
        loss = sur.loss_grad(dx, dt, n, h, w)
    finally:
        sur.set_loss_weight(None)
    g = sur.grad()
    ctx.free(dx)
    ctx.free(dt)
    loss_ref, g_ref, _ = sur_ref.loss_and_grad(p, x, t, weight_eps=eps)
    print(f"[sur weighted gradient] loss={loss:.6e} ref={loss_ref:.6e}", flush=True)
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref
    for name, sl in _layer_slices():
        a, b = g[sl].astype(np.float64), g_ref[sl]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"[sur weighted gradient] {name}: rel={rel:.3e}", flush=True)
        assert rel <= 3e-2, name


def test_patches_and_targets(ctx, sur):
    n, s = 8, 64
    print(f"[sur patches] n={n} size={s}", flush=True)
    dx, dp = ctx.alloc(4 * n * s * s), ctx.alloc(4 * n * s * s)
    sur.make_patches(n, s, 77, dx, dp)
    x = ctx.download(dx, np.empty((n, s, s), np.float32))
    pc = ctx.download(dp, np.empty((n, s, s), np.float32))
    ctx.free(dx)
    ctx.free(dp)
    y = x * 255.0
    assert np.max(np.abs(y - np.rint(y))) < 1e-3 and y.min() >= 0 and y.max() <= 255
    u8 = np.rint(y).astype(np.uint8)
    assert len({u8[i].tobytes() for i in range(n)}) == n
    fsim = iqm_amd.FSIM()
    for i in range(n // 2):
        r = fsim.compute_metric(ctx, u8[2 * i], u8[2 * i + 1], maps=("pc_in", "pc_ref"))
        for k, key in ((0, "pc_in"), (1, "pc_ref")):
            ref = np.nan_to_num(r.maps[key].reshape(s, s), nan=0.0)
            assert np.max(np.abs(pc[2 * i + k] - ref)) < 1e-6


def test_training_reduces_loss(ctx):
    n, s = 32, 64
    print(f"[sur train] n={n} size={s}", flush=True)
    sur = iqm_amd.Surrogate(ctx, seed=5)
    dx, dp = ctx.alloc(4 * n * s * s), ctx.alloc(4 * n * s * s)
    sur.make_patches(n, s, 1234, dx, dp)
    losses = [sur.train_step(dx, dp, n, s, s, 1e-3) for _ in range(80)]
    ctx.free(dx)
    ctx.free(dp)
    print(f"[sur train] loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0]
    assert sur.state() == (80, 80 * n)
    sur.close()


def test_checkpoint_resume_is_exact(ctx):
    n, s = 8, 32
    print(f"[sur resume] n={n} size={s}", flush=True)
    dx, dp = ctx.alloc(4 * n * s * s), ctx.alloc(4 * n * s * s)
    a = iqm_amd.Surrogate(ctx, seed=9)
    a.make_patches(n, s, 55, dx, dp)
    for _ in range(3):
        a.train_step(dx, dp, n, s, s, 1e-3)
    blob = a.save()
    for _ in range(4):
        a.train_step(dx, dp, n, s, s, 1e-3)
    b = iqm_amd.Surrogate(ctx, seed=123)
    b.load(blob)
    assert b.state() == (3, 3 * n)
    for _ in range(4):
        b.train_step(dx, dp, n, s, s, 1e-3)
    assert np.array_equal(a.params(), b.params())
    assert a.state() == b.state()
    with pytest.raises(iqm_amd.IqmError):
        b.load(blob[:-4])
    ctx.free(dx)
    ctx.free(dp)
    a.close()
    b.close()


def _sm64(z):
    M = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def _spot_err(pc_net, pc_exact, wd, hd, j):
    """Largest |difference| over the 4 seeded 32x32 tiles of both images of spot pair j."""
    tw, th = min(32, wd), min(32, hd)
    m = 0.0
    for im in range(2):
        for t in range(4):
            z = _sm64(0x5EED5EED ^ (((j * 8 + im * 4 + t) * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)))
            x0 = (z & 0xFFFFFFFF) % (wd - tw + 1)
            y0 = (z >> 32) % (hd - th + 1)
            d = np.abs(pc_net[im][y0:y0 + th, x0:x0 + tw] - pc_exact[im][y0:y0 + th, x0:x0 + tw])
            d = d[np.isfinite(d)]
            if d.size:
                m = max(m, float(d.max()))
    return m


def _host_fsim(pc, grad, yiq):
    """F16 + F17 in FP64 from PC, gradient and YIQ planes (fsim_final_multiply.glsl:26-57)."""
    p1, p2 = pc[0].astype(np.float64), pc[1].astype(np.float64)
    g1, g2 = grad[0].astype(np.float64), grad[1].astype(np.float64)
    i1, q1 = yiq[0][0].astype(np.float64), yiq[0][1].astype(np.float64)
    i2, q2 = yiq[1][0].astype(np.float64), yiq[1][1].astype(np.float64)
    pcm = np.maximum(p1, p2)
    spc = (2 * p1 * p2 + 0.85) / (p1 * p1 + p2 * p2 + 0.85)
    sg = (2 * g1 * g2 + 160) / (g1 * g1 + g2 * g2 + 160)
    si = (2 * i1 * i2 + 200) / (i1 * i1 + i2 * i2 + 200)
    sq = (2 * q1 * q2 + 200) / (q1 * q1 + q2 * q2 + 200)
    sim = sg * spc * pcm
    simc = sg * spc * np.abs(si * sq) ** 0.03 * pcm
    return sim.sum() / pcm.sum(), simc.sum() / pcm.sum()


def test_fsim_with_surrogate_pc(ctx, sur):
    w, h, n = 320, 240, 3
    print(f"[sur fsim] {w}x{h} n={n}", flush=True)
    sur.set_params(sur_ref.he_uniform(31))
    pairs = [synth.rgb_pair(w, h, seed=700 + i) for i in range(n)]
    imgs = np.stack([p[0] for p in pairs])
    refs = np.stack([p[1] for p in pairs])
    di, dr = _dev(ctx, imgs), _dev(ctx, refs)
    stride = w * h * 3
    fs, fc, err, back = sur.fsim_batch_dev(di, dr, w, h, 3, stride, n, spot_every=2, spot_tol=1e9)
    assert not back
    fsim = iqm_amd.FSIM()
    _, wd, hd = iqm_amd.fsim_dims(w, h)
    worst = 0.0
    for i, (img, ref) in enumerate(pairs):
        r = fsim.compute_metric(ctx, img, ref, maps=("pc_in", "pc_ref"), detail=True)
        yiq = [r.detail["yiq_in"].reshape(3, hd, wd), r.detail["yiq_ref"].reshape(3, hd, wd)]
        grad = [r.detail["grad_in"].reshape(hd, wd), r.detail["grad_ref"].reshape(hd, wd)]
        x = np.stack([yiq[0][2], yiq[1][2]]) * np.float32(1.0 / 255.0)
        dx, dy = _dev(ctx, x.astype(np.float32)), ctx.alloc(x.nbytes)
        sur.infer(dx, dy, 2, hd, wd)
        pc_net = ctx.download(dy, np.empty((2, hd, wd), np.float32))
        ctx.free(dx)
        ctx.free(dy)
        hf, hfc = _host_fsim(pc_net, grad, yiq)
        print(f"[sur fsim] pair {i}: gpu {fs[i]:.8f}/{fc[i]:.8f} host {hf:.8f}/{hfc:.8f}", flush=True)
        assert abs(fs[i] - hf) < 1e-5 and abs(fc[i] - hfc) < 1e-5
        if i % 2 == 0:
            pc_ex = [r.maps["pc_in"].reshape(hd, wd), r.maps["pc_ref"].reshape(hd, wd)]
            worst = max(worst, _spot_err(pc_net, pc_ex, wd, hd, i // 2))
    print(f"[sur fsim] spot_err gpu={err:.6e} host={worst:.6e}", flush=True)
    assert abs(err - worst) <= 1e-6
    # a tolerance below any error forces the exact fallback for the whole
    # batch; the verdict is taken on the device and gates the exact kernels
    fs2, fc2, err2, back2 = sur.fsim_batch_dev(di, dr, w, h, 3, stride, n, spot_every=2, spot_tol=-1.0)
    ex_fs, ex_fc = fsim.compute_batch_dev(ctx, di, dr, w, h, 3, stride, n)
    assert back2 and err2 == err
    assert np.array_equal(fs2, ex_fs) and np.array_equal(fc2, ex_fc)
    # failure memory: this size failed, so a later checked batch runs exactly
    # even with a tolerance it would pass, and reports the remembered error
    fs3, fc3, err3, back3 = sur.fsim_batch_dev(di, dr, w, h, 3, stride, n, spot_every=2, spot_tol=1e9)
    assert back3 and err3 == err and np.array_equal(fs3, ex_fs) and np.array_equal(fc3, ex_fc)
    # memory off: the network runs again; repeated calls replay captured
    # graphs, passing (network scores) and failing (exact scores) alike
    sur.set_failure_memory(False)
    for _ in range(2):
        fs4, fc4, err4, back4 = sur.fsim_batch_dev(di, dr, w, h, 3, stride, n, spot_every=2, spot_tol=1e9)
        assert not back4 and err4 == err and np.array_equal(fs4, fs) and np.array_equal(fc4, fc)
        fs5, fc5, err5, back5 = sur.fsim_batch_dev(di, dr, w, h, 3, stride, n, spot_every=2, spot_tol=-1.0)
        assert back5 and err5 == err and np.array_equal(fs5, ex_fs) and np.array_equal(fc5, ex_fc)
        sur.set_failure_memory(False)
    sur.set_failure_memory(True)
    ctx.free(di)
    ctx.free(dr)


def test_trained_network_gain(ctx):
    """The committed network against the method without it.  FSIM: mean and
    max |FSIM_net - FSIM_exact| over 12 seeded 1920x1080 pairs (seeds
    4100..4111) for the trained weights and for a constant PC map (the FSIM
    formula with no phase-congruency estimate); the trained network must do
    better on the mean, and both are printed against the 1e-3 target, which
    the network does not reach on every pair.  PC maps: mean |PC_net -
    PC_exact| over both maps of the first 3 pairs for the trained weights, the
    same network untrained and the constant map."""
    import os

    w, h, n, n_pc = 1920, 1080, 12, 3
    print(f"[sur gain] {w}x{h} n={n}", flush=True)
    pb = w * h * 3
    di, dr = ctx.alloc(n * pb), ctx.alloc(n * pb)
    for i in range(n):
        ctx.synth_pair(w, h, 4100 + i, di + i * pb, dr + i * pb)
    ex_fs, _ = iqm_amd.FSIM().compute_batch_dev(ctx, di, dr, w, h, 3, pb, n)
    _, wd, hd = iqm_amd.fsim_dims(w, h)
    ys, pcs = [], []
    for i in range(n_pc):
        img = ctx.download(di + i * pb, np.empty((h, w, 3), np.uint8))
        ref = ctx.download(dr + i * pb, np.empty((h, w, 3), np.uint8))
        r = iqm_amd.FSIM().compute_metric(ctx, img, ref, maps=("pc_in", "pc_ref"), detail=True)
        ys += [r.detail["yiq_in"][2], r.detail["yiq_ref"][2]]
        pcs += [r.maps["pc_in"].reshape(hd, wd), r.maps["pc_ref"].reshape(hd, wd)]
    x = (np.stack(ys) * np.float32(1.0 / 255.0)).astype(np.float32)
    pc_ex = np.stack(pcs)
    dx, dy = _dev(ctx, x), ctx.alloc(x.nbytes)
    path = os.path.join(os.path.dirname(iqm_amd.__file__), "weights", "pc_surrogate.iqmsur")
    pc_err, fs_err = {}, {}
    for name in ("trained", "untrained", "constant"):
        s = iqm_amd.Surrogate(ctx, seed=7)
        if name == "trained":
            with open(path, "rb") as f:
                s.load(f.read())
        elif name == "constant":
            p = np.zeros(sur_ref.TOTAL, np.float32)
            p[sur_ref.OFF_B6] = 0.3
            s.set_params(p)
        s.infer(dx, dy, 2 * n_pc, hd, wd)
        pc = ctx.download(dy, np.empty_like(x))
        d = np.abs(pc - pc_ex)
        pc_err[name] = float(np.mean(d[np.isfinite(d)]))
        fs, _, _, _ = s.fsim_batch_dev(di, dr, w, h, 3, pb, n)
        df = np.abs(fs - ex_fs)
        fs_err[name] = (float(np.mean(df)), float(np.max(df)))
        s.close()
    for p_ in (di, dr, dx, dy):
        ctx.free(p_)
    print("[sur gain] mean |dPC|: " + " ".join(f"{k}={v:.3e}" for k, v in pc_err.items()), flush=True)
    print("[sur gain] |dFSIM| mean/max (target 1e-3): " +
          " ".join(f"{k}={m:.3e}/{x_:.3e}" for k, (m, x_) in fs_err.items()), flush=True)
    assert pc_err["trained"] < 0.5 * pc_err["untrained"] and pc_err["trained"] < 0.5 * pc_err["constant"]
    assert fs_err["trained"][0] < fs_err["constant"][0]
