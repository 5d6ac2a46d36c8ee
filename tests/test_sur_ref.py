"""Pins the CPU reference of the surrogate network (tests/sur_ref.py) on the CPU:
layer semantics against a brute-force loop, autograd against central finite
differences, and torch's BF16 rounding against the round-to-nearest-even
rule the HIP kernels implement."""
import numpy as np

import sur_ref


def _brute_forward(p, x):
    """Direct loops: out[co,y,x] = b[co] + sum_ci,ky,kx W[co,ci,ky,kx] * a[ci,y+ky-1,x+kx-1]."""
    layers = sur_ref.unpack(p.astype(np.float64))
    a = x[None].astype(np.float64)
    for i, (W, b) in enumerate(layers):
        co_n, ci_n = W.shape[:2]
        h, w = a.shape[1:]
        z = np.zeros((co_n, h, w))
        for co in range(co_n):
            for y in range(h):
                for xx in range(w):
                    s = b[co]
                    for ky in range(3):
                        for kx in range(3):
                            yy, xc = y + ky - 1, xx + kx - 1
                            if 0 <= yy < h and 0 <= xc < w:
                                s += np.dot(W[co, :, ky, kx], a[:, yy, xc])
                    z[co, y, xx] = s
        a = z if i == 5 else np.maximum(z, 0)
    return a[0]


def test_forward_matches_brute_force():
    p = sur_ref.he_uniform(3)
    x = np.random.default_rng(4).uniform(0, 1, (1, 5, 4))
    ref = sur_ref.predict(p, x, emulate=False)[0]
    brute = _brute_forward(p, x[0])
    assert np.max(np.abs(ref - brute)) < 1e-12


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = sur_ref.he_uniform(6).astype(np.float64)
    x = rng.uniform(0, 1, (2, 6, 5))
    t = rng.uniform(0, 1, (2, 6, 5))
    _, g, _ = sur_ref.loss_and_grad(p, x, t, emulate=False)

    def loss(q):
        return float(np.mean((sur_ref.predict(q, x, emulate=False) - t) ** 2))

    idx = np.concatenate([rng.integers(0, sur_ref.OFF_MID, 8), rng.integers(sur_ref.OFF_MID, sur_ref.OFF_W6, 16),
                          rng.integers(sur_ref.OFF_W6, sur_ref.TOTAL, 6), [sur_ref.OFF_B6]])
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp = loss(q)
        q[i] -= 2 * eps
        lm = loss(q)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def _rne_bits(f32: np.ndarray) -> np.ndarray:
    u = f32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def test_bf16_rounding_is_nearest_even():
    import torch

    rng = np.random.default_rng(7)
    v = np.concatenate([rng.normal(0, 3, 4000), [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8), 0.0, 1e-30]])
    v = v.astype(np.float32)
    ours = _rne_bits(v.copy())
    theirs = sur_ref.bf16(torch.tensor(v)).numpy().astype(np.float32)
    assert np.array_equal(ours.view(np.uint32), theirs.view(np.uint32))
