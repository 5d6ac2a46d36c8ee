"""Seeded synthetic image pairs for the SSIM / FSIM workloads.

The generator is counter based (SplitMix64 -> Box-Muller, float64), so every
caller -- tests, the benchmark and the CPU baseline -- sees identical bytes
for a given seed.  Images are uint8 arrays of shape (h, w, 3) (RGB) or (h, w)
(gray).

Workloads:
  * ssim_gray_pair(512, 512, seed=1): 1/f^2 reference, test = blur(1.2) +
    N(0, 5^2), clipped.
  * rgb_pair(w, h, seed): colour-correlated (0.8) 1/f^2 reference, test =
    distort(reference) with blur sigma U[0.5,3], contrast U[0.7,1.3], noise
    sigma U[2,20] and 16-level quantisation.
  * mixed_sizes(n, seed0): (w, h) draws for the uneven-size batch.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Outputs offset..offset+n-1 of the SplitMix64 stream started at seed."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Uniform doubles in (0, 1]."""
    z = splitmix64(seed, n, offset)
    return ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Standard normal doubles by Box-Muller on pairs of uniforms."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, 2 * offset)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:n]


def pink_field(h: int, w: int, seed: int) -> np.ndarray:
    """1/f^2-power field: white N(0,1) filtered by 1/|f| on the fftfreq grid (DC := 0)."""
    white = normal(seed, h * w).reshape(h, w)
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.fftfreq(w)[None, :]
    f = np.sqrt(fx * fx + fy * fy)
    f[0, 0] = np.inf
    return np.fft.ifft2(np.fft.fft2(white) / f).real


def _rescale_u8(x: np.ndarray) -> np.ndarray:
    lo, hi = float(x.min()), float(x.max())
    y = (x - lo) * (255.0 / (hi - lo if hi > lo else 1.0))
    return np.clip(np.round(y), 0, 255).astype(np.uint8)


def gaussian_blur(img: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian, radius ceil(3 sigma), replicate border; float64 out."""
    r = int(np.ceil(3.0 * sigma))
    t = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-(t * t) / (2.0 * sigma * sigma))
    k /= k.sum()
    x = img.astype(np.float64)
    pad = [(r, r), (0, 0)] + [(0, 0)] * (x.ndim - 2)
    xp = np.pad(x, pad, mode="edge")
    acc = np.zeros_like(x)
    for i, kv in enumerate(k):
        acc += kv * xp[i:i + x.shape[0]]
    pad = [(0, 0), (r, r)] + [(0, 0)] * (x.ndim - 2)
    xp = np.pad(acc, pad, mode="edge")
    out = np.zeros_like(x)
    for i, kv in enumerate(k):
        out += kv * xp[:, i:i + x.shape[1]]
    return out


def rgb_reference(w: int, h: int, seed: int) -> np.ndarray:
    """Three channels of 1/f^2 noise with colour correlation 0.8, joint min-max to u8."""
    shared = pink_field(h, w, seed * 4 + 0)
    ch = [np.sqrt(0.8) * shared + np.sqrt(0.2) * pink_field(h, w, seed * 4 + c) for c in (1, 2, 3)]
    return _rescale_u8(np.stack(ch, axis=-1))


def distort(ref: np.ndarray, seed: int) -> tuple[np.ndarray, dict]:
    """Blur, contrast about the mean, additive noise, 16-level quantisation, clip."""
    u = uniform(seed ^ 0x5DEECE66D, 4)
    sigma_b = 0.5 + 2.5 * u[0]
    contrast = 0.7 + 0.6 * u[1]
    sigma_n = 2.0 + 18.0 * u[2]
    v = gaussian_blur(ref, sigma_b)
    m = v.mean()
    v = m + contrast * (v - m)
    v = v + sigma_n * normal(seed ^ 0x2545F4914F6CDD1D, v.size).reshape(v.shape)
    v = 17.0 * np.round(v / 17.0)
    img = np.clip(np.round(v), 0, 255).astype(np.uint8)
    return img, {"blur_sigma": sigma_b, "contrast": contrast, "noise_sigma": sigma_n}


def rgb_pair(w: int, h: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """(test, reference) RGB uint8 pair of the FSIMc workloads."""
    ref = rgb_reference(w, h, seed)
    img, _ = distort(ref, seed + 1000003)
    return img, ref


def ssim_gray_pair(w: int = 512, h: int = 512, seed: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """(test, reference) gray pair of the plain-SSIM workload."""
    ref = _rescale_u8(pink_field(h, w, seed))
    v = gaussian_blur(ref, 1.2) + 5.0 * normal(seed ^ 0x2545F4914F6CDD1D, w * h).reshape(h, w)
    img = np.clip(np.round(v), 0, 255).astype(np.uint8)
    return img, ref


def mixed_sizes(n: int = 256, seed0: int = 3000) -> list[tuple[int, int]]:
    """(w, h) of the uneven-size batch: long side U[360,2160], aspect U[1,2]."""
    out = []
    for i in range(n):
        u = uniform(seed0 + i, 3)
        long_side = int(round(360 + (2160 - 360) * u[0]))
        aspect = 1.0 + u[1]
        short = max(360, int(round(long_side / aspect)))
        out.append((long_side, short) if u[2] < 0.5 else (short, long_side))
    return out


def rgba(img: np.ndarray) -> np.ndarray:
    """The reference's RGBA8 view of an image (main.cpp:37-40): gray -> R=G=B, A=255."""
    if img.ndim == 2:
        img = np.repeat(img[:, :, None], 3, axis=2)
    a = np.full(img.shape[:2] + (1,), 255, np.uint8)
    return np.ascontiguousarray(np.concatenate([img[:, :, :3], a], axis=2))
