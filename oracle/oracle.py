"""ctypes front end of the CPU oracle (oracle/liborc.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's cpu_baseline leg import this module;
the product path (iqm_amd) never does.  The C++ restatement it wraps follows the
reference's shaders and host code file by file (see iqm_oracle.cpp).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None


def build() -> str:
    """Compile liborc.so with the committed Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class _FsimOut(ctypes.Structure):
    _fields_ = [
        ("F", ctypes.c_int), ("wd", ctypes.c_int), ("hd", ctypes.c_int),
        ("fsim", ctypes.c_double), ("fsimc", ctypes.c_double),
        ("em_n", ctypes.c_double * 4), ("est", ctypes.c_double * 8),
        ("median", ctypes.c_double * 8), ("noise_power", ctypes.c_double * 8),
        ("T", ctypes.c_double * 8), ("sums", ctypes.c_double * 3),
        ("yiq_in", ctypes.c_void_p), ("yiq_ref", ctypes.c_void_p),
        ("grad_in", ctypes.c_void_p), ("grad_ref", ctypes.c_void_p),
        ("an_e", ctypes.c_void_p), ("keys", ctypes.c_void_p),
        ("pc_in", ctypes.c_void_p), ("pc_ref", ctypes.c_void_p),
        ("pcm", ctypes.c_void_p), ("sim", ctypes.c_void_p), ("simc", ctypes.c_void_p),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        u8 = ctypes.c_void_p
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_fft1d.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_fft2d.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_ssim.argtypes = [ctypes.c_int, u8, u8, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               dp, dp]
        L.orc_fsim.argtypes = [ctypes.c_int, u8, u8, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(_FsimOut)]
        L.orc_fsim_filters.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.orc_fsim_noise_consts.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_fsim_dims.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _img(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    if a.ndim == 2:
        a = a[:, :, None]
    return a


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return lib().orc_get_threads()


def fft1d(x, inverse=False, normalize=False, prec=64):
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    buf = x.view(np.float64)
    rc = lib().orc_fft1d(prec, _dp(buf), x.size, 1 if inverse else -1, int(normalize))
    assert rc == 0
    return x


def fft2d(x, inverse=False, normalize=False, prec=64):
    """2-D DFT of an (h, w) complex array (x contiguous)."""
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    h, w = x.shape
    rc = lib().orc_fft2d(prec, _dp(x.view(np.float64)), w, h, 1 if inverse else -1, int(normalize))
    assert rc == 0
    return x


def ssim(img, ref, prec=64, kernel_size=11, k1=0.01, k2=0.03, sigma=1.5, want_map=True):
    """Returns (mssim, map) for two (h, w[, ch]) uint8 images."""
    a, b = _img(img), _img(ref)
    h, w, ch = a.shape
    m = np.empty((h, w), np.float64) if want_map else None
    ms = np.zeros(1, np.float64)
    rc = lib().orc_ssim(prec, a.ctypes.data, b.ctypes.data, w, h, ch, kernel_size, k1, k2, sigma,
                        _dp(m), _dp(ms))
    if rc != 0:
        raise ValueError(f"orc_ssim failed ({rc})")
    return float(ms[0]), m


def dims(w, h):
    F, wd, hd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().orc_fsim_dims(w, h, ctypes.byref(F), ctypes.byref(wd), ctypes.byref(hd))
    return F.value, wd.value, hd.value


@dataclass
class FsimResult:
    fsim: float
    fsimc: float
    F: int
    wd: int
    hd: int
    em_n: np.ndarray
    est: np.ndarray
    median: np.ndarray
    noise_power: np.ndarray
    T: np.ndarray
    sums: np.ndarray
    maps: dict = field(default_factory=dict)


_MAP_SHAPES = {
    "yiq_in": 3, "yiq_ref": 3, "grad_in": 1, "grad_ref": 1, "an_e": 16, "keys": 8,
    "pc_in": 1, "pc_ref": 1, "pcm": 1, "sim": 1, "simc": 1,
}


def fsim(img, ref, prec=64, maps=("simc",)):
    """Full FSIM/FSIMc with the requested per-pixel intermediates."""
    a, b = _img(img), _img(ref)
    h, w, ch = a.shape
    _, wd, hd = dims(w, h)
    out = _FsimOut()
    keep = {}
    for name in maps:
        planes = _MAP_SHAPES[name]
        arr = np.empty((planes, hd, wd), np.float64) if planes > 1 else np.empty((hd, wd), np.float64)
        keep[name] = arr
        setattr(out, name, arr.ctypes.data)
    rc = lib().orc_fsim(prec, a.ctypes.data, b.ctypes.data, w, h, ch, ctypes.byref(out))
    if rc != 0:
        raise ValueError(f"orc_fsim failed ({rc})")
    return FsimResult(
        fsim=out.fsim, fsimc=out.fsimc, F=out.F, wd=out.wd, hd=out.hd,
        em_n=np.array(out.em_n[:]), est=np.array(out.est[:]), median=np.array(out.median[:]),
        noise_power=np.array(out.noise_power[:]), T=np.array(out.T[:]),
        sums=np.array(out.sums[:]), maps=keep)


def fsim_batch(pairs, prec=32):
    """(fsim, fsimc) of same-size pairs, one pair per OpenMP thread (the CPU
    baseline's throughput form); pairs: list of (img, ref) uint8 arrays."""
    imgs = [_img(a) for a, _ in pairs]
    refs = [_img(b) for _, b in pairs]
    h, w, ch = imgs[0].shape
    n = len(pairs)
    L = lib()
    L.orc_fsim_batch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    pi = (ctypes.c_void_p * n)(*[a.ctypes.data for a in imgs])
    pr = (ctypes.c_void_p * n)(*[b.ctypes.data for b in refs])
    outs = (_FsimOut * n)()
    if L.orc_fsim_batch(prec, n, pi, pr, w, h, ch, outs) != 0:
        raise ValueError("orc_fsim_batch failed")
    return np.array([[o.fsim, o.fsimc] for o in outs])


def fsim_filters(wd, hd, prec=64):
    lp = np.empty((hd, wd), np.float64)
    lg = np.empty((4, hd, wd), np.float64)
    an = np.empty((4, hd, wd), np.float64)
    assert lib().orc_fsim_filters(prec, wd, hd, _dp(lp), _dp(lg), _dp(an)) == 0
    return lp, lg, an


def fsim_noise_consts(wd, hd, prec=64):
    em = np.empty(4, np.float64)
    est = np.empty(8, np.float64)
    assert lib().orc_fsim_noise_consts(prec, wd, hd, _dp(em), _dp(est)) == 0
    return em, est
