"""MI355X-native SSIM and FSIM/FSIMc for the cadik/IQM method interface.

This module is the ctypes binding of libiqm_hip.so (include/iqm/iqm.h).  It
mirrors the reference's per-method interface -- ``SSIM().compute_metric(ctx,
image, ref)`` and ``FSIM().compute_metric(ctx, image, ref)``, after
src/gpu/ssim/ssim.h:25-35 and src/gpu/fsim/fsim.h:33-36 -- and raises the
reference's error messages.  There is no CPU fallback: when the HIP library
is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libiqm_hip.so")

IQM_OK = 0
IQM_ERR_INVALID = 1
IQM_ERR_SIZE_MISMATCH = 2
IQM_ERR_TOO_SMALL = 3
IQM_ERR_UNSUPPORTED = 4
IQM_ERR_DEVICE = 5
IQM_ERR_NOMEM = 6

# every symbol include/iqm/iqm.h declares
EXPORTS = (
    "iqm_create", "iqm_destroy", "iqm_last_error", "iqm_device_name", "iqm_ssim_default_params",
    "iqm_ssim", "iqm_ssim_batch_dev", "iqm_fsim_dims", "iqm_fsim", "iqm_fsim_batch_dev",
    "iqm_fsim_batch_dev_async",
    "iqm_fft2d", "iqm_dev_alloc", "iqm_dev_free", "iqm_memcpy_h2d", "iqm_memcpy_d2h",
    "iqm_stream", "iqm_sync", "iqm_set_graphs", "iqm_autotune_fsim", "iqm_autotune_lines", "iqm_tune_export",
    "iqm_tune_import", "iqm_set_stage_timing", "iqm_stage_times", "iqm_reset_stage_times",
    "iqm_sur_param_count", "iqm_sur_create", "iqm_sur_destroy", "iqm_sur_get_params",
    "iqm_sur_set_params", "iqm_sur_get_grad", "iqm_sur_state", "iqm_sur_make_patches", "iqm_sur_make_patches_scaled",
    "iqm_sur_loss_grad", "iqm_sur_train_step", "iqm_sur_infer", "iqm_sur_save", "iqm_sur_load",
    "iqm_fsim_surrogate_batch_dev", "iqm_synth_pair_dev", "iqm_synth_distort_dev", "iqm_memcpy2d_dev",
    "iqm_batch_dev", "iqm_synth_reference_dev", "iqm_ssim_batch_dev_noalloc",
    "iqm_graph_cache_sizes", "iqm_fsim_batch_host", "iqm_host_alloc", "iqm_host_free",
    "iqm_set_marks", "iqm_marks", "iqm_sur_set_failure_memory", "iqm_sur_set_loss_weight",
)

JOB_SSIM = 1
JOB_FSIM = 2


class IqmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class Image(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("channels", ctypes.c_int32), ("row_stride", ctypes.c_int32)]


class SsimParams(ctypes.Structure):
    _fields_ = [("kernel_size", ctypes.c_int32), ("k1", ctypes.c_float), ("k2", ctypes.c_float),
                ("sigma", ctypes.c_float)]


class FsimMaps(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("pc_in", "pc_ref", "pcm", "sim", "simc")]


class PairJob(ctypes.Structure):
    _fields_ = [("d_img", ctypes.c_void_p), ("d_ref", ctypes.c_void_p), ("w", ctypes.c_int32),
                ("h", ctypes.c_int32), ("ch", ctypes.c_int32), ("methods", ctypes.c_int32),
                ("d_ssim_map", ctypes.c_void_p), ("d_simc_map", ctypes.c_void_p)]


class PairResult(ctypes.Structure):
    _fields_ = [("mssim", ctypes.c_double), ("fsim", ctypes.c_double), ("fsimc", ctypes.c_double)]


class FsimDetail(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("yiq_in", "yiq_ref", "grad_in", "grad_ref", "an_e", "keys")] + [
        ("em_n", ctypes.c_double * 4), ("est", ctypes.c_double * 8), ("median", ctypes.c_double * 8),
        ("noise_power", ctypes.c_double * 8), ("T", ctypes.c_double * 8), ("sums", ctypes.c_double * 3)]


_lib = None


def lib():
    """Loads libiqm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C iqm_amd/csrc` "
                          "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)
    L.iqm_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.iqm_destroy.argtypes = [vp]
    L.iqm_destroy.restype = None
    L.iqm_last_error.restype = ctypes.c_char_p
    L.iqm_device_name.argtypes = [vp]
    L.iqm_device_name.restype = ctypes.c_char_p
    L.iqm_ssim_default_params.argtypes = [ctypes.POINTER(SsimParams)]
    L.iqm_ssim_default_params.restype = None
    L.iqm_ssim.argtypes = [vp, ctypes.POINTER(Image), ctypes.POINTER(Image), ctypes.POINTER(SsimParams), vp, dp]
    L.iqm_ssim_batch_dev.argtypes = [vp, vp, vp, i32, i32, i32, i64, i32, ctypes.POINTER(SsimParams), vp, dp]
    L.iqm_ssim_batch_dev_noalloc.argtypes = [vp, vp, vp, i32, i32, i32, i64, i32, ctypes.POINTER(SsimParams),
                                             vp, vp]
    L.iqm_fsim_dims.argtypes = [i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.iqm_fsim_dims.restype = None
    L.iqm_fsim.argtypes = [vp, ctypes.POINTER(Image), ctypes.POINTER(Image), dp, dp,
                           ctypes.POINTER(FsimMaps), ctypes.POINTER(FsimDetail)]
    L.iqm_fsim_batch_dev.argtypes = [vp, vp, vp, i32, i32, i32, i64, i32, vp, dp, dp]
    L.iqm_fsim_batch_dev_async.argtypes = [vp, vp, vp, i32, i32, i32, i64, i32, vp, vp]
    L.iqm_fsim_batch_host.argtypes = [vp, vp, vp, i32, i32, i32, i64, i32, i32, dp, dp]
    L.iqm_host_alloc.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(vp)]
    L.iqm_host_free.argtypes = [vp, vp]
    L.iqm_fft2d.argtypes = [vp, vp, i32, i32, i32, i32, i32]
    L.iqm_dev_alloc.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(vp)]
    L.iqm_dev_free.argtypes = [vp, vp]
    L.iqm_memcpy_h2d.argtypes = [vp, vp, vp, ctypes.c_size_t]
    L.iqm_memcpy_d2h.argtypes = [vp, vp, vp, ctypes.c_size_t]
    L.iqm_stream.argtypes = [vp]
    L.iqm_stream.restype = vp
    L.iqm_sync.argtypes = [vp]
    L.iqm_set_graphs.argtypes = [vp, i32]
    L.iqm_graph_cache_sizes.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.iqm_autotune_fsim.argtypes = [vp, i32, i32, i32, i32, dp]
    L.iqm_autotune_lines.argtypes = [vp, ctypes.POINTER(PairJob), i32, ctypes.POINTER(i32)]
    L.iqm_tune_export.argtypes = [vp, ctypes.c_char_p, i64]
    L.iqm_tune_export.restype = i64
    L.iqm_tune_import.argtypes = [vp, ctypes.c_char_p]
    L.iqm_set_stage_timing.argtypes = [vp, i32]
    L.iqm_set_marks.argtypes = [vp, i32]
    L.iqm_marks.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(i64)]
    L.iqm_marks.restype = i32
    L.iqm_stage_times.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_char_p), dp, ctypes.POINTER(i64)]
    L.iqm_stage_times.restype = i32
    L.iqm_reset_stage_times.argtypes = [vp]
    L.iqm_reset_stage_times.restype = None
    u64 = ctypes.c_uint64
    fp = ctypes.POINTER(ctypes.c_float)
    L.iqm_sur_param_count.argtypes = []
    L.iqm_sur_param_count.restype = i32
    L.iqm_sur_create.argtypes = [vp, u64, ctypes.POINTER(vp)]
    L.iqm_sur_destroy.argtypes = [vp]
    L.iqm_sur_destroy.restype = None
    L.iqm_sur_get_params.argtypes = [vp, fp]
    L.iqm_sur_set_params.argtypes = [vp, fp]
    L.iqm_sur_get_grad.argtypes = [vp, fp]
    L.iqm_sur_state.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(u64)]
    L.iqm_sur_make_patches.argtypes = [vp, i32, i32, u64, vp, vp]
    L.iqm_sur_make_patches_scaled.argtypes = [vp, i32, i32, i32, u64, vp, vp]
    L.iqm_sur_loss_grad.argtypes = [vp, vp, vp, i32, i32, i32, dp]
    L.iqm_sur_train_step.argtypes = [vp, vp, vp, i32, i32, i32, ctypes.c_float, dp]
    L.iqm_sur_infer.argtypes = [vp, vp, vp, i32, i32, i32, i32]
    L.iqm_sur_save.argtypes = [vp, ctypes.POINTER(ctypes.c_uint8), i64]
    L.iqm_sur_save.restype = i64
    L.iqm_sur_load.argtypes = [vp, ctypes.POINTER(ctypes.c_uint8), i64]
    L.iqm_fsim_surrogate_batch_dev.argtypes = [vp, vp, vp, vp, i32, i32, i32, i64, i32, i32,
                                               ctypes.c_float, dp, dp, fp, ctypes.POINTER(i32)]
    L.iqm_sur_set_failure_memory.argtypes = [vp, i32]
    L.iqm_sur_set_loss_weight.argtypes = [vp, ctypes.c_float]
    L.iqm_batch_dev.argtypes = [vp, ctypes.POINTER(PairJob), i32, ctypes.POINTER(SsimParams),
                                ctypes.POINTER(PairResult)]
    L.iqm_synth_pair_dev.argtypes = [vp, i32, i32, u64, vp, vp]
    L.iqm_synth_distort_dev.argtypes = [vp, vp, i64, i32, i32, u64, vp]
    L.iqm_synth_reference_dev.argtypes = [vp, i32, i32, u64, vp]
    L.iqm_memcpy2d_dev.argtypes = [vp, vp, i64, vp, i64, i64, i64]
    for name in EXPORTS:
        getattr(L, name)
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != IQM_OK:
        raise IqmError(rc, lib().iqm_last_error().decode())


def _image(a: np.ndarray) -> tuple[Image, np.ndarray]:
    a = np.ascontiguousarray(a, dtype=np.uint8)
    if a.ndim == 2:
        h, w = a.shape
        ch = 1
    elif a.ndim == 3:
        h, w, ch = a.shape
    else:
        raise ValueError("image must be (h, w) or (h, w, ch)")
    return Image(a.ctypes.data, w, h, ch, w * ch), a


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fsim_dims(w: int, h: int) -> tuple[int, int, int]:
    F, wd, hd = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    lib().iqm_fsim_dims(w, h, ctypes.byref(F), ctypes.byref(wd), ctypes.byref(hd))
    return F.value, wd.value, hd.value


class Context:
    """One HIP device and stream (stands where VulkanRuntime stands in the reference)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().iqm_create(int(device), ctypes.byref(h)))
        self.handle = h

    @property
    def device_name(self) -> str:
        return lib().iqm_device_name(self.handle).decode()

    @property
    def stream(self) -> int:
        return lib().iqm_stream(self.handle) or 0

    def sync(self) -> None:
        _check(lib().iqm_sync(self.handle))

    def close(self) -> None:
        if self.handle:
            lib().iqm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # device memory for device-resident batches
    def alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        _check(lib().iqm_dev_alloc(self.handle, int(nbytes), ctypes.byref(p)))
        return p.value

    def free(self, ptr: int) -> None:
        _check(lib().iqm_dev_free(self.handle, ptr))

    def host_alloc(self, nbytes: int) -> np.ndarray:
        """Pinned host memory as a uint8 numpy array (freed by host_free)."""
        p = ctypes.c_void_p()
        _check(lib().iqm_host_alloc(self.handle, int(nbytes), ctypes.byref(p)))
        return np.ctypeslib.as_array((ctypes.c_uint8 * int(nbytes)).from_address(p.value))

    def host_free(self, a: np.ndarray) -> None:
        _check(lib().iqm_host_free(self.handle, a.ctypes.data))

    def upload(self, dst: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        _check(lib().iqm_memcpy_h2d(self.handle, dst, a.ctypes.data, a.nbytes))

    def download(self, src: int, out: np.ndarray) -> np.ndarray:
        _check(lib().iqm_memcpy_d2h(self.handle, out.ctypes.data, src, out.nbytes))
        return out

    def synth_pair(self, w: int, h: int, seed: int, d_img: int, d_ref: int) -> None:
        """Device copy of iqm_amd.synth.rgb_pair(w, h, seed) into packed RGB buffers."""
        _check(lib().iqm_synth_pair_dev(self.handle, w, h, int(seed), d_img, d_ref))

    def synth_reference(self, w: int, h: int, seed: int, d_out: int) -> None:
        """Device copy of iqm_amd.synth.rgb_reference(w, h, seed)."""
        _check(lib().iqm_synth_reference_dev(self.handle, w, h, int(seed), d_out))

    def synth_distort(self, d_src: int, row_stride: int, w: int, h: int, seed: int, d_out: int) -> None:
        _check(lib().iqm_synth_distort_dev(self.handle, d_src, row_stride, w, h, int(seed), d_out))

    def copy2d(self, dst: int, dst_pitch: int, src: int, src_pitch: int, width_bytes: int, rows: int) -> None:
        _check(lib().iqm_memcpy2d_dev(self.handle, dst, dst_pitch, src, src_pitch, width_bytes, rows))

    def batch_dev(self, jobs, ssim_params: "SsimParams | None" = None) -> np.ndarray:
        """Scores a mixed batch; jobs are dicts with d_img, d_ref, w, h, ch, methods and
        optional d_ssim_map / d_simc_map.  Returns an (n, 3) array [mssim, fsim, fsimc]."""
        n = len(jobs)
        arr = (PairJob * n)()
        for i, j in enumerate(jobs):
            arr[i] = PairJob(j["d_img"], j["d_ref"], j["w"], j["h"], j.get("ch", 3),
                             j.get("methods", JOB_SSIM | JOB_FSIM), j.get("d_ssim_map"), j.get("d_simc_map"))
        res = (PairResult * n)()
        _check(lib().iqm_batch_dev(self.handle, arr, n, ctypes.byref(ssim_params) if ssim_params else None, res))
        return np.array([[r.mssim, r.fsim, r.fsimc] for r in res], np.float64)

    def set_graphs(self, on: bool) -> None:
        _check(lib().iqm_set_graphs(self.handle, 1 if on else 0))

    def graph_cache_sizes(self) -> tuple[int, int]:
        """(same-size FSIM graphs, mixed-batch graphs) the context holds."""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().iqm_graph_cache_sizes(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune_fsim(self, w: int, h: int, ch: int = 3, n: int = 8) -> float:
        """Tunes the FSIM tiles for w x h inputs; returns the best us/pair."""
        us = np.zeros(1, np.float64)
        _check(lib().iqm_autotune_fsim(self.handle, w, h, ch, n, _dptr(us)))
        return float(us[0])

    def autotune_lines(self, jobs) -> int:
        """Tunes the FFT plan (two-level split) of every downscaled line length of the
        batch's FSIM pairs, as columns and as rows; returns the number of entries timed.
        jobs as for batch_dev."""
        n = len(jobs)
        arr = (PairJob * n)()
        for i, j in enumerate(jobs):
            arr[i] = PairJob(j.get("d_img"), j.get("d_ref"), j["w"], j["h"], j.get("ch", 3),
                             j.get("methods", JOB_SSIM | JOB_FSIM), None, None)
        cnt = ctypes.c_int32()
        _check(lib().iqm_autotune_lines(self.handle, arr, n, ctypes.byref(cnt)))
        return cnt.value

    def export_tunes(self) -> str:
        n = lib().iqm_tune_export(self.handle, None, 0)
        if n < 0:
            raise IqmError(IQM_ERR_INVALID, lib().iqm_last_error().decode())
        buf = ctypes.create_string_buffer(n + 1)
        lib().iqm_tune_export(self.handle, buf, n + 1)
        return buf.value.decode()

    def import_tunes(self, plan: str) -> None:
        _check(lib().iqm_tune_import(self.handle, plan.encode()))

    def set_marks(self, on: bool) -> None:
        """Host stage marks of iqm_ssim / iqm_fsim (src/timestamps.h)."""
        _check(lib().iqm_set_marks(self.handle, 1 if on else 0))

    def marks(self) -> list:
        """[(name, ns)] of the last iqm_ssim / iqm_fsim call."""
        n = lib().iqm_marks(self.handle, 0, None, None)
        if n <= 0:
            return []
        names = (ctypes.c_char_p * n)()
        ns = (ctypes.c_int64 * n)()
        lib().iqm_marks(self.handle, n, names, ns)
        return [(names[i].decode(), int(ns[i])) for i in range(n)]

    def set_stage_timing(self, on: bool) -> None:
        _check(lib().iqm_set_stage_timing(self.handle, 1 if on else 0))

    def reset_stage_times(self) -> None:
        lib().iqm_reset_stage_times(self.handle)

    def stage_times(self) -> dict:
        n = lib().iqm_stage_times(self.handle, 0, None, None, None)
        if n <= 0:
            return {}
        names = (ctypes.c_char_p * n)()
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int64 * n)()
        lib().iqm_stage_times(self.handle, n, names, ms, cnt)
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(n)}

    def fft2d(self, x: np.ndarray, inverse: bool = False, normalize: bool = False) -> np.ndarray:
        """Batched 2-D FFT of (..., h, w) complex data in float32."""
        a = np.ascontiguousarray(x, dtype=np.complex64).copy()
        h, w = a.shape[-2:]
        batch = int(np.prod(a.shape[:-2])) if a.ndim > 2 else 1
        _check(lib().iqm_fft2d(self.handle, a.ctypes.data, w, h, batch, 1 if inverse else -1,
                               int(normalize)))
        return a


@dataclass
class SSIMResult:
    """src/gpu/ssim/ssim.h:17-23"""
    image_data: np.ndarray | None
    width: int
    height: int
    mssim: float


@dataclass
class FSIMResult:
    """src/gpu/fsim/fsim.h:27-31, extended with the per-pixel maps."""
    fsim: float
    fsimc: float
    maps: dict = field(default_factory=dict)
    detail: dict = field(default_factory=dict)


class SSIM:
    """src/gpu/ssim/ssim.h:25-35"""

    def __init__(self, ctx: Context | None = None):
        self.kernel_size = 11
        self.k_1 = 0.01
        self.k_2 = 0.03
        self.sigma = 1.5

    def params(self) -> SsimParams:
        return SsimParams(self.kernel_size, self.k_1, self.k_2, self.sigma)

    def compute_metric(self, ctx: Context, image: np.ndarray, ref: np.ndarray,
                       want_map: bool = True) -> SSIMResult:
        a, ka = _image(image)
        b, kb = _image(ref)
        m = np.empty((a.height, a.width), np.float32) if want_map else None
        ms = np.zeros(1, np.float64)
        p = self.params()
        _check(lib().iqm_ssim(ctx.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(p),
                              m.ctypes.data if m is not None else None, _dptr(ms)))
        return SSIMResult(m, a.width, a.height, float(ms[0]))

    def compute_batch_dev(self, ctx: Context, d_img: int, d_ref: int, w: int, h: int, ch: int,
                          pair_stride: int, n: int, d_maps: int | None = None) -> np.ndarray:
        out = np.zeros(n, np.float64)
        p = self.params()
        _check(lib().iqm_ssim_batch_dev(ctx.handle, d_img, d_ref, w, h, ch, pair_stride, n,
                                        ctypes.byref(p), d_maps, _dptr(out)))
        return out

    def compute_batch_dev_noalloc(self, ctx: Context, d_img: int, d_ref: int, w: int, h: int, ch: int,
                                  pair_stride: int, n: int, d_mssim: int, d_maps: int | None = None) -> None:
        """Allocation-free form: enqueues the batch with mssim to d_mssim (n device
        doubles) and its partials in the library's static scratch; ctx.sync() waits."""
        p = self.params()
        _check(lib().iqm_ssim_batch_dev_noalloc(ctx.handle, d_img, d_ref, w, h, ch, pair_stride, n,
                                                ctypes.byref(p), d_maps, d_mssim))


class FSIM:
    """src/gpu/fsim/fsim.h:33-36"""

    MAPS = ("pc_in", "pc_ref", "pcm", "sim", "simc")

    def __init__(self, ctx: Context | None = None):
        pass

    def compute_metric(self, ctx: Context, image: np.ndarray, ref: np.ndarray, maps=(),
                       detail: bool = False) -> FSIMResult:
        a, ka = _image(image)
        b, kb = _image(ref)
        _, wd, hd = fsim_dims(a.width, a.height)
        N = wd * hd
        fs = np.zeros(1, np.float64)
        fc = np.zeros(1, np.float64)
        keep = {}
        mp = None
        if maps:
            mp = FsimMaps()
            for name in maps:
                arr = np.empty((hd, wd), np.float32)
                keep[name] = arr
                setattr(mp, name, arr.ctypes.data)
        det = None
        dk = {}
        if detail:
            det = FsimDetail()
            for name, planes in (("yiq_in", 3), ("yiq_ref", 3), ("grad_in", 1), ("grad_ref", 1),
                                 ("an_e", 10), ("keys", 8)):
                arr = np.empty((planes, hd, wd) if planes > 1 else (hd, wd), np.float32)
                dk[name] = arr
                setattr(det, name, arr.ctypes.data)
        _check(lib().iqm_fsim(ctx.handle, ctypes.byref(a), ctypes.byref(b), _dptr(fs), _dptr(fc),
                              ctypes.byref(mp) if mp is not None else None,
                              ctypes.byref(det) if det is not None else None))
        if det is not None:
            for name in ("em_n", "est", "median", "noise_power", "T", "sums"):
                dk[name] = np.array(getattr(det, name)[:])
        return FSIMResult(float(fs[0]), float(fc[0]), keep, dk)

    def compute_batch_dev(self, ctx: Context, d_img: int, d_ref: int, w: int, h: int, ch: int,
                          pair_stride: int, n: int, d_simc_maps: int | None = None):
        fs = np.zeros(n, np.float64)
        fc = np.zeros(n, np.float64)
        _check(lib().iqm_fsim_batch_dev(ctx.handle, d_img, d_ref, w, h, ch, pair_stride, n,
                                        d_simc_maps, _dptr(fs), _dptr(fc)))
        return fs, fc

    def compute_batch_host(self, ctx: Context, img: np.ndarray, ref: np.ndarray, chunk: int = 32):
        """FSIM/FSIMc of n host pairs, arrays of shape (n, h, w[, ch]) uint8 (pinned
        ones from ctx.host_alloc are copied by DMA directly); returns (fsim, fsimc)."""
        if img.shape != ref.shape or img.dtype != np.uint8 or ref.dtype != np.uint8:
            raise ValueError("img and ref must be uint8 arrays of one shape")
        if not (img.flags.c_contiguous and ref.flags.c_contiguous):
            raise ValueError("img and ref must be C-contiguous")
        n, h, w = img.shape[:3]
        ch = img.shape[3] if img.ndim == 4 else 1
        fs = np.zeros(n, np.float64)
        fc = np.zeros(n, np.float64)
        _check(lib().iqm_fsim_batch_host(ctx.handle, img.ctypes.data, ref.ctypes.data, w, h, ch, w * h * ch, n,
                                         int(chunk), _dptr(fs), _dptr(fc)))
        return fs, fc

    def compute_batch_dev_async(self, ctx: Context, d_img: int, d_ref: int, w: int, h: int, ch: int,
                                pair_stride: int, n: int, d_scores: int, d_simc_maps: int | None = None):
        """Enqueues the batch; d_scores (device, n x 2 doubles [fsim, fsimc]) is
        filled when the context's stream gets there (ctx.sync() waits)."""
        _check(lib().iqm_fsim_batch_dev_async(ctx.handle, d_img, d_ref, w, h, ch, pair_stride, n,
                                              d_simc_maps, d_scores))


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Surrogate:
    """Learned phase-congruency network (no counterpart in the reference).

    A 6-layer 3x3 convolutional network, 32 channels, from Y/255 on FSIM's
    downscaled grid to the PC map; BF16 on the matrix cores, trained on the
    device with Adam.  Parameters are one flat FP32 vector in PyTorch OIHW
    order: w1[32,1,3,3], b1[32], 4 x (w[32,32,3,3], b[32]), w6[1,32,3,3], b6[1].
    """

    def __init__(self, ctx: Context, seed: int = 1):
        h = ctypes.c_void_p()
        _check(lib().iqm_sur_create(ctx.handle, int(seed), ctypes.byref(h)))
        self.ctx = ctx
        self.handle = h

    @staticmethod
    def param_count() -> int:
        return int(lib().iqm_sur_param_count())

    def close(self) -> None:
        if self.handle:
            lib().iqm_sur_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(self.param_count(), np.float32)
        _check(lib().iqm_sur_get_params(self.handle, _fptr(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.param_count():
            raise ValueError("parameter vector has the wrong length")
        _check(lib().iqm_sur_set_params(self.handle, _fptr(p)))

    def grad(self) -> np.ndarray:
        g = np.zeros(self.param_count(), np.float32)
        _check(lib().iqm_sur_get_grad(self.handle, _fptr(g)))
        return g

    def state(self) -> tuple[int, int]:
        st, cur = ctypes.c_int64(), ctypes.c_uint64()
        _check(lib().iqm_sur_state(self.handle, ctypes.byref(st), ctypes.byref(cur)))
        return st.value, cur.value

    def make_patches(self, n: int, size: int, seed: int, d_x: int, d_pc: int, scale: int = 1) -> None:
        if scale == 1:
            _check(lib().iqm_sur_make_patches(self.handle, n, size, int(seed), d_x, d_pc))
        else:
            _check(lib().iqm_sur_make_patches_scaled(self.handle, n, size, scale, int(seed), d_x, d_pc))

    def loss_grad(self, d_x: int, d_t: int, n: int, h: int, w: int) -> float:
        loss = np.zeros(1, np.float64)
        _check(lib().iqm_sur_loss_grad(self.handle, d_x, d_t, n, h, w, _dptr(loss)))
        return float(loss[0])

    def train_step(self, d_x: int, d_t: int, n: int, h: int, w: int, lr: float) -> float:
        loss = np.zeros(1, np.float64)
        _check(lib().iqm_sur_train_step(self.handle, d_x, d_t, n, h, w, float(lr), _dptr(loss)))
        return float(loss[0])

    def infer(self, d_x: int, d_out: int, n: int, h: int, w: int, fused: bool = False) -> None:
        """Layer-by-layer forward (the training kernels), or with fused=True the
        single all-layer kernel the FSIM path uses."""
        _check(lib().iqm_sur_infer(self.handle, d_x, d_out, n, h, w, 1 if fused else 0))

    def save(self) -> bytes:
        n = int(lib().iqm_sur_save(self.handle, None, 0))
        if n < 0:
            raise IqmError(IQM_ERR_INVALID, lib().iqm_last_error().decode())
        buf = (ctypes.c_uint8 * n)()
        if lib().iqm_sur_save(self.handle, buf, n) != n:
            raise IqmError(IQM_ERR_INVALID, lib().iqm_last_error().decode())
        return bytes(buf)

    def load(self, blob: bytes) -> None:
        buf = (ctypes.c_uint8 * len(blob)).from_buffer_copy(blob)
        _check(lib().iqm_sur_load(self.handle, buf, len(blob)))

    def set_loss_weight(self, eps: float | None) -> None:
        """Training loss: None (or eps < 0) mean squared error; eps >= 0 each
        pixel's squared error weighted by max(PC_a, PC_b) + eps of its FSIM
        pair (images 2i, 2i+1), as FSIM weights the pixel."""
        _check(lib().iqm_sur_set_loss_weight(self.handle, -1.0 if eps is None else float(eps)))

    def set_failure_memory(self, on: bool) -> None:
        """Whether a size whose spot check failed goes straight to exact PC
        afterwards (default on); setting it clears the memory."""
        _check(lib().iqm_sur_set_failure_memory(self.handle, 1 if on else 0))

    def fsim_batch_dev(self, d_img: int, d_ref: int, w: int, h: int, ch: int, pair_stride: int,
                       n: int, spot_every: int = 0, spot_tol: float = 0.02):
        """FSIM/FSIMc with PC from the network; returns (fsim, fsimc, spot_err, fell_back)."""
        fs = np.zeros(n, np.float64)
        fc = np.zeros(n, np.float64)
        err = np.zeros(1, np.float32)
        back = ctypes.c_int32()
        _check(lib().iqm_fsim_surrogate_batch_dev(self.ctx.handle, self.handle, d_img, d_ref, w, h, ch,
                                                  pair_stride, n, spot_every, float(spot_tol),
                                                  _dptr(fs), _dptr(fc), _fptr(err), ctypes.byref(back)))
        return fs, fc, float(err[0]), bool(back.value)
