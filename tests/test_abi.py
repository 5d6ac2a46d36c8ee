"""C-ABI checks that need no GPU: the library loads and exports every entry
point include/iqm/iqm.h declares; host-side validation answers without a
device."""
import ctypes
import os
import re

import iqm_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "iqm", "iqm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(iqm_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert header_symbols() == sorted(iqm_amd.EXPORTS)


def test_library_exports_every_symbol():
    L = iqm_amd.lib()
    for name in header_symbols():
        assert hasattr(L, name), name


def test_fsim_dims_host_only():
    assert iqm_amd.fsim_dims(1920, 1080) == (4, 480, 270)
    assert iqm_amd.fsim_dims(3840, 2160) == (8, 480, 270)
    assert iqm_amd.fsim_dims(384, 700) == (2, 192, 350)


def test_default_params():
    p = iqm_amd.SsimParams()
    iqm_amd.lib().iqm_ssim_default_params(ctypes.byref(p))
    assert (p.kernel_size, round(p.k1, 6), round(p.k2, 6), p.sigma) == (11, 0.01, 0.03, 1.5)


def test_null_context_is_rejected():
    L = iqm_amd.lib()
    assert L.iqm_sync(None) == iqm_amd.IQM_ERR_INVALID
    assert b"null context" in L.iqm_last_error()
    assert L.iqm_create(0, None) == iqm_amd.IQM_ERR_INVALID
