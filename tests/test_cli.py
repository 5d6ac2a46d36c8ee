"""The iqm command line: the reference's argument parsing and messages
(src/args.cpp:8-61, src/main.cpp:37-57, 264-293) without a GPU, and the
metric lines on a GPU."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "iqm_amd", "bin", "iqm")


def run(*args):
    p = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout, p.stderr


def write_png(path, img):
    """Minimal PNG encoder (8-bit gray, RGB or RGBA) for the fixtures."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    ctype = {1: 0, 3: 2, 4: 6}[ch]
    raw = b"".join(b"\x00" + img[y].tobytes() for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
    data += chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b"")
    open(path, "wb").write(data)


def read_png_gray(path):
    data = open(path, "rb").read()
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        t = data[pos + 4:pos + 8]
        d = data[pos + 8:pos + 8 + n]
        if t == b"IHDR":
            w, h = struct.unpack(">II", d[:8])
            assert d[8] == 8 and d[9] == 0
        elif t == b"IDAT":
            idat += d
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = [np.frombuffer(raw[y * (w + 1) + 1:(y + 1) * (w + 1)], np.uint8) for y in range(h)]
    assert all(raw[y * (w + 1)] == 0 for y in range(h))
    return np.stack(rows)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CLI):
        pytest.skip("CLI not built (make -C iqm_amd/csrc)")


def test_unknown_method():
    rc, out, err = run("--method", "SSIM_GPU", "--input", "a", "--ref", "b")
    assert rc != 0 and err.strip() == "Unknown method"


@pytest.mark.parametrize("argv,msg", [
    (["--input", "a", "--ref", "b"], "missing method"),
    (["--method", "SSIM", "--ref", "b"], "missing input"),
    (["--method", "FSIM", "--input", "a"], "missing reference"),
])
def test_missing_arguments(argv, msg):
    rc, out, err = run(*argv)
    assert rc != 0 and err.strip() == msg


def test_unreadable_image(tmp_path):
    rc, out, err = run("--method", "SSIM", "--input", str(tmp_path / "nope.png"), "--ref", "x")
    assert rc != 0 and err.startswith(f"Failed to load image '{tmp_path / 'nope.png'}'")


def test_size_mismatch_before_device(tmp_path):
    write_png(tmp_path / "a.png", np.zeros((10, 12, 3), np.uint8))
    write_png(tmp_path / "b.png", np.zeros((10, 13, 3), np.uint8))
    rc, out, err = run("--method", "FSIM", "--input", str(tmp_path / "a.png"), "--ref", str(tmp_path / "b.png"))
    assert rc != 0 and err.strip() == "Compared images must have the same size"


def test_methods_outside_this_build(tmp_path):
    write_png(tmp_path / "a.png", np.zeros((8, 8), np.uint8))
    rc, out, err = run("--method", "FLIP", "--input", str(tmp_path / "a.png"), "--ref", str(tmp_path / "a.png"))
    assert rc != 0 and "FLIP support was not compiled" in err
    rc, out, err = run("--method", "CW_SSIM_CPU", "--input", "a", "--ref", "b")
    assert rc == 0 and out == ""


@pytest.mark.gpu
def test_cli_metrics_match_oracle(tmp_path):
    """The CLI's stdout lines and map PNG against the FP64 oracle on the same
    decoded pixels (the CLI prints floats with 6 significant digits)."""
    from iqm_amd import synth
    from oracle import oracle

    img, ref = synth.rgb_pair(400, 300, seed=33)
    write_png(tmp_path / "a.png", img)
    write_png(tmp_path / "b.png", synth.rgba(ref))
    rc, out, err = run("--method", "SSIM", "--input", str(tmp_path / "a.png"), "--ref",
                       str(tmp_path / "b.png"), "--output", str(tmp_path / "m.png"))
    assert rc == 0, err
    ms, smap = oracle.ssim(img, ref, prec=64)
    assert out.startswith("MSSIM: ") and abs(float(out.split()[1]) - ms) < 1e-5
    m = read_png_gray(tmp_path / "m.png")
    expect = np.clip(smap, 0, 1) * 255.0
    # the map is truncated to u8; the FP32 map is within 1e-4 of FP64, so the
    # byte can differ from the FP64 truncation only where it sits on an edge
    d = m.astype(np.float64) - np.floor(expect)
    assert m.shape == (300, 400) and np.all(np.abs(d) <= 1)
    assert np.mean(d != 0) < 1e-3
    rc, out, err = run("--method", "FSIM", "--input", str(tmp_path / "a.png"), "--ref", str(tmp_path / "b.png"),
                       "--output", str(tmp_path / "f.png"))
    assert rc == 0, err
    o = oracle.fsim(img, synth.rgba(ref)[:, :, :3], prec=64, maps=("simc",))
    lines = out.strip().splitlines()
    assert lines[0].startswith("FSIM: ") and lines[1].startswith("FSIMc: ")
    assert abs(float(lines[0].split()[1]) - o.fsim) < 1e-5 and abs(float(lines[1].split()[1]) - o.fsimc) < 1e-5
    fm = read_png_gray(tmp_path / "f.png")
    de = fm.astype(np.float64) - np.floor(np.clip(o.maps["simc"], 0, 1) * 255.0)
    assert fm.shape == o.maps["simc"].shape and np.all(np.abs(de) <= 1)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libstb_ref.so")),
                    reason="reference image loader not built")
def test_cli_jpeg_input_matches_oracle(tmp_path):
    """A JPEG pair (4:2:0 and progressive): the CLI decodes it as the reference's
    stbi_load does (oracle/_ref), and its scores equal the FP64 oracle's on
    those pixels."""
    from PIL import Image

    import test_image_decode as tid
    from iqm_amd import synth
    from oracle import oracle

    img, ref = synth.rgb_pair(320, 240, seed=36)
    Image.fromarray(img).save(tmp_path / "a.jpg", quality=80, subsampling="4:2:0")
    Image.fromarray(ref).save(tmp_path / "b.jpg", quality=92, progressive=True)
    a = tid.stb_load(tmp_path / "a.jpg")[:, :, :3]
    b = tid.stb_load(tmp_path / "b.jpg")[:, :, :3]
    rc, out, err = run("--method", "FSIM", "--input", str(tmp_path / "a.jpg"), "--ref", str(tmp_path / "b.jpg"))
    assert rc == 0, err
    o = oracle.fsim(a, b, prec=64)
    lines = out.strip().splitlines()
    assert abs(float(lines[0].split()[1]) - o.fsim) < 1e-5 and abs(float(lines[1].split()[1]) - o.fsimc) < 1e-5
    rc, out, err = run("--method", "SSIM", "--input", str(tmp_path / "a.jpg"), "--ref", str(tmp_path / "b.jpg"))
    assert rc == 0, err
    ms, _ = oracle.ssim(a, b, prec=64)
    assert abs(float(out.split()[1]) - ms) < 1e-5


@pytest.mark.gpu
def test_cli_verbose_timestamps_and_repeat(tmp_path):
    """-v prints the reference's Timestamps table (src/timestamps.h:18-51) with
    the stage names of src/gpu/fsim/fsim.cpp:85-155 and ssim.cpp:78-235;
    --repeat N prints the mean time of N evaluations."""
    from iqm_amd import synth

    img, ref = synth.rgb_pair(320, 240, seed=34)
    write_png(tmp_path / "a.png", img)
    write_png(tmp_path / "b.png", ref)
    rc, out, err = run("--method", "FSIM", "--input", str(tmp_path / "a.png"), "--ref", str(tmp_path / "b.png"), "-v")
    assert rc == 0, err
    names = [ln.split(":")[0].strip() for ln in out.splitlines() if "us |" in ln]
    assert names == ["downscale factor computed", "images sent to gpu", "FFT library initialized",
                     "pre median work recorded", "noise powers computed", "FSIM, FSIMc computed"], out
    assert any(ln.strip().startswith("TOTAL:") for ln in out.splitlines())
    rc, out, err = run("--method", "SSIM", "--input", str(tmp_path / "a.png"), "--ref", str(tmp_path / "b.png"), "-v",
                       "--repeat", "3")
    assert rc == 0, err
    names = [ln.split(":")[0].strip() for ln in out.splitlines() if "us |" in ln]
    assert names == ["start GPU pipeline", "end GPU pipeline", "end copy from GPU", "end MSSIM compute"], out
    assert "mean time over 3 runs" in out
