"""The CLI's image decoders (iqm_amd/cli/image_io.hpp) against the loader the
reference calls, stbi_load(..., 4) (src/main.cpp:37-57), on the CPU.

Checkers:
  * oracle/_ref/libstb_ref.so -- the reference's own stb_image 2.30 compiled
    from its source under the reference tree (oracle/Makefile); every file
    must decode to the same RGBA bytes (skipped where that build is absent);
  * for the lossless formats, the pixels the files were written from (PNG,
    BMP, TGA, PNM) or Pillow's decoder (GIF), so the comparison does not rest
    on stb alone;
  * for JPEG, Pillow's libjpeg decode within a few levels (a different IDCT
    and upsampler; this catches gross errors such as a wrong colour transform
    or block order).
Files are written here by Pillow and by a small PNG encoder below (Adam7
interlace, every filter type, bit depths 1-16, tRNS), with seeded contents.
"""
import ctypes
import io
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest
from PIL import Image

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STB = os.path.join(ROOT, "oracle", "_ref", "libstb_ref.so")


@pytest.fixture(scope="module")
def decoder(tmp_path_factory):
    exe = tmp_path_factory.mktemp("dec") / "decode_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "iqm_amd", "cli"), "-o", str(exe),
                    os.path.join(ROOT, "tests", "native", "decode_check.cpp"), "-lz"], check=True, timeout=300)
    return str(exe)


def decode_many(decoder, paths):
    p = subprocess.run([decoder, *map(str, paths)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    errs = {ln.split(":")[0][4:]: ln.split(": ", 1)[1] for ln in p.stdout.splitlines() if ln.startswith("ERR ")}
    out = {}
    for q in paths:
        if str(q) in errs:
            out[str(q)] = errs[str(q)]
            continue
        raw = open(str(q) + ".rgba", "rb").read()
        w, h = struct.unpack("<II", raw[:8])
        out[str(q)] = np.frombuffer(raw[8:], np.uint8).reshape(h, w, 4)
    return out


_stb = None


def stb_load(path):
    global _stb
    if _stb is None:
        _stb = ctypes.CDLL(STB)
        _stb.stbi_load_from_memory.restype = ctypes.POINTER(ctypes.c_uint8)
        _stb.stbi_load_from_memory.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                               ctypes.c_int]
        _stb.stbi_failure_reason.restype = ctypes.c_char_p
        _stb.stbi_image_free.argtypes = [ctypes.c_void_p]
    data = open(path, "rb").read()
    x, y, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    p = _stb.stbi_load_from_memory(data, len(data), ctypes.byref(x), ctypes.byref(y), ctypes.byref(n), 4)
    if not p:
        return _stb.stbi_failure_reason().decode()
    a = np.ctypeslib.as_array(p, shape=(y.value, x.value, 4)).copy()
    _stb.stbi_image_free(p)
    return a


# ---- a PNG encoder with the features Pillow does not write -----------------

def _chunk(t, d):
    return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)


def _pack_row(samples, depth):
    if depth == 16:
        return samples.astype(">u2").tobytes()
    if depth == 8:
        return samples.astype(np.uint8).tobytes()
    per = 8 // depth
    s = np.concatenate([samples.astype(np.uint8), np.zeros((-len(samples)) % per, np.uint8)]).reshape(-1, per)
    b = np.zeros(len(s), np.uint16)
    for k in range(per):
        b = (b << depth) | s[:, k]
    return b.astype(np.uint8).tobytes()


def _filter(row, prev, bpp, ft):
    row = np.frombuffer(row, np.uint8).astype(np.int32)
    prev = np.frombuffer(prev, np.uint8).astype(np.int32)
    a = np.concatenate([np.zeros(bpp, np.int32), row[:-bpp]]) if len(row) > bpp else np.zeros_like(row)
    c = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]]) if len(prev) > bpp else np.zeros_like(prev)
    if ft == 0:
        out = row
    elif ft == 1:
        out = row - a
    elif ft == 2:
        out = row - prev
    elif ft == 3:
        out = row - ((a + prev) >> 1)
    else:
        p = a + prev - c
        pa, pb, pc = np.abs(p - a), np.abs(p - prev), np.abs(p - c)
        pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, prev, c))
        out = row - pred
    return bytes([ft]) + (out & 255).astype(np.uint8).tobytes()


def encode_png(samples, ctype, depth, interlace=False, plte=None, trns=None, seed=0):
    """samples: (h, w, comps) integer array of the raw sample values."""
    rng = np.random.default_rng(seed)
    h, w, comps = samples.shape
    bpp = max(1, comps * depth // 8)
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)] \
        if interlace else [(0, 0, 1, 1)]
    raw = b""
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        prev = bytes(len(_pack_row(sub[0].reshape(-1), depth)))
        for r in sub:
            row = _pack_row(r.reshape(-1), depth)
            raw += _filter(row, prev, bpp, int(rng.integers(0, 5)))
            prev = row
    data = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    if plte is not None:
        data += _chunk(b"PLTE", plte.astype(np.uint8).tobytes())
    if trns is not None:
        data += _chunk(b"tRNS", trns)
    data += _chunk(b"IDAT", zlib.compress(raw)) + _chunk(b"IEND", b"")
    return data


def _expected_png(samples, ctype, depth, plte=None, trns=None):
    """RGBA the reference's loader gives for these samples."""
    h, w, comps = samples.shape
    s = samples.astype(np.int64)
    if ctype == 3:
        pal = plte.reshape(-1, 3)
        rgb = pal[s[:, :, 0]]
        a = np.full((h, w), 255)
        if trns is not None:
            t = np.frombuffer(trns, np.uint8)
            idx = s[:, :, 0]
            a = np.where(idx < len(t), t[np.minimum(idx, len(t) - 1)], 255)
        return np.dstack([rgb, a]).astype(np.uint8)
    if depth == 16:
        s8 = s >> 8
    elif depth == 8:
        s8 = s
    else:
        scale = {1: 0xFF, 2: 0x55, 4: 0x11}[depth] if ctype == 0 else 1
        s8 = s * scale
    if ctype in (0, 4):
        g = s8[:, :, 0]
        rgb = np.dstack([g, g, g])
        a = s8[:, :, 1] if ctype == 4 else np.full((h, w), 255)
    else:
        rgb = s8[:, :, :3]
        a = s8[:, :, 3] if ctype == 6 else np.full((h, w), 255)
    if trns is not None:
        key = np.frombuffer(trns, ">u2").astype(np.int64)
        if depth == 16:
            hit = np.all(s[:, :, :comps] == key[:comps], axis=2)
        else:
            sc = {1: 0xFF, 2: 0x55, 4: 0x11, 8: 1}[depth]
            hit = np.all(s8[:, :, :comps] == ((key[:comps] & 255) * sc) % 256, axis=2)
        a = np.where(hit, 0, 255)
    return np.dstack([rgb, a]).astype(np.uint8)


def png_cases(tmp):
    rng = np.random.default_rng(1234)
    cases = []
    sizes = [(37, 23), (1, 1), (9, 2), (64, 17)]
    for ctype, depths in ((0, (1, 2, 4, 8, 16)), (2, (8, 16)), (3, (1, 2, 4, 8)), (4, (8, 16)), (6, (8, 16))):
        comps = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
        for depth in depths:
            for interlace in (False, True):
                for (w, h) in sizes[: 2 if depth == 16 else 4]:
                    hi = (1 << depth) if ctype != 3 else min(1 << depth, 200)
                    smp = rng.integers(0, hi, (h, w, comps))
                    plte = trns = None
                    if ctype == 3:
                        plte = rng.integers(0, 256, (hi, 3))
                        if rng.random() < 0.5:
                            trns = rng.integers(0, 256, int(rng.integers(1, hi + 1))).astype(np.uint8).tobytes()
                    elif ctype in (0, 2) and w * h > 1:
                        key = smp[0, 0] if rng.random() < 0.7 else rng.integers(0, 1 << depth, comps)
                        trns = np.array(key, ">u2").tobytes()
                    name = tmp / f"png_c{ctype}_d{depth}_{'i' if interlace else 'n'}_{w}x{h}.png"
                    name.write_bytes(encode_png(smp, ctype, depth, interlace, plte, trns, seed=w * h + depth))
                    cases.append((name, _expected_png(smp, ctype, depth, plte, trns)))
    return cases


def pillow_cases(tmp):
    """(path, expected RGBA or None, kind) for files Pillow writes."""
    rng = np.random.default_rng(99)
    out = []
    rgb = rng.integers(0, 256, (45, 61, 3), dtype=np.uint8)
    smooth = (np.add.outer(np.arange(45), np.arange(61)) * 2 % 256).astype(np.uint8)
    photo = np.dstack([smooth, np.roll(smooth, 7, 1), 255 - smooth])
    rgba = np.dstack([rgb, rng.integers(0, 256, (45, 61), dtype=np.uint8)])
    gray = rng.integers(0, 256, (45, 61), dtype=np.uint8)

    def save(name, im, fmt, expect, kind, **kw):
        p = tmp / name
        im.save(p, fmt, **kw)
        out.append((p, expect, kind))

    a255 = np.full((45, 61, 1), 255, np.uint8)
    # BMP: 24-bit, 32-bit with alpha, 8-bit palette (and gray as palette), 1-bit
    save("rgb.bmp", Image.fromarray(rgb), "BMP", np.dstack([rgb, a255]), "lossless")
    save("rgba.bmp", Image.fromarray(rgba), "BMP", rgba, "lossless")
    save("gray.bmp", Image.fromarray(gray), "BMP", np.dstack([gray, gray, gray, a255]), "lossless")
    pim = Image.fromarray(rgb).quantize(64)
    save("pal.bmp", pim, "BMP", np.asarray(pim.convert("RGBA")), "lossless")
    bw = Image.fromarray(gray).convert("1")
    save("bw.bmp", bw, "BMP", np.asarray(bw.convert("RGBA")), "lossless")
    # TGA: plain and RLE; RGB, RGBA, gray, gray+alpha, palette; both origins
    for rle in (None, "tga_rle"):
        kw = {"compression": rle} if rle else {}
        tag = "rle" if rle else "raw"
        save(f"rgb_{tag}.tga", Image.fromarray(rgb), "TGA", np.dstack([rgb, a255]), "lossless", **kw)
        save(f"rgba_{tag}.tga", Image.fromarray(rgba), "TGA", rgba, "lossless", **kw)
        save(f"gray_{tag}.tga", Image.fromarray(gray), "TGA", np.dstack([gray, gray, gray, a255]), "lossless", **kw)
        la = Image.fromarray(np.dstack([gray, rgb[:, :, 0]]), "LA")
        save(f"la_{tag}.tga", la, "TGA", np.asarray(la.convert("RGBA")), "lossless", **kw)
        save(f"pal_{tag}.tga", pim, "TGA", np.asarray(pim.convert("RGBA")), "lossless", **kw)
        save(f"rgb_{tag}_tl.tga", Image.fromarray(photo), "TGA", np.dstack([photo, a255]), "lossless",
             orientation=1, **kw)
    # GIF: palette image, interlaced or not, with a transparent index
    for inter in (True, False):
        save(f"pal_{int(inter)}.gif", pim, "GIF", None, "gif", interlace=inter)
    tim = pim.copy()
    tim.info["transparency"] = 3
    save("pal_t.gif", tim, "GIF", None, "gif", transparency=3)
    # PNM: binary gray and RGB, 8 and 16 bit
    save("gray.pgm", Image.fromarray(gray), "PPM", np.dstack([gray, gray, gray, a255]), "lossless")
    save("rgb.ppm", Image.fromarray(rgb), "PPM", np.dstack([rgb, a255]), "lossless")
    g16 = rng.integers(0, 65536, (45, 61)).astype(np.uint16)
    lo = (g16 & 255).astype(np.uint8)  # the byte the reference keeps
    save("gray16.pgm", Image.fromarray(g16, "I;16"), "PPM", np.dstack([lo, lo, lo, a255]), "lossless")
    # JPEG: sampling 4:4:4 / 4:2:2 / 4:2:0 / 4:1:1, gray, progressive, restarts,
    # odd sizes, CMYK
    for size in ((61, 45), (33, 17), (8, 8), (1, 1), (130, 67)):
        im = Image.fromarray(photo).resize(size)
        for ss in ("4:4:4", "4:2:2", "4:2:0", "4:1:1"):
            for prog in (False, True):
                save(f"j_{size[0]}x{size[1]}_{ss.replace(':', '')}_{int(prog)}.jpg", im, "JPEG",
                     np.asarray(im.convert("RGBA")), "jpeg", subsampling=ss, progressive=prog, quality=85)
        save(f"j_{size[0]}x{size[1]}_gray.jpg", im.convert("L"), "JPEG", np.asarray(im.convert("L").convert("RGBA")),
             "jpeg", quality=90)
        save(f"j_{size[0]}x{size[1]}_rst.jpg", im, "JPEG", np.asarray(im.convert("RGBA")), "jpeg",
             restart_marker_blocks=3, subsampling="4:2:0")
        save(f"j_{size[0]}x{size[1]}_rstp.jpg", im, "JPEG", np.asarray(im.convert("RGBA")), "jpeg",
             restart_marker_rows=1, progressive=True)
    cm = Image.fromarray(photo).convert("CMYK")
    save("cmyk.jpg", cm, "JPEG", np.asarray(cm.convert("RGB").convert("RGBA")), "jpeg", quality=95)
    save("noise_q30.jpg", Image.fromarray(rgb), "JPEG", np.asarray(Image.fromarray(rgb).convert("RGBA")), "noisy",
         quality=30)
    return out


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("imgs")
    return png_cases(tmp), pillow_cases(tmp)


def test_png_against_source_pixels(decoder, files):
    pngs, _ = files
    got = decode_many(decoder, [p for p, _ in pngs])
    for p, want in pngs:
        g = got[str(p)]
        assert not isinstance(g, str), (p.name, g)
        assert np.array_equal(g, want), p.name


def test_lossless_formats_against_source_pixels(decoder, files):
    _, others = files
    cases = [(p, e) for p, e, k in others if k == "lossless"]
    got = decode_many(decoder, [p for p, _ in cases])
    for p, want in cases:
        g = got[str(p)]
        assert not isinstance(g, str), (p.name, g)
        assert g.shape == want.shape and np.array_equal(g, want), p.name


def test_gif_against_pillow(decoder, files):
    _, others = files
    cases = [p for p, _, k in others if k == "gif"]
    got = decode_many(decoder, cases)
    for p in cases:
        g = got[str(p)]
        want = np.asarray(Image.open(p).convert("RGBA"))
        assert not isinstance(g, str), (p.name, g)
        opaque = want[:, :, 3] > 0
        # transparent pixels: the reference leaves the canvas (0, 0, 0, 0)
        assert np.array_equal(g[opaque], want[opaque]), p.name
        assert np.all(g[~opaque] == 0), p.name


def test_jpeg_close_to_pillow(decoder, files):
    _, others = files
    cases = [(p, e, k) for p, e, k in others if k in ("jpeg", "noisy")]
    got = decode_many(decoder, [p for p, _, _ in cases])
    for p, _, kind in cases:
        g = got[str(p)]
        assert not isinstance(g, str), (p.name, g)
        want = np.asarray(Image.open(p).convert("RGBA")).astype(int)
        d = np.abs(g.astype(int) - want)
        # different IDCTs and upsamplers; chroma edges of subsampled noise move most
        assert d[:, :, 3].max() == 0 and d.mean() < (3.0 if kind == "noisy" else 1.5), (p.name, d.mean(), d.max())


@pytest.mark.skipif(not os.path.exists(STB), reason="reference loader not built (no reference tree)")
def test_every_file_matches_reference_loader(decoder, files):
    pngs, others = files
    paths = [p for p, _ in pngs] + [p for p, _, _ in others]
    got = decode_many(decoder, paths)
    for p in paths:
        want = stb_load(p)
        g = got[str(p)]
        assert not isinstance(want, str), (p.name, want)
        assert not isinstance(g, str), (p.name, g)
        assert g.shape == want.shape and np.array_equal(g, want), (p.name, int(np.sum(g != want)))


@pytest.mark.skipif(not os.path.exists(STB), reason="reference loader not built (no reference tree)")
def test_rejections_match_reference_loader(decoder, tmp_path):
    bad = {
        "empty.png": b"",
        "text.png": b"hello world, not an image",
        "trunc.png": encode_png(np.zeros((4, 4, 1), int), 0, 8)[:40],
        "huge.png": b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", 1 << 25, 1, 8, 0, 0, 0, 0)),
        "bad_ihdr.png": b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", b"\x00" * 12),
        "trunc.jpg": b"\xff\xd8\xff\xdb\x00",
    }
    paths = []
    for name, data in bad.items():
        (tmp_path / name).write_bytes(data)
        paths.append(tmp_path / name)
    got = decode_many(decoder, paths)
    for p in paths:
        want = stb_load(p)
        assert isinstance(want, str), p.name
        assert isinstance(got[str(p)], str), p.name
    assert got[str(tmp_path / "text.png")] == stb_load(tmp_path / "text.png") == \
        "Image not of any known type, or corrupt"
