"""Image files (lrt_image_write, image_write.cpp:10-100) on the CPU.

The 8-bit mapping is checked against its closed form (NaN -> 0, gamma 1/2.2,
clamp to [0, 0.999], x 256, truncated) and against the reference's own
ImageWriter (oracle/_ref, built with the reference's stb) where that is
built: PNG pixels after decoding both files, PPM files byte for byte, HDR
pixels after decoding both files (RGBE; stb writes run-length scanlines).
PFM and raw float32 round-trip exactly.
"""
import json
import struct
import zlib

import numpy as np
import pytest

import linrt
import oracle_lib as OL


def frame(h=13, w=21, seed=3):
    rng = np.random.default_rng(seed)
    img = rng.uniform(0.0, 1.5, (h, w, 3)).astype(np.float32)
    img[0, 0] = [np.nan, 0.0, 1e-30]
    img[0, 1] = [np.inf, 1.0, 0.999]
    img[1, 2] = [0.5, 2.0, 1e-6]
    img[2, 3] = [0.0, 0.0, 0.0]
    return img


def decode_png(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    at, idat, w, h = 8, b"", None, None
    while at < len(data):
        n, = struct.unpack(">I", data[at:at + 4])
        typ, body = data[at + 4:at + 8], data[at + 8:at + 8 + n]
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2
        elif typ == b"IDAT":
            idat += body
        at += 12 + n
    raw = zlib.decompress(idat)
    stride = 3 * w
    out = np.zeros((h, stride), np.int32)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f = raw[y * (stride + 1)]
        row = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        for i in range(stride):
            a = cur[i - 3] if i >= 3 else 0
            b = prev[i]
            c = prev[i - 3] if i >= 3 else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[i] = (row[i] + p) & 0xff
        out[y] = cur
        prev = cur
    return out.reshape(h, w, 3).astype(np.uint8)


def decode_hdr(data):
    head, _, body = data.partition(b"\n\n")
    line, _, px = body.partition(b"\n")
    _, h, _, w = line.split()
    h, w = int(h), int(w)
    out = np.zeros((h, w, 4), np.uint8)
    at = 0
    for y in range(h):
        if 8 <= w < 32768 and px[at] == 2 and px[at + 1] == 2:  # run-length scanline
            at += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = px[at]
                    at += 1
                    if n > 128:
                        out[y, x:x + n - 128, c] = px[at]
                        at += 1
                        x += n - 128
                    else:
                        out[y, x:x + n, c] = np.frombuffer(px, np.uint8, n, at)
                        at += n
                        x += n
        else:
            out[y] = np.frombuffer(px, np.uint8, 4 * w, at).reshape(w, 4)
            at += 4 * w
    return out


def closed_form(img):
    v = np.nan_to_num(img, nan=0.0, posinf=np.inf).astype(np.float32)
    v = np.power(v, np.float32(1.0) / np.float32(2.2), dtype=np.float32)
    v = np.clip(v, np.float32(0.0), np.float32(0.999))
    return (v * np.float32(256.0)).astype(np.uint8)


def test_to_byte_closed_form():
    img = frame()
    np.testing.assert_array_equal(linrt.image_to_bytes(img), closed_form(img))
    np.testing.assert_array_equal(linrt.image_to_bytes(img, flip=True), closed_form(img)[::-1])


@pytest.mark.parametrize("flip", [False, True])
def test_png_ppm_hdr_like_the_reference(tmp_path, flip):
    img = frame()
    h, w = img.shape[:2]
    linrt.write_image(tmp_path / "a.png", img, flip=flip)
    linrt.write_image(tmp_path / "a.ppm", img, flip=flip)
    linrt.write_image(tmp_path / "a.hdr", img, flip=flip)
    png = decode_png((tmp_path / "a.png").read_bytes())
    np.testing.assert_array_equal(png, closed_form(img)[::-1] if flip else closed_form(img))
    if OL.REF is None:
        return
    for fmt, name in ((0, "r.png"), (2, "r.ppm"), (3, "r.hdr")):
        assert OL.REF.ref_image_save(str(tmp_path / name).encode(), w, h, OL._p(img), fmt, int(flip)) == 0
    np.testing.assert_array_equal(png, decode_png((tmp_path / "r.png").read_bytes()))
    assert (tmp_path / "a.ppm").read_bytes() == (tmp_path / "r.ppm").read_bytes()
    np.testing.assert_array_equal(decode_hdr((tmp_path / "a.hdr").read_bytes()),
                                  decode_hdr((tmp_path / "r.hdr").read_bytes()))


def test_float_formats_round_trip(tmp_path):
    img = frame()
    h, w = img.shape[:2]
    linrt.write_image(tmp_path / "a.pfm", img)
    data = (tmp_path / "a.pfm").read_bytes()
    assert data.startswith(f"PF\n{w} {h}\n-1.0\n".encode())
    back = np.frombuffer(data[len(f"PF\n{w} {h}\n-1.0\n"):], "<f4").reshape(h, w, 3)
    np.testing.assert_array_equal(back, img)
    linrt.write_image(tmp_path / "a.raw", img, flip=True)
    side = json.loads((tmp_path / "a.raw.json").read_text())
    assert side["width"] == w and side["height"] == h and side["dtype"] == "float32"
    np.testing.assert_array_equal(np.fromfile(tmp_path / "a.raw", "<f4").reshape(h, w, 3), img[::-1])


def test_jpeg_is_not_written(tmp_path):
    with pytest.raises(linrt.LinrtError) as e:
        linrt.write_image(tmp_path / "a.jpg", frame(), fmt="jpeg")
    assert e.value.code == linrt.ERR_UNSUPPORTED
