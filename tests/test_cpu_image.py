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


def read_exr(data):
    """Minimal OpenEXR reader (single-part scanline, uncompressed, FLOAT or
    HALF channels) written from the file layout: magic, version, attributes
    (name, type, size, value) until an empty name, the line offset table,
    then (y, size, channel rows) blocks.  Returns (header dict, {channel:
    [H, W] array}) with rows in file order."""
    assert struct.unpack("<I", data[:4])[0] == 20000630
    ver, = struct.unpack("<I", data[4:8])
    assert ver & 0xFF == 2 and not ver & 0x1200  # version 2, scanline, single part
    at, hdr = 8, {}
    while data[at] != 0:
        e = data.index(b"\0", at)
        name = data[at:e].decode()
        t = data.index(b"\0", e + 1)
        typ = data[e + 1:t].decode()
        size, = struct.unpack("<i", data[t + 1:t + 5])
        hdr[name] = (typ, data[t + 5:t + 5 + size])
        at = t + 5 + size
    at += 1
    chans, p, raw = [], 0, hdr["channels"][1]
    while raw[p] != 0:
        e = raw.index(b"\0", p)
        pt, = struct.unpack("<i", raw[e + 1:e + 5])
        xs, ys = struct.unpack("<ii", raw[e + 9:e + 17])
        assert xs == 1 and ys == 1
        chans.append((raw[p:e].decode(), pt))
        p = e + 17
    assert hdr["compression"][1] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<4i", hdr["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack(f"<{h}Q", data[at:at + 8 * h])
    out = {c: np.zeros((h, w), np.float32) for c, _ in chans}
    for i, off in enumerate(offsets):
        y, n = struct.unpack("<ii", data[off:off + 8])
        q = off + 8
        for c, pt in chans:
            dt, sz = (np.dtype("<f4"), 4) if pt == 2 else (np.dtype("<f2"), 2)
            out[c][y - y0] = np.frombuffer(data[q:q + sz * w], dt)
            q += sz * w
        assert q - off - 8 == n
    return hdr, out


def test_exr_roundtrip(tmp_path):
    """EXR holds the frame exactly (float32), upright: file row 0 is the
    frame's top row (frame rows are bottom first), or the input's row 0 when
    the input is flipped; the required header attributes are present."""
    img = frame()
    h, w, _ = img.shape
    for flip in (False, True):
        p = tmp_path / f"f{int(flip)}.exr"
        linrt.write_image(p, img, flip=flip)
        hdr, ch = read_exr(p.read_bytes())
        for k in ("channels", "compression", "dataWindow", "displayWindow", "lineOrder", "pixelAspectRatio",
                  "screenWindowCenter", "screenWindowWidth"):
            assert k in hdr
        assert list(ch) == ["B", "G", "R"]
        got = np.stack([ch["R"], ch["G"], ch["B"]], axis=-1)
        want = img if flip else img[::-1]
        assert np.array_equal(got, want, equal_nan=True)


def test_accum_roundtrip_and_resume(tmp_path):
    """FP64 accumulators and their sidecar round-trip exactly; the sidecar
    says where a resumed render continues; sums without squares read back
    with zero squares; a missing or foreign sidecar is an error."""
    rng = np.random.default_rng(9)
    s = rng.normal(size=(7, 5, 3))
    q = s * s + rng.random((7, 5, 3))
    p = tmp_path / "acc.f64"
    linrt.write_accum(p, s, q, spp_done=4096, next_sample=4096, seed=(1 << 40) + 3, max_depth=16, rr_depth=5)
    side = json.loads((tmp_path / "acc.f64.json").read_text())
    assert side["dtype"] == "float64" and side["arrays"] == ["sum", "sumsq"] and side["seed"] == (1 << 40) + 3
    assert p.stat().st_size == 2 * s.size * 8
    s2, q2, info = linrt.read_accum(p)
    assert np.array_equal(s, s2) and np.array_equal(q, q2)
    assert (info["width"], info["height"], info["spp_done"], info["next_sample"]) == (5, 7, 4096, 4096)
    assert (info["seed"], info["max_depth"], info["rr_depth"], info["has_sumsq"]) == ((1 << 40) + 3, 16, 5, 1)
    linrt.write_accum(p, s, None, spp_done=8)
    s3, q3, info = linrt.read_accum(p)
    assert q3 is None and info["has_sumsq"] == 0 and np.array_equal(s, s3)
    with pytest.raises(linrt.LinrtError):
        linrt.read_accum(tmp_path / "missing.f64")
    (tmp_path / "bad.f64.json").write_text('{"width": 5, "height": 7, "dtype": "float32"}')
    with pytest.raises(linrt.LinrtError):
        linrt.read_accum(tmp_path / "bad.f64")
