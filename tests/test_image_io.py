"""Environment image decoding (shiro_hip_load_image), the part of
EnvironmentMap::Load (EnvironmentMap.cpp:49-92) that OpenImageIO provides.

OpenImageIO is absent here, so decoding is pinned by the formats'
specifications rather than by the reference's own loader: files are written
by encoders in this test (numpy + zlib, written from the published PFM,
Radiance RGBE and OpenEXR layouts, independently of the C++ decoder), plus
known answers fixed by the formats (every one of the 65,536 half-float bit
patterns against numpy's IEEE binary16, the RGBE exponent rule
c * 2^(e - 136), a byte-by-byte hand-assembled OpenEXR file, orientation
flags).  Malformed and truncated files must fail with an error, never crash.
Host code only: no device is used.
"""
import os
import struct
import zlib

import numpy as np
import pytest

from shiro_amd import abi, hip


def _load(data):
    return hip.load_image(bytes(data))


# ---------------------------------------------------------------------------
# PFM


def pfm_bytes(img, little=True):
    h, w = img.shape[:2]
    grey = img.ndim == 2
    head = f"{'Pf' if grey else 'PF'}\n{w} {h}\n{-1.0 if little else 1.0}\n".encode()
    return head + np.ascontiguousarray(img[::-1]).astype("<f4" if little else ">f4").tobytes()


@pytest.mark.parametrize("little", [True, False])
def test_pfm_colour_both_byte_orders(little):
    rng = np.random.default_rng(1)
    img = rng.normal(size=(5, 7, 3)).astype(np.float32) * 100
    px, ch, fmt = _load(pfm_bytes(img, little))
    assert fmt == abi.IMAGE_PFM and ch == 3
    assert np.array_equal(px, img)


def test_pfm_grey_expands_to_rgb():
    img = np.arange(12, dtype=np.float32).reshape(3, 4)
    px, ch, _ = _load(pfm_bytes(img))
    assert ch == 1
    assert np.array_equal(px, np.repeat(img[:, :, None], 3, axis=2))


# ---------------------------------------------------------------------------
# Radiance RGBE


def rgbe_rle_plane(v):
    """Adaptive run-length encoding of one component plane: runs of >= 4
    equal bytes as (128 + n, byte), the rest as literal blocks (n, bytes)."""
    out, i, n = bytearray(), 0, len(v)
    while i < n:
        j = i
        while j < n and j - i < 127 and v[j] == v[i]:
            j += 1
        if j - i >= 4:
            out += bytes([128 + j - i, v[i]])
            i = j
            continue
        k = i
        while k < n and k - i < 128:
            r = k
            while r < n and r - k < 4 and v[r] == v[k]:
                r += 1
            if r - k >= 4:
                break
            k += 1
        out += bytes([k - i]) + bytes(v[i:k])
        i = k
    return bytes(out)


def rgbe_bytes(q, rle, res="-Y {h} +X {w}", extra_header=()):
    """q: (rows, cols, 4) uint8 in file order."""
    h, w = q.shape[:2]
    out = bytearray(b"#?RADIANCE\n# written by the test\n")
    for line in extra_header:
        out += line.encode() + b"\n"
    out += b"FORMAT=32-bit_rle_rgbe\n\n" + res.format(h=h, w=w).encode() + b"\n"
    for r in range(h):
        if rle:
            out += bytes([2, 2, w >> 8, w & 255])
            for k in range(4):
                out += rgbe_rle_plane(bytes(q[r, :, k]))
        else:
            out += q[r].tobytes()
    return bytes(out)


def rgbe_expected(q):
    e = q[..., 3].astype(np.int32)
    c = q[..., :3].astype(np.float32)
    return np.where(e[..., None] > 0, np.ldexp(c, (e - 136)[..., None]), np.float32(0)).astype(np.float32)


def test_rgbe_known_answer():
    q = np.array([[[128, 64, 32, 129], [255, 1, 0, 136], [7, 7, 7, 0], [1, 2, 3, 1]]], np.uint8)
    px, ch, fmt = _load(rgbe_bytes(q, rle=False))
    assert fmt == abi.IMAGE_RGBE and ch == 3
    want = np.array([[[1.0, 0.5, 0.25], [255.0, 1.0, 0.0], [0.0, 0.0, 0.0],
                      [2.0 ** -135, 2.0 ** -134, 3 * 2.0 ** -135]]], np.float32)
    assert np.array_equal(px, want)


@pytest.mark.parametrize("rle", [False, True], ids=["flat", "rle"])
def test_rgbe_random_image(rle):
    rng = np.random.default_rng(2)
    h, w = 9, 37
    q = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    q[:, 5:20] = q[:, 5:6]  # long runs
    q[2, :, 3] = 0  # zero exponents
    data = rgbe_bytes(q, rle)
    if rle:
        assert len(data) < h * w * 4  # the runs were encoded
    px, _, _ = _load(data)
    assert np.array_equal(px, rgbe_expected(q))


def test_rgbe_orientations():
    rng = np.random.default_rng(3)
    q = rng.integers(0, 256, (4, 9, 4)).astype(np.uint8)
    top_down = rgbe_expected(q)
    for res, want in [("-Y {h} +X {w}", top_down), ("+Y {h} +X {w}", top_down[::-1]),
                      ("-Y {h} -X {w}", top_down[:, ::-1]), ("+Y {h} -X {w}", top_down[::-1, ::-1])]:
        px, _, _ = _load(rgbe_bytes(q, rle=True, res=res, extra_header=("EXPOSURE=1.0",)))
        assert np.array_equal(px, want), res


# ---------------------------------------------------------------------------
# OpenEXR

HALF, FLOAT, UINT = 1, 2, 0
NONE, RLE, ZIPS, ZIP, PIZ = 0, 1, 2, 3, 4


def _attr(name, type_, value):
    return name.encode() + b"\0" + type_.encode() + b"\0" + struct.pack("<i", len(value)) + value


def _split_predict(raw):
    """Even/odd byte split, then the byte predictor (OpenEXR zip/rle)."""
    t = np.frombuffer(raw, np.uint8)
    t = np.concatenate([t[0::2], t[1::2]]).astype(np.int32)
    out = t.copy()
    out[1:] = (t[1:] - t[:-1] + 128) & 255
    return out.astype(np.uint8).tobytes()


def _exr_rle(b):
    out, i, n = bytearray(), 0, len(b)
    while i < n:
        j = i
        while j < n and j - i < 128 and b[j] == b[i]:
            j += 1
        if j - i >= 3:
            out += struct.pack("<b", j - i - 1) + bytes([b[i]])
            i = j
            continue
        k = i
        while k < n and k - i < 127 and not (k + 2 < n and b[k] == b[k + 1] == b[k + 2]):
            k += 1
        out += struct.pack("<b", -(k - i)) + bytes(b[i:k])
        i = k
    return bytes(out)


def exr_bytes(channels, comp, xmin=0, ymin=0, reverse_chunks=False):
    """channels: {name: (pixel type, (H, W) array)}; stored in name order."""
    names = sorted(channels)
    h, w = channels[names[0]][1].shape
    chl = b""
    for n in names:
        chl += n.encode() + b"\0" + struct.pack("<iB3xii", channels[n][0], 0, 1, 1)
    chl += b"\0"
    box = struct.pack("<iiii", xmin, ymin, xmin + w - 1, ymin + h - 1)
    head = struct.pack("<ii", 20000630, 2)
    head += _attr("channels", "chlist", chl)
    head += _attr("compression", "compression", bytes([comp]))
    head += _attr("dataWindow", "box2i", box)
    head += _attr("displayWindow", "box2i", box)
    head += _attr("lineOrder", "lineOrder", b"\0")
    head += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    head += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
    head += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    head += b"\0"
    lpb = 16 if comp == ZIP else 1
    dt = {HALF: "<f2", FLOAT: "<f4", UINT: "<u4"}
    chunks = []
    for y0 in range(0, h, lpb):
        raw = b""
        for y in range(y0, min(h, y0 + lpb)):
            for n in names:
                t, a = channels[n]
                raw += np.ascontiguousarray(a[y]).astype(dt[t]).tobytes()
        if comp == NONE:
            data = raw
        elif comp == RLE:
            data = _exr_rle(_split_predict(raw))
        else:
            data = zlib.compress(_split_predict(raw))
        if len(data) >= len(raw):
            data = raw  # stored uncompressed when the codec does not shrink it
        chunks.append(struct.pack("<ii", ymin + y0, len(data)) + data)
    order = list(range(len(chunks)))[::-1] if reverse_chunks else list(range(len(chunks)))
    base = len(head) + 8 * len(chunks)
    offsets, body, pos = [0] * len(chunks), b"", base
    for k in order:
        offsets[k] = pos
        body += chunks[k]
        pos += len(chunks[k])
    return head + struct.pack(f"<{len(chunks)}Q", *offsets) + body


def test_exr_all_half_bit_patterns():
    """Every binary16 pattern (zeros, subnormals, normals, infinities, NaNs)
    decodes to numpy's float16 -> float32 conversion."""
    bits = np.arange(65536, dtype=np.uint16).reshape(256, 256)
    ref = bits.view(np.float16).astype(np.float32)
    data = exr_bytes({"Y": (HALF, bits.view(np.float16))}, NONE)
    px, ch, fmt = _load(data)
    assert fmt == abi.IMAGE_EXR and ch == 1
    g = px[:, :, 0]
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(g), nan)
    assert np.array_equal(g[~nan].view(np.uint32), ref[~nan].view(np.uint32))  # signed zeros too
    assert np.array_equal(px[:, :, 1].view(np.uint32), g.view(np.uint32))
    assert np.array_equal(px[:, :, 2].view(np.uint32), g.view(np.uint32))


def test_exr_hand_assembled_file():
    """A 2x1 RGB FLOAT file, NONE compression, laid out byte by byte from the
    OpenEXR file-layout document (magic, version, attributes, offset table,
    one chunk per scanline, channels in alphabetical order B, G, R)."""
    f32 = lambda *v: struct.pack("<" + "f" * len(v), *v)  # noqa: E731
    i32 = lambda *v: struct.pack("<" + "i" * len(v), *v)  # noqa: E731
    chl = b"B\0" + i32(2) + b"\0\0\0\0" + i32(1, 1)
    chl += b"G\0" + i32(2) + b"\0\0\0\0" + i32(1, 1)
    chl += b"R\0" + i32(2) + b"\0\0\0\0" + i32(1, 1) + b"\0"
    head = bytes([0x76, 0x2F, 0x31, 0x01]) + bytes([2, 0, 0, 0])
    head += b"channels\0chlist\0" + i32(len(chl)) + chl
    head += b"compression\0compression\0" + i32(1) + b"\0"
    head += b"dataWindow\0box2i\0" + i32(16) + i32(0, 0, 1, 0)
    head += b"displayWindow\0box2i\0" + i32(16) + i32(0, 0, 1, 0)
    head += b"lineOrder\0lineOrder\0" + i32(1) + b"\0"
    head += b"pixelAspectRatio\0float\0" + i32(4) + f32(1.0)
    head += b"screenWindowCenter\0v2f\0" + i32(8) + f32(0.0, 0.0)
    head += b"screenWindowWidth\0float\0" + i32(4) + f32(1.0) + b"\0"
    chunk = i32(0, 24) + f32(0.25, 4.0) + f32(0.5, 5.0) + f32(1.0, 6.0)  # B row, G row, R row
    data = head + struct.pack("<Q", len(head) + 8) + chunk
    px, ch, _ = _load(data)
    assert ch == 3
    assert np.array_equal(px, np.array([[[1.0, 0.5, 0.25], [6.0, 5.0, 4.0]]], np.float32))


@pytest.mark.parametrize("comp", [NONE, RLE, ZIPS, ZIP], ids=["none", "rle", "zips", "zip"])
@pytest.mark.parametrize("ptype", [HALF, FLOAT], ids=["half", "float"])
def test_exr_codecs(comp, ptype):
    rng = np.random.default_rng(10 + comp)
    h, w = 37, 23
    base = np.exp(rng.normal(size=(h, w, 3))).astype(np.float32)
    base[5:12, 3:17] = base[5, 3]  # runs for RLE, repeats for zip
    if ptype == HALF:
        base = base.astype(np.float16)
    chans = {"R": (ptype, base[:, :, 0]), "G": (ptype, base[:, :, 1]), "B": (ptype, base[:, :, 2]),
             "A": (ptype, np.ones((h, w), base.dtype))}
    data = exr_bytes(chans, comp, xmin=-3, ymin=5, reverse_chunks=comp == ZIP)
    px, ch, _ = _load(data)
    assert ch == 4
    assert np.array_equal(px, base.astype(np.float32))


def test_exr_uint_and_partial_colour():
    """UINT samples convert to float; R and G without B expand b = g, as
    Load does for a two-channel image."""
    r = np.arange(6, dtype=np.uint32).reshape(2, 3) * 1000
    g = np.arange(6, dtype=np.uint32).reshape(2, 3) + 7
    px, ch, _ = _load(exr_bytes({"R": (UINT, r), "G": (UINT, g)}, ZIPS))
    assert ch == 2
    assert np.array_equal(px[:, :, 0], r.astype(np.float32))
    assert np.array_equal(px[:, :, 1], g.astype(np.float32))
    assert np.array_equal(px[:, :, 2], g.astype(np.float32))


def test_exr_rejects_unsupported():
    y = {"Y": (HALF, np.ones((2, 2), np.float16))}
    good = exr_bytes(y, NONE)
    for mutate, msg in [
        (lambda b: b[:4] + struct.pack("<i", 2 | 0x200) + b[8:], "tiled"),
        (lambda b: b[:4] + struct.pack("<i", 2 | 0x1000) + b[8:], "multi-part"),
        (lambda b: b.replace(b"compression\0compression\0\x01\0\0\0\0", b"compression\0compression\0\x01\0\0\0\x04"),
         "PIZ"),
    ]:
        with pytest.raises(hip.HipError, match=msg):
            _load(mutate(good))
    with pytest.raises(hip.HipError, match="no R or Y"):
        _load(exr_bytes({"Z": (HALF, np.ones((2, 2), np.float16))}, NONE))


def test_malformed_files_fail_cleanly():
    """Truncations and byte corruption of valid files of every format return
    an error or a decoded image; nothing reads out of range."""
    rng = np.random.default_rng(20)
    img = np.exp(rng.normal(size=(20, 33, 3))).astype(np.float32)
    q = rng.integers(0, 256, (20, 33, 4)).astype(np.uint8)
    q[:, 4:14] = q[:, 4:5]
    files = [pfm_bytes(img), rgbe_bytes(q, rle=True),
             exr_bytes({"R": (HALF, img[..., 0].astype(np.float16)), "G": (FLOAT, img[..., 1]),
                        "B": (HALF, img[..., 2].astype(np.float16))}, ZIP),
             exr_bytes({"Y": (FLOAT, img[..., 0])}, RLE)]
    for f in files:
        for cut in sorted(set(rng.integers(0, len(f), 60).tolist())):
            with pytest.raises(hip.HipError):
                _load(f[:cut])
        for _ in range(300):
            b = bytearray(f)
            for p in rng.integers(0, len(b), rng.integers(1, 4)):
                b[p] = rng.integers(0, 256)
            try:
                px, _, _ = _load(b)
                assert px.ndim == 3 and px.shape[2] == 3
            except hip.HipError:
                pass
    with pytest.raises(hip.HipError, match="unrecognised"):
        _load(b"GIF89a....")


def test_load_from_path_and_dome_light(tmp_path):
    img = np.exp(np.random.default_rng(4).normal(size=(16, 32, 3))).astype(np.float32)
    p = tmp_path / "sky.pfm"
    p.write_bytes(pfm_bytes(img))
    px, _, _ = hip.load_image(str(p))
    assert np.array_equal(px, img)
    d = hip.load_dome_light(str(p), abi.ENV_LATLONG, radiance=(2.0, 2.0, 2.0))
    assert d.layout == abi.ENV_LATLONG and np.array_equal(d.pixels, img)
    with pytest.raises(hip.HipError, match="cannot open"):
        hip.load_image(os.path.join(str(tmp_path), "missing.exr"))
