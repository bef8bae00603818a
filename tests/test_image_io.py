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


def test_exr_rejects_files_without_colour():
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


# ---------------------------------------------------------------------------
# OpenEXR PIZ, PXR24, B44/B44A, tiled and multi-part files.  The encoders
# below follow the OpenEXR codec descriptions (PIZ: value table, wavelet,
# Huffman with a run-length symbol; PXR24: byte planes of sample
# differences, the float's top 24 bits; B44: 4x4 blocks of 14 or 3 bytes)
# and are written independently of the C++ decoder.

PXR24, B44, B44A, DWAA = 5, 6, 7, 8


def _wenc14(a, b):
    a, b = a - 65536 if a >= 32768 else a, b - 65536 if b >= 32768 else b
    return ((a + b) >> 1) & 0xFFFF, (a - b) & 0xFFFF


def _wenc16(a, b):
    ao = (a + 0x8000) & 0xFFFF
    m = (ao + b) >> 1
    d = ao - b
    if d < 0:
        m = (m + 0x8000) & 0xFFFF
    return m, d & 0xFFFF


def _wav2_encode(buf, start, nx, ox, ny, oy, mx):
    """Forward 2D wavelet in place over buf[start + x * ox + y * oy]."""
    enc = _wenc14 if mx < (1 << 14) else _wenc16
    n = min(nx, ny)
    p, p2 = 1, 2
    while p2 <= n:
        oy1, oy2, ox1, ox2 = oy * p, oy * p2, ox * p, ox * p2
        ey = start + oy * (ny - p2)
        py = start
        while py <= ey:
            px = py
            ex = py + ox * (nx - p2)
            while px <= ex:
                p01, p10 = px + ox1, px + oy1
                p11 = p10 + ox1
                i00, i01 = enc(buf[px], buf[p01])
                i10, i11 = enc(buf[p10], buf[p11])
                buf[px], buf[p10] = enc(i00, i10)
                buf[p01], buf[p11] = enc(i01, i11)
                px += ox2
            if nx & p:
                p10 = px + oy1
                buf[px], buf[p10] = enc(buf[px], buf[p10])
            py += oy2
        if ny & p:
            px = py
            ex = py + ox * (nx - p2)
            while px <= ex:
                p01 = px + ox1
                buf[px], buf[p01] = enc(buf[px], buf[p01])
                px += ox2
        p, p2 = p2, p2 << 1


class _Bits:
    def __init__(self):
        self.bits = []

    def put(self, n, v):
        self.bits.extend((v >> (n - 1 - i)) & 1 for i in range(n))

    def bytes(self):
        b = self.bits + [0] * (-len(self.bits) % 8)
        return bytes(int("".join(map(str, b[i:i + 8])), 2) for i in range(0, len(b), 8))


def _huf_encode(words, min_run=3):
    """OpenEXR Huffman stream of 16-bit words (header, packed code lengths,
    canonical codes with the longest first, run-length symbol = iM)."""
    import heapq
    freq = np.bincount(np.asarray(words, np.int64), minlength=65536)
    present = np.nonzero(freq)[0]
    im, rlc = int(present[0]), int(present[-1]) + 1
    f = {int(s): int(freq[s]) for s in present}
    f[rlc] = 1
    heap = [(c, i, [s]) for i, (s, c) in enumerate(sorted(f.items()))]
    heapq.heapify(heap)
    length = dict.fromkeys(f, 0)
    k = len(heap)
    while len(heap) > 1:
        c1, _, a = heapq.heappop(heap)
        c2, _, b = heapq.heappop(heap)
        for s in a + b:
            length[s] += 1
        heapq.heappush(heap, (c1 + c2, k, a + b))
        k += 1
    assert max(length.values()) <= 58
    count = [0] * 59
    for l in length.values():
        count[l] += 1
    first, c = [0] * 59, 0
    for l in range(58, 0, -1):
        first[l] = c
        c = (c + count[l]) >> 1
    code, nxt = {}, list(first)
    for s in sorted(length):
        code[s] = nxt[length[s]]
        nxt[length[s]] += 1
    table = _Bits()
    i = im
    while i <= rlc:
        l = length.get(i, 0)
        if l == 0:
            run = 1
            while i + run <= rlc and length.get(i + run, 0) == 0 and run < 255 + 6:
                run += 1
            if run >= 6:
                table.put(6, 63)
                table.put(8, run - 6)
                i += run
                continue
            if run >= 2:
                table.put(6, 59 + run - 2)
                i += run
                continue
        table.put(6, l)
        i += 1
    tbytes = table.bytes()
    out = _Bits()
    words = [int(w) for w in words]
    j = 0
    while j < len(words):
        s = words[j]
        r = 1
        while j + r < len(words) and words[j + r] == s and r < 256:
            r += 1
        out.put(length[s], code[s])
        if r >= min_run:
            out.put(length[rlc], code[rlc])
            out.put(8, r - 1)
            j += r
        else:
            j += 1
    nbits = len(out.bits)
    return struct.pack("<IIIII", im, rlc, len(tbytes), nbits, 0) + tbytes + out.bytes()


def _channel_words(t, a):
    """A channel's samples of a block as 16-bit words, (h, w * size)."""
    a = np.ascontiguousarray(a)
    if t == HALF:
        return a.astype("<f2").view("<u2")
    return a.astype("<f4" if t == FLOAT else "<u4").view("<u2")


def _piz_block(chans, names):
    parts = [_channel_words(*chans[n]) for n in names]
    buf = np.concatenate([p.ravel() for p in parts]).astype(np.int64)
    used = np.unique(buf)
    bitmap = np.zeros(8192, np.uint8)
    np.bitwise_or.at(bitmap, used >> 3, (1 << (used & 7)).astype(np.uint8))
    bitmap[0] &= 0xFE  # value 0 is always in the table
    nz = np.nonzero(bitmap)[0]
    mn, mx = (int(nz[0]), int(nz[-1])) if len(nz) else (8191, 0)
    keep = np.concatenate([[0], used[used != 0]])
    lut = np.zeros(65536, np.int64)
    lut[keep] = np.arange(len(keep))
    maxv = len(keep) - 1
    buf = lut[buf].tolist()
    off = 0
    for n, p in zip(names, parts):
        h, ws = p.shape
        w = chans[n][1].shape[1]
        size = ws // w
        for j in range(size):
            _wav2_encode(buf, off + j, w, size, h, w * size, maxv)
        off += p.size
    huf = _huf_encode(buf)
    out = struct.pack("<HH", mn, mx)
    if mn <= mx:
        out += bitmap[mn:mx + 1].tobytes()
    return out + struct.pack("<i", len(huf)) + huf


def _float24(f):
    """The float's top 24 bits, rounded (PXR24)."""
    i = int(np.float32(f).view(np.uint32))
    s, e, m = i & 0x80000000, i & 0x7F800000, i & 0x007FFFFF
    if e == 0x7F800000:
        r = (e >> 8) | (m >> 8) | (1 if m and not (m >> 8) else 0)
    else:
        r = ((e | m) + (m & 0x80)) >> 8
        if r >= 0x7F8000:
            r = (e | m) >> 8
    return (s >> 8) | r


def _pxr24_block(chans, names):
    h = chans[names[0]][1].shape[0]
    out = bytearray()
    for y in range(h):
        for n in names:
            t, a = chans[n]
            row = np.ascontiguousarray(a[y])
            prev = 0
            if t == HALF:
                v = row.astype("<f2").view("<u2").astype(np.int64)
                planes = [[], []]
                for x in v:
                    d = (int(x) - prev) & 0xFFFF
                    prev = int(x)
                    planes[0].append(d >> 8)
                    planes[1].append(d & 255)
            elif t == FLOAT:
                planes = [[], [], []]
                for x in row.astype(np.float32):
                    f24 = _float24(x)
                    d = (f24 - prev) & 0xFFFFFFFF
                    prev = f24
                    planes[0].append((d >> 16) & 255)
                    planes[1].append((d >> 8) & 255)
                    planes[2].append(d & 255)
            else:
                planes = [[], [], [], []]
                for x in row.astype(np.uint32):
                    d = (int(x) - prev) & 0xFFFFFFFF
                    prev = int(x)
                    for k in range(4):
                        planes[k].append((d >> (24 - 8 * k)) & 255)
            for pl in planes:
                out += bytes(pl)
    return zlib.compress(bytes(out))


def _shift_round(x, shift):
    x <<= 1
    a = (1 << shift) - 1
    shift += 1
    b = (x >> shift) & 1
    return (x + a + b) >> shift


def _b44_pack(s, flat):
    t = []
    for v in s:
        if (v & 0x7C00) == 0x7C00:
            t.append(0x8000)
        elif v & 0x8000:
            t.append(~v & 0xFFFF)
        else:
            t.append(v | 0x8000)
    tmax = max(t)
    pairs = [(0, 4), (4, 8), (8, 12), (0, 1), (4, 5), (8, 9), (12, 13), (1, 2), (5, 6), (9, 10), (13, 14),
             (2, 3), (6, 7), (10, 11), (14, 15)]
    shift = -1
    while True:
        shift += 1
        d = [_shift_round(tmax - x, shift) for x in t]
        r = [d[a] - d[b] + 0x20 for a, b in pairs]
        if min(r) >= 0 and max(r) <= 0x3F:
            break
    if flat and min(r) == 0x20 and max(r) == 0x20:
        return bytes([t[0] >> 8, t[0] & 255, 0xFC])
    t0 = (tmax - (d[0] << shift)) & 0xFFFF
    b = [t0 >> 8, t0 & 255, (shift << 2) | (r[0] >> 4), (r[0] << 4) | (r[1] >> 2), (r[1] << 6) | r[2],
         (r[3] << 2) | (r[4] >> 4), (r[4] << 4) | (r[5] >> 2), (r[5] << 6) | r[6],
         (r[7] << 2) | (r[8] >> 4), (r[8] << 4) | (r[9] >> 2), (r[9] << 6) | r[10],
         (r[11] << 2) | (r[12] >> 4), (r[12] << 4) | (r[13] >> 2), (r[13] << 6) | r[14]]
    return bytes(x & 255 for x in b)


def _b44_unpack(b):
    """Decoded 16 samples of one block (the C++ decoder's expected output)."""
    if b[2] >= 52:
        v = (b[0] << 8) | b[1]
        v = v & 0x7FFF if v & 0x8000 else ~v & 0xFFFF
        return [v] * 16
    s = [0] * 16
    s[0] = (b[0] << 8) | b[1]
    shift = b[2] >> 2
    bias = 0x20 << shift
    spec = [(4, 0, ((b[2] << 4) | (b[3] >> 4)) & 0x3F), (8, 4, ((b[3] << 2) | (b[4] >> 6)) & 0x3F),
            (12, 8, b[4] & 0x3F), (1, 0, b[5] >> 2), (5, 4, ((b[5] << 4) | (b[6] >> 4)) & 0x3F),
            (9, 8, ((b[6] << 2) | (b[7] >> 6)) & 0x3F), (13, 12, b[7] & 0x3F), (2, 1, b[8] >> 2),
            (6, 5, ((b[8] << 4) | (b[9] >> 4)) & 0x3F), (10, 9, ((b[9] << 2) | (b[10] >> 6)) & 0x3F),
            (14, 13, b[10] & 0x3F), (3, 2, b[11] >> 2), (7, 6, ((b[11] << 4) | (b[12] >> 4)) & 0x3F),
            (11, 10, ((b[12] << 2) | (b[13] >> 6)) & 0x3F), (15, 14, b[13] & 0x3F)]
    for i, f, r in spec:
        s[i] = (s[f] + (r << shift) - bias) & 0xFFFF
    return [v & 0x7FFF if v & 0x8000 else ~v & 0xFFFF for v in s]


def _b44_block(chans, names, flat, expect):
    """B44 data of a block; `expect` receives each channel's decoded samples."""
    out = b""
    for n in names:
        t, a = chans[n]
        if t != HALF:
            out += np.ascontiguousarray(a).astype("<f4" if t == FLOAT else "<u4").tobytes()
            expect[n] = np.asarray(a).astype(np.float32)
            continue
        v = np.ascontiguousarray(a).astype("<f2").view("<u2").astype(np.int64)
        h, w = v.shape
        dec = np.zeros((h, w), np.uint16)
        for y in range(0, h, 4):
            for x in range(0, w, 4):
                s = [int(v[min(y + i // 4, h - 1), min(x + i % 4, w - 1)]) for i in range(16)]
                blk = _b44_pack(s, flat)
                out += blk
                d = _b44_unpack(blk)
                for i in range(16):
                    if y + i // 4 < h and x + i % 4 < w:
                        dec[y + i // 4, x + i % 4] = d[i]
        expect[n] = dec.view(np.float16).astype(np.float32)
    return out


def _encode_block(chans, names, comp, expect=None):
    raw = b""
    h = chans[names[0]][1].shape[0]
    for y in range(h):
        for n in names:
            t, a = chans[n]
            raw += np.ascontiguousarray(a[y]).astype({HALF: "<f2", FLOAT: "<f4", UINT: "<u4"}[t]).tobytes()
    if comp == NONE:
        data = raw
    elif comp == RLE:
        data = _exr_rle(_split_predict(raw))
    elif comp in (ZIPS, ZIP):
        data = zlib.compress(_split_predict(raw))
    elif comp == PIZ:
        data = _piz_block(chans, names)
    elif comp == PXR24:
        data = _pxr24_block(chans, names)
    elif comp in (B44, B44A):
        data = _b44_block(chans, names, comp == B44A, expect if expect is not None else {})
    else:
        raise ValueError(comp)
    if len(data) >= len(raw):
        if expect is not None:
            expect.clear()
        return raw  # stored uncompressed when the codec does not shrink it
    return data


LINES = {NONE: 1, RLE: 1, ZIPS: 1, ZIP: 16, PIZ: 32, PXR24: 16, B44: 32, B44A: 32}


def _header(channels, comp, w, h, xmin=0, ymin=0, extra=b""):
    names = sorted(channels)
    chl = b""
    for n in names:
        chl += n.encode() + b"\0" + struct.pack("<iB3xii", channels[n][0], 0, 1, 1)
    chl += b"\0"
    box = struct.pack("<iiii", xmin, ymin, xmin + w - 1, ymin + h - 1)
    return (_attr("channels", "chlist", chl) + _attr("compression", "compression", bytes([comp])) +
            _attr("dataWindow", "box2i", box) + _attr("displayWindow", "box2i", box) +
            _attr("lineOrder", "lineOrder", b"\0") + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
            _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0)) +
            _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + extra)


def _chunks(channels, comp, tile=None, expect=None):
    """Chunk payloads (without part numbers): scanline blocks, or level-0
    tiles in row-major order as (tx, ty, data)."""
    names = sorted(channels)
    h, w = channels[names[0]][1].shape
    out = []
    if tile is None:
        lpb = LINES[comp]
        for y0 in range(0, h, lpb):
            sub = {n: (t, a[y0:y0 + lpb]) for n, (t, a) in channels.items()}
            e = {}
            data = _encode_block(sub, names, comp, e)
            if expect is not None:
                for n in names:
                    expect.setdefault(n, np.zeros((h, w), np.float32))[y0:y0 + lpb] = (
                        e[n] if n in e else np.asarray(sub[n][1]).astype(np.float32))
            out.append((y0, data))
    else:
        tw, th = tile
        for ty in range(0, (h + th - 1) // th):
            for tx in range(0, (w + tw - 1) // tw):
                sub = {n: (t, a[ty * th:(ty + 1) * th, tx * tw:(tx + 1) * tw]) for n, (t, a) in channels.items()}
                e = {}
                data = _encode_block(sub, names, comp, e)
                if expect is not None:
                    for n in names:
                        expect.setdefault(n, np.zeros((h, w), np.float32))[
                            ty * th:(ty + 1) * th, tx * tw:(tx + 1) * tw] = (
                            e[n] if n in e else np.asarray(sub[n][1]).astype(np.float32))
                out.append((tx, ty, data))
    return out


def exr_file(channels, comp, tile=None, level_mode=0, xmin=0, ymin=0, reverse=False, expect=None):
    """Single-part scanline or tiled file.  Tiled mipmap/ripmap files get
    level-1 tiles (of a 2x-decimated image) after level 0's."""
    names = sorted(channels)
    h, w = channels[names[0]][1].shape
    extra = b""
    if tile is not None:
        extra = _attr("tiles", "tiledesc", struct.pack("<IIB", tile[0], tile[1], level_mode))
    head = struct.pack("<ii", 20000630, 2 | (0x200 if tile else 0)) + _header(channels, comp, w, h, xmin, ymin,
                                                                              extra) + b"\0"
    chunks = _chunks(channels, comp, tile, expect)
    blobs = []
    for ck in chunks:
        if tile is None:
            blobs.append(struct.pack("<ii", ymin + ck[0], len(ck[1])) + ck[1])
        else:
            blobs.append(struct.pack("<iiiii", ck[0], ck[1], 0, 0, len(ck[2])) + ck[2])
    if tile is not None and level_mode:
        half = {n: (t, a[::2, ::2]) for n, (t, a) in channels.items()}
        for tx, ty, data in _chunks(half, comp, tile):
            lx, ly = (1, 1) if level_mode == 1 else (1, 0)
            blobs.append(struct.pack("<iiiii", tx, ty, lx, ly, len(data)) + data)
    order = list(range(len(blobs)))[::-1] if reverse else list(range(len(blobs)))
    base = len(head) + 8 * len(blobs)
    offsets, body, pos = [0] * len(blobs), b"", base
    for k in order:
        offsets[k] = pos
        body += blobs[k]
        pos += len(blobs[k])
    return head + struct.pack(f"<{len(blobs)}Q", *offsets) + body


def exr_multipart(parts, expect=None):
    """parts: list of (channels, comp, tile or None, part name)."""
    head = struct.pack("<ii", 20000630, 2 | 0x1000)
    all_chunks = []
    for i, (channels, comp, tile, name) in enumerate(parts):
        names = sorted(channels)
        h, w = channels[names[0]][1].shape
        chunks = _chunks(channels, comp, tile, expect if i == 0 else None)
        extra = (_attr("name", "string", name.encode()) +
                 _attr("type", "string", b"tiledimage" if tile else b"scanlineimage") +
                 _attr("chunkCount", "int", struct.pack("<i", len(chunks))))
        if tile:
            extra += _attr("tiles", "tiledesc", struct.pack("<IIB", tile[0], tile[1], 0))
        head += _header(channels, comp, w, h, extra=extra) + b"\0"
        blobs = []
        for ck in chunks:
            if tile is None:
                blobs.append(struct.pack("<iii", i, ck[0], len(ck[1])) + ck[1])
            else:
                blobs.append(struct.pack("<iiiiii", i, ck[0], ck[1], 0, 0, len(ck[2])) + ck[2])
        all_chunks.append(blobs)
    head += b"\0"
    n = sum(len(b) for b in all_chunks)
    pos = len(head) + 8 * n
    table, body = b"", b""
    for blobs in all_chunks:
        for b in blobs:
            table += struct.pack("<Q", pos)
            body += b
            pos += len(b)
    return head + table + body


def _rgb_channels(rng, h, w, ptype, spread=1.0):
    base = np.exp(spread * rng.normal(size=(h, w, 3))).astype(np.float32)
    base[3:9, 2:11] = base[min(3, h - 1), min(2, w - 1)]  # flat areas: runs, flat blocks
    if ptype == HALF:
        base = base.astype(np.float16)
    if ptype == UINT:
        base = (base * 1000).astype(np.uint32)
    return {"R": (ptype, base[:, :, 0]), "G": (ptype, base[:, :, 1]), "B": (ptype, base[:, :, 2])}, base


@pytest.mark.parametrize("ptype", [HALF, FLOAT, UINT], ids=["half", "float", "uint"])
@pytest.mark.parametrize("shape", [(37, 23), (70, 45), (1, 9), (33, 1)], ids=["37x23", "70x45", "1x9", "33x1"])
def test_exr_piz(ptype, shape):
    """PIZ is lossless: every sample comes back bit for bit (few distinct
    words use the 14-bit wavelet step, float noise the 16-bit one)."""
    rng = np.random.default_rng(30 + ptype + shape[0])
    chans, base = _rgb_channels(rng, *shape, ptype)
    data = exr_file(chans, PIZ, xmin=2, ymin=-7)
    px, ch, _ = _load(data)
    assert ch == 3
    assert np.array_equal(px.view(np.uint32), base.astype(np.float32).view(np.uint32))


def test_exr_piz_wavelet_forms():
    """A block with fewer than 2^14 distinct words (14-bit wavelet step) and
    one with more (16-bit modular step), both round trips exact."""
    rng = np.random.default_rng(44)
    small = rng.integers(0, 9000, (40, 52)).astype(np.uint16).view(np.float16)
    big = rng.integers(0, 65536, (160, 130)).astype(np.uint16)
    big[(big & 0x7C00) == 0x7C00] = 0x3C00  # keep finite values for a plain comparison
    big = big.view(np.float16)
    for a, distinct in ((small, lambda n: n < 1 << 14), (big, lambda n: n > 1 << 14)):
        assert distinct(len(np.unique(a.view(np.uint16))))
        px, _, _ = _load(exr_file({"Y": (HALF, a)}, PIZ))
        assert np.array_equal(px[:, :, 0], a.astype(np.float32))


@pytest.mark.parametrize("ptype", [HALF, FLOAT, UINT], ids=["half", "float", "uint"])
def test_exr_pxr24(ptype):
    """PXR24: HALF and UINT exact; FLOAT to its top 24 bits, rounded."""
    rng = np.random.default_rng(50 + ptype)
    chans, base = _rgb_channels(rng, 41, 19, ptype)
    px, _, _ = _load(exr_file(chans, PXR24))
    if ptype == FLOAT:
        want = np.vectorize(lambda f: np.uint32(_float24(f) << 8))(base).astype(np.uint32).view(np.float32)
        assert np.array_equal(px.view(np.uint32), want.view(np.uint32))
        assert np.allclose(px, base, rtol=2.0 ** -15, atol=0)
    else:
        assert np.array_equal(px, base.astype(np.float32))


@pytest.mark.parametrize("comp", [B44, B44A], ids=["b44", "b44a"])
def test_exr_b44(comp):
    """B44/B44A: each 4x4 block decodes to what the block's bytes state
    (checked with an independent unpacking), within B44's error of the
    original; flat blocks (3 bytes in B44A) and slowly varying blocks
    (shift 0) are exact; FLOAT channels pass through unchanged."""
    rng = np.random.default_rng(60 + comp)
    h, w = 37, 22
    smooth = (1.0 + 0.001 * np.add.outer(np.arange(h), np.arange(w))).astype(np.float16)
    noisy = np.exp(0.05 * rng.normal(size=(h, w))).astype(np.float16)
    noisy[4:12, 4:12] = noisy[4, 4]
    f = rng.normal(size=(h, w)).astype(np.float32)
    chans = {"R": (HALF, noisy), "G": (HALF, smooth), "B": (FLOAT, f)}
    expect = {}
    data = exr_file(chans, comp, expect=expect)
    px, _, _ = _load(data)
    assert np.array_equal(px[:, :, 0], expect["R"])
    assert np.array_equal(px[:, :, 1], expect["G"])
    assert np.array_equal(px[:, :, 2], f)
    assert np.array_equal(px[:, :, 1], smooth.astype(np.float32))
    assert np.array_equal(px[4:12, 4:12, 0], noisy[4:12, 4:12].astype(np.float32))
    rel = np.abs(px[:, :, 0] - noisy.astype(np.float32)) / np.abs(noisy.astype(np.float32))
    assert rel.max() < 0.02
    if comp == B44A:
        assert len(data) < len(exr_file(chans, B44))  # flat blocks in 3 bytes


@pytest.mark.parametrize("comp", [NONE, ZIP, PIZ, PXR24, B44A], ids=["none", "zip", "piz", "pxr24", "b44a"])
@pytest.mark.parametrize("level_mode", [0, 1, 2], ids=["one", "mip", "rip"])
def test_exr_tiled(comp, level_mode):
    """Tiled files: level-0 tiles (edge tiles clipped to the data window)
    placed by their coordinates, in any chunk order; mipmap and ripmap
    levels beyond 0 are skipped."""
    rng = np.random.default_rng(70 + comp)
    chans, base = _rgb_channels(rng, 45, 38, HALF)
    expect = {}
    data = exr_file(chans, comp, tile=(16, 12), level_mode=level_mode, reverse=True, expect=expect)
    px, ch, _ = _load(data)
    assert ch == 3
    for i, n in enumerate("RGB"):
        assert np.array_equal(px[:, :, i], expect[n]), n


def test_exr_multipart():
    """Multi-part files decode their first part (scanline or tiled), as
    OIIO's subimage 0."""
    rng = np.random.default_rng(80)
    a, base_a = _rgb_channels(rng, 20, 30, FLOAT)
    b, _ = _rgb_channels(rng, 9, 7, HALF)
    px, _, _ = _load(exr_multipart([(a, PIZ, None, "beauty"), (b, ZIP, None, "other")]))
    assert np.array_equal(px, base_a)
    px, _, _ = _load(exr_multipart([(a, ZIPS, (8, 8), "tiles"), (b, NONE, None, "other")]))
    assert np.array_equal(px, base_a)


def test_exr_rejects_unsupported_forms():
    y = {"Y": (HALF, np.ones((4, 4), np.float16))}
    good = exr_file(y, NONE)
    for mutate, msg in [
        (lambda b: b[:4] + struct.pack("<i", 2 | 0x800) + b[8:], "deep"),
        (lambda b: b.replace(b"compression\0compression\0\x01\0\0\0\0",
                             b"compression\0compression\0\x01\0\0\0\x08"), "DWAA"),
        (lambda b: b.replace(b"compression\0compression\0\x01\0\0\0\0",
                             b"compression\0compression\0\x01\0\0\0\x09"), "DWAB"),
        (lambda b: b[:4] + struct.pack("<i", 2 | 0x200) + b[8:], "tile description"),
    ]:
        with pytest.raises(hip.HipError, match=msg):
            _load(mutate(good))
    lin = exr_file({"Y": (HALF, np.ones((4, 4), np.float16) * 3)}, B44)
    lin = lin.replace(b"Y\0\x01\0\0\0\0", b"Y\0\x01\0\0\0\x01")  # pLinear
    with pytest.raises(hip.HipError, match="pLinear"):
        _load(lin)
    deep = exr_multipart([(y, NONE, None, "d")]).replace(b"scanlineimage", b"deepscanline\0")
    with pytest.raises(hip.HipError):
        _load(deep)


def _fuzz_corpus(rng, files, mutations):
    out = bytearray()
    for f in files:
        recs = [f] + [f[:c] for c in sorted(set(rng.integers(0, len(f), 40).tolist()))]
        for _ in range(mutations):
            b = bytearray(f)
            for p in rng.integers(0, len(b), rng.integers(1, 5)):
                b[p] = rng.integers(0, 256) if rng.random() < 0.7 else b[p] ^ (1 << rng.integers(0, 8))
            recs.append(bytes(b))
        for r in recs:
            out += struct.pack("<I", len(r)) + r
    return bytes(out)


def test_decoders_under_sanitizers(tmp_path):
    """Every codec's files, truncated and with random byte and bit
    corruption, decoded by the library's image decoder built with
    AddressSanitizer and UndefinedBehaviorSanitizer on the host: no
    out-of-range access, no undefined behaviour, whatever the input."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "image_fuzz"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I" + os.path.join(root, "include"), "-o", str(exe),
                    os.path.join(root, "shiro_amd", "csrc", "image_io.cpp"),
                    os.path.join(root, "tests", "native", "image_fuzz.cpp"), "-lz"], check=True)
    rng = np.random.default_rng(90)
    img = np.exp(0.3 * rng.normal(size=(21, 19, 3))).astype(np.float32)
    h16 = {n: (HALF, img[..., i].astype(np.float16)) for i, n in enumerate("RGB")}
    f32 = {n: (FLOAT, img[..., i]) for i, n in enumerate("RGB")}
    mixed = {"R": (HALF, img[..., 0].astype(np.float16)), "G": (FLOAT, img[..., 1]),
             "B": (UINT, (img[..., 2] * 100).astype(np.uint32))}
    q = rng.integers(0, 256, (12, 17, 4)).astype(np.uint8)
    files = [pfm_bytes(img), rgbe_bytes(q, rle=True)]
    for comp in (RLE, ZIP, PIZ, PXR24, B44, B44A):
        files.append(exr_file(mixed if comp != B44 else h16, comp))
    files += [exr_file(f32, PIZ, tile=(8, 8), level_mode=1), exr_file(h16, B44A, tile=(16, 4)),
              exr_multipart([(h16, PIZ, None, "a"), (f32, ZIPS, (8, 8), "b")]),
              exr_multipart([(f32, PXR24, (4, 8), "a")])]
    corpus = tmp_path / "corpus.bin"
    corpus.write_bytes(_fuzz_corpus(rng, files, 400))
    r = subprocess.run([str(exe), str(corpus)], capture_output=True, text=True, timeout=600,
                       env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert r.stdout.startswith("records")
