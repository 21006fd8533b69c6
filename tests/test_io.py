"""Image files either side of the path (io.rs) through the library's C ABI:
the decoders against files this test encodes itself (its own PNG row filters,
TIFF layouts, an LZW coder with the TIFF code-width convention, the horizontal
predictor), the reference's committed PNG fixture against a decoder written
here, and the writers against this test's readers.  Host code only: no GPU."""
import struct
import zlib

import numpy as np
import pytest

import lapstack_amd as la

GOLDEN = "tests/golden/reference_expected_depth.png"


# ---------------------------------------------------------------- a PNG coder of the test's own
def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter_rows(raw, row, bpp, h):
    """Each row with filter type y % 5 (ISO/IEC 15948 section 9), written from the definitions."""
    out = bytearray()
    prev = bytes(row)
    for y in range(h):
        cur = raw[y * row:(y + 1) * row]
        ft = y % 5
        f = bytearray(row)
        for i in range(row):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][ft]
            f[i] = (cur[i] - pred) & 0xFF
        out.append(ft)
        out += f
        prev = cur
    return bytes(out)


def _png(samples, bits, ctype):
    h, w = samples.shape[:2]
    ch = {0: 1, 2: 3, 4: 2, 6: 4}[ctype]
    arr = samples.reshape(h, w * ch)
    raw = arr.astype(">u2").tobytes() if bits == 16 else arr.astype(np.uint8).tobytes()
    bpp = ch * bits // 8

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    ihdr = struct.pack(">IIBBBBB", w, h, bits, ctype, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr)
            + chunk(b"IDAT", zlib.compress(_filter_rows(raw, w * bpp, bpp, h), 6)) + chunk(b"IEND", b""))


def _png_read(path):
    """A PNG reader of the test's own (8/16-bit, colour types 0/2/4/6, not interlaced)."""
    b = open(path, "rb").read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    pos, z = 8, b""
    w = h = bits = ctype = None
    while pos < len(b):
        n, = struct.unpack(">I", b[pos:pos + 4])
        t, d = b[pos + 4:pos + 8], b[pos + 8:pos + 8 + n]
        if t == b"IHDR":
            w, h, bits, ctype = struct.unpack(">IIBB", d[:10])
        elif t == b"IDAT":
            z += d
        pos += 12 + n
    ch = {0: 1, 2: 3, 4: 2, 6: 4}[ctype]
    bpp = ch * bits // 8
    row = w * bpp
    raw = zlib.decompress(z)
    out = bytearray()
    prev = bytearray(row)
    for y in range(h):
        ft = raw[y * (row + 1)]
        cur = bytearray(raw[y * (row + 1) + 1:(y + 1) * (row + 1)])
        for i in range(row):
            a = cur[i - bpp] if i >= bpp else 0
            bb = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            cur[i] = (cur[i] + [0, a, bb, (a + bb) // 2, _paeth(a, bb, c)][ft]) & 0xFF
        out += cur
        prev = cur
    dt = ">u2" if bits == 16 else np.uint8
    return np.frombuffer(bytes(out), dt).reshape(h, w, ch).astype(np.uint16), bits


def _rgb16(samples, bits, ctype):
    """The decoder's expected output: RGB u16, gray replicated, alpha dropped, 8-bit times 257."""
    s = samples.astype(np.uint32) * (257 if bits == 8 else 1)
    rgb = np.repeat(s[..., :1], 3, axis=2) if ctype in (0, 4) else s[..., :3]
    return rgb.astype(np.uint16)


@pytest.mark.parametrize("bits", [8, 16])
@pytest.mark.parametrize("ctype", [0, 2, 4, 6])
def test_png_decodes_every_filter_and_colour_type(tmp_path, bits, ctype):
    rng = np.random.default_rng(bits * 10 + ctype)
    ch = {0: 1, 2: 3, 4: 2, 6: 4}[ctype]
    h, w = 23, 37
    s = rng.integers(0, 1 << bits, (h, w, ch))
    p = tmp_path / "f.png"
    p.write_bytes(_png(s, bits, ctype))
    got, b = la.load_rgb(str(p))
    assert b == bits and np.array_equal(got, _rgb16(s, bits, ctype))
    assert la.image_info(str(p)) == (w, h, bits)


def test_the_reference_png_fixture_decodes_like_the_test_reader():
    got, bits = la.load_rgb(GOLDEN)
    want, wb = _png_read(GOLDEN)
    assert bits == wb == 8 and got.shape == (768, 1024, 3)
    assert np.array_equal(got, np.repeat(want, 3, axis=2) * 257)


# ---------------------------------------------------------------- TIFF files of the test's own
def _lzw(data):
    """LZW with the TIFF conventions (clear 256, end 257, 9-12 bit codes MSB
    first, the width growing one code early), written from the specification."""
    out_bits = []
    table = {bytes([i]): i for i in range(256)}
    nxt, width = 258, 9

    def put(code):
        out_bits.extend((code >> (width - 1 - i)) & 1 for i in range(width))

    put(256)
    w = b""
    for byte in data:
        wc = w + bytes([byte])
        if wc in table:
            w = wc
            continue
        put(table[w])
        table[wc] = nxt
        nxt += 1
        if nxt + 1 > (1 << width) and width < 12:
            width += 1
        if nxt >= 4094:
            put(256)
            table = {bytes([i]): i for i in range(256)}
            nxt, width = 258, 9
        w = bytes([byte])
    if w:
        put(table[w])
    put(257)
    out_bits.extend([0] * (-len(out_bits) % 8))
    return np.packbits(np.array(out_bits, np.uint8)).tobytes()


def _tiff(samples, bits, comp=1, pred=1, planar=1, le=True, rows_per_strip=7):
    """A baseline TIFF: strips of rows_per_strip rows, one IFD after the data."""
    h, w, ch = samples.shape
    e = "<" if le else ">"
    dt = np.dtype(e + "u2") if bits == 16 else np.dtype(np.uint8)
    planes = [samples[..., c:c + 1] for c in range(ch)] if planar == 2 else [samples]
    strips = []
    for pl in planes:
        a = pl.astype(np.int64)
        if pred == 2:
            a = a.copy()
            a[:, 1:, :] = (pl[:, 1:, :].astype(np.int64) - pl[:, :-1, :].astype(np.int64)) % (1 << bits)
        a = a.astype(dt)
        for y0 in range(0, h, rows_per_strip):
            raw = a[y0:y0 + rows_per_strip].tobytes()
            strips.append(raw if comp == 1 else (zlib.compress(raw) if comp == 8 else _lzw(raw)))
    body = b"".join(strips)
    offs, o = [], 8
    for s_ in strips:
        offs.append(o)
        o += len(s_)
    ifd_at = 8 + len(body)
    ents = [(256, 4, [w]), (257, 4, [h]), (258, 3, [bits] * ch), (259, 3, [comp]), (262, 3, [2 if ch >= 3 else 1]),
            (273, 4, offs), (277, 3, [ch]), (278, 4, [rows_per_strip]), (279, 4, [len(s_) for s_ in strips]),
            (284, 3, [planar]), (317, 3, [pred])]
    extra_at = ifd_at + 2 + 12 * len(ents) + 4
    ifd, extra = struct.pack(e + "H", len(ents)), b""
    for tag, typ, vals in ents:
        fmt = e + ("H" if typ == 3 else "I") * len(vals)
        data = struct.pack(fmt, *vals)
        if len(data) <= 4:
            ifd += struct.pack(e + "HHI", tag, typ, len(vals)) + data.ljust(4, b"\0")
        else:
            ifd += struct.pack(e + "HHII", tag, typ, len(vals), extra_at + len(extra))
            extra += data
    ifd += struct.pack(e + "I", 0)
    head = (b"II" if le else b"MM") + struct.pack(e + "HI", 42, ifd_at)
    return head + body + ifd + extra


@pytest.mark.parametrize("bits", [8, 16])
@pytest.mark.parametrize("comp,pred", [(1, 1), (8, 1), (8, 2), (5, 1), (5, 2)])
@pytest.mark.parametrize("le", [True, False])
def test_tiff_decodes_compressions_predictors_and_byte_orders(tmp_path, bits, comp, pred, le):
    rng = np.random.default_rng(bits + comp + pred + le)
    s = rng.integers(0, 1 << bits, (19, 29, 3))
    s[5:9] = s[5:6]  # repeated rows: LZW strings longer than one byte
    p = tmp_path / "f.tif"
    p.write_bytes(_tiff(s, bits, comp, pred, 1, le))
    got, b = la.load_rgb(str(p))
    assert b == bits and np.array_equal(got, _rgb16(s, bits, 2))


def test_tiff_planar_gray_and_alpha(tmp_path):
    rng = np.random.default_rng(3)
    s = rng.integers(0, 65536, (11, 13, 3))
    p = tmp_path / "planar.tif"
    p.write_bytes(_tiff(s, 16, 1, 1, planar=2))
    assert np.array_equal(la.load_rgb(str(p))[0], s.astype(np.uint16))
    g = rng.integers(0, 256, (11, 13, 1))
    p2 = tmp_path / "gray.tif"
    p2.write_bytes(_tiff(g, 8))
    assert np.array_equal(la.load_rgb(str(p2))[0], _rgb16(g, 8, 0))
    a = rng.integers(0, 65536, (11, 13, 4))
    p3 = tmp_path / "rgba.tif"
    p3.write_bytes(_tiff(a, 16, 8, 2))
    assert np.array_equal(la.load_rgb(str(p3))[0], a[..., :3].astype(np.uint16))


def test_lzw_long_run_crosses_every_code_width(tmp_path):
    # enough distinct strings to pass 511, 1023, 2047 and a clear code
    rng = np.random.default_rng(9)
    s = rng.integers(0, 65536, (64, 200, 3))
    s[:, 100:] = s[:, :100]
    p = tmp_path / "long.tif"
    p.write_bytes(_tiff(s, 16, 5, 2, rows_per_strip=64))
    assert np.array_equal(la.load_rgb(str(p))[0], s.astype(np.uint16))


def test_bad_files_are_errors_not_crashes(tmp_path):
    p = tmp_path / "t.png"
    full = _png(np.zeros((8, 8, 3), np.int64), 16, 2)
    p.write_bytes(full[:40])
    with pytest.raises(la.HipError, match="cannot decode"):
        la.load_rgb(str(p))
    q = tmp_path / "x.bin"
    q.write_bytes(b"hello world")
    with pytest.raises(la.HipError, match="neither PNG nor TIFF"):
        la.load_rgb(str(q))
    with pytest.raises(la.HipError, match="cannot open"):
        la.load_rgb(str(tmp_path / "missing.tif"))


# ---------------------------------------------------------------- the writers (io.rs:273-346)
def _tiff_read(path):
    """A reader of the writer's own TIFF layout (little-endian, chunky, strips)."""
    b = open(path, "rb").read()
    assert b[:4] == b"II*\0"
    ifd, = struct.unpack("<I", b[4:8])
    n, = struct.unpack("<H", b[ifd:ifd + 2])
    tags = {}
    for i in range(n):
        tag, typ, cnt = struct.unpack("<HHI", b[ifd + 2 + 12 * i:ifd + 10 + 12 * i])
        sz = {3: 2, 4: 4, 2: 1}[typ]
        at = ifd + 10 + 12 * i if cnt * sz <= 4 else struct.unpack("<I", b[ifd + 10 + 12 * i:ifd + 14 + 12 * i])[0]
        fmt = {3: "H", 4: "I", 2: "B"}[typ]
        tags[tag] = struct.unpack("<" + fmt * cnt, b[at:at + cnt * sz])
    w, h, bits = tags[256][0], tags[257][0], tags[258][0]
    data = b"".join(b[o:o + c] for o, c in zip(tags[273], tags[279]))
    assert tags[259] == (1,) and tags[262] == (2,) and tags[277] == (3,) and tags[284] == (1,)
    return np.frombuffer(data, "<u2" if bits == 16 else np.uint8).reshape(h, w, 3), bits, tags


def _quant(img3, bits):
    m = 65535.0 if bits == 16 else 255.0
    q = np.clip(img3, 0, 1).astype(np.float32) * np.float32(m) + np.float32(0.5)
    return np.minimum(q, np.float32(m)).astype(np.uint32).transpose(1, 2, 0)


@pytest.mark.parametrize("ext,bits", [("tif", 16), ("tiff", 8), ("png", 16), ("png", 8)])
def test_save_rgb_writes_the_reference_quantisation(tmp_path, ext, bits):
    rng = np.random.default_rng(bits)
    img = rng.uniform(-0.1, 1.1, (3, 21, 34)).astype(np.float32)
    img[0, 0, :4] = [0.0, 1.0, 0.5, 1.0 / 65535.0]
    p = str(tmp_path / f"out.{ext}")
    la.save_rgb(img, p, bits)
    got, b = (_tiff_read(p)[:2] if ext.startswith("tif") else _png_read(p))
    assert b == bits and np.array_equal(got.astype(np.uint32), _quant(img, bits))
    back, bb = la.load_rgb(p)
    assert bb == bits and np.array_equal(back, got.astype(np.uint16) * (257 if bits == 8 else 1))


def test_tiff_writer_strips_of_about_8mb(tmp_path):
    img = np.random.default_rng(1).random((3, 1500, 1000)).astype(np.float32)  # 9 MB of 16-bit samples
    p = str(tmp_path / "big.tif")
    la.save_rgb(img, p, 16)
    got, _, tags = _tiff_read(p)
    assert tags[278][0] == (8 << 20) // (1000 * 6) and len(tags[273]) == 2
    assert np.array_equal(got.astype(np.uint32), _quant(img, 16))


def test_save_rgb_u16_and_the_gray_writers(tmp_path):
    rng = np.random.default_rng(4)
    rgb = rng.integers(0, 65536, (9, 14, 3)).astype(np.uint16)
    p = str(tmp_path / "u16.png")
    la.save_rgb_u16(rgb, p)
    assert np.array_equal(_png_read(p)[0], rgb)
    d = rng.uniform(-1, 9, (9, 14)).astype(np.float32)
    p16 = str(tmp_path / "d16.png")
    la.save_gray16(d, p16, 7.0)
    k = np.float32(65535.0) / np.float32(7.0)
    want = np.minimum(np.clip(d, 0, 7).astype(np.float32) * k + np.float32(0.5), 65535).astype(np.uint16)
    assert np.array_equal(_png_read(p16)[0][..., 0], want)
    p8 = str(tmp_path / "d8.png")
    la.save_gray(d, p8)
    lo, hi = d.min(), d.max()
    want8 = (((d - lo) / np.float32(max(hi - lo, 1e-6))) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    assert np.array_equal(_png_read(p8)[0][..., 0], want8)
    with pytest.raises(la.HipError, match="PNG or TIFF"):
        la.save_rgb(np.zeros((3, 2, 2), np.float32), str(tmp_path / "x.jpg"), 8)
