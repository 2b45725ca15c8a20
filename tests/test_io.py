"""File formats of the task queue's callbacks and readers (host code, no GPU):
readPNG / writeOutPNG (readInFuncs.cpp:108-157, writeOutFuncs.cpp:69-114),
readTiff (readInFuncs.cpp:160-195, ReadTiff.cpp) and readTxt
(readInFuncs.cpp:56-104) through the C ABI.

Pins: the reference's own round-trip test (testReadWrite.cpp:55-185: random
data from srand(876512789), log-spaced square sizes, repeated write/read
must not degrade), an independent PNG decoder/encoder written here with
zlib and numpy (every colour type, bit depth and row filter), the
reference's 8-bit test image (tests/golden/lattice-aspect-ratio-2.png, from
examples/testData) and, where the reference is mounted, its 16-bit one.
"""
import math
import os
import struct
import zlib

import numpy as np
import pytest

import imresh_amd as im
from helpers import RAND_MAX, log_spaced_points, rand, srand

L = im.lib
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
REF_PNG16 = "/root/reference/examples/testData/imresh.png"


# ---------------------------------------------------------------- an independent PNG codec
def _chunks(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    p, out = 8, []
    while p < len(data):
        n = struct.unpack(">I", data[p:p + 4])[0]
        t, body = data[p + 4:p + 8], data[p + 8:p + 8 + n]
        assert struct.unpack(">I", data[p + 8 + n:p + 12 + n])[0] == zlib.crc32(t + body)
        out.append((t, body))
        p += 12 + n
    return out


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def decode_png(path):
    """Red channel as the library defines it: 16-bit value (8-bit x 256),
    palette looked up, gray as red; returned top row first (file order)."""
    data = open(path, "rb").read()
    ch = _chunks(data)
    w, h, depth, ctype, _, _, il = struct.unpack(">IIBBBBB", ch[0][1])
    assert il == 0
    plte = b"".join(b for t, b in ch if t == b"PLTE")
    raw = zlib.decompress(b"".join(b for t, b in ch if t == b"IDAT"))
    nch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bits = nch * depth
    rb = (w * bits + 7) // 8
    bpp = max(1, bits // 8)
    prev = bytearray(rb)
    out = np.zeros((h, w), np.int64)
    for r in range(h):
        f, line = raw[r * (rb + 1)], bytearray(raw[r * (rb + 1) + 1:(r + 1) * (rb + 1)])
        for i in range(rb):
            a = line[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            line[i] = (line[i] + [0, a, b, (a + b) // 2, _paeth(a, b, c)][f]) & 0xFF
        if depth == 16:
            v = np.frombuffer(bytes(line), ">u2").reshape(w, nch)[:, 0].astype(np.int64)
        elif depth == 8 and ctype != 3:
            v = np.frombuffer(bytes(line), np.uint8).reshape(w, nch)[:, 0].astype(np.int64) * 256
        else:
            bitstr = np.unpackbits(np.frombuffer(bytes(line), np.uint8))
            s = np.array([int("".join(map(str, bitstr[x * depth:(x + 1) * depth])), 2) for x in range(w)])
            if ctype == 3:
                v = np.array([plte[3 * k] for k in s], np.int64) * 256
            else:
                v = s * (255 // ((1 << depth) - 1)) * 256
        out[r] = v
        prev = line
    return out


def encode_png(path, samples, depth, ctype, plte=None, filters=(0, 1, 2, 3, 4), rng=None):
    """samples: h x w x channels integer array (palette: indices)."""
    h, w, nch = samples.shape
    rows = []
    for r in range(h):
        if depth == 16:
            line = samples[r].astype(">u2").tobytes()
        elif depth == 8:
            line = samples[r].astype(np.uint8).tobytes()
        else:
            bits = "".join(format(int(x), f"0{depth}b") for x in samples[r].reshape(-1))
            bits += "0" * (-len(bits) % 8)
            line = bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))
        rows.append(bytearray(line))
    bpp = max(1, nch * depth // 8)
    raw = bytearray()
    prev = bytearray(len(rows[0]))
    for r, line in enumerate(rows):
        f = filters[r % len(filters)] if rng is None else int(rng.choice(filters))
        enc = bytearray(len(line))
        for i in range(len(line)):
            a = line[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            enc[i] = (line[i] - [0, a, b, (a + b) // 2, _paeth(a, b, c)][f]) & 0xFF
        raw += bytes([f]) + enc
        prev = line

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body))

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if plte is not None:
        data += chunk(b"PLTE", bytes(plte))
    data += chunk(b"tEXt", b"Comment\x00ancillary chunks are skipped")
    data += chunk(b"IDAT", zlib.compress(bytes(raw), 9)) + chunk(b"IEND", b"")
    open(path, "wb").write(data)


def read(path, kind="png"):
    fn = getattr(L(), f"imresh_read_{kind}")
    w, h = im.ctypes.c_uint32(), im.ctypes.c_uint32()
    st = fn(os.fsencode(path), None, 0, im.ctypes.byref(w), im.ctypes.byref(h))
    if st:
        return st, None
    out = np.empty((h.value, w.value), np.float32)
    st = fn(os.fsencode(path), out.ctypes.data, out.size, im.ctypes.byref(w), im.ctypes.byref(h))
    return st, out


# ---------------------------------------------------------------- PNG
def test_png_round_trip_like_the_reference(tmp_path):
    """testReadWrite.cpp:55-185: write random data, read it back; further
    write/read rounds reproduce the first read exactly.  The first read is
    the 16-bit quantisation int(v/max*65535 + 0.5f)/65535 in float."""
    p = str(tmp_path / "blfh8921721afh")
    for n in log_spaced_points(2, 1024 * 1024, 20):
        nx = int(math.floor(math.sqrt(n)))
        if nx == 1:
            continue
        srand(876512789)
        data = np.array([np.float32(rand()) / np.float32(RAND_MAX) for _ in range(nx * nx)],
                        np.float32).reshape(nx, nx)
        im.check(L().imresh_write_png(os.fsencode(p), data.ctypes.data, nx, nx))
        st, first = read(p)
        assert st == 0 and first.shape == (nx, nx)
        mx = data.max()
        q = np.floor((data / mx) * np.float32(65535) + np.float32(0.5)).astype(np.int64)
        assert np.array_equal(first, (q.astype(np.float32) / np.float32(65535)).astype(np.float32))
        for _ in range(2):
            im.check(L().imresh_write_png(os.fsencode(p), first.ctypes.data, nx, nx))
            st, again = read(p)
            assert st == 0 and np.array_equal(again, first), nx


def test_png_file_layout_and_scaling(tmp_path):
    """The written file decodes (independently) as 16-bit RGB with R = G = B,
    the buffer's last row on top (pngwriter's bottom-left origin), negatives
    and NaN clamped to 0, an all-zero image to 0."""
    p = str(tmp_path / "a.png")
    a = np.arange(12, dtype=np.float32).reshape(3, 4) - 2
    a[1, 1] = np.nan
    im.write_png(p, a)
    ch = _chunks(open(p, "rb").read())
    w, h, depth, ctype = struct.unpack(">IIBB", ch[0][1][:10])
    assert (w, h, depth, ctype) == (4, 3, 16, 2)
    raw = zlib.decompress(b"".join(b for t, b in ch if t == b"IDAT"))
    px = np.frombuffer(b"".join(raw[r * 25 + 1:(r + 1) * 25] for r in range(3)), ">u2").reshape(3, 4, 3)
    assert np.all(px[..., 0] == px[..., 1]) and np.all(px[..., 1] == px[..., 2])
    exp = np.floor(np.nan_to_num(a / np.float32(9)) * np.float32(65535) + np.float32(0.5))
    exp = np.clip(np.where(np.isnan(a), 0, exp), 0, 65535)
    assert np.array_equal(px[..., 0][::-1].astype(np.int64), exp.astype(np.int64))
    im.write_png(p, np.zeros((2, 2), np.float32))
    assert not im.read_image(p).any()


@pytest.mark.parametrize("depth,ctype", [(1, 0), (2, 0), (4, 0), (8, 0), (16, 0), (1, 3), (2, 3), (4, 3),
                                         (8, 3), (8, 2), (16, 2), (8, 4), (16, 4), (8, 6), (16, 6)])
def test_png_reader_all_layouts(tmp_path, depth, ctype):
    rng = np.random.default_rng(depth * 10 + ctype)
    h, w = 7, 13
    nch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    top = (1 << depth) - 1
    plte = None
    if ctype == 3:
        plte = rng.integers(0, 256, 3 * (top + 1)).astype(np.uint8)
    s = rng.integers(0, top + 1, (h, w, nch))
    p = str(tmp_path / "t.png")
    encode_png(p, s, depth, ctype, plte, rng=rng)
    st, got = read(p)
    assert st == 0
    exp = decode_png(p)
    assert np.array_equal(got, (exp[::-1].astype(np.float32) / np.float32(65535)).astype(np.float32))
    # and the expected red values follow from the samples directly
    if ctype == 3:
        red = plte.reshape(-1, 3)[s[..., 0], 0].astype(np.int64) * 256
    elif depth == 16:
        red = s[..., 0]
    elif depth == 8:
        red = s[..., 0] * 256
    else:
        red = s[..., 0] * (255 // top) * 256
    assert np.array_equal(exp, red)


def test_png_reference_test_images():
    """examples/testData: the 8-bit RGB lattice (fixture) and, when the
    reference is mounted, the 16-bit RGB imresh.png."""
    for path in (os.path.join(GOLDEN, "lattice-aspect-ratio-2.png"), REF_PNG16):
        if not os.path.exists(path):
            continue
        st, got = read(path)
        assert st == 0
        exp = decode_png(path)
        assert got.shape == exp.shape
        assert np.array_equal(got, (exp[::-1].astype(np.float32) / np.float32(65535)).astype(np.float32))
    st, lat = read(os.path.join(GOLDEN, "lattice-aspect-ratio-2.png"))
    assert lat.shape == (256, 512) and lat.max() == np.float32(255 * 256) / np.float32(65535)


def test_png_errors(tmp_path):
    p = str(tmp_path / "t.png")
    encode_png(p, np.zeros((4, 4, 1), np.int64), 8, 0)
    data = bytearray(open(p, "rb").read())
    bad = str(tmp_path / "bad.png")
    for mutate in (lambda d: d.__setitem__(20, d[20] ^ 1),   # IHDR CRC
                   lambda d: d.__delitem__(slice(len(d) - 20, None)),   # truncated
                   lambda d: d.__setitem__(0, 0)):   # signature
        d = bytearray(data)
        mutate(d)
        open(bad, "wb").write(bytes(d))
        assert read(bad)[0] == 10   # IMRESH_ERR_IO
    assert read(str(tmp_path / "missing.png"))[0] == 10
    # interlaced images are rejected
    d = bytearray(data)
    d[28] = 1
    d[29:33] = struct.pack(">I", zlib.crc32(bytes(d[12:29])))
    open(bad, "wb").write(bytes(d))
    assert read(bad)[0] == 10
    w, h = im.ctypes.c_uint32(), im.ctypes.c_uint32()
    small = np.empty(3, np.float32)
    assert L().imresh_read_png(os.fsencode(p), small.ctypes.data, 3, im.ctypes.byref(w), im.ctypes.byref(h)) == 2
    assert (w.value, h.value) == (4, 4)
    assert L().imresh_write_png(os.fsencode(str(tmp_path / "no" / "dir.png")), small.ctypes.data, 1, 3) == 10
    assert L().imresh_write_png(None, small.ctypes.data, 1, 3) == 1


# ---------------------------------------------------------------- TIFF
def write_tiff(path, a, big_endian=False, rows_per_strip=3, bits=32, fmt=3, comp=1):
    e = ">" if big_endian else "<"
    h, w = a.shape
    body = a.astype(e + "f4").tobytes() if bits == 32 else a.astype(e + "u2").tobytes()
    row = w * bits // 8
    strips = [body[r * row:min(h, r + rows_per_strip) * row] for r in range(0, h, rows_per_strip)]
    data = bytearray(b"MM" if big_endian else b"II") + struct.pack(e + "HI", 42, 0)
    offs = []
    for s in strips:
        offs.append(len(data))
        data += s
    arr_off = len(data)
    data += struct.pack(e + "%dI" % len(offs), *offs)
    cnt_off = len(data)
    data += struct.pack(e + "%dI" % len(strips), *[len(s) for s in strips])
    if len(data) % 2:
        data += b"\0"
    ifd = len(data)
    entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, bits), (259, 3, 1, comp), (262, 3, 1, 1),
               (273, 4, len(offs), arr_off if len(offs) > 1 else offs[0]), (277, 3, 1, 1),
               (278, 4, 1, rows_per_strip),
               (279, 4, len(strips), cnt_off if len(strips) > 1 else len(strips[0])), (339, 3, 1, fmt)]
    data += struct.pack(e + "H", len(entries))
    for tag, typ, n, v in entries:
        if typ == 3 and n == 1:
            data += struct.pack(e + "HHIHH", tag, typ, n, v, 0)
        else:
            data += struct.pack(e + "HHII", tag, typ, n, v)
    data += struct.pack(e + "I", 0)
    data[4:8] = struct.pack(e + "I", ifd)
    open(path, "wb").write(bytes(data))


@pytest.mark.parametrize("big_endian,rps", [(False, 3), (True, 3), (False, 100), (True, 1)])
def test_tiff_float_reader(tmp_path, big_endian, rps):
    rng = np.random.default_rng(rps)
    a = (rng.standard_normal((10, 7)) * 1e3).astype(np.float32)
    a[0, 0] = -0.0
    a[1, 1] = np.inf
    p = str(tmp_path / "t.tif")
    write_tiff(p, a, big_endian, rps)
    st, got = read(p, "tiff")
    assert st == 0 and got.shape == a.shape
    assert np.array_equal(got.view(np.uint32), a.view(np.uint32))
    assert np.array_equal(im.read_image(p).view(np.uint32), a.view(np.uint32))


def _peak_rss_mib():
    import resource
    return resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1024


def test_corrupt_sizes_rejected_before_allocating(tmp_path):
    """A header that claims 40000 x 40000 pixels over a few bytes of data is
    rejected at once: the PNG reader bounds the inflated size by deflate's
    1032:1 ratio before allocating it, the TIFF reader checks the strips
    hold W x H floats before reserving them (a 12.8 GB zero-filled buffer
    otherwise)."""
    import time
    p = str(tmp_path / "big.png")

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body))

    ihdr = struct.pack(">IIBBBBB", 40000, 40000, 16, 6, 0, 0, 0)
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(b"\0" * 64)) + chunk(b"IEND", b"")
    open(p, "wb").write(data)
    rss0, t0 = _peak_rss_mib(), time.perf_counter()
    assert read(p)[0] == 10
    t = str(tmp_path / "big.tif")
    write_tiff(t, np.ones((2, 2), np.float32), rows_per_strip=2)
    d = bytearray(open(t, "rb").read())
    ifd = struct.unpack("<I", d[4:8])[0]
    for i in range(struct.unpack("<H", d[ifd:ifd + 2])[0]):
        e = ifd + 2 + 12 * i
        if struct.unpack("<H", d[e:e + 2])[0] in (256, 257):   # width, height
            d[e + 8:e + 12] = struct.pack("<I", 40000)
    open(t, "wb").write(bytes(d))
    assert read(t, "tiff")[0] == 10
    assert time.perf_counter() - t0 < 2.0
    assert _peak_rss_mib() - rss0 < 512


def test_tiff_unsupported_layouts(tmp_path):
    p = str(tmp_path / "t.tif")
    write_tiff(p, np.ones((4, 4), np.float32), bits=16, fmt=1)
    assert read(p, "tiff")[0] == 10
    write_tiff(p, np.ones((4, 4), np.float32), comp=5)
    assert read(p, "tiff")[0] == 10
    open(p, "wb").write(b"II*\0garbage")
    assert read(p, "tiff")[0] == 10


# ---------------------------------------------------------------- text
def test_txt_reader(tmp_path):
    p = str(tmp_path / "a.txt")
    open(p, "w").write("1 2 3.5\n-4 5e2 6\n7 8 9\n")
    st, got = read(p, "txt")
    assert st == 0
    assert np.array_equal(got, np.array([[1, 2, 3.5], [-4, 500, 6], [7, 8, 9]], np.float32))
    open(p, "w").write("1 2\n3 4\n5\n")   # height = count / width (truncating)
    assert read(p, "txt")[1].shape == (2, 2)
    open(p, "w").write("1 x 3\n")
    assert read(p, "txt")[0] == 10
