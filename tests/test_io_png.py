"""PNG files (include/nli_io.h) on the CPU: the reader against PIL, an independent PNG
decoder, on files PIL writes (every colour type, palette bit depths, tRNS) and on files
written here with every row filter, Adam7 interlacing and 16-bit samples; the reader's
expansion rules of iio/libpng (io_utility.cpp:119-168, iio.c:1071-1130); the writer's
rounding (iio.c:676) read back through PIL; the reference's data/*.png where the
reference sources are present."""
import os
import struct
import zlib

import numpy as np
import pytest

PIL = pytest.importorskip("PIL.Image")
REF_DATA = "/root/reference/data"


def pil_array(path, mode=None):
    im = PIL.open(path)
    if mode:
        im = im.convert(mode)
    a = np.asarray(im)
    return a[:, :, None] if a.ndim == 2 else a


# ---- a small PNG writer with every row filter and Adam7, for the reader's tests ----
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def _chunk(t, d):
    return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xffffffff)


def _pack_row(samples, depth):
    """samples: 1-D ints of one row (all channels interleaved)."""
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in samples)
    if depth == 8:
        return bytes(int(v) for v in samples)
    bits = "".join(format(int(v), f"0{depth}b") for v in samples)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


def _filter_row(row, prev, bpp, t):
    out = bytearray(len(row))
    for i in range(len(row)):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        if t == 0:
            pred = 0
        elif t == 1:
            pred = a
        elif t == 2:
            pred = b
        elif t == 3:
            pred = (a + b) >> 1
        else:
            p = a + b - c
            pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
            pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
        out[i] = (row[i] - pred) & 0xff
    return bytes(out)


def write_png(path, samples, ctype, depth, interlace=False, rng=None, plte=None, trns=None, extra=b""):
    """samples: (H, W, channels) integers of the stored colour type (palette indices for
    type 3).  Row filters are drawn at random from all five."""
    rng = rng or np.random.default_rng(0)
    H, W, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    raw = bytearray()
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        prev = bytes(len(_pack_row(sub[0].reshape(-1), depth)))
        for r in range(sub.shape[0]):
            row = _pack_row(sub[r].reshape(-1), depth)
            t = int(rng.integers(0, 5))
            raw += bytes([t]) + _filter_row(row, prev, bpp, t)
            prev = row
    ihdr = struct.pack(">IIBBBBB", W, H, depth, ctype, 0, 0, 1 if interlace else 0)
    data = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr)
    if plte is not None:
        data += _chunk(b"PLTE", bytes(np.asarray(plte, np.uint8).reshape(-1)))
    if trns is not None:
        data += _chunk(b"tRNS", trns)
    data += extra + _chunk(b"IDAT", zlib.compress(bytes(raw), 9)) + _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


CTYPE_CH = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("ctype,depth", [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2),
                                         (3, 4), (3, 8), (4, 8), (4, 16), (6, 8), (6, 16)])
def test_reader_every_colour_type_depth_filter_interlace(nli, tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth + 7 * interlace)
    H, W = 13, 11   # ragged against the 8x8 Adam7 grid; some passes are 1 pixel wide
    ch = CTYPE_CH[ctype]
    top = (1 << depth) - 1
    plte = None
    if ctype == 3:
        n = 1 << depth
        plte = rng.integers(0, 256, (n, 3))
    s = rng.integers(0, top + 1, (H, W, ch))
    path = str(tmp_path / f"t{ctype}_{depth}_{int(interlace)}.png")
    write_png(path, s, ctype, depth, interlace, rng, plte=plte)
    got = nli.read_image(path)
    if ctype == 3:
        want = plte[s[:, :, 0]]
    elif ctype == 0 and depth < 8:
        want = s * (255 // top)
    else:
        want = s
    assert nli.image_info(path) == (W, H, want.shape[2], 16 if depth == 16 else 8)
    np.testing.assert_array_equal(got, want.astype(np.float32))
    if depth == 8 or ctype == 3 or (ctype == 0 and depth < 8):
        # PIL decodes the same file: pins this test's writer and the reader together.
        mode = {0: "L", 2: "RGB", 3: "RGB", 4: "LA", 6: "RGBA"}[ctype]
        np.testing.assert_array_equal(got, pil_array(path, mode).astype(np.float32))


def test_reader_trns_becomes_alpha(nli, tmp_path):
    rng = np.random.default_rng(3)
    # grey with a tRNS key: alpha 0 on the key value, 255 elsewhere; the key keeps the
    # low `depth` bits of its 16-bit field
    g = rng.integers(0, 16, (9, 10, 1))
    p = str(tmp_path / "g4.png")
    write_png(p, g, 0, 4, trns=struct.pack(">H", 0x0105))
    got = nli.read_image(p)
    np.testing.assert_array_equal(got[:, :, 0], g[:, :, 0] * 17)
    np.testing.assert_array_equal(got[:, :, 1], np.where(g[:, :, 0] == 5, 0, 255))
    # RGB 16-bit with a key colour
    c = rng.integers(0, 3, (7, 8, 3)) * 20000
    key = (20000, 0, 40000)
    p = str(tmp_path / "rgb16.png")
    write_png(p, c, 2, 16, trns=struct.pack(">HHH", *key))
    got = nli.read_image(p)
    assert got.shape == (7, 8, 4)
    np.testing.assert_array_equal(got[:, :, :3], c)
    hit = np.all(c == np.array(key), axis=2)
    np.testing.assert_array_equal(got[:, :, 3], np.where(hit, 0, 65535))
    # palette with a short alpha table: entries beyond it are opaque
    idx = rng.integers(0, 4, (6, 5, 1))
    plte = [[10, 20, 30], [40, 50, 60], [70, 80, 90], [100, 110, 120]]
    p = str(tmp_path / "p2.png")
    write_png(p, idx, 3, 2, plte=plte, trns=bytes([0, 128]))
    got = nli.read_image(p)
    np.testing.assert_array_equal(got[:, :, :3], np.array(plte)[idx[:, :, 0]])
    np.testing.assert_array_equal(got[:, :, 3], np.array([0, 128, 255, 255])[idx[:, :, 0]])
    np.testing.assert_array_equal(got, pil_array(p, "RGBA").astype(np.float32))


@pytest.mark.parametrize("mode", ["1", "L", "LA", "RGB", "RGBA", "P", "I;16"])
def test_reader_matches_pil_written_files(nli, tmp_path, mode):
    rng = np.random.default_rng(11)
    H, W = 37, 53
    if mode == "I;16":
        im = PIL.fromarray(rng.integers(0, 65536, (H, W)).astype(np.uint16))
    else:
        base = PIL.fromarray(rng.integers(0, 256, (H, W, 4)).astype(np.uint8), "RGBA")
        im = base.convert(mode) if mode != "P" else base.convert("RGB").quantize(37)
    path = str(tmp_path / "pil.png")
    im.save(path)
    got = nli.read_image(path)
    if mode == "I;16":
        want = np.asarray(im).astype(np.float32)[:, :, None]
    else:
        want = pil_array(path, {"1": "L", "P": "RGB"}.get(mode, mode)).astype(np.float32)
    np.testing.assert_array_equal(got, want)


def test_pil_transparency_key_files(nli, tmp_path):
    rng = np.random.default_rng(5)
    g = rng.integers(0, 4, (20, 30)).astype(np.uint8) * 60
    p = str(tmp_path / "lt.png")
    PIL.fromarray(g, "L").save(p, transparency=120)
    got = nli.read_image(p)
    np.testing.assert_array_equal(got[:, :, 0], g)
    np.testing.assert_array_equal(got[:, :, 1], np.where(g == 120, 0, 255))
    c = rng.integers(0, 2, (20, 30, 3)).astype(np.uint8) * 200
    p = str(tmp_path / "rgbt.png")
    PIL.fromarray(c, "RGB").save(p, transparency=(200, 0, 200))
    got = nli.read_image(p)
    np.testing.assert_array_equal(got[:, :, :3], c)
    np.testing.assert_array_equal(got[:, :, 3], np.where(np.all(c == [200, 0, 200], axis=2), 0, 255))


def test_read_rgb_and_mono_rules(nli, tmp_path):
    rng = np.random.default_rng(9)
    H, W = 15, 17
    rgba = rng.integers(0, 256, (H, W, 4)).astype(np.uint8)
    for mode in ["L", "LA", "RGB", "RGBA"]:
        p = str(tmp_path / f"{mode}.png")
        PIL.fromarray(rgba, "RGBA").convert(mode).save(p)
        a = pil_array(p).astype(np.float64)
        rgb = nli.read_rgb_image(p)
        # io_utility.cpp:160-162: channel 1 from the second channel if there is one,
        # channel 2 from the third, otherwise from the first
        want = np.stack([a[..., 0], a[..., 1] if a.shape[2] > 1 else a[..., 0],
                         a[..., 2] if a.shape[2] > 2 else a[..., 0]], axis=2)
        np.testing.assert_array_equal(rgb, want.astype(np.float32))
        if mode == "LA":
            with pytest.raises(OSError, match="non-scalarizable"):
                nli.read_mono_image(p)
            continue
        mono = nli.read_mono_image(p)
        if a.shape[2] == 1:
            np.testing.assert_array_equal(mono, a[..., 0].astype(np.float32))
        else:
            luma = np.trunc(.299 * a[..., 0] + .587 * a[..., 1] + .114 * a[..., 2])
            np.testing.assert_array_equal(mono, luma.astype(np.float32))
    # the weights sum to 1 in exact arithmetic; a grey colour comes back as itself or
    # one below (truncation of the double sum), never above
    grey = np.repeat(np.arange(256, dtype=np.uint8).reshape(16, 16, 1), 3, axis=2)
    p = str(tmp_path / "grey_rgb.png")
    PIL.fromarray(grey, "RGB").save(p)
    mono = nli.read_mono_image(p)
    d = grey[..., 0].astype(np.float32) - mono
    assert set(np.unique(d)) <= {0.0, 1.0}
    assert mono[0, 0] == 0 and mono[15, 15] in (254.0, 255.0)
    # 16-bit colour cannot be made grey (iio refuses it)
    p = str(tmp_path / "rgb16.png")
    write_png(p, rng.integers(0, 65536, (4, 4, 3)), 2, 16)
    with pytest.raises(OSError, match="uncolorize"):
        nli.read_mono_image(p)
    # 16-bit grey: values as stored
    s = rng.integers(0, 65536, (4, 5, 1))
    p = str(tmp_path / "g16.png")
    write_png(p, s, 0, 16)
    np.testing.assert_array_equal(nli.read_mono_image(p), s[:, :, 0].astype(np.float32))


def t8(v):
    """iio.c:676: (uint8_t)T8(0.5 + v) with the sum in double; NaN gives 0."""
    t = 0.5 + np.asarray(v, np.float64)
    c = np.where(t > 0, np.where(t < 255, t, 255), 0)
    return np.trunc(c).astype(np.uint8)


@pytest.mark.parametrize("C", [1, 2, 3, 4])
def test_writer_rounding_read_back_by_pil(nli, tmp_path, C):
    rng = np.random.default_rng(C)
    H, W = 29, 31
    v = rng.uniform(-20, 280, (H, W, C)).astype(np.float32)
    # the rounding boundaries, NaN and infinities
    edge = np.array([-0.5, -0.4999, 0.0, 0.49999997, 0.5, 1.5, 2.5, 254.49998, 254.5, 255.0, 1e30, -1e30,
                     np.nan, np.inf, -np.inf], np.float32)
    v.reshape(-1)[:edge.size] = edge
    p = str(tmp_path / f"w{C}.png")
    nli.write_image(p, v)
    got = pil_array(p)
    want = t8(v)
    np.testing.assert_array_equal(got, want)
    assert want.reshape(-1)[:edge.size].tolist() == [0, 0, 0, 0, 1, 2, 3, 254, 255, 255, 255, 0, 0, 255, 0]
    # our reader returns the written bytes
    np.testing.assert_array_equal(nli.read_image(p), want.astype(np.float32))


def test_write_rgb_and_mono(nli, tmp_path):
    rng = np.random.default_rng(4)
    img4 = rng.uniform(0, 255, (12, 14, 4)).astype(np.float32)
    p = str(tmp_path / "rgb.png")
    nli.write_rgb_image(p, img4)
    np.testing.assert_array_equal(pil_array(p), t8(img4[:, :, :3]))
    rgb = img4[:, :, :3].copy()
    p = str(tmp_path / "mono.png")
    nli.write_mono_image(p, rgb)
    luma = (0.2126 * rgb[..., 0].astype(np.float64) + 0.7152 * rgb[..., 1] + 0.0722 * rgb[..., 2]).astype(np.float32)
    np.testing.assert_array_equal(pil_array(p)[:, :, 0], t8(luma))
    nli.write_mono_image(p, rgb[:, :, 0])
    np.testing.assert_array_equal(pil_array(p)[:, :, 0], t8(rgb[:, :, 0]))
    with pytest.raises(ValueError, match="3 channels"):
        nli.write_rgb_image(str(tmp_path / "x.png"), rgb[:, :, :2])
    with pytest.raises(ValueError, match="1 or 3"):
        nli.write_mono_image(str(tmp_path / "x.png"), img4)
    with pytest.raises(ValueError, match="only PNG"):
        nli.write_image(str(tmp_path / "x.tif"), rgb)


def test_round_trip_bytes_large(nli, tmp_path):
    rng = np.random.default_rng(8)
    img = rng.integers(0, 256, (300, 517, 3)).astype(np.float32)
    img[100:200, 100:300] = 77.0   # flat area: exercises the filter choice
    p = str(tmp_path / "big.PNG")
    nli.write_image(p, img)
    np.testing.assert_array_equal(nli.read_image(p), img)
    np.testing.assert_array_equal(pil_array(p).astype(np.float32), img)


def test_damaged_files(nli, tmp_path):
    rng = np.random.default_rng(2)
    p = str(tmp_path / "ok.png")
    write_png(p, rng.integers(0, 256, (5, 6, 3)), 2, 8)
    good = open(p, "rb").read()

    def variant(name, data):
        q = str(tmp_path / name)
        with open(q, "wb") as f:
            f.write(data)
        return q

    with pytest.raises(OSError, match="not a PNG"):
        nli.read_image(variant("sig.png", b"GIF89a" + good[6:]))
    bad = bytearray(good)
    bad[8 + 8 + 3] ^= 1   # IHDR data byte: the CRC no longer matches
    with pytest.raises(OSError, match="CRC"):
        nli.read_image(variant("crc.png", bytes(bad)))
    with pytest.raises(OSError):
        nli.read_image(variant("cut.png", good[:len(good) - 20]))
    with pytest.raises(OSError, match="cannot read"):
        nli.read_rgb_image(str(tmp_path / "missing.png"))
    # an ancillary chunk with a wrong CRC is dropped, as libpng drops it
    anc = struct.pack(">I", 4) + b"tEXt" + b"abcd" + b"\0\0\0\0"
    q = str(tmp_path / "anc.png")
    write_png(q, rng.integers(0, 256, (5, 6, 3)), 2, 8, extra=anc)
    assert nli.read_image(q).shape == (5, 6, 3)
    # a header far larger than its data is refused before the image is allocated
    raw = zlib.compress(b"\0" * 100)
    huge = (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", 46000, 46000, 16, 6, 0, 0, 0))
            + _chunk(b"IDAT", raw) + _chunk(b"IEND", b""))
    q = variant("huge.png", huge)
    assert nli.image_info(q) == (46000, 46000, 4, 16)
    one = np.zeros(1, np.float32)
    with pytest.raises(OSError, match="not enough image data"):   # checked before the capacity
        nli._io_check(nli.load_plugins().nli_read_image(q.encode(), one.ctypes.data_as(nli._f32p), 1))
    # a buffer too small is refused before anything is written
    lib = nli.load_plugins()
    out = np.zeros(10, np.float32)
    assert lib.nli_read_image(p.encode(), out.ctypes.data_as(nli._f32p), out.size) == nli.NLI_E_ARGUMENT


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="the reference's data/*.png are not present")
@pytest.mark.parametrize("name", ["img.png", "mask.png", "depth.png", "kom07.png", "kom07_msk.png"])
def test_reference_data_files(nli, name):
    path = os.path.join(REF_DATA, name)
    im = PIL.open(path)
    want = pil_array(path, "L" if im.mode == "1" else None).astype(np.float32)
    np.testing.assert_array_equal(nli.read_image(path), want)
    rgb = nli.read_rgb_image(path)
    assert rgb.shape == want.shape[:2] + (3,)
    mono = nli.read_mono_image(path)
    if want.shape[2] == 1:
        np.testing.assert_array_equal(mono, want[:, :, 0])
        if name in ("mask.png", "kom07_msk.png"):
            assert set(np.unique(mono)) <= {0.0, 255.0} and (mono > 0).any()


def test_cpp_ioutility_file_functions(nli, tmp_path):
    """IOUtility's file functions in include/nli_plugins.hpp, compiled into a program:
    read_rgb_image + write_rgb_image and read_mono_image + write_mono_image copy 8-bit
    files unchanged; a missing file throws the reader's message."""
    import shutil
    import subprocess
    from conftest import ROOT
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    pkg = os.path.join(ROOT, "non-local-inpainting_amd")
    src = tmp_path / "io.cpp"
    src.write_text(r'''
#include "nli_plugins.hpp"
#include <cstdio>
int main(int argc, char **argv) {
    nli::Image<float> rgb = nli::IOUtility::read_rgb_image(argv[1]);
    nli::IOUtility::write_rgb_image(argv[2], rgb);
    nli::Image<float> mono = nli::IOUtility::read_mono_image(argv[1]);
    nli::IOUtility::write_mono_image(argv[3], mono);
    nli::Mask m(5, 4);
    m.mask(1, 0); m.mask(4, 3);
    nli::IOUtility::write_mono_image(argv[5], nli::IOUtility::mask_to_greyscale(m));
    nli::Image<float> pr(4, 1);
    pr(0, 0) = -1.0f; pr(1, 0) = 0.25f; pr(2, 0) = 2.0f; pr(3, 0) = 0.5f;
    nli::IOUtility::write_mono_image(argv[6], nli::IOUtility::probability_to_greyscale(pr));
    try { nli::IOUtility::read_rgb_image(argv[4]); return 3; }
    catch (const std::runtime_error &e) { std::printf("%s\n", e.what()); }
    return 0;
}
''')
    exe = str(tmp_path / "io")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", exe,
                    "-L", pkg, "-lnli_plugins", "-lnli", f"-Wl,-rpath,{pkg}"], check=True)
    rng = np.random.default_rng(12)
    a = rng.integers(0, 256, (21, 19, 3)).astype(np.uint8)
    PIL.fromarray(a, "RGB").save(tmp_path / "a.png")
    r = subprocess.run([exe, str(tmp_path / "a.png"), str(tmp_path / "b.png"), str(tmp_path / "m.png"),
                        str(tmp_path / "none.png"), str(tmp_path / "mask.png"), str(tmp_path / "prob.png")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "read_rgb_image: cannot read" in r.stdout
    np.testing.assert_array_equal(pil_array(tmp_path / "b.png"), a)
    luma = np.trunc(.299 * a[..., 0].astype(np.float64) + .587 * a[..., 1] + .114 * a[..., 2])
    np.testing.assert_array_equal(pil_array(tmp_path / "m.png")[:, :, 0], luma.astype(np.uint8))
    want = np.zeros((4, 5), np.uint8)
    want[0, 1] = want[3, 4] = 255
    np.testing.assert_array_equal(pil_array(tmp_path / "mask.png")[:, :, 0], want)
    # 255 * clamp(p, 0, 1): 0, 63.75 -> 64 (rounded on writing), 255, 127.5 -> 128
    np.testing.assert_array_equal(pil_array(tmp_path / "prob.png")[:, :, 0], [[0, 64, 255, 128]])


def test_mutated_files_are_decoded_or_refused(nli, tmp_path):
    """Damaged files (bit flips, truncation, rewritten header fields; CRCs repaired on most
    so the damage reaches the decoder) come back decoded or refused with a code, never a
    crash."""
    rng = np.random.default_rng(21)
    seeds = []
    for k, (ctype, depth, il) in enumerate([(0, 2, True), (2, 16, False), (3, 4, True), (6, 8, False)]):
        p = str(tmp_path / f"seed{k}.png")
        plte = rng.integers(0, 256, (16, 3)) if ctype == 3 else None
        write_png(p, rng.integers(0, (1 << depth), (9, 10, CTYPE_CH[ctype])), ctype, depth, il, rng, plte=plte)
        seeds.append(open(p, "rb").read())

    def fix_crcs(m):
        at = 8
        while at + 12 <= len(m):
            n = struct.unpack(">I", bytes(m[at:at + 4]))[0]
            if n > len(m) - at - 12:
                break
            m[at + 8 + n:at + 12 + n] = struct.pack(">I", zlib.crc32(bytes(m[at + 4:at + 8 + n])) & 0xffffffff)
            at += 12 + n

    lib = nli.load_plugins()
    buf = np.zeros(4 * 64 * 64, np.float32)
    q = str(tmp_path / "m.png")
    counts = {}
    for it in range(1500):
        m = bytearray(seeds[it % len(seeds)])
        kind = int(rng.integers(0, 3))
        if kind == 0:
            for _ in range(int(rng.integers(1, 6))):
                m[int(rng.integers(8, len(m)))] ^= 1 << int(rng.integers(0, 8))
        elif kind == 1:
            m = m[:int(rng.integers(8, len(m)))]
        else:
            m[int(rng.integers(16, 29))] = int(rng.integers(0, 256))
        if rng.random() < 0.75:
            fix_crcs(m)
        with open(q, "wb") as f:
            f.write(m)
        rc = lib.nli_read_image(q.encode(), buf.ctypes.data_as(nli._f32p), buf.size)
        assert rc in (nli.NLI_OK, nli.NLI_E_IO, nli.NLI_E_ARGUMENT)
        counts[rc] = counts.get(rc, 0) + 1
    assert counts.get(nli.NLI_OK, 0) > 50 and counts.get(nli.NLI_E_IO, 0) > 50


# ---- PGM (io_utility.cpp:26-117), against the reference's own code built in place ----
def _ref_pgm():
    import ctypes as ct
    import oracle as O
    if not os.path.exists(O.REF_COLOR_SO):
        pytest.skip("reference io_utility.cpp not built in place (oracle/_ref)")
    lib = ct.CDLL(O.REF_COLOR_SO, mode=ct.RTLD_LOCAL)
    f32p = ct.POINTER(ct.c_float)

    def read(path):
        W, H = ct.c_int(), ct.c_int()
        if lib.ref_read_pgm(path.encode(), ct.byref(W), ct.byref(H), None, 0) != 0:
            return None
        out = np.zeros((H.value, W.value), np.float32)
        assert lib.ref_read_pgm(path.encode(), ct.byref(W), ct.byref(H), out.ctypes.data_as(f32p), out.size) == 0
        return out

    def write(path, img):
        img = np.ascontiguousarray(img, np.float32)
        img = img[:, :, None] if img.ndim == 2 else img
        H, W, C = img.shape
        assert lib.ref_write_pgm(path.encode(), img.ctypes.data_as(f32p), W, H, C) == 0
    return read, write


PGM_FILES = {
    "ascii_comments.pgm": b"P2\n# a comment\n  4 \t3\n#another\n255\n 0 1 2 3\n10 200 255 7\n  99 0 12 13\n",
    "binary.pgm": b"P5\n5 2\n255\n" + bytes([200, 1, 2, 3, 4, 250, 6, 7, 8, 9]),
    # leading samples that are whitespace or '#' are skipped by the reference's reader,
    # and the missing samples at the end read as EOF (-1)
    "binary_space_first.pgm": b"P5 3 2 255 " + bytes([32, 10, 35]) + b"xx\n" + bytes([5, 6, 7, 8]),
    "binary_short.pgm": b"P5\n4 4\n# c\n255\n" + bytes(range(100, 110)),
    "binary_16bit_maxval.pgm": b"P5\n3 1\n65535\n" + bytes([1, 2, 3, 4, 5, 6]),
    "bad_magic.pgm": b"P6\n2 2\n255\n" + bytes(12),
}


@pytest.mark.parametrize("name", sorted(PGM_FILES))
def test_pgm_reader_equals_reference(nli, tmp_path, name):
    ref_read, _ = _ref_pgm()
    p = str(tmp_path / name)
    with open(p, "wb") as f:
        f.write(PGM_FILES[name])
    want = ref_read(p)
    if want is None:
        with pytest.raises(OSError):
            nli.read_pgm_image(p)
        return
    got = nli.read_pgm_image(p)
    np.testing.assert_array_equal(got, want)
    if name == "binary_space_first.pgm":
        # the quirk, stated: the bytes 32, 10 are skipped as whitespace, 35 ('#') starts a
        # comment to the newline, and that newline is pushed back and read as the first
        # sample; the last sample is past the end (EOF, -1)
        assert got.reshape(-1).tolist() == [10, 5, 6, 7, 8, -1]


def test_pgm_writer_equals_reference(nli, tmp_path):
    _, ref_write = _ref_pgm()
    rng = np.random.default_rng(17)
    one = rng.uniform(-400, 400, (6, 7)).astype(np.float32)
    one.reshape(-1)[:9] = [200, -1, 300, 1e10, np.nan, -1e10, 127.9, -0.5, 2147483520]
    three = rng.uniform(-50, 350, (5, 4, 3)).astype(np.float32)
    two = rng.uniform(0, 255, (3, 3, 2)).astype(np.float32)
    for k, img in enumerate([one, three, two]):
        a, b = str(tmp_path / f"mine{k}.pgm"), str(tmp_path / f"ref{k}.pgm")
        nli.write_pgm_image(a, img)
        ref_write(b, img)
        assert open(a, "rb").read() == open(b, "rb").read()
    body = open(str(tmp_path / "mine0.pgm"), "rb").read()[len(b"P5\n7 6\n255\n"):]
    assert list(body[:9]) == [200, 255, 44, 0, 0, 0, 127, 0, 128]
    assert open(str(tmp_path / "mine2.pgm"), "rb").read() == b"P5\n3 3\n255\n"


@pytest.mark.parametrize("header", [b"P5\n4294967297 4294967295\n255\n", b"P2\n2147483648 1\n255\n",
                                    b"P5\n65536 65537\n255\n", b"P5\n1099511627776 2\n255\n"])
def test_pgm_header_sizes_beyond_int_are_refused(nli, tmp_path, header):
    """Header sides beyond int, or whose product overflows 64 bits, are refused before
    any allocation (no image of a wrapped size)."""
    p = str(tmp_path / "huge.pgm")
    with open(p, "wb") as f:
        f.write(header + bytes(16))
    with pytest.raises(OSError):
        nli.read_pgm_image(p)
