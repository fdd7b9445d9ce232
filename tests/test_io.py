"""The reference's file formats (host code of the library, no GPU needed):
DAT/RAW descriptors and time series, the raw noise file, 8-bit PNG and the
transfer-function PNG pair.  PNG is checked against Pillow (an independent
codec) and against rows encoded here with each of the five PNG row filters."""
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

import vectorvisualization_amd as vv
from vectorvisualization_amd import _lib as L


# ---- PNG ---------------------------------------------------------------------
@pytest.mark.parametrize("mode,ch", [("L", 1), ("LA", 2), ("RGB", 3), ("RGBA", 4)])
def test_png_read_matches_pillow(tmp_path, mode, ch):
    rng = np.random.default_rng(ch)
    # smooth + noisy content so Pillow's adaptive filtering uses several row filters
    y, x = np.mgrid[0:37, 0:53]
    base = (x * 3 + y * 5)[..., None] + np.arange(ch)[None, None, :] * 40
    px = ((base + rng.integers(0, 4, base.shape)) % 256).astype(np.uint8)
    img = Image.fromarray(px[..., 0] if ch == 1 else px, mode)
    f = tmp_path / "a.png"
    img.save(f, optimize=True)
    got = L.png_read(f)
    assert got.shape == (37, 53, ch)
    np.testing.assert_array_equal(got, px.reshape(37, 53, ch))


@pytest.mark.parametrize("ch", [1, 2, 3, 4])
def test_png_write_read_by_pillow(tmp_path, ch):
    rng = np.random.default_rng(10 + ch)
    px = rng.integers(0, 256, (19, 31, ch), dtype=np.uint8)
    f = tmp_path / "b.png"
    L.png_write(f, px)
    got = np.asarray(Image.open(f))
    np.testing.assert_array_equal(got.reshape(19, 31, ch), px)
    L.png_write(f, px, flip_rows=True)
    got = np.asarray(Image.open(f))
    np.testing.assert_array_equal(got.reshape(19, 31, ch), px[::-1])


def _png_bytes(px: np.ndarray, filt: int) -> bytes:
    """PNG encoder for the test: every row with row filter `filt` (PNG spec 9.2)."""
    h, w, ch = px.shape
    color = {1: 0, 2: 4, 3: 2, 4: 6}[ch]
    raw = bytearray()
    prev = np.zeros(w * ch, np.int64)
    for yy in range(h):
        cur = px[yy].reshape(-1).astype(np.int64)
        left = np.concatenate([np.zeros(ch, np.int64), cur[:-ch]])
        upleft = np.concatenate([np.zeros(ch, np.int64), prev[:-ch]])
        if filt == 0:
            f = cur
        elif filt == 1:
            f = cur - left
        elif filt == 2:
            f = cur - prev
        elif filt == 3:
            f = cur - (left + prev) // 2
        else:
            p = left + prev - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
            f = cur - pred
        raw += bytes([filt]) + bytes((f % 256).astype(np.uint8))
        prev = cur

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    ihdr = struct.pack(">IIBBBBB", w, h, 8, color, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(bytes(raw)))
            + chunk(b"IEND", b""))


@pytest.mark.parametrize("filt", [0, 1, 2, 3, 4])
def test_png_every_row_filter(tmp_path, filt):
    rng = np.random.default_rng(filt)
    px = rng.integers(0, 256, (9, 13, 3), dtype=np.uint8)
    f = tmp_path / "c.png"
    f.write_bytes(_png_bytes(px, filt))
    np.testing.assert_array_equal(np.asarray(Image.open(f)), px)  # the encoder is right
    np.testing.assert_array_equal(L.png_read(f), px)


def test_png_rejects_what_pngread_rejects(tmp_path):
    f = tmp_path / "p.png"
    Image.fromarray(np.zeros((4, 4), np.uint8), "L").convert("P").save(f)
    with pytest.raises(vv.VVError) as e:
        L.png_read(f)
    assert e.value.status == 4  # palette: unsupported
    Image.fromarray(np.full((4, 4), 300, np.uint16)).save(f)
    with pytest.raises(vv.VVError) as e:
        L.png_read(f)
    assert e.value.status == 4  # 16 bits
    good = _png_bytes(np.zeros((3, 3, 1), np.uint8), 0)
    bad = bytearray(good)
    bad[40] ^= 0xFF  # inside IDAT: CRC mismatch
    f.write_bytes(bytes(bad))
    with pytest.raises(vv.VVError) as e:
        L.png_read(f)
    assert e.value.status == 6
    f.write_bytes(good[:30])
    with pytest.raises(vv.VVError):
        L.png_read(f)
    with pytest.raises(vv.VVError):
        L.png_read(tmp_path / "missing.png")
    # a header claiming 65536 x 65536 RGBA (16 GiB of pixels) is refused
    # before anything is allocated
    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    huge = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 65536, 65536, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(b"\x00" * 16)) + chunk(b"IEND", b""))
    f.write_bytes(huge)
    with pytest.raises(vv.VVError) as e:
        L.png_read(f)
    assert e.value.status == 4 and "2^28" in str(e.value)


def test_image_write_png(tmp_path):
    w, h = 7, 5
    rgba = np.zeros((h, w, 4), np.float32)
    rgba[..., 0] = np.linspace(0, 1, w)[None, :]
    rgba[..., 1] = np.linspace(0, 1, h)[:, None]
    rgba[..., 2] = -0.5
    rgba[..., 3] = 2.0
    f = tmp_path / "frame.png"
    L.image_write_png(f, rgba.ravel(), w, h)
    got = np.asarray(Image.open(f))
    # row 0 of the image is the bottom row: the file's first row is the top
    ref = np.rint(np.clip(rgba[::-1], 0, 1) * 255).astype(np.uint8)
    np.testing.assert_array_equal(got, ref)


# ---- transfer function PNGs -----------------------------------------------------
def test_tf_save_load_roundtrip(tmp_path):
    rng = np.random.default_rng(5)
    tf5 = rng.integers(0, 256, 1280, dtype=np.uint8)
    L.tf_save_png(tmp_path / "tf.png", tf5)
    rgba = np.asarray(Image.open(tmp_path / "tf_rgba.png"))
    la = np.asarray(Image.open(tmp_path / "tf_alpha.png"))
    assert rgba.shape == (10, 256, 4) and la.shape == (10, 256, 2)
    t = tf5.reshape(256, 5)
    for row in range(10):
        np.testing.assert_array_equal(rgba[row], t[:, :4])
        np.testing.assert_array_equal(la[row], t[:, 3:5])
    # loading reads the RGBA file only (loadRGBATF() || loadAlphaOpacTF()):
    # the opacity column keeps what the table held before
    start = np.zeros(1280, np.uint8)
    got = L.tf_load_png(tmp_path / "tf.png", start).reshape(256, 5)
    np.testing.assert_array_equal(got[:, :4], t[:, :4])
    np.testing.assert_array_equal(got[:, 4], 0)
    # without the RGBA file the alpha/opacity pair is read
    (tmp_path / "tf_rgba.png").unlink()
    got = L.tf_load_png(tmp_path / "tf.png", start).reshape(256, 5)
    np.testing.assert_array_equal(got[:, 3:5], t[:, 3:5])
    np.testing.assert_array_equal(got[:, :3], 0)


def test_tf_names_without_extension_and_gray(tmp_path):
    # "name" -> name_rgba.png; a gray image sets r = g = b (alpha untouched)
    g = np.tile(np.arange(256, dtype=np.uint8)[::-1], (2, 1))
    Image.fromarray(g, "L").save(tmp_path / "gray_rgba.png")
    start = np.full(1280, 7, np.uint8)
    got = L.tf_load_png(tmp_path / "gray", start).reshape(256, 5)
    np.testing.assert_array_equal(got[:, 0], g[0])
    np.testing.assert_array_equal(got[:, 1], g[0])
    np.testing.assert_array_equal(got[:, 2], g[0])
    np.testing.assert_array_equal(got[:, 3:], 7)
    with pytest.raises(vv.VVError):
        L.tf_load_png(tmp_path / "nothing.png")


def test_tf_too_small_falls_back_to_alpha(tmp_path):
    Image.fromarray(np.zeros((1, 100, 4), np.uint8), "RGBA").save(tmp_path / "s_rgba.png")
    Image.fromarray(np.full((1, 256, 2), 9, np.uint8), "LA").save(tmp_path / "s_alpha.png")
    got = L.tf_load_png(tmp_path / "s.png", np.zeros(1280, np.uint8)).reshape(256, 5)
    np.testing.assert_array_equal(got[:, 3:], 9)
    np.testing.assert_array_equal(got[:, :3], 0)


# ---- DAT / RAW -------------------------------------------------------------------
def _write_dat(path, lines):
    path.write_text("\n".join(lines) + "\n")


def test_dat_single_float3(tmp_path):
    v = np.random.default_rng(1).normal(size=(4, 5, 6, 3)).astype(np.float32)  # z, y, x, c
    (tmp_path / "vel.raw").write_bytes(v.tobytes())
    _write_dat(tmp_path / "vel.dat", [
        "# a comment line with Resolution: 1 1 1",
        "ObjectFileName: vel.raw",
        "TaggedFileName: ---",
        "Resolution:     6 5 4",
        "SliceThickness: 1.0 0.5 2.0",
        "Format:         FLOAT3",
        "ObjectModel:    RGB",
    ])
    info = L.dat_parse(tmp_path / "vel.dat")  # RAW found next to the DAT file
    assert tuple(info.dims) == (6, 5, 4)
    assert tuple(info.spacing) == (1.0, 0.5, 2.0)
    assert info.raw_type == L.RAW_FLOAT and info.components == 3
    assert (info.t_begin, info.t_end) == (0, 0)
    assert info.raw_pattern.decode().endswith("vel.raw")
    np.testing.assert_array_equal(L.dat_read(info), v)
    np.testing.assert_array_equal(L.dat_read_field(info), v)


def test_dat_time_series_and_uchar(tmp_path):
    rng = np.random.default_rng(2)
    steps = [rng.integers(0, 256, (3, 4, 5, 3), dtype=np.uint8) for _ in range(3)]
    (tmp_path / "d").mkdir()
    for t, s in zip(range(2, 5), steps):
        (tmp_path / "d" / f"flow_{t:03d}.raw").write_bytes(s.tobytes())
    _write_dat(tmp_path / "d" / "flow.dat", [
        "ObjectFileName: flow_%03d.raw",
        "Resolution: 5 4 3",
        "Format: UCHAR 3",
        "TimeDependent: 2 4",
    ])
    info = L.dat_parse(tmp_path / "d" / "flow.dat")
    assert (info.t_begin, info.t_end) == (2, 4)
    assert info.raw_type == L.RAW_UCHAR and info.components == 3
    for t, s in zip(range(2, 5), steps):
        np.testing.assert_array_equal(L.dat_read(info, t), s)
        np.testing.assert_array_equal(L.dat_read_field(info, t), s.astype(np.float32) - 128.0)
    with pytest.raises(vv.VVError):
        L.dat_read(info, 5)
    # a missing time step fails the parse (reader.cpp:243-255)
    (tmp_path / "d" / "flow_003.raw").unlink()
    with pytest.raises(vv.VVError) as e:
        L.dat_parse(tmp_path / "d" / "flow.dat")
    assert e.value.status == 6


def test_dat_errors(tmp_path):
    (tmp_path / "x.raw").write_bytes(b"\0" * 8)
    cases = {
        "nores": ["ObjectFileName: x.raw", "Format: UCHAR"],
        "badres": ["ObjectFileName: x.raw", "Resolution: 2 2", "Format: UCHAR"],
        "badfmt": ["ObjectFileName: x.raw", "Resolution: 2 2 2", "Format: INT"],
        "badtime": ["ObjectFileName: x.raw", "Resolution: 2 2 2", "Format: UCHAR", "TimeDependent: 3 1"],
        "pattern": ["ObjectFileName: x%s.raw", "Resolution: 2 2 2", "Format: UCHAR"],
        "noraw": ["ObjectFileName: y.raw", "Resolution: 2 2 2", "Format: UCHAR"],
    }
    for name, lines in cases.items():
        _write_dat(tmp_path / f"{name}.dat", lines)
        with pytest.raises(vv.VVError):
            L.dat_parse(tmp_path / f"{name}.dat")
    _write_dat(tmp_path / "ok.dat", ["ObjectFileName: x.raw", "Resolution: 2 2 2", "Format: UCHAR"])
    info = L.dat_parse(tmp_path / "ok.dat")
    assert info.components == 1
    with pytest.raises(vv.VVError):
        L.dat_read_field(info)  # not a vector data set
    # a RAW file shorter than the resolution says
    _write_dat(tmp_path / "short.dat", ["ObjectFileName: x.raw", "Resolution: 4 4 4", "Format: UCHAR"])
    with pytest.raises(vv.VVError):
        L.dat_read(L.dat_parse(tmp_path / "short.dat"))


# ---- noise file ----------------------------------------------------------------
def test_noise_file_roundtrip_and_layout(tmp_path):
    n = np.random.default_rng(3).integers(0, 256, 6 * 5 * 4, dtype=np.uint8)
    f = tmp_path / "noise.raw"
    L.noise_file_write(f, n, (6, 5, 4))
    b = f.read_bytes()
    assert struct.unpack("<3i", b[:12]) == (6, 5, 4)
    assert b[12:] == n.tobytes()
    got, dims = L.noise_file_read(f)
    assert dims == (6, 5, 4)
    np.testing.assert_array_equal(got, n)
    f.write_bytes(b[:-1])
    with pytest.raises(vv.VVError):
        L.noise_file_read(f)
