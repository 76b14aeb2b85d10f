"""Image export (rendertoy_amd.image; rt_image_to_bgra8): the OpenCL UNORM_INT8 write conversion
of the reference's BGRA render targets (rendering/_core.py:340; OpenCL 1.2 section 8.3.1.1:
convert_uchar_sat_rte(f * 255)) and PNG files.

CPU: the oracle's conversion against an independent numpy statement and hand-worked values;
PNG write/read round trips (including files with Sub/Up/Average/Paeth filters written by zlib+numpy
here). GPU: the device conversion equals the oracle's bit for bit; a rendered frame captured by the
offline presenter reads back as the same pixels."""
import math
import struct
import zlib

import numpy as np
import pytest

from rendertoy_amd import image


def numpy_unorm8(f):
    x = np.asarray(f, np.float32) * np.float32(255.0)
    with np.errstate(invalid="ignore"):
        y = np.where(x > 0, np.minimum(np.rint(x), 255), 0)
    return np.nan_to_num(y, nan=0).astype(np.uint8)


SPECIAL = np.array([0.0, -0.0, 1.0, 2.0, -1.0, np.nan, np.inf, -np.inf, 0.5, 0.5 / 255, 1.5 / 255,
                    2.5 / 255, 254.5 / 255, 1e-30, 0.999, 0.0019607843], np.float32)


def test_twin_conversion_known_values():
    import oracle
    rgb = np.stack([SPECIAL, SPECIAL[::-1], np.roll(SPECIAL, 3)], 1)
    out = oracle.rgb_to_bgra8(rgb)
    assert out.shape == (len(SPECIAL), 4)
    np.testing.assert_array_equal(out[:, 3], 255)
    r = out[:, 2]
    # 0 -> 0, 1 -> 255, saturation, NaN -> 0, 0.5 -> 127.5 -> 128 (ties to even), 0.5/255 -> 0
    assert list(r[:10]) == [0, 0, 255, 255, 0, 0, 255, 0, 128, 0]
    np.testing.assert_array_equal(out[:, [2, 1, 0]], numpy_unorm8(rgb))


def test_twin_conversion_random_against_numpy():
    import oracle
    rng = np.random.default_rng(5)
    rgb = rng.uniform(-0.2, 1.2, (4000, 3)).astype(np.float32)
    rgb[::97] = (np.arange(0, 256, 1)[: len(rgb[::97])] + 0.5)[:, None] / np.float32(255)  # near-ties
    np.testing.assert_array_equal(oracle.rgb_to_bgra8(rgb)[:, [2, 1, 0]], numpy_unorm8(rgb))


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(6)
    bgra = rng.integers(0, 256, (37, 53, 4), dtype=np.uint8)
    p = tmp_path / "a.png"
    image.save_png(p, bgra)
    back = image.read_png(p)
    np.testing.assert_array_equal(back, bgra[:, :, [2, 1, 0, 3]])


def _png_filtered(rgb, filters):
    """An RGB PNG whose scanlines use the given filter types (encoded here, by definition)."""
    h, w, _ = rgb.shape
    ch = 3
    img = rgb.reshape(h, w * ch).astype(np.int32)
    rows = []
    for y in range(h):
        ft = filters[y % len(filters)]
        cur, prev = img[y], img[y - 1] if y else np.zeros(w * ch, np.int32)
        a = np.concatenate([np.zeros(ch, np.int32), cur[:-ch]])
        c = np.concatenate([np.zeros(ch, np.int32), prev[:-ch]])
        if ft == 0:
            pred = 0
        elif ft == 1:
            pred = a
        elif ft == 2:
            pred = prev
        elif ft == 3:
            pred = (a + prev) >> 1
        else:
            pa, pb, pc = np.abs(prev - c), np.abs(a - c), np.abs(a + prev - 2 * c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, prev, c))
        rows.append(bytes([ft]) + ((cur - pred) & 0xFF).astype(np.uint8).tobytes())
    chunk = lambda k, d: struct.pack(">I", len(d)) + k + d + struct.pack(">I", zlib.crc32(k + d) & 0xFFFFFFFF)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(b"".join(rows))) + chunk(b"IEND", b""))


def test_png_reader_filters(tmp_path):
    rng = np.random.default_rng(7)
    rgb = rng.integers(0, 256, (10, 9, 3), dtype=np.uint8)
    p = tmp_path / "f.png"
    p.write_bytes(_png_filtered(rgb, [0, 1, 2, 3, 4]))
    np.testing.assert_array_equal(image.read_png(p), rgb)


@pytest.mark.gpu
def test_device_conversion_equals_twin(ren):
    import oracle
    import torch
    rng = np.random.default_rng(8)
    rgb = rng.uniform(-0.2, 1.2, (123, 77, 3)).astype(np.float32)
    rgb.reshape(-1, 3)[: len(SPECIAL)] = SPECIAL[:, None]
    dev = image.to_bgra8(torch.from_numpy(rgb).cuda()).cpu().numpy()
    np.testing.assert_array_equal(dev, oracle.rgb_to_bgra8(rgb))


@pytest.mark.gpu
def test_offline_presenter_captures_a_render(ren, tmp_path):
    import oracle
    from rendertoy_amd import scenes
    d = scenes.icosphere(32, 9)
    rc = ren.Raycaster([ren.mesh_from_arrays(d["pos"], d["idx"], d["nrm"], d["uv"])])
    W, H = 160, 120
    cam = ren.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 4, W / H)
    light = (np.array([1, 1, 1], np.float32) / np.float32(np.sqrt(3.0))).astype(np.float32)
    rgb = rc.render(cam, W, H, light=light)
    pr = image.create_presenter(W, H)
    assert not pr.is_alive() and pr.poll_events()[0] == image.Event.CLOSED
    pr.write(rgb)
    pr.present()
    pr.capture(tmp_path / "frame.png")
    back = image.read_png(tmp_path / "frame.png")
    want = oracle.rgb_to_bgra8(rgb.cpu().numpy())
    np.testing.assert_array_equal(back, want[:, :, [2, 1, 0, 3]])
    assert (back[:, :, :3] > 0).any()
