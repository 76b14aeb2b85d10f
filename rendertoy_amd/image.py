"""Image export for the ray-casting path: the render's float RGB to the reference's BGRA8 render
target format, PNG files, and an offline stand-in for its presenter.

- `to_bgra8(rgb)`: float32 RGB [H, W, 3] device tensor -> uint8 [H, W, 4] in B, G, R, A order with
  the OpenCL UNORM_INT8 write conversion (rt_image_to_bgra8; rendering/_core.py:340 makes RGBA
  images CL_BGRA / CL_UNORM_INT8).
- `save_png(path, image)`: writes a BGRA8 image (or float RGB, converted first) as an 8-bit RGBA
  PNG, swapping B and R as the reference's capture does (Class2022/Team Camilo-Javier-Karel/
  scene.py:74-79); `read_png(path)` reads such files back (8-bit RGB/RGBA, no interlace).
- `Presenter` / `create_presenter`: the offline mode of rendering/_presentation.py:17-87 (no
  window: `present()` does nothing, `poll_events()` reports CLOSED); its render target is a BGRA8
  device tensor, and `capture(path)` saves it.
"""
import struct
import zlib
from enum import Enum

import numpy as np
import torch

from . import _native as N


def to_bgra8(rgb: torch.Tensor, out: torch.Tensor = None, stream=None) -> torch.Tensor:
    """float32 RGB [..., 3] on the device -> uint8 [..., 4] (B, G, R, 255), on `stream`."""
    lib = N.require_gpu()
    if rgb.device.type != "cuda" or rgb.dtype != torch.float32 or not rgb.is_contiguous() or rgb.shape[-1] != 3:
        raise TypeError("rgb must be a contiguous float32 device tensor [..., 3]")
    shape = tuple(rgb.shape[:-1]) + (4,)
    ts = stream if stream is not None else torch.cuda.current_stream()
    if out is None:
        with torch.cuda.stream(ts):
            out = torch.empty(shape, dtype=torch.uint8, device=rgb.device)
    elif out.dtype != torch.uint8 or tuple(out.shape) != shape or not out.is_contiguous():
        raise TypeError(f"out must be a contiguous uint8 device tensor {shape}")
    N.check(lib.rt_image_to_bgra8(rgb.data_ptr(), rgb.numel() // 3, out.data_ptr(), N.stream_handle(ts)),
            "rt_image_to_bgra8")
    return out


def _png_chunk(kind: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def save_png(path, image):
    """Writes `image` as an 8-bit RGBA PNG. `image`: uint8 BGRA [H, W, 4] (numpy or tensor), or
    float RGB [H, W, 3] (a device tensor, converted with to_bgra8)."""
    if isinstance(image, torch.Tensor):
        if image.dtype == torch.float32:
            image = to_bgra8(image.contiguous())
        image = image.cpu().numpy()
    img = np.ascontiguousarray(image, np.uint8)
    if img.ndim != 3 or img.shape[2] != 4:
        raise TypeError("save_png takes a BGRA8 [H, W, 4] image or a float RGB [H, W, 3] tensor")
    h, w = img.shape[:2]
    rgba = img[:, :, [2, 1, 0, 3]]
    raw = np.concatenate([np.zeros((h, 1), np.uint8), rgba.reshape(h, 4 * w)], axis=1)  # filter 0
    data = (b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            _png_chunk(b"IDAT", zlib.compress(raw.tobytes(), 6)) + _png_chunk(b"IEND", b""))
    with open(path, "wb") as f:
        f.write(data)


def read_png(path) -> np.ndarray:
    """Reads an 8-bit RGB or RGBA, non-interlaced PNG: uint8 [H, W, channels] in file order."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, hdr = 8, [], None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
        pos += 12 + n
    w, h, depth, color, _, _, interlace = hdr
    if depth != 8 or color not in (2, 6) or interlace:
        raise ValueError(f"{path}: only 8-bit RGB/RGBA non-interlaced PNGs are read")
    ch = 4 if color == 6 else 3
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8).reshape(h, 1 + w * ch)
    out = np.zeros((h, w * ch), np.int32)
    for y in range(h):  # PNG filters (None, Sub, Up, Average, Paeth) per scanline
        ft, line = raw[y, 0], raw[y, 1:].astype(np.int32)
        prev = out[y - 1] if y else np.zeros(w * ch, np.int32)
        if ft == 0:
            out[y] = line
        elif ft == 2:
            out[y] = (line + prev) & 0xFF
        else:
            cur = np.zeros(w * ch, np.int32)
            for x in range(w * ch):
                a = cur[x - ch] if x >= ch else 0
                b = prev[x]
                c = prev[x - ch] if x >= ch else 0
                if ft == 1:
                    p = a
                elif ft == 3:
                    p = (a + b) >> 1
                else:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                cur[x] = (line[x] + p) & 0xFF
            out[y] = cur
    return out.astype(np.uint8).reshape(h, w, ch)


class Event(Enum):
    NONE = 0
    CLOSED = 1


class Presenter:
    """Offline presenter (rendering/_presentation.py:17-87 with offline=True): a BGRA8 render
    target on the device and no window. An on-screen presenter needs SDL, which this package does
    not use."""

    def __init__(self, width: int, height: int, offline: bool = True):
        if not offline:
            raise RuntimeError("rendertoy_amd presents offline only (no SDL window); use capture()")
        self.width = width
        self.height = height
        self.offline = True
        self.render_target = torch.zeros((height, width, 4), dtype=torch.uint8, device=torch.device("cuda"))
        self.frames = 0

    def get_render_target(self) -> torch.Tensor:
        return self.render_target

    def is_alive(self) -> bool:
        return not self.offline

    def poll_events(self):
        return Event.CLOSED, None

    def write(self, rgb: torch.Tensor):
        """Converts a float RGB [H, W, 3] frame into the render target."""
        to_bgra8(rgb.contiguous(), out=self.render_target)

    def present(self):
        self.frames += 1

    def capture(self, path):
        """Saves the render target as a PNG (RGBA)."""
        save_png(path, self.render_target)


def create_presenter(width: int, height: int) -> Presenter:
    """rendering/_presentation.py:90-91, offline."""
    return Presenter(width, height, True)
