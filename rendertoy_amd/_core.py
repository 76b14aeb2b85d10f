"""Buffers, structs and host math of the rendertoy API (rendering/_core.py), on torch device memory.

What is kept from rendering/_core.py:
  - vector dtypes float2..float4x4, int2..uint4 with the OpenCL layout pyopencl registers
    (3-component vectors padded to 16 bytes; component titles x, y, z, w) (:100-114);
  - `kernel_struct` -> numpy structured dtype with the OpenCL C layout (:302-310);
  - `create_buffer`, `create_buffer_from`, `create_struct`, `create_struct_from`, `mapped`,
    `clear` (:13-14, :323-332, :376-418) returning arrays with `.get()`, `.shape`, `.dtype`;
  - host camera math `identity/translate/scale/rotate/matmul/dot/normalize/cross/direction/
    look_at/perspective` (:421-548), float32 like the reference;
  - `Texture2D` descriptor and `create_texture2D` (:313-317, :551-578), here over a pool of device
    memory with an RGBA8 texel format added for the ray-cast shading path.
What is not: pyopencl contexts, OpenCL images and the JIT of OpenCL-C docstrings
(`kernel_function`/`kernel_main` compile user kernels; the ray-casting path ships precompiled
HIP kernels instead, see `Dispatcher`).
"""
import contextlib
import math
import typing

import numpy as np
import torch

# --------------------------------------------------------------------------------------------
# device selection: buffers live on the HIP device; without one (CPU-only hosts) buffers are
# host tensors so the host-side API can still be exercised, and every ray-casting call raises.
# --------------------------------------------------------------------------------------------


def device():
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


# --------------------------------------------------------------------------------------------
# vector dtypes (pyopencl.cltypes layout: fields s0..sN, titles x,y,z,w; 3-vectors padded to 4)
# --------------------------------------------------------------------------------------------

_TITLES = ["x", "y", "z", "w"]


def _vector_dtype(base, count):
    padded = 4 if count == 3 else count
    names = [f"s{i}" for i in range(count)] + [f"padding{i}" for i in range(padded - count)]
    titles = _TITLES[:count] + [None] * (padded - min(count, 4))
    titles = titles[:padded]
    return np.dtype({"names": names, "formats": [base] * padded, "titles": titles})


float2 = _vector_dtype(np.float32, 2)
float3 = _vector_dtype(np.float32, 3)
float4 = _vector_dtype(np.float32, 4)
int2 = _vector_dtype(np.int32, 2)
int3 = _vector_dtype(np.int32, 3)
int4 = _vector_dtype(np.int32, 4)
uint2 = _vector_dtype(np.uint32, 2)
uint3 = _vector_dtype(np.uint32, 3)
uint4 = _vector_dtype(np.uint32, 4)
float4x4 = _vector_dtype(np.float32, 16)
RGBA = _vector_dtype(np.uint8, 4)

# OpenCL alignment of each registered vector type = its padded size
_VECTOR_ALIGN = {float2: 8, float3: 16, float4: 16, int2: 8, int3: 16, int4: 16, uint2: 8,
                 uint3: 16, uint4: 16, float4x4: 64, RGBA: 4}


def _align_of(dt: np.dtype) -> int:
    dt = np.dtype(dt)
    for k, a in _VECTOR_ALIGN.items():
        if dt == k:
            return a
    if dt.fields:
        return getattr(dt, "_rt_align", max(_align_of(f[0]) for f in dt.fields.values()))
    if dt.subdtype is not None:
        return _align_of(dt.subdtype[0])
    return dt.itemsize


def _as_dtype(t):
    if isinstance(t, np.dtype):
        return t
    if t is int:
        return np.dtype(np.int32)  # OpenCL `int` (the reference maps Python ints to int32, :276)
    if t is float:
        return np.dtype(np.float32)
    return np.dtype(t)


def kernel_struct(cls):
    """Class decorator -> numpy structured dtype laid out by OpenCL C rules (rendering/_core.py:302).

    Each field sits at the next multiple of its alignment; the struct size is padded to the largest
    field alignment. This is what `pyopencl.tools.match_dtype_to_c_struct` reports for a device.
    """
    fields = cls.__dict__.get("__annotations__", {})
    assert all(k in fields for k in cls.__dict__ if k[0] != "_"), \
        "A public field was declared without annotation"
    names, formats, offsets = [], [], []
    off, max_align = 0, 1
    for name, t in fields.items():
        dt = _as_dtype(t)
        a = _align_of(dt)
        off = (off + a - 1) // a * a
        names.append(name)
        formats.append(dt)
        offsets.append(off)
        off += dt.itemsize
        max_align = max(max_align, a)
    size = (off + max_align - 1) // max_align * max_align
    dtype = np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size})
    _STRUCT_ALIGN[dtype] = max_align
    _STRUCT_NAMES[dtype] = cls.__name__
    return dtype


_STRUCT_ALIGN: typing.Dict[np.dtype, int] = {}
_STRUCT_NAMES: typing.Dict[np.dtype, str] = {}


def struct_layout(dtype):
    """[(name, offset, itemsize)] of a kernel_struct dtype, for layout checks against C headers."""
    return [(n, dtype.fields[n][1], dtype.fields[n][0].itemsize) for n in dtype.names]


# --------------------------------------------------------------------------------------------
# device arrays
# --------------------------------------------------------------------------------------------


class DeviceArray:
    """A typed device buffer: torch uint8 storage viewed through a numpy dtype and shape.

    Stands in for `pyopencl.array.Array` on the calls the reference makes: `.get()`, `.shape`,
    `.dtype`, indexing a struct out of a 1-element buffer (`create_struct`, :327-328).
    """

    def __init__(self, storage: torch.Tensor, dtype, shape, offset=0):
        self.storage = storage          # flat torch.uint8 tensor
        self.dtype = np.dtype(dtype)
        self.shape = tuple(int(s) for s in shape)
        self.offset = int(offset)       # byte offset into storage

    # ---- properties ----
    @property
    def size(self):
        return int(np.prod(self.shape)) if self.shape else 1

    @property
    def nbytes(self):
        return self.size * self.dtype.itemsize

    @property
    def itemsize(self):
        return self.dtype.itemsize

    @property
    def device(self):
        return self.storage.device

    def data_ptr(self) -> int:
        return self.storage.data_ptr() + self.offset

    def bytes_tensor(self) -> torch.Tensor:
        return self.storage[self.offset:self.offset + self.nbytes]

    def tensor(self, torch_dtype=None, shape=None) -> torch.Tensor:
        """A torch view of the bytes (e.g. float32 components of a float3 buffer)."""
        t = self.bytes_tensor()
        if torch_dtype is not None:
            t = t.view(torch_dtype)
        if shape is not None:
            t = t.view(*shape)
        return t

    def __len__(self):
        return self.shape[0] if self.shape else 1

    def __getitem__(self, i):
        if self.shape and isinstance(i, int):
            n = self.shape[0]
            if i < 0:
                i += n
            if not 0 <= i < n:
                raise IndexError(i)
            inner = int(np.prod(self.shape[1:])) if len(self.shape) > 1 else 1
            return DeviceArray(self.storage, self.dtype, self.shape[1:],
                               self.offset + i * inner * self.dtype.itemsize)
        if isinstance(i, slice):
            start, stop, step = i.indices(self.shape[0])
            if step != 1:
                raise IndexError("only contiguous slices of device arrays are supported")
            inner = int(np.prod(self.shape[1:])) if len(self.shape) > 1 else 1
            return DeviceArray(self.storage, self.dtype, (max(0, stop - start),) + self.shape[1:],
                               self.offset + start * inner * self.dtype.itemsize)
        raise IndexError(i)

    def view(self, dtype):
        dtype = np.dtype(dtype)
        nb = self.nbytes
        if nb % dtype.itemsize:
            raise ValueError("buffer size is not a multiple of the new dtype")
        return DeviceArray(self.storage, dtype, (nb // dtype.itemsize,), self.offset)

    def reshape(self, *shape):
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        if int(np.prod(shape)) != self.size:
            raise ValueError("reshape changes the element count")
        return DeviceArray(self.storage, self.dtype, shape, self.offset)

    # ---- transfers ----
    def get(self) -> np.ndarray:
        """Device -> host copy as a numpy array of the buffer's dtype (blocking)."""
        host = self.bytes_tensor().cpu().numpy()
        return host.view(self.dtype).reshape(self.shape)

    def set(self, ary):
        """Host -> device copy of `ary` (same byte size)."""
        a = np.ascontiguousarray(ary)
        b = a.view(np.uint8).reshape(-1)
        if b.size != self.nbytes:
            raise ValueError(f"set: {b.size} bytes for a {self.nbytes}-byte buffer")
        self.bytes_tensor().copy_(torch.from_numpy(b.copy()))

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype}, device={self.device})"


Buffer = DeviceArray


def _alloc(nbytes):
    # 256-byte aligned allocations (torch's caching allocator returns >= 512-byte aligned blocks)
    return torch.zeros(max(int(nbytes), 1), dtype=torch.uint8, device=device())


def create_buffer(count: int, dtype) -> DeviceArray:
    """Zero-filled device buffer of `count` elements (rendering/_core.py:13-14)."""
    dtype = _as_dtype(dtype)
    return DeviceArray(_alloc(count * dtype.itemsize), dtype, (int(count),))


def create_buffer_from(ary: np.ndarray) -> DeviceArray:
    """Device copy of a host array (rendering/_core.py:323-324)."""
    ary = np.ascontiguousarray(ary)
    buf = DeviceArray(_alloc(ary.nbytes), ary.dtype, ary.shape)
    if ary.nbytes:
        buf.set(ary)
    return buf


def create_struct(dtype) -> DeviceArray:
    """Zero-filled single struct on the device (rendering/_core.py:327-328)."""
    return create_buffer(1, dtype)[0]


def create_struct_from(ary: np.ndarray) -> DeviceArray:
    """Device copy of one struct value (rendering/_core.py:331-332)."""
    a = np.asarray(ary)
    return create_buffer_from(a.reshape(1))[0]


@contextlib.contextmanager
def mapped(b: DeviceArray):
    """Host view of a device buffer, written back on exit (rendering/_core.py:391-418)."""
    host = b.get().copy()
    yield host
    b.set(host)


def clear(b: DeviceArray, value=np.float32(0)):
    """Fills a buffer with a repeated scalar/struct bit pattern (rendering/_core.py:376-388)."""
    if isinstance(value, float):
        value = np.float32(value)
    if isinstance(value, int):
        value = np.int32(value)
    pat = np.frombuffer(np.ascontiguousarray(value).tobytes(), dtype=np.uint8)
    nb = b.nbytes
    if nb % pat.size:
        raise ValueError("buffer size is not a multiple of the fill pattern")
    t = b.bytes_tensor().view(-1, pat.size)
    t.copy_(torch.from_numpy(pat.copy()).to(t.device).expand_as(t))


# --------------------------------------------------------------------------------------------
# vector constructors (rendering/_core.py:127-181)
# --------------------------------------------------------------------------------------------


def make_int2(*args):
    if len(args) == 1 and isinstance(args[0], np.ndarray):
        return args[0].ravel().view(int2).item()
    return np.array(tuple(args), dtype=int2)


def make_int3(*args):
    if len(args) == 1 and isinstance(args[0], np.ndarray):
        args = args[0].ravel().view(np.int32)
    return np.array(tuple([*args[:3], 0]), dtype=int3)


def make_int4(*args):
    if len(args) == 1 and isinstance(args[0], np.ndarray):
        return args[0].ravel().view(int4).item()
    return np.array(tuple(args), dtype=int4)


def make_float2(*args):
    if len(args) == 1 and isinstance(args[0], np.ndarray):
        return np.ascontiguousarray(args[0], dtype=np.float32).ravel().view(float2).item()
    return np.array(tuple(args), dtype=float2)


def make_float3(*args):
    if len(args) == 1 and isinstance(args[0], np.ndarray):
        args = np.ascontiguousarray(args[0]).ravel().astype(np.float32)
    return np.array(tuple([*[float(a) for a in args[:3]], 0.0]), dtype=float3)


def make_float4(*args):
    if len(args) == 1 and isinstance(args[0], np.ndarray):
        a = np.ascontiguousarray(args[0], dtype=np.float32).ravel()
        return np.array(tuple(float(x) for x in a[:4]), dtype=float4)
    return np.array(tuple(args), dtype=float4)


def make_float4x4(*args):
    if len(args) == 1 and isinstance(args[0], np.ndarray):
        a = np.ascontiguousarray(args[0], dtype=np.float32).ravel()
        return np.array(tuple(float(x) for x in a[:16]), dtype=float4x4)
    return np.array(tuple(args), dtype=float4x4)


def to_array(v):
    """Vector dtype -> float32 components (float3 drops the padding) (rendering/_core.py:163)."""
    v = np.asarray(v)
    shape = v.shape
    if v.shape == ():
        v = np.expand_dims(v, 0)
    if v.dtype == float2:
        return v.view(np.float32).reshape(*shape, 2)
    if v.dtype == float3:
        return v.view(np.float32).reshape(*shape, 4)[..., 0:3]
    if v.dtype == float4:
        return v.view(np.float32).reshape(*shape, 4)
    if v.dtype == float4x4:
        return v.view(np.float32).reshape(*shape, 4, 4)
    return v


# --------------------------------------------------------------------------------------------
# host math (rendering/_core.py:421-548): row-vector convention, v' = v @ M
# --------------------------------------------------------------------------------------------


def identity():
    return make_float4x4(np.eye(4, dtype=np.float32))


def translate(*args):
    x, y, z = to_array(args[0]) if len(args) == 1 else args
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = (x, y, z)
    return make_float4x4(m)


def scale(*args):
    if len(args) == 1:
        x = y = z = args[0] if np.isscalar(args[0]) else None
        if x is None:
            x, y, z = to_array(args[0])
    else:
        x, y, z = args
    return make_float4x4(np.diag(np.array([x, y, z, 1.0], dtype=np.float32)))


def rotate(angle, axis):
    """Rotation by `angle` about the unit `axis` (row-vector form, rendering/_core.py:459)."""
    c, s = np.cos(angle), np.sin(angle)
    x, y, z = (float(a) for a in to_array(axis))
    k = 1.0 - c
    m = np.array([
        [x * x * k + c, y * x * k + z * s, z * x * k - y * s, 0.0],
        [x * y * k - z * s, y * y * k + c, z * y * k + x * s, 0.0],
        [x * z * k + y * s, y * z * k - x * s, z * z * k + c, 0.0],
        [0.0, 0.0, 0.0, 1.0]], dtype=np.float32)
    return make_float4x4(m)


def matmul(a, b):
    assert a.dtype == float4 or a.dtype == float4x4, "First vector must be a float4 or a matrix float4x4"
    assert b.dtype == float4x4, "Second argument must be a matrix"
    c = to_array(a) @ to_array(b)
    return make_float4(c) if a.dtype == float4 else make_float4x4(c)


def dot(v1, v2):
    assert v1.dtype == v2.dtype, "Can not apply dot product between different vector types"
    assert v1.shape == v2.shape, "Can not apply dot product between different vector types"
    n = {float2: 2, float3: 3, float4: 4}.get(v1.dtype)
    if n is None:
        raise Exception("Not valid dtype")
    a, b = to_array(v1), to_array(v2)
    return (a[..., :n] * b[..., :n]).sum(-1).item()


def normalize(v):
    l = np.float32(np.sqrt(dot(v, v)))
    a = to_array(v) / l
    return make_float3(a) if v.dtype == float3 else make_float4(a) if v.dtype == float4 else make_float2(a)


def cross(v1, v2):
    a, b = to_array(v1), to_array(v2)
    return make_float3(np.cross(a, b).astype(np.float32))


def direction(f, t):
    return normalize(make_float3(to_array(t) - to_array(f)))


def look_at(camera, target, up_vector):
    """Left-handed view matrix: z = forward, x = normalize(up x z), y = z x x (rendering/_core.py:528)."""
    zaxis = direction(camera, target)
    xaxis = normalize(cross(up_vector, zaxis))
    yaxis = cross(zaxis, xaxis)
    xa, ya, za = to_array(xaxis), to_array(yaxis), to_array(zaxis)
    return make_float4x4(
        xa[0], ya[0], za[0], 0.0,
        xa[1], ya[1], za[1], 0.0,
        xa[2], ya[2], za[2], 0.0,
        -dot(xaxis, camera), -dot(yaxis, camera), -dot(zaxis, camera), 1.0)


def perspective(fov=3.141593 / 4, aspect_ratio=1.0, znear=.01, zfar=100.0):
    """Vertical-fov projection, NDC z in [0,1] (rendering/_core.py:540-548)."""
    hs = 1.0 / np.tan(fov / 2)
    ws = hs / aspect_ratio
    return make_float4x4(
        ws, 0, 0, 0,
        0, hs, 0, 0,
        0, 0, zfar / (zfar - znear), 1.0,
        0, 0, -znear * zfar / (zfar - znear), 0)


# --------------------------------------------------------------------------------------------
# textures (rendering/_core.py:313-317, :551-578)
# --------------------------------------------------------------------------------------------


@kernel_struct
class Texture2D:
    width: np.int32
    height: np.int32
    offset: np.int32


class MemoryPool:
    """Bump allocator of texture memory: a texture is (texels, Texture2D{width, height, offset}),
    offset in bytes from the pool base. The reference fixes float4 texels (16 B); `RGBA` (4 B)
    texels are the format the ray-cast shading kernel samples."""

    def __init__(self, max_size=256 * 1024 * 1024):
        self.max_size = int(max_size)
        self.buffer = None
        self.malloc_ptr = 0

    def _ensure(self):
        if self.buffer is None:
            self.buffer = _alloc(self.max_size)

    def allocate_texture(self, width, height, texel=float4):
        texel = np.dtype(texel)
        nbytes = width * height * texel.itemsize
        start = (self.malloc_ptr + 255) // 256 * 256
        if start + nbytes > self.max_size:
            raise Exception("Memory out!")
        self._ensure()
        memory = DeviceArray(self.buffer, texel, (height, width), start)
        desc = create_struct(Texture2D)
        with mapped(desc) as m:
            m["width"] = width
            m["height"] = height
            m["offset"] = start
        self.malloc_ptr = start + nbytes
        return memory, desc

    def pointer_of(self, offset):
        self._ensure()
        return self.buffer.data_ptr() + int(offset)

    def get_buffer(self):
        self._ensure()
        return self.buffer


__MEMORY_POOL__ = MemoryPool()


def create_texture2D(width: int, height: int, texel=float4):
    """(texels [height, width], descriptor) from the texture pool (rendering/_core.py:577-578)."""
    return __MEMORY_POOL__.allocate_texture(width, height, texel)


def texture_pointer(descriptor) -> int:
    """Device address of a texture's texels from its Texture2D descriptor."""
    d = descriptor.get() if isinstance(descriptor, DeviceArray) else np.asarray(descriptor)
    return __MEMORY_POOL__.pointer_of(int(d["offset"]))


# --------------------------------------------------------------------------------------------
# launch convention (rendering/_core.py:261-290): k[n](*args)
# --------------------------------------------------------------------------------------------


class Dispatcher:
    """`k[n](*args)` launch of a precompiled entry point: `n` is an int or a shape (product taken,
    as Dispatcher.__getitem__ does in rendering/_core.py:265-267); the call runs `fn(n, *args)`."""

    def __init__(self, fn, name=None):
        self._fn = fn
        self.name = name or getattr(fn, "__name__", "kernel")

    def __getitem__(self, num_threads):
        if isinstance(num_threads, (list, tuple)):
            num_threads = math.prod(num_threads)
        n = int(num_threads)

        def dispatch_call(*args):
            return self._fn(n, *args)

        return dispatch_call
