"""ctypes binding of the C ABI in include/vvlic.h (libvvlic.so).

The library is the only compute path: importing this module fails loudly when
libvvlic.so is missing, and every volume operation runs in its HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("VVLIC_LIB", _PKG / "lib" / "libvvlic.so"))

VV_OK, VV_EINVAL, VV_ENOMEM, VV_EDEVICE, VV_EUNSUPPORTED, VV_ESTATE = range(6)
STATUS_NAMES = {0: "VV_OK", 1: "VV_EINVAL", 2: "VV_ENOMEM", 3: "VV_EDEVICE",
                4: "VV_EUNSUPPORTED", 5: "VV_ESTATE", 6: "VV_EIO"}
STORE_FP32, STORE_FP16, STORE_UNORM8 = 0, 1, 2
ENCODE_INTERP, ENCODE_RAW_MAGNITUDE = 0, 1
FILTER_BOX, FILTER_TRIANGLE, FILTER_COS2, FILTER_GAUSSIAN = range(4)
NOISE_BINARY, NOISE_UNIFORM = 0, 1
ILLUM_LIC, ILLUM_GRADIENT = 0, 1
TF_INDEX_B, TF_INDEX_MAG = 0, 1


class LicParams(C.Structure):
    """vv_lic_params, mirror of LICParams (VV/types.h:91-109)."""

    _fields_ = [
        ("step_size_vol", C.c_float),
        ("gradient_scale", C.c_float),
        ("illum_scale", C.c_float),
        ("freq_scale", C.c_float),
        ("num_iterations", C.c_int),
        ("steps_forward", C.c_int),
        ("steps_backward", C.c_int),
        ("step_size_lic", C.c_float),
        ("eye_factor", C.c_float),
        ("illum_mode", C.c_int),
        ("tf_index", C.c_int),
        ("gate_enable", C.c_int),
        ("gate_lo", C.c_float),
        ("gate_hi", C.c_float),
        ("speed_of_flow", C.c_int),
    ]


class Camera(C.Structure):
    """vv_camera, mirror of Camera (VV/camera.cpp:42-68) plus the light."""

    _fields_ = [
        ("width", C.c_int),
        ("height", C.c_int),
        ("fovy_deg", C.c_float),
        ("dist", C.c_float),
        ("pos", C.c_float * 3),
        ("quat", C.c_float * 4),
        ("light_pos", C.c_float * 3),
        ("light_diffuse", C.c_float),
        ("light_specular", C.c_float),
        ("light_ambient", C.c_float),
        ("spec_exponent", C.c_int),
    ]


class VVError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


ILLUM_MALLO, ILLUM_ZOECKLER = 2, 3
RAW_NONE, RAW_UCHAR, RAW_USHORT, RAW_FLOAT = range(4)


class LineMaterial(C.Structure):
    """vv_line_material (LineMat, VV/illumination.h:40-62)."""

    _fields_ = [
        ("light_color", C.c_float * 4),
        ("ambient", C.c_float * 4),
        ("diffuse", C.c_float * 4),
        ("specular", C.c_float * 4),
        ("spec_exp", C.c_float),
        ("diff_exp", C.c_float),
    ]
_RAW_DTYPE = {RAW_UCHAR: np.uint8, RAW_USHORT: np.uint16, RAW_FLOAT: np.float32}


class DatInfo(C.Structure):
    """vv_dat_info: a parsed .dat descriptor (DatFile, VV/reader.h:38-77)."""

    _fields_ = [
        ("raw_pattern", C.c_char * 1024),
        ("dims", C.c_int * 3),
        ("spacing", C.c_float * 3),
        ("raw_type", C.c_int),
        ("components", C.c_int),
        ("t_begin", C.c_int),
        ("t_end", C.c_int),
    ]


def _load() -> C.CDLL:
    # torch ships its own HIP runtime under the same SONAME as the system one
    # (libamdhip64.so.7).  Loading torch first makes libvvlic bind to that
    # runtime, so a process that also uses torch has exactly one HIP runtime;
    # loading libvvlic first would leave torch pairing the system HIP runtime
    # with its bundled HSA runtime, and torch's device initialisation fails.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not LIB_PATH.exists():
        raise ImportError(
            f"libvvlic.so not found at {LIB_PATH}; build it with "
            "`python vectorvisualization_amd/build.py` (hipcc, gfx950)")
    lib = C.CDLL(str(LIB_PATH))
    P, I, F, D = C.c_void_p, C.c_int, C.c_float, C.c_double
    IP = C.POINTER(C.c_int)
    FP = C.POINTER(C.c_float)
    u8p = C.c_void_p
    sig = {
        "vv_create": (I, [C.POINTER(P), I]),
        "vv_destroy": (I, [P]),
        "vv_last_error": (C.c_char_p, [P]),
        "vv_abi_version": (I, []),
        "vv_stream": (P, [P]),
        "vv_synchronize": (I, [P]),
        "vv_default_lic_params": (I, [C.POINTER(LicParams)]),
        "vv_default_camera": (I, [C.POINTER(Camera), I, I, FP]),
        "vv_upload_field": (I, [P, P, P, F, IP, FP, I, I]),
        "vv_upload_field_ex": (I, [P, P, P, F, IP, FP, I, I, I]),
        "vv_field_geometry": (I, [P, FP, FP, FP]),
        "vv_download_field": (I, [P, P, I]),
        "vv_upload_noise": (I, [P, u8p, IP, I, u8p, I]),
        "vv_generate_noise": (I, [P, C.c_uint64, I, D, IP, I]),
        "vv_download_noise": (I, [P, u8p, I]),
        "vv_noise_info": (I, [P, IP, IP]),
        "vv_make_filter": (I, [I, I, F, u8p, IP, FP]),
        "vv_set_filter": (I, [P, u8p, I, F]),
        "vv_set_tf": (I, [P, P, P]),
        "vv_set_tf_u8": (I, [P, u8p]),
        "vv_default_tf_u8": (I, [u8p]),
        "vv_upload_scalar": (I, [P, u8p, IP, I]),
        "vv_set_lic_params": (I, [P, C.POINTER(LicParams)]),
        "vv_get_lic_params": (I, [P, C.POINTER(LicParams)]),
        "vv_compute_lic": (I, [P]),
        "vv_lic_info": (I, [P, IP, IP]),
        "vv_download_lic": (I, [P, P, I]),
        "vv_upload_lic": (I, [P, P, I, IP, I]),
        "vv_render": (I, [P, C.POINTER(Camera), P, I, C.POINTER(C.c_uint64)]),
        "vv_render_fused": (I, [P, C.POINTER(Camera), P, I, C.POINTER(C.c_uint64)]),
        "vv_render_slices": (I, [P, C.POINTER(Camera), P, I, C.POINTER(C.c_uint64)]),
        "vv_render_slices_blend": (I, [P, C.POINTER(Camera), P, I, C.POINTER(C.c_uint64)]),
        "vv_set_opacity_scale": (I, [P, P, I]),
        "vv_render_adjoint": (I, [P, C.POINTER(Camera), P, P, P, I]),
        "vv_render_adjoint_noalloc": (I, [P, C.POINTER(Camera), P, P, P]),
        "vv_render_record": (I, [P, C.POINTER(Camera), P, I, C.POINTER(C.c_uint64)]),
        "vv_record_ray_samples": (I, [P, C.POINTER(Camera), P, I]),
        "vv_batch_view_span": (I, [P, C.POINTER(C.c_double)]),
        "vv_view_block": (I, [P, C.POINTER(Camera), FP]),
        "vv_lic_weights": (I, [P, FP]),
        "vv_batch_create": (I, [P, C.POINTER(P)]),
        "vv_batch_destroy": (I, [P]),
        "vv_batch_add_timestep": (I, [P, P, IP, FP, I, C.c_uint64, I, D, I]),
        "vv_batch_add_view": (I, [P, I, C.POINTER(Camera), P]),
        "vv_batch_info": (I, [P, IP, IP]),
        "vv_batch_autotune": (I, [P]),
        "vv_batch_get_tuning": (I, [P, I, IP]),
        "vv_batch_run": (I, [P]),
        "vv_batch_view_latency": (I, [P, C.POINTER(C.c_double)]),
        "vv_batch_samples": (I, [P, C.POINTER(C.c_uint64)]),
        "vv_batch_download_lic": (I, [P, I, P, I]),
        "vv_dat_parse": (I, [C.c_char_p, C.POINTER(DatInfo)]),
        "vv_dat_read": (I, [C.POINTER(DatInfo), I, P, C.c_size_t]),
        "vv_dat_read_field": (I, [C.POINTER(DatInfo), I, FP]),
        "vv_load_field_dat": (I, [P, C.c_char_p, I, I, F, I]),
        "vv_load_scalar_dat": (I, [P, C.c_char_p]),
        "vv_noise_file_read": (I, [C.c_char_p, IP, u8p, C.c_size_t]),
        "vv_noise_file_write": (I, [C.c_char_p, u8p, IP]),
        "vv_load_noise_file": (I, [P, C.c_char_p, I, I]),
        "vv_load_noise_or_generate": (I, [P, C.c_char_p, I, I, C.c_uint64, IP]),
        "vv_png_read": (I, [C.c_char_p, u8p, C.c_size_t, IP, IP, IP]),
        "vv_png_write": (I, [C.c_char_p, u8p, I, I, I, I]),
        "vv_image_write_png": (I, [C.c_char_p, FP, I, I]),
        "vv_load_filter_png": (I, [P, C.c_char_p]),
        "vv_tf_load_png": (I, [C.c_char_p, u8p]),
        "vv_tf_save_png": (I, [C.c_char_p, u8p]),
        "vv_batch_set_field": (I, [P, I, P, P, F, I]),
        "vv_anim_create": (I, [P, C.POINTER(P), I, IP, FP, I, I, I, C.POINTER(P)]),
        "vv_anim_load_dat": (I, [P, C.c_char_p, I, I, C.POINTER(P)]),
        "vv_anim_next": (I, [P, IP, FP]),
        "vv_anim_seek": (I, [P, C.c_longlong]),
        "vv_anim_destroy": (I, [P]),
        "vv_default_line_material": (None, [C.POINTER(LineMaterial)]),
        "vv_set_line_material": (I, [P, C.POINTER(LineMaterial)]),
        "vv_get_line_material": (I, [P, C.POINTER(LineMaterial)]),
        "vv_make_illum_tables": (I, [C.POINTER(LineMaterial), FP, FP, FP]),
        "vv_set_tuning": (I, [P, C.c_char_p, I]),
        "vv_get_tuning": (I, [P, C.c_char_p, IP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()

# every function declared in include/vvlic.h
EXPORTED = [
    "vv_create", "vv_destroy", "vv_last_error", "vv_abi_version", "vv_stream", "vv_synchronize",
    "vv_default_lic_params", "vv_default_camera", "vv_upload_field", "vv_upload_field_ex", "vv_field_geometry",
    "vv_download_field", "vv_upload_noise", "vv_generate_noise", "vv_download_noise",
    "vv_noise_info", "vv_make_filter", "vv_set_filter", "vv_set_tf", "vv_set_tf_u8",
    "vv_default_tf_u8", "vv_upload_scalar", "vv_set_lic_params", "vv_get_lic_params",
    "vv_compute_lic", "vv_lic_info", "vv_download_lic", "vv_upload_lic", "vv_render",
    "vv_set_opacity_scale", "vv_render_adjoint", "vv_view_block", "vv_lic_weights",
    "vv_set_tuning", "vv_get_tuning", "vv_batch_create", "vv_batch_destroy",
    "vv_batch_add_timestep", "vv_batch_add_view", "vv_batch_info", "vv_batch_autotune",
    "vv_batch_get_tuning", "vv_batch_run", "vv_batch_view_latency", "vv_batch_samples",
    "vv_batch_download_lic", "vv_dat_parse", "vv_dat_read", "vv_dat_read_field",
    "vv_load_field_dat", "vv_load_scalar_dat", "vv_noise_file_read", "vv_noise_file_write",
    "vv_load_noise_file", "vv_load_noise_or_generate", "vv_png_read", "vv_png_write", "vv_image_write_png",
    "vv_load_filter_png", "vv_tf_load_png", "vv_tf_save_png", "vv_batch_set_field",
    "vv_anim_create", "vv_anim_load_dat", "vv_anim_next", "vv_anim_seek", "vv_anim_destroy",
    "vv_default_line_material", "vv_set_line_material", "vv_get_line_material",
    "vv_make_illum_tables", "vv_render_fused", "vv_render_slices", "vv_render_adjoint_noalloc",
    "vv_render_slices_blend", "vv_render_record", "vv_record_ray_samples", "vv_batch_view_span",
]


def _i3(d) -> C.Array:
    return (C.c_int * 3)(*[int(x) for x in d])


def _f3(d) -> C.Array:
    return (C.c_float * 3)(*[float(x) for x in d])


def _ptr(a) -> tuple[int, int]:
    """(address, on_device) of a numpy array or a torch tensor."""
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data, 0
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr(), 1 if a.is_cuda else 0
    raise TypeError(f"unsupported buffer type {type(a)}")


def _buf(a, dtype: str, count: int, what: str) -> tuple[int, int]:
    """(address, on_device) of a numpy array or torch tensor holding at least
    `count` elements of `dtype` ("float32" or "uint8"); the library reads or
    writes that many, so a shorter or differently typed buffer is refused
    here (VV_EINVAL) instead of being overrun."""
    if isinstance(a, np.ndarray):
        got_type, n = str(a.dtype), a.size
    elif hasattr(a, "data_ptr"):
        got_type, n = str(a.dtype).replace("torch.", ""), a.numel()
    else:
        raise TypeError(f"{what}: unsupported buffer type {type(a)}")
    if got_type != dtype:
        raise VVError(1, f"{what}: {dtype} buffer expected, got {got_type}")
    if n < count:
        raise VVError(1, f"{what}: {count} elements needed, the buffer holds {n}")
    return _ptr(a)


def _one_side(named, what: str) -> None:
    """Refuse a call whose buffers mix host memory and device memory.  The C
    entry points take one on_device flag for all their buffers, so a host
    buffer passed beside a device one would be dereferenced on the wrong side
    (a GPU page fault or a host segfault).  `named` holds (name, buffer)
    pairs; None buffers are skipped."""
    sides = [(n, _ptr(b)[1]) for n, b in named if b is not None]
    if len({d for _, d in sides}) > 1:
        desc = ", ".join(f"{n} on {'device' if d else 'host'}" for n, d in sides)
        raise ValueError(f"{what}: buffers must all be host memory or all device memory ({desc})")


def _nvox(dims) -> int:
    return int(dims[0]) * int(dims[1]) * int(dims[2])


def _is_cuda(a) -> bool:
    return a is not None and getattr(a, "is_cuda", False)


class _StreamOrder:
    """Orders the context's stream after the caller's current torch stream on
    entry, and the caller's stream after the context's stream on exit, when
    any argument is a CUDA tensor (device events; the host does not wait).
    Inputs produced by torch are then complete before the library reads
    them, and torch work queued after the call sees the library's outputs."""

    def __init__(self, ctx: "Context", arrays):
        self._ctx = ctx
        self._on = any(_is_cuda(a) for a in arrays)

    def __enter__(self):
        if self._on:
            import torch

            self._cur = torch.cuda.current_stream(self._ctx.device)
            self._ext = self._ctx._torch_stream()
            self._on = self._cur.cuda_stream != self._ext.cuda_stream
            if self._on:
                ev = torch.cuda.Event()
                ev.record(self._cur)
                self._ext.wait_event(ev)
        return self

    def __exit__(self, *exc):
        if self._on:
            import torch

            ev = torch.cuda.Event()
            ev.record(self._ext)
            self._cur.wait_event(ev)
        return False


def default_lic_params() -> LicParams:
    p = LicParams()
    LIB.vv_default_lic_params(C.byref(p))
    return p


def default_camera(width: int, height: int, center=(0.5, 0.5, 0.5)) -> Camera:
    c = Camera()
    LIB.vv_default_camera(C.byref(c), int(width), int(height), _f3(center))
    return c


def make_filter(kind: int, length: int = 0, sigma: float = 0.0):
    """Synthesised 256-entry LIC filter LUT and its inverse area."""
    lut = np.zeros(256, np.uint8)
    w = C.c_int()
    inv = C.c_float()
    st = LIB.vv_make_filter(int(kind), int(length), float(sigma), lut.ctypes.data, C.byref(w),
                            C.byref(inv))
    if st:
        raise VVError(st, "vv_make_filter")
    return lut[: w.value].copy(), inv.value


def default_tf_u8() -> np.ndarray:
    t = np.zeros(256 * 5, np.uint8)
    LIB.vv_default_tf_u8(t.ctypes.data)
    return t


def default_line_material() -> LineMaterial:
    m = LineMaterial()
    LIB.vv_default_line_material(C.byref(m))
    return m


def make_illum_tables(mat: LineMaterial | None = None):
    """(zoeckler 256x256x2, mallo_diff 256x256x3, mallo_spec 256x256x3) float32."""
    m = default_line_material() if mat is None else mat
    z = np.empty(256 * 256 * 2, np.float32)
    md = np.empty(256 * 256 * 3, np.float32)
    ms = np.empty(256 * 256 * 3, np.float32)
    fp = C.POINTER(C.c_float)
    _io_chk(LIB.vv_make_illum_tables(C.byref(m), z.ctypes.data_as(fp), md.ctypes.data_as(fp),
                                     ms.ctypes.data_as(fp)))
    return z, md, ms


def _io_chk(st: int) -> None:
    if st:
        raise VVError(st, LIB.vv_last_error(None).decode(errors="replace"))


def _path(p) -> bytes:
    return os.fsencode(p)


def dat_parse(path) -> DatInfo:
    """Parse a .dat descriptor; every time step's RAW file must exist."""
    info = DatInfo()
    _io_chk(LIB.vv_dat_parse(_path(path), C.byref(info)))
    return info


def dat_read(info: DatInfo, timestep: int | None = None) -> np.ndarray:
    """Raw voxels of one time step, shape (z, y, x[, components])."""
    t = info.t_begin if timestep is None else int(timestep)
    shape = (info.dims[2], info.dims[1], info.dims[0]) + ((info.components,) if info.components > 1 else ())
    out = np.empty(shape, _RAW_DTYPE[info.raw_type])
    _io_chk(LIB.vv_dat_read(C.byref(info), t, out.ctypes.data, out.nbytes))
    return out


def dat_read_field(info: DatInfo, timestep: int | None = None) -> np.ndarray:
    """One time step of a vector data set as float32 (z, y, x, 3)."""
    t = info.t_begin if timestep is None else int(timestep)
    out = np.empty((info.dims[2], info.dims[1], info.dims[0], 3), np.float32)
    _io_chk(LIB.vv_dat_read_field(C.byref(info), t, out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


def noise_file_read(path):
    d = (C.c_int * 3)()
    _io_chk(LIB.vv_noise_file_read(_path(path), d, None, 0))
    out = np.empty(d[0] * d[1] * d[2], np.uint8)
    _io_chk(LIB.vv_noise_file_read(_path(path), d, out.ctypes.data, out.nbytes))
    return out, tuple(d)


def noise_file_write(path, noise: np.ndarray, dims) -> None:
    noise = np.ascontiguousarray(noise, np.uint8)
    _io_chk(LIB.vv_noise_file_write(_path(path), noise.ctypes.data, _i3(dims)))


def png_read(path) -> np.ndarray:
    """8-bit PNG as (height, width, channels), rows in file order."""
    w, h, ch = C.c_int(), C.c_int(), C.c_int()
    _io_chk(LIB.vv_png_read(_path(path), None, 0, C.byref(w), C.byref(h), C.byref(ch)))
    out = np.empty((h.value, w.value, ch.value), np.uint8)
    _io_chk(LIB.vv_png_read(_path(path), out.ctypes.data, out.nbytes, C.byref(w), C.byref(h),
                            C.byref(ch)))
    return out


def png_write(path, pixels: np.ndarray, flip_rows: bool = False) -> None:
    """pixels: (height, width[, channels]) uint8."""
    px = np.ascontiguousarray(pixels, np.uint8)
    ch = 1 if px.ndim == 2 else px.shape[2]
    _io_chk(LIB.vv_png_write(_path(path), px.ctypes.data, px.shape[1], px.shape[0], ch, int(flip_rows)))


def image_write_png(path, rgba, width: int, height: int) -> None:
    """A rendered image (width*height RGBA float32, bottom row first) as PNG."""
    rgba = np.ascontiguousarray(rgba, np.float32)
    _io_chk(LIB.vv_image_write_png(_path(path), rgba.ctypes.data_as(C.POINTER(C.c_float)), int(width),
                                   int(height)))


def tf_load_png(name, tf5: np.ndarray | None = None) -> np.ndarray:
    """TransferEdit::loadTF on a 256x5 byte table (starting from tf5, or the default)."""
    t = default_tf_u8() if tf5 is None else np.array(tf5, np.uint8).ravel().copy()
    _io_chk(LIB.vv_tf_load_png(_path(name), t.ctypes.data))
    return t


def tf_save_png(name, tf5: np.ndarray) -> None:
    t = np.ascontiguousarray(tf5, np.uint8).ravel()
    _io_chk(LIB.vv_tf_save_png(_path(name), t.ctypes.data))


class Context:
    """One engine context on one GPU (vv_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = LIB.vv_create(C.byref(h), int(device))
        if st:
            raise VVError(st, LIB.vv_last_error(None).decode(errors="replace"))
        self._h = h
        self.device = int(device)
        self._ext = None
        self.sigma_serial = 0  # count of set_opacity_scale calls

    def _torch_stream(self):
        """The context's stream as a torch stream object."""
        if self._ext is None:
            import torch

            self._ext = torch.cuda.ExternalStream(self.stream, device=self.device)
        return self._ext

    def ordered(self, *arrays) -> _StreamOrder:
        """Context manager ordering this context's stream against the current
        torch stream around work that touches the given tensors."""
        return _StreamOrder(self, arrays)

    # -- plumbing --
    def _chk(self, st: int) -> None:
        if st:
            raise VVError(st, LIB.vv_last_error(self._h).decode(errors="replace"))

    def close(self) -> None:
        if getattr(self, "_h", None):
            LIB.vv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return LIB.vv_stream(self._h) or 0

    def synchronize(self) -> None:
        self._chk(LIB.vv_synchronize(self._h))

    def last_error(self) -> str:
        return LIB.vv_last_error(self._h).decode()

    # -- inputs --
    def upload_field(self, v0, dims, v1=None, lerp_t: float = 0.0, spacing=(1.0, 1.0, 1.0),
                     storage: int = STORE_FP32, encoding: int = ENCODE_INTERP) -> None:
        p0, dev = _buf(v0, "float32", 3 * _nvox(dims), "field")
        p1 = _buf(v1, "float32", 3 * _nvox(dims), "second field")[0] if v1 is not None else None
        with self.ordered(v0, v1):
            if encoding == ENCODE_INTERP:
                self._chk(LIB.vv_upload_field(self._h, p0, p1, float(lerp_t), _i3(dims), _f3(spacing),
                                              int(storage), dev))
            else:
                self._chk(LIB.vv_upload_field_ex(self._h, p0, p1, float(lerp_t), _i3(dims), _f3(spacing),
                                                 int(storage), int(encoding), dev))

    def set_line_material(self, m: LineMaterial) -> None:
        self._chk(LIB.vv_set_line_material(self._h, C.byref(m)))

    def get_line_material(self) -> LineMaterial:
        m = LineMaterial()
        self._chk(LIB.vv_get_line_material(self._h, C.byref(m)))
        return m

    def load_field_dat(self, path, timestep: int | None = None, next_timestep: int | None = None,
                       lerp_t: float = 0.0, storage: int = STORE_FP32) -> None:
        info = dat_parse(path)
        t = info.t_begin if timestep is None else int(timestep)
        n = t if next_timestep is None else int(next_timestep)
        self._chk(LIB.vv_load_field_dat(self._h, _path(path), t, n, float(lerp_t), int(storage)))

    def load_scalar_dat(self, path) -> None:
        self._chk(LIB.vv_load_scalar_dat(self._h, _path(path)))

    def load_noise_file(self, path, with_gradients: bool = False, use_gradient_cache: bool = True) -> None:
        self._chk(LIB.vv_load_noise_file(self._h, _path(path), int(bool(with_gradients)),
                                         int(bool(use_gradient_cache))))

    def load_noise_or_generate(self, path, with_gradients: bool = False, use_gradient_cache: bool = True,
                               seed: int = 1) -> bool:
        """NoiseDataSet::loadData: the noise file, or generated 256^3 white noise
        when it cannot be read; returns True when the noise was generated."""
        g = C.c_int(0)
        p = None if path is None else _path(path)
        self._chk(LIB.vv_load_noise_or_generate(self._h, p, int(bool(with_gradients)),
                                                int(bool(use_gradient_cache)), C.c_uint64(seed), C.byref(g)))
        return bool(g.value)

    def load_filter_png(self, path) -> None:
        self._chk(LIB.vv_load_filter_png(self._h, _path(path)))

    def field_geometry(self):
        e, s, c = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
        self._chk(LIB.vv_field_geometry(self._h, e, s, c))
        return list(e), list(s), list(c)

    def download_field(self, dims) -> np.ndarray:
        out = np.empty(int(np.prod(dims)) * 4, np.float32)
        self._chk(LIB.vv_download_field(self._h, out.ctypes.data, 0))
        return out

    def upload_noise(self, noise, dims, with_gradients: bool = False, grad_rgb=None) -> None:
        p, dev = _buf(noise, "uint8", _nvox(dims), "noise")
        g = _buf(grad_rgb, "uint8", 3 * _nvox(dims), "noise gradients")[0] if grad_rgb is not None else None
        with self.ordered(noise, grad_rgb):
            self._chk(LIB.vv_upload_noise(self._h, p, _i3(dims), int(bool(with_gradients)), g, dev))

    def generate_noise(self, seed: int, kind: int, density: float, dims,
                       with_gradients: bool = False) -> None:
        self._chk(LIB.vv_generate_noise(self._h, C.c_uint64(seed), int(kind), float(density),
                                        _i3(dims), int(bool(with_gradients))))

    def noise_info(self):
        d = (C.c_int * 3)()
        ch = C.c_int()
        self._chk(LIB.vv_noise_info(self._h, d, C.byref(ch)))
        return list(d), ch.value

    def download_noise(self) -> np.ndarray:
        d, ch = self.noise_info()
        out = np.empty(int(np.prod(d)) * ch, np.uint8)
        self._chk(LIB.vv_download_noise(self._h, out.ctypes.data, 0))
        return out

    def set_filter(self, lut: np.ndarray, inv_area: float = 0.0) -> None:
        lut = np.ascontiguousarray(lut, np.uint8)
        self._chk(LIB.vv_set_filter(self._h, lut.ctypes.data, int(lut.size), float(inv_area)))

    def set_tf(self, rgba: np.ndarray, opacity: np.ndarray) -> None:
        rgba = np.ascontiguousarray(rgba, np.float32)
        opacity = np.ascontiguousarray(opacity, np.float32)
        self._chk(LIB.vv_set_tf(self._h, rgba.ctypes.data, opacity.ctypes.data))

    def set_tf_u8(self, tf5: np.ndarray) -> None:
        tf5 = np.ascontiguousarray(tf5, np.uint8)
        self._chk(LIB.vv_set_tf_u8(self._h, tf5.ctypes.data))

    def upload_scalar(self, scalar, dims) -> None:
        p, dev = _buf(scalar, "uint8", _nvox(dims), "scalar volume")
        with self.ordered(scalar):
            self._chk(LIB.vv_upload_scalar(self._h, p, _i3(dims), dev))

    def set_lic_params(self, p: LicParams) -> None:
        self._chk(LIB.vv_set_lic_params(self._h, C.byref(p)))

    def get_lic_params(self) -> LicParams:
        p = LicParams()
        self._chk(LIB.vv_get_lic_params(self._h, C.byref(p)))
        return p

    # -- LIC --
    def compute_lic(self) -> None:
        self._chk(LIB.vv_compute_lic(self._h))

    def lic_info(self):
        d = (C.c_int * 3)()
        ch = C.c_int()
        self._chk(LIB.vv_lic_info(self._h, d, C.byref(ch)))
        return list(d), ch.value

    def _lic_voxels(self) -> int:
        """Voxels of the LIC grid (the opacity scale's size); 0 without a
        LIC volume, which the library then reports itself."""
        try:
            return _nvox(self.lic_info()[0])
        except VVError:
            return 0

    def download_lic(self, out=None) -> np.ndarray:
        d, ch = self.lic_info()
        if out is None:
            out = np.empty(int(np.prod(d)) * ch, np.float32)
        p, dev = _buf(out, "float32", _nvox(d) * ch, "LIC output")
        with self.ordered(out):
            self._chk(LIB.vv_download_lic(self._h, p, dev))
        return out

    def upload_lic(self, lic, channels: int, dims) -> None:
        p, dev = _buf(lic, "float32", _nvox(dims) * int(channels), "LIC volume")
        with self.ordered(lic):
            self._chk(LIB.vv_upload_lic(self._h, p, int(channels), _i3(dims), dev))

    def lic_weights(self) -> np.ndarray:
        p = self.get_lic_params()
        out = np.empty(1 + p.steps_forward + p.steps_backward, np.float32)
        self._chk(LIB.vv_lic_weights(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_tuning(self, key: str, value: int) -> None:
        self._chk(LIB.vv_set_tuning(self._h, key.encode(), int(value)))

    def get_tuning(self, key: str) -> int:
        v = C.c_int()
        self._chk(LIB.vv_get_tuning(self._h, key.encode(), C.byref(v)))
        return v.value

    # -- render --
    def render(self, cam: Camera, out=None, count_samples: bool = False):
        if out is None:
            out = np.empty(cam.width * cam.height * 4, np.float32)
        p, dev = _buf(out, "float32", cam.width * cam.height * 4, "image")
        ns = C.c_uint64()
        with self.ordered(out):
            self._chk(LIB.vv_render(self._h, C.byref(cam), p, dev,
                                    C.byref(ns) if count_samples else None))
        return (out, ns.value) if count_samples else out

    def render_record(self, cam: Camera, out=None, count_samples: bool = False):
        """vv_render that keeps the adjoint's forward pass for a following
        render_adjoint with the same camera and scene."""
        if out is None:
            out = np.empty(cam.width * cam.height * 4, np.float32)
        p, dev = _buf(out, "float32", cam.width * cam.height * 4, "image")
        ns = C.c_uint64()
        with self.ordered(out):
            self._chk(LIB.vv_render_record(self._h, C.byref(cam), p, dev,
                                           C.byref(ns) if count_samples else None))
        return (out, ns.value) if count_samples else out

    def record_ray_samples(self, cam: Camera, out=None):
        """vv_record_ray_samples: per-ray sample counts (int32, height*width)
        of the last render_record of this camera."""
        if out is None:
            out = np.empty(cam.width * cam.height, np.int32)
        p, dev = _buf(out, "int32", cam.width * cam.height, "ray sample counts")
        with self.ordered(out):
            self._chk(LIB.vv_record_ray_samples(self._h, C.byref(cam), p, dev))
        return out

    def render_fused(self, cam: Camera, out=None, count_samples: bool = False):
        """Fused ray cast: the LIC of every sample computed on the fly."""
        if out is None:
            out = np.empty(cam.width * cam.height * 4, np.float32)
        p, dev = _buf(out, "float32", cam.width * cam.height * 4, "image")
        ns = C.c_uint64()
        with self.ordered(out):
            self._chk(LIB.vv_render_fused(self._h, C.byref(cam), p, dev,
                                          C.byref(ns) if count_samples else None))
        return (out, ns.value) if count_samples else out

    def render_slices(self, cam: Camera, out=None, count_samples: bool = False):
        """View-aligned slicing with on-the-fly LIC (FP16 render target)."""
        if out is None:
            out = np.empty(cam.width * cam.height * 4, np.float32)
        p, dev = _buf(out, "float32", cam.width * cam.height * 4, "image")
        ns = C.c_uint64()
        with self.ordered(out):
            self._chk(LIB.vv_render_slices(self._h, C.byref(cam), p, dev,
                                           C.byref(ns) if count_samples else None))
        return (out, ns.value) if count_samples else out

    def render_slices_blend(self, cam: Camera, out=None, count_samples: bool = False):
        """The reference's default slicing output: the slices blended into an
        8-bit RGBA framebuffer over a white background (uint8, 4 per pixel)."""
        if out is None:
            out = np.empty(cam.width * cam.height * 4, np.uint8)
        p, dev = _buf(out, "uint8", cam.width * cam.height * 4, "image")
        ns = C.c_uint64()
        with self.ordered(out):
            self._chk(LIB.vv_render_slices_blend(self._h, C.byref(cam), p, dev,
                                                 C.byref(ns) if count_samples else None))
        return (out, ns.value) if count_samples else out

    def view_block(self, cam: Camera) -> np.ndarray:
        out = (C.c_float * 24)()
        self._chk(LIB.vv_view_block(self._h, C.byref(cam), out))
        return np.array(out, np.float32)

    def set_opacity_scale(self, sigma) -> None:
        self.sigma_serial += 1
        if sigma is None:
            self._chk(LIB.vv_set_opacity_scale(self._h, None, 0))
            return
        p, dev = _buf(sigma, "float32", self._lic_voxels(), "opacity scale")
        with self.ordered(sigma):
            self._chk(LIB.vv_set_opacity_scale(self._h, p, dev))

    def render_adjoint(self, cam: Camera, dimage, dtf=None, dsigma=None, want_sigma: bool = True):
        _one_side([("dimage", dimage), ("dtf", dtf), ("dsigma", dsigma)], "render_adjoint")
        host = isinstance(dimage, np.ndarray)
        if host:
            dimage = np.ascontiguousarray(dimage, np.float32)
            if dtf is None:
                dtf = np.zeros(1024, np.float32)
            if dsigma is None and want_sigma:
                d, _ = self.lic_info()
                dsigma = np.zeros(int(np.prod(d)), np.float32)
        pi, dev = _buf(dimage, "float32", cam.width * cam.height * 4, "image gradient")
        pt = _buf(dtf, "float32", 1024, "TF gradient")[0]
        ps = _buf(dsigma, "float32", self._lic_voxels(), "opacity-scale gradient")[0] if dsigma is not None else None
        with self.ordered(dimage, dtf, dsigma):
            self._chk(LIB.vv_render_adjoint(self._h, C.byref(cam), pi, pt, ps, dev))
        return dtf, dsigma


    def render_adjoint_noalloc(self, cam: Camera, dimage, dtf, dsigma=None):
        """vv_render_adjoint_noalloc: device tensors only; no device allocation."""
        _one_side([("dimage", dimage), ("dtf", dtf), ("dsigma", dsigma)], "render_adjoint_noalloc")
        pi, dev = _buf(dimage, "float32", cam.width * cam.height * 4, "image gradient")
        if not dev:
            raise ValueError("render_adjoint_noalloc takes device tensors")
        pt = _buf(dtf, "float32", 1024, "TF gradient")[0]
        ps = _buf(dsigma, "float32", self._lic_voxels(), "opacity-scale gradient")[0] if dsigma is not None else None
        with self.ordered(dimage, dtf, dsigma):
            self._chk(LIB.vv_render_adjoint_noalloc(self._h, C.byref(cam), pi, pt, ps))
        return dtf, dsigma


class Batch:
    """Many timesteps and views in one captured launch (vv_batch)."""

    def __init__(self, ctx: Context):
        h = C.c_void_p()
        ctx._chk(LIB.vv_batch_create(ctx.handle, C.byref(h)))
        self._ctx = ctx
        self._h = h
        self._images = []
        self._dims = []  # per time step, for the buffer checks of set_field

    def close(self) -> None:
        if getattr(self, "_h", None):
            LIB.vv_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_timestep(self, v, dims, spacing=(1.0, 1.0, 1.0), storage: int = STORE_FP32,
                     noise_seed: int = 1, noise_kind: int = NOISE_UNIFORM, density: float = 0.0) -> int:
        p, dev = _buf(v, "float32", 3 * _nvox(dims), "time step field")
        with self._ctx.ordered(v):
            self._ctx._chk(LIB.vv_batch_add_timestep(self._h, p, _i3(dims), _f3(spacing),
                                                     int(storage), C.c_uint64(noise_seed),
                                                     int(noise_kind), float(density), dev))
        self._dims.append(tuple(int(x) for x in dims))
        return self.info()[0] - 1

    def add_view(self, timestep: int, cam: Camera, image) -> int:
        p, dev = _buf(image, "float32", cam.width * cam.height * 4, "batch image")
        if not dev:
            raise ValueError("batch images must be device tensors")
        self._ctx._chk(LIB.vv_batch_add_view(self._h, int(timestep), C.byref(cam), p))
        self._images.append(image)
        return self.info()[1] - 1

    def info(self):
        t, v = C.c_int(), C.c_int()
        LIB.vv_batch_info(self._h, C.byref(t), C.byref(v))
        return t.value, v.value

    def autotune(self) -> None:
        self._ctx._chk(LIB.vv_batch_autotune(self._h))

    def tuning(self, timestep: int):
        """LIC tile x,y,z and brick tiles x,y,z of a timestep, then the ray
        caster's pixel tile x,y (16 x 8: 2 lanes per ray, 8 x 8: 4 lanes per
        ray, any other shape: one lane per pixel)."""
        out = (C.c_int * 8)()
        self._ctx._chk(LIB.vv_batch_get_tuning(self._h, int(timestep), out))
        return list(out)

    def run(self) -> None:
        with self._ctx.ordered(*self._images):
            self._ctx._chk(LIB.vv_batch_run(self._h))

    def view_latency_ms(self) -> np.ndarray:
        n = self.info()[1]
        out = (C.c_double * n)()
        self._ctx._chk(LIB.vv_batch_view_latency(self._h, out))
        return np.array(out)

    def view_span_ms(self) -> np.ndarray:
        """Per view: its own render span, first tile start to last tile end."""
        n = self.info()[1]
        out = (C.c_double * n)()
        self._ctx._chk(LIB.vv_batch_view_span(self._h, out))
        return np.array(out)

    def samples(self) -> int:
        t = C.c_uint64()
        self._ctx._chk(LIB.vv_batch_samples(self._h, C.byref(t)))
        return t.value

    def set_field(self, timestep: int, v0, v1=None, lerp_t: float = 0.0) -> None:
        n = 3 * _nvox(self._dims[timestep]) if 0 <= timestep < len(self._dims) else 0
        _one_side([("v0", v0), ("v1", v1)], "Batch.set_field")
        p0, dev = _buf(v0, "float32", n, "time step field")
        p1 = _buf(v1, "float32", n, "second time step field")[0] if v1 is not None else None
        with self._ctx.ordered(v0, v1):
            self._ctx._chk(LIB.vv_batch_set_field(self._h, int(timestep), p0, p1, float(lerp_t), dev))

    def download_lic(self, timestep: int, dims) -> np.ndarray:
        out = np.empty(int(np.prod(dims)), np.float32)
        self._ctx._chk(LIB.vv_batch_download_lic(self._h, int(timestep), out.ctypes.data, 0))
        return out


class Anim:
    """An unsteady series resident on the GPU; next() encodes the next frame's
    lerped field into the context (vv_anim)."""

    def __init__(self, ctx: Context, steps=None, dims=None, spacing=(1.0, 1.0, 1.0), interp_size: int = 10,
                 storage: int = STORE_FP32, dat_path=None):
        h = C.c_void_p()
        self._ctx = ctx
        if dat_path is not None:
            ctx._chk(LIB.vv_anim_load_dat(ctx.handle, _path(dat_path), int(interp_size), int(storage),
                                          C.byref(h)))
        else:
            ptrs = [_buf(s, "float32", 3 * _nvox(dims), f"time step {i}") for i, s in enumerate(steps)]
            devs = {d for _, d in ptrs}
            if len(devs) != 1:
                raise ValueError("all time steps must be host arrays or all device tensors")
            arr = (C.c_void_p * len(ptrs))(*[p for p, _ in ptrs])
            with ctx.ordered(*steps):
                ctx._chk(LIB.vv_anim_create(ctx.handle, arr, len(ptrs), _i3(dims), _f3(spacing),
                                            int(interp_size), int(storage), devs.pop(), C.byref(h)))
        self._h = h

    def next(self):
        s, t = C.c_int(), C.c_float()
        self._ctx._chk(LIB.vv_anim_next(self._h, C.byref(s), C.byref(t)))
        return s.value, t.value

    def seek(self, frame: int) -> None:
        self._ctx._chk(LIB.vv_anim_seek(self._h, int(frame)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            LIB.vv_anim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
