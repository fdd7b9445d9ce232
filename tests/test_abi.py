"""C-ABI checks that need no GPU: the library loads, exports every function
include/vvlic.h declares, and its host-only entry points and error paths
behave as documented."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import vectorvisualization_amd as vv
from vectorvisualization_amd import _lib
from oracle import oracle as orc

HEADER = Path(__file__).resolve().parents[1] / "include" / "vvlic.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vv_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_what_the_binding_lists():
    assert declared_functions() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_lib.LIB_PATH))
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_abi_version():
    assert vv.LIB.vv_abi_version() == 2  # vv_lic_params gained speed_of_flow


def test_default_params_mirror_licparams():
    p = vv.default_lic_params()
    # VV/types.h:93-98
    assert p.step_size_vol == np.float32(1.0 / 128.0)
    assert (p.gradient_scale, p.illum_scale, p.freq_scale) == (30.0, 1.0, 1.0)
    assert (p.num_iterations, p.steps_forward, p.steps_backward) == (255, 32, 32)
    assert p.step_size_lic == np.float32(0.01)
    assert p.eye_factor == np.float32(0.3)


def test_default_camera_and_light():
    c = vv.default_camera(640, 480, (0.5, 0.25, 0.5))
    assert (c.fovy_deg, c.dist, list(c.quat)) == (35.0, 4.0, [0, 0, 0, 1])
    np.testing.assert_allclose(list(c.light_pos), orc.light_position((0.5, 0.25, 0.5)), rtol=1e-7)
    assert c.spec_exponent == 40


def test_default_tf_matches_transferedit_constructor():
    t = vv.default_tf_u8().reshape(256, 5)
    i = np.arange(256)
    assert np.all(t[:, :3] == i[:, None])
    assert np.all(t[:, 3:] == np.where(i < 20, 0, i - 20)[:, None])


@pytest.mark.parametrize("kind,L", [(vv.FILTER_BOX, 20), (vv.FILTER_TRIANGLE, 30),
                                    (vv.FILTER_COS2, 60), (vv.FILTER_GAUSSIAN, 40)])
def test_make_filter_matches_oracle(kind, L):
    a, ia = vv.make_filter(kind, L)
    b, ib = orc.make_filter(kind, L)
    assert np.array_equal(a, b) and ia == ib


def test_make_filter_rejects_bad_kind():
    lut = np.zeros(256, np.uint8)
    w, inv = C.c_int(), C.c_float()
    assert vv.LIB.vv_make_filter(9, 10, 0.0, lut.ctypes.data, C.byref(w), C.byref(inv)) == 1


def test_create_without_gpu_reports_error(monkeypatch):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(vv.VVError) as e:
        vv.Context(0)
    assert e.value.status == 3  # VV_EDEVICE
    assert "device" in vv.LIB.vv_last_error(None).decode()


def test_null_arguments_rejected():
    assert vv.LIB.vv_default_lic_params(None) == 1
    assert vv.LIB.vv_default_camera(None, 1, 1, None) == 1
    assert vv.LIB.vv_synchronize(None) == 1
    assert vv.LIB.vv_destroy(None) == 0


def test_binding_refuses_short_or_mistyped_buffers():
    # the library reads or writes as many elements as the dims say; the
    # binding checks every buffer before its address reaches the C ABI
    a = np.zeros(12, np.float32)
    assert _lib._buf(a, "float32", 12, "field")[0] == a.ctypes.data
    with pytest.raises(vv.VVError) as e:
        _lib._buf(a, "float32", 13, "field")
    assert e.value.status == 1 and "13 elements needed" in str(e.value)
    with pytest.raises(vv.VVError) as e:
        _lib._buf(a.astype(np.float64), "float32", 12, "field")
    assert "float32 buffer expected, got float64" in str(e.value)
    with pytest.raises(vv.VVError):
        _lib._buf(np.zeros(8, np.uint8), "float32", 2, "image")
    torch = pytest.importorskip("torch")
    t = torch.zeros(12, dtype=torch.float32)
    assert _lib._buf(t, "float32", 12, "field") == (t.data_ptr(), 0)
    with pytest.raises(vv.VVError):
        _lib._buf(t, "uint8", 12, "noise")
    with pytest.raises(ValueError):
        _lib._buf(np.zeros((4, 4), np.float32)[:, ::2], "float32", 8, "field")  # not contiguous


def test_every_entry_point_keeps_exceptions_inside():
    # the C ABI lets no C++ exception through: every vv_status entry point
    # runs its body inside guarded(), which maps std::bad_alloc to VV_ENOMEM
    csrc = Path(__file__).resolve().parent.parent / "vectorvisualization_amd" / "csrc"
    bare = []
    n = 0
    for f in sorted(csrc.glob("*.cpp")):
        text = f.read_text()
        for m in re.finditer(r"^vv_status (vv_\w+)\([^;{]*?\)\n\{\n(.*)$", text, re.M):
            n += 1
            if not m.group(2).strip().startswith("return guarded("):
                bare.append(f"{f.name}:{m.group(1)}")
    assert n >= 60, n  # 69 entry points return vv_status
    assert not bare, bare


class _FakeDeviceTensor:
    """Stands in for a CUDA tensor on a machine without a GPU: the binding
    reads only these attributes before it reaches the C ABI."""

    is_cuda = True
    dtype = "torch.float32"

    def __init__(self, n):
        self._n = n

    def numel(self):
        return self._n

    def data_ptr(self):
        return 0x1000

    def is_contiguous(self):
        return True


class _NoLib:
    def __getattr__(self, name):
        raise AssertionError(f"{name} reached the C ABI with mixed host and device buffers")


def test_adjoint_and_batch_refuse_mixed_host_and_device_buffers(monkeypatch):
    # one on_device flag covers every buffer of vv_render_adjoint and
    # vv_batch_set_field: a host buffer beside a device one must be refused
    # before any library call
    monkeypatch.setattr(_lib, "LIB", _NoLib())
    ctx = object.__new__(_lib.Context)
    ctx._h = None
    ctx._lic_voxels = lambda: 8
    cam = _lib.Camera()
    cam.width, cam.height = 2, 2
    dev_img, host_img = _FakeDeviceTensor(16), np.zeros(16, np.float32)
    for dimage, dtf, dsigma in [(dev_img, np.zeros(1024, np.float32), None),
                                (dev_img, _FakeDeviceTensor(1024), np.zeros(8, np.float32)),
                                (host_img, _FakeDeviceTensor(1024), None)]:
        with pytest.raises(ValueError, match="all be host memory or all device memory"):
            ctx.render_adjoint(cam, dimage, dtf, dsigma)
    b = object.__new__(_lib.Batch)
    b._ctx, b._dims, b._h = ctx, [(2, 2, 2)], None
    with pytest.raises(ValueError, match="v0 on device, v1 on host"):
        b.set_field(0, _FakeDeviceTensor(24), np.zeros(24, np.float32), 0.5)
    with pytest.raises(ValueError, match="v0 on host, v1 on device"):
        b.set_field(0, np.zeros(24, np.float32), _FakeDeviceTensor(24), 0.5)
