"""vectorvisualization_amd — MI355X-native 3D LIC and volume ray casting.

The engine lives in libvvlic.so (HIP kernels for gfx950 behind the C ABI of
include/vvlic.h).  This package binds it with ctypes (``_lib``), generates
the synthetic inputs of the benchmark workloads on the GPU or on the host
(``fields``), and adds a ``torch.autograd.Function`` over the transfer-function
adjoint (``autograd``, imported lazily because it needs torch).
"""
from ._lib import (  # noqa: F401
    LIB,
    Anim,
    Batch,
    LIB_PATH,
    Camera,
    Context,
    LicParams,
    VVError,
    default_camera,
    default_lic_params,
    default_tf_u8,
    make_filter,
    STORE_FP16,
    STORE_FP32,
    STORE_UNORM8,
    ENCODE_INTERP,
    ENCODE_RAW_MAGNITUDE,
    FILTER_BOX,
    FILTER_COS2,
    FILTER_GAUSSIAN,
    FILTER_TRIANGLE,
    NOISE_BINARY,
    NOISE_UNIFORM,
    ILLUM_GRADIENT,
    ILLUM_LIC,
    ILLUM_MALLO,
    ILLUM_ZOECKLER,
    LineMaterial,
    default_line_material,
    TF_INDEX_B,
    TF_INDEX_MAG,
)

__all__ = [
    "Context", "Camera", "LicParams", "VVError", "default_camera", "default_lic_params",
    "default_tf_u8", "make_filter",
]
