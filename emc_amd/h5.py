"""Optional HDF5 file formats (include/emc_h5.h) through libemc_h5.so, present where the library was
built against an HDF5 installation.  Host code only."""
import ctypes as C
import os

import numpy as np

from ._lib import check, lib as _emc_lib

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "libemc_h5.so")
_lib = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def available() -> bool:
    return os.path.exists(_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"{_PATH} is not built (no HDF5 library where the package was built)")
        _emc_lib()  # libemc.so first: libemc_h5.so reports errors through it
        L = C.CDLL(_PATH)
        L.emc_h5_image_dims.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.emc_h5_read_pattern.argtypes = [C.c_char_p, C.c_int32, C.c_int32, f32p, i32p]
        L.emc_h5_read_mask.argtypes = [C.c_char_p, C.c_int32, C.c_int32, i32p]
        L.emc_h5_write_image.argtypes = [C.c_char_p, C.c_int32, C.c_int32, f32p, C.c_void_p]
        L.emc_h5_write_model.argtypes = [C.c_char_p, C.c_int32, f32p, f32p]
        L.emc_h5_read_model.argtypes = [C.c_char_p, C.c_int32, f32p]
        for n in ("emc_h5_image_dims", "emc_h5_read_pattern", "emc_h5_read_mask", "emc_h5_write_image",
                  "emc_h5_write_model", "emc_h5_read_model"):
            getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


def image_dims(path):
    nx, ny = C.c_int32(0), C.c_int32(0)
    check(lib().emc_h5_image_dims(str(path).encode(), C.byref(nx), C.byref(ny)))
    return nx.value, ny.value


def read_pattern(path, L: int, read_stride: int):
    """(pattern [L*L], mask [L*L]) binned as read_images does (emc_helper.cpp:198-248)."""
    p = np.zeros(L * L, np.float32)
    m = np.zeros(L * L, np.int32)
    check(lib().emc_h5_read_pattern(str(path).encode(), L, read_stride, p, m))
    return p, m


def read_mask(path, L: int, read_stride: int):
    m = np.zeros(L * L, np.int32)
    check(lib().emc_h5_read_mask(str(path).encode(), L, read_stride, m))
    return m


def write_image(path, real, mask=None):
    """real [ny][nx] in the engine's layout (pixel y * nx + x); mask likewise or None."""
    real = np.ascontiguousarray(real, np.float32)
    ny, nx = real.shape
    mk = None if mask is None else np.ascontiguousarray(mask, np.float32).reshape(ny, nx)
    check(lib().emc_h5_write_image(str(path).encode(), nx, ny, real,
                                   None if mk is None else mk.ctypes.data_as(C.c_void_p)))


def write_model(path, L: int, model, weight):
    check(lib().emc_h5_write_model(str(path).encode(), L, np.ascontiguousarray(model, np.float32).ravel(),
                                   np.ascontiguousarray(weight, np.float32).ravel()))


def read_model(path, L: int):
    m = np.zeros(L ** 3, np.float32)
    check(lib().emc_h5_read_model(str(path).encode(), L, m))
    return m
