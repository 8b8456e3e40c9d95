"""Host binding of libsmx_io.so (include/smx_io.h): SETSM's image and model
formats.

    Readtiff_T            Template.hpp:211-360        -> read_tiff
    ReadGeotiff_info(_dxy) setsmgeo.cpp:150-197       -> tiff_info
    WriteGeotiff(_round)  setsmgeo.cpp:22-116         -> write_geotiff
    OpenXMLFile (RPC)     SubFunctions.cpp:1302-1582  -> read_dg_xml

Failures raise IOError with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .records import rpc_struct
from .records import SmxRpc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsmx_io.so")

UCHAR, FLOAT, UINT16 = 1, 4, 12
_NP = {UCHAR: np.uint8, UINT16: np.uint16, FLOAT: np.float32}
EXPORTS = ["smx_io_last_error", "smx_tiff_info_read", "smx_tiff_read", "smx_geotiff_write", "smx_dg_xml_read",
           "smx_dg_xml_write"]


class TiffInfo(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("bits", C.c_uint32), ("sample_format", C.c_uint32),
                ("samples", C.c_uint32), ("compression", C.c_uint32), ("predictor", C.c_uint32),
                ("tiled", C.c_uint32), ("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("bigtiff", C.c_uint32),
                ("has_geo", C.c_int32), ("min_x", C.c_double), ("max_y", C.c_double), ("grid_dx", C.c_double),
                ("grid_dy", C.c_double)]


class DgInfo(C.Structure):
    _fields_ = [("gsd_row", C.c_double), ("gsd_col", C.c_double), ("gsd", C.c_double),
                ("abscalfactor", C.c_double), ("effbw", C.c_double), ("tdi", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IOError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.smx_io_last_error.restype = C.c_char_p
        L.smx_tiff_info_read.argtypes = [C.c_char_p, C.POINTER(TiffInfo)]
        L.smx_tiff_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_long), C.c_void_p,
                                    C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.smx_geotiff_write.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.smx_dg_xml_read.argtypes = [C.c_char_p, C.POINTER(SmxRpc), C.POINTER(C.c_double), C.POINTER(DgInfo)]
        L.smx_dg_xml_write.argtypes = [C.c_char_p, C.POINTER(SmxRpc), C.POINTER(C.c_double), C.POINTER(DgInfo)]
        _lib = L
    return _lib


def _check(r):
    if r != 0:
        raise IOError(lib().smx_io_last_error().decode())


def tiff_info(path) -> TiffInfo:
    info = TiffInfo()
    _check(lib().smx_tiff_info_read(os.fsencode(path), C.byref(info)))
    return info


def read_tiff(path, dtype_code=None, cols=None, rows=None):
    """Readtiff_T: returns (array, (c0, c1), (r0, r1)) — the window actually
    read (tiled files widen it to whole tiles, as the reference does)."""
    info = tiff_info(path)
    if dtype_code is None:
        dtype_code = {(8, 1): UCHAR, (16, 1): UINT16, (32, 3): FLOAT}.get((info.bits, info.sample_format))
        if dtype_code is None:
            raise IOError(f"{path}: {info.bits}-bit samples of format {info.sample_format}")
    c = (C.c_long * 2)(*(cols if cols is not None else (0, info.width)))
    r = (C.c_long * 2)(*(rows if rows is not None else (0, info.height)))
    # the widest window a tiled file can return
    cw = min(info.width, (c[1] // max(info.tile_w, 1) + 1) * max(info.tile_w, 1)) if info.tiled else c[1] - c[0]
    rh = min(info.height, (r[1] // max(info.tile_h, 1) + 1) * max(info.tile_h, 1)) if info.tiled else r[1] - r[0]
    out = np.empty(max(1, cw * rh), dtype=_NP[dtype_code])
    w, h = C.c_uint32(), C.c_uint32()
    _check(lib().smx_tiff_read(os.fsencode(path), dtype_code, c, r, out.ctypes.data, out.nbytes, C.byref(w),
                               C.byref(h)))
    return out[: w.value * h.value].reshape(h.value, w.value).copy(), (c[0], c[1]), (r[0], r[1])


def write_geotiff(path, array, scale, min_x, max_y, projection=0, zone=1, north=True, round_128=False):
    """WriteGeotiff / WriteGeotiff_round of a 2-D uint8 / uint16 / float32 array."""
    a = np.ascontiguousarray(array)
    code = {np.dtype(np.uint8): UCHAR, np.dtype(np.uint16): UINT16, np.dtype(np.float32): FLOAT}.get(a.dtype)
    if code is None or a.ndim != 2:
        raise IOError(f"write_geotiff: 2-D uint8 / uint16 / float32 expected, got {a.dtype} {a.shape}")
    _check(lib().smx_geotiff_write(os.fsencode(path), a.ctypes.data, a.shape[1], a.shape[0], float(scale),
                                   float(min_x), float(max_y), int(projection), int(zone), 1 if north else 0, code,
                                   1 if round_128 else 0))


def read_dg_xml(path):
    """OpenXMLFile (RPC branch): returns (rpc90, err[2], DgInfo); rpc90 in the
    library's order (offsets and scales {line, samp, lon, lat, height}, then
    the four 20-term coefficient rows)."""
    rpc = SmxRpc()
    err = (C.c_double * 2)()
    info = DgInfo()
    _check(lib().smx_dg_xml_read(os.fsencode(path), C.byref(rpc), err, C.byref(info)))
    flat = np.frombuffer(bytes(rpc), dtype=np.float64).copy()
    return flat, np.array(err[:]), info


def write_dg_xml(path, rpc90, err=(0.0, 0.0), gsd=(0.5, 0.5, 0.5)):
    rpc = rpc_struct(np.asarray(rpc90, np.float64))
    e = (C.c_double * 2)(*err)
    info = DgInfo(gsd[0], gsd[1], gsd[2], 0.0, 0.0, 0.0)
    _check(lib().smx_dg_xml_write(os.fsencode(path), C.byref(rpc), e, C.byref(info)))
