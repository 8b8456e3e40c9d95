"""ctypes / numpy mirrors of include/smx_types.h.

UGRID_DTYPE, NCCRESULT_DTYPE and the VOXEL word helpers follow SETSM's
records (Typedefine.hpp:248-293) byte for byte; SmxLevel mirrors smx_level.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

UGRID_DTYPE = np.dtype([
    ("Height", "<f4"), ("minHeight", "<i2"), ("maxHeight", "<i2"), ("roh", "<i2"),
    ("ortho_ncc", "<i2", (2,)), ("Mean_ortho_ncc", "<i2"), ("Matched_flag", "u1"),
    ("anchor_flag", "u1"), ("_pad", "V2"),
])
assert UGRID_DTYPE.itemsize == 20

NCCRESULT_DTYPE = np.dtype([
    ("result2", "<f4"), ("result3", "<f4"), ("minHeight", "<i2"), ("maxHeight", "<i2"),
    ("result0", "<i2"), ("result1", "<i2"), ("max_WNCC", "<i2"), ("GNCC", "<i2"),
    ("NumOfHeight", "<u2"), ("result4", "u1"), ("check_height_change", "u1"),
])
assert NCCRESULT_DTYPE.itemsize == 24


# D3DPOINT (Typedefine.hpp:140-194) and UI3DPOINT (Typedefine.hpp:74-97)
POINT3_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("flag", "u1"), ("_pad", "V7")])
assert POINT3_DTYPE.itemsize == 32
TRI_DTYPE = np.dtype([("a", "<u4"), ("b", "<u4"), ("c", "<u4")])
assert TRI_DTYPE.itemsize == 12


class SmxMpsArgs(C.Structure):
    """smx_mps_args: SelectMPs' per-call arguments (setsm_code.cpp:8535)."""
    _fields_ = [("th_roh", C.c_double), ("th_roh_min", C.c_double), ("th_roh_start", C.c_double),
                ("th_roh_next", C.c_double), ("mpp", C.c_double), ("mpp_stereo_angle", C.c_double),
                ("iteration", C.c_int32), ("final_level_iteration", C.c_int32)]


class SmxHrArgs(C.Structure):
    """smx_hr_args: SetHeightRange's per-call arguments (setsm_code.cpp:9851)."""
    _fields_ = [("mpp", C.c_double), ("seed_dem_sigma", C.c_double), ("iteration", C.c_int32),
                ("check_matching_rate", C.c_int32)]


class SmxBdArgs(C.Structure):
    """smx_bd_args: blunder_detection_TIN's per-call arguments (setsm_code.cpp:9158)."""
    _fields_ = [("hinterval", C.c_double), ("iteration", C.c_int32), ("flag_blunder", C.c_int32),
                ("count_bl", C.c_int32), ("pad_", C.c_int32)]


def voxel_flag(words):
    return (np.asarray(words, dtype=np.uint32) & 0xFF).astype(np.uint8)


def voxel_incc(words):
    return (np.asarray(words, dtype=np.uint32) >> 16).astype(np.uint16).view(np.int16)


class SmxRpc(C.Structure):
    _fields_ = [("off", C.c_double * 5), ("scale", C.c_double * 5), ("coef", (C.c_double * 20) * 4)]


class SmxView(C.Structure):
    _fields_ = [("img", C.c_void_p), ("mag", C.c_void_p), ("ori", C.c_void_p),
                ("width", C.c_uint32), ("height", C.c_uint32),
                ("start_x", C.c_double), ("start_y", C.c_double)]


class SmxLevel(C.Structure):
    _fields_ = [
        ("view", SmxView * 2), ("next", SmxView * 2), ("rpc", SmxRpc * 2),
        ("adjust", (C.c_double * 2) * 2),
        ("level", C.c_int32), ("template_size", C.c_int32), ("py_combined_level", C.c_int32),
        ("sgm_py", C.c_int32), ("is_ra", C.c_int32), ("check_matchtag", C.c_int32),
        ("pre_dem", C.c_int32), ("check_matching_rate", C.c_int32),
        ("resolution", C.c_double), ("dem_resolution", C.c_double), ("bin_angle", C.c_double),
        ("height_step", C.c_double), ("grid_resolution", C.c_double), ("boundary", C.c_double * 4),
        ("grid_w", C.c_uint32), ("grid_h", C.c_uint32),
        ("grid_pts", C.c_void_p), ("grid_wgs", C.c_void_p),
        ("reference_id", C.c_int32), ("ortho_w", C.c_uint32), ("ortho_h", C.c_uint32),
        ("ortho_wgs", C.c_void_p),
    ]


def rpc_struct(rpc90) -> SmxRpc:
    r = np.ascontiguousarray(rpc90, dtype=np.float64)
    assert r.shape == (90,)
    s = SmxRpc()
    C.memmove(C.addressof(s), r.ctypes.data, 90 * 8)
    return s


def height_step(level: int, resolution: float) -> float:
    """GetHeightStep (setsm_code.cpp:5741-5757)."""
    im = resolution * (1 << level)
    hs = im / 2.0
    hs = np.floor(hs * 1000.0 + 0.1) / 1000.0
    return float(min(hs, 3.0))


class LevelInputs:
    """Host arrays for one (tile, level) plus the SmxLevel that points at them.

    Keeps every numpy array alive as long as the struct is used.
    """

    def __init__(self, pyr, level, rpcs, grid, *, resolution=0.5, dem_resolution=2.0,
                 grid_resolution=2.0, template_size=15, sgm_py=-1, py_combined_level=0,
                 start=((0.0, 0.0), (0.0, 0.0)), adjust=((0.0, 0.0), (0.0, 0.0)),
                 is_ra=0, check_matchtag=0, pre_dem=0, check_matching_rate=0, ortho_wgs=None, reference_id=0):
        self.keep = []
        L = SmxLevel()
        for ti in range(2):
            for dst, lv in ((L.view, level), (L.next, level - 1)):
                if lv < 0 or lv >= len(pyr[ti]) or (dst is L.next and level <= py_combined_level):
                    continue
                img, mag, ori = (np.ascontiguousarray(a) for a in pyr[ti][lv])
                self.keep += [img, mag, ori]
                v = dst[ti]
                v.img, v.mag, v.ori = img.ctypes.data, mag.ctypes.data, ori.ctypes.data
                v.height, v.width = img.shape
                v.start_x = start[ti][0] / (1 << lv)
                v.start_y = start[ti][1] / (1 << lv)
            L.rpc[ti] = rpc_struct(rpcs[ti])
            L.adjust[ti][0], L.adjust[ti][1] = adjust[ti]
        L.level = level
        L.template_size = template_size
        L.py_combined_level = py_combined_level
        L.sgm_py = sgm_py
        L.is_ra, L.check_matchtag, L.pre_dem, L.check_matching_rate = is_ra, check_matchtag, pre_dem, check_matching_rate
        L.resolution = resolution
        L.dem_resolution = dem_resolution
        L.bin_angle = 360.0 / 18.0
        L.height_step = height_step(level, resolution)
        L.grid_resolution = grid_resolution
        for i in range(4):
            L.boundary[i] = float(grid["boundary"][i])
        gp = np.ascontiguousarray(grid["grid_pts"], dtype=np.float64)
        gw = np.ascontiguousarray(grid["grid_wgs"], dtype=np.float64)
        self.keep += [gp, gw]
        n = int(round(np.sqrt(gp.shape[0]))) if "shape" not in grid else None
        gh, gwid = grid.get("shape", (n, n))
        L.grid_w, L.grid_h = gwid, gh
        L.grid_pts, L.grid_wgs = gp.ctypes.data, gw.ctypes.data
        L.reference_id = reference_id
        if ortho_wgs is not None:
            ow = np.ascontiguousarray(ortho_wgs, dtype=np.float64)
            assert ow.ndim == 3 and ow.shape[2] == 2
            self.keep.append(ow)
            L.ortho_h, L.ortho_w = ow.shape[0], ow.shape[1]
            L.ortho_wgs = ow.ctypes.data
        self.level = L
        self.ncells = gh * gwid


def window_bits(half: int, row: int, col: int) -> int:
    """Nested-window membership of a window offset: bit 0 outer disc
    (setsm_code.cpp:6434), bits 1/2 the middle / inner windows
    (SubFunctions.cpp:3274-3312)."""
    r2 = row * row + col * col
    if abs(row) > half or abs(col) > half or r2 > (half + 1) ** 2:
        return 0
    bits = 1
    s1 = half // 2
    if r2 <= (half - s1 + 1) ** 2 and -half + s1 <= row <= half - s1 and -half + s1 <= col <= half - s1:
        bits |= 2
    s2 = s1 + int(s1 / 2.0 + 0.5)
    if r2 <= (half - s2 + 1) ** 2 and -half + s2 <= row <= half - s2 and -half + s2 <= col <= half - s2:
        bits |= 4
    return bits


def window_counts(half: int):
    """Samples of the outer, middle and inner windows at half-size `half`."""
    n = [0, 0, 0]
    for r in range(-half, half + 1):
        for c in range(-half, half + 1):
            b = window_bits(half, r, c)
            for k in range(3):
                if b & (1 << k):
                    n[k] += 1
    return n


def flops_per_hypothesis(half: int) -> int:
    """Algorithmic FP32 flops of one hypothesis: bilinear taps (4 channels x
    3 FMA per outer-window sample) + moments (5 FMA per sample per window per
    channel pair); the FP64 RPC work (~400 flops) is not counted."""
    n0, n1, n2 = window_counts(half)
    return 24 * n0 + 20 * (n0 + n1 + n2)
