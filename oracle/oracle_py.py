"""ctypes binding of the test oracle (libsmx_oracle.so) and of the reference
harnesses (_ref/libsmx_ref.so, _ref/libref_tin.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from setsm_amd.records import (NCCRESULT_DTYPE, POINT3_DTYPE, TRI_DTYPE, UGRID_DTYPE, SmxBdArgs, SmxHrArgs,
                               SmxLevel, SmxMpsArgs)

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None
_tin = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libsmx_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        L.orc_voxels_create.restype = vp
        L.orc_voxels_create.argtypes = [C.c_long, ip]
        L.orc_voxels_destroy.argtypes = [vp]
        L.orc_voxels_export.restype = C.c_long
        L.orc_voxels_export.argtypes = [vp, vp, vp, vp]
        L.orc_voxels_import.argtypes = [vp, vp, vp]
        L.orc_rpc_ground_to_image.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_height_step.restype = C.c_double
        L.orc_height_step.argtypes = [ip, C.c_double]
        L.orc_vll_half_template.restype = ip
        L.orc_vll_half_template.argtypes = [ip, ip, C.c_double]
        L.orc_window_bits.restype = ip
        L.orc_window_bits.argtypes = [ip, ip, ip]
        L.orc_correlate.restype = C.c_double
        L.orc_correlate.argtypes = [vp, vp, ip]
        L.orc_check_image_boundary.restype = ip
        L.orc_check_image_boundary.argtypes = [vp, C.c_long, C.c_double, C.c_double, ip]
        L.orc_ncc_alpha.restype = C.c_double
        L.orc_ncc_alpha.argtypes = [ip, ip, ip]
        L.orc_initialize_voxel.argtypes = [vp, vp, vp, vp, ip, vp]
        L.orc_vertical_line_locus.restype = C.c_long
        L.orc_vertical_line_locus.argtypes = [vp, vp, vp, vp, ip, vp]
        L.orc_awncc_peak.argtypes = [vp, vp, vp, vp, ip]
        L.orc_awncc.argtypes = [vp, vp, vp, vp, ip, vp]
        L.orc_sgm_aggregate.argtypes = [vp, vp, vp, vp, vp]
        L.orc_gauss_pyramid.argtypes = [vp, C.c_uint32, C.c_uint32, vp]
        L.orc_sobel.argtypes = [vp, C.c_uint32, C.c_uint32, vp, vp]
        L.orc_orientation.argtypes = [vp, vp, C.c_uint32, C.c_uint32, ip, vp]
        L.orc_select_mps.restype = C.c_long
        L.orc_select_mps.argtypes = [vp, vp, vp, vp, vp, dp]
        L.orc_set_height_range.restype = ip
        L.orc_set_height_range.argtypes = [vp, vp, vp, C.c_long, vp, C.c_long, vp, dp, dp]
        L.orc_set_height_range_blunder.restype = None
        L.orc_set_height_range_blunder.argtypes = [vp, vp, C.c_long, vp, C.c_long, vp]
        L.orc_blunder_half_template.restype = ip
        L.orc_blunder_half_template.argtypes = [vp, ip, ip, ip]
        L.orc_vll_blunder.restype = ip
        L.orc_vll_blunder.argtypes = [vp, vp, ip, vp, ip, ip, vp, vp]
        L.orc_blunder_detection_tin.restype = ip
        L.orc_blunder_detection_tin.argtypes = [vp, vp, vp, vp, C.c_long, vp, C.c_long, vp, vp, vp, dp, dp, dp]
        L.orc_set_threads.argtypes = [ip]
        L.orc_get_threads.restype = ip
        _lib = L
    return _lib


def ref():
    """The reference's own header code, or None where it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libsmx_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        R.ref_rpc_ground_to_image.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double, dp, dp]
        R.ref_to_pyramid.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp]
        R.ref_gauss_pyramid.argtypes = [vp, C.c_uint, C.c_uint, vp]
        R.ref_q_result.restype = C.c_short
        R.ref_q_result.argtypes = [C.c_double]
        R.ref_q_voxel.restype = C.c_short
        R.ref_q_voxel.argtypes = [C.c_double]
        R.ref_layout.restype = C.c_long
        R.ref_layout.argtypes = [C.c_int]
        _ref = R
    return _ref


def tin():
    """The reference's TIN (grid_triangulation.cpp via ref_tin.cpp), or None
    where it was not built."""
    global _tin
    if _tin is None:
        path = os.path.join(HERE, "_ref", "libref_tin.so")
        if not os.path.exists(path):
            return None
        T = C.CDLL(path)
        T.ref_tin_create.restype = C.c_long
        T.ref_tin_create.argtypes = [C.c_void_p, C.c_long, C.POINTER(C.c_double), C.c_double, C.c_void_p, C.c_long]
        _tin = T
    return _tin


def tin_create(pts, boundary, resolution):
    """TINCreate_list (SubFunctions.cpp:2768-2846) with the reference's own
    triangulation: triangles (TRI_DTYPE) over the matched points."""
    T = tin()
    if T is None:
        raise RuntimeError("oracle/_ref/libref_tin.so is not built (needs the reference sources)")
    pts = np.ascontiguousarray(pts, POINT3_DTYPE)
    out = np.zeros(2 * len(pts) + 16, TRI_DTYPE)
    mm = (C.c_double * 4)(*[float(b) for b in boundary])
    n = T.ref_tin_create(_p(pts) if len(pts) else None, len(pts), mm, float(resolution), _p(out), len(out))
    if n < 0:
        raise RuntimeError(f"reference TIN failed ({n})")
    return out[:n].copy()


def select_mps(level_inputs, ncc, ugrid, iteration, th_roh, th_roh_min, th_roh_start, th_roh_next, mpp,
               mpp_stereo_angle, final_level_iteration):
    """SelectMPs (setsm_code.cpp:8535-8889): returns (points, minimum_th);
    ugrid is updated in place."""
    a = SmxMpsArgs(th_roh, th_roh_min, th_roh_start, th_roh_next, mpp, mpp_stereo_angle, int(iteration),
                   int(final_level_iteration))
    ncc = np.ascontiguousarray(ncc, NCCRESULT_DTYPE)
    assert ugrid.dtype == UGRID_DTYPE and ugrid.flags.c_contiguous
    pts = np.zeros(max(len(ncc), 1), POINT3_DTYPE)
    th = C.c_double(0)
    n = lib().orc_select_mps(C.addressof(level_inputs.level), _p(ncc), _p(ugrid), C.addressof(a), _p(pts),
                             C.byref(th))
    if n < 0:
        raise RuntimeError("SelectMPs has no threshold at this level")
    return pts[:n].copy(), th.value


def set_height_range(level_inputs, ugrid, pts, tris, iteration, mpp, check_matching_rate=False,
                     seed_dem_sigma=0.0):
    """SetHeightRange (setsm_code.cpp:9851-10261): ugrid becomes the returned
    grid; returns (minH_grid, maxH_grid)."""
    a = SmxHrArgs(mpp, seed_dem_sigma, int(iteration), int(bool(check_matching_rate)))
    pts = np.ascontiguousarray(pts, POINT3_DTYPE)
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    assert ugrid.dtype == UGRID_DTYPE and ugrid.flags.c_contiguous
    lo, hi = C.c_double(0), C.c_double(0)
    r = lib().orc_set_height_range(C.addressof(level_inputs.level), C.addressof(a),
                                   _p(pts) if len(pts) else None, len(pts), _p(tris) if len(tris) else None,
                                   len(tris), _p(ugrid), C.byref(lo), C.byref(hi))
    if r < 0:
        raise MemoryError("oracle SetHeightRange ran out of memory")
    return lo.value, hi.value


def set_height_range_blunder(level_inputs, ugrid, pts, tris):
    """SetHeightRange_blunder (setsm_code.cpp:10393-10491): ugrid's Height in place."""
    pts = np.ascontiguousarray(pts, POINT3_DTYPE)
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    assert ugrid.dtype == UGRID_DTYPE and ugrid.flags.c_contiguous
    lib().orc_set_height_range_blunder(C.addressof(level_inputs.level), _p(pts) if len(pts) else None, len(pts),
                                       _p(tris) if len(tris) else None, len(tris), _p(ugrid))


def blunder_half_template(level_inputs, blunder_level, iteration, bblunder):
    return lib().orc_blunder_half_template(C.addressof(level_inputs.level), int(blunder_level), int(iteration),
                                           int(bool(bblunder)))


def vll_blunder(level_inputs, pyr, blunder_level, ortho_wgs, ugrid, iteration, bblunder, start=((0.0, 0.0), (0.0, 0.0))):
    """VerticalLineLocus_blunder (setsm_code.cpp:7882-8013) on the pyramid
    level blunder_level of pyr (per image (img, mag, ori) levels): returns the
    per-cell ortho NCC; ugrid's ortho_ncc[1] / Mean_ortho_ncc in place."""
    from setsm_amd.records import SmxView
    views = (SmxView * 2)()
    keep = []
    for k in range(2):
        img, mag, ori = (np.ascontiguousarray(a) for a in pyr[k][blunder_level])
        keep += [img, mag, ori]
        v = views[k]
        v.img, v.mag, v.ori = img.ctypes.data, mag.ctypes.data, ori.ctypes.data
        v.height, v.width = img.shape
        v.start_x = start[k][0] / (1 << blunder_level)
        v.start_y = start[k][1] / (1 << blunder_level)
    ow = np.ascontiguousarray(ortho_wgs, np.float64)
    assert ugrid.dtype == UGRID_DTYPE and ugrid.flags.c_contiguous
    out = np.zeros(len(ugrid), np.float32)
    r = lib().orc_vll_blunder(C.addressof(level_inputs.level), views, int(blunder_level), _p(ow), int(iteration),
                              int(bool(bblunder)), _p(ugrid), _p(out))
    if r < 0:
        raise MemoryError("oracle VerticalLineLocus_blunder ran out of memory")
    return out


def blunder_detection_tin(level_inputs, ugrid, ortho_ncc, pts, tris, iteration, flag_blunder, hinterval,
                          count_bl=0, blunder_count=(0, 0), minz=100000.0, maxz=-100000.0):
    """blunder_detection_TIN (setsm_code.cpp:9158-9782): pts' flags and
    ugrid's anchor_flag in place. Returns (detected, blunder_count, minz,
    maxz, height_th); the counts and heights continue from the values given,
    as DecisionMPs' running totals do."""
    a = SmxBdArgs(float(hinterval), int(iteration), int(bool(flag_blunder)), int(count_bl), 0)
    assert pts.dtype == POINT3_DTYPE and pts.flags.c_contiguous
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    assert ugrid.dtype == UGRID_DTYPE and ugrid.flags.c_contiguous
    on = np.ascontiguousarray(ortho_ncc, np.float32)
    assert len(on) == len(ugrid)
    det = np.zeros(max(len(pts), 1), np.uint8)
    bc = (C.c_long * 2)(*blunder_count)
    lo, hi, th = C.c_double(minz), C.c_double(maxz), C.c_double(float("nan"))
    r = lib().orc_blunder_detection_tin(C.addressof(level_inputs.level), C.addressof(a), _p(on),
                                        _p(pts) if len(pts) else None, len(pts), _p(tris) if len(tris) else None,
                                        len(tris), _p(ugrid), _p(det), bc, C.byref(lo), C.byref(hi), C.byref(th))
    if r < 0:
        raise MemoryError("oracle blunder_detection_TIN ran out of memory")
    return det[:len(pts)].astype(bool), (bc[0], bc[1]), lo.value, hi.value, th.value


class ChainJob:
    """select_mps / set_height_range with the signatures of `smx.Job`, on the
    oracle's NCCresult and GridPT3 (ugrid updated in place), for driving
    `setsm_amd.host_loop.iteration_chain` on the CPU."""

    def __init__(self, level_inputs, ncc, ugrid):
        self.li, self.ncc, self.ug = level_inputs, ncc, ugrid
        self.ncells = len(ugrid)

    def select_mps(self, iteration, mpp, **kw):
        return select_mps(self.li, self.ncc, self.ug, iteration, mpp=mpp, **kw)

    def set_height_range(self, pts, tris, iteration, mpp, check_matching_rate=False):
        return set_height_range(self.li, self.ug, pts, tris, iteration, mpp, check_matching_rate)


class ChainOps:
    """DecisionMPs' three steps (the `ops` of `setsm_amd.host_loop`) on the
    oracle, the blunder level's ortho NCC kept between them."""

    def __init__(self, level_inputs, ugrid, pyr, blunder_level, ortho_wgs):
        self.li, self.ug, self.pyr, self.B, self.ow = level_inputs, ugrid, pyr, blunder_level, ortho_wgs
        self.on = None

    def set_height_range_blunder(self, pts, tris):
        set_height_range_blunder(self.li, self.ug, pts, tris)

    def vll_blunder(self, iteration, bblunder):
        self.on = vll_blunder(self.li, self.pyr, self.B, self.ow, self.ug, iteration, bblunder)

    def blunder_detection(self, pts, tris, iteration, flag_blunder, hinterval, count_bl, zero_ortho, minz, maxz):
        on = np.zeros(len(self.ug), np.float32) if zero_ortho else self.on
        det, bc, lo, hi, _ = blunder_detection_tin(self.li, self.ug, on, pts, tris, iteration, flag_blunder,
                                                   hinterval, count_bl=count_bl, minz=minz, maxz=maxz)
        return det, bc, lo, hi


def _p(a):
    return a.ctypes.data


# ----------------------------------------------------------------- pyramid
def gauss_pyramid(img):
    h, w = img.shape
    out = np.zeros((h // 2, w // 2), np.uint16)
    lib().orc_gauss_pyramid(_p(np.ascontiguousarray(img, np.uint16)), w, h, _p(out))
    return out


def sobel(img):
    img = np.ascontiguousarray(img, np.uint16)
    h, w = img.shape
    mag = np.zeros_like(img)
    d = np.zeros((h, w), np.int16)
    lib().orc_sobel(_p(img), w, h, _p(mag), _p(d))
    return mag, d


def orientation(mag, d, template_size=15):
    h, w = mag.shape
    ori = np.zeros((h, w), np.uint8)
    lib().orc_orientation(_p(np.ascontiguousarray(mag)), _p(np.ascontiguousarray(d)), w, h, template_size, _p(ori))
    return ori


def build_pyramid(img0, levels):
    """SetPyramidImages (SubFunctions.cpp:3383-3406): per level (img, mag, ori)."""
    out = []
    img = np.ascontiguousarray(img0, np.uint16)
    for lv in range(levels):
        if lv > 0:
            img = gauss_pyramid(img)
        mag, d = sobel(img)
        ori = orientation(mag, d, 15)
        out.append((img, mag, ori))
    return out


# -------------------------------------------------------------- hot path
class OracleState:
    """nccresult + voxel rows for one (tile, level), driven like SETSM."""

    def __init__(self, level_inputs, keep_rho=True):
        self.li = level_inputs
        self.n = level_inputs.ncells
        self.ncc = np.zeros(self.n, NCCRESULT_DTYPE)
        self.vox = lib().orc_voxels_create(self.n, 1 if keep_rho else 0)

    def __del__(self):
        try:
            lib().orc_voxels_destroy(self.vox)
        except Exception:
            pass

    @property
    def L(self):
        return C.addressof(self.li.level)

    def initialize_voxel(self, ugrid, iteration, minmax):
        mm = np.ascontiguousarray(minmax, np.float64)
        lib().orc_initialize_voxel(self.L, self.vox, _p(ugrid), _p(self.ncc), iteration, _p(mm))

    def vertical_line_locus(self, ugrid, iteration, minmax):
        mm = np.ascontiguousarray(minmax, np.float64)
        r = lib().orc_vertical_line_locus(self.L, self.vox, _p(self.ncc), _p(ugrid), iteration, _p(mm))
        if r < 0:
            raise RuntimeError(f"oracle VerticalLineLocus refused the input (code {r})")
        return r

    def awncc(self, ugrid, iteration):
        """AWNCC as the reference runs it: SGM aggregation + SGM peak mode at
        levels <= SGM_py (SumCost kept in self.sumcost), plain scan above."""
        self.sumcost = None
        if self.li.level.level <= self.li.level.sgm_py:
            total = int(self.ncc["NumOfHeight"].astype(np.int64).sum())
            self.sumcost = np.zeros(max(total, 1), np.float32)
            lib().orc_awncc(self.L, self.vox, _p(ugrid), _p(self.ncc), iteration, _p(self.sumcost))
            self.sumcost = self.sumcost[:total]
        else:
            lib().orc_awncc(self.L, self.vox, _p(ugrid), _p(self.ncc), iteration, None)

    def sgm_aggregate(self, ugrid):
        """SumCost of the current cube (CSR order) without the peak scan."""
        total = int(self.ncc["NumOfHeight"].astype(np.int64).sum())
        out = np.zeros(max(total, 1), np.float32)
        lib().orc_sgm_aggregate(self.L, self.vox, _p(ugrid), _p(self.ncc), _p(out))
        return out[:total]

    def import_voxels(self, ncc, words):
        """Seed NCCresult and the voxel rows (CSR words) directly."""
        self.ncc[:] = ncc
        w = np.ascontiguousarray(words, np.uint32)
        lib().orc_voxels_import(self.vox, _p(self.ncc), _p(w))

    def export(self):
        total = int(self.ncc["NumOfHeight"].astype(np.int64).sum())
        words = np.zeros(max(total, 1), np.uint32)
        rho = np.zeros(max(total, 1), np.float64)
        lib().orc_voxels_export(self.vox, _p(self.ncc), _p(words), _p(rho))
        return words[:total], rho[:total]


def set_threads(n):
    lib().orc_set_threads(int(n))


def threads():
    return lib().orc_get_threads()


def rpc_ground_to_image(rpc90, adjust, lon, lat, h):
    x, y = C.c_double(), C.c_double()
    r = np.ascontiguousarray(rpc90, np.float64)
    a = np.ascontiguousarray(adjust, np.float64)
    lib().orc_rpc_ground_to_image(_p(r), _p(a), lon, lat, h, C.byref(x), C.byref(y))
    return x.value, y.value


__all__ = ["lib", "ref", "build", "OracleState", "build_pyramid", "gauss_pyramid", "sobel",
           "orientation", "set_threads", "threads", "rpc_ground_to_image", "UGRID_DTYPE", "SmxLevel"]
