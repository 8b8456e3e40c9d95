"""Host binding of libsmx.so (include/smx_vll.h) over ctypes.

`Job` owns one tile resident in HBM; its methods keep the names and meaning of
the reference entry points they stand in for:

    InitializeVoxel    setsm_code.cpp:5759-6012   -> Job.initialize_voxel
    VerticalLineLocus  setsm_code.cpp:6049-6548   -> Job.vertical_line_locus
    AWNCC (no SGM)     setsm_code.cpp:7511-7702   -> Job.awncc

Every failure raises SmxError with the library's message. There is no CPU
fallback: without the built library (setsm_amd/lib/libsmx.so) or a GPU the
calls fail.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .records import (NCCRESULT_DTYPE, POINT3_DTYPE, TRI_DTYPE, UGRID_DTYPE, SmxBdArgs, SmxHrArgs, SmxLevel,
                      SmxMpsArgs)

HERE = os.path.dirname(os.path.abspath(__file__))
# SMX_LIBRARY names another build of the same library (e.g. a kernel variant)
LIB_PATH = os.environ.get("SMX_LIBRARY") or os.path.join(HERE, "lib", "libsmx.so")

PREC_FP64 = 0
PREC_FAST = 1

EXPORTS = [
    "smx_last_error", "smx_version", "smx_set_device", "smx_job_create", "smx_job_destroy",
    "smx_job_device_bytes", "smx_job_set_budget", "smx_job_set_view", "smx_job_build_pyramid",
    "smx_job_get_view", "smx_job_set_level", "smx_job_set_searchspace", "smx_job_get_searchspace",
    "smx_job_get_nccresult", "smx_job_set_nccresult", "smx_initialize_voxel", "smx_vertical_line_locus",
    "smx_awncc", "smx_job_num_hypotheses", "smx_job_get_voxels", "smx_job_get_voxel_rows", "smx_job_set_voxels",
    "smx_job_get_sumcost", "smx_job_set_ortho_wgs", "smx_select_mps", "smx_set_height_range",
    "smx_set_height_range_blunder", "smx_job_set_blunder_level", "smx_vertical_line_locus_blunder",
    "smx_blunder_detection_tin", "smx_tin_create",
    "smx_job_last_kernel_ms", "smx_job_last_evaluated", "smx_job_last_redo", "smx_job_sync", "smx_fast_num_configs",
    "smx_job_set_fast_config", "smx_fast_config_desc", "smx_job_last_stats", "smx_vertical_line_locus_batch",
    "smx_vertical_line_locus_noalloc", "smx_release_cached", "smx_job_trim", "smx_job_half", "smx_graph_create", "smx_graph_launch", "smx_graph_last_ms", "smx_graph_destroy",
]


class SmxError(RuntimeError):
    pass


_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "-j4"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SmxError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, i, ll = C.c_void_p, C.c_int, C.c_longlong
        sig = {
            "smx_last_error": (C.c_char_p, []),
            "smx_version": (i, []),
            "smx_set_device": (i, [i]),
            "smx_job_create": (i, [C.POINTER(vp), i]),
            "smx_job_destroy": (i, [vp]),
            "smx_job_device_bytes": (ll, [vp]),
            "smx_job_set_budget": (i, [vp, ll]),
            "smx_job_set_view": (i, [vp, i, i, vp, vp, vp, C.c_uint32, C.c_uint32, C.c_double, C.c_double]),
            "smx_job_build_pyramid": (i, [vp, i, vp, C.c_uint32, C.c_uint32, C.c_double, C.c_double]),
            "smx_job_get_view": (i, [vp, i, i, vp, vp, vp]),
            "smx_job_set_level": (i, [vp, vp]),
            "smx_job_set_searchspace": (i, [vp, vp, vp]),
            "smx_job_get_searchspace": (i, [vp, vp]),
            "smx_job_get_nccresult": (i, [vp, vp]),
            "smx_job_set_nccresult": (i, [vp, vp]),
            "smx_initialize_voxel": (i, [vp, i]),
            "smx_vertical_line_locus": (ll, [vp, i, i, i]),
            "smx_awncc": (i, [vp, i]),
            "smx_job_num_hypotheses": (ll, [vp]),
            "smx_job_get_voxels": (i, [vp, vp, vp, vp]),
            "smx_job_get_voxel_rows": (i, [vp, ll, ll, vp, vp]),
            "smx_job_set_voxels": (i, [vp, vp]),
            "smx_job_get_sumcost": (i, [vp, vp]),
            "smx_job_set_ortho_wgs": (i, [vp, vp, C.c_uint32, C.c_uint32]),
            "smx_select_mps": (i, [vp, vp, vp, C.c_long, C.POINTER(C.c_long), C.POINTER(C.c_double)]),
            "smx_set_height_range": (i, [vp, vp, C.c_long, vp, C.c_long, vp, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]),
            "smx_set_height_range_blunder": (i, [vp, vp, C.c_long, vp, C.c_long]),
            "smx_job_set_blunder_level": (i, [vp, i, vp, C.c_uint32, C.c_uint32]),
            "smx_vertical_line_locus_blunder": (i, [vp, i, i, vp]),
            "smx_blunder_detection_tin": (i, [vp, vp, vp, vp, C.c_long, vp, C.c_long, vp, C.POINTER(C.c_long),
                                              C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "smx_tin_create": (C.c_long, [vp, C.c_long, vp, C.c_double, vp, C.c_long]),
            "smx_job_last_kernel_ms": (C.c_double, [vp]),
            "smx_job_last_evaluated": (ll, [vp]),
            "smx_job_last_redo": (i, [vp, vp, vp, vp]),
            "smx_job_sync": (i, [vp]),
            "smx_fast_num_configs": (i, []),
            "smx_job_set_fast_config": (i, [vp, i]),
            "smx_fast_config_desc": (i, [i, C.c_char_p, i]),
            "smx_job_last_stats": (i, [vp, vp]),
            "smx_vertical_line_locus_batch": (i, [vp, i, i, i, i, vp]),
            "smx_vertical_line_locus_noalloc": (ll, [vp, i, i]),
            "smx_release_cached": (ll, []),
            "smx_job_trim": (i, [vp]),
            "smx_job_half": (i, [vp]),
            "smx_graph_create": (i, [C.POINTER(vp), vp, i, i, i, i]),
            "smx_graph_launch": (i, [vp, vp]),
            "smx_graph_last_ms": (C.c_double, [vp]),
            "smx_graph_destroy": (i, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        msg = lib().smx_last_error().decode(errors="replace")
        raise SmxError(f"libsmx error {rc}: {msg}")
    return rc


def tin_create(pts, min_max, resolution):
    """Delaunay TIN of matched points on the level's lattice (host code; the
    role of SETSM's TINCreate_list): triangles (TRI_DTYPE, clockwise,
    indices into pts)."""
    pts = np.ascontiguousarray(pts, POINT3_DTYPE)
    mm = (C.c_double * 4)(*[float(v) for v in min_max[:4]])
    out = np.zeros(2 * len(pts) + 1, TRI_DTYPE)
    n = lib().smx_tin_create(_p(pts) if len(pts) else None, len(pts), mm, float(resolution), _p(out), len(out))
    if n < 0:
        raise SmxError(f"smx_tin_create failed ({n}): " + {-1: "lattice beyond 32767 cells or a point outside it",
                                                            -2: "triangle buffer too small"}.get(n, "bad argument"))
    return out[:n].copy()


def _p(a):
    return a.ctypes.data if a is not None else None


class Job:
    """One tile resident in HBM (smx_job)."""

    def __init__(self, num_levels: int, budget_bytes: int = 0):
        h = C.c_void_p()
        _check(lib().smx_job_create(C.byref(h), int(num_levels)))
        self.h = h
        self.num_levels = num_levels
        self.ncells = 0
        if budget_bytes:
            _check(lib().smx_job_set_budget(self.h, int(budget_bytes)))

    def close(self):
        if self.h:
            lib().smx_job_destroy(self.h)
            self.h = None

    def trim(self):
        """Free the per-call working lists (run / tile lists, hand-over queue,
        batch descriptors); the next call reserves them at the size it needs."""
        _check(lib().smx_job_trim(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------- pyramid
    def set_view(self, img, level, intensity, magnitude, orientation, start=(0.0, 0.0)):
        a = np.ascontiguousarray(intensity, np.uint16)
        m = np.ascontiguousarray(magnitude, np.uint16)
        o = np.ascontiguousarray(orientation, np.uint8)
        h, w = a.shape
        _check(lib().smx_job_set_view(self.h, img, level, _p(a), _p(m), _p(o), w, h, float(start[0]), float(start[1])))

    def build_pyramid(self, img, level0, start0=(0.0, 0.0)):
        """level0: uint16 (H, W) numpy array, or a contiguous uint16 torch
        tensor already on the GPU (copied device to device)."""
        if hasattr(level0, "data_ptr"):
            if not level0.is_contiguous() or level0.element_size() != 2 or level0.dim() != 2:
                raise SmxError("build_pyramid: a 2-D contiguous 16-bit tensor is required")
            h, w = level0.shape
            ptr = level0.data_ptr()
        else:
            a = np.ascontiguousarray(level0, np.uint16)
            h, w = a.shape
            ptr = _p(a)
        _check(lib().smx_job_build_pyramid(self.h, img, ptr, w, h, float(start0[0]), float(start0[1])))

    def get_view(self, img, level, shape):
        h, w = shape
        a = np.zeros((h, w), np.uint16)
        m = np.zeros((h, w), np.uint16)
        o = np.zeros((h, w), np.uint8)
        _check(lib().smx_job_get_view(self.h, img, level, _p(a), _p(m), _p(o)))
        return a, m, o

    # ------------------------------------------------------------ level
    def set_level(self, level: SmxLevel):
        _check(lib().smx_job_set_level(self.h, C.addressof(level)))
        self.ncells = int(level.grid_w) * int(level.grid_h)

    def set_ortho_wgs(self, lonlat):
        """(rows, cols, 2) lon/lat of the level's ortho grid (GNCC term)."""
        a = np.ascontiguousarray(lonlat, np.float64)
        _check(lib().smx_job_set_ortho_wgs(self.h, _p(a), a.shape[1], a.shape[0]))

    def set_searchspace(self, ugrid, minmax):
        ug = np.ascontiguousarray(ugrid, UGRID_DTYPE)
        assert ug.shape[0] == self.ncells
        mm = np.ascontiguousarray(minmax, np.float64)
        _check(lib().smx_job_set_searchspace(self.h, _p(ug), _p(mm)))

    def get_searchspace(self):
        out = np.zeros(self.ncells, UGRID_DTYPE)
        _check(lib().smx_job_get_searchspace(self.h, _p(out)))
        return out

    def get_nccresult(self):
        out = np.zeros(self.ncells, NCCRESULT_DTYPE)
        _check(lib().smx_job_get_nccresult(self.h, _p(out)))
        return out

    def set_nccresult(self, ncc):
        a = np.ascontiguousarray(ncc, NCCRESULT_DTYPE)
        _check(lib().smx_job_set_nccresult(self.h, _p(a)))

    # -------------------------------------------------- reference trio
    def initialize_voxel(self, iteration: int):
        _check(lib().smx_initialize_voxel(self.h, int(iteration)))

    def vertical_line_locus(self, iteration: int, precision: int = PREC_FP64, keep_rho: bool = False) -> int:
        return _check(lib().smx_vertical_line_locus(self.h, int(iteration), int(precision), 1 if keep_rho else 0))

    def vertical_line_locus_noalloc(self, iteration: int, precision: int = PREC_FP64) -> int:
        """smx_vertical_line_locus without device allocation (the library's
        static scratch array; no mean NCC kept)."""
        return _check(lib().smx_vertical_line_locus_noalloc(self.h, int(iteration), int(precision)))

    def awncc(self, iteration: int):
        _check(lib().smx_awncc(self.h, int(iteration)))

    # ------------------------------------------------- between iterations
    def select_mps(self, iteration: int, th_roh: float, th_roh_min: float, th_roh_start: float,
                   th_roh_next: float, mpp: float, mpp_stereo_angle: float, final_level_iteration: int):
        """SelectMPs (setsm_code.cpp:8535-8889) on the resident NCCresult and
        GridPT3: returns (matched points as POINT3_DTYPE in cell order, the
        level's minimum peak threshold). GridPT3 is updated in place."""
        a = SmxMpsArgs(th_roh, th_roh_min, th_roh_start, th_roh_next, mpp, mpp_stereo_angle, int(iteration),
                       int(final_level_iteration))
        pts = np.empty(max(self.ncells, 1), POINT3_DTYPE)
        count, th = C.c_long(0), C.c_double(0)
        _check(lib().smx_select_mps(self.h, C.byref(a), _p(pts), C.c_long(len(pts)), C.byref(count),
                                    C.byref(th)))
        return pts[: count.value], th.value

    def set_height_range(self, pts, tris, iteration: int, mpp: float, check_matching_rate: bool = False,
                         seed_dem_sigma: float = 0.0):
        """SetHeightRange (setsm_code.cpp:9851-10261): the TIN's triangles over
        the matched points set the next search space (GridPT3 is replaced).
        Returns (minH_grid, maxH_grid)."""
        pts = np.ascontiguousarray(pts, POINT3_DTYPE)
        tris = np.ascontiguousarray(tris, TRI_DTYPE)
        a = SmxHrArgs(mpp, seed_dem_sigma, int(iteration), int(bool(check_matching_rate)))
        lo, hi = C.c_double(0), C.c_double(0)
        _check(lib().smx_set_height_range(self.h, _p(pts) if len(pts) else None, C.c_long(len(pts)),
                                          _p(tris) if len(tris) else None, C.c_long(len(tris)), C.byref(a),
                                          C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def set_height_range_blunder(self, pts, tris):
        """SetHeightRange_blunder (setsm_code.cpp:10393-10491) on GridPT3."""
        pts = np.ascontiguousarray(pts, POINT3_DTYPE)
        tris = np.ascontiguousarray(tris, TRI_DTYPE)
        _check(lib().smx_set_height_range_blunder(self.h, _p(pts) if len(pts) else None, C.c_long(len(pts)),
                                                  _p(tris) if len(tris) else None, C.c_long(len(tris))))

    def set_blunder_level(self, blunder_level: int, ortho_lonlat):
        """The blunder level's ortho grid: (h, w, 2) lon/lat per ortho pixel."""
        ow = np.ascontiguousarray(ortho_lonlat, np.float64)
        assert ow.ndim == 3 and ow.shape[2] == 2
        self._bortho = ow
        _check(lib().smx_job_set_blunder_level(self.h, int(blunder_level), _p(ow), ow.shape[1], ow.shape[0]))

    def vertical_line_locus_blunder(self, iteration: int, bblunder: bool):
        """VerticalLineLocus_blunder (setsm_code.cpp:7882-8013): per-cell ortho
        NCC (float); GridPT3's ortho_ncc[1] / Mean_ortho_ncc updated."""
        out = np.zeros(max(self.ncells, 1), np.float32)
        _check(lib().smx_vertical_line_locus_blunder(self.h, int(iteration), int(bool(bblunder)), _p(out)))
        return out[: self.ncells]

    def blunder_detection_tin(self, pts, tris, iteration: int, flag_blunder: bool, hinterval: float,
                              count_bl: int = 0, ortho_ncc=None, blunder_count=(0, 0), minz=100000.0,
                              maxz=-100000.0):
        """blunder_detection_TIN (setsm_code.cpp:9158-9782): pts (POINT3
        records) get their new flags in place; GridPT3's anchor flags are
        updated. ortho_ncc None reads the last vertical_line_locus_blunder
        output kept on the device. Returns (detected, blunder_count, minz,
        maxz), the counts and heights continued from the values given."""
        assert pts.dtype == POINT3_DTYPE and pts.flags.c_contiguous
        tris = np.ascontiguousarray(tris, TRI_DTYPE)
        on = None
        if ortho_ncc is not None:
            on = np.ascontiguousarray(ortho_ncc, np.float32)
            assert len(on) == self.ncells
        a = SmxBdArgs(float(hinterval), int(iteration), int(bool(flag_blunder)), int(count_bl), 0)
        det = np.zeros(max(len(pts), 1), np.uint8)
        bc = (C.c_long * 2)(*blunder_count)
        lo, hi = C.c_double(minz), C.c_double(maxz)
        _check(lib().smx_blunder_detection_tin(self.h, C.byref(a), _p(on) if on is not None else None,
                                               _p(pts) if len(pts) else None, C.c_long(len(pts)),
                                               _p(tris) if len(tris) else None, C.c_long(len(tris)), _p(det), bc,
                                               C.byref(lo), C.byref(hi)))
        return det[: len(pts)].astype(bool), (bc[0], bc[1]), lo.value, hi.value

    # ------------------------------------------------------------- cube
    @property
    def num_hypotheses(self) -> int:
        return int(lib().smx_job_num_hypotheses(self.h))

    def get_voxels(self, with_rho=False):
        n = self.num_hypotheses
        words = np.zeros(max(n, 1), np.uint32)
        rho = np.zeros(max(n, 1), np.float32) if with_rho else None
        off = np.zeros(self.ncells + 1, np.int64)
        _check(lib().smx_job_get_voxels(self.h, _p(words), _p(rho), _p(off)))
        return words[:n], (rho[:n] if with_rho else None), off

    def get_voxel_rows(self, cell0, cell1, nwords, with_rho=False):
        words = np.zeros(max(nwords, 1), np.uint32)
        rho = np.zeros(max(nwords, 1), np.float32) if with_rho else None
        _check(lib().smx_job_get_voxel_rows(self.h, int(cell0), int(cell1), _p(words), _p(rho)))
        return words[:nwords], (rho[:nwords] if with_rho else None)

    def set_voxels(self, words):
        w = np.ascontiguousarray(words, np.uint32)
        _check(lib().smx_job_set_voxels(self.h, _p(w)))

    def get_sumcost(self):
        """SumCost (float32, CSR order) of the last SGM-level AWNCC."""
        out = np.zeros(max(self.num_hypotheses, 1), np.float32)
        _check(lib().smx_job_get_sumcost(self.h, _p(out)))
        return out[:self.num_hypotheses]

    @property
    def last_kernel_ms(self) -> float:
        return float(lib().smx_job_last_kernel_ms(self.h))

    @property
    def last_evaluated(self) -> int:
        return int(lib().smx_job_last_evaluated(self.h))

    @property
    def device_bytes(self) -> int:
        return int(lib().smx_job_device_bytes(self.h))

    @property
    def last_stats(self) -> dict:
        out = np.zeros(4, np.int64)
        _check(lib().smx_job_last_stats(self.h, _p(out)))
        runs, hyps, ms = C.c_longlong(), C.c_longlong(), C.c_double()
        _check(lib().smx_job_last_redo(self.h, C.byref(runs), C.byref(hyps), C.byref(ms)))
        return dict(evaluated=int(out[0]), tiles_lds=int(out[1]), waves_hbm=int(out[2]), fast_config=int(out[3]),
                    redo_runs=runs.value, redo_hyps=hyps.value, redo_ms=ms.value)

    @property
    def half(self) -> int:
        """Half_template_size of the current level (the window size batches group by)."""
        return _check(lib().smx_job_half(self.h))

    def set_fast_config(self, config: int):
        _check(lib().smx_job_set_fast_config(self.h, int(config)))

    def sync(self):
        _check(lib().smx_job_sync(self.h))


def release_cached() -> int:
    """Return the library's cached device blocks to the runtime (bytes)."""
    return int(lib().smx_release_cached())


def fast_configs():
    """[(index, (tx, ty, lds_texels, occ, stage, wide, run_len))] of the
    SMX_PREC_FAST launch shapes: cells per tile, LDS texels of both boxes,
    register cap (waves per SIMD, 0 = compiler's), LDS staging, FP64 window
    totals, heights per run (64 / run_len runs per wavefront)."""
    out = []
    buf = C.create_string_buffer(64)
    for i in range(lib().smx_fast_num_configs()):
        _check(lib().smx_fast_config_desc(i, buf, 64))
        out.append((i, tuple(int(x) for x in buf.value.split())))
    return out


def vertical_line_locus_batch(jobs, iteration: int, precision: int = PREC_FAST, keep_rho: bool = False):
    """One batched VerticalLineLocus over `jobs`; returns Accessable_grid per job."""
    arr = (C.c_void_p * len(jobs))(*[j.h.value for j in jobs])
    acc = np.zeros(len(jobs), np.int64)
    _check(lib().smx_vertical_line_locus_batch(arr, len(jobs), int(iteration), int(precision),
                                               1 if keep_rho else 0, _p(acc)))
    return acc


class Graph:
    """One iteration's VerticalLineLocus (+ AWNCC) over `jobs`, captured as a
    hipGraph (smx_graph_create) and replayed by launch()."""

    def __init__(self, jobs, iteration: int, precision: int = PREC_FAST, with_awncc: bool = True):
        self.jobs = list(jobs)
        arr = (C.c_void_p * len(self.jobs))(*[j.h.value for j in self.jobs])
        h = C.c_void_p()
        _check(lib().smx_graph_create(C.byref(h), arr, len(self.jobs), int(iteration), int(precision),
                                      1 if with_awncc else 0))
        self.h = h

    def launch(self):
        acc = np.zeros(len(self.jobs), np.int64)
        _check(lib().smx_graph_launch(self.h, _p(acc)))
        return acc

    @property
    def last_ms(self) -> float:
        return float(lib().smx_graph_last_ms(self.h))

    def close(self):
        if self.h:
            lib().smx_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
