"""The caller's side of SETSM's per-iteration loop, over the GPU steps.

Between two iterations of a level, Matching_SETSM (setsm_code.cpp:3860-4222)
runs SelectMPs, DecisionMPs' blunder loop (for the anchor points and then
for all points) and SetHeightRange over a TIN of the points that survive.
Every data-parallel step of that chain runs on the GPU through `smx.Job`.
This module holds the control flow around them: DecisionMPs' loop
(setsm_code.cpp:8891-9093), restated step for step, and the chain. The
TIN is the caller's. SETSM uses its own grid triangulation. A caller without
one can use `native_tin` (libsmx's `smx_tin_create`, a radial-sweep
Delaunay on the lattice with exact integer predicates) or `qhull_tin`
(scipy's Qhull, slower on lattices), both listed clockwise as SETSM's TIN
lists its triangles. Among cocircular lattice points each splits quads its
own way, so their triangles are not the reference's, but they are Delaunay
TINs of the same points.

The loop takes its steps from an `ops` object, so the same control flow
drives the GPU (`JobOps`) and, in the tests, the oracle.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .records import POINT3_DTYPE, TRI_DTYPE


def qhull_tin(pts, boundary, grid_res):
    """Delaunay TIN of the matched points (TINCreate_list's role,
    SubFunctions.cpp:2768-2846): the points snapped to the level's lattice
    as the reference snaps them (0.5 + (x - Boundary[0]) / resolution,
    truncated), triangulated by Qhull, each triangle clockwise. Fewer than
    3 points, or all on one line: no triangles."""
    n = len(pts)
    if n <= 2:
        return np.zeros(0, TRI_DTYPE)
    col = np.floor(0.5 + (pts["x"] - boundary[0]) / grid_res)
    row = np.floor(0.5 + (pts["y"] - boundary[1]) / grid_res)
    xy = np.stack([col, row], 1)
    from scipy.spatial import Delaunay
    from scipy.spatial import QhullError
    try:
        tri = Delaunay(xy).simplices.astype(np.int64)
    except QhullError:
        return np.zeros(0, TRI_DTYPE)
    a, b, c = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    tri = tri[cross != 0]
    ccw = cross[cross != 0] > 0
    tri[ccw] = tri[ccw][:, [0, 2, 1]]
    out = np.zeros(len(tri), TRI_DTYPE)
    out["a"], out["b"], out["c"] = tri[:, 0], tri[:, 1], tri[:, 2]
    return out


def native_tin(pts, boundary, grid_res):
    """Delaunay TIN of the matched points by libsmx's `smx_tin_create`
    (TINCreate_list's role): boundary = (min_x, min_y, max_x, max_y) of the
    level's lattice."""
    from . import smx
    return smx.tin_create(pts, boundary, grid_res)


class JobOps:
    """DecisionMPs' three steps on a job's resident GridPT3, on the GPU; the
    ortho NCC between them stays on the device."""

    def __init__(self, job):
        self.job = job

    def set_height_range_blunder(self, pts, tris):
        self.job.set_height_range_blunder(pts, tris)

    def vll_blunder(self, iteration, bblunder):
        self.job.vertical_line_locus_blunder(iteration, bblunder)

    def blunder_detection(self, pts, tris, iteration, flag_blunder, hinterval, count_bl, zero_ortho, minz, maxz):
        on = np.zeros(self.job.ncells, np.float32) if zero_ortho else None
        det, bc, lo, hi = self.job.blunder_detection_tin(pts, tris, iteration, flag_blunder, hinterval,
                                                          count_bl=count_bl, ortho_ncc=on, minz=minz, maxz=maxz)
        return det, bc, lo, hi


@dataclass
class DecisionResult:
    count_results: tuple = (0, 0)    # (points, triangles) of the last TIN
    minz: float = 100000.0
    maxz: float = -100000.0
    passes: int = 0
    blunders: int = 0
    history: list = field(default_factory=list)


def decision_mps(ops, pts, level, iteration, hinterval, flag_blunder, boundary, grid_res, is_ra=False,
                 tin=native_tin):
    """DecisionMPs (setsm_code.cpp:8891-9093) on the points pts
    (POINT3_DTYPE, flags updated in place). Each pass: SetHeightRange_blunder
    and VerticalLineLocus_blunder (not in RA mode, whose ortho NCC stays
    zero), blunder_detection_TIN, then a new TIN over the points still in
    play when new blunders were found. Stops when a pass finds fewer than one
    new blunder or no kept point, when the counts repeat, after 10 / 20 / 30
    passes (anchor points; blunders at levels 2-0 / 3 / 4), or when the TIN
    falls to 20 triangles. Ends with SetHeightRange_blunder and a
    VerticalLineLocus_blunder without the loop's window shrink."""
    res = DecisionResult()
    if level == 0 and iteration == 3:
        return res
    th_pts = 1
    if not flag_blunder:
        max_count = 10
    elif level >= 4:
        max_count = 30
    elif level >= 3:
        max_count = 20
    else:
        max_count = 10
    tris = tin(pts, boundary, grid_res)
    res.count_results = (len(pts), len(tris))
    flag = True
    count = 0
    pre_blunder = pre_mps = -100
    while flag and count < max_count and len(tris) > 20:
        if not is_ra:
            ops.set_height_range_blunder(pts, tris)
            ops.vll_blunder(iteration, True)
        det, bc, res.minz, res.maxz = ops.blunder_detection(pts, tris, iteration, flag_blunder, hinterval,
                                                            count, bool(is_ra), res.minz, res.maxz)
        count_blunders = abs(int(bc[1]) - pre_blunder) if count > 0 else int(bc[1])
        if count_blunders < th_pts:
            flag = False
        if bc[0] < th_pts:
            flag = False
        if pre_blunder == bc[1] and pre_mps == bc[0]:
            flag = False
        else:
            pre_mps, pre_blunder = int(bc[0]), int(bc[1])
        new = int(np.count_nonzero(det))
        res.blunders += new
        if new > 0:
            # setsm_code.cpp:8991-9041: the next TIN over the points still
            # in play (anchors, flag 2, leave it while the loop goes on)
            f = pts["flag"]
            keep = np.nonzero((f != 1) & (f != 2) if flag else (f != 1))[0]
            sub = tin(pts[keep], boundary, grid_res)
            tris = np.zeros(len(sub), TRI_DTYPE)
            for k in ("a", "b", "c"):
                tris[k] = keep[sub[k]]
        res.count_results = (len(pts), len(tris))
        res.history.append(dict(kept=int(bc[0]), blunders=int(bc[1]), new=new, triangles=len(tris)))
        count += 1
        if count >= max_count or len(tris) <= 20:
            flag = False
    res.passes = count
    ops.set_height_range_blunder(pts, tris)
    ops.vll_blunder(iteration, False)
    return res


@dataclass
class ChainResult:
    points: int
    minimum_th: float
    anchors: DecisionResult | None
    blunders: DecisionResult | None
    kept: int
    triangles: int
    min_max_height: tuple
    tin_s: float = 0.0      # host time in the TIN calls
    tins: int = 0


def iteration_chain(job, level, iteration, boundary, grid_res, mps_args, hinterval, mpp, dem_resolution=2.0,
                    is_ra=False, check_matching_rate=False, tin=native_tin, ops=None):
    """SelectMPs -> DecisionMPs (anchor points, then all points) ->
    SetHeightRange over the TIN of the points that are not blunders
    (setsm_code.cpp:3862-3945, 4192-4222), on the job's level after its
    AWNCC. The job's blunder level must be set (smx_job_set_blunder_level)
    unless is_ra. The job's GridPT3 becomes the next iteration's search
    space."""
    ops = ops or JobOps(job)
    spent = [0.0, 0]

    def timed_tin(p, b, r):
        t0 = time.perf_counter()
        out = tin(p, b, r)
        spent[0] += time.perf_counter() - t0
        spent[1] += 1
        return out

    pts, th = job.select_mps(iteration, mpp=mpp, **mps_args)
    pts = np.ascontiguousarray(pts.copy(), POINT3_DTYPE)
    anchors = blunders = None
    n = len(pts)
    if n > 10:
        ortho_level = 3 if dem_resolution >= 8 else 2
        check_ortho = level >= ortho_level and not (level == 2 and iteration > 5)
        if check_ortho and not is_ra:
            anchors = decision_mps(ops, pts, level, iteration, hinterval, False, boundary, grid_res, is_ra,
                                   timed_tin)
            pts["flag"] = 0
        blunders = decision_mps(ops, pts, level, iteration, hinterval, True, boundary, grid_res, is_ra, timed_tin)
        kept = pts[pts["flag"] != 1]
    else:
        kept = pts
    tris = timed_tin(kept, boundary, grid_res)
    lohi = job.set_height_range(kept, tris, iteration, mpp, check_matching_rate)
    return ChainResult(points=n, minimum_th=th, anchors=anchors, blunders=blunders, kept=len(kept),
                       triangles=len(tris), min_max_height=lohi, tin_s=spent[0], tins=spent[1])
