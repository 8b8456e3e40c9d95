"""SelectMPs and SetHeightRange on the GPU against the oracle, bit for bit.

Each level runs SETSM's chain between two iterations: AWNCC's peaks ->
SelectMPs (matched points, grid update) -> the TIN over the points (the
reference's own grid_triangulation.cpp, built into oracle/_ref) ->
SetHeightRange (the next search space), then a second iteration from that
search space so the matched-cell branches run too. Every grid field, every
point and the range bounds must equal the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_py as O
from setsm_amd.records import LevelInputs, NCCRESULT_DTYPE, POINT3_DTYPE, UGRID_DTYPE
from tests.helpers import GpuRun, level_grid, scene, upload_pyramid

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(O.tin() is None, reason="oracle/_ref/libref_tin.so not built")]

SEL = dict(th_roh=0.6, th_roh_min=0.35, th_roh_start=0.6, th_roh_next=0.5, mpp_stereo_angle=10.0,
           final_level_iteration=1)


def grids_equal(a, b):
    for f in ("Height", "minHeight", "maxHeight", "roh", "ortho_ncc", "Mean_ortho_ncc", "Matched_flag",
              "anchor_flag"):
        x, y = a[f], b[f]
        if f == "Height":
            x, y = x.view(np.uint32), y.view(np.uint32)
        if not np.array_equal(x, y):
            bad = np.nonzero(x != y)[0]
            raise AssertionError(f"{f} differs in {len(bad)} cells, first {bad[:5]}: {x[bad[:3]]} vs {y[bad[:3]]}")


def points_equal(a, b):
    assert len(a) == len(b), (len(a), len(b))
    for f in ("x", "y", "z", "flag"):
        assert np.array_equal(a[f], b[f]), f


def chain(job, li, g, it, mpp):
    """SelectMPs -> TIN -> SetHeightRange on the GPU and on the oracle from
    the job's current NCCresult and GridPT3; returns the new grid."""
    ncc = job.get_nccresult()
    ug_o = job.get_searchspace()
    pts_g, th_g = job.select_mps(it, mpp=mpp, **SEL)
    pts_o, th_o = O.select_mps(li, ncc, ug_o, it, mpp=mpp, **SEL)
    assert th_g == th_o
    points_equal(pts_g, pts_o)
    grids_equal(job.get_searchspace(), ug_o)
    tris = O.tin_create(pts_o, g["boundary"], g["grid_res"])
    lo_g, hi_g = job.set_height_range(pts_g, tris, it, mpp)
    lo_o, hi_o = O.set_height_range(li, ug_o, pts_o, tris, it, mpp)
    assert (lo_g, hi_g) == (lo_o, hi_o)
    out = job.get_searchspace()
    grids_equal(out, ug_o)
    return out, len(pts_o), len(tris)


@pytest.mark.parametrize("level,sgm_py,mpp", [(3, -1, 1.5), (2, -1, 1.0), (1, 1, 0.8), (0, 1, 0.4)])
def test_select_and_set_height_range_match_oracle(require_gpu, level, sgm_py, mpp):
    from setsm_amd import smx
    sc, pyr = scene(image_px=1024, relief=90.0, seed=17, levels=4)
    g = level_grid(sc, level, 96, band=48.0 * (1 << level), seed=3)
    job = smx.Job(4)
    try:
        upload_pyramid(job, pyr)
        gr = GpuRun(job, level, pyr, sc.rpcs, g, sgm_py=sgm_py)
        li = gr.li
        ug = g["ugrid"]
        for it in (1, 2):
            gr.step(ug, it, g["minmax"], precision=smx.PREC_FP64)
            ug, npts, ntri = chain(job, li, g, it, mpp)
            print(f"level {level} iteration {it}: {npts} points, {ntri} triangles, "
                  f"{int((ug['Matched_flag'] != 0).sum())} cells matched", flush=True)
            assert npts > 10 and ntri > 10
    finally:
        job.close()


def synthetic(n, rng, level):
    ncc = np.zeros(n, NCCRESULT_DTYPE)
    ncc["NumOfHeight"] = rng.integers(0, 20, n)
    r0 = rng.integers(-200, 1000, n)
    ncc["result0"] = r0
    ncc["result1"] = r0 - rng.integers(0, 300, n)
    ncc["result2"] = (50 + rng.normal(0, 8, n)).astype(np.float32)
    ncc["result3"] = (50 + rng.normal(0, 8, n)).astype(np.float32)
    ncc["result4"] = rng.integers(0, 3, n)
    ug = np.zeros(n, UGRID_DTYPE)
    ug["Height"] = (50 + rng.normal(0, 5, n)).astype(np.float32)
    ug["minHeight"] = 30 + rng.integers(-5, 5, n)
    ug["maxHeight"] = 70 + rng.integers(-5, 5, n)
    ug["roh"] = rng.integers(0, 700, n)
    ug["Matched_flag"] = rng.choice(np.array([0, 1, 2, 4], np.uint8), n)
    ug["Mean_ortho_ncc"] = rng.integers(0, 900, n)
    ug["ortho_ncc"][:, 1] = rng.integers(0, 900, n)
    return ncc, ug


@pytest.mark.parametrize("level", [2, 0])
def test_million_cell_grid_matches_oracle(require_gpu, level):
    """1000 x 1000 cells of synthetic AWNCC output: the per-cell branches,
    the histogram threshold and the triangle claims at scale."""
    from setsm_amd import smx
    sc, pyr = scene(image_px=512, relief=60.0, seed=3, levels=3)
    W = 1000
    gres = 2.0 * (1 << level)
    c, r = np.meshgrid(np.arange(W), np.arange(W))
    g = dict(grid_pts=np.stack([100 + c.ravel() * gres, 200 + r.ravel() * gres], 1),
             grid_wgs=np.zeros((W * W, 2)), boundary=np.array([100, 200, 100 + (W - 1) * gres, 200 + (W - 1) * gres]),
             shape=(W, W), grid_res=gres)
    li = LevelInputs(pyr, level, sc.rpcs, g, grid_resolution=gres, dem_resolution=1.0, sgm_py=-1)
    rng = np.random.default_rng(40 + level)
    ncc, ug = synthetic(W * W, rng, level)
    job = smx.Job(3)
    try:
        upload_pyramid(job, pyr)
        job.set_level(li.level)
        job.set_searchspace(ug, np.array([0.0, 100.0]))
        job.set_nccresult(ncc)
        out, npts, ntri = chain(job, li, g, 2, 0.5)
        print(f"level {level}: {npts} points, {ntri} triangles", flush=True)
        assert npts > 1000
    finally:
        job.close()


@pytest.mark.parametrize("level,kw,hr", [
    (2, dict(is_ra=1), {}),                                   # RA: fixed 0.2 threshold
    (1, dict(sgm_py=1), {}),                                  # SGM level: fraction 0.85
    (3, dict(pre_dem=1), dict(seed_dem_sigma=6.0)),           # seed DEM caps the buffer
    (1, dict(dem_resolution=8.0), {}),                        # DEM >= 4 m: gap limit 1000
    (2, dict(check_matching_rate=1), dict(check_matching_rate=True)),
    (0, dict(dem_resolution=1.0), {}),                        # level 0, DEM < 2 m: gap limit 50
])
def test_branches_match_oracle(require_gpu, level, kw, hr):
    """The level-record and argument branches of both functions on a
    300 x 300 synthetic grid, two iterations apart (it 1 and it 5)."""
    from setsm_amd import smx
    sc, pyr = scene(image_px=512, relief=60.0, seed=3, levels=4)
    W = 300
    gres = 2.0 * (1 << level)
    c, r = np.meshgrid(np.arange(W), np.arange(W))
    g = dict(grid_pts=np.stack([50 + c.ravel() * gres, 80 + r.ravel() * gres], 1), grid_wgs=np.zeros((W * W, 2)),
             boundary=np.array([50, 80, 50 + (W - 1) * gres, 80 + (W - 1) * gres]), shape=(W, W), grid_res=gres)
    args = dict(dem_resolution=2.0)
    args.update(kw)
    li = LevelInputs(pyr, level, sc.rpcs, g, grid_resolution=gres, **args)
    for it, seed in ((1, 7), (5, 8)):
        ncc, ug = synthetic(W * W, np.random.default_rng(seed + 10 * level), level)
        job = smx.Job(4)
        try:
            upload_pyramid(job, pyr)
            job.set_level(li.level)
            job.set_searchspace(ug, np.array([0.0, 100.0]))
            job.set_nccresult(ncc)
            ug_o = ug.copy()
            pts_g, th_g = job.select_mps(it, mpp=0.7, **SEL)
            pts_o, th_o = O.select_mps(li, ncc, ug_o, it, mpp=0.7, **SEL)
            assert th_g == th_o
            points_equal(pts_g, pts_o)
            grids_equal(job.get_searchspace(), ug_o)
            tris = O.tin_create(pts_o, g["boundary"], gres)
            assert job.set_height_range(pts_g, tris, it, 0.7, **hr) == \
                O.set_height_range(li, ug_o, pts_o, tris, it, 0.7, **hr)
            grids_equal(job.get_searchspace(), ug_o)
        finally:
            job.close()


def test_edge_cases(require_gpu):
    """Triangles indexing past the points are skipped, no triangles at all,
    a point buffer too small, and a level with no threshold."""
    from setsm_amd import smx
    sc, pyr = scene(image_px=512, relief=60.0, seed=3, levels=3)
    W, gres = 40, 4.0
    c, r = np.meshgrid(np.arange(W), np.arange(W))
    g = dict(grid_pts=np.stack([c.ravel() * gres, r.ravel() * gres], 1), grid_wgs=np.zeros((W * W, 2)),
             boundary=np.array([0, 0, (W - 1) * gres, (W - 1) * gres]), shape=(W, W), grid_res=gres)
    li = LevelInputs(pyr, 2, sc.rpcs, g, grid_resolution=gres)
    ncc, ug = synthetic(W * W, np.random.default_rng(9), 2)
    job = smx.Job(3)
    try:
        upload_pyramid(job, pyr)
        job.set_level(li.level)
        job.set_searchspace(ug, np.array([0.0, 100.0]))
        job.set_nccresult(ncc)
        pts, _ = job.select_mps(2, mpp=1.0, **SEL)
        ug_o = ug.copy()
        O.select_mps(li, ncc, ug_o, 2, mpp=1.0, **SEL)
        tris = O.tin_create(pts, g["boundary"], gres)
        bad = tris.copy()[:5]
        bad["c"] = len(pts) + 3
        tris2 = np.concatenate([bad, tris])
        base = job.get_searchspace()
        for t in (tris2, tris[:0]):
            job.set_searchspace(base, np.array([0.0, 100.0]))
            ref = base.copy()
            assert job.set_height_range(pts, t, 2, 1.0) == O.set_height_range(li, ref, pts, t, 2, 1.0)
            grids_equal(job.get_searchspace(), ref)
        # a point buffer smaller than the count: SMX_E_ARG, the count reported
        job.set_searchspace(ug, np.array([0.0, 100.0]))
        a = smx.SmxMpsArgs(0.6, 0.35, 0.6, 0.5, 1.0, 10.0, 2, 1)
        small = np.zeros(3, POINT3_DTYPE)
        cnt, th = C.c_long(0), C.c_double(0)
        rc = smx.lib().smx_select_mps(job.h, C.byref(a), small.ctypes.data, 3, C.byref(cnt), C.byref(th))
        assert rc == -1 and cnt.value == len(pts)
    finally:
        job.close()
    # level 5 above SGM_py: no threshold -> SMX_E_ARG and the grid unchanged
    sc5, pyr5 = scene(image_px=1024, relief=60.0, seed=3, levels=6)
    gres5 = 64.0
    g5 = dict(grid_pts=np.stack([c.ravel() * gres5, r.ravel() * gres5], 1), grid_wgs=np.zeros((W * W, 2)),
              boundary=np.array([0, 0, (W - 1) * gres5, (W - 1) * gres5]), shape=(W, W), grid_res=gres5)
    li5 = LevelInputs(pyr5, 5, sc5.rpcs, g5, grid_resolution=gres5)
    job = smx.Job(6)
    try:
        upload_pyramid(job, pyr5)
        job.set_level(li5.level)
        job.set_searchspace(ug, np.array([0.0, 100.0]))
        job.set_nccresult(ncc)
        with pytest.raises(smx.SmxError):
            job.select_mps(1, mpp=1.0, **SEL)
        grids_equal(job.get_searchspace(), ug)
    finally:
        job.close()
