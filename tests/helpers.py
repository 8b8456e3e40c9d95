"""Scenes and drivers shared by the tests: the same synthetic inputs go to the
oracle (CPU FP64 restatement) and to libsmx (HIP), level by level and
iteration by iteration, the way Matching_SETSM drives the reference
(setsm_code.cpp:3774-3844)."""
from __future__ import annotations

import functools

import numpy as np

from oracle import oracle_py as O
from setsm_amd.records import LevelInputs, height_step
from setsm_amd.synth import GridSpec, SceneSpec, make_grid, make_scene


@functools.lru_cache(maxsize=8)
def scene(image_px=512, relief=60.0, seed=3, view=12.0, levels=1):
    sc = make_scene(SceneSpec(image_px=image_px, relief=relief, seed=seed, view_deg=(view, -view)))
    pyr = [O.build_pyramid(im, levels) for im in sc.images]
    return sc, pyr


def level_grid(sc, level, cells, band=None, seed=7, inset=None, res=0.5):
    """Cell grid for one pyramid level; the band is a whole number of metres
    holding `band / height_step` hypotheses around the true height."""
    grid_res = max(2.0, res * (1 << level))
    hs = height_step(level, res)
    if band is None:
        band = min(128 * hs, 256.0)
    band = float(max(1, int(round(band))))
    if inset is None:
        side = sc.spec.image_px * sc.spec.gsd
        inset = (side - cells * grid_res) / 2.0
    g = make_grid(sc, GridSpec(cells=cells, grid_res=grid_res, band=band, inset=inset, x_offset=0.0,
                               y_offset=0.0, seed=seed + level))
    g["boundary"] = np.array([inset, inset, inset + cells * grid_res, inset + cells * grid_res])
    g["shape"] = (cells, cells)
    g["grid_res"] = grid_res
    return g


def perturb_searchspace(ug, seed, level):
    """A stand-in for SetHeightRange between iterations: shift/narrow/widen
    the bands of some cells, leave others unchanged, zero a few flags."""
    rng = np.random.default_rng(seed)
    out = ug.copy()
    n = out.shape[0]
    pick = rng.random(n)
    lo = out["minHeight"].astype(np.int64)
    hi = out["maxHeight"].astype(np.int64)
    widen = pick < 0.25
    lo[widen] -= rng.integers(1, 5, widen.sum())
    hi[widen] += rng.integers(1, 5, widen.sum())
    narrow = (pick >= 0.25) & (pick < 0.45)
    lo[narrow] += 2
    hi[narrow] -= 2
    out["minHeight"] = lo
    out["maxHeight"] = hi
    if level <= 1:
        off = rng.random(n) < 0.03
        out["Matched_flag"][off] = 0
    return out


class OracleRun:
    def __init__(self, pyr, rpcs, level, grid, **kw):
        self.li = LevelInputs(pyr, level, rpcs, grid, grid_resolution=grid["grid_res"], **kw)
        self.st = O.OracleState(self.li)

    def step(self, ug, iteration, minmax):
        ug = ug.copy()
        self.st.initialize_voxel(ug, iteration, minmax)
        acc = self.st.vertical_line_locus(ug, iteration, minmax)
        ncc_after_vll = self.st.ncc.copy()
        words, rho = self.st.export()
        self.st.awncc(ug, iteration)
        return dict(ug=ug, acc=acc, ncc_vll=ncc_after_vll, words=words, rho=rho, ncc=self.st.ncc.copy())


class GpuRun:
    def __init__(self, job, level, pyr, rpcs, grid, **kw):
        from setsm_amd.smx import Job  # noqa: F401
        self.job = job
        self.li = LevelInputs(pyr, level, rpcs, grid, grid_resolution=grid["grid_res"], **kw)
        job.set_level(self.li.level)

    def step(self, ug, iteration, minmax, precision=0):
        j = self.job
        j.set_searchspace(ug, minmax)
        j.initialize_voxel(iteration)
        acc = j.vertical_line_locus(iteration, precision=precision, keep_rho=True)
        ncc_after_vll = j.get_nccresult()
        words, rho, off = j.get_voxels(with_rho=True)
        j.awncc(iteration)
        return dict(ug=j.get_searchspace(), acc=acc, ncc_vll=ncc_after_vll, words=words, rho=rho,
                    ncc=j.get_nccresult(), off=off, evaluated=j.last_evaluated, ms=j.last_kernel_ms)


def upload_pyramid(job, pyr):
    for ti in range(2):
        for lv, (img, mag, ori) in enumerate(pyr[ti]):
            job.set_view(ti, lv, img, mag, ori, start=(0.0, 0.0))
