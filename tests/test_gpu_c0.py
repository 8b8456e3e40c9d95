"""The reference's own CPU-runnable case at its specified size: a 2048^2
pair at 0.5 m over 1 km of fBm terrain (300 m relief, views +-12 deg), the
five matching levels L4..L0 with SETSM's grid spacings 8/8/6/3/2 m
(125^2/125^2/167^2/334^2/500^2 cells, setsm_code.cpp:5074-5198) and height
steps 3/2/1/0.5/0.25 m, 64 heights per cell: the FP64 path bit for bit
against the oracle at every level (cube, NCCresult after VerticalLineLocus
and after AWNCC, search space), the FAST path within the contract."""
import math

import numpy as np
import pytest

from setsm_amd.records import height_step
from setsm_amd.synth import GridSpec, make_grid
from tests.helpers import GpuRun, OracleRun, scene, upload_pyramid
from tests.test_gpu_parity import _compare, _fast_vs_fp64, _job

pytestmark = pytest.mark.gpu

SPACING = {4: 8.0, 3: 8.0, 2: 6.0, 1: 3.0, 0: 2.0}


def _grid(sc, level, nheights=64):
    side = sc.spec.image_px * sc.spec.gsd
    res = SPACING[level]
    cells = int(math.ceil(1000.0 / res))
    inset = (side - cells * res) / 2.0
    band = float(max(1, int(round(nheights * height_step(level, sc.spec.gsd)))))
    g = make_grid(sc, GridSpec(cells=cells, grid_res=res, band=band, inset=inset, x_offset=0.0, y_offset=0.0,
                               seed=40 + level))
    g["boundary"] = np.array([inset, inset, inset + cells * res, inset + cells * res])
    g["shape"] = (cells, cells)
    g["grid_res"] = res
    return g


@pytest.fixture(scope="module")
def c0():
    return scene(image_px=2048, relief=300.0, seed=21, view=12.0, levels=5)


@pytest.mark.parametrize("level", [4, 3, 2, 1, 0])
def test_c0_level(require_gpu, c0, level):
    from setsm_amd.smx import PREC_FAST
    sc, pyr = c0
    g = _grid(sc, level)
    print(f"[C0 L{level}] {g['shape'][0]}^2 cells at {SPACING[level]} m", flush=True)
    ro = OracleRun(pyr, sc.rpcs, level, g).step(g["ugrid"], 1, g["minmax"])
    job = _job(5)
    upload_pyramid(job, pyr)
    rg = GpuRun(job, level, pyr, sc.rpcs, g).step(g["ugrid"], 1, g["minmax"])
    _compare(ro, rg, f"C0 L{level} FP64")
    job2 = _job(5)
    upload_pyramid(job2, pyr)
    rf = GpuRun(job2, level, pyr, sc.rpcs, g).step(g["ugrid"], 1, g["minmax"], precision=PREC_FAST)
    _fast_vs_fp64(ro, rf, len(g["ugrid"]), f"C0 L{level} FAST")
    job.close()
    job2.close()
