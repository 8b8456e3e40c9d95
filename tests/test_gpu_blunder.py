"""DecisionMPs' GPU helpers against the oracle, bit for bit: one pass of the
blunder loop's data-parallel part (setsm_code.cpp:8958-8963) after a real
AWNCC: SelectMPs -> the reference's TIN -> SetHeightRange_blunder (TIN
heights into GridPT3) -> VerticalLineLocus_blunder (ortho NCC of every cell
on the blunder level's images)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from setsm_amd.records import UGRID_DTYPE
from setsm_amd.synth import ortho_wgs
from tests.helpers import GpuRun, level_grid, scene, upload_pyramid
from tests.test_gpu_mps import SEL, grids_equal

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(O.tin() is None, reason="oracle/_ref/libref_tin.so not built")]


@pytest.mark.parametrize("level,B,bblunder,it,cells", [
    (4, 4, 1, 1, 40),     # the blunder loop's window at level 4: half 15 - 2 = 13
    (3, 3, 1, 1, 80),     # level 3: half 15 - 2 - 2 = 11
    (2, 3, 1, 1, 80),
    (1, 2, 0, 1, 80),
    (0, 1, 0, 1, 80)])
def test_blunder_pass_matches_oracle(require_gpu, level, B, bblunder, it, cells):
    from setsm_amd import smx
    sc, pyr = scene(image_px=1024, relief=90.0, seed=17, levels=5)
    g = level_grid(sc, level, cells, band=min(48.0 * (1 << level), 400.0), seed=3)
    job = smx.Job(5)
    try:
        upload_pyramid(job, pyr)
        gr = GpuRun(job, level, pyr, sc.rpcs, g, sgm_py=-1)
        li = gr.li
        gr.step(g["ugrid"], it, g["minmax"], precision=smx.PREC_FP64)
        ncc = job.get_nccresult()
        ug_o = job.get_searchspace()
        pts, _ = job.select_mps(it, mpp=1.0, **SEL)
        O.select_mps(li, ncc, ug_o, it, mpp=1.0, **SEL)
        tris = O.tin_create(pts, g["boundary"], g["grid_res"])
        job.set_height_range_blunder(pts, tris)
        O.set_height_range_blunder(li, ug_o, pts, tris)
        grids_equal(job.get_searchspace(), ug_o)
        ow = ortho_wgs(sc, g["boundary"], B)
        job.set_blunder_level(B, ow)
        out_g = job.vertical_line_locus_blunder(it, bblunder)
        out_o = O.vll_blunder(li, pyr, B, ow, ug_o, it, bblunder)
        assert np.array_equal(out_g.view(np.uint32), out_o.view(np.uint32)), \
            np.nonzero(out_g.view(np.uint32) != out_o.view(np.uint32))[0][:5]
        grids_equal(job.get_searchspace(), ug_o)
        print(f"level {level} blunder level {B}: {len(pts)} points, {len(tris)} triangles, half "
              f"{O.blunder_half_template(li, B, it, bblunder)}, mean ortho NCC {out_o.mean():.3f}", flush=True)
        assert out_o.mean() > 0.3 and len(tris) > 0
    finally:
        job.close()


def test_blunder_errors(require_gpu):
    from setsm_amd import smx
    sc, pyr = scene(image_px=1024, relief=90.0, seed=17, levels=5)
    g = level_grid(sc, 2, 40, seed=3)
    job = smx.Job(5)
    try:
        upload_pyramid(job, pyr)
        gr = GpuRun(job, 2, pyr, sc.rpcs, g)
        job.set_searchspace(g["ugrid"], g["minmax"])
        with pytest.raises(smx.SmxError):          # no blunder level yet
            job.vertical_line_locus_blunder(1, True)
        ow = ortho_wgs(sc, g["boundary"], 3)
        with pytest.raises(smx.SmxError):          # the grid of another level
            job.set_blunder_level(2, ow)
        with pytest.raises(smx.SmxError):          # a level the job does not hold
            job.set_blunder_level(7, ow)
        job.set_blunder_level(3, ow)
        # no triangles: heights unchanged; the NCC still runs
        before = job.get_searchspace()
        job.set_height_range_blunder(np.zeros(0, O.POINT3_DTYPE), np.zeros(0, O.TRI_DTYPE))
        grids_equal(job.get_searchspace(), before)
        ug_o = before.copy()
        out_o = O.vll_blunder(gr.li, pyr, 3, ow, ug_o, 1, True)
        out_g = job.vertical_line_locus_blunder(1, True)
        assert np.array_equal(out_g.view(np.uint32), out_o.view(np.uint32))
    finally:
        job.close()
