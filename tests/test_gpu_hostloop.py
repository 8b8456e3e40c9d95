"""The per-iteration chain of SETSM's loop (setsm_code.cpp:3862-3945,
4192-4222) over the GPU steps against the same chain over the oracle, bit
for bit: SelectMPs, DecisionMPs for the anchor points and for all points
(setsm_code.cpp:8891-9093, each pass SetHeightRange_blunder,
VerticalLineLocus_blunder and blunder_detection_TIN, then a new TIN), and
SetHeightRange over the TIN of the points that are not blunders. The loop's
control flow is `setsm_amd.host_loop` for both; the TIN is Qhull's
(`host_loop.qhull_tin`) for both. Then a second iteration runs on the
chain's search space."""
import numpy as np
import pytest

from oracle import oracle_py as O
from setsm_amd import host_loop
from setsm_amd.synth import ortho_wgs
from tests.helpers import GpuRun, level_grid, scene, upload_pyramid
from tests.test_gpu_mps import SEL, grids_equal

pytestmark = pytest.mark.gpu


def dec_equal(a, b):
    if a is None or b is None:
        assert a is None and b is None
        return
    assert a.count_results == b.count_results and a.passes == b.passes and a.blunders == b.blunders
    assert a.history == b.history
    assert np.float64(a.minz).tobytes() == np.float64(b.minz).tobytes()
    assert np.float64(a.maxz).tobytes() == np.float64(b.maxz).tobytes()


@pytest.mark.parametrize("level,its,cells", [(4, 1, 40), (3, 1, 80), (2, 1, 80), (1, 1, 80), (0, 2, 80)])
def test_iteration_chain_matches_oracle(require_gpu, level, its, cells):
    """Iterations 1..its: the GPU step (FP64), then the chain on the GPU and on
    the oracle from the same NCCresult and GridPT3; the next iteration
    searches the chain's grid."""
    from setsm_amd import smx
    sc, pyr = scene(image_px=1024, relief=90.0, seed=17, levels=5)
    g = level_grid(sc, level, cells, band=min(48.0 * (1 << level), 400.0), seed=3)
    B = level if level >= 4 else level + 1
    ow = ortho_wgs(sc, g["boundary"], B)
    hint = float(g["minmax"][1] - g["minmax"][0])
    job = smx.Job(5)
    try:
        upload_pyramid(job, pyr)
        # the chain's VerticalLineLocus_blunder leaves ortho NCC in GridPT3, so
        # the next iteration's GNCC term needs the level's ortho grid
        gr = GpuRun(job, level, pyr, sc.rpcs, g, sgm_py=-1, ortho_wgs=ortho_wgs(sc, g["boundary"], level))
        job.set_blunder_level(B, ow)
        ug = g["ugrid"]
        for it in range(1, its + 2):
            r = gr.step(ug, it, g["minmax"], precision=smx.PREC_FP64)
            assert r["evaluated"] > 0
            if it > its:
                break
            ncc, ug_o = job.get_nccresult(), job.get_searchspace()
            kw = dict(mps_args=SEL, hinterval=hint, mpp=1.0, dem_resolution=2.0)
            rg = host_loop.iteration_chain(job, level, it, g["boundary"], g["grid_res"], **kw)
            ro = host_loop.iteration_chain(O.ChainJob(gr.li, ncc, ug_o), level, it, g["boundary"], g["grid_res"],
                                           ops=O.ChainOps(gr.li, ug_o, pyr, B, ow), **kw)
            assert (rg.points, rg.kept, rg.triangles, rg.minimum_th) == \
                (ro.points, ro.kept, ro.triangles, ro.minimum_th)
            assert rg.min_max_height == ro.min_max_height
            dec_equal(rg.anchors, ro.anchors)
            dec_equal(rg.blunders, ro.blunders)
            ug = job.get_searchspace()
            grids_equal(ug, ug_o)
            b = rg.blunders
            print(f"level {level} it {it}: {rg.points} points, anchor passes "
                  f"{rg.anchors.passes if rg.anchors else '-'}, blunder passes {b.passes if b else '-'} "
                  f"({b.blunders if b else 0} blunders), {rg.kept} kept, {rg.triangles} triangles, "
                  f"heights {rg.min_max_height}", flush=True)
            assert rg.points > 100 and rg.triangles > 100
    finally:
        job.close()
