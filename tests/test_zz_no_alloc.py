"""smx_vertical_line_locus_noalloc takes the call's scratch from the library's
static device array instead of the job: it must give exactly what
smx_vertical_line_locus gives on the same inputs (cube words, CSR offsets,
NCCresult array, accessible and evaluated counts, FP64 hand-over), on two
calls in a row over the one shared array, the second with a smaller input,
and refuse a level whose scratch would not fit, with a clear error."""
import numpy as np
import pytest

from tests.helpers import GpuRun, level_grid, perturb_searchspace, scene, upload_pyramid

pytestmark = pytest.mark.gpu


def _tile(level, cells, seed, px, band):
    from setsm_amd.smx import Job
    sc, pyr = scene(image_px=px, relief=80.0, seed=seed, levels=level + 1)
    g = level_grid(sc, level, cells, band=band, seed=seed)
    jobs = []
    for _ in range(2):
        j = Job(level + 1)
        upload_pyramid(j, pyr)
        GpuRun(j, level, pyr, sc.rpcs, g)
        jobs.append(j)
    return g, jobs


def _bring(j, g, level, iteration, precision):
    """Run the iterations before `iteration` and InitializeVoxel of it."""
    ug1 = g["ugrid"]
    seq = [ug1] if iteration == 1 else [ug1, perturb_searchspace(ug1, 7 + level, level)]
    for it, ug in enumerate(seq, start=1):
        j.set_searchspace(ug, g["minmax"])
        j.initialize_voxel(it)
        if it < iteration:
            j.vertical_line_locus(it, precision=precision)
            j.awncc(it)


@pytest.mark.parametrize("precision,level,iteration", [(0, 0, 1), (1, 0, 1), (1, 1, 2)])
def test_noalloc_equals_allocating_call(require_gpu, precision, level, iteration):
    big = _tile(level, 200, 71 + level, 1024, 32)
    small = _tile(level, 40, 72 + level, 512, 16)
    for tag, (g, (ref, na)) in (("big", big), ("small", small)):   # the second call is the smaller one
        _bring(ref, g, level, iteration, precision)
        _bring(na, g, level, iteration, precision)
        acc_r = ref.vertical_line_locus(iteration, precision=precision)
        st_r = ref.last_stats
        acc_n = na.vertical_line_locus_noalloc(iteration, precision=precision)
        st_n = na.last_stats
        print(f"{tag}: precision {precision} level {level} iteration {iteration}: {na.num_hypotheses} hypotheses, "
              f"accessible {acc_n}/{acc_r}, evaluated {st_n['evaluated']}/{st_r['evaluated']}", flush=True)
        assert acc_n == acc_r
        wr, _, offr = ref.get_voxels()
        wn, _, offn = na.get_voxels()
        assert np.array_equal(offr, offn)
        assert np.array_equal(wr, wn)
        assert np.array_equal(ref.get_nccresult(), na.get_nccresult())
        assert st_n["evaluated"] == st_r["evaluated"] > 0
        assert (st_n["redo_runs"], st_n["redo_hyps"]) == (st_r["redo_runs"], st_r["redo_hyps"])
    for g, jobs in (big, small):
        for j in jobs:
            j.close()


def test_noalloc_rejects_beyond_its_scratch(require_gpu):
    """500^2 cells x 992 heights in runs of 32: ~7.8 M height runs (16 B each
    in the precision form), beyond the static array."""
    from setsm_amd import smx
    g, (ref, na) = _tile(0, 500, 77, 2304, 248)
    runs32 = [i for i, d in smx.fast_configs() if d == (1, 16, 7680, 0, 1, 1, 32)]
    na.set_fast_config(runs32[0])
    _bring(na, g, 0, 1, smx.PREC_FAST)
    before = na.get_nccresult()
    with pytest.raises(smx.SmxError, match="static array holds"):
        na.vertical_line_locus_noalloc(1, precision=smx.PREC_FAST)
    assert np.array_equal(na.get_nccresult(), before)   # nothing was queued
    ref.close()
    na.close()
