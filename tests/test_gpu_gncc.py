"""GNCC term on the GPU (k_ortho_coords, k_gncc and the INCC/GNCC blend in
max_WNCC) against the oracle: FP64 path bit for bit (GNCC, max_WNCC, every
NCCresult field, the cube), FAST path with GNCC identical (it is computed in
FP64 on both paths) and max_WNCC within one quantum."""
import numpy as np
import pytest

from setsm_amd.synth import ortho_wgs
from tests.helpers import GpuRun, OracleRun, level_grid, perturb_searchspace, scene, upload_pyramid
from tests.test_gpu_parity import _compare

pytestmark = pytest.mark.gpu


def _setup(level, cells, seed=9):
    if level == 0:
        sc, pyr = scene(image_px=512, levels=1)
        nlev = 1
    else:
        sc, pyr = scene(image_px=1024, relief=120.0, seed=5, levels=5)
        nlev = 5
    g = level_grid(sc, level, cells)
    rng = np.random.default_rng(seed + level)
    on = rng.random(len(g["ugrid"])) < 0.8
    g["ugrid"]["ortho_ncc"][:, 1] = np.where(on, 900, 50)
    g["ugrid"]["Height"] += rng.normal(0, 0.5, len(g["ugrid"])).astype(np.float32)
    g["ugrid"]["Height"][rng.random(len(g["ugrid"])) < 0.02] = -1000      # holes in the previous surface
    kw = dict(ortho_wgs=ortho_wgs(sc, g["boundary"], level), sgm_py=-1)
    return sc, pyr, g, nlev, kw


@pytest.mark.parametrize("level,cells,iters", [(0, 56, (2, 3)), (2, 40, (1, 2)), (3, 28, (1, 2)), (4, 16, (2,))])
def test_gncc_fp64_bit_exact(require_gpu, level, cells, iters):
    from setsm_amd.smx import Job
    sc, pyr, g, nlev, kw = _setup(level, cells)
    o = OracleRun(pyr, sc.rpcs, level, g, **kw)
    job = Job(nlev)
    upload_pyramid(job, pyr)
    gr = GpuRun(job, level, pyr, sc.rpcs, g, **kw)
    ug = g["ugrid"]
    for it in iters:
        if it != iters[0]:
            ug = perturb_searchspace(ug, 700 + it, level)
        ro = o.step(ug, it, g["minmax"])
        rg = gr.step(ug, it, g["minmax"])
        gn = ro["ncc_vll"]["GNCC"]
        print(f"[GNCC L{level} it{it}] cells with GNCC: {int((gn > -1000).sum())} of {len(gn)}", flush=True)
        assert (gn > -1000).sum() > len(gn) // 4
        _compare(ro, rg, f"GNCC L{level} it{it}")
        ug = ro["ug"]
    job.close()


def test_gncc_fast(require_gpu):
    from setsm_amd.smx import PREC_FAST, Job
    sc, pyr, g, nlev, kw = _setup(2, 40)
    o = OracleRun(pyr, sc.rpcs, 2, g, **kw)
    job = Job(nlev)
    upload_pyramid(job, pyr)
    gr = GpuRun(job, 2, pyr, sc.rpcs, g, **kw)
    ro = o.step(g["ugrid"], 2, g["minmax"])
    rg = gr.step(g["ugrid"], 2, g["minmax"], precision=PREC_FAST)
    assert np.array_equal(ro["ncc_vll"]["GNCC"], rg["ncc_vll"]["GNCC"])
    d = np.abs(ro["ncc_vll"]["max_WNCC"].astype(int) - rg["ncc_vll"]["max_WNCC"].astype(int))
    print(f"[GNCC FAST] max_WNCC max diff {d.max()}", flush=True)
    assert d.max() <= 1
    job.close()
