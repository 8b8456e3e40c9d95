"""SGM aggregation + SGM-mode AWNCC on the GPU (sgm.hip, k_awncc) against the
oracle: SumCost and every NCCresult field bit for bit.

* real cubes: levels 0 and 1 with SGM_py = 1 after the FP64
  VerticalLineLocus, iterations 1-3 (level 0 from iteration 2 on takes the
  INCC-at-peak result mode);
* synthetic cubes seeded through smx_job_set_nccresult / smx_job_set_voxels:
  random bands, voxel flags and INCC, exercising every neighbour branch of
  SGM_con_pos, and a cell with more heights than the LDS path buffers hold
  (global-scratch variant)."""
import numpy as np
import pytest

from setsm_amd.records import NCCRESULT_DTYPE, UGRID_DTYPE
from tests.helpers import OracleRun, level_grid, perturb_searchspace, scene, upload_pyramid
from tests.test_gpu_parity import NCC_FIELDS

pytestmark = pytest.mark.gpu


def _cmp_ncc(a, b, tag):
    for f in NCC_FIELDS:
        bad = np.flatnonzero(a[f] != b[f])
        assert bad.size == 0, f"{tag}: field {f} differs at {bad[:5]}: {a[f][bad[:5]]} vs {b[f][bad[:5]]}"


def _cmp_sum(a, b, tag):
    assert a.shape == b.shape, tag
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
    assert bad.size == 0, f"{tag}: {bad.size} SumCost values differ, first {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


@pytest.mark.parametrize("level,iters", [(0, (1, 2, 3)), (1, (1, 2))])
def test_sgm_levels_bit_exact(require_gpu, level, iters):
    from setsm_amd.records import LevelInputs
    from setsm_amd.smx import PREC_FP64, Job
    if level == 0:
        sc, pyr = scene(image_px=512, levels=1)
        g = level_grid(sc, 0, 64)
        nlev = 1
    else:
        sc, pyr = scene(image_px=1024, relief=120.0, seed=5, levels=5)
        g = level_grid(sc, 1, 48)
        nlev = 5
    o = OracleRun(pyr, sc.rpcs, level, g, sgm_py=1)
    job = Job(nlev)
    upload_pyramid(job, pyr)
    li = LevelInputs(pyr, level, sc.rpcs, g, grid_resolution=g["grid_res"], sgm_py=1)
    job.set_level(li.level)
    ug = g["ugrid"]
    for it in iters:
        if it > 1:
            ug = perturb_searchspace(ug, 600 + it, level)
        ro = o.step(ug, it, g["minmax"])
        job.set_searchspace(ug, g["minmax"])
        job.initialize_voxel(it)
        job.vertical_line_locus(it, precision=PREC_FP64)
        words, _, _ = job.get_voxels()
        assert np.array_equal(words, ro["words"]), f"L{level} it{it}: cube differs before SGM"
        job.awncc(it)
        sg = job.get_sumcost()
        print(f"[L{level} it{it}] hyps={len(words)} sumcost range {sg.min() if sg.size else 0:.3f}.."
              f"{sg.max() if sg.size else 0:.3f}", flush=True)
        _cmp_sum(o.st.sumcost, sg, f"L{level} it{it}")
        _cmp_ncc(ro["ncc"], job.get_nccresult(), f"L{level} it{it}")
        ug = ro["ug"]
    job.close()


def _synthetic(W, H, seed, nh_max=40, big=0):
    rng = np.random.default_rng(seed)
    n = W * H
    ncc = np.zeros(n, NCCRESULT_DTYPE)
    ug = np.zeros(n, UGRID_DTYPE)
    base = rng.integers(100, 104, n)
    nh = rng.integers(1, nh_max + 1, n)
    nh[rng.random(n) < 0.1] = 0
    if big:
        nh[n // 2] = big
    ncc["minHeight"] = base
    ncc["NumOfHeight"] = nh
    ncc["maxHeight"] = base + np.ceil(nh * 0.25).astype(int)
    ncc["max_WNCC"] = rng.integers(-200, 1000, n)
    ncc["GNCC"] = -1000
    ug["minHeight"] = base + rng.integers(0, 3, n)
    ug["maxHeight"] = base + rng.integers(2, 12, n)
    if big:
        ug["maxHeight"][n // 2] = base[n // 2] + big // 4
    total = int(nh.sum())
    incc = rng.integers(-1000, 1000, total).astype(np.int16)
    flag = (rng.random(total) < 0.75).astype(np.uint32)
    words = flag | (incc.astype(np.uint16).astype(np.uint32) << 16)
    return ncc, ug, words


@pytest.mark.parametrize("W,H,seed,big", [(37, 23, 1, 0), (9, 64, 2, 0), (70, 5, 3, 0), (5, 7, 4, 7000)])
def test_sgm_synthetic_cubes(require_gpu, W, H, seed, big):
    import types

    from oracle import oracle_py as O
    from setsm_amd.records import SmxLevel
    from setsm_amd.smx import Job
    ncc, ug, words = _synthetic(W, H, seed, big=big)
    L = SmxLevel()
    L.level, L.sgm_py, L.height_step = 0, 1, 0.25
    L.grid_w, L.grid_h = W, H
    L.resolution, L.dem_resolution, L.template_size, L.bin_angle = 0.5, 2.0, 15, 20.0
    gw = np.zeros((W * H, 2))
    L.grid_wgs = gw.ctypes.data
    L.grid_pts = gw.ctypes.data
    st = O.OracleState(types.SimpleNamespace(level=L, ncells=W * H), keep_rho=False)
    st.import_voxels(ncc, words)
    st.awncc(ug, 2)

    job = Job(1)
    z16 = np.ones((32, 32), np.uint16)
    for ti in range(2):
        job.set_view(ti, 0, z16, z16, np.zeros((32, 32), np.uint8))
    job.set_level(L)
    job.set_searchspace(ug, np.array([0.0, 1000.0]))
    job.set_nccresult(ncc)
    job.set_voxels(words)
    job.awncc(2)
    print(f"[{W}x{H} seed {seed}] hyps={len(words)} max NumOfHeight={int(ncc['NumOfHeight'].max())}", flush=True)
    _cmp_sum(st.sumcost, job.get_sumcost(), f"{W}x{H}")
    _cmp_ncc(st.ncc, job.get_nccresult(), f"{W}x{H}")
    job.close()
