"""Edge cases on the GPU against the oracle: nodata (DN 0) regions in both
images, flat (constant) patches, grids reaching the image borders,
degenerate search spaces (empty and single-height bands, every cell
blundered). FP64 bit for bit; FAST within the 1e-5 contract with the same
evaluated set (its degenerate windows decided by the FP64 pass)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import GpuRun, OracleRun, level_grid, scene
from tests.test_gpu_parity import _compare, _fast_vs_fp64

pytestmark = pytest.mark.gpu


def _nodata_pyramid(sc, levels, seed=3, value=None, patches=6):
    """Rectangles of DN 0 (value None) or of one constant DN per rectangle."""
    rng = np.random.default_rng(seed)
    imgs = []
    for ti in range(2):
        im = sc.images[ti].copy()
        h, w = im.shape
        for _ in range(patches):
            y0, x0 = rng.integers(0, h - 40), rng.integers(0, w - 40)
            dn = 0 if value is None else int(rng.integers(*value))
            im[y0:y0 + rng.integers(8, 40), x0:x0 + rng.integers(8, 40)] = dn
        imgs.append(im)
    return imgs, [O.build_pyramid(im, levels) for im in imgs]


def _job_with(pyr):
    from setsm_amd.smx import Job
    job = Job(len(pyr[0]))
    for ti in range(2):
        for lv, (img, mag, ori) in enumerate(pyr[ti]):
            job.set_view(ti, lv, img, mag, ori)
    return job


@pytest.mark.parametrize("level", [0, 1])
def test_nodata_regions(require_gpu, level):
    from setsm_amd.smx import PREC_FAST
    sc, _ = scene(image_px=512, levels=2) if level else scene(image_px=512, levels=1)
    _, pyr = _nodata_pyramid(sc, level + 1)
    g = level_grid(sc, level, 64 if level == 0 else 48)
    o = OracleRun(pyr, sc.rpcs, level, g)
    ro = o.step(g["ugrid"], 1, g["minmax"])
    job = _job_with(pyr)
    rg = GpuRun(job, level, pyr, sc.rpcs, g).step(g["ugrid"], 1, g["minmax"])
    _compare(ro, rg, f"nodata L{level}")
    o2 = OracleRun(pyr, sc.rpcs, level, g)
    ro2 = o2.step(g["ugrid"], 1, g["minmax"])
    job2 = _job_with(pyr)
    rf = GpuRun(job2, level, pyr, sc.rpcs, g).step(g["ugrid"], 1, g["minmax"], precision=PREC_FAST)
    _fast_vs_fp64(ro2, rf, len(g["ugrid"]), f"FAST nodata L{level}")
    print(f"[FAST nodata L{level}] runs redone in FP64: {job2.last_stats['redo_runs']}", flush=True)
    job.close()
    job2.close()


def test_flat_patches_fast(require_gpu):
    """Constant non-zero patches: the reference's window variance is an exact
    zero there; the FAST kernel hands those runs to the FP64 pass instead of
    correlating FP32 rounding noise."""
    from setsm_amd.smx import PREC_FAST
    sc, _ = scene(image_px=512, levels=1)
    _, pyr = _nodata_pyramid(sc, 1, seed=5, value=(50, 3000), patches=10)
    g = level_grid(sc, 0, 64)
    o = OracleRun(pyr, sc.rpcs, 0, g)
    ro = o.step(g["ugrid"], 1, g["minmax"])
    job = _job_with(pyr)
    rf = GpuRun(job, 0, pyr, sc.rpcs, g).step(g["ugrid"], 1, g["minmax"], precision=PREC_FAST)
    redo = job.last_stats["redo_runs"]
    print(f"[FAST flat patches] runs redone in FP64: {redo}", flush=True)
    assert redo > 0
    _fast_vs_fp64(ro, rf, len(g["ugrid"]), "FAST flat patches")
    job.close()


def test_grid_at_image_borders(require_gpu):
    """Cells whose windows reach the image edge: per-sample bounds tests in
    both kernels, the FAST kernel's checked and HBM paths."""
    from setsm_amd.smx import PREC_FAST
    sc, pyr = scene(image_px=512, levels=1)
    side = sc.spec.image_px * sc.spec.gsd
    g = level_grid(sc, 0, int((side - 4.0) / 2.0), inset=2.0)
    o = OracleRun(pyr, sc.rpcs, 0, g)
    ro = o.step(g["ugrid"], 1, g["minmax"])
    job = _job_with(pyr)
    gr = GpuRun(job, 0, pyr, sc.rpcs, g)
    rg = gr.step(g["ugrid"], 1, g["minmax"])
    _compare(ro, rg, "borders FP64")
    o2 = OracleRun(pyr, sc.rpcs, 0, g)
    ro2 = o2.step(g["ugrid"], 1, g["minmax"])
    job2 = _job_with(pyr)
    rf = GpuRun(job2, 0, pyr, sc.rpcs, g).step(g["ugrid"], 1, g["minmax"], precision=PREC_FAST)
    _fast_vs_fp64(ro2, rf, len(g["ugrid"]), "FAST borders")
    job.close()
    job2.close()


@pytest.mark.parametrize("case", ["empty_bands", "all_blunders"])
def test_degenerate_search_spaces(require_gpu, case):
    from setsm_amd.smx import PREC_FAST
    sc, pyr = scene(image_px=512, levels=1)
    g = level_grid(sc, 0, 40)
    ug = g["ugrid"].copy()
    if case == "empty_bands":
        ug["maxHeight"][::2] = ug["minHeight"][::2]          # NumOfHeight 0
        ug["maxHeight"][1::4] = ug["minHeight"][1::4] + 1     # 4 heights
    else:
        ug["Matched_flag"][:] = 0                             # level 0: every cell is a blunder cell
    for prec in (0, PREC_FAST):
        o = OracleRun(pyr, sc.rpcs, 0, g)
        ro = o.step(ug, 1, g["minmax"])
        job = _job_with(pyr)
        rg = GpuRun(job, 0, pyr, sc.rpcs, g).step(ug, 1, g["minmax"], precision=prec)
        if case == "all_blunders":
            assert ro["acc"] == 0 and rg["acc"] == 0
        if prec == 0:
            _compare(ro, rg, f"{case} FP64")
        else:
            assert ro["acc"] == rg["acc"]
            _fast_vs_fp64(ro, rg, len(ug), f"{case} FAST")
        job.close()


@pytest.mark.parametrize("band", [249.0, 262.0])
def test_max_height_band(require_gpu, band):
    """Bands at the reference's 1000-height limit at level 0: 996 heights
    (long voxel rows, 16 runs of 64 per cell) and 1048 heights, which
    InitializeVoxel drops (setsm_code.cpp:5999-6009)."""
    from setsm_amd.smx import PREC_FAST
    sc, pyr = scene(image_px=512, levels=1)
    g = level_grid(sc, 0, 16, band=band)
    nh = g["ugrid"]["maxHeight"].astype(int) - g["ugrid"]["minHeight"].astype(int)
    print(f"[max band] band {nh.min()}..{nh.max()} m", flush=True)
    o = OracleRun(pyr, sc.rpcs, 0, g)
    ro = o.step(g["ugrid"], 1, g["minmax"])
    if band < 250:
        assert ro["ncc"]["NumOfHeight"].max() > 900
    else:
        assert len(ro["words"]) == 0
    job = _job_with(pyr)
    rg = GpuRun(job, 0, pyr, sc.rpcs, g).step(g["ugrid"], 1, g["minmax"])
    _compare(ro, rg, "max band FP64")
    o2 = OracleRun(pyr, sc.rpcs, 0, g)
    ro2 = o2.step(g["ugrid"], 1, g["minmax"])
    job2 = _job_with(pyr)
    rf = GpuRun(job2, 0, pyr, sc.rpcs, g).step(g["ugrid"], 1, g["minmax"], precision=PREC_FAST)
    _fast_vs_fp64(ro2, rf, len(g["ugrid"]), "FAST max band")
    job.close()
    job2.close()


def test_batch_with_degenerate_windows(require_gpu):
    """A batched launch over jobs with nodata and flat patches: each job's cube
    equals the one it gets alone (the FP64 pass over handed-over runs
    included), and the oracle within the contract."""
    from setsm_amd import smx
    sc, _ = scene(image_px=512, levels=1)
    pyrs = [_nodata_pyramid(sc, 1, seed=3)[1], _nodata_pyramid(sc, 1, seed=5, value=(50, 3000), patches=10)[1]]
    grids = [level_grid(sc, 0, 48, seed=11), level_grid(sc, 0, 40, band=16.0, seed=12)]
    jobs, alone = [], []
    for pyr, g in zip(pyrs, grids):
        job = _job_with(pyr)
        GpuRun(job, 0, pyr, sc.rpcs, g)
        job.set_searchspace(g["ugrid"], g["minmax"])
        job.initialize_voxel(1)
        job.vertical_line_locus(1, precision=smx.PREC_FAST, keep_rho=True)
        alone.append((job.get_voxels(with_rho=True), job.last_stats["redo_runs"]))
        jobs.append(job)
    acc = smx.vertical_line_locus_batch(jobs, 1, precision=smx.PREC_FAST, keep_rho=True)
    for k, (job, pyr, g) in enumerate(zip(jobs, pyrs, grids)):
        (w0, rho0, _), redo0 = alone[k]
        w, rho, _ = job.get_voxels(with_rho=True)
        redo = job.last_stats["redo_runs"]
        print(f"[batch degenerate] job {k}: {len(w)} hypotheses, redo runs alone {redo0} batched {redo}", flush=True)
        assert redo == redo0
        assert np.array_equal(w, w0)
        ev = ~np.isnan(rho)
        assert np.array_equal(ev, ~np.isnan(rho0))
        assert np.array_equal(rho[ev], rho0[ev])
        o = OracleRun(pyr, sc.rpcs, 0, g)
        ro = o.step(g["ugrid"], 1, g["minmax"])
        assert acc[k] == ro["acc"]
        assert np.array_equal(np.isfinite(ro["rho"]), ev)
        assert np.max(np.abs(rho[ev].astype(np.float64) - ro["rho"][ev])) <= 1e-5
    assert sum(a[1] for a in alone) > 0
    for job in jobs:
        job.close()


@pytest.mark.parametrize("run_len", [None, 32, 16])
def test_nodata_combined_iteration(require_gpu, run_len):
    """Level 1, iterations 1 and 2 (the second adds the level-0 term: two FAST
    passes, the level-0 window staged on its own tiles) over pyramids with
    nodata rectangles and flat patches: the FAST contract against the oracle,
    the degenerate windows of either pass decided by the FP64 pass. With runs
    of 32 or 16 heights a wave of the level-below pass holds runs the level
    pass handed over beside runs it did not: each is finished exactly once."""
    from setsm_amd.smx import PREC_FAST
    from tests.helpers import perturb_searchspace
    from tests.test_gpu_batch import _config
    sc, _ = scene(image_px=512, levels=2)
    imgs, _ = _nodata_pyramid(sc, 2, seed=7)
    rng = np.random.default_rng(9)
    for im in imgs:
        for _ in range(6):
            y0, x0 = rng.integers(0, 472, 2)
            im[y0:y0 + rng.integers(10, 40), x0:x0 + rng.integers(10, 40)] = int(rng.integers(50, 3000))
    pyr = [O.build_pyramid(im, 2) for im in imgs]
    g = level_grid(sc, 1, 48)
    o = OracleRun(pyr, sc.rpcs, 1, g)
    job = _job_with(pyr)
    if run_len:
        job.set_fast_config(_config(run_len))
    gr = GpuRun(job, 1, pyr, sc.rpcs, g)
    ug = g["ugrid"]
    for it in (1, 2):
        if it > 1:
            ug = perturb_searchspace(ug, 200 + it, 1)
        ro = o.step(ug, it, g["minmax"])
        rf = gr.step(ug, it, g["minmax"], precision=PREC_FAST)
        print(f"[FAST nodata combined R{run_len}] it{it}: runs redone in FP64 {job.last_stats['redo_runs']}, "
              f"evaluated {rf['evaluated']}", flush=True)
        _fast_vs_fp64(ro, rf, len(ug), f"FAST nodata L1 it{it}")
    job.close()
