"""Batched FAST launch: one persistent kernel over the tiles of several jobs
(different sizes, height ranges, pyramids) gives each job exactly the cube it
gets alone, and agrees with the FP64 oracle within 1e-5."""
import numpy as np
import pytest

from tests.helpers import GpuRun, OracleRun, level_grid, scene, upload_pyramid

pytestmark = pytest.mark.gpu


def _make(level, cells, seed, px=512, relief=60.0, band=None):
    from setsm_amd.smx import Job
    sc, pyr = scene(image_px=px, relief=relief, seed=seed, levels=max(1, level + 1))
    g = level_grid(sc, level, cells, band=band, seed=seed)
    job = Job(max(1, level + 1))
    upload_pyramid(job, pyr)
    gr = GpuRun(job, level, pyr, sc.rpcs, g)
    return sc, pyr, g, job, gr


def test_batch_equals_alone(require_gpu):
    from setsm_amd import smx
    specs = [(0, 60, 3, 512, 60.0, 32), (0, 24, 4, 512, 80.0, 16), (0, 44, 5, 512, 40.0, 64)]
    made = [_make(*s[:5], band=s[5]) for s in specs]
    alone = []
    for sc, pyr, g, job, gr in made:
        job.set_searchspace(g["ugrid"], g["minmax"])
        job.initialize_voxel(1)
        job.vertical_line_locus(1, precision=smx.PREC_FAST, keep_rho=True)
        alone.append(job.get_voxels(with_rho=True))
    jobs = [m[3] for m in made]
    acc = smx.vertical_line_locus_batch(jobs, 1, precision=smx.PREC_FAST, keep_rho=True)
    for k, (sc, pyr, g, job, gr) in enumerate(made):
        w, rho, off = job.get_voxels(with_rho=True)
        print(f"job {k}: {len(w)} hypotheses, accessible {acc[k]}", flush=True)
        assert np.array_equal(w, alone[k][0])
        assert np.array_equal(np.isnan(rho), np.isnan(alone[k][1]))
        ev = ~np.isnan(rho)
        assert np.array_equal(rho[ev], alone[k][1][ev])
        # and the batch agrees with the FP64 oracle
        o = OracleRun(pyr, sc.rpcs, 0, g)
        ro = o.step(g["ugrid"], 1, g["minmax"])
        evo = np.isfinite(ro["rho"])
        assert np.array_equal(evo, ev)
        assert np.max(np.abs(rho[ev].astype(np.float64) - ro["rho"][ev])) <= 1e-5
        assert acc[k] == ro["acc"]
    for m in made:
        m[3].close()


def test_batch_fp64_is_per_job(require_gpu):
    from setsm_amd import smx
    made = [_make(0, 30, 6), _make(0, 20, 7)]
    for sc, pyr, g, job, gr in made:
        job.set_searchspace(g["ugrid"], g["minmax"])
        job.initialize_voxel(1)
    smx.vertical_line_locus_batch([m[3] for m in made], 1, precision=smx.PREC_FP64, keep_rho=False)
    for sc, pyr, g, job, gr in made:
        o = OracleRun(pyr, sc.rpcs, 0, g)
        ro = o.step(g["ugrid"], 1, g["minmax"])
        w, _, _ = job.get_voxels()
        assert np.array_equal(w, ro["words"])
        job.close()
