"""Batched FAST launch: one persistent kernel over the tiles of several jobs
(different sizes, height ranges, pyramids, levels) gives each job exactly the
cube, NCCresult array, evaluated count and FP64 hand-over it gets alone, and
agrees with the FP64 oracle within 1e-5.

The batch always runs on fresh jobs (their own InitializeVoxel, nothing left
in the cube by an earlier call), so a launch that skipped a tile or a job
could not pass on stale words."""
import numpy as np
import pytest

from tests.helpers import GpuRun, OracleRun, level_grid, perturb_searchspace, scene, upload_pyramid

pytestmark = pytest.mark.gpu


def _make(level, cells, seed, px=512, relief=60.0, band=None):
    """Scene, pyramid, grid of one tile and two jobs on it (alone / batched)."""
    from setsm_amd.smx import Job
    sc, pyr = scene(image_px=px, relief=relief, seed=seed, levels=level + 1)
    g = level_grid(sc, level, cells, band=band, seed=seed)
    jobs = []
    for _ in range(2):
        job = Job(level + 1)
        upload_pyramid(job, pyr)
        GpuRun(job, level, pyr, sc.rpcs, g)
        jobs.append(job)
    return dict(sc=sc, pyr=pyr, g=g, level=level, alone=jobs[0], batch=jobs[1])


def _state(job):
    w, rho, off = job.get_voxels(with_rho=True)
    st = job.last_stats
    return dict(words=w, rho=rho, off=off, ncc=job.get_nccresult(), evaluated=job.last_evaluated,
                redo_runs=st["redo_runs"], redo_hyps=st["redo_hyps"])


def _assert_same(a, b, tag):
    assert np.array_equal(a["off"], b["off"]), tag
    assert np.array_equal(a["words"], b["words"]), tag
    assert np.array_equal(np.isnan(a["rho"]), np.isnan(b["rho"])), tag
    ev = ~np.isnan(a["rho"])
    assert np.array_equal(a["rho"][ev], b["rho"][ev]), tag
    assert np.array_equal(a["ncc"], b["ncc"]), tag
    assert a["evaluated"] == b["evaluated"] > 0, tag
    assert (a["redo_runs"], a["redo_hyps"]) == (b["redo_runs"], b["redo_hyps"]), tag


def _oracle_check(m, ug_seq, minmax, rho, acc):
    """The oracle's last step after running the iterations in ug_seq."""
    o = OracleRun(m["pyr"], m["sc"].rpcs, m["level"], m["g"])
    for it, ug in enumerate(ug_seq, start=1):
        ro = o.step(ug, it, minmax)
    ev = ~np.isnan(rho)
    evo = np.isfinite(ro["rho"])
    assert np.array_equal(evo, ev)
    if ev.any():
        assert np.max(np.abs(rho[ev].astype(np.float64) - ro["rho"][ev])) <= 1e-5
    assert acc == ro["acc"]


def _config(run_len, tx=1, ty=16, lds=7680):
    """Index of the launch shape (tx, ty, lds, compiler's registers, staged,
    FP64 totals, run_len)."""
    from setsm_amd.smx import fast_configs
    for i, desc in fast_configs():
        if desc == (tx, ty, lds, 0, 1, 1, run_len):
            return i
    raise LookupError((tx, ty, lds, run_len))


@pytest.mark.parametrize("level,specs,run_len", [
    (0, [(60, 3, 512, 60.0, 32), (24, 4, 512, 80.0, 16), (44, 5, 512, 40.0, 64)], None),
    (0, [(60, 3, 512, 60.0, 33), (24, 4, 512, 80.0, 17), (44, 5, 512, 40.0, 9)], 32),   # runs of 32 heights
    (0, [(60, 3, 512, 60.0, 33), (24, 4, 512, 80.0, 17), (44, 5, 512, 40.0, 9)], 16),   # runs of 16 heights
    (3, [(20, 8, 1024, 120.0, 96), (12, 9, 1024, 60.0, 48)], None),      # half window 6
    (4, [(10, 10, 1024, 150.0, 120), (8, 11, 1024, 90.0, 64)], None),     # half window 7
])
def test_batch_equals_alone(require_gpu, level, specs, run_len):
    from setsm_amd import smx
    made = [_make(level, *s[:4], band=s[4]) for s in specs]
    if run_len:
        cfg = _config(run_len)
        for m in made:
            m["alone"].set_fast_config(cfg)
            m["batch"].set_fast_config(cfg)
    alone = []
    for m in made:
        j = m["alone"]
        j.set_searchspace(m["g"]["ugrid"], m["g"]["minmax"])
        j.initialize_voxel(1)
        j.vertical_line_locus(1, precision=smx.PREC_FAST, keep_rho=True)
        alone.append(_state(j))
    jobs = []
    for m in made:
        j = m["batch"]
        j.set_searchspace(m["g"]["ugrid"], m["g"]["minmax"])
        j.initialize_voxel(1)
        jobs.append(j)
    acc = smx.vertical_line_locus_batch(jobs, 1, precision=smx.PREC_FAST, keep_rho=True)
    for k, m in enumerate(made):
        b = _state(m["batch"])
        print(f"level {level} job {k}: {len(b['words'])} hypotheses, evaluated {b['evaluated']}, "
              f"accessible {acc[k]}, redo {b['redo_runs']}", flush=True)
        _assert_same(alone[k], b, f"level {level} job {k}")
        _oracle_check(m, [m["g"]["ugrid"]], m["g"]["minmax"], b["rho"], acc[k])
    for m in made:
        m["alone"].close()
        m["batch"].close()


def test_batch_combined_iteration(require_gpu):
    """Iteration 2 at levels 1 and 2 adds the level-below term
    (setsm_code.cpp:6087-6088): those jobs take the two-pass persistent
    launches, the level-0 job in the same batch the single-pass one."""
    from setsm_amd import smx
    made = [_make(1, 40, 21, 1024, 80.0, 24), _make(2, 30, 22, 1024, 120.0, 48), _make(0, 36, 23, 512, 60.0, 20)]
    seqs = []
    for m in made:
        ug1 = m["g"]["ugrid"]
        ug2 = perturb_searchspace(ug1, 5 + m["level"], m["level"])
        seqs.append([ug1, ug2])
        for j in (m["alone"], m["batch"]):
            j.set_searchspace(ug1, m["g"]["minmax"])
            j.initialize_voxel(1)
            j.vertical_line_locus(1, precision=smx.PREC_FAST, keep_rho=True)
            j.awncc(1)
            j.set_searchspace(ug2, m["g"]["minmax"])
            j.initialize_voxel(2)
    alone = []
    for m in made:
        m["alone"].vertical_line_locus(2, precision=smx.PREC_FAST, keep_rho=True)
        alone.append(_state(m["alone"]))
    acc = smx.vertical_line_locus_batch([m["batch"] for m in made], 2, precision=smx.PREC_FAST, keep_rho=True)
    for k, m in enumerate(made):
        b = _state(m["batch"])
        print(f"combined batch job {k} (level {m['level']}): evaluated {b['evaluated']}, redo {b['redo_runs']}",
              flush=True)
        _assert_same(alone[k], b, f"job {k}")
        _oracle_check(m, seqs[k], m["g"]["minmax"], b["rho"], acc[k])
        m["alone"].close()
        m["batch"].close()


def test_batch_refuses_before_queueing(require_gpu):
    """Jobs of different window sizes cannot share the persistent launch: the
    call fails before anything is queued, every job keeps its state."""
    from setsm_amd import smx
    made = [_make(0, 20, 31), _make(3, 12, 32, 1024, 80.0, 48)]
    before = []
    for m in made:
        j = m["batch"]
        j.set_searchspace(m["g"]["ugrid"], m["g"]["minmax"])
        j.initialize_voxel(1)
        before.append((j.get_nccresult(), j.get_voxels()[0]))
    assert [m["batch"].half for m in made] == [5, 6]
    with pytest.raises(smx.SmxError, match="one window size"):
        smx.vertical_line_locus_batch([m["batch"] for m in made], 1, precision=smx.PREC_FAST)
    for m, (ncc, words) in zip(made, before):
        assert np.array_equal(m["batch"].get_nccresult(), ncc)
        assert np.array_equal(m["batch"].get_voxels()[0], words)
        assert m["batch"].vertical_line_locus(1, precision=smx.PREC_FAST) >= 0   # still usable
        m["alone"].close()
        m["batch"].close()


def test_batch_fp64_is_per_job(require_gpu):
    from setsm_amd import smx
    made = [_make(0, 30, 6), _make(0, 20, 7)]
    for m in made:
        m["batch"].set_searchspace(m["g"]["ugrid"], m["g"]["minmax"])
        m["batch"].initialize_voxel(1)
    smx.vertical_line_locus_batch([m["batch"] for m in made], 1, precision=smx.PREC_FP64, keep_rho=False)
    for m in made:
        o = OracleRun(m["pyr"], m["sc"].rpcs, 0, m["g"])
        ro = o.step(m["g"]["ugrid"], 1, m["g"]["minmax"])
        w, _, _ = m["batch"].get_voxels()
        assert np.array_equal(w, ro["words"])
        m["alone"].close()
        m["batch"].close()


def test_trim_then_rerun(require_gpu):
    """smx_job_trim frees the per-call lists; the next calls (alone and
    batched, with a shape of short runs in between) reserve them again and
    give the same cube and counts."""
    from setsm_amd import smx
    m = _make(0, 40, 41)
    j = m["alone"]

    def fresh():
        j.set_searchspace(m["g"]["ugrid"], m["g"]["minmax"])
        j.initialize_voxel(1)

    states = []
    for cfg in (_config(64), _config(16), _config(64)):
        j.set_fast_config(cfg)
        fresh()
        j.vertical_line_locus(1, precision=smx.PREC_FAST, keep_rho=True)
        states.append(_state(j))
        before = j.device_bytes
        j.trim()
        assert j.device_bytes < before
    _assert_same(states[0], states[2], "rerun after trim")
    assert states[1]["evaluated"] == states[0]["evaluated"]
    fresh()
    smx.vertical_line_locus_batch([j], 1, precision=smx.PREC_FAST, keep_rho=True)
    _assert_same(states[0], _state(j), "batched after trim")
    m["alone"].close()
    m["batch"].close()
