"""One iteration's VerticalLineLocus + AWNCC captured as a hipGraph
(smx_graph_create) gives, on every replay, exactly what the eager calls give
on the same inputs: cube words, mean NCC, NCCresult array, accessible and
evaluated counts. Replays are interleaved with eager work on other jobs and
host synchronisations, and each one is checked."""
import numpy as np
import pytest

from tests.helpers import GpuRun, level_grid, perturb_searchspace, scene, upload_pyramid

pytestmark = pytest.mark.gpu


def _pair(level, cells, seed, px=512, relief=60.0, band=None, sgm_py=-1):
    from setsm_amd.smx import Job
    sc, pyr = scene(image_px=px, relief=relief, seed=seed, levels=level + 1)
    g = level_grid(sc, level, cells, band=band, seed=seed)
    jobs = []
    for _ in range(2):
        j = Job(level + 1)
        upload_pyramid(j, pyr)
        GpuRun(j, level, pyr, sc.rpcs, g, sgm_py=sgm_py)
        jobs.append(j)
    return dict(g=g, level=level, eager=jobs[0], graph=jobs[1])


def _prepare(m, iteration, precision):
    """Bring both jobs to the start of `iteration` the same way."""
    ug1 = m["g"]["ugrid"]
    seq = [ug1] if iteration == 1 else [ug1, perturb_searchspace(ug1, 3 + m["level"], m["level"])]
    for j in (m["eager"], m["graph"]):
        for it, ug in enumerate(seq, start=1):
            j.set_searchspace(ug, m["g"]["minmax"])
            j.initialize_voxel(it)
            if it < iteration:
                j.vertical_line_locus(it, precision=precision)
                j.awncc(it)


def _snap(j):
    w, _, off = j.get_voxels()
    return dict(words=w, off=off, ncc=j.get_nccresult(), evaluated=j.last_evaluated)


def _same(a, b, tag):
    assert np.array_equal(a["off"], b["off"]), tag
    assert np.array_equal(a["words"], b["words"]), tag
    assert np.array_equal(a["ncc"], b["ncc"]), tag
    assert a["evaluated"] == b["evaluated"], tag


@pytest.mark.parametrize("level,iteration,precision,sgm_py", [
    (0, 1, 0, -1),   # FP64
    (0, 1, 1, -1),   # FAST
    (1, 2, 1, -1),   # FAST, combined-level iteration (two passes)
    (2, 2, 0, -1),   # FP64, combined-level iteration
    (0, 1, 1, 1),    # FAST + SGM aggregation before the peak scan
])
def test_graph_replays_equal_eager(require_gpu, level, iteration, precision, sgm_py):
    from setsm_amd import smx
    m = _pair(level, 40 if level else 48, 41 + level, 1024 if level else 512, 80.0, 24, sgm_py)
    other = _pair(0, 20, 49, 512, 40.0, 16)
    _prepare(m, iteration, precision)
    _prepare(other, 1, 0)
    acc_e = m["eager"].vertical_line_locus(iteration, precision=precision)
    m["eager"].awncc(iteration)
    want = _snap(m["eager"])
    g = smx.Graph([m["graph"]], iteration, precision=precision, with_awncc=True)
    for rep in range(3):
        acc = g.launch()
        print(f"level {level} it {iteration} prec {precision} sgm {sgm_py}: replay {rep} {g.last_ms:.3f} ms, "
              f"accessible {acc[0]}", flush=True)
        assert acc[0] == acc_e
        _same(want, _snap(m["graph"]), f"replay {rep}")
        # eager work of another job between replays
        other["eager"].vertical_line_locus(1, precision=smx.PREC_FAST)
        other["eager"].awncc(1)
    g.close()
    for d in (m, other):
        d["eager"].close()
        d["graph"].close()


def test_batched_graph_equals_eager_batch(require_gpu):
    from setsm_amd import smx
    made = [_pair(0, 50, 51, 512, 60.0, 32), _pair(0, 24, 52, 512, 90.0, 16), _pair(0, 36, 53, 512, 40.0, 48)]
    for m in made:
        _prepare(m, 1, 1)
    acc_e = smx.vertical_line_locus_batch([m["eager"] for m in made], 1, precision=smx.PREC_FAST)
    want = []
    for m in made:
        m["eager"].awncc(1)
        want.append(_snap(m["eager"]))
    g = smx.Graph([m["graph"] for m in made], 1, precision=smx.PREC_FAST, with_awncc=True)
    for rep in range(3):
        acc = g.launch()
        print(f"batched graph replay {rep}: {g.last_ms:.3f} ms, accessible {list(acc)}", flush=True)
        assert list(acc) == list(acc_e)
        for k, m in enumerate(made):
            _same(want[k], _snap(m["graph"]), f"replay {rep} job {k}")
        made[0]["eager"].vertical_line_locus(1, precision=smx.PREC_FP64)   # eager work between replays
    g.close()
    for m in made:
        m["eager"].close()
        m["graph"].close()


def test_stale_graph_refuses(require_gpu):
    from setsm_amd import smx
    m = _pair(0, 20, 61)
    _prepare(m, 1, 1)
    g = smx.Graph([m["graph"]], 1, precision=smx.PREC_FAST)
    g.launch()
    m["graph"].set_searchspace(m["g"]["ugrid"], m["g"]["minmax"])
    with pytest.raises(smx.SmxError, match="capture again"):
        g.launch()
    g.close()
    m["eager"].close()
    m["graph"].close()
