"""Where the precision form's AWNCC peaks can move.

AWNCC (setsm_code.cpp:7511-7702) ranks the heights of a cell by the
quantised INCC word ((short)(rho*1000), Template.hpp:984-987), so a peak can
only move where the FAST word and the FP64 word of some height differ, and
the precision contract bounds that difference by one quantum. This test runs
two iterations of SETSM's schedule at level 0 on both paths and checks:

  * every INCC word differs by at most one quantum, the flags are equal;
  * a cell whose words are all equal gets the same peaks from AWNCC
    (result0..4: the peak scan adds no divergence of its own);
  * the cells where the peak moved are therefore all cells holding a
    one-quantum word difference (counted and printed).
"""
import numpy as np
import pytest

from tests.helpers import GpuRun, level_grid, perturb_searchspace, scene, upload_pyramid

pytestmark = pytest.mark.gpu


def test_peaks_move_only_at_one_quantum_words(require_gpu):
    from setsm_amd import smx
    sc, pyr = scene(image_px=1024, relief=90.0, seed=17, levels=1)
    g = level_grid(sc, 0, 200, band=32.0, seed=5)
    ug1 = g["ugrid"]
    ug2 = perturb_searchspace(ug1, 23, 0)
    runs = {}
    for prec in (smx.PREC_FP64, smx.PREC_FAST):
        job = smx.Job(1)
        upload_pyramid(job, pyr)
        gr = GpuRun(job, 0, pyr, sc.rpcs, g)
        runs[prec] = [gr.step(ug, it, g["minmax"], precision=prec) for it, ug in ((1, ug1), (2, ug2))]
        job.close()
    for k in range(2):
        a, b = runs[smx.PREC_FP64][k], runs[smx.PREC_FAST][k]
        assert np.array_equal(a["off"], b["off"])
        assert np.array_equal(a["words"] & 0xff, b["words"] & 0xff)
        ia = (a["words"] >> 16).astype(np.uint16).view(np.int16).astype(np.int64)
        ib = (b["words"] >> 16).astype(np.uint16).view(np.int16).astype(np.int64)
        dq = np.abs(ia - ib)
        assert dq.max(initial=0) <= 1
        # cells whose row holds a differing word
        lens = np.diff(a["off"])
        cell_of = np.repeat(np.arange(len(lens)), lens)
        touched = np.zeros(len(lens), bool)
        touched[cell_of[dq > 0]] = True
        na, nb = a["ncc"], b["ncc"]
        # AWNCC's outputs (max_WNCC comes from VerticalLineLocus' unquantised
        # values, within the contract, and is not compared here)
        fields = ("result0", "result1", "result2", "result3", "result4")
        same = np.logical_and.reduce([na[f] == nb[f] for f in fields])
        assert same[~touched].all(), "AWNCC differs in a cell whose words are identical"
        moved = (na["result2"] != nb["result2"]) & (na["result4"] > 0) & (nb["result4"] > 0)
        assert not (moved & ~touched).any()
        print(f"iteration {k + 1}: {int(dq.astype(bool).sum())} of {len(dq)} words one quantum apart, "
              f"{int(touched.sum())} cells touched, {int(moved.sum())} peaks moved", flush=True)
