"""RA and matching-rate passes of VerticalLineLocus on the GPU against the
oracle (setsm_code.cpp:6308-6313, 6468-6496 and FindPeakNcc :6672-6752):
heights counted from the search-space bottom, no voxel cube, NCCresult
(result0..4 from FindPeakNcc in the matching-rate pass, max_WNCC,
NumOfHeight, check_height_change) bit for bit."""
import numpy as np
import pytest

from setsm_amd.records import LevelInputs
from setsm_amd.synth import ortho_wgs
from tests.helpers import level_grid, scene, upload_pyramid
from tests.test_gpu_parity import NCC_FIELDS

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,level,iteration,gncc", [("cmr", 3, 1, False), ("cmr", 4, 2, False),
                                                      ("cmr", 2, 2, True), ("ra", 4, 1, False),
                                                      ("ra", 3, 2, False)])
def test_ra_and_matching_rate_bit_exact(require_gpu, mode, level, iteration, gncc):
    from oracle import oracle_py as O
    from setsm_amd.smx import PREC_FAST, Job
    sc, pyr = scene(image_px=1024, relief=120.0, seed=5, levels=5)
    g = level_grid(sc, level, {2: 36, 3: 28, 4: 16}[level])
    kw = dict(is_ra=1) if mode == "ra" else dict(check_matching_rate=1)
    if gncc:
        g["ugrid"]["ortho_ncc"][:, 1] = 900
        kw["ortho_wgs"] = ortho_wgs(sc, g["boundary"], level)
    li = LevelInputs(pyr, level, sc.rpcs, g, grid_resolution=g["grid_res"], **kw)
    st = O.OracleState(li, keep_rho=False)
    ug_o = g["ugrid"].copy()
    job = Job(5)
    upload_pyramid(job, pyr)
    job.set_level(li.level)
    job.set_searchspace(g["ugrid"], g["minmax"])
    if mode == "ra":
        st.initialize_voxel(ug_o, iteration, g["minmax"])
        job.initialize_voxel(iteration)
    acc_o = st.vertical_line_locus(ug_o, iteration, g["minmax"])
    acc_g = job.vertical_line_locus(iteration, precision=PREC_FAST)   # these passes run the FP64 kernel
    ng = job.get_nccresult()
    print(f"[{mode} L{level} it{iteration} gncc={gncc}] acc {acc_o}/{acc_g}, peaks in "
          f"{int((st.ncc['result4'] > 0).sum())} cells", flush=True)
    assert acc_o == acc_g
    for f in NCC_FIELDS:
        a, b = st.ncc[f], ng[f]
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{mode} L{level}: field {f} differs at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"
    if mode == "cmr":
        assert (st.ncc["result4"] > 0).sum() > 50
    job.close()
