"""Oracle matching-rate pass (VerticalLineLocus with check_matching_rate,
setsm_code.cpp:6308-6313, 6492-6496, FindPeakNcc :6672-6752) pinned on the
CPU: the pass evaluates the heights start_H + h*step of the search space,
so its NCC sequence is a sub-sequence of the voxel pass's (same heights, same
NCC); an independent numpy FindPeakNcc over that sequence must give the
pass's result0..4 and max_WNCC."""
import numpy as np

from oracle import oracle_py as O
from setsm_amd.records import LevelInputs
from tests.helpers import level_grid, scene


def q_result(v):
    return int(np.int16(np.int32(np.round(v * 1000.0))))


def find_peak(seq):
    """setsm_code.cpp:6672-6752 over (height, rho) pairs in height order."""
    r0 = r1 = q_result(-1.0)
    r2 = r3 = -1000.0
    r4 = 0
    pre_rho, pre_h, direction, check = -1.0, 0.0, 0, False
    mx = -100.0
    for h, rho in seq:
        if rho >= 0 and pre_rho >= 0:
            d = rho - pre_rho
        elif rho < 0 and pre_rho < 0:
            d = -(abs(rho) - abs(pre_rho))
        elif rho >= 0 and pre_rho < 0:
            d = 1
        else:
            d = -1
        t = 1 if d > 0 else (-1 if d < 0 else 0)
        if not check:
            check, t = True, -1
        elif pre_rho != -1 and direction > 0 and t < 0:
            r4 = min(r4 + 1, 10) if r4 < 10 else r4
            if r0 < q_result(pre_rho):
                t1 = r0 / 1000.0
                r0 = q_result(pre_rho)
                t2 = r2
                r2 = pre_h
                if r1 < q_result(t1):
                    r1, r3 = q_result(t1), t2
            elif r1 < q_result(pre_rho):
                r1, r3 = q_result(pre_rho), pre_h
        pre_rho, pre_h, direction = rho, h, t
        mx = max(mx, rho)
    return r0, r1, r2, r3, r4, q_result(mx)


def test_matching_rate_pass_matches_numpy_findpeak():
    sc, pyr = scene(image_px=1024, relief=120.0, seed=5, levels=5)
    level = 3
    g = level_grid(sc, level, 24)
    step = O.lib().orc_height_step(level, 0.5)
    # voxel pass: the NCC of every height of the band
    li = LevelInputs(pyr, level, sc.rpcs, g, grid_resolution=g["grid_res"])
    st = O.OracleState(li, keep_rho=True)
    ug = g["ugrid"].copy()
    st.initialize_voxel(ug, 1, g["minmax"])
    st.vertical_line_locus(ug, 1, g["minmax"])
    _, rho = st.export()
    off = np.concatenate([[0], np.cumsum(st.ncc["NumOfHeight"].astype(np.int64))])
    # matching-rate pass on the same search space
    lm = LevelInputs(pyr, level, sc.rpcs, g, grid_resolution=g["grid_res"], check_matching_rate=1)
    sm = O.OracleState(lm, keep_rho=False)
    ug2 = g["ugrid"].copy()
    sm.vertical_line_locus(ug2, 1, g["minmax"])
    checked = 0
    for c in range(len(ug)):
        nm = sm.ncc[c]
        if not nm["check_height_change"]:
            continue
        s_h, e_h = int(ug2["minHeight"][c]), int(ug2["maxHeight"][c])
        n = int((e_h - s_h) / step)
        assert nm["NumOfHeight"] == n
        base = int(st.ncc["minHeight"][c])
        nn = int(st.ncc["NumOfHeight"][c])
        if nn == 0 or int(round((s_h - base) / step)) < 0 or int(round((s_h + (n - 1) * step - base) / step)) >= nn:
            continue   # the voxel pass did not lay out these heights (e.g. a band ending at 0)
        seq = []
        for k in range(n):
            hh = np.float32(s_h + k * step)
            j = int(round((float(hh) - base) / step))
            assert np.float32(base + j * step) == hh
            r = rho[off[c] + j]
            if np.isfinite(r):
                seq.append((float(hh), float(r)))
        want = find_peak(seq)
        got = (int(nm["result0"]), int(nm["result1"]), float(nm["result2"]), float(nm["result3"]),
               int(nm["result4"]), int(nm["max_WNCC"]))
        assert got == want, (c, got, want)
        checked += 1
    assert checked > 200
