"""Pins of the oracle's SelectMPs and SetHeightRange (CPU).

The reference's own setsm_code.cpp does not build here, so each expectation
below is worked out by hand from the rules of setsm_code.cpp:8535-8889
(SelectMPs) and :9851-10261 with SubFunctions.cpp:2915-3024 (SetHeightRange)
for inputs small enough to follow, and the TIN comes from the reference's own
grid_triangulation.cpp (oracle/_ref/libref_tin.so) where it is built.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle_py as O
from setsm_amd.records import NCCRESULT_DTYPE, POINT3_DTYPE, TRI_DTYPE, UGRID_DTYPE, SmxLevel


class Lv:
    """An smx_level with only the fields these two functions read."""

    def __init__(self, w, h, level, gres=1.0, b0=0.0, b1=0.0, sgm_py=-1, is_ra=0, resolution=0.5,
                 dem_resolution=2.0, pre_dem=0):
        L = SmxLevel()
        L.level, L.sgm_py, L.is_ra = level, sgm_py, is_ra
        L.resolution, L.dem_resolution, L.pre_dem = resolution, dem_resolution, pre_dem
        L.grid_resolution = gres
        L.boundary[0], L.boundary[1] = b0, b1
        L.boundary[2], L.boundary[3] = b0 + (w - 1) * gres, b1 + (h - 1) * gres
        L.grid_w, L.grid_h = w, h
        c, r = np.meshgrid(np.arange(w), np.arange(h))
        self.gp = np.ascontiguousarray(np.stack([b0 + c.ravel() * gres, b1 + r.ravel() * gres], 1), np.float64)
        L.grid_pts = self.gp.ctypes.data
        self.level = L
        self.ncells = w * h


def mk_ncc(n, r0=0, r1=0, h1=0.0, h2=0.0, peaks=1, nh=10):
    a = np.zeros(n, NCCRESULT_DTYPE)
    a["result0"], a["result1"], a["result2"], a["result3"] = r0, r1, h1, h2
    a["result4"], a["NumOfHeight"] = peaks, nh
    return a


SEL = dict(th_roh=0.5, th_roh_min=0.45, th_roh_start=0.5, th_roh_next=0.3, mpp=1.0, mpp_stereo_angle=10.0,
           final_level_iteration=1)


# ---------------------------------------------------------------- SelectMPs

@pytest.mark.parametrize("level,it,sgm,values,expect", [
    # level 2: the tail must hold more than 60%; half the cells in bin 91
    # (ceil(90.5)), half in bin 51 (ceil(50.5)): 50% at 91, 100% at 51
    (2, 1, -1, [905] * 50 + [505] * 50, 0.51),
    # level 4 iteration 1: more than 5% -> already at bin 91
    (4, 1, -1, [905] * 50 + [505] * 50, 0.91),
    # level 4 iteration 3: more than 7%; 6% in bin 99, 94% in bin 21 -> 0.21
    (4, 3, -1, [990] * 6 + [205] * 94, 0.21),
    # everything in bin 99 -> 0.99, capped at 0.95
    (3, 1, -1, [999] * 10, 0.95),
    # level 1 at an SGM level: fraction 0.95 - 0.1 = 0.85; 80% in bin 70,
    # 20% in bin 40 -> 0.40
    (1, 1, 1, [700] * 80 + [400] * 20, 0.40),
    # level 1 not an SGM level: 0.80 is not exceeded by 80% -> 0.40 too
    (1, 1, -1, [700] * 80 + [400] * 20, 0.40),
    # level 1, 81% in bin 70 -> 0.70
    (1, 1, -1, [700] * 81 + [400] * 19, 0.70),
    # level 0 and negative peaks: 0.2 without a histogram
    (0, 1, -1, [900] * 10, 0.2),
])
def test_select_mps_threshold(level, it, sgm, values, expect):
    n = len(values)
    L = Lv(n, 1, level, sgm_py=sgm)
    ncc = mk_ncc(n, r0=np.array(values), peaks=0)
    ug = np.zeros(n, UGRID_DTYPE)
    _, th = O.select_mps(L, ncc, ug, it, **SEL)
    assert th == pytest.approx(expect, abs=1e-12)


def test_select_mps_threshold_counts_only_cells_with_three_heights():
    # 10 cells in bin 91 with NumOfHeight 3, 90 cells in bin 11 with only 2
    # heights (not counted): at level 2 the tail at 91 is 100% -> 0.91
    L = Lv(100, 1, 2)
    ncc = mk_ncc(100, r0=np.array([905] * 10 + [105] * 90), peaks=0)
    ncc["NumOfHeight"][10:] = 2
    _, th = O.select_mps(L, ncc, np.zeros(100, UGRID_DTYPE), 1, **SEL)
    assert th == pytest.approx(0.91)


def test_select_mps_negative_peaks_count_in_total_only():
    # 40 cells at rho -0.3 (no bin) and 60 at 0.805 (bin 81): at level 2 the
    # tail at 81 is 60% of 100, not > 60% -> no bin qualifies -> 0.2
    L = Lv(100, 1, 2)
    ncc = mk_ncc(100, r0=np.array([-300] * 40 + [805] * 60), peaks=0)
    _, th = O.select_mps(L, ncc, np.zeros(100, UGRID_DTYPE), 1, **SEL)
    assert th == pytest.approx(0.2)


def test_select_mps_undefined_threshold_level_5():
    L = Lv(4, 1, 5)
    with pytest.raises(RuntimeError):
        O.select_mps(L, mk_ncc(4, r0=500), np.zeros(4, UGRID_DTYPE), 1, **SEL)


def one_cell(level, it, r0, r1, h1, h2, peaks=1, roh=0, flag=0, height=0.0, lo=0, hi=0, nh=10, **kw):
    """SelectMPs on cell 0 beside 29 weak cells (rho 0.1, no peak): the
    target is 1/30 of the histogram, under every level's fraction (the
    smallest is level 4's 5%), so the threshold falls to the fillers' bin,
    0.10 (0.2 at level 0)."""
    L = Lv(30, 1, level)
    ncc = mk_ncc(30, r0=100, peaks=0)
    ncc[0] = mk_ncc(1, r0=r0, r1=r1, h1=h1, h2=h2, peaks=peaks, nh=nh)[0]
    ug = np.zeros(30, UGRID_DTYPE)
    ug["roh"][0], ug["Matched_flag"][0], ug["Height"][0] = roh, flag, height
    ug["minHeight"][0], ug["maxHeight"][0] = lo, hi
    args = dict(SEL)
    args.update(kw)
    pts, th = O.select_mps(L, ncc, ug, it, **args)
    assert th == (0.2 if level == 0 else 0.10)
    return pts, th, ug[0]


def test_select_mps_level0_takes_strong_peak():
    # level 0, iteration 3, first iteration of the final level < 3:
    # ROR = (0.8 - 0.5)/0.8 = 0.375 >= 0.01, 0.8 > 0.2, 0.8 > 0 - 0.05, one peak
    pts, th, g = one_cell(0, 3, 800, 500, 101.5, 97.0)
    assert th == 0.2
    assert len(pts) == 1 and pts[0]["z"] == 101.5 and pts[0]["x"] == 0.0 and pts[0]["flag"] == 0
    assert g["roh"] == 800


def test_select_mps_no_peak_resets_roh_and_ends_iterations():
    # level 0, result4 = 0 -> not taken; roh = round(0.5*1000) = 500, and since
    # Th_roh - 0.10 = 0.40 < Th_roh_min 0.45 and Th_roh_start == 0.5 -> Th_roh_next
    pts, _, g = one_cell(0, 3, 800, 500, 101.5, 97.0, peaks=0)
    assert len(pts) == 0 and g["roh"] == 300
    # Th_roh_min low enough: no end of iterations -> 500
    pts, _, g = one_cell(0, 3, 800, 500, 101.5, 97.0, peaks=0, th_roh_min=0.1)
    assert g["roh"] == 500


def test_select_mps_level1_rescue_needs_stereo_angle():
    # level 1, iteration 3: ROR = (0.9 - 0.85)/0.9 = 0.056 < 0.1 fails, but
    # with MPP_stereo_angle > 5 a first peak above the threshold and above the
    # previous rho is taken anyway (index_1); at 3 degrees it is not. A second,
    # weak cell holds the threshold down to 0.10.
    L = Lv(2, 1, 1)
    ncc = mk_ncc(2, r0=np.array([900, 100]), r1=np.array([850, 0]), h1=50.0, h2=40.0)
    for angle, expect in ((10.0, 1), (3.0, 0)):
        ug = np.zeros(2, UGRID_DTYPE)
        args = dict(SEL, mpp_stereo_angle=angle)
        pts, th = O.select_mps(L, ncc, ug, 3, **args)
        # level 1: more than 80% of 2 cells -> both bins -> ceil(10) = bin 10
        assert th == pytest.approx(0.10)
        assert int((pts["x"] == 0.0).sum()) == expect


def test_select_mps_matched_cell_keeps_both_peaks():
    # level 1, matched cell (flag 1): ROR = (0.9-0.85)/0.9 < 0.1 and second peak
    # 0.85 > 0 - 0.05 -> range spans both peaks: result2 100.3 > result3 95.6
    # -> min = floor(95.6 - 0.5) = 95, max = ceil(100.3 + 0.5) = 101, flag 4
    _, _, g = one_cell(1, 3, 900, 850, 100.3, 95.6, flag=1, mpp_stereo_angle=3.0)
    assert (g["minHeight"], g["maxHeight"], g["Matched_flag"]) == (95, 101, 4)
    # the other order: result2 95.6 < result3 100.3 -> floor(95.1), ceil(100.8)
    _, _, g = one_cell(1, 3, 900, 850, 95.6, 100.3, flag=1, mpp_stereo_angle=3.0)
    assert (g["minHeight"], g["maxHeight"], g["Matched_flag"]) == (95, 101, 4)


def test_select_mps_level0_final_iterations_need_no_peak_count():
    # level 0 with final_level_iteration >= 3: index_2 & index_3 only, so a
    # cell without a counted peak (result4 = 0) is still taken
    pts, _, _ = one_cell(0, 3, 800, 500, 101.5, 97.0, peaks=0, final_level_iteration=3)
    assert len(pts) == 1
    pts, _, _ = one_cell(0, 3, 800, 500, 101.5, 97.0, peaks=0, final_level_iteration=2)
    assert len(pts) == 0


def test_select_mps_previous_height_gates_matched_grid():
    # level 0, previously matched grid (flag 2) at height 100, new peak at 103:
    # |100 - 103| = 3 > MPP*2*1 = 2 -> rejected (roh back to Th_roh = 500,
    # no end of iterations with Th_roh_min 0.1); at 101.5 it is kept
    pts, _, g = one_cell(0, 3, 800, 100, 103.0, 0.0, flag=2, height=100.0, lo=90, hi=110, th_roh_min=0.1)
    assert len(pts) == 0 and g["roh"] == 500
    pts, _, g = one_cell(0, 3, 800, 100, 101.5, 0.0, flag=2, height=100.0, lo=90, hi=110, th_roh_min=0.1)
    assert len(pts) == 1 and g["roh"] == 800
    # flag 1 (inside a triangle): 103 is within half the range (10) -> kept
    pts, _, _ = one_cell(0, 3, 800, 100, 103.0, 0.0, flag=1, height=100.0, lo=90, hi=110)
    assert len(pts) == 1


def test_select_mps_height_limits_and_level4_reset():
    # level 3 takes any peak in (-100, 10000); 20000 is not a point but the
    # cell's roh is still the peak's
    pts, _, g = one_cell(3, 3, 800, 100, 20000.0, 0.0)
    assert len(pts) == 0 and g["roh"] == 800
    # level 4 iteration 1 resets Height to -1000 (cells with > 2 heights only)
    _, _, g = one_cell(4, 1, 800, 100, 50.0, 0.0, height=12.5)
    assert g["Height"] == -1000.0
    _, _, g = one_cell(4, 1, 800, 100, 50.0, 0.0, height=12.5, nh=2)
    assert g["Height"] == 12.5 and g["roh"] == 0


def test_select_mps_roh_floor_is_threshold():
    # a taken peak below the level's threshold is stored as the threshold:
    # level 2, iteration 3; cells at 0.3 (x4) and 0.9 (x1): tail > 60% only at
    # bin 30 -> th 0.30; the 0.3 cell (ROR 1 - 0.1/0.3 >= 0.1, 0.30 > 0.30 is
    # false) is not taken; the 0.9 cell is, roh 900
    L = Lv(5, 1, 2)
    ncc = mk_ncc(5, r0=np.array([300, 300, 300, 300, 900]), r1=100, h1=10.0)
    ug = np.zeros(5, UGRID_DTYPE)
    pts, th = O.select_mps(L, ncc, ug, 3, **SEL)
    assert th == pytest.approx(0.30)
    assert len(pts) == 1 and pts[0]["x"] == 4.0
    assert list(ug["roh"]) == [300, 300, 300, 300, 900]


# ----------------------------------------------------------- SetHeightRange

def pts3(*xyz):
    a = np.zeros(len(xyz), POINT3_DTYPE)
    for i, (x, y, z) in enumerate(xyz):
        a[i]["x"], a[i]["y"], a[i]["z"] = x, y, z
    return a


def tris3(*abc):
    t = np.zeros(len(abc), TRI_DTYPE)
    for i, (a, b, c) in enumerate(abc):
        t[i]["a"], t[i]["b"], t[i]["c"] = a, b, c
    return t


def test_set_height_range_planar_triangle_level2():
    # plane z = 10 + x + 2y over the counter-clockwise triangle (0,0) (10,0)
    # (0,10) on a 12x12 grid of 1 m cells; level 2, iteration 1, MPP 1:
    # BufferOfHeight = 1*4*4 = 16, BF = ceil(16) = 16; the slope only weights
    # levels >= 3. Cell (2, 3) is inside: Height = 10 + 2 + 6 = 18; flag 1;
    # Mean_ortho_ncc 0 is not above 500, so at level 2 the range is
    # [floor(min(10, 18) - 16), ceil(max(30, 18) + 16)] = [-6, 46].
    L = Lv(12, 12, 2)
    p = pts3((0, 0, 10), (10, 0, 20), (0, 10, 30))
    ug = np.zeros(144, UGRID_DTYPE)
    lo, hi = O.set_height_range(L, ug, p, tris3((0, 1, 2)), 1, 1.0)
    c = 3 * 12 + 2
    assert ug[c]["Height"] == np.float32(18.0)
    assert (ug[c]["minHeight"], ug[c]["maxHeight"], ug[c]["Matched_flag"]) == (-6, 46, 1)
    # every strictly interior cell has the plane's height
    for r in range(1, 9):
        for col in range(1, 10 - r):
            assert ug[r * 12 + col]["Height"] == np.float32(10 + col + 2 * r)
    # counter-clockwise: points on the edges give a zero cross product, which
    # counts as outside (Sum = 2): the cell (5, 5) on the hypotenuse and the
    # vertices are not claimed by the triangle
    assert ug[5 * 12 + 5]["Height"] == -1000.0 and ug[0]["Height"] == -1000.0
    # but they border claimed cells -> the neighbours' range, flag 1 (m_bHeight 2)
    assert (ug[5 * 12 + 5]["minHeight"], ug[5 * 12 + 5]["maxHeight"], ug[5 * 12 + 5]["Matched_flag"]) == (-6, 46, 1)
    # a far cell (11, 11) has no claimed neighbour: level 2 keeps its range
    # (0, 0), which the returned grid turns into (-1, 1), flag 0
    far = ug[11 * 12 + 11]
    assert (far["minHeight"], far["maxHeight"], far["Matched_flag"], far["Height"]) == (-1, 1, 0, -1000.0)
    # minH_grid / maxH_grid: over every cell before the zero fix-up
    assert (lo, hi) == (-6.0, 46.0)


def test_set_height_range_clockwise_edges_and_vertex_idw():
    # the same triangle listed clockwise: edge points now have all three cross
    # products <= 0 (Sum = 0) and are inside. At level 1 a vertex cell takes
    # the vertex height (IDW distance 0) and Matched_flag 2 with range
    # [floor(z - BF), ceil(z + BF)]; level 1, iteration 1, MPP 1 ->
    # BufferOfHeight = 3, BF = 3.
    L = Lv(12, 12, 1)
    p = pts3((0, 0, 10), (10, 0, 20), (0, 10, 30))
    ug = np.zeros(144, UGRID_DTYPE)
    O.set_height_range(L, ug, p, tris3((0, 2, 1)), 1, 1.0)
    v = ug[0]
    assert (v["Height"], v["minHeight"], v["maxHeight"], v["Matched_flag"]) == (10.0, 7, 13, 2)
    e = ug[5 * 12 + 5]   # on the hypotenuse: z = 10 + 5 + 10 = 25
    assert e["Height"] == 25.0 and e["Matched_flag"] == 1
    # level 1, not matched, Mean_ortho_ncc 0: the range stays [floor(t1 - BF),
    # ceil(t2 + BF)] from the t1 / t2 branch = [floor(10 - 3), ceil(30 + 3)]
    assert (e["minHeight"], e["maxHeight"]) == (7, 33)


def test_set_height_range_first_triangle_claims():
    # two overlapping triangles over the same cells, different planes: the
    # first in the list claims every shared cell (m_bHeight), so cell (2, 2)
    # gets the first plane (z = 5) whichever order the second would give
    L = Lv(8, 8, 2)
    p = pts3((0, 0, 5), (0, 7, 5), (7, 0, 5), (0, 0, 50), (0, 7, 50), (7, 0, 50))
    for first, z in ((0, 5.0), (1, 50.0)):
        order = [(0, 1, 2), (3, 4, 5)] if first == 0 else [(3, 4, 5), (0, 1, 2)]
        ug = np.zeros(64, UGRID_DTYPE)
        O.set_height_range(L, ug, p, tris3(*order), 1, 1.0)
        assert ug[2 * 8 + 2]["Height"] == z


def test_set_height_range_skips_bad_indices_and_gap_limit():
    # a triangle indexing past the points is skipped; level 0, iteration 2,
    # DEM resolution < 2: th_HG = 50, and a steep triangle (z 0..200) gives
    # ranges wider than 50 -> (-100, -100)
    L = Lv(12, 12, 0, dem_resolution=1.0)
    p = pts3((0, 0, 0), (0, 10, 200), (10, 0, 0))
    ug = np.zeros(144, UGRID_DTYPE)
    O.set_height_range(L, ug, p, tris3((0, 1, 7), (0, 1, 2)), 2, 1.0)
    c = ug[2 * 12 + 2]
    assert c["Height"] == np.float32(40.0)
    assert (c["minHeight"], c["maxHeight"]) == (-100, -100)


def test_set_height_range_level0_far_cells_take_tin_span():
    # level 0, iteration 1, MPP 0.2: BufferOfHeight = max(0.4, 0.5) -> ceil 1.
    # A far cell with no claimed neighbour takes the TIN's height span: the
    # span/3 = (30 - 10)/3 = 6.67 > 1, so BF = 1, raised to 2^0 * 0.5 = 0.5? no
    # (1 >= 0.5): its range (-500, 500) is clamped to [floor(10 - 1), ceil(30 + 1)]
    L = Lv(20, 20, 0, resolution=0.5)
    p = pts3((0, 0, 10), (6, 0, 20), (0, 6, 30))
    ug = np.zeros(400, UGRID_DTYPE)
    ug["minHeight"], ug["maxHeight"] = -500, 500
    O.set_height_range(L, ug, p, tris3((0, 1, 2)), 1, 0.2)
    far = ug[19 * 20 + 19]
    assert (far["minHeight"], far["maxHeight"], far["Matched_flag"]) == (9, 31, 0)


def test_set_height_range_keeps_fields():
    # roh, Mean_ortho_ncc and ortho_ncc[1] are carried, ortho_ncc[0] and
    # anchor_flag cleared; a claimed cell whose Mean_ortho_ncc is above the
    # level's threshold (level 2 it 1: 0.5 -> 500) takes [floor(z - BF), ceil(z + BF)]
    L = Lv(12, 12, 2)
    p = pts3((0, 0, 10), (10, 0, 20), (0, 10, 30))
    ug = np.zeros(144, UGRID_DTYPE)
    ug["roh"], ug["Mean_ortho_ncc"], ug["anchor_flag"] = 321, 700, 1
    ug["ortho_ncc"][:, 0], ug["ortho_ncc"][:, 1] = 11, 22
    O.set_height_range(L, ug, p, tris3((0, 1, 2)), 1, 1.0)
    c = ug[3 * 12 + 2]
    assert (c["minHeight"], c["maxHeight"]) == (18 - 16, 18 + 16)
    assert (c["roh"], c["Mean_ortho_ncc"], c["anchor_flag"]) == (321, 700, 0)
    assert tuple(c["ortho_ncc"]) == (0, 22)


def test_tin_slope_weight_at_level3():
    # level 3, iteration 1, MPP 1: BufferOfHeight = 1*4*8 = 32. A flat
    # triangle has slope 0 -> weight 1 -> BF = 32; the plane z = x (45 degrees)
    # -> weight 1.3 -> BF = ceil(41.6) = 42. Cell (2, 2) inside both.
    L = Lv(12, 12, 3)
    ug = np.zeros(144, UGRID_DTYPE)
    O.set_height_range(L, ug, pts3((0, 0, 5), (10, 0, 5), (0, 10, 5)), tris3((0, 1, 2)), 1, 1.0)
    assert (ug[26]["minHeight"], ug[26]["maxHeight"]) == (5 - 32, 5 + 32)
    ug = np.zeros(144, UGRID_DTYPE)
    O.set_height_range(L, ug, pts3((0, 0, 0), (10, 0, 10), (0, 10, 0)), tris3((0, 1, 2)), 1, 1.0)
    # Height 2; t1 = min(0, 2) = 0, t2 = max(10, 2) = 10 -> [-42, 52]
    assert ug[26]["Height"] == 2.0
    assert (ug[26]["minHeight"], ug[26]["maxHeight"]) == (-42, 52)


# ------------------------------------------------------------ reference TIN

needs_tin = pytest.mark.skipif(O.tin() is None, reason="oracle/_ref/libref_tin.so not built")


@needs_tin
def test_reference_tin_is_delaunay_and_deterministic():
    rng = np.random.default_rng(5)
    W = 120
    sel = rng.random((W, W)) < 0.25
    rr, cc = np.nonzero(sel)
    p = np.zeros(len(rr), POINT3_DTYPE)
    p["x"], p["y"], p["z"] = 10 + cc * 2.0, 20 + rr * 2.0, rng.random(len(rr))
    b = [10, 20, 10 + (W - 1) * 2.0, 20 + (W - 1) * 2.0]
    t1 = O.tin_create(p, b, 2.0)
    t2 = O.tin_create(p, b, 2.0)
    assert np.array_equal(t1, t2)
    # Euler: a triangulation of n points with h on the hull has 2n - 2 - h triangles
    assert len(t1) <= 2 * len(p) - 5
    # empty circumcircles (integer grid coordinates, exact in float64)
    X, Y = cc.astype(np.float64), rr.astype(np.float64)
    from scipy.spatial import cKDTree
    kd = cKDTree(np.c_[X, Y])
    for t in t1[::7]:
        a, bb, c = int(t["a"]), int(t["b"]), int(t["c"])
        ax, ay, bx, by, cx, cy = X[a], Y[a], X[bb], Y[bb], X[c], Y[c]
        d = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
        ux = ((ax * ax + ay * ay) * (by - cy) + (bx * bx + by * by) * (cy - ay) + (cx * cx + cy * cy) * (ay - by)) / d
        uy = ((ax * ax + ay * ay) * (cx - bx) + (bx * bx + by * by) * (ax - cx) + (cx * cx + cy * cy) * (bx - ax)) / d
        r = math.hypot(ax - ux, ay - uy)
        inside = [q for q in kd.query_ball_point([ux, uy], r - 1e-7) if q not in (a, bb, c)]
        assert not inside


@needs_tin
def test_reference_tin_small_cases():
    # fewer than three points: no triangles; a unit square: two triangles
    assert len(O.tin_create(pts3((0, 0, 0), (1, 0, 0)), [0, 0, 1, 1], 1.0)) == 0
    t = O.tin_create(pts3((0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)), [0, 0, 1, 1], 1.0)
    assert len(t) == 2
    used = sorted(set(int(v) for tri in t for v in (tri["a"], tri["b"], tri["c"])))
    assert used == [0, 1, 2, 3]
