"""Independent pins of the oracle's branchy functions.

Every expected value here was worked out by hand from the reference text
(the setsm_code.cpp lines named per table) and is written down as a literal;
none is computed by restating the oracle's own formula. A dropped term, a
swapped comparison (< for <=, > for >=) or a wrong margin in the oracle fails
one of these rows.

  * SetNCC_alpha          setsm_code.cpp:6014-6047   -> test_ncc_alpha_table
  * check_image_boundary  setsm_code.cpp:11002-11083 -> test_check_image_boundary_closed_form
  * VLL blunder / range gating  setsm_code.cpp:6160-6204, 6308-6345
                                                     -> test_vll_gate_table
  * InitializeVoxel       setsm_code.cpp:5759-6012   -> test_initialize_voxel_lattice
  * FindPeakNcc           setsm_code.cpp:6672-6752   -> test_find_peak_hand_sequences
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_py as O
from setsm_amd.records import NCCRESULT_DTYPE, UGRID_DTYPE, LevelInputs
from tests.helpers import level_grid, scene

FRESH = 1 | (np.uint32(np.uint16(np.int16(-1000))) << 16)   # flag_cal 1, INCC -1000 (a new row)


def word(flag, incc):
    return np.uint32(flag) | (np.uint32(np.uint16(np.int16(incc))) << 16)


# ------------------------------------------------------------ SetNCC_alpha
# Hand-evaluated setsm_code.cpp:6014-6047 for levels 0..5 and iterations 1..6.
ALPHA = {
    5: [1.2, 1.15, 1.1, 1.05, 1.0, 0.95],   # (4-5)*0.2 = -0.2: alpha above 1, no upper clamp
    4: [1.0, 0.95, 0.9, 0.85, 0.8, 0.8],    # clamped at 0.8
    3: [0.8, 0.75, 0.7, 0.65, 0.6, 0.6],    # clamped at 0.6
    2: [0.6, 0.55, 0.5, 0.45, 0.4, 0.4],    # clamped at 0.4
    1: [0.4, 0.35, 0.3, 0.25, 0.2, 0.15],   # else branch, no clamp reached
    0: [0.2, 0.15, 0.1, 0.1, 0.1, 0.1],     # 0.05, 0.0, -0.05 clamped to 0.1
}


def test_ncc_alpha_table():
    L = O.lib()
    for level, row in ALPHA.items():
        for it, want in enumerate(row, start=1):
            assert L.orc_ncc_alpha(level, it, 0) == pytest.approx(want, abs=1e-12), (level, it)
            assert L.orc_ncc_alpha(level, it, 1) == 1.0   # IsRA


# ---------------------------------------------------- check_image_boundary
def _affine_rpc(samp_off, c):
    """RPC00B with every term zero but: line numerator P + c*H, sample
    numerator L, denominators 1; offsets 0 (sample: samp_off), scales 1024.
    With lon, lat, h dyadic, every step of CoordConversion.hpp:26-87 is exact:
    line = lat + c*h, sample = lon + samp_off (both images' clamps at 1228.8)."""
    r = np.zeros(90)
    r[0:5] = [0.0, samp_off, 0.0, 0.0, 0.0]
    r[5:10] = [1024.0, 1024.0, 1024.0, 1024.0, 1024.0]
    coef = r[10:].reshape(4, 20)
    coef[0, 2] = 1.0      # line num: P
    coef[0, 3] = c        #           + c * H
    coef[1, 0] = 1.0      # line den
    coef[2, 1] = 1.0      # samp num: L
    coef[3, 0] = 1.0      # samp den
    return r


def _boundary_level(level, sizes, starts, samp_offs, cells):
    """A level whose pyramid images have the given (w, h) per image and
    level-0 start offsets; cells = [(lon, lat)]."""
    pyr = []
    for ti in range(2):
        w, h = sizes[ti]
        lv = [(np.zeros((4, 4), np.uint16), np.zeros((4, 4), np.uint16), np.zeros((4, 4), np.uint8))] * level
        lv.append((np.zeros((h, w), np.uint16), np.zeros((h, w), np.uint16), np.zeros((h, w), np.uint8)))
        pyr.append(lv)
    rpcs = [_affine_rpc(samp_offs[0], 1.0 / 16), _affine_rpc(samp_offs[1], -1.0 / 16)]
    n = len(cells)
    wgs = np.array(cells, np.float64)
    grid = dict(grid_pts=np.zeros((n, 2)), grid_wgs=wgs, boundary=np.array([0.0, 0.0, 1.0, 1.0]), shape=(1, n))
    return LevelInputs(pyr, level, rpcs, grid, start=tuple(starts))


def test_check_image_boundary_closed_form():
    """Image 0: x = lon, y = lat + h/16; image 1: x = lon + off1, y = lat - h/16,
    then x / 2^level - start. A cell passes when, for both images at both
    band ends, H+1 < x < w-H-1 and H+1 < y < h-H-1 (strict, buff_pixel 1)."""
    f = O.lib().orc_check_image_boundary
    # Geometry A, level 0, H = 5 (margin 6): image 0 is 64 x 48, image 1 128 x 96
    A = [  # (lon, lat, minH, maxH, expected, what)
        (30.0, 24.0, 0.0, 16.0, 1, "interior"),
        (6.0, 24.0, 0.0, 16.0, 0, "image 0 x exactly at the left margin"),
        (6.25, 24.0, 0.0, 16.0, 1, "image 0 x a quarter pixel inside"),
        (58.0, 24.0, 0.0, 16.0, 0, "image 0 x exactly at w-H-1"),
        (57.75, 24.0, 0.0, 16.0, 1, "image 0 x a quarter pixel inside the right margin"),
        (30.0, 6.0, 0.0, 16.0, 0, "image 0 y at the top margin at minH"),
        (30.0, 6.25, 0.0, 16.0, 0, "image 1 y = 5.25 at maxH (target fails at maxH only)"),
        (30.0, 8.0, 0.0, 16.0, 1, "both images inside at both heights"),
        (30.0, 41.0, 0.0, 16.0, 0, "image 0 y = 42 = h-H-1 at maxH (reference fails at maxH)"),
        (30.0, 41.0, 0.0, 15.0, 1, "image 0 y = 41.9375 at maxH"),
        (30.0, 5.5, 0.0, 16.0, 0, "image 0 y = 5.5 at minH, 6.5 at maxH"),
        (30.0, 0.5, 0.0, 16.0, 0, "image 1 y negative at maxH, clamped to 0"),
    ]
    li = _boundary_level(0, [(64, 48), (128, 96)], [(0.0, 0.0), (0.0, 0.0)], [0.0, 8.0],
                         [(c[0], c[1]) for c in A])
    for k, (lon, lat, lo, hi, want, what) in enumerate(A):
        assert f(C.addressof(li.level), k, lo, hi, 5) == want, what
    # the margin follows the H_template_size argument: 7 (InitializeVoxel's) moves it to 8
    assert f(C.addressof(li.level), 2, 0.0, 16.0, 7) == 0   # x = 6.25 < 8
    # Geometry B, level 0: image 1 (24 x 48) starts at x = 48, so x1 = lon + 8 - 48
    B = [(46.0, 24.0, 0.0, 16.0, 0, "image 1 x = 6 at its left margin"),
         (46.25, 24.0, 0.0, 16.0, 1, "image 1 x = 6.25"),
         (57.75, 24.0, 0.0, 16.0, 1, "image 0 x = 57.75, image 1 x = 17.75 < 18"),
         (58.0, 24.0, 0.0, 16.0, 0, "both at their right margins")]
    li = _boundary_level(0, [(64, 48), (24, 48)], [(0.0, 0.0), (48.0, 0.0)], [0.0, 8.0],
                         [(c[0], c[1]) for c in B])
    for k, (lon, lat, lo, hi, want, what) in enumerate(B):
        assert f(C.addressof(li.level), k, lo, hi, 5) == want, what
    # Geometry C, level 2, H = 6 (margin 7): image 0 32 x 32 starting at level-0
    # (16, 8) -> pyramid start (4, 2): x0 = lon/4 - 4, y0 = (lat + h/16)/4 - 2
    Cg = [(44.0, 64.0, 0.0, 16.0, 0, "x0 = 7 exactly"),
          (45.0, 64.0, 0.0, 16.0, 1, "x0 = 7.25"),
          (116.0, 64.0, 0.0, 16.0, 0, "x0 = 25 = w-H-1"),
          (115.0, 64.0, 0.0, 16.0, 1, "x0 = 24.75")]
    li = _boundary_level(2, [(32, 32), (64, 64)], [(16.0, 8.0), (0.0, 0.0)], [0.0, 8.0],
                         [(c[0], c[1]) for c in Cg])
    for k, (lon, lat, lo, hi, want, what) in enumerate(Cg):
        assert f(C.addressof(li.level), k, lo, hi, 6) == want, what


# ------------------------------------------------------- one-cell drivers
@pytest.fixture(scope="module")
def cell_scene():
    # 3 degree views: 1300 m of band moves a point ~70 px at level 1, so every
    # band below stays inside the images
    return scene(image_px=512, relief=60.0, seed=13, view=3.0, levels=3)


def _cell_state(cell_scene, level, dem=2.0, matchtag=0, edge=False):
    sc, pyr = cell_scene
    g = level_grid(sc, level, 1)
    if edge:   # a cell at the image corner: the window cannot fit
        g = level_grid(sc, level, 1, inset=1.0)
    li = LevelInputs(pyr, level, sc.rpcs, g, grid_resolution=g["grid_res"], dem_resolution=dem,
                     check_matchtag=matchtag)
    return O.OracleState(li), g


def _ug(lo, hi, flag=1):
    u = np.zeros(1, UGRID_DTYPE)
    u["minHeight"], u["maxHeight"], u["Matched_flag"] = lo, hi, flag
    u["Height"] = 0.5 * (lo + hi)
    return u


def _ncc(lo, hi, num, chc, gncc=77):
    n = np.zeros(1, NCCRESULT_DTYPE)
    n["minHeight"], n["maxHeight"], n["NumOfHeight"], n["check_height_change"], n["GNCC"] = lo, hi, num, chc, gncc
    return n


WIDE = (-500.0, 2000.0)


# ----------------------------------------------------- VLL gating table
# (level, dem, matchtag, iteration, G (min, max, flag), minmax, ncc (min, chc),
#  expected accessible, expected evaluated-mask of the 8 heights, what)
GATE = [
    (2, 2.0, 0, 1, (220, 240, 0), WIDE, (220, 1), 1, [1] * 8, "L>=2 ignores Matched_flag"),
    (1, 2.0, 0, 1, (220, 240, 0), WIDE, (220, 1), 0, [0] * 8, "L1, Matched_flag 0: blunder"),
    (0, 2.0, 0, 1, (220, 240, 1), WIDE, (220, 1), 1, [1] * 8, "L0 band 20 m"),
    (0, 2.0, 0, 1, (-100, 400, 1), WIDE, (-100, 1), 0, [0] * 8, "L0 band exactly th_height 500: not < 500"),
    (0, 2.0, 0, 1, (-99, 400, 1), WIDE, (-99, 1), 1, [1] * 8, "L0 band 499"),
    (0, 5.0, 0, 1, (-100, 500, 1), WIDE, (-100, 1), 1, [1] * 8, "DEM 5 m: th_height 1000, band 600"),
    (1, 2.0, 0, 1, (220, 240, 1), (-500.0, 230.0), (220, 1), 0, [0] * 8, "L1 maxH above minmax[1]"),
    (1, 2.0, 0, 1, (220, 240, 1), (225.0, 2000.0), (220, 1), 0, [0] * 8, "L1 minH below minmax[0]"),
    (2, 2.0, 0, 1, (220, 240, 1), (-500.0, 230.0), (220, 1), 1, [1] * 8, "L2: no minmax test"),
    (0, 2.0, 1, 2, (220, 240, 0), WIDE, (220, 1), 1, [1] * 8, "matchtag, iteration 2: flag reset"),
    (0, 2.0, 1, 3, (220, 240, 0), WIDE, (220, 1), 0, [0] * 8, "matchtag, iteration 3: flag 0 stays a blunder"),
    (2, 2.0, 1, 1, (220, 240, 1), (-500.0, 230.0), (220, 1), 0, [0] * 8, "matchtag: minmax at any level"),
    (0, 2.0, 1, 2, (-100, 400, 1), WIDE, (-100, 1), 1, [1] * 8, "matchtag: band 500 is not > 500"),
    (0, 2.0, 1, 2, (-101, 400, 1), WIDE, (-101, 1), 0, [0] * 8, "matchtag: band 501"),
    (0, 2.0, 0, 1, (220, 240, 1), WIDE, (220, 0), 1, [0] * 8, "no height change, DEM 2: accessible, not evaluated"),
    (0, 1.0, 0, 1, (220, 240, 1), WIDE, (220, 0), 1, [1] * 8, "no height change, L0 DEM 1 < 2: evaluated"),
    (1, 1.0, 0, 1, (220, 240, 1), WIDE, (220, 0), 1, [0] * 8, "no height change, L1 DEM 1: not evaluated"),
    (0, 2.0, 0, 1, (220, 240, 1), WIDE, (219, 1), 1, [0, 0, 0, 0, 1, 1, 1, 1],
     "heights below minHeight (219..219.75) skipped"),
]


@pytest.mark.parametrize("case", GATE, ids=[c[-1] for c in GATE])
def test_vll_gate_table(cell_scene, case):
    level, dem, mt, it, G, minmax, (nmin, chc), want_acc, want_eval, _ = case
    st, g = _cell_state(cell_scene, level, dem, mt)
    ug = _ug(*G)
    st.import_voxels(_ncc(nmin, G[1], 8, chc), np.full(8, word(1, 5000), np.uint32))
    acc = st.vertical_line_locus(ug, it, np.array(minmax))
    words, _ = st.export()
    evaluated = ((words >> 16).astype(np.uint16).view(np.int16) != 5000).astype(int)
    assert acc == want_acc
    assert list(evaluated) == want_eval
    # an accessible cell gets GNCC = -1 (the term is off here); a gated one keeps its value
    assert int(st.ncc["GNCC"][0]) == (-1000 if want_acc else 77)


def test_vll_gate_boundary(cell_scene):
    st, g = _cell_state(cell_scene, 0, edge=True)
    st.import_voxels(_ncc(220, 240, 8, 1), np.full(8, word(1, 5000), np.uint32))
    assert st.vertical_line_locus(_ug(220, 240), 1, np.array(WIDE)) == 0
    assert int(st.ncc["GNCC"][0]) == 77


# ------------------------------------------------ InitializeVoxel lattice
# (level, iteration, dem, matchtag, G (min, max, flag), previous (min, max, num, chc) or None, minmax,
#  expected ncc (min, max, num, chc, GNCC), expected G after (min, max), row: fresh / kept / none, what)
# Height steps: 0.25 m at level 0, 0.5 m at level 1, 1 m at level 2.
INIT = [
    (0, 1, 2.0, 0, (200, 232, 1), None, WIDE, (200, 232, 128, 1, -1000), (200, 232), "fresh", "iteration 1"),
    (0, 1, 2.0, 0, (200, 801, 1), None, WIDE, (200, 801, 0, 0, -1000), (-100, -100), "none",
     "L0 iteration 1, band 601 > th_height 500: search space rewritten to -100"),
    (1, 1, 2.0, 0, (200, 1201, 1), None, WIDE, (200, 1201, 0, 0, -1000), (-100, -100), "none",
     "L1 iteration 1, band 1001 > 1000"),
    (1, 1, 2.0, 0, (200, 1200, 1), None, WIDE, (200, 1200, 2000, 1, -1000), (200, 1200), "fresh",
     "L1 iteration 1, band exactly 1000 kept (2000 heights: no cap above L0)"),
    (0, 1, 2.0, 0, (0, 32, 1), None, WIDE, (0, 32, 0, 0, -1000), (0, 32), "none", "an endpoint at 0"),
    (0, 1, 2.0, 0, (200, 200, 1), None, WIDE, (200, 200, 0, 0, -1000), (200, 200), "none", "empty band"),
    (0, 1, 2.0, 0, (200, 451, 1), None, WIDE, (200, 451, 0, 0, -1000), (200, 451), "none",
     "1004 heights > 1000 at L0"),
    (0, 1, 2.0, 0, (200, 450, 1), None, WIDE, (200, 450, 1000, 1, -1000), (200, 450), "fresh",
     "exactly 1000 heights at L0"),
    (0, 1, 5.0, 0, (200, 801, 1), None, WIDE, (200, 801, 0, 0, -1000), (200, 801), "none",
     "DEM 5 m: th_height 1000 keeps the band, 2404 heights > 1000 drop it"),
    (0, 2, 2.0, 0, (200, 232, 0), (200, 232, 128, 1), WIDE, (200, 232, 0, 0, 77), (200, 232), "none",
     "iteration 2, Matched_flag 0"),
    (0, 2, 2.0, 0, (200, 200, 1), (200, 232, 128, 1), WIDE, (200, 232, 0, 0, 77), (200, 200), "none",
     "iteration 2, empty band"),
    (0, 2, 2.0, 0, (200, 232, 1), (200, 232, 128, 1), WIDE, (200, 232, 128, 0, 77), (200, 232), "kept",
     "iteration 2, unchanged band"),
    (0, 2, 2.0, 0, (200, 232, 1), (204, 228, 96, 1), WIDE, (196, 236, 160, 1, 77), (200, 232), "fresh",
     "extension by 16 steps each side, applied twice (setsm_code.cpp:5912-5913 and 5925-5926)"),
    (0, 2, 2.0, 0, (210, 230, 1), (200, 232, 0, 0), WIDE, (200, 232, 128, 1, 77), (210, 230), "fresh",
     "new search height: the previous NCCresult bounds are used"),
    (1, 2, 2.0, 0, (200, 1300, 1), (200, 232, 64, 1), WIDE, (200, 232, 0, 0, 77), (200, 1300), "none",
     "L1 iteration 2, band 1100 >= 1000"),
    (1, 2, 2.0, 0, (200, 240, 1), (200, 240, 80, 1), (-500.0, 220.0), (200, 240, 80, 0, 77), (200, 240), "kept",
     "L1: the band test overrides the minmax test"),
    (0, 2, 2.0, 0, (200, 232, 1), (200, 232, 128, 1), (-500.0, 220.0), (200, 232, 128, 0, 77), (200, 232),
     "kept", "L0: the band test overrides the minmax test"),
    (0, 2, 2.0, 0, (200, 232, 1), (100, 700, 0, 0), WIDE, (100, 700, 0, 0, 77), (200, 232), "none",
     "L0 iteration 2, previous band 600 >= 500"),
    (0, 3, 2.0, 1, (200, 232, 1), (200, 232, 128, 1), WIDE, (200, 232, 128, 0, -1000), (200, 232), "kept",
     "matchtag at iteration 3: bounds reset, GNCC -1, then no extension"),
    (0, 2, 1.0, 0, (200, 232, 1), (200, 232, 128, 1), WIDE, (200, 232, 128, 1, -1000), (200, 232), "fresh",
     "L0 with DEM 1 m < 2: rebuilt every iteration"),
    (2, 2, 2.0, 0, (200, 232, 0), (200, 232, 32, 1), WIDE, (200, 232, 32, 0, 77), (200, 232), "kept",
     "L2 ignores Matched_flag"),
]


@pytest.mark.parametrize("case", INIT, ids=[c[-1] for c in INIT])
def test_initialize_voxel_lattice(cell_scene, case):
    level, it, dem, mt, G, prev, minmax, want, want_g, row, _ = case
    st, g = _cell_state(cell_scene, level, dem, mt)
    prev_words = None
    if prev is not None:
        prev_words = (np.arange(prev[2], dtype=np.uint32) << 16) | np.uint32(1)
        st.import_voxels(_ncc(*prev), prev_words)
    ug = _ug(*G)
    st.initialize_voxel(ug, it, np.array(minmax))
    nr = st.ncc[0]
    got = (int(nr["minHeight"]), int(nr["maxHeight"]), int(nr["NumOfHeight"]), int(nr["check_height_change"]),
           int(nr["GNCC"]))
    assert got == want
    assert (int(ug["minHeight"][0]), int(ug["maxHeight"][0])) == want_g
    words, _ = st.export()
    if row == "fresh":
        assert len(words) == want[2] and np.all(words == FRESH)
    elif row == "kept":
        assert np.array_equal(words, prev_words)
    else:
        assert len(words) == 0


def test_initialize_voxel_outside_images(cell_scene):
    st, g = _cell_state(cell_scene, 0, edge=True)
    st.import_voxels(_ncc(200, 232, 128, 1), np.ones(128, np.uint32))
    ug = _ug(200, 232)
    st.initialize_voxel(ug, 2, np.array(WIDE))
    assert int(st.ncc["NumOfHeight"][0]) == 0 and int(st.ncc["check_height_change"][0]) == 0


# ------------------------------------------------------------ FindPeakNcc
class PeakState(C.Structure):
    _fields_ = [("check_rho", C.c_int), ("pre_rho", C.c_double), ("pre_height", C.c_float),
                ("direction", C.c_int), ("max_wncc", C.c_double)]


def _peaks(rhos, heights=None):
    L = O.lib()
    L.orc_peak_state_init.argtypes = [C.c_void_p]
    L.orc_find_peak_ncc.argtypes = [C.c_double, C.c_float, C.c_void_p, C.c_void_p]
    st = PeakState()
    L.orc_peak_state_init(C.byref(st))
    nr = np.zeros(1, NCCRESULT_DTYPE)
    # the values VerticalLineLocus sets before its height loop (setsm_code.cpp:6348-6352)
    nr["result0"], nr["result1"], nr["result2"], nr["result3"], nr["result4"] = -1000, -1000, -1000, -1000, 0
    heights = heights if heights is not None else list(range(len(rhos)))
    for r, h in zip(rhos, heights):
        L.orc_find_peak_ncc(float(r), float(h), C.byref(st), nr.ctypes.data)
    n = nr[0]
    return (int(n["result0"]), int(n["result1"]), float(n["result2"]), float(n["result3"]), int(n["result4"]),
            round(st.max_wncc, 9))


# (rho sequence, heights, expected (result0, result1, result2, result3, result4, max_WNCC), what)
PEAKS = [
    ([0.1, 0.3, 0.5, 0.4, 0.2], [10, 11, 12, 13, 14], (500, -1000, 12.0, -1000.0, 1, 0.5), "one peak"),
    ([0.2, 0.6, 0.3, 0.7, 0.1], None, (700, 600, 3.0, 1.0, 2, 0.7), "second peak higher: first becomes second"),
    ([0.2, 0.7, 0.3, 0.6, 0.1], None, (700, 600, 1.0, 3.0, 2, 0.7), "second peak lower"),
    ([0.1, 0.5, 0.5, 0.2], None, (-1000, -1000, -1000.0, -1000.0, 0, 0.5), "a plateau is not a peak"),
    ([0.1, 0.2, 0.3], None, (-1000, -1000, -1000.0, -1000.0, 0, 0.3), "monotone: no peak"),
    ([-0.5, -0.2, -0.4], None, (-200, -1000, 1.0, -1000.0, 1, -0.2), "all negative: peak at -0.2"),
    ([-1.5, -1.0, -1.2], None, (-1000, -1000, -1000.0, -1000.0, 0, -1.0), "pre_rho == -1 never marks a peak"),
    ([0.0, 1.0] * 12 + [0.0], None, (1000, 1000, 1.0, 3.0, 10, 1.0), "peak count saturates at 10"),
    ([0.1, 0.4994, 0.2, 0.4996, 0.1], None, (500, 499, 3.0, 1.0, 2, 0.4996), "result rounding: 0.4996 -> 500"),
    ([0.1, 0.6, 0.2, 0.6004, 0.1], None, (600, 600, 1.0, 3.0, 2, 0.6004), "equal quantised peaks: first keeps result2"),
]


@pytest.mark.parametrize("case", PEAKS, ids=[c[-1] for c in PEAKS])
def test_find_peak_hand_sequences(case):
    rhos, heights, want, _ = case
    assert _peaks(rhos, heights) == want
