"""Pins of the oracle's blunder_detection_TIN (CPU).

setsm_code.cpp does not build here, so each expectation is worked out by hand
from the rules of setsm_code.cpp:9158-9782 (with SetNormalAngle,
SubFunctions.cpp:2997-3022) on inputs small enough to follow: the slope of a
symmetric spike from its closed form, the height threshold from numpy's mean
and population deviation of the edge height differences or from counting
bins, and which point a steep triangle blames from the ordering of its
heights.
"""
import math

import numpy as np
import pytest

from oracle import oracle_py as O
from setsm_amd.records import POINT3_DTYPE, TRI_DTYPE, UGRID_DTYPE
from tests.test_oracle_mps import Lv

W = 40


def pts_of(rows):
    """(col, row, z[, flag]) on a 1 m grid at the origin."""
    p = np.zeros(len(rows), POINT3_DTYPE)
    for i, r in enumerate(rows):
        p[i]["x"], p[i]["y"], p[i]["z"] = r[0], r[1], r[2]
        p[i]["flag"] = r[3] if len(r) > 3 else 0
    return p


def tris_of(rows):
    t = np.zeros(len(rows), TRI_DTYPE)
    for i, (a, b, c) in enumerate(rows):
        t[i]["a"], t[i]["b"], t[i]["c"] = a, b, c
    return t


def cell(p):
    return int(round(p["y"])) * W + int(round(p["x"]))


def run(level, it, pts, tris, oncc=0.9, flag_blunder=True, hinterval=100.0, is_ra=0, count_bl=0, ug=None,
        gres=1.0, **kw):
    L = Lv(W, W, level, gres=gres, is_ra=is_ra)
    ug = np.zeros(W * W, UGRID_DTYPE) if ug is None else ug
    on = np.full(W * W, 0.9, np.float32)
    if isinstance(oncc, dict):
        for i, v in oncc.items():
            on[cell(pts[i])] = v
    else:
        on[:] = oncc
    det, bc, lo, hi, th = O.blunder_detection_tin(L, ug, on, pts, tris, it, flag_blunder, hinterval,
                                                  count_bl=count_bl, **kw)
    return det, bc, lo, hi, th, ug


# a far, flat triangle at list index 0: the reference skips triangle 0 in the
# per-point pass (savenode > 0), so the spikes below never use index 0
DUMMY = [(30, 30, 0.0), (31, 30, 0.0), (30, 31, 0.0)]


def spike(h, centre_flag=0):
    """A centre at height h over its four axis neighbours at 0, in four
    triangles (after the dummy at index 0)."""
    p = pts_of(DUMMY + [(5, 5, h, centre_flag), (6, 5, 0.0), (5, 6, 0.0), (4, 5, 0.0), (5, 4, 0.0)])
    t = tris_of([(0, 1, 2), (3, 4, 5), (3, 5, 6), (3, 6, 7), (3, 7, 4)])
    return p, t


def test_slope_threshold_closed_form():
    # the four faces have normal (h, h, 1): slope acos(1 / sqrt(2 h^2 + 1)),
    # above 30 degrees iff h > sqrt(1/6) = 0.4082. Level 2, iteration 3: the
    # count rule needs ortho NCC < 0.6 - (4 - 2) * 0.1 = 0.4, and the
    # neighbour rule (ortho NCC < 0.4 - 2 * 0.01 = 0.38) stays off at 0.39.
    for h, flagged in ((0.40, False), (0.42, True), (5.0, True)):
        slope = math.degrees(math.acos(1 / math.sqrt(2 * h * h + 1)))
        assert (slope > 30) == flagged
        p, t = spike(h)
        det, bc, *_ = run(2, 3, p, t, oncc={3: 0.39})
        assert list(np.nonzero(det)[0]) == ([3] if flagged else [])
        assert p[3]["flag"] == (1 if flagged else 0)
        assert bc == ((7, 1) if flagged else (8, 0))


def test_count_rule_ortho_threshold_per_level():
    # 0.6 - (4 - level) * 0.1 below level 1's fixed 0.6; level 2: 0.4
    for level, it, on, flagged in ((2, 3, 0.39, True), (2, 3, 0.41, False), (3, 1, 0.49, True),
                                   (3, 1, 0.51, False), (1, 1, 0.59, True), (1, 1, 0.61, False)):
        p, t = spike(5.0)
        # keep the neighbour rule out: ortho NCC thresholds 0.38 (level 2, it 3),
        # 0.5 (level 3, it 1) and 0.3 (level 1) are all below the centre's values
        # except level 3 at 0.49, where both rules blame the centre anyway
        det, *_ = run(level, it, p, t, oncc={3: on})
        assert bool(det[3]) == flagged, (level, it, on)
        assert det.sum() == int(flagged)


def test_triangle_zero_is_skipped():
    # a centre with two steep triangles: both counted -> count 2 > 1 and
    # blunder; with one of them at index 0 only one counts and count > 1 fails
    p = pts_of([(5, 5, 5.0), (6, 5, 0.0), (5, 6, 0.0), (4, 5, 0.0)] + DUMMY)
    ok = tris_of([(4, 5, 6), (0, 1, 2), (0, 2, 3)])
    det, *_ = run(2, 3, p.copy(), ok, oncc={0: 0.39})
    assert det[0]
    first = tris_of([(0, 1, 2), (0, 2, 3), (4, 5, 6)])
    det, *_ = run(2, 3, p.copy(), first, oncc={0: 0.39})
    assert not det[0]


def fan(steep_first):
    """A centre at 1000 m with 40 triangles of their own rim pairs: 24 steep
    (rims at 0) and 16 flat (rims at 1000). steep_first=False puts 14 steep +
    16 flat first and the other 10 steep last."""
    rows = DUMMY + [(20, 20, 1000.0)]
    kinds = (["s"] * 24 + ["f"] * 16) if steep_first else (["s"] * 14 + ["f"] * 16 + ["s"] * 10)
    tris = [(0, 1, 2)]
    for k, kind in enumerate(kinds):
        # rim pair k around (20, 20): (dx, dy) and its quarter turn (-dy, dx),
        # never collinear with the centre
        ang = 2 * math.pi * k / 40
        d = 3 + (k % 3)
        dx, dy = round(d * math.cos(ang)), round(d * math.sin(ang))
        ax, ay, bx, by = 20 + dx, 20 + dy, 20 - dy, 20 + dx
        z = 0.0 if kind == "s" else 1000.0
        rows += [(ax, ay, z), (bx, by, z)]
        tris.append((3, len(rows) - 2, len(rows) - 1))
    return pts_of(rows), tris_of(tris)


def test_first_thirty_triangles_only():
    # count rule at level 2: positives >= count * 0.5. The first 30 triangles
    # decide: 24 steep of 30 -> blunder; 14 steep of 30 -> not (all 40: 24 of
    # 40 would be one)
    p, t = fan(True)
    det, *_ = run(2, 3, p, t, oncc={3: 0.39})
    assert det[3]
    p, t = fan(False)
    det, *_ = run(2, 3, p, t, oncc={3: 0.39})
    assert not det[3]


def one_triangle(z):
    p = pts_of(DUMMY + [(5, 5, z[0]), (8, 5, z[1]), (5, 9, z[2])])
    return p, tris_of([(0, 1, 2), (3, 4, 5)])


def test_height_threshold_th80_and_cap():
    # level 2: bins of floor(|dh|) below Hinterval; heights 0/10/100 give
    # |dh| {10, 100, 90} plus the dummy's three zeros. With Hinterval 200 the
    # cumulative share first passes 99 % at bin 100; mean + 3 sd is larger, so
    # the threshold is 100, capped at 50.
    p, t = one_triangle((0.0, 10.0, 100.0))
    *_, th, _ = run(2, 1, p, t, oncc=0.5, hinterval=200.0)
    d = np.array([0, 0, 0, 10, 100, 90], float)
    assert d.mean() + 3 * d.std() > 100
    assert th == 50.0
    # Hinterval 60: bins 0..59 hold only 4 of 6 -> th80 is the last bin, 59,
    # still capped
    p, t = one_triangle((0.0, 10.0, 100.0))
    *_, th, _ = run(2, 1, p, t, oncc=0.5, hinterval=60.0)
    assert th == 50.0
    # small differences: 903 zeros and 6 twenties -> bin 0 already holds
    # 99.3 % -> th80 = 0, below mean + 3 sd: the threshold is 0
    rows, tris = [], []
    for k in range(303):
        x, y = (k % 15) * 2, (k // 15) * 2 % 38
        rows += [(x, y, 0.0), (x + 1, y, 0.0), (x, y + 1, 20.0 if k < 3 else 0.0)]
        tris.append((3 * k, 3 * k + 1, 3 * k + 2))
    *_, th, _ = run(2, 1, pts_of(rows), tris_of(tris), hinterval=100.0)
    assert th == 0.0
    # 10 tall of 303: bin 0 holds 889 / 909 = 97.8 %, bin 20 completes it ->
    # th80 = 20 > mean + 3 sd, which becomes the threshold
    rows = [(r[0], r[1], 20.0 if (i % 3 == 2 and i // 3 < 10) else 0.0) for i, r in enumerate(rows)]
    *_, th, _ = run(2, 1, pts_of(rows), tris_of(tris), hinterval=100.0)
    d = np.array([0.0] * 889 + [20.0] * 20)
    assert th == pytest.approx(d.mean() + 3 * d.std(), rel=1e-12)
    # ... unless the bins stop at Hinterval 10: th80 = 9 < 9.24
    *_, th, _ = run(2, 1, pts_of(rows), tris_of(tris), hinterval=10.0)
    assert d.mean() + 3 * d.std() > 9 and th == 9.0
    # exactly 99 %: 99 flat triangles and one of heights 0 / 5 / 10 -> 297 of
    # 300 in bin 0 is not above 99 %, so th80 = 5 (bin 5 reaches 99.7 %),
    # above mean + 3 sd = 2.18, which is the threshold
    rows, tris = [], []
    for k in range(100):
        x, y = (k % 12) * 3, (k // 12) * 3
        rows += [(x, y, 0.0), (x + 1, y, 5.0 if k == 0 else 0.0), (x, y + 1, 10.0 if k == 0 else 0.0)]
        tris.append((3 * k, 3 * k + 1, 3 * k + 2))
    *_, th, _ = run(2, 1, pts_of(rows), tris_of(tris), hinterval=100.0)
    d = np.array([0.0] * 297 + [5.0, 10.0, 5.0])
    assert 297 / 300 * 100 == 99.0 and d.mean() + 3 * d.std() < 5
    assert th == pytest.approx(d.mean() + 3 * d.std(), rel=1e-12)


def test_height_threshold_levels_3_and_4():
    # level >= 3: min(Hinterval, GSD * 10) with GSD = max(grid spacing, 2^(level-1)),
    # then mean + 3 sd when it is lower, but never below half
    rows, tris = [], []
    for k in range(100):
        x, y = (k % 12) * 3, (k // 12) * 3
        rows += [(x, y, 0.0), (x + 1, y, 0.0), (x, y + 1, float(k % 7))]
        tris.append((3 * k, 3 * k + 1, 3 * k + 2))
    p, t = pts_of(rows), tris_of(tris)
    d = np.array([v for k in range(100) for v in (0.0, float(k % 7), float(k % 7))])
    up = d.mean() + 3 * d.std()              # 8.42
    # level 3, grid 1 m: GSD 4 -> 40; up < 20 -> 20
    *_, th, _ = run(3, 1, p.copy(), t, hinterval=1000.0)
    assert th == 20.0
    # level 3, Hinterval 15: min(15, 40) = 15, up in [7.5, 15) -> up
    *_, th, _ = run(3, 1, p.copy(), t, hinterval=15.0)
    assert th == pytest.approx(up, rel=1e-12)
    # level 4, iteration 1, RA: GSD 8 * 30 = 240; Hinterval 12 -> 12, up >= 6 -> up
    *_, th, _ = run(4, 1, p.copy(), t, hinterval=12.0, is_ra=1)
    assert th == pytest.approx(up, rel=1e-12)
    # level 3, Hinterval 16: up 8.42 >= 8 -> up; Hinterval 17: up < 8.5 -> 8.5
    *_, th, _ = run(3, 1, p.copy(), t, hinterval=16.0)
    assert th == pytest.approx(up, rel=1e-12)
    *_, th, _ = run(3, 1, p.copy(), t, hinterval=17.0)
    assert th == 8.5


@pytest.mark.parametrize("z,blamed", [((0.0, 10.0, 100.0), 5), ((0.0, 90.0, 100.0), 3),
                                      ((100.0, 0.0, 10.0), 3), ((10.0, 100.0, 90.0), 3)])
def test_neighbour_rule_blames_the_odd_vertex(z, blamed):
    # one triangle, threshold 50 (see above). Sorted heights lo <= mid <= hi:
    # (mid - lo) - (hi - mid) < -50 blames the highest vertex, > 50 the lowest.
    # Level 2 iteration 1: ortho NCC 0.3 < 0.4 lets the rule act; the own rule
    # needs < 0.1 and the count rule count > 1, so nothing else fires.
    p, t = one_triangle(z)
    lo, mid, hi = sorted(z)
    dh = (mid - lo) - (hi - mid)
    order = np.argsort(z, kind="stable")
    want = 3 + (order[2] if dh < 0 else order[0])
    assert want == blamed
    det, bc, *_ = run(2, 1, p, t, oncc=0.3, hinterval=200.0)
    assert list(np.nonzero(det)[0]) == [blamed]
    assert p[blamed]["flag"] == 1 and bc == (5, 1)
    # the blamed vertex's ortho NCC at or above 0.4 spares it
    p, t = one_triangle(z)
    det, *_ = run(2, 1, p, t, oncc={**{i: 0.3 for i in range(6)}, blamed: 0.4}, hinterval=200.0)
    assert not det.any()


def test_anchor_mode_flags_every_vertex_of_a_steep_triangle():
    # flag_blunder off (anchor points): check_neigh -> blunder when the ortho
    # NCC is below 100, i.e. always
    p, t = one_triangle((0.0, 10.0, 100.0))
    det, bc, *_ = run(2, 1, p, t, oncc=0.95, flag_blunder=False, hinterval=200.0)
    assert list(np.nonzero(det)[0]) == [3, 4, 5] and bc == (3, 3)


def test_ra_count_rule():
    # IsRA: positives >= (int)(count * 0.7 + 0.5) with count > 1, whatever the
    # ortho NCC. Four steep faces -> 4 >= 3. Three rims raised to the centre's
    # height leave two steep faces -> 2 < 3.
    p, t = spike(5.0)
    det, *_ = run(2, 3, p, t, oncc=0.99, is_ra=1)
    assert det[3]
    p, t = spike(5.0)
    p[4]["z"] = p[5]["z"] = p[6]["z"] = 5.0    # faces (3,4,5), (3,5,6) flat
    det, *_ = run(2, 3, p, t, oncc=0.99, is_ra=1)
    slopes = []
    for a, b, c in ((3, 4, 5), (3, 5, 6), (3, 6, 7), (3, 7, 4)):
        P = [np.array([p[i]["x"], p[i]["y"], p[i]["z"]]) for i in (a, b, c)]
        n = np.cross(P[1] - P[0], P[2] - P[0])
        pos = P[0][2] - P[1][2] >= 0 and P[0][2] - P[2][2] >= 0
        slopes.append(pos and math.degrees(math.acos(abs(n[2]) / np.linalg.norm(n))) > 30)
    assert sum(slopes) == 2
    assert not det[3]


def test_own_ortho_rule():
    # not RA, level >= 2: ortho NCC below min(0.1 + (it - 1) * 0.05, level
    # threshold) marks the point by itself; level 1 never does
    p = pts_of(DUMMY + [(10, 10, 0.0)])
    t = tris_of([(0, 1, 2)])
    for level, it, on, flagged in ((2, 1, 0.09, True), (2, 1, 0.11, False), (2, 5, 0.29, True),
                                   (2, 5, 0.31, False), (4, 11, 0.49, True), (4, 11, 0.51, False),
                                   (1, 1, 0.0, False)):
        # level 4 iteration 11: min(0.1 + 0.5, 0.6 - 0.1) = 0.5
        q = p.copy()
        det, *_ = run(level, it, q, t, oncc={3: on})
        assert bool(det[3]) == flagged, (level, it, on)


def test_anchor_flags():
    # level >= 3 and iteration >= 4 in the blunder pass: ortho NCC >= 0.95
    # sets the tested point's cell anchor flag, otherwise it is reset; old
    # blunders (flag 1) are not tested and keep theirs
    p = pts_of(DUMMY + [(10, 10, 0.0), (12, 10, 0.0), (14, 10, 0.0, 1)])
    t = tris_of([(0, 1, 2)])
    ug = np.zeros(W * W, UGRID_DTYPE)
    ug["anchor_flag"][[cell(p[4]), cell(p[5])]] = 1
    run(3, 4, p, t, oncc={3: 0.96, 4: 0.94, 5: 0.99}, ug=ug)
    assert ug["anchor_flag"][cell(p[3])] == 1
    assert ug["anchor_flag"][cell(p[4])] == 0
    assert ug["anchor_flag"][cell(p[5])] == 1
    ug["anchor_flag"][:] = 0
    run(3, 3, p, t, oncc=0.99, ug=ug)
    assert ug["anchor_flag"].sum() == 0


def test_counts_flags_and_heights_accumulate():
    # flags 1 and 3 count as blunders (3 becomes 1), 0 and 2 as kept points
    # whose heights extend the running min / max
    p = pts_of(DUMMY + [(10, 10, 7.0, 1), (12, 10, -3.0, 3), (14, 10, 2.5, 2), (16, 10, 40.0, 0)])
    t = tris_of([(0, 1, 2)])
    det, bc, lo, hi, *_ = run(2, 1, p, t, blunder_count=(5, 7), minz=100000.0, maxz=-100000.0)
    assert bc == (5 + 5, 7 + 2)
    assert list(p["flag"]) == [0, 0, 0, 1, 1, 2, 0]
    assert lo == 0.0 and hi == 40.0 and not det.any()
    p = pts_of(DUMMY + [(10, 10, 7.0)])
    _, bc, lo, hi, *_ = run(2, 1, p, t, minz=-5.0, maxz=3.0)
    assert (lo, hi) == (-5.0, 7.0) and bc == (4, 0)


def test_level0_iteration3_does_nothing():
    p, t = spike(5.0)
    before = p.copy()
    ug = np.zeros(W * W, UGRID_DTYPE)
    ug["anchor_flag"][:] = 1
    det, bc, lo, hi, th, _ = run(0, 3, p, t, oncc=0.0, ug=ug, blunder_count=(1, 2), minz=9.0, maxz=10.0)
    assert bc == (1, 2) and (lo, hi) == (9.0, 10.0) and not det.any()
    assert np.array_equal(p, before) and ug["anchor_flag"].all()
