"""CPU tests of the caller-side loop (`setsm_amd.host_loop`): both TINs (the
library's own `smx_tin_create` and Qhull's) are clockwise Delaunay TINs of
the lattice points, and DecisionMPs' control flow (setsm_code.cpp:8891-9093)
stops where the reference's does."""
import numpy as np
import pytest

from setsm_amd import host_loop
from setsm_amd.records import POINT3_DTYPE


def lattice_points(W, frac, seed):
    rng = np.random.default_rng(seed)
    cells = np.nonzero(rng.random(W * W) < frac)[0]
    pts = np.zeros(len(cells), POINT3_DTYPE)
    pts["x"], pts["y"] = 10.0 + (cells % W) * 2.0, 20.0 + (cells // W) * 2.0
    return pts


def delaunay_checks(pts, t):
    """Clockwise, none flat, tiling the hull, every point used, empty
    circumcircles (exact in float64 at these lattice sizes)."""
    from scipy.spatial import ConvexHull
    xy = np.stack([(pts["x"] - 10.0) / 2.0, (pts["y"] - 20.0) / 2.0], 1)
    a, b, c = xy[t["a"]], xy[t["b"]], xy[t["c"]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    assert (cross < 0).all()
    assert abs(-cross.sum() / 2 - ConvexHull(xy).volume) < 1e-6
    used = np.zeros(len(pts), bool)
    used[t["a"]] = used[t["b"]] = used[t["c"]] = True
    assert used.all()
    for k in range(len(t)):
        P = (a[k], c[k], b[k])   # counter-clockwise
        ax, ay = P[0][0] - xy[:, 0], P[0][1] - xy[:, 1]
        bx, by = P[1][0] - xy[:, 0], P[1][1] - xy[:, 1]
        cx, cy = P[2][0] - xy[:, 0], P[2][1] - xy[:, 1]
        inc = ((ax * ax + ay * ay) * (bx * cy - cx * by) - (bx * bx + by * by) * (ax * cy - cx * ay)
               + (cx * cx + cy * cy) * (ax * by - bx * ay))
        assert (inc <= 0).all(), k


BOX = (10.0, 20.0, 10.0 + 2.0 * 99, 20.0 + 2.0 * 99)


@pytest.mark.parametrize("tin", [host_loop.native_tin, host_loop.qhull_tin])
@pytest.mark.parametrize("W,frac,seed", [(60, 0.35, 5), (30, 1.0, 1), (40, 0.08, 2), (99, 0.02, 7), (7, 0.9, 3)])
def test_tin_is_a_clockwise_delaunay_tin(tin, W, frac, seed):
    pts = lattice_points(W, frac, seed)
    delaunay_checks(pts, tin(pts, BOX, 2.0))


def test_native_tin_special_inputs():
    from setsm_amd import smx
    # a collinear run through the centre first, then points off it
    pts = np.zeros(12, POINT3_DTYPE)
    pts["x"] = 10.0 + 2.0 * np.array([5, 6, 4, 7, 3, 8, 2, 5, 5, 9, 1, 5])
    pts["y"] = 20.0 + 2.0 * np.array([5, 5, 5, 5, 5, 5, 5, 7, 2, 5, 5, 9])
    t = host_loop.native_tin(pts, BOX, 2.0)
    delaunay_checks(pts, t)
    assert len(t) == len(host_loop.qhull_tin(pts, BOX, 2.0))
    # all on one line, or fewer than three points: no triangles
    assert len(host_loop.native_tin(pts[pts["y"] == 30.0], BOX, 2.0)) == 0
    assert len(host_loop.native_tin(pts[:2], BOX, 2.0)) == 0
    # two points in one lattice cell: the later one stands for it
    dup = lattice_points(12, 0.6, 4)
    more = np.concatenate([dup, dup[3:4]])
    more["x"][-1] += 0.4      # same cell after snapping
    t = host_loop.native_tin(more, BOX, 2.0)
    used = set(t["a"]) | set(t["b"]) | set(t["c"])
    assert len(dup) in used and 3 not in used
    # a point off the lattice's box is an error, as the reference rejects it
    bad = dup.copy()
    bad["x"][0] = 0.0
    with pytest.raises(smx.SmxError):
        host_loop.native_tin(bad, BOX, 2.0)
    # through the C entry point: a buffer too small, a bad resolution
    import ctypes as C
    mm = (C.c_double * 4)(*BOX)
    small = np.zeros(3, host_loop.TRI_DTYPE)
    lib = smx.lib()
    assert lib.smx_tin_create(dup.ctypes.data, len(dup), mm, 2.0, small.ctypes.data, len(small)) == -2
    assert lib.smx_tin_create(dup.ctypes.data, len(dup), mm, 0.0, small.ctypes.data, len(small)) == -3


def test_qhull_tin_degenerate_inputs():
    pts = lattice_points(10, 0.0, 0)
    assert len(host_loop.qhull_tin(pts, BOX, 2.0)) == 0
    line = np.zeros(5, POINT3_DTYPE)
    line["x"] = 10.0 + 2.0 * np.arange(5)
    line["y"] = 20.0
    assert len(host_loop.qhull_tin(line, BOX, 2.0)) == 0


class ScriptedOps:
    """blunder_detection_TIN returns scripted (kept, blunders) counts and
    marks `new` fresh points as blunders; the other steps are recorded."""

    def __init__(self, script):
        self.script = list(script)
        self.calls = []

    def set_height_range_blunder(self, pts, tris):
        self.calls.append("shrb")

    def vll_blunder(self, iteration, bblunder):
        self.calls.append(f"vllb{int(bblunder)}")

    def blunder_detection(self, pts, tris, iteration, flag_blunder, hinterval, count_bl, zero_ortho, minz, maxz):
        self.calls.append(f"bd{count_bl}{'z' if zero_ortho else ''}")
        kept, blunders, new = self.script.pop(0)
        det = np.zeros(len(pts), bool)
        fresh = np.nonzero(pts["flag"] == 0)[0][:new]
        det[fresh] = True
        pts["flag"][fresh] = 1
        return det, (kept, blunders), min(minz, 1.0), max(maxz, 2.0)


def test_decision_loop_stops_like_the_reference():
    pts = lattice_points(40, 0.6, 3)
    n = len(pts)
    # new blunders in each pass, counts change: runs to max_count (level 2, blunders: 10)
    ops = ScriptedOps([(n - 5 * (k + 1), 5 * (k + 1), 5) for k in range(12)])
    r = host_loop.decision_mps(ops, pts.copy(), 2, 1, 100.0, True, BOX, 2.0)
    assert r.passes == 10 and r.blunders == 50
    assert ops.calls[:3] == ["shrb", "vllb1", "bd0"] and ops.calls[-2:] == ["shrb", "vllb0"]
    # level 4: 30 passes allowed; level 3: 20; anchors: 10
    for level, flag_blunder, want in ((4, True, 30), (3, True, 20), (4, False, 10)):
        ops = ScriptedOps([(n - (k + 1), k + 1, 1) for k in range(40)])
        r = host_loop.decision_mps(ops, pts.copy(), level, 1, 100.0, flag_blunder, BOX, 2.0)
        assert r.passes == want, (level, flag_blunder)
    # the same counts twice: stops after the repeat
    ops = ScriptedOps([(n - 3, 3, 3), (n - 3, 3, 0), (n - 3, 3, 0)])
    assert host_loop.decision_mps(ops, pts.copy(), 2, 1, 100.0, True, BOX, 2.0).passes == 2
    # no blunder in the first pass: one pass
    ops = ScriptedOps([(n, 0, 0)])
    assert host_loop.decision_mps(ops, pts.copy(), 2, 1, 100.0, True, BOX, 2.0).passes == 1
    # RA: no blunder NCC, a zero ortho NCC
    ops = ScriptedOps([(n, 0, 0)])
    host_loop.decision_mps(ops, pts.copy(), 2, 1, 100.0, True, (10.0, 20.0), 2.0, is_ra=True)
    assert ops.calls == ["bd0z", "shrb", "vllb0"]
    # level 0 iteration 3: the reference returns at once
    ops = ScriptedOps([])
    r = host_loop.decision_mps(ops, pts.copy(), 0, 3, 100.0, True, BOX, 2.0)
    assert r.passes == 0 and ops.calls == []
    # a TIN of at most 20 triangles: no pass
    few = lattice_points(5, 1.0, 0)[:12]
    ops = ScriptedOps([])
    r = host_loop.decision_mps(ops, few, 2, 1, 100.0, True, BOX, 2.0)
    assert r.passes == 0 and ops.calls == ["shrb", "vllb0"]
