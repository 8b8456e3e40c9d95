"""Oracle SGM aggregation and SGM-mode peak (setsm_code.cpp:6755-7509,
7511-7702) pinned on the CPU:

* closed forms: a 1x1 grid receives twelve start-cell contributions (four
  straight directions, and each diagonal's corner path twice — the reference
  runs it from both walls); a 1xW row of uniform voxels follows the
  one-dimensional recursion L_k = c + max(L_{k-1}, m - P2) - m;
* brute force: a literal transcription of the reference's loops and branch
  cascade (including the branch that compares V3 = 0 when only the two
  neighbouring voxels are computed) on small random cubes, compared bit for
  bit with the oracle's reduced max-of-candidates form;
* SGM peak mode: the peak of SumCost, and at level 0 from iteration 2 on the
  INCC at that peak with result1 = -1 and result3 = Nodata.
"""
import types

import numpy as np
import pytest

from oracle import oracle_py as O
from setsm_amd.records import NCCRESULT_DTYPE, UGRID_DTYPE, SmxLevel

P1, P2 = 0.3, 0.6


def _state(W, H, level=0, sgm_py=1, step=0.25):
    L = SmxLevel()
    L.level = level
    L.sgm_py = sgm_py
    L.height_step = step
    L.grid_w, L.grid_h = W, H
    li = types.SimpleNamespace(level=L, ncells=W * H)
    return O.OracleState(li, keep_rho=False)


def _random_cube(W, H, seed, nh_max=12, step=0.25):
    rng = np.random.default_rng(seed)
    n = W * H
    ncc = np.zeros(n, NCCRESULT_DTYPE)
    ug = np.zeros(n, UGRID_DTYPE)
    base = rng.integers(100, 103, n)
    nh = rng.integers(1, nh_max + 1, n)
    nh[rng.random(n) < 0.1] = 0
    ncc["minHeight"] = base
    ncc["NumOfHeight"] = nh
    ncc["maxHeight"] = base + np.ceil(nh * step).astype(int)
    ncc["max_WNCC"] = rng.integers(-200, 1000, n)
    ug["minHeight"] = base + rng.integers(0, 2, n)
    ug["maxHeight"] = base + rng.integers(1, 4, n)
    total = int(nh.sum())
    incc = rng.integers(-1000, 1000, total).astype(np.int16)
    flag = (rng.random(total) < 0.8).astype(np.uint32)
    words = flag | (incc.astype(np.uint16).astype(np.uint32) << 16)
    return ncc, ug, words


def _brute_sumcost(W, H, ncc, ug, words, step):
    """setsm_code.cpp:7017-7509 transcribed loop by loop."""
    n = W * H
    off = np.concatenate([[0], np.cumsum(ncc["NumOfHeight"].astype(np.int64))])
    incc = (words >> 16).astype(np.uint16).view(np.int16)
    flag = (words & 0xFF) != 0
    S = [np.zeros(int(ncc["NumOfHeight"][c]), np.float32) for c in range(n)]
    maxH = max(1, int(ncc["NumOfHeight"].max()))
    v_row = [0, 0, -1, 1, -1, -1, 1, 1]
    u_col = [-1, 1, 0, 0, -1, 1, -1, 1]

    def in_band(c, hs):
        ih = np.float32(float(ncc["minHeight"][c]) + hs * step)
        return ih, (ih >= ug["minHeight"][c] and ih <= ug["maxHeight"][c])

    def start(c, pre):
        for hs in range(int(ncc["NumOfHeight"][c])):
            ih, ok = in_band(c, hs)
            if ok:
                pre[hs] = np.float32(incc[off[c] + hs] / 1000.0)
                S[c][hs] = np.float32(S[c][hs] + pre[hs])

    def con(col, row, d, c, pre, cur):
        for hs in range(int(ncc["NumOfHeight"][c])):
            ih, ok = in_band(c, hs)
            if not ok:
                continue
            wsum = incc[off[c] + hs] / 1000.0
            t = (row + v_row[d]) * W + (col + u_col[d])
            mw = float(np.float32(ncc["max_WNCC"][t] / 1000.0))
            nt = int(ncc["NumOfHeight"][t])
            i0 = int(float(np.float32(ih - np.float32(ncc["minHeight"][t]))) / step)
            i1, i2 = i0 - 1, i0 + 1
            fl = lambda i: bool(flag[off[t] + i])
            V1 = V2 = V3 = 0.0
            V4 = mw - P2
            if i0 == 0 and 0 <= i2 < nt:
                if fl(i0) and fl(i2):
                    V2, V3 = float(pre[i2]) - P1, float(pre[i0])
                    a = V2 if V2 > V3 else V3
                    m = a if a > V4 else V4
                elif fl(i0):
                    V3 = float(pre[i0]); m = V3 if V3 > V4 else V4
                elif fl(i2):
                    V2 = float(pre[i2]) - P1; m = V2 if V2 > V4 else V4
                else:
                    m = V4
            elif i0 == nt - 1 and 0 <= i1 < nt:
                if fl(i0) and fl(i1):
                    V1, V3 = float(pre[i1]) - P1, float(pre[i0])
                    a = V1 if V1 > V3 else V3
                    m = a if a > V4 else V4
                elif fl(i0):
                    V3 = float(pre[i0]); m = V3 if V3 > V4 else V4
                elif fl(i1):
                    V1 = float(pre[i1]) - P1; m = V1 if V1 > V4 else V4
                else:
                    m = V4
            elif 0 < i0 < nt - 1 and 0 <= i1 < nt and 0 <= i2 < nt:
                if fl(i0) and fl(i1) and fl(i2):
                    V1, V2, V3 = float(pre[i1]) - P1, float(pre[i2]) - P1, float(pre[i0])
                    a = V1 if V1 > V2 else V2; b = a if a > V3 else V3; m = b if b > V4 else V4
                elif fl(i0) and fl(i1):
                    V1, V3 = float(pre[i1]) - P1, float(pre[i0])
                    b = V1 if V1 > V3 else V3; m = b if b > V4 else V4
                elif fl(i1) and fl(i2):
                    V1, V2 = float(pre[i1]) - P1, float(pre[i2]) - P1
                    a = V1 if V1 > V2 else V2; b = a if a > V3 else V3; m = b if b > V4 else V4
                elif fl(i0) and fl(i2):
                    V2, V3 = float(pre[i2]) - P1, float(pre[i0])
                    b = V2 if V2 > V3 else V3; m = b if b > V4 else V4
                elif fl(i0):
                    V3 = float(pre[i0]); m = V3 if V3 > V4 else V4
                elif fl(i1):
                    V1 = float(pre[i1]) - P1; m = V1 if V1 > V4 else V4
                elif fl(i2):
                    V2 = float(pre[i2]) - P1; m = V2 if V2 > V4 else V4
                else:
                    m = V4
            else:
                m = V4
            cur[hs] = np.float32(wsum + (m - mw))
            S[c][hs] = np.float32(S[c][hs] + cur[hs])

    def walk(cells, d):
        pre = np.zeros(maxH, np.float32)
        cur = np.zeros(maxH, np.float32)
        for k, (r, cl) in enumerate(cells):
            c = r * W + cl
            if k == 0:
                pre[:int(ncc["NumOfHeight"][c])] = 0
                start(c, pre)
            else:
                cur[:int(ncc["NumOfHeight"][c])] = 0
                con(cl, r, d, c, pre, cur)
                pre, cur = cur, pre

    start_row = [0, 0, 0, H - 1, 0, 0, H - 1, H - 1]
    start_col = [0, W - 1, 0, 0, 0, W - 1, 0, W - 1]
    for r in range(H):
        walk([(r, c) for c in range(W)], 0)
    for r in range(H):
        walk([(r, c) for c in range(W - 1, -1, -1)], 1)
    for c in range(W):
        walk([(r, c) for r in range(H)], 2)
    for c in range(W):
        walk([(r, c) for r in range(H - 1, -1, -1)], 3)
    for d in (4, 5, 6, 7):
        dr, dc = -v_row[d], -u_col[d]
        rows = range(H) if dr > 0 else range(H - 1, -1, -1)
        cols = range(W) if dc > 0 else range(W - 1, -1, -1)
        for r in rows:        # wall of start_col[d], one path per row
            cells, rr, cc = [], r, start_col[d]
            while 0 <= rr < H and 0 <= cc < W:
                cells.append((rr, cc)); rr += dr; cc += dc
            walk(cells, d)
        for c in cols:        # wall of start_row[d], one path per column
            cells, rr, cc = [], start_row[d], c
            while 0 <= rr < H and 0 <= cc < W:
                cells.append((rr, cc)); rr += dr; cc += dc
            walk(cells, d)
    return np.concatenate(S) if n else np.zeros(0, np.float32)


def test_sgm_single_cell_twelve_starts():
    st = _state(1, 1)
    ncc = np.zeros(1, NCCRESULT_DTYPE)
    ncc["minHeight"], ncc["maxHeight"], ncc["NumOfHeight"] = 100, 102, 8
    ug = np.zeros(1, UGRID_DTYPE)
    ug["minHeight"], ug["maxHeight"] = 100, 101      # heights 100.0 .. 101.0 in band: h = 0..4
    incc = np.array([-1000, 123, 457, 999, -5, 700, 1, 2], np.int16)
    words = 1 | (incc.astype(np.uint16).astype(np.uint32) << 16)
    st.import_voxels(ncc, words)
    got = st.sgm_aggregate(ug)
    for h in range(8):
        want = np.float32(0)
        if h <= 4:
            for _ in range(12):
                want = np.float32(want + np.float32(incc[h] / 1000.0))
        assert got[h] == want, (h, got[h], want)


def test_sgm_uniform_row_recursion():
    W, c, m = 7, 400, 900
    st = _state(W, 1)
    ncc = np.zeros(W, NCCRESULT_DTYPE)
    ncc["minHeight"], ncc["maxHeight"], ncc["NumOfHeight"], ncc["max_WNCC"] = 100, 101, 4, m
    ug = np.zeros(W, UGRID_DTYPE)
    ug["minHeight"], ug["maxHeight"] = 100, 101
    words = np.full(4 * W, 1 | (c << 16), np.uint32)
    st.import_voxels(ncc, words)
    got = st.sgm_aggregate(ug).reshape(W, 4)
    cd = c / 1000.0
    mw = float(np.float32(m / 1000.0))

    def run(k):          # path cost after k steps from a start
        L = np.float32(cd)
        for _ in range(k):
            L = np.float32(cd + (max(float(L), mw - P2) - mw))
        return L

    for col in range(W):
        contrib = [run(col), run(W - 1 - col), np.float32(cd), np.float32(cd)]   # dirs 0..3
        for wall in (0, W - 1, 0, W - 1):                                           # dirs 4..7
            if col == wall:
                contrib.append(np.float32(cd))                                      # row-wall start
            contrib.append(np.float32(cd))                                          # column-wall start
        want = np.float32(0)
        for v in contrib:
            want = np.float32(want + v)
        assert np.all(got[col] == want), (col, got[col], want)


@pytest.mark.parametrize("W,H,seed", [(5, 4, 1), (3, 7, 2), (6, 6, 3), (1, 5, 4), (8, 2, 5)])
def test_sgm_matches_literal_transcription(W, H, seed):
    step = 0.25
    ncc, ug, words = _random_cube(W, H, seed, step=step)
    st = _state(W, H, step=step)
    st.import_voxels(ncc, words)
    got = st.sgm_aggregate(ug)
    want = _brute_sumcost(W, H, ncc, ug, words, step)
    assert got.shape == want.shape
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, f"{bad.size} SumCost values differ, first {bad[:5]}: {got[bad[:5]]} vs {want[bad[:5]]}"


@pytest.mark.parametrize("iteration", [1, 2])
def test_sgm_peak_mode(iteration):
    """One cell: SumCost = 12 x INCC, so the SGM peak is the INCC peak; at
    level 0 from iteration 2 on, result0 is the INCC at the peak, result1 = -1
    and result3 = Nodata; otherwise result0/1 are the SumCost peaks."""
    st = _state(1, 1, level=0, sgm_py=1)
    ncc = np.zeros(1, NCCRESULT_DTYPE)
    ncc["minHeight"], ncc["maxHeight"], ncc["NumOfHeight"] = 100, 103, 12
    ncc["GNCC"] = -1000                      # no ortho term (VerticalLineLocus leaves -1)
    ug = np.zeros(1, UGRID_DTYPE)
    ug["minHeight"], ug["maxHeight"] = 100, 103
    incc = np.array([100, 200, 300, 250, 100, 50, 80, 400, 600, 500, 100, 90], np.int16)   # peaks at h=2, h=8
    words = 1 | (incc.astype(np.uint16).astype(np.uint32) << 16)
    st.import_voxels(ncc, words)
    st.awncc(ug, iteration)
    r = st.ncc[0]
    s = st.sumcost
    assert r["result2"] == np.float32(100 + 8 * 0.25)
    if iteration >= 2:
        assert r["result0"] == 600 and r["result1"] == -1000 and r["result3"] == -9999
    else:
        assert r["result3"] == np.float32(100 + 2 * 0.25)
        assert r["result0"] == int(np.round(float(s[8]) * 1000.0))
        assert r["result1"] == int(np.round(float(s[2]) * 1000.0))
    assert r["result4"] == 2
