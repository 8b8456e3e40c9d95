"""Pins of the CPU oracle (oracle/smx_oracle.c), the reference every GPU
result is compared against. Each test checks the restatement against something it
does not share code with: the reference's own header code (golden vectors
from oracle/_ref, and the live library where it is built), closed forms,
numpy, brute force on tiny inputs, or hand-made sequences.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle import oracle_py as O
from setsm_amd.records import NCCRESULT_DTYPE, UGRID_DTYPE, LevelInputs, height_step, window_bits

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _rpc(rpc, adj, lon, lat, h):
    return O.rpc_ground_to_image(rpc, adj, lon, lat, h)


# ------------------------------------------------------------------- RPC
def test_rpc_matches_reference_golden():
    """GetObjectToImageRPC_single, bit for bit, incl. wrap and clamp branches."""
    z = np.load(os.path.join(GOLD, "rpc_vectors.npz"))
    for rpc, a, p, o in zip(z["rpc"], z["adjust"], z["pts"], z["out"]):
        x, y = _rpc(rpc, a, *p)
        assert (x, y) == (o[0], o[1])


def test_rpc_matches_reference_live():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(5)
    from setsm_amd.synth import SceneSpec, make_scene
    sc = make_scene(SceneSpec(image_px=128, seed=4))
    for _ in range(200):
        rpc = sc.rpcs[rng.integers(0, 2)]
        a = rng.normal(0, 1, 2)
        lon = rpc[2] + rng.normal(0, rpc[7])
        lat = rpc[3] + rng.normal(0, rpc[8])
        h = rpc[4] + rng.normal(0, rpc[9])
        x, y = C.c_double(), C.c_double()
        R.ref_rpc_ground_to_image(rpc.ctypes.data, a.ctypes.data, lon, lat, h, C.byref(x), C.byref(y))
        assert _rpc(rpc, a, lon, lat, h) == (x.value, y.value)


def test_rpc_closed_form():
    """Line = P, Samp = L (numerators linear, denominators 1) gives the
    affine map back; the clamp to [0, off + 1.2 scale] applies."""
    rpc = np.zeros(90)
    rpc[0:5] = [500.0, 400.0, -118.0, 35.0, 100.0]
    rpc[5:10] = [500.0, 400.0, 0.01, 0.01, 50.0]
    rpc[10 + 2] = 1.0          # line num: P
    rpc[30] = 1.0              # line den: 1
    rpc[50 + 1] = 1.0          # samp num: L
    rpc[70] = 1.0              # samp den: 1
    for lon, lat in [(-118.004, 35.003), (-117.995, 34.999), (-118.0, 35.0)]:
        x, y = _rpc(rpc, np.array([0.5, -0.25]), lon, lat, 130.0)
        ey = (lat - 35.0) / 0.01 * 500.0 + 500.0 + 0.5
        ex = (lon + 118.0) / 0.01 * 400.0 + 400.0 - 0.25
        assert abs(y - ey) < 1e-8 and abs(x - ex) < 1e-8
    x, y = _rpc(rpc, np.zeros(2), -118.02, 35.02, 100.0)   # beyond 1.2 scale -> clamp
    assert y == 500.0 + 1.2 * 500.0 and x == 0.0


def test_pyramid_coordinates_golden():
    z = np.load(os.path.join(GOLD, "pyr_vectors.npz"))
    for (x, y, sx, sy, lv), (px, py) in zip(z["inp"], z["out"]):
        assert (x / (1 << int(lv))) - sx == px and (y / (1 << int(lv))) - sy == py


def test_rpc_fit_reproduces_camera():
    """The generator's RPC fit reproduces its pushbroom model to < 0.01 px."""
    from setsm_amd.synth import SceneSpec, make_scene
    sc = make_scene(SceneSpec(image_px=256, seed=8))
    assert max(sc.fit_residual_px) < 0.01


# ------------------------------------------------------- level geometry
def test_height_step_values():
    L = O.lib()
    for lv, hs in enumerate([0.25, 0.5, 1.0, 2.0, 3.0]):
        assert L.orc_height_step(lv, 0.5) == hs == height_step(lv, 0.5)
    assert L.orc_height_step(4, 2.0) == 3.0            # clamp at 3 m


def test_window_geometry():
    L = O.lib()
    assert [L.orc_vll_half_template(lv, 15, 0.5) for lv in range(5)] == [5, 5, 5, 6, 7]
    expect = {5: (109, 45, 25), 6: (145, 45, 9), 7: (193, 77, 25)}
    for half, counts in expect.items():
        n = [0, 0, 0]
        for r in range(-half - 1, half + 2):
            for c in range(-half - 1, half + 2):
                b = L.orc_window_bits(half, r, c)
                assert b == window_bits(half, r, c)
                for k in range(3):
                    n[k] += (b >> k) & 1
        assert tuple(n) == counts
    # nested windows: inner in middle in outer
    for r in range(-5, 6):
        for c in range(-5, 6):
            b = window_bits(5, r, c)
            assert b in (0, 1, 3, 7)


# ------------------------------------------------------------ Correlate
def _corr(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return O.lib().orc_correlate(a.ctypes.data, b.ctypes.data, len(a))


def test_correlate_against_numpy():
    rng = np.random.default_rng(0)
    for n in (1, 2, 9, 25, 45, 109, 193):
        a = rng.uniform(0, 2047, n)
        b = 0.3 * a + rng.normal(0, 100, n)
        if n >= 2:
            assert abs(_corr(a, b) - np.corrcoef(a, b)[0, 1]) < 1e-12
    a = rng.uniform(0, 2047, 109)
    assert abs(_corr(a, 3.0 * a + 7.0) - 1.0) < 1e-12
    assert abs(_corr(a, -0.5 * a + 100.0) + 1.0) < 1e-12
    assert _corr(np.full(25, 812.0), a[:25]) == -99.0        # zero variance
    assert _corr(np.zeros(0), np.zeros(0)) == -99.0           # N = 0
    assert _corr([5.0], [6.0]) == -99.0


def q(v):
    """DoubleToSignedChar_result: C round() (half away from zero), via int32."""
    return int(np.int16(np.int32(np.sign(v * 1000.0) * np.floor(abs(v * 1000.0) + 0.5))))


def test_quantisers_golden():
    """The reference's quantisers: result = round half away, voxel =
    truncation, both through a 32-bit integer (so -100 -> 31072)."""
    z = np.load(os.path.join(GOLD, "quant_vectors.npz"))
    assert [q(v) for v in z["vals"]] == list(z["q_result"])
    assert np.array_equal(np.trunc(z["vals"] * 1000.0).astype(np.int32).astype(np.int16), z["q_voxel"])
    assert z["q_result"][list(z["vals"]).index(-100.0)] == 31072


# -------------------------------------------------------------- pyramid
def test_gauss_pyramid_golden():
    z = np.load(os.path.join(GOLD, "gauss_vectors.npz"))
    for k in range(3):
        assert np.array_equal(O.gauss_pyramid(z[f"img{k}"]), z[f"out{k}"])


def test_gauss_pyramid_live():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(3)
    for h, w in [(31, 45), (90, 70)]:
        img = rng.integers(0, 2048, (h, w)).astype(np.uint16)
        out = np.zeros((h // 2, w // 2), np.uint16)
        R.ref_gauss_pyramid(img.ctypes.data, w, h, out.ctypes.data)
        assert np.array_equal(O.gauss_pyramid(img), out)


def test_sobel_brute_force():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 2048, (17, 23)).astype(np.uint16)
    mag, d = O.sobel(img)
    a = img.astype(np.int64)
    for i in range(17):
        for j in range(23):
            if i in (0, 16) or j in (0, 22):
                assert mag[i, j] == 0 and d[i, j] == 0
                continue
            p = a[i - 1:i + 2, j - 1:j + 2]
            gx = (p[:, 2] - p[:, 0]) @ np.array([1, 2, 1])
            gy = np.array([1, 2, 1]) @ (p[0] - p[2])
            assert mag[i, j] == int(np.sqrt(float(gx * gx + gy * gy)) + 0.5)
            assert d[i, j] == int(np.floor(np.arctan2(float(gy), float(gx)) * 180 / 3.141592653589793))


def _orientation_brute(mag, d):
    """Direct loops over the 18-bin histogram of SubFunctions.cpp:2473-2571."""
    h, w = mag.shape
    numcols = [6, 5, 5, 5, 4, 3, 0]
    flat_m = mag.ravel().astype(np.float64)
    flat_d = d.ravel().astype(np.float64)
    out = np.zeros((h, w), np.uint8)
    for mr in range(h):
        for mc in range(w):
            b = [0.0] * 18
            sb = [0.0] * 4
            for row in range(max(-mr, -6), min(h - mr - 1, 6) + 1):
                dc = numcols[abs(row)]
                for col in range(max(-mr, -dc), min(dc, w - mc - 1) + 1):
                    idx = (mr + row) * w + (mc + col)
                    m = flat_m[idx] if idx >= 0 else 0.0
                    t = flat_d[idx] if idx >= 0 else 0.0
                    if t < 0:
                        t += 360
                    elif t > 360:
                        t -= 360
                    k = int(t / 20.0)
                    if 0 <= k < 18:
                        g = np.exp(-float(row * row + col * col) / 4.5)
                        b[k] += m * g
                        sb[int(k / 4.6)] += m * g
            best, sub = -100.0, 0
            for i in range(4):
                if sb[i] > best:
                    best, sub = sb[i], i
            best, th = -100.0, 0
            for i in range(int(sub * 4.6), int(sub * 4.6 + 4.6)):
                if b[i] > best:
                    best, th = b[i], i
            out[mr, mc] = th
    return out


def test_orientation_brute_force():
    rng = np.random.default_rng(2)
    img = rng.integers(1, 2048, (14, 19)).astype(np.uint16)
    mag, d = O.sobel(img)
    assert np.array_equal(O.orientation(mag, d), _orientation_brute(mag, d))


# ------------------------------------------------------------ peak scan
def _peak_level(n_heights, level=0, iteration=1, step=0.25, minh=100):
    """One-cell level for the AWNCC scan (no images needed)."""
    pyr = [[(np.ones((4, 4), np.uint16), np.ones((4, 4), np.uint16), np.zeros((4, 4), np.uint8))] * 5] * 2
    grid = {"grid_pts": np.zeros((1, 2)), "grid_wgs": np.zeros((1, 2)), "boundary": np.zeros(4), "shape": (1, 1)}
    li = LevelInputs(pyr, level, [np.zeros(90), np.zeros(90)], grid)
    li.level.height_step = step
    st = O.OracleState(li, keep_rho=False)
    st.ncc["NumOfHeight"] = n_heights
    st.ncc["minHeight"] = minh
    st.ncc["GNCC"] = -1000
    ug = np.zeros(1, UGRID_DTYPE)
    ug["minHeight"] = minh
    ug["maxHeight"] = minh + 1000
    return st, ug


def _scan(incc, level=0, iteration=1):
    st, ug = _peak_level(len(incc), level, iteration)
    words = ((np.array(incc, np.int16).view(np.uint16).astype(np.uint32) << 16) | 1).astype(np.uint32)
    O.lib().orc_voxels_import(st.vox, st.ncc.ctypes.data, words.ctypes.data)
    st.awncc(ug, iteration)
    return st.ncc[0]


def test_awncc_single_peak():
    r = _scan([100, 300, 500, 400, 200])
    assert r["result0"] == 500 and r["result2"] == 100 + 2 * 0.25 and r["result4"] == 1
    assert r["result1"] == q(-100.0) and r["result3"] == -1000


def test_awncc_two_peaks_and_order():
    r = _scan([100, 600, 200, 300, 800, 100])
    assert r["result0"] == 800 and r["result2"] == 100 + 4 * 0.25
    assert r["result1"] == 600 and r["result3"] == 100 + 1 * 0.25
    assert r["result4"] == 2


def test_awncc_monotone_and_flat():
    r = _scan([100, 200, 300, 400])                 # no interior peak
    assert r["result0"] == q(-100.0) and r["result4"] == 0
    r = _scan([300, 300, 300])                      # plateau: no +/- direction change
    assert r["result4"] == 0
    r = _scan([-500, -200, -400])                   # negative peak counts too
    assert r["result0"] == -200 and r["result4"] == 1


def test_awncc_max_mode_coarse_level():
    """L4 iterations <= 3 pick the plain maximum (setsm_code.cpp:7620-7635)."""
    r = _scan([100, 200, 300, 400], level=4, iteration=1)
    assert r["result0"] == 400 and r["result2"] == 100 + 3 * 0.25 and r["result4"] == 3
    assert r["result1"] == q(-1.0) and r["result3"] == -9999


# ------------------------------------------------- VLL end-to-end pins
@pytest.fixture(scope="module")
def small_level():
    from tests.helpers import level_grid, scene
    sc, pyr = scene(image_px=512, levels=1)
    g = level_grid(sc, 0, 40)
    li = LevelInputs(pyr, 0, sc.rpcs, g, grid_resolution=g["grid_res"])
    st = O.OracleState(li)
    ug = g["ugrid"].copy()
    st.initialize_voxel(ug, 1, g["minmax"])
    acc = st.vertical_line_locus(ug, 1, g["minmax"])
    words, rho = st.export()
    return dict(sc=sc, pyr=pyr, g=g, st=st, ug=ug, acc=acc, words=words, rho=rho)


def _numpy_rho(sc, pyr, g, cell, hgt):
    """Independent numpy evaluation of one hypothesis: RPC pair, window,
    rotation, bilinear samples, six Pearson correlations, their mean."""
    img0, mag0, ori0 = pyr[0][0]
    img1, mag1, ori1 = pyr[1][0]
    lon, lat = g["grid_wgs"][cell]
    z = np.zeros(2)
    rx, ry = O.rpc_ground_to_image(sc.rpcs[0], z, lon, lat, hgt)
    tx, ty = O.rpc_ground_to_image(sc.rpcs[1], z, lon, lat, hgt)
    dth = int(ori0[int(ry), int(rx)]) - int(ori1[int(ty), int(tx)])
    th = dth * 20.0 * np.pi / 180.0
    c0, s0 = np.cos(-th), np.sin(-th)
    H = 5
    vals = {k: ([], [], [], []) for k in range(3)}

    def bil(I, x, y):
        x0, y0 = int(np.floor(x)), int(np.floor(y))
        dx, dy = x - x0, y - y0
        I = I.astype(np.float64)
        return (I[y0, x0] * (1 - dx) * (1 - dy) + I[y0, x0 + 1] * dx * (1 - dy) + I[y0 + 1, x0] * (1 - dx) * dy
                + I[y0 + 1, x0 + 1] * dx * dy)

    for row in range(-H, H + 1):
        for col in range(-H, H + 1):
            b = window_bits(H, row, col)
            if not b:
                continue
            lx, ly = rx + col, ry + row
            px, py = tx + c0 * col - s0 * row, ty + s0 * col + c0 * row
            li, lm = bil(img0, lx, ly), bil(mag0, lx, ly)
            ri, rm = bil(img1, px, py), bil(mag1, px, py)
            if li > 0 and ri > 0:
                for k in range(3):
                    if b >> k & 1:
                        for arr, v in zip(vals[k], (li, lm, ri, rm)):
                            arr.append(v)
    rhos = []
    for k in range(3):
        li, lm, ri, rm = (np.array(a) for a in vals[k])
        for a, c in ((li, ri), (lm, rm)):
            if np.var(a) > 0 and np.var(c) > 0:
                rhos.append(np.corrcoef(a, c)[0, 1])
    return np.mean(rhos)


def test_vll_matches_independent_numpy(small_level):
    d = small_level
    rng = np.random.default_rng(11)
    off = np.concatenate([[0], np.cumsum(d["st"].ncc["NumOfHeight"].astype(np.int64))])
    checked = 0
    for cell in rng.choice(len(d["ug"]), 12, replace=False):
        n = int(d["st"].ncc["NumOfHeight"][cell])
        for h in rng.choice(n, 3, replace=False) if n else []:
            rho = d["rho"][off[cell] + h]
            if not np.isfinite(rho):
                continue
            hgt = float(np.float32(d["st"].ncc["minHeight"][cell] + h * 0.25))
            assert abs(_numpy_rho(d["sc"], d["pyr"], d["g"], cell, hgt) - rho) < 1e-9
            checked += 1
    assert checked > 10


def test_vll_recovers_terrain(small_level):
    """Physics invariant: the NCC peak lies near the true height for most
    well-textured cells (synthetic scene with exact RPCs)."""
    d = small_level
    st, ug = d["st"], d["ug"]
    st.awncc(ug, 1)
    ok = st.ncc["NumOfHeight"] > 0
    err = np.abs(st.ncc["result2"][ok] - d["g"]["truth"][ok])
    assert np.mean(err < 2 * 0.25 + 0.5) > 0.6
    assert np.median(err) < 0.75


def test_incc_is_truncated_mean(small_level):
    d = small_level
    ev = np.isfinite(d["rho"])
    incc = (d["words"][ev] >> 16).astype(np.uint16).view(np.int16)
    assert np.array_equal(incc, np.trunc(d["rho"][ev] * 1000.0).astype(np.int32).astype(np.int16))


# ------------------------------------------------------ InitializeVoxel
def test_initialize_voxel_lattice():
    from tests.helpers import level_grid, scene
    sc, pyr = scene(image_px=512, levels=1)
    g = level_grid(sc, 0, 24)
    li = LevelInputs(pyr, 0, sc.rpcs, g, grid_resolution=g["grid_res"])
    st = O.OracleState(li)
    ug = g["ugrid"].copy()
    ug["minHeight"][0] = 0                       # an endpoint at 0 drops the cell
    ug["maxHeight"][1] = ug["minHeight"][1] + 501  # wider than th_height (500) at L0
    inside = O.lib().orc_check_image_boundary(C.addressof(li.level), 1, float(ug["minHeight"][1]),
                                               float(ug["maxHeight"][1]), 7)
    st.initialize_voxel(ug, 1, g["minmax"])
    n = st.ncc["NumOfHeight"].astype(int)
    band = (g["ugrid"]["maxHeight"] - g["ugrid"]["minHeight"]).astype(int)
    assert n[0] == 0
    assert n[1] == 0
    if inside:
        assert ug["minHeight"][1] == -100 and ug["maxHeight"][1] == -100   # band rewritten
    else:
        assert ug["minHeight"][1] != -100
    assert np.all(n[2:] == (band[2:] / 0.25).astype(int))
    assert np.all(st.ncc["check_height_change"][2:] == 1)
    # iteration 2 with the same bands: rows kept, no height change
    words1, _ = st.export()
    st.initialize_voxel(ug, 2, g["minmax"])
    assert np.all(st.ncc["check_height_change"][2:] == 0)
    words2, _ = st.export()
    assert np.array_equal(words1, words2)
    # widen one band: the cell is re-initialised. The reference moves the
    # endpoint by the extension twice (setsm_code.cpp:5912-5913, then again
    # in the common block at 5925-5926), so 3 m of widening moves it 6 m.
    old_min = int(st.ncc["minHeight"][5])
    ug2 = ug.copy()
    ug2["minHeight"][5] -= 3
    st.initialize_voxel(ug2, 3, g["minmax"])
    k = int((old_min - int(ug2["minHeight"][5])) / 0.25 + 0.5)
    once = np.floor(old_min - k * 0.25)
    assert st.ncc["check_height_change"][5] == 1
    assert st.ncc["minHeight"][5] == np.floor(once - k * 0.25) == old_min - 6
    assert st.ncc["NumOfHeight"][5] == int((float(st.ncc["maxHeight"][5]) - float(st.ncc["minHeight"][5])) / 0.25)


def test_layout_matches_reference():
    with open(os.path.join(GOLD, "layout.json")) as f:
        lay = json.load(f)
    assert lay["sizeof_UGRID"] == UGRID_DTYPE.itemsize
    assert lay["UGRID.Matched_flag"] == UGRID_DTYPE.fields["Matched_flag"][1]
    assert lay["UGRID.ortho_ncc"] == UGRID_DTYPE.fields["ortho_ncc"][1]
    assert lay["sizeof_NCCresult"] == NCCRESULT_DTYPE.itemsize
    for f in ("NumOfHeight", "check_height_change", "max_WNCC", "GNCC", "minHeight", "result4"):
        assert lay[f"NCCresult.{f}"] == NCCRESULT_DTYPE.fields[f][1]
    assert lay["sizeof_VOXEL"] == 4 and lay["VOXEL.INCC"] == 2
    R = O.ref()
    if R is not None:
        assert R.ref_layout(0) == 20 and R.ref_layout(3) == 24
