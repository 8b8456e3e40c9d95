"""The FAST path's window arithmetic (setsm_amd/csrc/fast_window.hpp, built
for the host by tests/fastwin) against the oracle's FP64 mean NCC, on the
CPU: same hypotheses (positions from the oracle's RPC and pyramid
coordinates, rotation from the orientation images), windows with nodata
(DN 0) samples and windows crossing the image edge (checked path), and
interior windows through the straight-line unchecked path. Contract: the
mean NCC within 1e-5 of FP64, or the hypothesis handed to the FP64 path
(kRedoRho) when a window's spread is not clear of FP32 rounding."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import oracle_py as O
from setsm_amd.records import LevelInputs
from tests.helpers import level_grid, scene

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "fastwin", "_build", "libfastwin.so")
REDO = 3e300   # kRedoRho


def _lib():
    if not os.path.exists(LIB):
        pytest.skip("host build of the window arithmetic missing (run __graft_entry__.build())")
    L = C.CDLL(LIB)
    vp, d, f, i = C.c_void_p, C.c_double, C.c_float, C.c_int
    L.fw_window.restype = i
    L.fw_window.argtypes = [i, i, i, vp, i, i, vp, i, i, d, d, d, d, f, f, C.POINTER(d)]
    return L


def _pack(img, mag):
    return np.ascontiguousarray(img.astype(np.uint32) | (mag.astype(np.uint32) << 16))


def _run(pyr, sc, g, level=0, checked=True, limit=None, stage=0):
    L = _lib()
    li = LevelInputs(pyr, level, sc.rpcs, g, grid_resolution=g["grid_res"])
    st = O.OracleState(li, keep_rho=True)
    ug = g["ugrid"].copy()
    st.initialize_voxel(ug, 1, g["minmax"])
    st.vertical_line_locus(ug, 1, g["minmax"])
    _, rho = st.export()
    off = np.concatenate([[0], np.cumsum(st.ncc["NumOfHeight"].astype(np.int64))])
    step = O.lib().orc_height_step(level, 0.5)
    half = O.lib().orc_vll_half_template(level, 15, 0.5)
    tex = [_pack(pyr[k][level][0], pyr[k][level][1]) for k in range(2)]
    ori = [pyr[k][level][2] for k in range(2)]
    hgt, wid = tex[0].shape
    out = C.c_double()
    worst, n, redo = 0.0, 0, 0
    cells = range(len(ug)) if limit is None else range(0, len(ug), max(1, len(ug) // limit))
    for c in cells:
        lon, lat = g["grid_wgs"][c]
        for h in range(int(st.ncc["NumOfHeight"][c])):
            r = rho[off[c] + h]
            if not np.isfinite(r):
                continue
            z = float(np.float32(st.ncc["minHeight"][c] + h * step))
            xs = []
            for k in range(2):
                x, y = O.rpc_ground_to_image(sc.rpcs[k], [0.0, 0.0], lon, lat, z)
                xs.append((x / (1 << level), y / (1 << level)))
            (rx, ry), (tx, ty) = xs
            d = int(ori[0][int(ry), int(rx)]) - int(ori[1][int(ty), int(tx)])
            th = float(d) * 20.0 * math.pi / 180.0
            ok = L.fw_window(half, 1 if checked else 0, stage, tex[0].ctypes.data, wid, hgt, tex[1].ctypes.data, wid, hgt,
                             rx, ry, tx, ty, math.cos(-th), math.sin(-th), C.byref(out))
            n += 1
            if ok and out.value == REDO:
                redo += 1
                continue
            got = out.value if (ok and out.value != 1e300) else -1.0
            worst = max(worst, abs(got - r))
    return worst, n, redo


def test_window_math_nodata_checked():
    sc, _ = scene(image_px=512, levels=1)
    rng = np.random.default_rng(3)
    imgs = []
    for ti in range(2):
        im = sc.images[ti].copy()
        for _ in range(6):
            y0, x0 = rng.integers(0, 472, 2)
            im[y0:y0 + rng.integers(8, 40), x0:x0 + rng.integers(8, 40)] = 0
        imgs.append(im)
    pyr = [O.build_pyramid(im, 1) for im in imgs]
    g = level_grid(sc, 0, 40, band=8.0)
    worst, n, redo = _run(pyr, sc, g, checked=True)
    print(f"nodata, checked path: {n} hypotheses, {redo} to FP64, max |drho| {worst:.2e}")
    assert n > 1000
    assert worst <= 1e-5
    assert redo < 0.05 * n


def test_window_math_interior_unchecked():
    sc, pyr = scene(image_px=512, levels=1)
    g = level_grid(sc, 0, 40, band=8.0)
    for stage in (0, 24):
        worst, n, redo = _run(pyr, sc, g, checked=False, stage=stage)
        print(f"interior, unchecked path, pivot offset {stage}: {n} hypotheses, {redo} to FP64, "
              f"max |drho| {worst:.2e}")
        assert n > 1000
        assert worst <= 1e-5
        assert redo < 0.01 * n


def test_window_math_flat_patches():
    """Flat (constant, non-zero) patches: the reference's variance is an exact
    zero there (-99 windows); the FP32 moments must not turn rounding noise
    into a correlation."""
    sc, _ = scene(image_px=512, levels=1)
    rng = np.random.default_rng(5)
    imgs = []
    for ti in range(2):
        im = sc.images[ti].copy()
        for _ in range(10):
            y0, x0 = rng.integers(0, 472, 2)
            im[y0:y0 + rng.integers(10, 40), x0:x0 + rng.integers(10, 40)] = rng.integers(50, 3000)
        imgs.append(im)
    pyr = [O.build_pyramid(im, 1) for im in imgs]
    g = level_grid(sc, 0, 40, band=8.0)
    for checked, stage in ((True, 0), (False, 0), (True, 24), (False, 24)):
        worst, n, redo = _run(pyr, sc, g, checked=checked, stage=stage)
        print(f"flat patches, checked={checked} pivot offset {stage}: {n} hypotheses, {redo} to FP64, "
              f"max |drho| {worst:.2e}")
        assert n > 1000
        assert worst <= 1e-5
