"""Oracle GNCC term (setsm_code.cpp:6240-6305 with SetOrthoImageCoord
:6550-6670) pinned on the CPU:

* identical views (one image and one RPC for both sides): the ortho patches
  coincide, so every computed GNCC is exactly 1 (1000 quantised), and cells
  below the ortho_ncc threshold keep -1;
* geometry: on a rendered pair with the true surface in GridPT3.Height the
  ortho patches correlate strongly, and a 40 m error in that surface lowers
  the correlation — an index, offset or level mix-up in the ortho grid
  would not show this;
* blend: max_WNCC of cells with a GNCC term is the maximum over heights of
  alpha*INCC + (1-alpha)*GNCC (checked against the exported mean NCC to one
  quantum, GNCC being quantised in NCCresult).
"""
import numpy as np

from oracle import oracle_py as O
from setsm_amd.records import LevelInputs
from setsm_amd.synth import ortho_wgs
from tests.helpers import level_grid, scene


def _run(sc, pyr, rpcs, level, g, iteration=1, **kw):
    li = LevelInputs(pyr, level, rpcs, g, grid_resolution=g["grid_res"],
                     ortho_wgs=ortho_wgs(sc, g["boundary"], level), **kw)
    st = O.OracleState(li)
    ug = g["ugrid"].copy()
    st.initialize_voxel(ug, iteration, g["minmax"])
    acc = st.vertical_line_locus(ug, iteration, g["minmax"])
    return st, ug, acc


def test_gncc_identical_views_is_one():
    sc, pyr = scene(image_px=1024, relief=120.0, seed=5, levels=5)
    g = level_grid(sc, 2, 40)
    g["ugrid"]["ortho_ncc"][:, 1] = 900
    g["ugrid"]["ortho_ncc"][::3, 1] = 100          # below ortho_th = 0.5 at level 2
    same = [pyr[0], pyr[0]]
    st, ug, acc = _run(sc, same, [sc.rpcs[0], sc.rpcs[0]], 2, g, iteration=1)
    gn = st.ncc["GNCC"]
    on = np.ones(len(gn), bool)
    on[::3] = False
    assert acc > 0
    assert np.all(gn[~on] == -1000)
    assert np.all(gn[on] == 1000), np.unique(gn[on])


def test_gncc_tracks_the_surface():
    sc, pyr = scene(image_px=1024, relief=120.0, seed=5, levels=5)
    g = level_grid(sc, 2, 40)
    g["ugrid"]["ortho_ncc"][:, 1] = 900
    st, _, _ = _run(sc, pyr, sc.rpcs, 2, g, iteration=1)
    good = st.ncc["GNCC"].astype(float) / 1000
    g2 = dict(g)
    g2["ugrid"] = g["ugrid"].copy()
    g2["ugrid"]["Height"] += 40.0
    st2, _, _ = _run(sc, pyr, sc.rpcs, 2, g2, iteration=1)
    bad = st2.ncc["GNCC"].astype(float) / 1000
    print(f"GNCC at the true surface: mean {good.mean():.3f}; 40 m off: {bad.mean():.3f}")
    assert good.mean() > 0.6
    assert good.mean() - bad.mean() > 0.2


def test_gncc_blend_in_max_wncc():
    sc, pyr = scene(image_px=1024, relief=120.0, seed=5, levels=5)
    g = level_grid(sc, 3, 24)
    g["ugrid"]["ortho_ncc"][:, 1] = 900
    st, _, _ = _run(sc, pyr, sc.rpcs, 3, g, iteration=1)
    words, rho = st.export()
    a = O.lib().orc_ncc_alpha(3, 1, 0)
    off = np.concatenate([[0], np.cumsum(st.ncc["NumOfHeight"].astype(np.int64))])
    checked = 0
    for c in range(len(st.ncc)):
        gq = st.ncc["GNCC"][c]
        r = rho[off[c]:off[c + 1]]
        r = r[np.isfinite(r)]
        if gq == -1000 or r.size == 0:
            continue
        want = np.max(r * a + (gq / 1000.0) * (1 - a))
        assert abs(st.ncc["max_WNCC"][c] - want * 1000) <= 1.0 + 1000 * (1 - a) * 0.0005, c
        checked += 1
    assert checked > 100


def _bilinear(img, x, y):
    ix, iy = int(x), int(y)
    dx, dy = x - np.floor(x), y - np.floor(y)
    a, b = float(img[iy, ix]), float(img[iy, ix + 1])
    c, d = float(img[iy + 1, ix]), float(img[iy + 1, ix + 1])
    return a * (1 - dx) * (1 - dy) + b * dx * (1 - dy) + c * (1 - dx) * dy + d * dx * dy


def _pearson(a, b):
    a = np.asarray(a) - np.mean(a)
    b = np.asarray(b) - np.mean(b)
    va, vb = np.sum(a * a), np.sum(b * b)
    return np.sum(a * b) / np.sqrt(va * vb) if va > 1e-8 and vb > 1e-8 else None


def _brute_gncc(sc, pyr, level, g, ow, cell, half):
    """GNCC of one cell built from scratch: window offsets on the ortho grid,
    each snapped to its ortho pixel, projected at the height of the pixel's
    grid cell, sampled bilinearly in both images; above level 0 the same
    again on the level - 1 pyramid, averaged with the level term."""
    from setsm_amd.records import window_bits
    im_res = 0.5 * (1 << level)
    gr = g["grid_res"]
    b0, b1 = g["boundary"][0], g["boundary"][1]
    W = g["shape"][1]
    Hh = g["shape"][0]
    sh, sw = ow.shape[:2]
    heights = g["ugrid"]["Height"]
    win = [[[], [], [], []] for _ in range(3)]
    nxt = [[[], [], [], []] for _ in range(3)]
    combined = level >= 1
    im_next = 0.5 * (1 << (level - 1)) if combined else im_res
    swn = int((g["boundary"][2] - b0) / im_next) + 1
    shn = int((g["boundary"][3] - b1) / im_next) + 1

    def project(pc, pr, lv):
        pX, pY = b0 + pc * im_res, b1 + pr * im_res
        zc, zr = int((pX - b0) / gr), int((pY - b1) / gr)
        if not (zc < W and zr < Hh) or heights[zr * W + zc] == -1000:
            return [(0.0, 0.0), (0.0, 0.0)]
        Z = float(heights[zr * W + zc])
        lon, lat = ow[pr, pc]
        out = []
        for k in range(2):
            x, y = O.rpc_ground_to_image(sc.rpcs[k], [0.0, 0.0], lon, lat, Z)
            out.append((x / (1 << lv), y / (1 << lv)))
        return out

    def sample(pos, lv):
        vals = []
        for k in range(2):
            x, y = pos[k]
            img, mag, _ = pyr[k][lv]
            if not (0 <= x and x + 1 < img.shape[1] and 0 <= y and y + 1 < img.shape[0]):
                return None
            vals.append((_bilinear(img, x, y), _bilinear(mag, x, y)))
        return vals if vals[0][0] > 0 and vals[1][0] > 0 else None

    for row in range(-half, half + 1):
        for col in range(-half, half + 1):
            bits = window_bits(half, row, col)
            if not bits:
                continue
            tX = g["grid_pts"][cell, 0] + col * im_res
            tY = g["grid_pts"][cell, 1] + row * im_res
            pc, pr = int((tX - b0) / im_res), int((tY - b1) / im_res)
            dc, dr = int((tX - b0) / gr), int((tY - b1) / gr)
            if not (0 <= pc < sw and 0 <= pr < sh and 0 <= dc < W and 0 <= dr < Hh):
                continue
            if heights[dr * W + dc] == -1000:
                continue
            # the pixel's own grid cell gives the projection height
            vals = sample(project(pc, pr, level), level)
            if vals is not None:
                for w in range(3):
                    if bits & (1 << w):
                        win[w][0].append(vals[0][0]); win[w][1].append(vals[1][0])
                        win[w][2].append(vals[0][1]); win[w][3].append(vals[1][1])
            if combined:
                # the level - 1 ortho grid is filled only where a level pixel
                # lands (every second index); elsewhere it holds (0, 0)
                nc, nr = int((tX - b0) / im_next), int((tY - b1) / im_next)
                if not (nr * swn + nc >= 0 and nr * swn + nc < swn * shn and nc < swn and nr < shn):
                    continue
                if nc % 2 == 0 and nr % 2 == 0 and nc // 2 < sw and nr // 2 < sh:
                    pos = project(nc // 2, nr // 2, level - 1)
                else:
                    pos = [(0.0, 0.0), (0.0, 0.0)]
                vals = sample(pos, level - 1)
                if vals is not None:
                    for w in range(3):
                        if bits & (1 << w):
                            nxt[w][0].append(vals[0][0]); nxt[w][1].append(vals[1][0])
                            nxt[w][2].append(vals[0][1]); nxt[w][3].append(vals[1][1])

    def multi(wn):
        if not all(len(wn[w][0]) for w in range(3)):
            return None
        rs = [r for w in range(3) for r in (_pearson(wn[w][0], wn[w][1]), _pearson(wn[w][2], wn[w][3]))
              if r is not None]
        return np.mean(rs) if rs else None

    terms = [t for t in (multi(win), multi(nxt) if combined else None) if t is not None]
    return np.mean(terms) if terms else None


def test_gncc_matches_brute_force():
    """Level 0 (ortho pixel 0.5 m, grid cell 2 m: the pixel snapping and the
    per-pixel height matter) and level 2, SGM off."""
    for level, cells, pick in ((0, 48, (100, 777, 1500, 2000)), (2, 40, (50, 400, 801, 1111))):
        if level == 0:
            sc, pyr = scene(image_px=512, levels=1)
        else:
            sc, pyr = scene(image_px=1024, relief=120.0, seed=5, levels=5)
        g = level_grid(sc, level, cells)
        g["ugrid"]["ortho_ncc"][:, 1] = 900
        rng = np.random.default_rng(level)
        g["ugrid"]["Height"] += rng.normal(0, 0.7, len(g["ugrid"])).astype(np.float32)
        ow = ortho_wgs(sc, g["boundary"], level)
        st, _, _ = _run(sc, pyr, sc.rpcs, level, g, iteration=2, sgm_py=-1)
        half = O.lib().orc_vll_half_template(level, 15, 0.5)
        for c in pick:
            want = _brute_gncc(sc, pyr, level, g, ow, c, half)
            got = st.ncc["GNCC"][c]
            print(f"L{level} cell {c}: oracle {got} brute {want}")
            assert want is not None
            assert abs(got - want * 1000) <= 0.5 + 1e-6, (level, c, got, want)
