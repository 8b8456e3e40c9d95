"""Full-size workloads on the GPU, checked against the oracle on stripes.

  * the headline level: a 16384^2 level-0 pair over 8 km, 4000^2 cells at 2 m, 128
    heights per cell (2.05e9 hypotheses) in FP64 — a 64-row stripe (32.8 M
    hypotheses) bit for bit against the CPU restatement;
  * the same level in the precision form — mean NCC within 1e-5
    of FP64 on the stripe, INCC within one quantum, the per-cell peak
    (first argmax of the mean NCC) the same except near-ties (FP64 top two
    closer than 1e-5), which are counted;
  * the mixed-tile distribution: a 512^2-cell and a 4096^2-cell tile drawn
    from it (relief, convergence and heights seeded), both precisions against the
    oracle on a stripe.
Over the whole level, FP64 and FAST must evaluate the same hypotheses and
gate the same cells."""
import numpy as np
import pytest

from oracle import oracle_py as O
from setsm_amd.records import LevelInputs

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _job(sc, g):
    from setsm_amd import smx
    job = smx.Job(1)
    for ti in range(2):
        job.build_pyramid(ti, sc.images[ti])
    stub = [[(sc.images[ti], sc.images[ti], np.zeros((1, 1), np.uint8))] for ti in range(2)]
    li = LevelInputs(stub, 0, sc.rpcs, g, grid_resolution=g["grid_res"])
    job.set_level(li.level)
    job.set_searchspace(g["ugrid"], g["minmax"])
    job.initialize_voxel(1)
    return job


def _stripe(job, g, r0, rows):
    cols = g["shape"][1]
    c0, c1 = r0 * cols, (r0 + rows) * cols
    lens = job.get_nccresult()["NumOfHeight"].astype(np.int64)
    nw = int(lens[c0:c1].sum())
    w, rho = job.get_voxel_rows(c0, c1, nw, with_rho=True)
    return w, rho, lens[c0:c1]


def _oracle_stripe(imgs, sc, g, r0, rows):
    """The CPU restatement on cell rows [r0, r0 + rows) over the GPU-built
    pyramid (downloaded as `imgs`)."""
    cols = g["shape"][1]
    c0, c1 = r0 * cols, (r0 + rows) * cols
    sub = {k: g[k][c0:c1] for k in ("grid_pts", "grid_wgs", "ugrid")}
    sub["boundary"] = g["boundary"]
    sub["shape"] = (rows, cols)
    li = LevelInputs([[imgs[0]], [imgs[1]]], 0, sc.rpcs, sub, grid_resolution=g["grid_res"])
    st = O.OracleState(li, keep_rho=True)
    ug = sub["ugrid"].copy()
    st.initialize_voxel(ug, 1, g["minmax"])
    acc = st.vertical_line_locus(ug, 1, g["minmax"])
    w, rho = st.export()
    return w, rho, acc


def _peaks(rho64, rhof, lens):
    """Cells whose first argmax moved, split into near-ties and flips."""
    off = np.concatenate([[0], np.cumsum(lens)])
    ties = flips = 0
    for k in range(len(lens)):
        a, b = off[k], off[k + 1]
        if b == a:
            continue
        r64 = np.where(np.isfinite(rho64[a:b]), rho64[a:b], -np.inf)
        rf = np.where(np.isfinite(rhof[a:b]), rhof[a:b], -np.inf)
        if not np.isfinite(r64.max()):
            continue
        i, j = int(np.argmax(r64)), int(np.argmax(rf))
        if i != j:
            if r64[i] - r64[j] < TOL:
                ties += 1
            else:
                flips += 1
    return ties, flips


def _check_level(sc, g, r0, rows, tag):
    from setsm_amd import smx
    job = _job(sc, g)
    print(f"{tag}: {job.num_hypotheses} hypotheses, {job.device_bytes / 2**30:.1f} GiB resident", flush=True)
    acc64 = job.vertical_line_locus(1, precision=smx.PREC_FP64, keep_rho=True)
    ev64 = job.last_evaluated
    w64, rho64, lens = _stripe(job, g, r0, rows)
    accf = job.vertical_line_locus(1, precision=smx.PREC_FAST, keep_rho=True)
    evf = job.last_evaluated
    wf, rhof, _ = _stripe(job, g, r0, rows)
    imgs = [job.get_view(ti, 0, sc.images[ti].shape) for ti in range(2)]
    job.close()
    assert accf == acc64 and evf == ev64 > 0, (acc64, accf, ev64, evf)
    wo, rho_o, acc_o = _oracle_stripe(imgs, sc, g, r0, rows)
    assert np.array_equal(w64, wo), f"{tag}: FP64 stripe differs from the oracle"
    ev = np.isfinite(rho64)
    assert np.array_equal(ev, np.isfinite(rhof))
    d = np.abs(rhof[ev].astype(np.float64) - rho_o[ev])
    incc_f = (wf >> 16).astype(np.uint16).view(np.int16).astype(int)
    incc_o = (wo >> 16).astype(np.uint16).view(np.int16).astype(int)
    ties, flips = _peaks(rho_o, rhof.astype(np.float64), lens)
    print(f"{tag}: stripe {ev.sum()} hypotheses, accessible {acc64}, max |drho| {d.max():.2e}, "
          f"rms {np.sqrt(np.mean(d * d)):.2e}, INCC off by one {(incc_f != incc_o).sum()}, "
          f"near-ties {ties}, flips {flips}", flush=True)
    assert d.max() <= TOL
    assert np.abs(incc_f - incc_o).max() <= 1
    assert flips == 0


def test_configs1_and_2_full_level(require_gpu):
    import torch
    from setsm_amd.synth import GridSpec, SceneSpec, make_grid, make_scene
    spec = SceneSpec(image_px=16384, gsd=0.5, relief=300.0, base=200.0, seed=11)
    sc = make_scene(spec, device="cuda")
    torch.cuda.empty_cache()
    cells = 4000
    inset = (16384 * 0.5 - cells * 2.0) / 2.0
    g = make_grid(sc, GridSpec(cells=cells, grid_res=2.0, band=32.0, inset=inset, x_offset=0.0, y_offset=0.0,
                               seed=12))
    g["boundary"] = np.array([inset, inset, inset + cells * 2.0, inset + cells * 2.0])
    g["shape"] = (cells, cells)
    g["grid_res"] = 2.0
    _check_level(sc, g, cells // 2, 64, "headline level")


@pytest.mark.parametrize("side,seed", [(512, 301), (4096, 302)])
def test_configs3_tiles(require_gpu, side, seed):
    import torch
    from setsm_amd.synth import draw_tiles, render_tile
    d = draw_tiles(1, seed)[0]
    d.side = side
    print(f"tile: {d}", flush=True)
    sc, g = render_tile(d, device="cuda")
    torch.cuda.empty_cache()
    _check_level(sc, g, side // 2, 16 if side > 1024 else 32, f"mixed tile {side}^2")
