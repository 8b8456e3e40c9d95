#!/usr/bin/env python3
"""Benchmark of the VerticalLineLocus hot path on one MI355X.

Headline line (`value`): one pyramid level at full scale — a 16384 x 16384
uint16 11-bit stereo pair rendered from a seeded fBm terrain over 8 km (Hurst
0.8, 300 m relief, 0.5 m GSD, along-track views +12 / -12 degrees,
80-coefficient RPCs), a 4000 x 4000 cell grid at 2 m, 128 height hypotheses
per cell at the level-0 step of 0.25 m (a 32 m band around the true height),
windows of the reference geometry (11x11 disc / 7x7 disc / 5x5, intensity +
Sobel magnitude), in the precision form (FP32 moments, FP64 RPC and window
totals). One step = one VerticalLineLocus call over the whole level (gate,
run lists, NCC kernel, FP64 hand-over, max_WNCC) on inputs resident in HBM.
The metric is evaluated height hypotheses per second.

The same run reports beside it:
  * the plain FP64 kernel (bit-identical to the CPU path) on that level, and
    its roofline;
  * the CPU FP64 restatement of the reference's OpenMP loop on a 64-row
    stripe of the level (median of 3 runs, the reported CPU baseline),
    checked word for word against the GPU cube;
  * SGM aggregation + peak scan (AWNCC at a level <= SGM_py) on the level;
  * `reference_cpu_case`: the reference's CPU-runnable case (2048^2 pair, levels
    L4..L0, two iterations each) on the CPU restatement, timed beside the
    GPU FP64 and FAST paths;
  * `batched`: the mixed-size tile workload — 64 tiles with sides log-uniform
    over 512..4096 cells, 32..256 heights, 50..1500 m relief, 15..40 degree
    convergence, each its own seeded scene — in ONE persistent FAST launch
    replayed from a captured hipGraph, against one FAST launch per tile and
    the plain FP64 kernel per tile, with the roofline of the persistent
    kernel and the FAST-vs-FP64 error on a stripe of every tile;
  * `served`: a Poisson stream at 20 requests/s of those tiles as requests,
    4 pyramid levels x 2 iterations each, 200 GB HBM admission budget,
    continuous batching against the plain FP64 server (one request at a
    time); the arrival window is shortened so the run takes minutes and is
    stated in the line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def save_scene(d, sc):
    os.makedirs(d, exist_ok=True)
    np.savez(os.path.join(d, "scene.npz"), img0=sc.images[0], img1=sc.images[1], rpc0=sc.rpcs[0], rpc1=sc.rpcs[1],
             terrain=sc.terrain, origin=np.array(sc.terrain_origin), extent=np.array(sc.extent),
             resid=np.array(sc.fit_residual_px))


def load_scene(d, spec):
    """The headline scene saved by an earlier run with the same --scene-cache
    (profiled runs load it instead of rendering on the GPU)."""
    from setsm_amd.synth import Scene
    path = os.path.join(d, "scene.npz")
    if not os.path.exists(path):
        return None
    z = np.load(path)
    return Scene(spec, [z["img0"], z["img1"]], [z["rpc0"], z["rpc1"]], z["terrain"], tuple(z["origin"]),
                 tuple(z["extent"]), list(z["resid"]))


def flops_per_hypothesis(half):
    from setsm_amd.records import flops_per_hypothesis as f
    return f(half)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measured_traffic():
    """HBM bytes per FAST launch at the headline workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated): the
    counters cannot be read from inside the timed run."""
    path = os.path.join(ROOT, "profiles", "vll_c1_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def peak_heights(rho64, rhof, lens, step, tol=1e-5):
    """Per-cell peak height (first argmax of the mean NCC) of the FAST cube
    against the FP64 cube: RMS height difference in metres and the cells
    whose peak moved, split into near-ties (the FP64 top two closer than
    tol) and real flips."""
    lens = lens.astype(np.int64)
    keep = lens > 0
    off = np.concatenate([[0], np.cumsum(lens)])[:-1][keep]
    ro = np.where(np.isfinite(rho64), rho64.astype(np.float64), -np.inf)
    rf = np.where(np.isfinite(rhof), rhof.astype(np.float64), -np.inf)
    idx = np.arange(len(ro)) - np.repeat(np.concatenate([[0], np.cumsum(lens)])[:-1], lens)

    def first_argmax(v):
        m = np.maximum.reduceat(v, off)
        hit = v == np.repeat(m, lens[keep])
        return m, np.minimum.reduceat(np.where(hit, idx, np.iinfo(np.int64).max), off)

    if not keep.any():
        return dict(peak_cells=0, peak_height_rms_m=0.0, peak_near_ties=0, peak_flips=0)
    mo, io = first_argmax(ro)
    _, jf = first_argmax(rf)
    ok = np.isfinite(mo)
    io, jf = io[ok], jf[ok]
    dh = (jf - io) * step
    moved = io != jf
    base = off[ok]
    gap = ro[base + io] - ro[base + jf]
    ties = int(np.sum(moved & (gap < tol)))
    return dict(peak_cells=int(ok.sum()), peak_height_rms_m=float(np.sqrt(np.mean(dh * dh))) if dh.size else 0.0,
                peak_near_ties=ties, peak_flips=int(moved.sum()) - ties)


def rho_error(rho64, rhof, w64, wf):
    ev = np.isfinite(rho64)
    d = np.abs(rhof[ev].astype(np.float64) - rho64[ev].astype(np.float64))
    return dict(max_abs=float(d.max()) if d.size else 0.0, rms=float(np.sqrt(np.mean(d * d))) if d.size else 0.0,
                incc_diff=int(np.sum((wf >> 16) != (w64 >> 16))), n=int(ev.sum()))


def level_job(sc, g, levels=1, grid_pts=True):
    """A job holding one tile's level-0 pyramid, grid and search space, after
    InitializeVoxel of iteration 1."""
    from setsm_amd import smx
    from setsm_amd.records import LevelInputs
    job = smx.Job(levels)
    for ti in range(2):
        job.build_pyramid(ti, sc.images[ti])
    stub = [[(sc.images[ti], sc.images[ti], np.zeros((1, 1), np.uint8))] for ti in range(2)]
    li = LevelInputs(stub, 0, sc.rpcs, g, grid_resolution=g["grid_res"])
    if not grid_pts:
        li.level.grid_pts = None    # GridPts feed only the GNCC term, off at level 0
    job.set_level(li.level)
    job.set_searchspace(g["ugrid"], g["minmax"])
    job.initialize_voxel(1)
    return job


def sub_grid(g, r0, rows):
    cols = g["shape"][1]
    c0, c1 = r0 * cols, (r0 + rows) * cols
    sub = {k: g[k][c0:c1].copy() for k in ("grid_pts", "grid_wgs", "ugrid")}
    for k in ("boundary", "minmax", "grid_res"):
        sub[k] = g[k]
    sub["shape"] = (rows, cols)
    return sub


# ------------------------------------------------------------ headline level
def headline(args, out):
    import torch
    from setsm_amd import smx
    from setsm_amd.records import LevelInputs
    from setsm_amd.synth import GridSpec, SceneSpec, make_grid, make_scene

    t0 = time.time()
    spec = SceneSpec(image_px=args.image_px, gsd=0.5, relief=300.0, base=200.0, seed=args.seed)
    sc = load_scene(args.scene_cache, spec) if args.scene_cache else None
    if sc is None:
        sc = make_scene(spec, device="cuda")
        if args.scene_cache:
            save_scene(args.scene_cache, sc)
    torch.cuda.empty_cache()
    log(f"scene {args.image_px}^2 x2 ready in {time.time() - t0:.1f}s, RPC fit residual {max(sc.fit_residual_px):.2e} px")
    side = args.image_px * spec.gsd
    cells = args.cells
    inset = (side - cells * 2.0) / 2.0
    g = make_grid(sc, GridSpec(cells=cells, grid_res=2.0, band=32.0, inset=inset, x_offset=0.0, y_offset=0.0,
                               seed=args.seed + 1))
    g["boundary"] = np.array([inset, inset, inset + cells * 2.0, inset + cells * 2.0])
    g["shape"] = (cells, cells)
    g["grid_res"] = 2.0
    job = level_job(sc, g, levels=2)   # level 1 feeds VerticalLineLocus_blunder
    log(f"job resident: {job.device_bytes / 2**30:.1f} GiB, {job.num_hypotheses} hypotheses")
    fph = flops_per_hypothesis(5)

    def run(prec, steps, warmup, keep_rho=False):
        for _ in range(warmup):
            job.vertical_line_locus(1, precision=prec, keep_rho=keep_rho)
        job.sync()
        torch.cuda.synchronize()
        kms = []
        t = time.perf_counter()
        for _ in range(steps):
            acc = job.vertical_line_locus(1, precision=prec, keep_rho=keep_rho)
            kms.append(job.last_kernel_ms)
        job.sync()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t) / steps
        return dict(acc=acc, evaluated=job.last_evaluated, wall_s=wall, kernel_ms=float(np.mean(kms)),
                    stats=job.last_stats)

    log("timing plain FP64 kernel")
    p64 = run(smx.PREC_FP64, max(1, args.fp64_steps), 1 if args.fp64_steps > 1 else 0, keep_rho=True)
    log(f"fp64: {p64['evaluated']} hyps, {p64['wall_s'] * 1e3:.1f} ms/step, kernel {p64['kernel_ms']:.1f} ms")
    stripe_rows = max(1, args.cpu_rows)
    r0 = cells // 2
    c0, c1 = r0 * cells, (r0 + stripe_rows) * cells
    lens = job.get_nccresult()["NumOfHeight"].astype(np.int64)
    nw = int(lens[c0:c1].sum())
    w64, rho64 = job.get_voxel_rows(c0, c1, nw, with_rho=True)

    from setsm_amd import autotune
    key = f"vll_L0_{args.image_px}px_{cells}cells_h128"
    tuned = None
    if args.autotune:
        best, rows = autotune.tune_fast(job, 1, log=log)
        autotune.save(key, best, rows, path=args.tuning_out or autotune.TUNING_PATH)
        tuned = best
    else:
        entry = autotune.load(key)
        if entry:
            job.set_fast_config(entry["best"]["config"])
            tuned = entry["best"]
    log(f"FAST launch shape: {tuned}")
    log("timing FAST kernel")
    fast = run(smx.PREC_FAST, args.steps, args.warmup)
    job.vertical_line_locus(1, precision=smx.PREC_FAST, keep_rho=True)
    wf, rhof = job.get_voxel_rows(c0, c1, nw, with_rho=True)
    fast_err = rho_error(rho64, rhof, w64, wf)
    fast_err.update(peak_heights(rho64, rhof, lens[c0:c1], 0.25))
    log(f"fast: {fast['wall_s'] * 1e3:.1f} ms/step, kernel {fast['kernel_ms']:.1f} ms, err {fast_err}")

    # SGM aggregation + peak scan (AWNCC at a level <= SGM_py) on this cube
    sgm = between = None
    if args.sgm:
        stub = [[(sc.images[ti], sc.images[ti], np.zeros((1, 1), np.uint8))] for ti in range(2)]
        li = LevelInputs(stub, 0, sc.rpcs, g, grid_resolution=2.0, sgm_py=1)
        job.set_level(li.level)
        job.set_searchspace(g["ugrid"], g["minmax"])
        job.initialize_voxel(1)
        job.vertical_line_locus(1, precision=smx.PREC_FAST)
        job.awncc(1)   # warm-up
        ts = []
        for _ in range(2):
            job.sync()
            t = time.perf_counter()
            job.awncc(1)
            ts.append(time.perf_counter() - t)
        sgm = dict(ms=1e3 * min(ts), hypotheses=job.num_hypotheses, paths=12,
                   hyp_per_s=job.num_hypotheses / min(ts), cells=cells * cells)
        log(f"SGM + peak scan: {sgm}")
        if args.between:
            between = between_run(args, job, sc, g, li, r0)
            log(f"between iterations: {between}")

    cpu = parity = None
    if not args.no_cpu:
        from oracle import oracle_py as O
        imgs = [job.get_view(ti, 0, sc.images[ti].shape) for ti in range(2)]
        sub = sub_grid(g, r0, stripe_rows)
        lio = LevelInputs([[imgs[0]], [imgs[1]]], 0, sc.rpcs, sub, grid_resolution=2.0)
        times, wo, acc_cpu, hyps_cpu = [], None, 0, 0
        threads = O.threads()
        for _ in range(3):
            st = O.OracleState(lio, keep_rho=False)
            ug = sub["ugrid"].copy()
            st.initialize_voxel(ug, 1, g["minmax"])
            t = time.perf_counter()
            acc_cpu = st.vertical_line_locus(ug, 1, g["minmax"])
            times.append(time.perf_counter() - t)
            wo, _ = st.export()
            hyps_cpu = int(st.ncc["NumOfHeight"].astype(np.int64).sum())
            del st
        med = float(np.median(times))
        cpu = dict(value=hyps_cpu / med, unit="height-hypotheses/s", cores=threads, kind="port",
                   cpu_model=cpu_model(), runs_s=[round(x, 3) for x in times],
                   sample=f"{stripe_rows} x {cells} cell stripe (rows {r0}..{r0 + stripe_rows - 1}) of the headline "
                          f"level, {hyps_cpu} hypotheses, oracle FP64 OpenMP dynamic,1 over cells, median of 3 "
                          f"runs ({med:.1f} s)")
        parity = dict(cells=c1 - c0, words=int(len(wo)), fp64_words_equal=bool(np.array_equal(wo, w64)),
                      accessible_cpu=int(acc_cpu))
        log(f"cpu: {cpu['value']:.3e} hyp/s on {threads} threads ({cpu['cpu_model']}); fp64 parity {parity}")

    job.close()
    del job
    torch.cuda.empty_cache()

    value = fast["evaluated"] / fast["wall_s"]
    redo = fast["stats"]
    log(f"FP64 pass over handed-over runs: {redo['redo_runs']} runs, {redo['redo_hyps']} hypotheses, "
        f"{redo['redo_ms']:.2f} ms")
    achieved = (fast["evaluated"] - redo["redo_hyps"]) * fph / (fast["kernel_ms"] / 1e3) / 1e12
    traffic = measured_traffic() if (args.image_px, cells) == (16384, 4000) else None
    p64_rate = p64["evaluated"] / p64["wall_s"]
    out.update({
        "metric": "height-hypotheses/s (VerticalLineLocus, one level, 16384^2 pair, 4000^2 cells x 128 heights)",
        "value": value,
        "unit": "height-hypotheses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fast["wall_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 moments, f64 window totals + rpc",
        "data": "synthetic (seeded fBm terrain, rendered stereo pair, fitted RPCs)",
        "config": {"workload": f"VLL level 0: {args.image_px}^2 uint16 pair, {cells}^2 cells at 2 m, 128 heights/cell",
                   "precision": "fast", "hypotheses_per_step": fast["evaluated"], "flops_per_hypothesis": fph},
        "roofline": dict(bound="valu", achieved=achieved, peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                         frac=achieved / PEAK_FP32_TFLOPS, traffic=traffic,
                         traffic_from="profiles/vll_c1_hbm_traffic.md (rocprofv3 FETCH_SIZE + WRITE_SIZE of this "
                                      "kernel at this workload, copy-calibrated); not read in this run"),
        "cpu_baseline": cpu,
        "plain_fp64": {"value": p64_rate, "ms_per_step": p64["wall_s"] * 1e3, "kernel_ms": p64["kernel_ms"],
                       "roofline_frac_fp64_peak":
                           p64["evaluated"] * fph / (p64["kernel_ms"] / 1e3) / 1e12 / PEAK_FP64_TFLOPS},
        "speedup_vs_plain_fp64": value / p64_rate,
        "fast_vs_fp64_error": fast_err,
        "parity_stripe": parity,
        "fast_launch": dict(tuned or {}, **fast["stats"]),
        "kernel_ms": fast["kernel_ms"],
        "sgm": sgm,
        "between_iterations": between,
    })


# ------------------------------------------------------------ between iterations
MPS_ARGS = dict(th_roh=0.4, th_roh_min=0.1, th_roh_start=0.4, th_roh_next=0.3, mpp=1.0, mpp_stereo_angle=10.0,
                final_level_iteration=1)
# Hinterval of blunder_detection_TIN: the span of the level's height bands
BD_HINTERVAL = 500.0
# SelectMPs' thresholds in the reference CPU case's chain (MPP passed apart)
C0_MPS = {k: v for k, v in MPS_ARGS.items() if k != "mpp"}


def lattice_tris(w, h):
    """The triangles of the full cell lattice, each grid square split along
    one diagonal, in row order and clockwise, as the reference's TIN lists
    its triangles (so vertex and edge cells are inside): a TIN over every
    cell of the grid."""
    c, r = np.meshgrid(np.arange(w - 1, dtype=np.uint32), np.arange(h - 1, dtype=np.uint32))
    i = (r * w + c).ravel()
    t = np.empty((2 * len(i), 3), np.uint32)
    t[0::2] = np.stack([i, i + w + 1, i + 1], 1)
    t[1::2] = np.stack([i, i + w, i + w + 1], 1)
    from setsm_amd.records import TRI_DTYPE
    return t.view(TRI_DTYPE).ravel()


def between_run(args, job, sc, g, li, r0):
    """The data-parallel steps between two iterations of SETSM's loop on the
    headline level after its AWNCC: SelectMPs, SetHeightRange and DecisionMPs'
    SetHeightRange_blunder / VerticalLineLocus_blunder (blunder level 1) /
    blunder_detection_TIN. The TIN is the full cell lattice (a point per cell
    at its peak height, two triangles per grid square). Per call: the kernel
    time on the job stream (HIP events; cells_per_s from it) and the wall
    clock with the host copies and host passes (the points come back for the
    caller's TIN, the TIN goes up from pageable memory; blunder_detection_TIN
    sums its edge statistics on the host in triangle order and writes the
    flags back into the caller's point records). The oracle is timed beside
    it on a stripe of the same level."""
    from setsm_amd.records import POINT3_DTYPE
    from setsm_amd.synth import ortho_wgs
    W, H = g["shape"][1], g["shape"][0]
    out = {}
    ncc = job.get_nccresult()
    ug0 = job.get_searchspace()
    pts_all = np.zeros(W * H, POINT3_DTYPE)
    pts_all["x"], pts_all["y"] = g["grid_pts"][:, 0], g["grid_pts"][:, 1]
    pts_all["z"] = ncc["result2"]
    tris = lattice_tris(W, H)
    t0 = time.time()
    ow = ortho_wgs(sc, g["boundary"], 1)
    log(f"blunder ortho grid {ow.shape[1]}x{ow.shape[0]} ready in {time.time() - t0:.1f}s")
    job.set_blunder_level(1, ow)
    pts_bd = pts_all.copy()
    steps = {"select_mps": lambda: job.select_mps(1, **MPS_ARGS),
             "set_height_range": lambda: job.set_height_range(pts_all, tris, 1, 1.0),
             "set_height_range_blunder": lambda: job.set_height_range_blunder(pts_all, tris),
             "vertical_line_locus_blunder": lambda: job.vertical_line_locus_blunder(1, False),
             # on the ortho NCC the previous step left on the device
             "blunder_detection_tin": lambda: job.blunder_detection_tin(pts_bd, tris, 1, True, BD_HINTERVAL)}
    for name, fn in steps.items():
        ts, ks = [], []
        for _ in range(3):
            job.set_searchspace(ug0, g["minmax"])
            pts_bd["flag"] = 0
            job.sync()
            t = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - t)
            ks.append(job.last_kernel_ms)
        out[name] = dict(kernel_ms=min(ks), cells_per_s=W * H / (min(ks) / 1e3), wall_ms=1e3 * min(ts))
        if name == "select_mps":
            out[name]["points"] = int(len(r[0]))
            out[name]["minimum_th"] = r[1]
        if name == "blunder_detection_tin":
            out[name]["points"] = int(len(pts_bd))
            out[name]["new_blunders"] = int(r[0].sum())
    out["cells"] = W * H
    out["triangles"] = int(len(tris))
    out["blunder_ortho_px"] = int(ow.shape[0] * ow.shape[1])
    if not args.no_cpu:
        from oracle import oracle_py as O
        from setsm_amd.records import LevelInputs
        rows = max(1, args.between_cpu_rows)
        # the stripe as a level of its own: its boundary starts at its first row
        sub = sub_grid(g, r0, rows)
        b = g["boundary"]
        sub["boundary"] = np.array([b[0], b[1] + r0 * g["grid_res"], b[2], b[1] + (r0 + rows) * g["grid_res"]])
        stub = [[(sc.images[ti], sc.images[ti], np.zeros((1, 1), np.uint8))] for ti in range(2)]
        lio = LevelInputs(stub, 0, sc.rpcs, sub, grid_resolution=g["grid_res"], sgm_py=1)
        c0, c1 = r0 * W, (r0 + rows) * W
        ncc_s, ug_s, pts_s = ncc[c0:c1].copy(), ug0[c0:c1].copy(), pts_all[c0:c1].copy()
        tris_s = lattice_tris(W, rows)
        ow_s = ortho_wgs(sc, sub["boundary"], 1)
        lv1 = [job.get_view(ti, 1, (sc.images[ti].shape[0] // 2, sc.images[ti].shape[1] // 2)) for ti in range(2)]
        pyr1 = [[None, lv1[ti]] for ti in range(2)]
        cpu = {}
        for name, fn in (("select_mps", lambda u: O.select_mps(lio, ncc_s, u, 1, **MPS_ARGS)),
                         ("set_height_range", lambda u: O.set_height_range(lio, u, pts_s, tris_s, 1, 1.0)),
                         ("set_height_range_blunder", lambda u: O.set_height_range_blunder(lio, u, pts_s, tris_s)),
                         ("vertical_line_locus_blunder", lambda u: O.vll_blunder(lio, pyr1, 1, ow_s, u, 1, False)),
                         ("blunder_detection_tin",
                          lambda u: O.blunder_detection_tin(lio, u, on_s, pts_s.copy(), tris_s, 1, True,
                                                            BD_HINTERVAL))):
            u = ug_s.copy()
            t = time.perf_counter()
            r = fn(u)
            dt = time.perf_counter() - t
            cpu[name] = dict(ms=1e3 * dt, cells_per_s=(c1 - c0) / dt)
            if name == "vertical_line_locus_blunder":
                on_s = r
        out["cpu_baseline"] = dict(cpu, cores=O.threads(), kind="port",
                                   sample=f"{rows} x {W} cell stripe (rows {r0}..{r0 + rows - 1}) of the same level "
                                          f"as a level of its own, its lattice and ortho grid; oracle restatement "
                                          f"(SelectMPs / SetHeightRange / blunder_detection_TIN single-threaded, "
                                          f"the blunder NCC and ortho grid on OpenMP)")
    return out


# ------------------------------------------------------------ the reference's CPU-runnable case
def same_records(a, b):
    """Field-wise equality of two record arrays (padding bytes ignored)."""
    return len(a) == len(b) and all(np.array_equal(a[f], b[f]) for f in a.dtype.names if not f.startswith("_"))


def c0_run(args):
    """The reference's CPU-runnable case: a 2048^2 pair over 1 km
    (300 m relief, views +-12 degrees), matching levels L4..L0 at SETSM's grid
    spacings (8/8/6/3/2 m), 64 heights per cell, two iterations per level
    (the second adds the level-below term above level 0). The first
    iteration of a level searches a seeded band about the terrain; the
    second searches the grid SETSM's loop makes from the first: SelectMPs,
    DecisionMPs (anchor points, then all points, with the ortho-NCC blunder
    tests at the blunder level), SetHeightRange over a TIN of the surviving
    points (the library's host TIN, as the caller's), run on the GPU FP64 path and
    checked grid for grid against the same chain on the oracle. CPU FP64
    restatement (median of 3 passes) against the GPU FP64 and FAST paths,
    InitializeVoxel + VerticalLineLocus + AWNCC per step, with word parity
    per step.

    The pair goes through the reference's formats first: each image is
    written as a GeoTIFF (WriteGeotiff, uint16) with its RPCs as a
    DigitalGlobe .XML, and the run reads them back (Readtiff_T, OpenXMLFile).
    The level-0 peak heights of the last step (AWNCC result2, before the
    host's blunder filter and TIN) are written as a float GeoTIFF DEM per
    path; the line reports their RMS against the FP64 run and the terrain."""
    import math
    from oracle import oracle_py as O
    from setsm_amd import smx
    from setsm_amd.records import LevelInputs, height_step
    from setsm_amd import host_loop
    from setsm_amd.serve import setgrids_spacing
    from setsm_amd.synth import GridSpec, SceneSpec, make_grid, make_scene, ortho_wgs
    import tempfile
    from setsm_amd import io as sio
    sc = make_scene(SceneSpec(image_px=2048, relief=300.0, seed=21, view_deg=(12.0, -12.0)), device="cuda")
    fdir = tempfile.mkdtemp(prefix="smx_c0_")
    files = dict(dir=fdir)
    for t in range(2):
        tif, xml = os.path.join(fdir, f"img{t}.tif"), os.path.join(fdir, f"img{t}.xml")
        sio.write_geotiff(tif, sc.images[t], sc.spec.gsd, sc.extent[0], sc.extent[3], projection=0, zone=14)
        sio.write_dg_xml(xml, sc.rpcs[t], gsd=(sc.spec.gsd,) * 3)
        img, _, _ = sio.read_tiff(tif)
        rpc, _, _ = sio.read_dg_xml(xml)
        assert img.tobytes() == sc.images[t].tobytes() and rpc.tobytes() == np.asarray(sc.rpcs[t]).tobytes()
        sc.images[t], sc.rpcs[t] = img, rpc
        files[f"img{t}_tif_bytes"] = os.path.getsize(tif)
    pyr = [O.build_pyramid(im, 5) for im in sc.images]   # the same views on both sides
    side = sc.spec.image_px * sc.spec.gsd
    steps, chains = [], []
    setup = smx.Job(5)
    for ti in range(2):
        for lv, (img, mag, ori) in enumerate(pyr[ti]):
            setup.set_view(ti, lv, img, mag, ori, start=(0.0, 0.0))
    for lv in (4, 3, 2, 1, 0):
        res = setgrids_spacing(lv)
        cells = int(math.ceil(1000.0 / res))
        inset = (side - cells * res) / 2.0
        band = float(max(1, int(round(64 * height_step(lv, 0.5)))))
        g = make_grid(sc, GridSpec(cells=cells, grid_res=res, band=band, inset=inset, x_offset=0.0, y_offset=0.0,
                                   seed=40 + lv))
        g["boundary"] = np.array([inset, inset, inset + cells * res, inset + cells * res])
        g["shape"] = (cells, cells)
        g["grid_res"] = res
        # the level's ortho grid: the chain leaves ortho NCC in GridPT3, which
        # turns the GNCC term on in the second iteration
        li = LevelInputs(pyr, lv, sc.rpcs, g, grid_resolution=res, ortho_wgs=ortho_wgs(sc, g["boundary"], lv))
        # iteration 1 on the GPU FP64 path, then the chain to iteration 2's grid
        B = lv if lv >= 4 else lv + 1
        ow = ortho_wgs(sc, g["boundary"], B)
        setup.set_level(li.level)
        setup.set_searchspace(g["ugrid"], g["minmax"])
        setup.initialize_voxel(1)
        setup.vertical_line_locus(1, precision=smx.PREC_FP64)
        setup.awncc(1)
        ncc1, ug1 = setup.get_nccresult(), setup.get_searchspace()
        setup.set_blunder_level(B, ow)
        kw = dict(mps_args=C0_MPS, hinterval=float(g["minmax"][1] - g["minmax"][0]), mpp=1.0)
        setup.sync()
        t0 = time.perf_counter()
        ch = host_loop.iteration_chain(setup, lv, 1, g["boundary"], res, **kw)
        gpu_chain_s = time.perf_counter() - t0
        ug2 = setup.get_searchspace()
        chains.append(dict(level=lv, ncc1=ncc1, ug1=ug1, ug2=ug2, B=B, ow=ow, kw=kw, gpu_s=gpu_chain_s,
                           gpu_tin_s=ch.tin_s, tins=ch.tins,
                           points=ch.points, kept=ch.kept, triangles=ch.triangles,
                           blunder_passes=ch.blunders.passes if ch.blunders else 0,
                           blunders=ch.blunders.blunders if ch.blunders else 0))
        steps.append((lv, g, li, [g["ugrid"], ug2]))
    setup.close()

    def cpu_chains():
        """The same chains on the oracle, from the oracle's own first
        iteration (its NCCresult and GridPT3 equal the GPU FP64 ones word for
        word): the grids must come out equal."""
        equal, t, t_tin, diff = True, 0.0, 0.0, []
        for (lv, g, li, ugs), c in zip(steps, chains):
            st = O.OracleState(li, keep_rho=False)
            ug = ugs[0].copy()
            st.initialize_voxel(ug, 1, g["minmax"])
            st.vertical_line_locus(ug, 1, g["minmax"])
            st.awncc(ug, 1)
            first = same_records(st.ncc, c["ncc1"]) and same_records(ug, c["ug1"])
            t0 = time.perf_counter()
            ch = host_loop.iteration_chain(O.ChainJob(li, st.ncc, ug), lv, 1, g["boundary"], g["grid_res"],
                                           ops=O.ChainOps(li, ug, pyr, c["B"], c["ow"]), **c["kw"])
            t += time.perf_counter() - t0
            t_tin += ch.tin_s
            ok = first and same_records(ug, c["ug2"])
            if not ok:
                diff.append(dict(level=lv, first_iteration_equal=bool(first),
                                 oracle=(ch.points, ch.kept, ch.triangles, ch.blunders.passes if ch.blunders else 0,
                                         ch.blunders.blunders if ch.blunders else 0),
                                 gpu=(c["points"], c["kept"], c["triangles"], c["blunder_passes"], c["blunders"]),
                                 cells={f: int(np.count_nonzero(ug[f] != c["ug2"][f])) for f in ug.dtype.names}))
            equal = equal and ok
        return equal, t, t_tin, diff

    def cpu_pass():
        words, t, hyps = [], 0.0, 0
        for lv, g, li, ugs in steps:
            st = O.OracleState(li, keep_rho=False)
            for it, ug0 in enumerate(ugs, start=1):
                ug = ug0.copy()
                t0 = time.perf_counter()
                st.initialize_voxel(ug, it, g["minmax"])
                st.vertical_line_locus(ug, it, g["minmax"])
                st.awncc(ug, it)
                t += time.perf_counter() - t0
                words.append(st.export()[0])
                hyps += int(st.ncc["NumOfHeight"].astype(np.int64).sum())
        return t, words, hyps

    def gpu_pass(prec):
        job = smx.Job(5)
        for ti in range(2):
            for lv, (img, mag, ori) in enumerate(pyr[ti]):
                job.set_view(ti, lv, img, mag, ori, start=(0.0, 0.0))
        words, t, ev = [], 0.0, 0
        for lv, g, li, ugs in steps:
            job.set_level(li.level)
            for it, ug in enumerate(ugs, start=1):
                job.sync()
                t0 = time.perf_counter()
                job.set_searchspace(ug, g["minmax"])
                job.initialize_voxel(it)
                job.vertical_line_locus(it, precision=prec)
                job.awncc(it)
                t += time.perf_counter() - t0
                ev += job.last_evaluated
                words.append(job.get_voxels()[0])
        nr = job.get_nccresult()
        job.close()
        # the level-0 peak heights of the last step as a DEM (no peak: -9999)
        n0 = steps[-1][1]["shape"][0]
        dem = np.where(nr["result4"] > 0, nr["result2"], -9999.0).astype(np.float32).reshape(n0, n0)
        return t, words, ev, dem

    gpu_pass(smx.PREC_FAST)   # warm-up
    t64, w64, ev64, dem64 = gpu_pass(smx.PREC_FP64)
    tf, wf, evf, demf = gpu_pass(smx.PREC_FAST)
    g0 = steps[-1][1]
    res0 = g0["grid_res"]
    b0 = g0["boundary"]
    for name, dem in (("dem_fp64", dem64), ("dem_fast", demf)):
        # north-up raster: row 0 is the northernmost cell row
        path = os.path.join(fdir, name + ".tif")
        sio.write_geotiff(path, dem[::-1].copy(), res0, b0[0], b0[1] + dem.shape[0] * res0, projection=0, zone=14)
        back, _, _ = sio.read_tiff(path)
        assert back[::-1].tobytes() == dem.tobytes()
    both = (dem64 > -9999) & (demf > -9999)
    truth = g0["truth"].reshape(dem64.shape)
    dem_rms = dict(cells=int(both.sum()), fast_vs_fp64_m=float(np.sqrt(np.mean((demf[both] - dem64[both]) ** 2))),
                   fp64_vs_terrain_m=float(np.sqrt(np.mean((dem64[both] - truth[both]) ** 2))),
                   fast_vs_terrain_m=float(np.sqrt(np.mean((demf[both] - truth[both]) ** 2))),
                   note="level-0 AWNCC peak heights of iteration 2 (before blunder filtering and TIN)")
    cpu = [cpu_pass() for _ in range(3)]
    chain_equal, chain_cpu_s, chain_cpu_tin_s, chain_diff = cpu_chains()
    tc = float(np.median([c[0] for c in cpu]))
    wc = cpu[0][1]
    parity = all(np.array_equal(a, b) for a, b in zip(w64, wc))
    dq = max(int(np.abs((a >> 16).astype(np.uint16).view(np.int16).astype(int)
                        - (b >> 16).astype(np.uint16).view(np.int16).astype(int)).max(initial=0))
             for a, b in zip(wf, wc))
    out = dict(config=dict(workload="reference CPU case: 2048^2 pair over 1 km, L4..L0 at 8/8/6/3/2 m, 64 heights, "
                                    "2 iterations per level", cells=[int(s[1]["shape"][0]) for s in steps]),
               hypotheses_evaluated=evf, cpu_s=tc, cpu_cores=O.threads(), cpu_runs_s=[round(c[0], 3) for c in cpu],
               cpu_hyp_per_s=ev64 / tc, gpu_fp64_s=t64, gpu_fast_s=tf, speedup_fp64_vs_cpu=tc / t64,
               speedup_fast_vs_cpu=tc / tf, fp64_words_equal_cpu=bool(parity), fast_incc_max_quanta=dq,
               evaluated_equal=bool(ev64 == evf), dem=dem_rms,
               chain=dict(levels=[{k: c[k] for k in ("level", "points", "kept", "triangles", "blunder_passes",
                                                     "blunders", "tins")} for c in chains],
                          gpu_s=sum(c["gpu_s"] for c in chains), cpu_s=chain_cpu_s,
                          tin_s=sum(c["gpu_tin_s"] for c in chains),
                          gpu_steps_s=sum(c["gpu_s"] - c["gpu_tin_s"] for c in chains),
                          cpu_steps_s=chain_cpu_s - chain_cpu_tin_s,
                          grids_equal_cpu=bool(chain_equal), differences=chain_diff,
                          note="iteration 1 -> 2 of every level: SelectMPs, DecisionMPs (anchor points, then "
                               "all points), SetHeightRange over a TIN of the kept points (smx_tin_create, "
                               "host); wall clock, and without the host TINs (the same on both sides: tin_s "
                               "is the GPU side's); the CPU side on the oracle (its blunder NCC on OpenMP, "
                               "the rest single-threaded)"),
               files=dict(files, note="pair read back from GeoTIFF + DigitalGlobe XML; DEMs written as GeoTIFF"))
    import shutil
    shutil.rmtree(fdir, ignore_errors=True)
    log(f"reference CPU case: {out}")
    return out


# ------------------------------------------------------------ mixed tiles
def render_pool(draws):
    import torch
    from setsm_amd.synth import SceneSpec, make_scene, tile_image_px
    scenes = []
    t = time.time()
    for d in draws:
        spec = SceneSpec(image_px=tile_image_px(d), gsd=0.5, relief=d.relief, base=200.0,
                         view_deg=(d.convergence / 2.0, -d.convergence / 2.0), seed=d.seed)
        scenes.append(make_scene(spec, device="cuda"))
    torch.cuda.empty_cache()
    px = sum(2 * s.images[0].size for s in scenes)
    log(f"{len(draws)} tile scenes rendered ({px / 1e9:.2f} G pixels) in {time.time() - t:.1f}s")
    return scenes


def batched_run(args, draws, scenes):
    """Mixed-size tiles: all tiles in one persistent FAST launch (captured with the
    peak scan as a hipGraph and replayed) vs one FAST launch per tile vs the
    plain FP64 kernel per tile."""
    import torch
    from setsm_amd import autotune, smx
    from setsm_amd.synth import level_grid
    # the earlier legs' device blocks go back to the runtime first: the 64
    # tiles take ~200 GiB of the 288 GB
    smx.release_cached()
    torch.cuda.empty_cache()
    t = time.time()
    jobs, stripes = [], []
    for d, sc in zip(draws, scenes):
        g = level_grid(sc, 0, d.side, d.nheights, seed=d.seed)
        rows = max(1, min(8, d.side // 64))
        stripes.append(sub_grid(g, d.side // 2, rows))
        jobs.append(level_job(sc, g, grid_pts=False))
        del g
    hyps_total = sum(j.num_hypotheses for j in jobs)
    resident = sum(j.device_bytes for j in jobs)
    log(f"{len(jobs)} tile jobs resident: {resident / 2**30:.1f} GiB, {hyps_total / 1e9:.2f} G hypotheses "
        f"({time.time() - t:.1f}s)")
    key = "vll_mixed_tiles"
    cfg = None
    if args.autotune:
        sub = list(range(0, len(jobs), 8))   # every 8th tile: the spread of sizes, heights and footprints
        best, rows = autotune.tune_batch([jobs[k] for k in sub], log=log)
        autotune.save(key, best, rows, path=args.tuning_out or autotune.TUNING_PATH)
        cfg = best["config"]
        # the sweep's small-tile and short-run shapes grew these jobs' lists
        for k in sub:
            jobs[k].trim()
        smx.release_cached()
    else:
        entry = autotune.load(key)
        cfg = entry["best"]["config"] if entry else None
    if cfg is not None:
        for j in jobs:
            j.set_fast_config(cfg)
    # eager batched call: the persistent kernel's own time
    smx.vertical_line_locus_batch(jobs, 1, precision=smx.PREC_FAST)
    t = time.perf_counter()
    smx.vertical_line_locus_batch(jobs, 1, precision=smx.PREC_FAST)
    eager_s = time.perf_counter() - t
    kernel_ms = jobs[0].last_kernel_ms
    evaluated = sum(j.last_evaluated for j in jobs)
    redo_hyps = sum(j.last_stats["redo_hyps"] for j in jobs)
    # the per-level sequence (fronts, persistent launch, hand-over, max_WNCC,
    # peak scan) captured as one hipGraph
    graph = smx.Graph(jobs, 1, precision=smx.PREC_FAST, with_awncc=True)
    graph.launch()
    g_ms = []
    for _ in range(max(1, args.batch_reps)):
        graph.launch()
        g_ms.append(graph.last_ms)
    graph.close()
    graph_s = min(g_ms) / 1e3
    log(f"batched: graph {graph_s * 1e3:.1f} ms, eager {eager_s * 1e3:.1f} ms, kernel {kernel_ms:.1f} ms, "
        f"{evaluated} hypotheses")
    t = time.perf_counter()
    for j in jobs:
        j.vertical_line_locus(1, precision=smx.PREC_FAST)
    each_s = time.perf_counter() - t
    t = time.perf_counter()
    k64 = 0.0
    ev64 = 0
    for j in jobs:
        j.vertical_line_locus(1, precision=smx.PREC_FP64)
        k64 += j.last_kernel_ms
        ev64 += j.last_evaluated
    fp64_s = time.perf_counter() - t
    log(f"per-tile FAST {each_s:.2f} s, plain FP64 per tile {fp64_s:.2f} s")
    for j in jobs:
        j.close()
    jobs.clear()
    smx.release_cached()
    # FAST vs FP64 on a stripe of every tile
    errs, worst = [], dict(max_abs=0.0)
    d_all, lens_all, r64_all, rf_all = [], [], [], []
    for d, sc, sub in zip(draws, scenes, stripes):
        j = level_job(sc, sub)
        j.vertical_line_locus(1, precision=smx.PREC_FP64, keep_rho=True)
        w64, rho64, _ = j.get_voxels(with_rho=True)
        j.vertical_line_locus(1, precision=smx.PREC_FAST, keep_rho=True)
        wf, rhof, _ = j.get_voxels(with_rho=True)
        lens_all.append(j.get_nccresult()["NumOfHeight"].astype(np.int64))
        j.close()
        e = rho_error(rho64, rhof, w64, wf)
        errs.append(e)
        r64_all.append(rho64)
        rf_all.append(rhof)
        if e["max_abs"] >= worst["max_abs"]:
            worst = dict(e, side=d.side, relief=round(d.relief, 1), convergence=round(d.convergence, 1))
    n = sum(e["n"] for e in errs)
    err = dict(max_abs=max(e["max_abs"] for e in errs), rms=float(np.sqrt(sum(e["rms"] ** 2 * e["n"] for e in errs) / n)),
               incc_diff=sum(e["incc_diff"] for e in errs), n=n, worst_tile=worst)
    err.update(peak_heights(np.concatenate(r64_all), np.concatenate(rf_all), np.concatenate(lens_all), 0.25))
    smx.release_cached()
    fph = flops_per_hypothesis(5)
    achieved = (evaluated - redo_hyps) * fph / (kernel_ms / 1e3) / 1e12
    plain = ev64 / fp64_s
    return dict(
        config=dict(workload="mixed-size tiles, one persistent launch", tiles=len(draws),
                    sides_cells=[min(d.side for d in draws), max(d.side for d in draws)],
                    heights=[min(d.nheights for d in draws), max(d.nheights for d in draws)],
                    relief_m=[round(min(d.relief for d in draws)), round(max(d.relief for d in draws))],
                    convergence_deg=[round(min(d.convergence for d in draws), 1),
                                     round(max(d.convergence for d in draws), 1)],
                    resident_gib=round(resident / 2**30, 1), fast_config=cfg, tile_seed=args.tile_seed),
        hypotheses=evaluated,
        batched_hyp_per_s=evaluated / graph_s,
        graph_ms=graph_s * 1e3,
        eager_batched_ms=eager_s * 1e3,
        persistent_kernel_ms=kernel_ms,
        per_tile_fast_hyp_per_s=evaluated / each_s,
        plain_fp64_hyp_per_s=plain,
        plain_fp64_roofline_frac_fp64_peak=ev64 * fph / (k64 / 1e3) / 1e12 / PEAK_FP64_TFLOPS,
        speedup_vs_plain=(evaluated / graph_s) / plain,
        roofline=dict(bound="valu", achieved=achieved, peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                      frac=achieved / PEAK_FP32_TFLOPS),
        fast_vs_fp64_error=err,
    )


def served_run(args, draws, scenes, cfg):
    """Served: Poisson requests (the tiles above, drawn without
    replacement) at `rate`/s, 4 levels x 2 iterations each, 200 GB budget."""
    from setsm_amd import smx
    from setsm_amd.serve import Request, Server
    rng = np.random.default_rng(args.tile_seed + 7)
    n = min(args.serve_requests, len(draws))
    pick = rng.permutation(len(draws))[:n]
    t = np.cumsum(rng.exponential(1.0 / args.serve_rate, n))
    stream = [Request(i, float(t[i]), int(k), draws[k].side, draws[k].nheights, int(draws[k].seed))
              for i, k in enumerate(pick)]
    tp = time.time()
    Server(scenes, levels=4, iterations=2).prepare(stream)   # request payloads, built before the clock starts
    log(f"{n} request payloads prepared in {time.time() - tp:.1f}s; arrival window {t[-1]:.2f} s")
    done_f, fast = Server(scenes, levels=4, iterations=2, fast_config=cfg).run(stream)
    smx.release_cached()
    log(f"served FAST: {fast}")
    done_p, plain = Server(scenes, levels=4, iterations=2, precision=smx.PREC_FP64, batching=False).run(stream)
    smx.release_cached()
    log(f"served plain FP64: {plain}")
    # the level-0 result of every request, FAST against the plain FP64 server:
    # peak heights (result2) and the peak scores (result0 / result1: the first
    # and second peak INCC, in units of 1e-3). AWNCC ranks peaks by the
    # quantised INCC, so a peak can move where the two peaks' INCC lie within
    # one quantum and the FAST word sits one quantum from the FP64 word
    # (a near-tie of the quantised scores).
    by = {d["rid"]: d["ncc"] for d in done_p}
    dh_all, dh_kept, dr_same, cells, moved, ties = [], [], [], 0, 0, 0
    for d in done_f:
        if d["rid"] not in by:
            continue
        a, b = d["ncc"], by[d["rid"]]
        both = (a["result4"] > 0) & (b["result4"] > 0)
        cells += int(both.sum())
        dh = (a["result2"][both] - b["result2"][both]).astype(np.float64)
        mv = a["result2"][both] != b["result2"][both]
        gap_p = np.abs(b["result0"][both].astype(np.int64) - b["result1"][both].astype(np.int64))
        gap_f = np.abs(a["result0"][both].astype(np.int64) - a["result1"][both].astype(np.int64))
        same_score = np.abs(a["result0"][both].astype(np.int64) - b["result0"][both].astype(np.int64)) <= 1
        tie = mv & ((gap_p <= 1) | (gap_f <= 1) | same_score)
        moved += int(mv.sum())
        ties += int(tie.sum())
        dh_all.append(dh)
        dh_kept.append(dh[~tie])
        dr_same.append(np.abs(a["result0"][both][~mv].astype(np.int64) - b["result0"][both][~mv].astype(np.int64)))
    cat = lambda xs: np.concatenate(xs) if xs else np.zeros(1)
    dh_all, dh_kept, dr_same = cat(dh_all), cat(dh_kept), cat(dr_same)
    err = dict(cells=cells, peak_moved_cells=moved, peak_moved_quantum_ties=ties,
               peak_height_rms_m=float(np.sqrt(np.mean(dh_all * dh_all))),
               peak_height_rms_m_without_ties=float(np.sqrt(np.mean(dh_kept * dh_kept))),
               peak_rho_max_quanta_unmoved=int(dr_same.max(initial=0)),
               note="level-0 AWNCC peaks of every request, FAST server vs plain FP64 server; a moved peak is a "
                    "quantum tie when either server's first and second peak INCC lie within one quantum (1e-3) or "
                    "the two servers' chosen peaks score within one quantum of each other")
    # VALU roofline of the served run: algorithmic flops of every evaluated
    # hypothesis (per level's window size) over the time in VerticalLineLocus calls
    for r in (fast, plain):
        r["achieved_tflops_in_vll_calls"] = r["flops"] / r["gpu_call_s"] / 1e12
    fast["roofline"] = dict(bound="valu", achieved=fast["achieved_tflops_in_vll_calls"], peak=PEAK_FP32_TFLOPS,
                            unit="TFLOP/s", frac=fast["achieved_tflops_in_vll_calls"] / PEAK_FP32_TFLOPS,
                            note="wall time of the batched VerticalLineLocus calls (launch + hand-over + "
                                 "max_WNCC included)")
    plain["roofline_frac_fp64_peak"] = plain["achieved_tflops_in_vll_calls"] / PEAK_FP64_TFLOPS
    return dict(config=dict(workload="Poisson tile requests, continuous batching", rate_per_s=args.serve_rate,
                            requests=n, arrival_window_s=round(float(t[-1]), 3), levels=4, iterations_per_level=2,
                            budget_gib=200, sizes_cells=[int(min(draws[k].side for k in pick)),
                                                         int(max(draws[k].side for k in pick))],
                            note="arrival window shortened from 300 s so the run takes minutes"),
                fast=fast, plain_fp64=plain, speedup_vs_plain=fast["hyp_per_s"] / plain["hyp_per_s"],
                fast_vs_fp64=err)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--image-px", type=int, default=16384)
    ap.add_argument("--cells", type=int, default=4000)
    ap.add_argument("--fp64-steps", type=int, default=2, help="timed plain FP64 steps (at least one runs; "
                    "below 2, without a warm-up)")
    ap.add_argument("--cpu-rows", type=int, default=64, help="cell rows of the CPU baseline stripe")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sgm", dest="sgm", action="store_false")
    ap.add_argument("--no-between", dest="between", action="store_false",
                    help="skip the between-iterations leg (SelectMPs, SetHeightRange, blunder helpers)")
    ap.add_argument("--between-cpu-rows", type=int, default=200, help="cell rows of its CPU baseline stripe")
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--scene-cache", default="", help="directory to load / save the headline scene")
    ap.add_argument("--autotune", action="store_true", help="time every FAST launch shape and store the best")
    ap.add_argument("--tuning-out", default=None,
                    help="file the --autotune table is written to (default: the package's tuning.json)")
    ap.add_argument("--no-c0", dest="c0", action="store_false", help="skip the reference CPU case (2048^2 pair, levels L4..L0)")
    ap.add_argument("--tiles", type=int, default=64, help="tiles of the mixed-size workload (0 = skip)")
    ap.add_argument("--tile-seed", type=int, default=2026)
    ap.add_argument("--batch-reps", type=int, default=2)
    ap.add_argument("--serve-requests", type=int, default=24, help="requests of the served run (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--copy-calib-gib", type=float, default=0.0,
                    help="first run one device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the VerticalLineLocus path shards by whole tiles with no "
              "collective, and runs over several GPUs are not implemented (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_calib_gib > 0:
        nbytes = int(args.copy_calib_gib * 2**30)
        src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {nbytes} bytes read, {nbytes} bytes written")
        del src, dst
        torch.cuda.empty_cache()

    out = {}
    headline(args, out)
    if args.c0:
        out["reference_cpu_case"] = c0_run(args)
    if args.tiles > 0 or args.serve_requests > 0:
        from setsm_amd import autotune
        from setsm_amd.synth import draw_tiles
        draws = draw_tiles(max(args.tiles, args.serve_requests), args.tile_seed)
        scenes = render_pool(draws)
        entry = autotune.load("vll_mixed_tiles")
        cfg = entry["best"]["config"] if entry else None
        if args.tiles > 0:
            out["batched"] = batched_run(args, draws[: args.tiles], scenes[: args.tiles])
            cfg = out["batched"]["config"]["fast_config"]
            log(f"batched: {json.dumps(out['batched'])}")
        if args.serve_requests > 0:
            out["served"] = served_run(args, draws, scenes, cfg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
