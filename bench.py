#!/usr/bin/env python3
"""Benchmark of the VerticalLineLocus hot path on one MI355X.

Workload (one pyramid level at full scale): a 16384 x 16384 uint16 11-bit
stereo pair rendered from a seeded fBm terrain over 8 km (Hurst 0.8, 300 m
relief, 0.5 m GSD, along-track views +12 / -12 degrees, 80-coefficient RPCs),
a 4000 x 4000 cell grid at 2 m, 128 height hypotheses per cell at the level-0
step of 0.25 m (a 32 m band around the true height), 5x... windows of the
reference geometry (11x11 disc / 7x7 / 5x5, intensity + Sobel magnitude).

One step = one VerticalLineLocus call over the whole level (gate + run list +
NCC kernel + max_WNCC) on inputs already resident in HBM. The metric is
evaluated height hypotheses per second. The same run times the plain FP64
kernel (bit-identical to the CPU path) and, on a bounded stripe of cells, the
CPU FP64 restatement of the reference's OpenMP loop (the reported CPU
baseline), and checks the GPU cube against it.

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


def window_counts(half):
    from setsm_amd.records import window_bits
    n = [0, 0, 0]
    for r in range(-half, half + 1):
        for c in range(-half, half + 1):
            b = window_bits(half, r, c)
            for k in range(3):
                if b & (1 << k):
                    n[k] += 1
    return n


def flops_per_hypothesis(half):
    """Algorithmic flops of one hypothesis: bilinear taps (4 channels x 3 FMA
    per outer-window sample) + moments (5 FMA per sample per window per
    channel pair). RPC FP64 work (~400 flops) is not counted."""
    n0, n1, n2 = window_counts(half)
    return 24 * n0 + 20 * (n0 + n1 + n2)


def save_scene(d, sc):
    os.makedirs(d, exist_ok=True)
    np.savez(os.path.join(d, "scene.npz"), img0=sc.images[0], img1=sc.images[1], rpc0=sc.rpcs[0], rpc1=sc.rpcs[1],
             terrain=sc.terrain, origin=np.array(sc.terrain_origin), extent=np.array(sc.extent),
             resid=np.array(sc.fit_residual_px))


def load_scene(d, spec):
    from setsm_amd.synth import Scene
    path = os.path.join(d, "scene.npz")
    if not os.path.exists(path):
        return None
    z = np.load(path)
    return Scene(spec, [z["img0"], z["img1"]], [z["rpc0"], z["rpc1"]], z["terrain"], tuple(z["origin"]),
                 tuple(z["extent"]), list(z["resid"]))


def batched_run(pool, ntiles, seed, cfg, tune_key=None, tuning_out=None):
    """Tiles of mixed size (log-uniform 128..1024 cells per side) and height
    range (32..256 per cell) from independently seeded scenes, one level:
    ONE persistent FAST launch over all tiles vs one FAST launch per tile vs
    the plain FP64 kernel one tile per launch. With tune_key, the launch shape
    is first tuned on these tiles and stored under that key."""
    import math
    from setsm_amd import smx
    from setsm_amd.records import LevelInputs
    from setsm_amd.synth import level_grid
    rng = np.random.default_rng(seed + 5)
    jobs = []
    for k in range(ntiles):
        sc = pool[k % len(pool)]
        cells = int(round(math.exp(rng.uniform(math.log(128), math.log(1024)))))
        side = sc.spec.image_px * sc.spec.gsd
        cells = min(cells, int((side - 100.0) / 2.0))
        nh = int(rng.integers(32, 257))
        g = level_grid(sc, 0, cells, nh, seed=int(rng.integers(0, 1 << 20)))
        j = smx.Job(1)
        if cfg is not None:
            j.set_fast_config(cfg)
        for ti in range(2):
            j.build_pyramid(ti, sc.images[ti])
        stub = [[(sc.images[ti], sc.images[ti], np.zeros((1, 1), np.uint8))] for ti in range(2)]
        li = LevelInputs(stub, 0, sc.rpcs, g, grid_resolution=g["grid_res"])
        j.set_level(li.level)
        j.set_searchspace(g["ugrid"], g["minmax"])
        j.initialize_voxel(1)
        jobs.append(j)

    def timed(fn, reps=3):
        fn()
        best = None
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        return best

    if tune_key:
        from setsm_amd import autotune
        best, rows = autotune.tune_batch(jobs, log=log)
        autotune.save(tune_key, best, rows, path=tuning_out or autotune.TUNING_PATH)
        cfg = best["config"]
    t_batch = timed(lambda: smx.vertical_line_locus_batch(jobs, 1, precision=smx.PREC_FAST))
    hyps = sum(j.last_evaluated for j in jobs)
    t_each = timed(lambda: [j.vertical_line_locus(1, precision=smx.PREC_FAST) for j in jobs])
    t_fp64 = timed(lambda: [j.vertical_line_locus(1, precision=smx.PREC_FP64) for j in jobs], reps=1)
    for j in jobs:
        j.close()
    return dict(tiles=ntiles, hypotheses=hyps, batched_hyp_per_s=hyps / t_batch, per_tile_fast_hyp_per_s=hyps / t_each,
                plain_fp64_hyp_per_s=hyps / t_fp64, speedup_vs_plain=t_fp64 / t_batch, fast_config=cfg)


def peak_heights(rho64, rhof, lens, step, tol=1e-5):
    """Per-cell peak height (argmax of the mean NCC, first maximum) of the
    FAST cube against the FP64 cube: RMS height difference in metres and the
    cells whose peak moved, split into near-ties (the FP64 top two closer
    than tol) and real flips."""
    lens = lens.astype(np.int64)
    keep = lens > 0
    off = np.concatenate([[0], np.cumsum(lens)])[:-1][keep]
    seg = np.repeat(np.arange(len(lens)), lens)
    ro = np.where(np.isfinite(rho64), rho64.astype(np.float64), -np.inf)
    rf = np.where(np.isfinite(rhof), rhof.astype(np.float64), -np.inf)
    idx = np.arange(len(ro)) - np.repeat(np.concatenate([[0], np.cumsum(lens)])[:-1], lens)

    def first_argmax(v):
        m = np.maximum.reduceat(v, off)
        hit = v == np.repeat(m, lens[keep])
        return m, np.minimum.reduceat(np.where(hit, idx, np.iinfo(np.int64).max), off)

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


def served_run(pool, nreq, rate, seed, cfg):
    """Poisson streams of 4-level tile requests: continuous batching on the
    FAST kernel vs the plain FP64 kernel serving one request at a time.
    Run twice: at `rate` (latency under the stated load) and with every
    request arriving within 0.25 s (saturation: throughput = capacity)."""
    from setsm_amd import smx
    from setsm_amd.serve import Server, make_stream
    out = dict(requests=nreq, levels=4, budget_gib=200, sizes_cells=[96, 640], heights=[32, 256])
    for tag, r in (("load", rate), ("saturated", nreq / 0.25)):
        stream = make_stream(nreq, r, sizes=(96, 640), heights=(32, 256), n_scenes=len(pool), seed=seed + 9)
        Server(pool, levels=4).prepare(stream)          # request payloads, built before the clock starts
        _, fast = Server(pool, levels=4, fast_config=cfg).run(stream)
        _, plain = Server(pool, levels=4, precision=smx.PREC_FP64, batching=False).run(stream)
        out[tag] = dict(rate_per_s=r, fast=fast, plain_fp64=plain,
                        speedup_vs_plain=fast["hyp_per_s"] / plain["hyp_per_s"])
    return out


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per FAST launch at this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated); the
    counters cannot be read from inside the timed run."""
    path = os.path.join(ROOT, "profiles", "vll_c1_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--image-px", type=int, default=16384)
    ap.add_argument("--cells", type=int, default=4000)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-rows", type=int, default=64, help="cell rows of the CPU baseline stripe")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--autotune", action="store_true", help="time every FAST launch shape and store the best")
    ap.add_argument("--tuning-out", default=None,
                    help="file the --autotune table is written to (default: the package's tuning.json)")
    ap.add_argument("--scene-cache", default="", help="directory to load / save the rendered scene")
    ap.add_argument("--batch-tiles", type=int, default=16, help="tiles of the batched run (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=80, help="requests of the served run (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--copy-calib-gib", type=float, default=0.0,
                    help="first run one device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the VerticalLineLocus path shards by whole tiles with no "
              "collective, and multi-GPU runs are not part of this build (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    from setsm_amd import smx
    from setsm_amd.records import LevelInputs
    from setsm_amd.synth import GridSpec, SceneSpec, make_grid, make_scene

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
    t0 = time.time()
    spec = SceneSpec(image_px=args.image_px, gsd=0.5, relief=300.0, base=200.0, seed=args.seed)
    sc = load_scene(args.scene_cache, spec) if args.scene_cache else None
    if sc is None:
        sc = make_scene(spec, device="cuda")
        torch.cuda.empty_cache()
        if args.scene_cache:
            save_scene(args.scene_cache, sc)
    log(f"scene {args.image_px}^2 x2 ready in {time.time() - t0:.1f}s, RPC fit residual {max(sc.fit_residual_px):.2e} px")

    side = args.image_px * spec.gsd
    grid_res = 2.0
    cells = args.cells
    inset = (side - cells * grid_res) / 2.0
    g = make_grid(sc, GridSpec(cells=cells, grid_res=grid_res, band=32.0, inset=inset, x_offset=0.0,
                               y_offset=0.0, seed=args.seed + 1))
    g["boundary"] = np.array([inset, inset, inset + cells * grid_res, inset + cells * grid_res])
    g["shape"] = (cells, cells)
    log(f"grid {cells}^2 cells ready at {time.time() - t0:.1f}s")

    job = smx.Job(1)
    for ti in range(2):
        job.build_pyramid(ti, sc.images[ti])
    pyr_stub = [[(sc.images[ti], sc.images[ti], np.zeros((1, 1), np.uint8))] for ti in range(2)]
    li = LevelInputs(pyr_stub, 0, sc.rpcs, g, grid_resolution=grid_res)
    job.set_level(li.level)
    job.set_searchspace(g["ugrid"], g["minmax"])
    job.initialize_voxel(1)
    nh = job.num_hypotheses
    log(f"job resident: {job.device_bytes / 2**30:.1f} GiB, {nh} hypotheses")

    half = 5
    fph = flops_per_hypothesis(half)

    def run(prec, steps, warmup, keep_rho=False):
        for _ in range(warmup):
            job.vertical_line_locus(1, precision=prec, keep_rho=keep_rho)
        torch.cuda.synchronize()
        job.sync()
        kms = []
        t = time.perf_counter()
        for _ in range(steps):
            acc = job.vertical_line_locus(1, precision=prec, keep_rho=keep_rho)
            kms.append(job.last_kernel_ms)
        job.sync()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t) / steps
        return dict(acc=acc, evaluated=job.last_evaluated, wall_s=wall, kernel_ms=float(np.mean(kms)))

    # plain FP64 kernel (bit-identical to the CPU path)
    if args.fp64_steps > 0:
        log("timing plain FP64 kernel")
        p64 = run(smx.PREC_FP64, args.fp64_steps, 1, keep_rho=True)
        log(f"fp64: {p64['evaluated']} hyps, {p64['wall_s'] * 1e3:.1f} ms/step, kernel {p64['kernel_ms']:.1f} ms")
    else:
        job.vertical_line_locus(1, precision=smx.PREC_FP64, keep_rho=True)   # reference values for the error check
        p64 = None

    # stripe of cells checked against the CPU restatement and timed there
    stripe_rows = max(1, args.cpu_rows)
    r0 = cells // 2
    c0, c1 = r0 * cells, (r0 + stripe_rows) * cells
    off = np.zeros(cells * cells + 1, np.int64)
    ncc_all = job.get_nccresult()
    lens = ncc_all["NumOfHeight"].astype(np.int64)
    nw = int(lens[c0:c1].sum())
    w64, rho64 = job.get_voxel_rows(c0, c1, nw, with_rho=True)

    fast = None
    fast_err = None
    tuned = None
    try:
        from setsm_amd import autotune
        key = f"vll_L0_{args.image_px}px_{cells}cells_h128"
        if args.autotune:
            best, rows = autotune.tune_fast(job, 1, log=log)
            out_path = args.tuning_out or autotune.TUNING_PATH
            os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
            autotune.save(key, best, rows, path=out_path)
            tuned = best
        else:
            entry = autotune.load(key)
            if entry:
                job.set_fast_config(entry["best"]["config"])
                tuned = entry["best"]
        log(f"FAST launch shape: {tuned}")
        log("timing FAST kernel")
        fast = run(smx.PREC_FAST, args.steps, args.warmup, keep_rho=False)
        fast["stats"] = job.last_stats
        job.vertical_line_locus(1, precision=smx.PREC_FAST, keep_rho=True)
        wf, rhof = job.get_voxel_rows(c0, c1, nw, with_rho=True)
        ev = np.isfinite(rho64)
        d = np.abs(rhof[ev].astype(np.float64) - rho64[ev].astype(np.float64))
        fast_err = dict(max_abs=float(d.max()) if d.size else 0.0, rms=float(np.sqrt(np.mean(d * d))) if d.size else 0.0,
                        incc_diff=int(np.sum((wf >> 16) != (w64 >> 16))), n=int(ev.sum()))
        fast_err.update(peak_heights(rho64, rhof, lens[c0:c1], li.level.height_step))
        log(f"fast: {fast['wall_s'] * 1e3:.1f} ms/step, kernel {fast['kernel_ms']:.1f} ms, err {fast_err}")
    except smx.SmxError as e:
        log(f"FAST path unavailable: {e}")

    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import oracle_py as O
        img0 = job.get_view(0, 0, sc.images[0].shape)
        img1 = job.get_view(1, 0, sc.images[1].shape)
        sub = {k: g[k][c0:c1] for k in ("grid_pts", "grid_wgs", "ugrid", "truth")}
        sub["boundary"] = g["boundary"]
        sub["shape"] = (stripe_rows, cells)
        lio = LevelInputs([[img0], [img1]], 0, sc.rpcs, sub, grid_resolution=grid_res)
        st = O.OracleState(lio, keep_rho=False)
        ug = sub["ugrid"].copy()
        st.initialize_voxel(ug, 1, g["minmax"])
        threads = O.threads()
        t = time.perf_counter()
        acc_cpu = st.vertical_line_locus(ug, 1, g["minmax"])
        cpu_s = time.perf_counter() - t
        wo, _ = st.export()
        hyps_cpu = int(st.ncc["NumOfHeight"].astype(np.int64).sum())
        cpu = dict(value=hyps_cpu / cpu_s, unit="height-hypotheses/s", cores=threads, kind="port",
                   sample=f"{stripe_rows} x {cells} cell stripe (rows {r0}..{r0 + stripe_rows - 1}), {hyps_cpu} hypotheses, "
                          f"oracle FP64 OpenMP dynamic,1 over cells, {cpu_s:.1f}s")
        parity = dict(cells=c1 - c0, words=int(len(wo)), fp64_words_equal=bool(np.array_equal(wo, w64)),
                      accessible_cpu=int(acc_cpu))
        log(f"cpu: {cpu['value']:.3e} hyp/s on {threads} threads; fp64 parity {parity}")

    if fast is None and p64 is None:
        raise SystemExit("no timed path (FAST failed and --fp64-steps 0)")
    best = fast if fast is not None else p64
    best_prec = "fast" if fast is not None else "fp64"
    value = best["evaluated"] / best["wall_s"]
    kernel_s = best["kernel_ms"] / 1e3
    if best_prec == "fast":
        # the FAST window kernel alone: the hypotheses of the runs it handed to
        # the FP64 pass, and that pass's time, are not its own
        redo = best.get("stats", {})
        log(f"FP64 pass over handed-over runs: {redo.get('redo_runs')} runs, {redo.get('redo_hyps')} hypotheses, "
            f"{redo.get('redo_ms', 0.0):.2f} ms")
        achieved = (best["evaluated"] - redo.get("redo_hyps", 0)) * fph / kernel_s / 1e12
        roof = dict(bound="valu", achieved=achieved, peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP32_TFLOPS,
                    traffic=measured_traffic() if (args.image_px, cells) == (16384, 4000) else None)
    else:
        achieved = best["evaluated"] * fph / kernel_s / 1e12
        roof = dict(bound="valu", achieved=achieved, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP64_TFLOPS, traffic=None)
    out = {
        "metric": "height-hypotheses/s (VerticalLineLocus, one level, 16384^2 pair, 4000^2 cells x 128 heights)",
        "value": value,
        "unit": "height-hypotheses/s",
        "n_gpus": 1,
        "steps": args.steps if fast is not None else args.fp64_steps,
        "warmup": args.warmup,
        "ms_per_step": best["wall_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-compensated+f64-rpc" if best_prec == "fast" else "f64",
        "data": "synthetic (seeded fBm terrain, rendered stereo pair, fitted RPCs)",
        "config": {"workload": f"VLL level 0: {args.image_px}^2 uint16 pair, {cells}^2 cells at 2 m, 128 heights/cell",
                   "precision": best_prec, "hypotheses_per_step": best["evaluated"], "flops_per_hypothesis": fph},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_fp64": ({"value": p64["evaluated"] / p64["wall_s"], "ms_per_step": p64["wall_s"] * 1e3,
                        "kernel_ms": p64["kernel_ms"],
                        "frac_fp64_peak": p64["evaluated"] * fph / (p64["kernel_ms"] / 1e3) / 1e12 / PEAK_FP64_TFLOPS}
                       if p64 else None),
        "speedup_vs_plain_fp64": ((value / (p64["evaluated"] / p64["wall_s"])) if fast is not None else 1.0)
                                 if p64 else None,
        "fast_vs_fp64_error": fast_err,
        "parity_stripe": parity,
        "fast_launch": (dict(tuned or {}, **fast.get("stats", {})) if fast else None),
        "kernel_ms": best["kernel_ms"],
    }
    job.close()
    del job
    torch.cuda.empty_cache()
    if args.batch_tiles > 0 or args.serve_requests > 0:
        from setsm_amd.serve import make_pool
        pool = make_pool(n_scenes=4, image_px=3072, device="cuda", seed=100 + args.seed)
        torch.cuda.empty_cache()
        # mixed small tiles (96..1024 cells per side) have their own tuned shape:
        # the level-0 one is sized for 4000^2 cells and leaves these launches
        # with too few, too large tiles
        from setsm_amd import autotune
        mixed_key = "vll_mixed_tiles_3072px"
        entry = autotune.load(mixed_key)
        mixed_cfg = entry["best"]["config"] if entry else None
        if args.batch_tiles > 0:
            out["batched"] = batched_run(pool, args.batch_tiles, args.seed, mixed_cfg,
                                         tune_key=mixed_key if args.autotune else None, tuning_out=args.tuning_out)
            mixed_cfg = out["batched"]["fast_config"]
            log(f"batched: {out['batched']}")
        if args.serve_requests > 0:
            out["served"] = served_run(pool, args.serve_requests, args.serve_rate, args.seed, mixed_cfg)
            log(f"served: {out['served']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
