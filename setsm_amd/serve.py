"""Continuous-batching server for tile-matching requests.

A request is one tile: a stereo pair, its two RPCs and a cell grid. It runs
SETSM's coarse-to-fine matching over its pyramid levels and, per level,
`iterations` iterations (Matching_SETSM's level and iteration loops,
setsm_code.cpp:3468-4603, 3774): InitializeVoxel, VerticalLineLocus and
AWNCC. Even iterations above level 0 add the level-below term
(setsm_code.cpp:6087-6088). The request's pyramid, search space, NCCresult
array and voxel cube stay resident in HBM (one smx job) from admission to
completion.

Grids follow SETSM's SetGrids spacings (setsm_code.cpp:5074-5198; with
0.5 m imagery and a 2 m DEM: 2, 3, 6, 8, 8 m at levels 0..4), each level's
grid covering the tile's extent.

Scheduling:
  * requests are admitted in arrival order while the HBM budget holds their
    projected footprint (pyramid and its build, the largest level's cell
    arrays, run lists, cube, the cube InitializeVoxel builds beside it and
    the level term of a combined iteration); each job is capped at its
    reservation (smx_job_set_budget), so a request that would exceed it
    fails alone; an admitted request's pyramid is built by admission workers
    on its own stream while the requests already admitted keep stepping;
  * every scheduling step takes all admitted requests, groups them by the
    window size (smx_job_half) and iteration of their next step, and runs ONE
    batched VerticalLineLocus per group (persistent kernels over all their
    tiles), so requests that arrive while others are on the GPU join the
    next step;
  * a request whose call fails (e.g. an allocation beyond its reservation)
    is rejected and released; the loop goes on with the others;
  * latency = completion - arrival = queue wait (admission - arrival) +
    service (completion - admission).

The plain baseline runs the same stream with the FP64 kernel, one request
and one launch at a time.

Between iterations SETSM's host step (SelectMPs, blunder detection, TIN,
SetHeightRange) sets the next search space. Here a seeded stand-in does:
a quarter of the cells widen their band by 1-4 m each side, a fifth narrow
it, the rest keep it, and at levels <= 1 a few lose their match flag; the
first level's bands come from the true surface of the synthetic scene.
"""
from __future__ import annotations

import math
import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field

import numpy as np

from . import smx
from .records import LevelInputs, flops_per_hypothesis, height_step
from .synth import GridSpec, Scene, SceneSpec, make_grid, make_scene


def setgrids_spacing(level, resolution=0.5, dem_resolution=2.0):
    """Grid spacing of a level as SetGrids computes it (setsm_code.cpp:
    5074-5198, imagery resolution >= 0.4 m, no full calculation)."""
    py = resolution * (1 << level)
    if py > 32:
        py = float(int(py / 4.0))
    elif resolution >= 0.4:
        if level > 0:
            if py * 3 > dem_resolution:
                if dem_resolution > 8:
                    py = dem_resolution
                else:
                    py = 8.0 if py * 3 > 8 else py * 3
            else:
                py = dem_resolution
        elif dem_resolution >= 2:
            py = dem_resolution
    return max(py, dem_resolution)


def next_searchspace(ug, seed, level):
    """Seeded stand-in for SETSM's host step between iterations: widen a
    quarter of the bands by 1-4 m each side, narrow a fifth by 2 m each side,
    keep the rest; at levels <= 1 clear 3% of the match flags."""
    rng = np.random.default_rng(seed)
    out = ug.copy()
    n = out.shape[0]
    pick = rng.random(n)
    lo = out["minHeight"].astype(np.int64)
    hi = out["maxHeight"].astype(np.int64)
    widen = pick < 0.25
    lo[widen] -= rng.integers(1, 5, widen.sum())
    hi[widen] += rng.integers(1, 5, widen.sum())
    narrow = (pick >= 0.25) & (pick < 0.45) & (hi - lo > 4)
    lo[narrow] += 2
    hi[narrow] -= 2
    out["minHeight"] = lo
    out["maxHeight"] = hi
    if level <= 1:
        out["Matched_flag"][rng.random(n) < 0.03] = 0
    return out


@dataclass
class Request:
    rid: int
    arrival: float        # seconds from the start of the stream
    scene: int            # index into the scene pool
    cells: int            # level-0 cells per side (2 m)
    nheights: int         # hypotheses per cell at every level
    seed: int
    plan: list = None     # [(level, iteration, LevelInputs, grid, ugrid)] (prepare())


@dataclass
class _Active:
    req: Request
    job: smx.Job
    stage: int = 0
    admitted: float = 0.0
    evaluated: int = 0
    flops: float = 0.0    # algorithmic flops of the evaluated hypotheses
    reserved: int = 0
    results: list = field(default_factory=list)


def make_stream(n, rate, sizes=(512, 4096), heights=(32, 256), n_scenes=4, seed=0):
    """Poisson arrivals; side log-uniform in `sizes`, heights uniform."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    lo, hi = math.log(sizes[0]), math.log(sizes[1])
    return [Request(i, float(t[i]), int(rng.integers(0, n_scenes)), int(round(math.exp(rng.uniform(lo, hi)))),
                    int(rng.integers(heights[0], heights[1] + 1)), int(rng.integers(0, 1 << 30)))
            for i in range(n)]


def make_pool(n_scenes=4, image_px=3072, device="cuda", seed=100):
    """Independently seeded scenes: relief 50-1500 m, convergence 15-40 deg."""
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(n_scenes):
        relief = float(rng.uniform(50, 1500))
        conv = float(rng.uniform(15, 40))
        spec = SceneSpec(image_px=image_px, relief=relief, base=200.0, view_deg=(conv / 2, -conv / 2),
                         seed=int(seed + k))
        pool.append(make_scene(spec, device=device))
    return pool


class Server:
    def __init__(self, pool: list, levels=4, iterations=2, hbm_budget=200 << 30, precision=smx.PREC_FAST,
                 batching=True, fast_config=None, resolution=0.5, dem_resolution=2.0):
        self.pool = pool
        self.levels = levels
        self.iterations = iterations
        self.budget = hbm_budget
        self.precision = precision
        self.batching = batching
        self.fast_config = fast_config
        self.resolution = resolution
        self.dem_resolution = dem_resolution
        self.in_use = 0

    # ------------------------------------------------------------ plan
    def _level_grid(self, sc: Scene, req: Request, lv: int):
        """The level's grid over the request's extent (the level-0 grid of
        req.cells 2 m cells, centred in the scene), bands of req.nheights
        heights at the level's step around the true surface."""
        spacing = setgrids_spacing(lv, self.resolution, self.dem_resolution)
        extent = req.cells * setgrids_spacing(0, self.resolution, self.dem_resolution)
        cells = max(4, int(math.ceil(extent / spacing)))
        side = sc.spec.image_px * sc.spec.gsd
        inset = (side - cells * spacing) / 2.0
        hs = height_step(lv, self.resolution)
        band = float(max(1, int(round(req.nheights * hs))))
        g = make_grid(sc, GridSpec(cells=cells, grid_res=spacing, band=band, inset=inset, x_offset=0.0,
                                   y_offset=0.0, seed=(req.seed + lv) % 1000003))
        g["boundary"] = np.array([inset, inset, inset + cells * spacing, inset + cells * spacing])
        g["shape"] = (cells, cells)
        g["grid_res"] = spacing
        return g

    def _projected_bytes(self, req: Request):
        """HBM the request may hold at its peak: pyramid (and its build
        buffers) plus its largest level."""
        sc = self.pool[req.scene]
        px = sc.images[0].size
        pyr = 2 * px * 5 * 4 // 3 + px * 9
        worst = 0
        for lv in range(self.levels):
            spacing = setgrids_spacing(lv, self.resolution, self.dem_resolution)
            cells = int(math.ceil(req.cells * 2.0 / spacing)) ** 2
            hs = height_step(lv, self.resolution)
            nh = int(round(req.nheights * hs) / hs) + 8
            runs = (nh + 63) // 64
            # cell arrays, run lists, cube + the cube InitializeVoxel builds beside
            # it, the level term of a combined iteration (8 B per hypothesis)
            worst = max(worst, cells * (128 + 24 * runs) + cells * nh * (4 + 4 + 8))
        return int((pyr + worst) * 1.05) + (64 << 20)

    def prepare(self, stream: list):
        """Attach each request's payload: per level and iteration the cell
        grid, lon/lat and search space, as a client would send them."""
        for req in stream:
            if req.plan is not None:
                continue
            sc: Scene = self.pool[req.scene]
            plan = []
            for lv in range(self.levels - 1, -1, -1):
                g = self._level_grid(sc, req, lv)
                pyr_stub = [[(sc.images[ti], sc.images[ti], np.zeros((1, 1), np.uint8))] * self.levels
                            for ti in range(2)]
                li = LevelInputs(pyr_stub, lv, sc.rpcs, g, grid_resolution=g["grid_res"],
                                 resolution=self.resolution, dem_resolution=self.dem_resolution)
                ug = g["ugrid"]
                for it in range(1, self.iterations + 1):
                    if it > 1:
                        ug = next_searchspace(ug, req.seed + 97 * it + lv, lv)
                    plan.append((lv, it, li, g, ug))
            req.plan = plan
        return stream

    # ------------------------------------------------------------ steps
    def _admit(self, req: Request, now: float):
        sc: Scene = self.pool[req.scene]
        a = _Active(req, None, admitted=now, reserved=self._projected_bytes(req))
        job = smx.Job(self.levels, budget_bytes=a.reserved)
        a.job = job
        try:
            if self.fast_config is not None:
                job.set_fast_config(self.fast_config)
            for ti in range(2):
                job.build_pyramid(ti, sc.images[ti])
        except smx.SmxError:
            job.close()
            raise
        return a

    def _begin(self, a: _Active):
        lv, it, li, g, ug = a.req.plan[a.stage]
        if it == 1:
            a.job.set_level(li.level)
        a.job.set_searchspace(ug, g["minmax"])
        a.job.initialize_voxel(it)
        return (a.job.half, it)

    def _end(self, a: _Active):
        _, it, _, _, _ = a.req.plan[a.stage]
        a.job.awncc(it)
        a.evaluated += a.job.last_evaluated
        a.flops += a.job.last_evaluated * flops_per_hypothesis(a.job.half)
        a.stage += 1

    def run(self, stream: list, log=None, workers=8):
        """Serve `stream`; returns per-request records and aggregate stats."""
        self.prepare(stream)
        with ThreadPoolExecutor(max_workers=workers) as ex, ThreadPoolExecutor(max_workers=2) as ax:
            return self._run(stream, ex, ax, log)

    def _admit_guarded(self, req: Request, now: float):
        try:
            return True, self._admit(req, now)
        except smx.SmxError as e:
            return False, str(e)

    def _try(self, ex, fn, items):
        """fn over items in the pool; returns (results, failed items)."""
        def guard(x):
            try:
                return True, fn(x)
            except smx.SmxError as e:
                return False, str(e)
        out = list(ex.map(guard, items))
        return [(x, r) for x, (ok, r) in zip(items, out) if ok], [(x, r) for x, (ok, r) in zip(items, out) if not ok]

    def _run(self, stream, ex, ax, log):
        pending = sorted(stream, key=lambda r: r.arrival)
        active: list = []
        admitting: list = []   # (future, request, reserved bytes): pyramid builds in flight
        done, rejected = [], []
        t0 = time.perf_counter()
        gpu_busy = 0.0
        phase = dict(admit_s=0.0, begin_s=0.0, end_s=0.0)

        def reject(a, why, now):
            rejected.append(dict(rid=a.req.rid, arrival=a.req.arrival, at=now, error=why))
            if a.job is not None:
                a.job.close()
            self.in_use -= a.reserved
            if a in active:
                active.remove(a)
            if log:
                log(f"request {a.req.rid} rejected: {why}")

        while pending or active or admitting:
            now = time.perf_counter() - t0
            # admissions: the request's pyramid is built on its own stream by
            # the admission workers while the admitted requests keep stepping
            while pending and pending[0].arrival <= now:
                if not self.batching and (active or admitting):
                    break        # the plain baseline serves one request at a time
                need = self._projected_bytes(pending[0])
                if need > self.budget:
                    req = pending.pop(0)
                    rejected.append(dict(rid=req.rid, arrival=req.arrival, at=now,
                                         error=f"needs {need} bytes, budget {self.budget}"))
                    continue
                if self.in_use + need > self.budget:
                    break
                self.in_use += need
                req = pending.pop(0)
                admitting.append((ax.submit(self._admit_guarded, req, now), req, need))
            tk = time.perf_counter()
            if admitting and not active:
                wait([f for f, _, _ in admitting], return_when=FIRST_COMPLETED)
            still = []
            for f, req, need in admitting:
                if not f.done():
                    still.append((f, req, need))
                    continue
                ok, a = f.result()
                if ok:
                    active.append(a)
                else:
                    self.in_use -= need
                    rejected.append(dict(rid=req.rid, arrival=req.arrival, at=now, error=a))
            admitting = still
            phase["admit_s"] += time.perf_counter() - tk
            if not active:
                if pending and not admitting:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                continue
            step = active if self.batching else active[:1]
            tk = time.perf_counter()
            ok, bad = self._try(ex, self._begin, step)
            for a, why in bad:
                reject(a, why, time.perf_counter() - t0)
            groups = {}
            for a, key in ok:
                groups.setdefault(key, []).append(a)
            phase["begin_s"] += time.perf_counter() - tk
            ran = []
            for (half, it), grp in groups.items():
                tk = time.perf_counter()
                try:
                    if self.batching and self.precision == smx.PREC_FAST and len(grp) > 1:
                        smx.vertical_line_locus_batch([a.job for a in grp], it, precision=self.precision)
                    else:
                        for a in grp:
                            a.job.vertical_line_locus(it, precision=self.precision)
                    ran += grp
                except smx.SmxError:
                    # find the request(s) that fail alone; the rest run
                    for a in grp:
                        try:
                            a.job.vertical_line_locus(it, precision=self.precision)
                            ran.append(a)
                        except smx.SmxError as e:
                            reject(a, str(e), time.perf_counter() - t0)
                gpu_busy += time.perf_counter() - tk
            tk = time.perf_counter()
            ok, bad = self._try(ex, self._end, ran)
            for a, why in bad:
                reject(a, why, time.perf_counter() - t0)
            phase["end_s"] += time.perf_counter() - tk
            t = time.perf_counter() - t0
            for a in [a for a, _ in ok if a.stage == len(a.req.plan)]:
                a.results = a.job.get_nccresult()
                done.append(dict(rid=a.req.rid, arrival=a.req.arrival, admitted=a.admitted, done=t,
                                 latency=t - a.req.arrival, queue_s=a.admitted - a.req.arrival,
                                 service_s=t - a.admitted, evaluated=a.evaluated, flops=a.flops, cells=a.req.cells,
                                 nheights=a.req.nheights, ncc=a.results))
                a.job.close()
                self.in_use -= a.reserved
                active.remove(a)
                if log:
                    log(f"request {a.req.rid} done: {a.req.cells}^2 cells x {a.req.nheights} h, "
                        f"latency {1e3 * (t - a.req.arrival):.1f} ms")
        wall = time.perf_counter() - t0
        lat = np.array([d["latency"] for d in done]) if done else np.zeros(1)
        qw = np.array([d["queue_s"] for d in done]) if done else np.zeros(1)
        sv = np.array([d["service_s"] for d in done]) if done else np.zeros(1)
        hyps = sum(d["evaluated"] for d in done)
        flops = sum(d["flops"] for d in done)
        return done, dict(requests=len(done), rejected=len(rejected), rejections=rejected[:8], wall_s=wall,
                          hyps=hyps, hyp_per_s=hyps / wall, flops=flops,
                          p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                          queue_p50_ms=float(np.percentile(qw, 50) * 1e3),
                          queue_p99_ms=float(np.percentile(qw, 99) * 1e3),
                          service_p50_ms=float(np.percentile(sv, 50) * 1e3),
                          service_p99_ms=float(np.percentile(sv, 99) * 1e3),
                          gpu_call_s=gpu_busy, **phase)
