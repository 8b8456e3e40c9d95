"""Continuous-batching server for tile-matching requests.

A request is one tile: a stereo pair, its two RPCs and a cell grid. It runs
SETSM's coarse-to-fine matching over its pyramid levels (Matching_SETSM's
level loop, setsm_code.cpp:3468-4603): per level InitializeVoxel,
VerticalLineLocus and AWNCC. Its pyramid, search space, NCCresult array and
voxel cube stay resident in HBM (one smx job) from admission to completion.

Scheduling:
  * requests are admitted in arrival order while the HBM budget holds their
    projected footprint (pyramid + largest level's cube and cell arrays);
  * every scheduling step takes all admitted requests, groups them by the
    window size of their next level, and runs ONE batched VerticalLineLocus
    per group (a persistent kernel over all their tiles), so requests that
    arrive while others are on the GPU join the next step;
  * each request's latency is completion time minus arrival time.

The plain baseline runs the same stream with the FP64 kernel, one request and
one launch at a time (no batching).

Between levels the host TIN / SetHeightRange step is not part of this build;
each level's band is drawn around the true terrain height (the same
synthetic search space the benchmark uses), so every level does real work.
"""
from __future__ import annotations

import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import smx
from .records import LevelInputs
from .synth import Scene, SceneSpec, level_grid, make_scene


@dataclass
class Request:
    rid: int
    arrival: float        # seconds from the start of the stream
    scene: int            # index into the scene pool
    cells: int            # level-0 cells per side
    nheights: int         # hypotheses per cell at every level
    seed: int
    plan: list = None     # per-level grid / search space payload (prepare())


@dataclass
class _Active:
    req: Request
    job: smx.Job
    plan: list            # [(level, LevelInputs, grid)] coarse to fine
    stage: int = 0
    admitted: float = 0.0
    evaluated: int = 0
    reserved: int = 0
    results: list = field(default_factory=list)


def make_stream(n, rate, sizes=(96, 640), heights=(32, 256), n_scenes=4, seed=0):
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
    def __init__(self, pool: list, levels=4, hbm_budget=200 << 30, precision=smx.PREC_FAST, batching=True,
                 fast_config=None):
        self.pool = pool
        self.levels = levels
        self.budget = hbm_budget
        self.precision = precision
        self.batching = batching
        self.fast_config = fast_config
        self.in_use = 0
        self.idle_jobs = []   # finished jobs, reused by later requests (their HBM stays with the process)

    # ------------------------------------------------------------ admit
    def _plan(self, req: Request):
        sc: Scene = self.pool[req.scene]
        plan = []
        side = sc.spec.image_px * sc.spec.gsd
        for lv in range(self.levels - 1, -1, -1):
            grid_res = max(2.0, 0.5 * (1 << lv))
            cells = max(8, int(req.cells * 2.0 / grid_res))
            cells = min(cells, int((side - 2 * 96.0 * (1 << lv) * 0.5) / grid_res))
            g = level_grid(sc, lv, cells, req.nheights, seed=req.seed % 100000)
            plan.append((lv, g))
        return plan

    def _projected_bytes(self, req: Request):
        sc = self.pool[req.scene]
        pyr = 2 * sc.spec.image_px ** 2 * 5 * 4 // 3
        cube = req.cells ** 2 * req.nheights * 4
        cells = req.cells ** 2 * 96
        return pyr + 2 * cube + cells

    def prepare(self, stream: list):
        """Attach each request's payload (per-level cell grid, lon/lat,
        search space), as a client would send it with the request."""
        for req in stream:
            if req.plan is not None:
                continue
            sc: Scene = self.pool[req.scene]
            plan = []
            for lv, g in self._plan(req):
                pyr_stub = [[(sc.images[ti], sc.images[ti], np.zeros((1, 1), np.uint8))] * self.levels
                            for ti in range(2)]
                li = LevelInputs(pyr_stub, lv, sc.rpcs, g, grid_resolution=g["grid_res"])
                plan.append((lv, li, g))
            req.plan = plan
        return stream

    def _admit(self, req: Request, now: float, job=None):
        sc: Scene = self.pool[req.scene]
        if job is None:
            job = smx.Job(self.levels)
            if self.fast_config is not None:
                job.set_fast_config(self.fast_config)
        for ti in range(2):
            job.build_pyramid(ti, sc.images[ti])
        a = _Active(req, job, req.plan, admitted=now)
        a.reserved = self._projected_bytes(req)
        return a

    # ------------------------------------------------------------- step
    def _begin_level(self, a: _Active):
        lv, li, g = a.plan[a.stage]
        a.job.set_level(li.level)
        a.job.set_searchspace(g["ugrid"], g["minmax"])
        a.job.initialize_voxel(1)
        return lv

    def _end_level(self, a: _Active):
        a.job.awncc(1)
        a.evaluated += a.job.last_evaluated
        a.stage += 1

    def run(self, stream: list, log=None, workers=8):
        """Serve `stream`; returns per-request records and aggregate stats."""
        self.prepare(stream)
        with ThreadPoolExecutor(max_workers=workers) as ex:
            return self._run(stream, ex, log)

    def _run(self, stream, ex, log):
        pending = sorted(stream, key=lambda r: r.arrival)
        active: list = []
        done = []
        t0 = time.perf_counter()
        gpu_busy = 0.0
        phase = dict(admit_s=0.0, begin_s=0.0, end_s=0.0)
        while pending or active:
            now = time.perf_counter() - t0
            admit = []
            while pending and pending[0].arrival <= now:
                if not self.batching and (active or admit):
                    break        # the plain baseline serves one request at a time
                need = self._projected_bytes(pending[0])
                if (active or admit) and self.in_use + need > self.budget:
                    break
                self.in_use += need
                admit.append(pending.pop(0))
            reuse = [self.idle_jobs.pop() if self.idle_jobs else None for _ in admit]
            tk = time.perf_counter()
            active += list(ex.map(lambda rj: self._admit(rj[0], now, rj[1]), zip(admit, reuse)))
            phase["admit_s"] += time.perf_counter() - tk
            if not active:
                time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                continue
            step = active if self.batching else active[:1]
            groups = {}
            tk = time.perf_counter()
            for a, lv in zip(step, ex.map(self._begin_level, step)):
                groups.setdefault(a.job.half, []).append(a)   # the library's Half_template_size
            phase["begin_s"] += time.perf_counter() - tk
            for half, grp in groups.items():
                tk = time.perf_counter()
                if self.batching and self.precision == smx.PREC_FAST:
                    smx.vertical_line_locus_batch([a.job for a in grp], 1, precision=self.precision)
                elif self.batching:
                    list(ex.map(lambda a: a.job.vertical_line_locus(1, precision=self.precision), grp))
                else:
                    for a in grp:
                        a.job.vertical_line_locus(1, precision=self.precision)
                gpu_busy += time.perf_counter() - tk
            tk = time.perf_counter()
            list(ex.map(self._end_level, step))
            phase["end_s"] += time.perf_counter() - tk
            t = time.perf_counter() - t0
            for a in [a for a in step if a.stage == len(a.plan)]:
                a.results = a.job.get_nccresult()
                done.append(dict(rid=a.req.rid, arrival=a.req.arrival, admitted=a.admitted, done=t,
                                 latency=t - a.req.arrival, evaluated=a.evaluated, cells=a.req.cells,
                                 nheights=a.req.nheights, ncc=a.results))
                self.idle_jobs.append(a.job)
                self.in_use -= a.reserved
                active.remove(a)
                if log:
                    log(f"request {a.req.rid} done: {a.req.cells}^2 cells x {a.req.nheights} h, "
                        f"latency {1e3 * (t - a.req.arrival):.1f} ms")
        wall = time.perf_counter() - t0
        for j in self.idle_jobs:
            j.close()
        self.idle_jobs = []
        lat = np.array([d["latency"] for d in done])
        hyps = sum(d["evaluated"] for d in done)
        return done, dict(requests=len(done), wall_s=wall, hyps=hyps, hyp_per_s=hyps / wall,
                          p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                          gpu_call_s=gpu_busy, **phase)
