"""Autotuner for the SMX_PREC_FAST launch shape.

The FAST kernel's shape is the cell tile a workgroup owns (which sets the LDS
footprint, the texel reuse and the number of workgroups), the LDS reserved
for the two images' boxes, the register budget (compiler's choice, or capped
for three or four waves per SIMD) and the run length: a wavefront evaluates
64 / run_len runs of run_len consecutive heights (64, 32 or 16), so short
height lists (32..256 heights per cell, not multiples of 64) leave fewer
lanes idle at shorter runs, while longer runs keep a wave on one cell.

`tune_fast` times every shape on the job's current level and keeps the
fastest; results are stored per workload in setsm_amd/tuning.json so the
benchmark, the profiler runs and production calls use the same shape.
"""
from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_PATH = os.path.join(HERE, "tuning.json")


def tune_fast(job, iteration=1, reps=2, log=None, only_wide=False):
    from .smx import PREC_FAST, fast_configs
    rows = []
    for i, desc in fast_configs():
        if only_wide and not desc[5]:
            continue
        job.set_fast_config(i)
        job.vertical_line_locus(iteration, precision=PREC_FAST)   # warm-up
        best = None
        for _ in range(reps):
            job.vertical_line_locus(iteration, precision=PREC_FAST)
            ms = job.last_kernel_ms
            best = ms if best is None else min(best, ms)
        st = job.last_stats
        rows.append(dict(config=i, tx=desc[0], ty=desc[1], lds_texels=desc[2], occ=desc[3], stage=desc[4], wide=desc[5],
                         run_len=desc[6], kernel_ms=best, tiles_lds=st["tiles_lds"], waves_hbm=st["waves_hbm"]))
        if log:
            log(f"tune config {i} {desc}: {best:.2f} ms (tiles in LDS {st['tiles_lds']}, HBM waves {st['waves_hbm']})")
    rows.sort(key=lambda r: r["kernel_ms"])
    pick = [r for r in rows if r["wide"]] or rows   # FP64 window totals are the precision contract
    job.set_fast_config(pick[0]["config"])
    return pick[0], rows


def tune_batch(jobs, iteration=1, reps=2, log=None):
    """The same table for one persistent launch over several jobs (mixed tile
    sizes): every LDS-staged shape with FP64 window totals, timed by the
    batched call."""
    from .smx import PREC_FAST, fast_configs, vertical_line_locus_batch
    rows = []
    for i, desc in fast_configs():
        if not desc[5] or not desc[4]:
            continue
        for j in jobs:
            j.set_fast_config(i)
        vertical_line_locus_batch(jobs, iteration, precision=PREC_FAST)   # warm-up
        best = None
        for _ in range(reps):
            vertical_line_locus_batch(jobs, iteration, precision=PREC_FAST)
            ms = jobs[0].last_kernel_ms
            best = ms if best is None else min(best, ms)
        rows.append(dict(config=i, tx=desc[0], ty=desc[1], lds_texels=desc[2], occ=desc[3], stage=desc[4], wide=desc[5],
                         run_len=desc[6], kernel_ms=best))
        if log:
            log(f"batch tune config {i} {desc}: {best:.2f} ms")
    rows.sort(key=lambda r: r["kernel_ms"])
    for j in jobs:
        j.set_fast_config(rows[0]["config"])
    return rows[0], rows


def load(key):
    """Stored best shape for `key`, with its current index in the library's
    shape table (matched by shape, not by position)."""
    try:
        with open(TUNING_PATH) as f:
            entry = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    if not entry:
        return None
    from .smx import fast_configs
    b = entry["best"]
    want = (b["tx"], b["ty"], b["lds_texels"], b["occ"], b["stage"], b["wide"], b.get("run_len", 64))
    for i, desc in fast_configs():
        if tuple(desc) == want:
            entry["best"] = dict(b, config=i)
            return entry
    return None


def save(key, best, rows, path=TUNING_PATH):
    data = {}
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    data[key] = dict(best=best, table=rows)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
