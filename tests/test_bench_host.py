"""Host-side logic of bench.py that needs no GPU: the per-cell peak-height
comparison of the FAST and FP64 cubes, against a per-cell loop."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _loop(rho64, rhof, lens, step, tol):
    off = np.concatenate([[0], np.cumsum(lens)])
    dh, ties, flips, cells = [], 0, 0, 0
    for c in range(len(lens)):
        a, b = off[c], off[c + 1]
        if b <= a or not np.isfinite(rho64[a:b]).any():
            continue
        cells += 1
        ro = np.where(np.isfinite(rho64[a:b]), rho64[a:b], -np.inf)
        rf = np.where(np.isfinite(rhof[a:b]), rhof[a:b], -np.inf)
        io = next(i for i in range(b - a) if ro[i] == ro.max())
        jf = next(i for i in range(b - a) if rf[i] == rf.max())
        dh.append((jf - io) * step)
        if io != jf:
            if ro[io] - ro[jf] < tol:
                ties += 1
            else:
                flips += 1
    dh = np.asarray(dh)
    return dict(peak_cells=cells, peak_height_rms_m=float(np.sqrt(np.mean(dh * dh))) if dh.size else 0.0,
                peak_near_ties=ties, peak_flips=flips)


def test_peak_heights_matches_loop():
    b = _bench()
    rng = np.random.default_rng(4)
    lens = rng.integers(0, 40, 300)
    n = int(lens.sum())
    rho64 = rng.uniform(-1, 1, n)
    rho64[rng.random(n) < 0.1] = np.nan            # not evaluated
    off = np.concatenate([[0], np.cumsum(lens)])
    for c in rng.choice(300, 20, replace=False):   # cells with nothing evaluated
        rho64[off[c]:off[c + 1]] = np.nan
    rhof = rho64 + rng.normal(0, 2e-6, n)          # FAST-like perturbation: some near-ties flip
    for c in rng.choice(300, 5, replace=False):    # a few real flips
        if lens[c] > 1:
            rhof[off[c]] = 2.0
    rhof = rhof.astype(np.float32)
    got = b.peak_heights(rho64, rhof, lens, 0.25)
    want = _loop(rho64, rhof.astype(np.float64), lens, 0.25, 1e-5)
    assert got["peak_cells"] == want["peak_cells"]
    assert got["peak_near_ties"] == want["peak_near_ties"]
    assert got["peak_flips"] == want["peak_flips"]
    assert abs(got["peak_height_rms_m"] - want["peak_height_rms_m"]) < 1e-12
    assert want["peak_flips"] > 0


def test_gpus_above_one_exits():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
