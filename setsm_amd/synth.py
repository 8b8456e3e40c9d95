"""Synthetic stereo tiles for the VerticalLineLocus engine.

A seeded fBm terrain (spectral synthesis), band-limited noise albedo and
Lambertian shading are rendered through two linear-pushbroom cameras that look
along-track with opposite view angles. Each camera gets an 80-coefficient
RPC00B model least-squares fitted to it (terrain independent, denominators with
a unit constant term), in the layout SETSM reads from DigitalGlobe XML
(SubFunctions.cpp:1420-1582): offsets/scales for line, sample, lon, lat,
height, then line-num, line-den, samp-num, samp-den.

Rendering runs in torch on any device (CPU for test-sized tiles, the GPU for
16384^2 pairs); the fit runs in float64 numpy. Images are uint16 11-bit DN in
[1, 2047] (0 is SETSM's nodata, SubFunctions.cpp:3274).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

EARTH_R = 6378137.0


@dataclass
class SceneSpec:
    image_px: int = 2048          # square image side, pixels
    gsd: float = 0.5              # metres per pixel
    terrain_post: float = 1.0     # fBm grid spacing, metres
    hurst: float = 0.8
    relief: float = 300.0         # max - min terrain height, metres
    base: float = 200.0           # minimum terrain height (keeps every band endpoint != 0)
    view_deg: tuple = (12.0, -12.0)
    albedo_sigma_px: float = 4.0
    sun_elev_deg: float = 40.0
    sun_azim_deg: float = 135.0
    lat0: float = 35.0
    lon0: float = -118.0
    orbit_h: float = 600000.0     # cross-track perspective centre height
    yaw: float = 1.0e-3           # small row/column coupling of the scan line
    seed: int = 1


@dataclass
class Scene:
    spec: SceneSpec
    images: list                  # two uint16 numpy arrays (H, W)
    rpcs: list                    # two float64 numpy arrays of 90
    terrain: np.ndarray           # float32 heights on the terrain grid
    terrain_origin: tuple         # (E, N) of terrain[0, 0]
    extent: tuple                 # ground (E0, N0, E1, N1) the images cover
    fit_residual_px: list = field(default_factory=list)

    # ---------------------------------------------------------- helpers
    def to_wgs(self, E, N):
        s = self.spec
        lat = s.lat0 + np.asarray(N) / EARTH_R * 180.0 / math.pi
        lon = s.lon0 + np.asarray(E) / (EARTH_R * math.cos(math.radians(s.lat0))) * 180.0 / math.pi
        return lon, lat

    def height_at(self, E, N):
        """Bilinear terrain height at ground points (numpy float64)."""
        post = self.spec.terrain_post
        x = (np.asarray(E, dtype=np.float64) - self.terrain_origin[0]) / post
        y = (np.asarray(N, dtype=np.float64) - self.terrain_origin[1]) / post
        h, w = self.terrain.shape
        x = np.clip(x, 0, w - 1.000001)
        y = np.clip(y, 0, h - 1.000001)
        x0 = np.floor(x).astype(np.int64)
        y0 = np.floor(y).astype(np.int64)
        fx = x - x0
        fy = y - y0
        t = self.terrain
        return ((t[y0, x0] * (1 - fx) + t[y0, x0 + 1] * fx) * (1 - fy)
                + (t[y0 + 1, x0] * (1 - fx) + t[y0 + 1, x0 + 1] * fx) * fy)


def _camera(spec: SceneSpec, view_deg: float, E, N, Z, zref, ec, nc):
    """Forward linear-pushbroom model -> (row, col) in float64 numpy."""
    W = spec.image_px
    c0 = (W - 1) / 2.0
    r0 = (W - 1) / 2.0
    dz = Z - zref
    col = c0 + (E - ec) / spec.gsd * spec.orbit_h / (spec.orbit_h - dz)
    row = r0 + (nc - N) / spec.gsd - dz * math.tan(math.radians(view_deg)) / spec.gsd + spec.yaw * (E - ec) / spec.gsd
    return row, col


def _rpc_terms(L, P, H):
    # RPC00B order (CoordConversion.hpp:58-65)
    return np.stack([np.ones_like(L), L, P, H, L * P, L * H, P * H, L * L, P * P, H * H,
                     P * L * H, L ** 3, L * P * P, L * H * H, L * L * P, P ** 3, P * H * H,
                     L * L * H, P * P * H, H ** 3], axis=-1)


def fit_rpc(lon, lat, hgt, row, col):
    """Least-squares RPC00B fit (linearised, den[0] = 1). Returns 90 doubles."""
    off = np.array([row.mean(), col.mean(), lon.mean(), lat.mean(), hgt.mean()])
    sc = np.array([np.abs(row - off[0]).max(), np.abs(col - off[1]).max(),
                   np.abs(lon - off[2]).max(), np.abs(lat - off[3]).max(), np.abs(hgt - off[4]).max()])
    L = (lon - off[2]) / sc[2]
    P = (lat - off[3]) / sc[3]
    H = (hgt - off[4]) / sc[4]
    T = _rpc_terms(L, P, H)
    out = np.zeros(90)
    out[0:5] = off
    out[5:10] = sc
    for j, tgt in enumerate([(row - off[0]) / sc[0], (col - off[1]) / sc[1]]):
        A = np.concatenate([T, -tgt[:, None] * T[:, 1:]], axis=1)
        colscale = np.sqrt((A * A).sum(0)) + 1e-300
        sol, *_ = np.linalg.lstsq(A / colscale, tgt, rcond=None)
        sol = sol / colscale
        num = sol[:20]
        den = np.concatenate([[1.0], sol[20:]])
        out[10 + 40 * j: 30 + 40 * j] = num
        out[30 + 40 * j: 50 + 40 * j] = den
    return out


def eval_rpc(rpc, lon, lat, hgt):
    """Plain float64 RPC evaluation (no clamping) for fit checks."""
    off, sc = rpc[0:5], rpc[5:10]
    T = _rpc_terms((lon - off[2]) / sc[2], (lat - off[3]) / sc[3], (hgt - off[4]) / sc[4])
    c = rpc[10:].reshape(4, 20)
    line = (T @ c[0]) / (T @ c[1]) * sc[0] + off[0]
    samp = (T @ c[2]) / (T @ c[3]) * sc[1] + off[1]
    return line, samp


def _fbm(n, hurst, gen, device):
    fx = torch.fft.fftfreq(n, device=device)
    fy = torch.fft.rfftfreq(n, device=device)
    f = torch.sqrt(fx[:, None] ** 2 + fy[None, :] ** 2)
    f[0, 0] = 1.0
    amp = f ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    re = torch.randn(f.shape, generator=gen, device=device)
    im = torch.randn(f.shape, generator=gen, device=device)
    spec = torch.complex(re * amp, im * amp)
    del re, im, amp, f
    z = torch.fft.irfft2(spec, s=(n, n))
    return z


def _gauss_blur(x, sigma):
    """Separable Gaussian blur (reflect padding) by shifted adds."""
    r = int(math.ceil(3 * sigma))
    t = torch.arange(-r, r + 1, dtype=torch.float64)
    k = torch.exp(-0.5 * (t / sigma) ** 2)
    k = (k / k.sum()).tolist()
    h, w = x.shape
    xp = torch.nn.functional.pad(x[None, None], (r, r, 0, 0), mode="reflect")[0, 0]
    out = torch.zeros_like(x)
    for i, kw in enumerate(k):
        out += kw * xp[:, i:i + w]
    xp = torch.nn.functional.pad(out[None, None], (0, 0, r, r), mode="reflect")[0, 0]
    out = torch.zeros_like(x)
    for i, kw in enumerate(k):
        out += kw * xp[i:i + h, :]
    return out


def _sample(grid, origin, post, E, N):
    """Bilinear sample of a 2-D torch grid at ground coordinates (torch)."""
    h, w = grid.shape
    x = ((E - origin[0]) / post).clamp(0, w - 1.0001)
    y = ((N - origin[1]) / post).clamp(0, h - 1.0001)
    x0 = x.floor().long()
    y0 = y.floor().long()
    fx = x - x0
    fy = y - y0
    flat = grid.reshape(-1)
    i00 = y0 * w + x0
    a = flat[i00]
    b = flat[i00 + 1]
    c = flat[i00 + w]
    d = flat[i00 + w + 1]
    return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy


def make_scene(spec: SceneSpec, device="cpu") -> Scene:
    gen = torch.Generator(device=device).manual_seed(spec.seed)
    W = spec.image_px
    side = W * spec.gsd
    tan_max = max(abs(math.tan(math.radians(v))) for v in spec.view_deg)
    margin = spec.relief * tan_max + 64.0
    ext = side + 2 * margin
    n = 1 << int(math.ceil(math.log2(ext / spec.terrain_post)))
    z = _fbm(n, spec.hurst, gen, device).float()
    z = (z - z.min()) / (z.max() - z.min()) * spec.relief + spec.base
    origin = (-margin, -margin)
    post = spec.terrain_post
    # Lambertian shading from the terrain normal
    gy, gx = torch.gradient(z, spacing=post)
    el, az = math.radians(spec.sun_elev_deg), math.radians(spec.sun_azim_deg)
    s = torch.tensor([math.cos(el) * math.sin(az), math.cos(el) * math.cos(az), math.sin(el)], device=device)
    nrm = torch.rsqrt(gx * gx + gy * gy + 1.0)
    shade = ((-gx * s[0] - gy * s[1] + s[2]) * nrm).clamp(min=0.0)
    # band-limited albedo on the image-pixel ground grid
    apost = spec.gsd
    an = int(math.ceil(ext / apost)) + 2
    noise = torch.randn((an, an), generator=gen, device=device)
    alb = _gauss_blur(noise, spec.albedo_sigma_px)
    alb = (alb - alb.mean()) / alb.std()
    alb = (0.55 + 0.22 * alb).clamp(0.05, 1.0)

    ec, nc = side / 2.0, side / 2.0
    zref = float(z.mean())
    images = []
    rows = torch.arange(W, device=device, dtype=torch.float64)
    c0 = (W - 1) / 2.0
    chunk = max(1, (1 << 24) // W)   # rows per pass: ~16 M pixels
    for vd in spec.view_deg:
        out = torch.empty((W, W), dtype=torch.int32, device=device)
        tv = math.tan(math.radians(vd))
        cols = torch.arange(W, device=device, dtype=torch.float64)
        for r0 in range(0, W, chunk):
            r = rows[r0:r0 + chunk][:, None].expand(-1, W)
            c = cols[None, :].expand(r.shape[0], -1)
            Z = torch.full(r.shape, zref, dtype=torch.float64, device=device)
            for _ in range(10):
                dz = Z - zref
                E = ec + (c - c0) * spec.gsd * (spec.orbit_h - dz) / spec.orbit_h
                N = nc - spec.gsd * (r - c0 - spec.yaw * (E - ec) / spec.gsd) - dz * tv
                Z = _sample(z, origin, post, E.float(), N.float()).double()
            dz = Z - zref
            E = ec + (c - c0) * spec.gsd * (spec.orbit_h - dz) / spec.orbit_h
            N = nc - spec.gsd * (r - c0 - spec.yaw * (E - ec) / spec.gsd) - dz * tv
            a = _sample(alb, origin, apost, E.float(), N.float())
            sh = _sample(shade, origin, post, E.float(), N.float())
            dn = 1.0 + torch.round(2046.0 * (0.08 + 0.92 * a * (0.25 + 0.75 * sh)).clamp(0.0, 1.0))
            out[r0:r0 + chunk] = dn.int()
        images.append(out.clamp(1, 2047).to(torch.int16).cpu().numpy().view(np.uint16))

    terrain = z.cpu().numpy()
    scene = Scene(spec, images, [], terrain, origin, (0.0, 0.0, side, side))
    # RPC fit on a terrain-independent control cube
    g = np.linspace(-margin * 0.5, side + margin * 0.5, 21)
    hh = np.linspace(spec.base - 60.0, spec.base + spec.relief + 60.0, 7)
    EE, NN, HH = np.meshgrid(g, g, hh, indexing="ij")
    EE, NN, HH = EE.ravel(), NN.ravel(), HH.ravel()
    lon, lat = scene.to_wgs(EE, NN)
    rng = np.random.default_rng(spec.seed + 99)
    Ev = rng.uniform(0, side, 4000)
    Nv = rng.uniform(0, side, 4000)
    Hv = rng.uniform(hh[0], hh[-1], 4000)
    lonv, latv = scene.to_wgs(Ev, Nv)
    for vd in spec.view_deg:
        row, col = _camera(spec, vd, EE, NN, HH, zref, ec, nc)
        rpc = fit_rpc(lon, lat, HH, row, col)
        rv, cv = _camera(spec, vd, Ev, Nv, Hv, zref, ec, nc)
        lr, lc = eval_rpc(rpc, lonv, latv, Hv)
        scene.rpcs.append(rpc)
        scene.fit_residual_px.append(float(max(np.abs(lr - rv).max(), np.abs(lc - cv).max())))
    return scene


@dataclass
class GridSpec:
    cells: int                    # cells per side
    grid_res: float               # metres
    band: float = 32.0            # height band width (integer metres)
    inset: float = 40.0           # metres kept from the image edge
    x_offset: float = 500000.0    # GridPts easting of ground E = 0
    y_offset: float = 3000000.0   # GridPts northing of ground N = 0
    seed: int = 7


def make_grid(scene: Scene, gs: GridSpec):
    """Cell grid, lon/lat per cell, truth heights and a band search space.

    Returns dict with grid_pts (n,2), grid_wgs (n,2), truth (n,), boundary,
    ugrid (structured numpy array in UGRID layout), minmax.
    """
    n = gs.cells
    E = gs.inset + np.arange(n) * gs.grid_res
    N = gs.inset + np.arange(n) * gs.grid_res
    EE, NN = np.meshgrid(E, N)           # row index = northing (SetDEMGrid order)
    EE = EE.ravel()
    NN = NN.ravel()
    lon, lat = scene.to_wgs(EE, NN)
    truth = scene.height_at(EE, NN)
    rng = np.random.default_rng(gs.seed)
    u = rng.uniform(0.0, gs.band, EE.shape[0])
    lo = np.floor(truth - u).astype(np.int64)
    hi = lo + int(gs.band)
    from .records import UGRID_DTYPE
    ug = np.zeros(EE.shape[0], dtype=UGRID_DTYPE)
    ug["Height"] = truth.astype(np.float32)
    ug["minHeight"] = lo
    ug["maxHeight"] = hi
    ug["Matched_flag"] = 1
    boundary = np.array([gs.x_offset + gs.inset, gs.y_offset + gs.inset,
                         gs.x_offset + gs.inset + n * gs.grid_res, gs.y_offset + gs.inset + n * gs.grid_res])
    return {
        "grid_pts": np.stack([EE + gs.x_offset, NN + gs.y_offset], 1),
        "grid_wgs": np.stack([lon, lat], 1),
        "truth": truth,
        "boundary": boundary,
        "ugrid": ug,
        "minmax": np.array([float(lo.min() - 1), float(hi.max() + 1)]),
    }


def level_grid(scene: Scene, level: int, cells: int, nheights: int, seed: int = 7, resolution: float = 0.5):
    """Centred cell grid at one pyramid level with `nheights` hypotheses per
    cell: grid spacing max(2 m, GSD * 2^level), height step GetHeightStep,
    band = nheights * step (whole metres) around the true height."""
    from .records import height_step
    grid_res = max(2.0, resolution * (1 << level))
    hs = height_step(level, resolution)
    band = float(max(1, int(round(nheights * hs))))
    side = scene.spec.image_px * scene.spec.gsd
    inset = (side - cells * grid_res) / 2.0
    g = make_grid(scene, GridSpec(cells=cells, grid_res=grid_res, band=band, inset=inset, x_offset=0.0,
                                  y_offset=0.0, seed=seed + level))
    g["boundary"] = np.array([inset, inset, inset + cells * grid_res, inset + cells * grid_res])
    g["shape"] = (cells, cells)
    g["grid_res"] = grid_res
    return g


def ortho_wgs(scene: Scene, boundary, level: int, resolution: float = 0.5, x_offset: float = 0.0,
              y_offset: float = 0.0):
    """(lon, lat) of every pixel of the ortho grid SETSM's SetOrthoImageCoord
    walks at one level (setsm_code.cpp:6552-6553, 6593-6594): columns
    (Boundary[2]-Boundary[0])/im_res + 1, rows likewise, pixel (row, col) at
    (Boundary[0] + col*im_res, Boundary[1] + row*im_res), im_res =
    resolution * 2^level. Returns an array (rows, cols, 2)."""
    im_res = resolution * float(1 << level)
    sw = int((boundary[2] - boundary[0]) / im_res) + 1
    sh = int((boundary[3] - boundary[1]) / im_res) + 1
    X = boundary[0] + np.arange(sw) * im_res
    Y = boundary[1] + np.arange(sh) * im_res
    XX, YY = np.meshgrid(X, Y)
    lon, lat = scene.to_wgs(XX - x_offset, YY - y_offset)
    return np.stack([lon, lat], -1)


@dataclass
class TileDraw:
    """One tile of the mixed-size workload: cells per side (2 m cells),
    hypotheses per cell, terrain relief, view convergence and scene seed."""
    side: int
    nheights: int
    relief: float
    convergence: float
    seed: int


def draw_tiles(n, seed, sides=(512, 4096), heights=(32, 256), relief=(50.0, 1500.0), convergence=(15.0, 40.0)):
    """Seeded draws: side log-uniform in `sides`, heights uniform integers in
    `heights`, relief and convergence uniform, an independent scene seed each."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        side = int(round(math.exp(rng.uniform(math.log(sides[0]), math.log(sides[1])))))
        out.append(TileDraw(side, int(rng.integers(heights[0], heights[1] + 1)), float(rng.uniform(*relief)),
                            float(rng.uniform(*convergence)), int(rng.integers(1, 1 << 30))))
    return out


def tile_image_px(d: TileDraw, gsd=0.5, grid_res=2.0, step=0.25):
    """Square image side that holds the tile's cell grid plus the parallax of
    every height of its bands (relief + band, each view at convergence / 2)
    and the window margins."""
    band = d.nheights * step
    tan = math.tan(math.radians(d.convergence / 2.0))
    margin = int(math.ceil((d.relief + band + 40.0) * tan / gsd)) + 32
    return int(round(d.side * grid_res / gsd)) + 2 * margin


def render_tile(d: TileDraw, device="cuda", base=200.0, gsd=0.5):
    """Scene and level-0 cell grid (2 m cells, d.nheights heights at the
    0.25 m step around the true surface) of one drawn tile."""
    spec = SceneSpec(image_px=tile_image_px(d, gsd), gsd=gsd, relief=d.relief, base=base,
                     view_deg=(d.convergence / 2.0, -d.convergence / 2.0), seed=d.seed)
    sc = make_scene(spec, device=device)
    return sc, level_grid(sc, 0, d.side, d.nheights, seed=d.seed)
