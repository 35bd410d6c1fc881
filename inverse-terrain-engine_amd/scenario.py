"""Solver parameters, forcing, and the synthetic workloads the evaluator is measured on.

SolverParams mirrors crates/itr-hydro/src/solver.rs:26-49 with the defaults of
validation::params (crates/itr-hydro/src/validation.rs:23-41).  The workload builders
produce DEMs and candidate earthworks from seeded generators (SplitMix64), so every run
of a configuration sees the same inputs.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field

import numpy as np

from .terrain import (FineTerrain, GeoTransform, Primitive, PreparedTerrain, TerrainBase, TerrainEdit,
                      materialize)


def mm_h(pts):
    """Hyetograph::from_mm_h (crates/itr-core/src/scenario.rs:534-536): mm/h -> m/s."""
    return [(float(t), float(r) / 1000.0 / 3600.0) for t, r in pts]


@dataclass
class SolverParams:
    duration_s: float
    sync_interval_s: float
    cfl: float = 0.35
    dt_max_s: float = 5.0
    h_dry: float = 1e-6
    h_eps: float = 1e-5
    rain: list = field(default_factory=list)     # [(t s, rate m/s)]
    infil_capacity: float = 0.0                   # m/s
    initial_depth: float = 0.0
    initial_stage: float | None = None
    skip_dry: bool = True
    locked_dt: np.ndarray | None = None
    cfl_hard: float = 0.5
    inflow_scale: float = 1.0
    max_steps: int = 0                            # test control: stop after this many steps
    kernel: str = "fused"                         # KernelKind (solver.rs:19-24): "fused" (HLL) | "inertial"


def scaled(p: SolverParams, rain: float, inflow: float) -> SolverParams:
    """Forcing of an ensemble member (crates/itr-opt/src/problem.rs:56-63): every rain rate
    times ``rain`` (f64), the inflow scale times ``inflow``."""
    q = copy.copy(p)
    q.rain = [(t, r * rain) for t, r in p.rain]
    q.inflow_scale = p.inflow_scale * inflow
    return q


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): the seeded generator of the workloads."""
    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = seed & self.MASK

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0) -> float:
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


def validation_terrain(nx, ny, dx, zfun, walls=None, n=0.035, segments=(), level=1) -> TerrainBase:
    """validation::terrain (validation.rs:44-65): z(x, y) at cell centres, origin (0, ny*dx)."""
    geo = GeoTransform(0.0, ny * dx, dx, dx)
    dem = np.zeros((ny, nx))
    wm = np.zeros((ny, nx), bool)
    for r in range(ny):
        for c in range(nx):
            x, y = geo.cell_center(c, r)
            dem[r, c] = zfun(x, y)
            if walls is not None:
                wm[r, c] = walls(c, r)
    return TerrainBase.build(FineTerrain.new(dem, wm, n, geo), level, segments)


def bench_case_terrain(n: int) -> TerrainBase:
    """The reference's `itr bench` case (crates/itr-cli/src/main.rs:682-702): dx = 2 m,
    z = 10 + 0.002 (L - x) + 0.4 sin(x/23) cos(y/17), n = 0.035, closed walls."""
    dx = 2.0
    L = n * dx
    return validation_terrain(n, n, dx, lambda x, y: 10.0 + 0.002 * (L - x) + 0.4 * math.sin(x / 23.0) * math.cos(y / 17.0))


def bench_case_params() -> SolverParams:
    """600 s, sync 60 s, initial depth 0.05 m, constant 50 mm/h (main.rs:689-691)."""
    return SolverParams(600.0, 60.0, initial_depth=0.05, rain=mm_h([[0.0, 50.0]]))


# ------------------------------------------------------------------ configurations
def hills_dem(nx, ny, dx, rng: SplitMix64, n_hills, slope=0.01, amp=(0.5, 2.0), sigma=(15.0, 40.0),
              noise=0.02, n_gullies=0, gully_depth=1.0, base=100.0) -> np.ndarray:
    """Plane sloping `slope` down to the east + Gaussian hills + seeded Gaussian-section
    gullies along random segments + uniform noise."""
    L = nx * dx
    x = (np.arange(nx) + 0.5) * dx
    y = ((ny - np.arange(ny)) - 0.5) * dx
    px, py = np.meshgrid(x, y)
    z = base + slope * (L - px)
    for _ in range(n_hills):
        cx, cy = rng.uniform(0, nx * dx), rng.uniform(0, ny * dx)
        a, s = rng.uniform(*amp), rng.uniform(*sigma)
        z = z + a * np.exp(-((px - cx) ** 2 + (py - cy) ** 2) / (2 * s * s))
    for _ in range(n_gullies):
        x0, y0 = rng.uniform(0, nx * dx), rng.uniform(0, ny * dx)
        ang = rng.uniform(0, math.pi)
        ln = rng.uniform(0.3, 0.8) * max(nx, ny) * dx
        x1, y1 = x0 + ln * math.cos(ang), y0 + ln * math.sin(ang)
        vx, vy = x1 - x0, y1 - y0
        t = np.clip(((px - x0) * vx + (py - y0) * vy) / (vx * vx + vy * vy), 0.0, 1.0)
        d2 = (x0 + t * vx - px) ** 2 + (y0 + t * vy - py) ** 2
        w = rng.uniform(3.0, 6.0) * dx
        z = z - gully_depth * np.exp(-d2 / (w * w))
    if noise > 0:
        u = np.array([[rng.uniform(-noise, noise) for _ in range(nx)] for _ in range(ny)])
        z = z + u
    return z


def random_primitives(rng: SplitMix64, k, geo: GeoTransform, nx, ny, kinds=("berm", "swale", "mound"), q=None):
    """k primitives: centre uniform on the grid, length 10-100 m, width 2-6 m, angle
    uniform in [0, pi), height uniform in [-0.5, 0.5] m; positions, lengths and widths
    quantized to cells and angles to 1 degree (design.rs:259-296)."""
    out = []
    dx = geo.dx
    q = dx if q is None else q
    for i in range(k):
        kind = kinds[i % len(kinds)]
        cx = round(rng.uniform(0, nx * dx) / q) * q + geo.origin_x
        cy = geo.origin_y - round(rng.uniform(0, ny * dx) / q) * q
        length = round(rng.uniform(10.0, 100.0) / q) * q
        width = max(round(rng.uniform(2.0, 6.0) / q) * q, 1e-6)
        ang = float(round(math.degrees(rng.uniform(0.0, math.pi))))
        h = rng.uniform(-0.5, 0.5)
        h = round(h / 0.01) * 0.01
        if kind == "mound":
            length, ang = 0.0, 0.0
        out.append(Primitive(kind, cx, cy, length, width, ang, h))
    return out


@dataclass
class Workload:
    name: str
    base: TerrainBase
    params: SolverParams
    candidates: list           # PreparedTerrain
    monitors: np.ndarray       # sorted unique cells
    guard_pos: np.ndarray      # positions into monitors
    asset_pos: list
    edits: list = field(default_factory=list)
    designs: list = field(default_factory=list)   # each candidate's primitives
    asset_boxes: list = field(default_factory=list)   # (c0, r0, w, h) cell boxes of the assets


def _ring_contains(ring, x, y):
    """Even-odd crossing test of one ring (crates/itr-core/src/raster.rs:122-138)."""
    inside = False
    n = len(ring)
    if n < 3:
        return False
    j = n - 1
    for i in range(n):
        xi, yi = ring[i]
        xj, yj = ring[j]
        if (yi > y) != (yj > y) and x < (xj - xi) * (y - yi) / (yj - yi) + xi:
            inside = not inside
        j = i
    return inside


def rasterize(polys, nx, ny, geo) -> np.ndarray:
    """Polygons (dicts with "exterior" and "holes" rings of (x, y)) to a [ny][nx] mask by
    the cell-centre rule (raster.rs:151-177): a cell belongs to a polygon iff its centre
    lies inside the exterior and outside every hole."""
    m = np.zeros((ny, nx), bool)
    for p in polys:
        ext = [tuple(v) for v in p["exterior"]]
        holes = [[tuple(v) for v in h] for h in p.get("holes", [])]
        xs, ys = [v[0] for v in ext], [v[1] for v in ext]
        c0f, r1f = geo.to_grid(min(xs), min(ys))
        c1f, r0f = geo.to_grid(max(xs), max(ys))
        c0, r0 = max(int(math.floor(c0f)), 0), max(int(math.floor(r0f)), 0)
        c1, r1 = min(max(int(math.ceil(c1f)), 0), nx), min(max(int(math.ceil(r1f)), 0), ny)
        for r in range(r0, r1):
            for c in range(c0, c1):
                x, y = geo.cell_center(c, r)
                if _ring_contains(ext, x, y) and not any(_ring_contains(h, x, y) for h in holes):
                    m[r, c] = True
    return m


def asset_and_guard_cells(nx, ny, boxes, guard_cols):
    cells = set()
    assets = []
    for (c0, r0, w, h) in boxes:
        a = [r * nx + c for r in range(r0, r0 + h) for c in range(c0, c0 + w)]
        assets.append(a)
        cells.update(a)
    guard = [r * nx + c for r in range(ny) for c in guard_cols]
    cells.update(guard)
    mon = np.array(sorted(cells), np.uint32)
    pos = {int(c): i for i, c in enumerate(mon)}
    return mon, np.array([pos[c] for c in guard], np.int32), [[pos[c] for c in a] for a in assets]


def example_workload(spec: dict, dem: np.ndarray):
    """A scenario given as data (the reference's example directory decoded: DEM, its
    georeferencing, hydrology, boundary segments, protected and guard polygons), assembled
    as Problem::build does for level 1 (crates/itr-opt/src/problem.rs:103-160): walls where
    the DEM is not finite, uniform Manning n, assets and guard rasterized by the cell-centre
    rule, monitors = asset cells and guard cells, the SolverParams defaults
    (crates/itr-core/src/scenario.rs:334-345, solver.rs:52-75).  Returns (Workload,
    assets [(positions, threshold_m, weight)], tau_m)."""
    dem = np.asarray(dem, np.float64)
    ny, nx = dem.shape
    g = spec["geo"]
    geo = GeoTransform(g["origin_x"], g["origin_y"], g["dx"], g["dy"])
    walls = ~np.isfinite(dem)
    fine = FineTerrain.new(np.where(walls, 0.0, dem), walls, spec["manning_n"], geo)
    base = TerrainBase.build(fine, 1, [dict(sg) for sg in spec["boundaries"]])
    p = SolverParams(float(spec["duration_s"]), float(spec["sync_interval_s"]), rain=mm_h(spec["rain_mm_h"]),
                     infil_capacity=spec["infil_capacity_mm_h"] / 1000.0 / 3600.0)
    amasks = [rasterize([a], nx, ny, geo) for a in spec["protected"]]
    gmask = rasterize(spec["guard"], nx, ny, geo)
    cells = set(int(k) for k in np.flatnonzero(gmask))
    for m in amasks:
        cells.update(int(k) for k in np.flatnonzero(m))
    mon = np.array(sorted(cells), np.uint32)
    pos = {int(c): i for i, c in enumerate(mon)}
    thr = spec["depth_threshold_m"]
    assets = [(np.array([pos[int(k)] for k in np.flatnonzero(m)], np.int32),
               a["properties"].get("threshold_m", thr), a["properties"].get("weight", 1.0))
              for a, m in zip(spec["protected"], amasks)]
    gpos = np.array([pos[int(k)] for k in np.flatnonzero(gmask)], np.int32)
    w = Workload("example", base, p, [PreparedTerrain(base)], mon, gpos, [a[0] for a in assets])
    return w, assets, spec["tau_m"]


def config1(cfl=0.9) -> Workload:
    """90x70 at 5 m; plane 1% + 6 hills + noise; 40 min triangular storm peaking at
    60 mm/h at 15 min; n 0.035; infiltration 5 mm/h; walls N/S/W, free outflow east;
    two 4x4 assets; 10x70 guard strip on the east side."""
    nx, ny, dx = 90, 70, 5.0
    rng = SplitMix64(1)
    dem = hills_dem(nx, ny, dx, rng, 6)
    geo = GeoTransform(0.0, ny * dx, dx, dx)
    base = TerrainBase.build(FineTerrain.new(dem, None, 0.035, geo), 1,
                             [{"kind": "transmissive", "side": "east"}])
    p = SolverParams(2400.0, 60.0, cfl=cfl, rain=mm_h([[0, 0], [900, 60], [2400, 0]]),
                     infil_capacity=5.0 / 1000.0 / 3600.0)
    mon, gpos, apos = asset_and_guard_cells(nx, ny, [(30, 20, 4, 4), (55, 45, 4, 4)], range(nx - 10, nx))
    return Workload("config1", base, p, [PreparedTerrain(base)], mon, gpos, apos)


def config2(n_cand=64, seed=2) -> Workload:
    """64 CMA-ES-like candidates on a 120x96 grid at 5 m (10 hills + 3 gullies), 3
    primitives each; 60 min storm peaking at 80 mm/h; guard abort at 0.01 m."""
    nx, ny, dx = 120, 96, 5.0
    rng = SplitMix64(seed)
    dem = hills_dem(nx, ny, dx, rng, 10, n_gullies=3)
    geo = GeoTransform(0.0, ny * dx, dx, dx)
    base = TerrainBase.build(FineTerrain.new(dem, None, 0.035, geo), 1,
                             [{"kind": "transmissive", "side": "east"}])
    p = SolverParams(3600.0, 60.0, rain=mm_h([[0, 0], [900, 80], [3600, 0]]),
                     infil_capacity=5.0 / 1000.0 / 3600.0)
    mon, gpos, apos = asset_and_guard_cells(nx, ny, [(40, 30, 4, 4), (70, 60, 4, 4)], range(nx - 10, nx))
    cands, edits = [], []
    for _ in range(n_cand):
        prims = random_primitives(rng, 3, geo, nx, ny)
        e = materialize(prims, geo, nx, ny, None, 1.0)
        edits.append(e)
        cands.append(PreparedTerrain(base).apply_edit(e))
    return Workload("config2", base, p, cands, mon, gpos, apos, edits)


@dataclass
class AdjointCase:
    """One design on one grid for the discrete adjoint."""
    base: TerrainBase
    params: SolverParams
    prims: list
    edit: TerrainEdit
    candidate: PreparedTerrain
    monitors: np.ndarray
    assets: list                # (positions, threshold_m, weight)
    tau_m: float
    earth_weight: float


def config4(seed=4, duration=2400.0, level=1) -> AdjointCase:
    """120x96 at 5 m (10 hills + 3 gullies), 40 min storm peaking at 80 mm/h at 15 min,
    infiltration 5 mm/h, CFL 0.9 (about 2000 steps), one design of 3 primitives; two 4x4
    assets (thresholds 0.02 and 0.04 m); tau 0.01 m, lambda 1e-3 per m^3.  The dt schedule is frozen before differentiating (the caller
    sets params.locked_dt from a forward run).  level=2: the same terrain and design on
    the 60x48 face-crest level at 10 m (assets 2x2 level cells over the same ground)."""
    nx, ny, dx = 120, 96, 5.0
    rng = SplitMix64(seed)
    dem = hills_dem(nx, ny, dx, rng, 10, n_gullies=3)
    geo = GeoTransform(0.0, ny * dx, dx, dx)
    base = TerrainBase.build(FineTerrain.new(dem, None, 0.035, geo), level, [{"kind": "transmissive", "side": "east"}])
    p = SolverParams(duration, 60.0, cfl=0.9, rain=mm_h([[0, 0], [900, 80], [2400, 0]]),
                     infil_capacity=5.0 / 1000.0 / 3600.0)
    p.cfl_hard = 1.0   # the locked schedule runs at CFL 0.9
    boxes = [(c // level, r // level, w // level, h // level) for c, r, w, h in [(40, 30, 4, 4), (70, 60, 4, 4)]]
    mon, _, apos = asset_and_guard_cells(base.nx, base.ny, boxes, range(base.nx - 10 // level, base.nx))
    prims = random_primitives(rng, 3, geo, nx, ny)
    edit = materialize(prims, geo, nx, ny, None, 1.0)
    assets = [(np.array(a, np.int32), th, 1.0) for a, th in zip(apos, (0.02, 0.04))]
    return AdjointCase(base, p, prims, edit, PreparedTerrain(base).apply_edit(edit), mon, assets, 0.01, 1e-3)


def adjoint_boundary_case(kind: str, seed=6, duration=1200.0) -> AdjointCase:
    """The adjoint on a grid with a special boundary: 60x48 at 5 m (6 hills, 2 gullies,
    sloping down to the east), a Stage (kind='stage': level 0.3 m above the lowest bed of
    the segment's boundary cells) or an Inflow (kind='inflow': 0 -> 1.5 -> 0.5 m^3/s)
    segment over rows 10-29 of the west side, the east side transmissive; 20 min storm
    peaking at 40 mm/h at 5 min, infiltration 5 mm/h, CFL 0.9.  One design: a mound on
    the segment's boundary cells, a berm across the inflow path and a swale; two 3x3
    assets downstream (thresholds 0.02 and 0.04 m); tau 0.01 m, lambda 1e-3 per m^3.
    kind='initial': no special segment; an initial stage 0.05 m above the bed's 30th
    percentile ponds the low ground, the two assets sit on its shore (initial depths
    nearest 0.03 and 0.12 m) and the mound on the first of them; the first 10 min of the
    storm (over all of it the second asset's pond fills past a sill, where its f32 and f64
    peaks part by 1.5 cm)."""
    nx, ny, dx = 60, 48, 5.0
    rng = SplitMix64(seed)
    dem = hills_dem(nx, ny, dx, rng, 6, n_gullies=2)
    geo = GeoTransform(0.0, ny * dx, dx, dx)
    rows = (10, 30)
    if kind == "initial":
        stage = float(np.quantile(dem, 0.3)) + 0.05
        d0 = stage - dem
        boxes = []
        for want in (0.03, 0.12):
            cost = np.abs(d0 - want)
            cost[:3, :] = cost[-3:, :] = np.inf
            cost[:, :3] = cost[:, -5:] = np.inf
            for (c0, r0, w, h) in boxes:   # keep the boxes apart
                cost[max(r0 - 4, 0):r0 + h + 4, max(c0 - 4, 0):c0 + w + 4] = np.inf
            r, c = np.unravel_index(int(np.argmin(cost)), cost.shape)
            boxes.append((int(c) - 1, int(r) - 1, 3, 3))
        base = TerrainBase.build(FineTerrain.new(dem, None, 0.035, geo), 1, [{"kind": "transmissive", "side": "east"}])
        p = SolverParams(min(duration, 600.0), 60.0, cfl=0.9, rain=mm_h([[0, 0], [300, 40], [1200, 0]]),
                         infil_capacity=5.0 / 1000.0 / 3600.0, initial_stage=stage)
        p.cfl_hard = 1.0
        y = lambda r: geo.origin_y - (r + 0.5) * dx
        (ca, ra, _, _), (cb, rb, _, _) = boxes
        prims = [Primitive("mound", (ca + 1.5) * dx, y(ra + 1), 0.0, 3.0 * dx, 0.0, 0.2),
                 Primitive("berm", (cb + 1.5) * dx, y(rb - 2), 8.0 * dx, 2.0 * dx, 10.0, 0.15),
                 Primitive("swale", (ca + 6.5) * dx, y(ra + 4), 8.0 * dx, 2.0 * dx, 60.0, -0.1)]
        edit = materialize(prims, geo, nx, ny, None, 1.0)
        mon, _, apos = asset_and_guard_cells(nx, ny, boxes, range(nx - 4, nx))
        assets = [(np.array(a, np.int32), th, 1.0) for a, th in zip(apos, (0.05, 0.15))]
        return AdjointCase(base, p, prims, edit, PreparedTerrain(base).apply_edit(edit), mon, assets, 0.01, 1e-3)
    if kind == "stage":
        seg = {"kind": "stage", "side": "west", "range_m": (rows[0] * dx, rows[1] * dx),
               "stage_m": float(dem[rows[0]:rows[1], 0].min()) + 0.3}
    else:
        seg = {"kind": "inflow", "side": "west", "range_m": (rows[0] * dx, rows[1] * dx),
               "hydrograph": [[0.0, 0.0], [300.0, 1.5], [1200.0, 0.5]]}
    base = TerrainBase.build(FineTerrain.new(dem, None, 0.035, geo), 1, [seg, {"kind": "transmissive", "side": "east"}])
    p = SolverParams(duration, 60.0, cfl=0.9, rain=mm_h([[0, 0], [300, 40], [1200, 0]]),
                     infil_capacity=5.0 / 1000.0 / 3600.0)
    p.cfl_hard = 1.0
    y = lambda r: geo.origin_y - (r + 0.5) * dx
    prims = [Primitive("mound", 1.5 * dx, y(20), 0.0, 4.0 * dx, 0.0, 0.25),
             Primitive("berm", 12.5 * dx, y(22), 12.0 * dx, 2.0 * dx, 80.0, 0.2),
             Primitive("swale", 20.5 * dx, y(18), 10.0 * dx, 2.0 * dx, 20.0, -0.15)]
    edit = materialize(prims, geo, nx, ny, None, 1.0)
    mon, _, apos = asset_and_guard_cells(nx, ny, [(8, 15, 3, 3), (16, 25, 3, 3)], range(nx - 4, nx))
    assets = [(np.array(a, np.int32), th, 1.0) for a, th in zip(apos, (0.02, 0.04))]
    return AdjointCase(base, p, prims, edit, PreparedTerrain(base).apply_edit(edit), mon, assets, 0.01, 1e-3)


def refine_dem(dem: np.ndarray, f: int) -> np.ndarray:
    """Upsample a DEM by an integer factor, bilinear between cell centres (the refinement
    the reference's revalidation uses, crates/itr-cli/src/revalidate.rs:19-38; every
    value here is finite, so its nearest-parent fallback never applies)."""
    ny, nx = dem.shape
    y = np.clip((np.arange(ny * f) + 0.5) / f - 0.5, 0.0, ny - 1.0)
    x = np.clip((np.arange(nx * f) + 0.5) / f - 0.5, 0.0, nx - 1.0)
    r0 = np.floor(y).astype(np.int64)
    c0 = np.floor(x).astype(np.int64)
    wy = (y - r0)[:, None]
    wx = (x - c0)[None, :]
    r1 = np.minimum(r0 + 1, ny - 1)
    c1 = np.minimum(c0 + 1, nx - 1)
    z00, z01 = dem[r0][:, c0], dem[r0][:, c1]
    z10, z11 = dem[r1][:, c0], dem[r1][:, c1]
    return (1.0 - wy) * ((1.0 - wx) * z00 + wx * z01) + wy * ((1.0 - wx) * z10 + wx * z11)


@dataclass
class Ladder:
    """A mixed fidelity-ladder batch: one group per grid (TerrainBase, candidates,
    monitors, guard-or-None), in the order the groups are uploaded."""
    name: str
    params: SolverParams
    groups: list
    edits: list


def config3(n_inst=256, seed=3, duration=2400.0) -> Ladder:
    """Uneven fidelity ladder: 45x35 at 10 m (face-crest coarse level 2), 90x70 at 5 m
    and 180x140 at 2.5 m (bilinear refinement of the 5 m DEM) in ratio 4:2:1 over
    n_inst instances (146/73/37 for 256); 2-4 primitives each, materialized on each
    grid's own fine raster; config-1 storm and infiltration."""
    nx, ny, dx = 90, 70, 5.0
    rng = SplitMix64(seed)
    dem = hills_dem(nx, ny, dx, rng, 6)
    geo = GeoTransform(0.0, ny * dx, dx, dx)
    segs = [{"kind": "transmissive", "side": "east"}]
    fine5 = FineTerrain.new(dem, None, 0.035, geo)
    geo2 = GeoTransform(0.0, ny * dx, dx / 2, dx / 2)
    fine25 = FineTerrain.new(refine_dem(dem, 2), None, 0.035, geo2)
    levels = [(fine5, 2), (fine5, 1), (fine25, 1)]
    n2 = round(n_inst * 2 / 7)
    n1 = round(n_inst * 1 / 7)
    counts = [n_inst - n2 - n1, n2, n1]
    p = SolverParams(duration, 60.0, rain=mm_h([[0, 0], [900, 60], [2400, 0]]),
                     infil_capacity=5.0 / 1000.0 / 3600.0)
    groups, edits = [], []
    for (fine, lv), n in zip(levels, counts):
        base = TerrainBase.build(fine, lv, segs)
        mon, _, _ = asset_and_guard_cells(base.nx, base.ny, [(base.nx // 3, base.ny // 3, 2, 2)],
                                          range(base.nx - 2, base.nx))
        cands = []
        for _ in range(n):
            k = 2 + int(rng.uniform() * 3) % 3
            prims = random_primitives(rng, k, fine.geo, fine.nx, fine.ny)
            e = materialize(prims, fine.geo, fine.nx, fine.ny, None, 1.0)
            edits.append(e)
            cands.append(PreparedTerrain(base).apply_edit(e))
        groups.append((base, cands, mon, None))
    return Ladder("config3", p, groups, edits)


def config5(n_cand=128, seed=5, n=512, duration=7200.0) -> Workload:
    """128 candidates of 4 primitives on a 512x512 grid at 2 m (40 hills, 8 gullies),
    2 h bimodal storm (peaks 50 and 90 mm/h)."""
    nx = ny = n
    dx = 2.0
    rng = SplitMix64(seed)
    dem = hills_dem(nx, ny, dx, rng, 40, n_gullies=8, sigma=(15.0, 60.0), noise=0.0)
    geo = GeoTransform(0.0, ny * dx, dx, dx)
    base = TerrainBase.build(FineTerrain.new(dem, None, 0.035, geo), 1,
                             [{"kind": "transmissive", "side": "east"}])
    p = SolverParams(duration, 60.0, rain=mm_h([[0, 0], [1800, 50], [3600, 10], [5400, 90], [7200, 0]]),
                     infil_capacity=5.0 / 1000.0 / 3600.0)
    boxes = [(200, 200, 8, 8), (300, 350, 8, 8)]
    mon, gpos, apos = asset_and_guard_cells(nx, ny, boxes, range(nx - 16, nx))
    cands, edits, designs = [], [], []
    for _ in range(n_cand):
        prims = random_primitives(rng, 4, geo, nx, ny)
        e = materialize(prims, geo, nx, ny, None, 1.0)
        designs.append(prims)
        edits.append(e)
        cands.append(PreparedTerrain(base).apply_edit(e))
    return Workload("config5", base, p, cands, mon, gpos, apos, edits, designs, boxes)
