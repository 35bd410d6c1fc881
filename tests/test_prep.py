"""Host-side preparation (the inputs the HIP path consumes) against the oracle's own
restatement of the reference's terrain code: bitwise.  CPU only."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import oracle_terrain, po
from inverse_terrain_engine_amd import scenario as sc
from inverse_terrain_engine_amd import terrain as tr


def faces_equal(base_or_prep, ot, edit=None):
    f = ot.faces(None if edit is None else edit.as_dict(), 32)
    nx, ny = ot.nx, ot.ny
    p = base_or_prep
    assert np.array_equal(f["fx0"][1:-1, 0:nx + 1].astype(np.float32).view(np.uint32), p.fx0.view(np.uint32))
    assert np.array_equal(f["fy0"][0:ny + 1, 1:-1].astype(np.float32).view(np.uint32), p.fy0.view(np.uint32))
    if ot.crest:
        assert np.array_equal(f["fx1"][1:-1, 0:nx + 1].astype(np.float32).view(np.uint32), p.fx1.view(np.uint32))
        assert np.array_equal(f["fy1"][0:ny + 1, 1:-1].astype(np.float32).view(np.uint32), p.fy1.view(np.uint32))
    z = p.z_level if hasattr(p, "z_level") else p.z
    assert np.array_equal(f["z_level"].view(np.uint64), np.asarray(z, np.float64).view(np.uint64))
    return f


@pytest.mark.parametrize("level", [1, 2, 3])
def test_terrain_build_matches_oracle(level):
    nx, ny, dx = 47, 33, 2.0
    rng = sc.SplitMix64(9)
    dem = sc.hills_dem(nx, ny, dx, rng, 5, n_gullies=1)
    walls = np.zeros((ny, nx), bool)
    walls[5:9, 10:13] = True
    dem[20, 20] = np.nan   # nodata becomes a wall
    geo = tr.GeoTransform(0.0, ny * dx, dx, dx)
    man = np.full((ny, nx), 0.035)
    man[:, :10] = 0.05
    base = tr.TerrainBase.build(tr.FineTerrain.new(dem, walls, man, geo), level,
                                [{"kind": "transmissive", "side": "east"}])
    ot = oracle_terrain(base)
    assert (ot.nx, ot.ny, ot.crest, ot.wet_count) == (base.nx, base.ny, base.crest, base.wet_count)
    f = faces_equal(base, ot)
    assert np.array_equal(f["gn2"][1:-1, 1:-1].astype(np.float32).view(np.uint32), base.gn2.view(np.uint32))
    assert np.array_equal(f["wet"][1:-1, 1:-1].astype(np.float32), base.wet)
    g = ot.ghosts()
    for side in ("north", "south", "west", "east"):
        assert np.array_equal(g[side], base.ghosts[side]["code"])


@pytest.mark.parametrize("level,nx,ny", [(1, 60, 48), (2, 60, 48), (2, 61, 47), (3, 61, 47)])
def test_edits_match_oracle(level, nx, ny):
    """Edits on signed and face-crest levels, fine grids divisible by the level or not, fine
    walls (a block that makes whole coarse cells walls: crest faces from z_wall)."""
    dx = 2.0
    rng = sc.SplitMix64(17)
    dem = sc.hills_dem(nx, ny, dx, rng, 6)
    geo = tr.GeoTransform(0.0, ny * dx, dx, dx)
    walls = np.zeros((ny, nx), bool)
    walls[12:20, 24:30] = True
    base = tr.TerrainBase.build(tr.FineTerrain.new(dem, walls, 0.035, geo), level, [])
    ot = oracle_terrain(base)
    editable = np.ones((ny, nx), bool)
    editable[:, :8] = False
    for _ in range(6):
        prims = sc.random_primitives(rng, 3, geo, nx, ny, kinds=("berm", "swale", "cut", "mound"))
        e = tr.materialize(prims, geo, nx, ny, editable, 0.75)
        oe = po.materialize([dict(kind=p.kind, cx=p.cx, cy=p.cy, length=p.length, width=p.width,
                                  angle_deg=p.angle_deg, height=p.height) for p in prims],
                            (geo.origin_x, geo.origin_y), geo.dx, geo.dy, nx, ny, editable.astype(np.uint8), 0.75)
        assert (e.c0, e.r0) == (oe["c0"], oe["r0"])
        assert np.array_equal(e.dz.view(np.uint64), oe["dz"].view(np.uint64))
        assert e.cut_m3 == oe["cut_m3"] and e.fill_m3 == oe["fill_m3"]
        pt = tr.PreparedTerrain(base).apply_edit(e)
        faces_equal(pt, ot, e)


def test_materialize_closed_forms():
    """design.rs:453-504: compact support, volume = cell sum, analytic ridge volume;
    masked cells untouched and the clamp holds."""
    geo = tr.GeoTransform(0.0, 100.0, 1.0, 1.0)
    p = tr.Primitive("berm", 50.0, 50.0, 20.0, 4.0, 0.0, 0.5)
    e = tr.materialize([p], geo, 100, 100)
    assert abs(e.dz.sum() - e.fill_m3) < 1e-9
    assert e.w <= 26 and e.h <= 6
    section = 0.5 * 4.0 * 2.0 * 8.0 / 15.0 * 0.5
    assert abs(e.fill_m3 / (section * 20.0) - 1.0) < 0.2
    geo = tr.GeoTransform(0.0, 50.0, 1.0, 1.0)
    m = np.ones((50, 50), bool)
    m[:, :25] = False
    e = tr.materialize([tr.Primitive("mound", 25.0, 25.0, 0.0, 10.0, 0.0, 1.0)], geo, 50, 50, m, 0.75)
    assert np.all(e.dense(50, 50)[:, :25] == 0.0)
    assert e.dz.max() <= 0.75


def test_splitmix64_reference_outputs():
    """SplitMix64 (Steele, Lea and Flood 2014) published outputs for seed 1234567."""
    r = sc.SplitMix64(1234567)
    assert [r.next_u64() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                               4593380528125082431, 16408922859458223821]


def test_bench_case_terrain_is_reference_terrain():
    """The product's bench terrain equals the oracle terrain of `itr bench` bitwise."""
    base = sc.bench_case_terrain(64)
    L = 64 * 2.0
    oy = 64 * 2.0
    z = np.array([[10.0 + 0.002 * (L - (c + 0.5) * 2.0) + 0.4 * math.sin((c + 0.5) * 2.0 / 23.0) *
                   math.cos((oy - (r + 0.5) * 2.0) / 17.0) for c in range(64)] for r in range(64)])
    ot = po.Terrain(z, manning=0.035, dx=2.0)
    faces_equal(base, ot)


def test_boundary_segments_match_oracle():
    """Stage / Inflow / transmissive segments (Boundaries::from_segments, terrain.rs:151-187):
    codes equal to the oracle's; the per-cell parameter is stage - z_ref (Stage) or
    1 / segment length (Inflow, cells snapped outward to whole cells), with the inflow's
    hydrograph index."""
    segs = [{"kind": "inflow", "side": "west", "hydrograph": [[0.0, 0.0], [200.0, 1.5]], "range_m": (21.0, 39.0)},
            {"kind": "inflow", "side": "north", "hydrograph": [[0.0, 0.3]], "range_m": (4.0, 10.0)},
            {"kind": "transmissive", "side": "east"},
            {"kind": "stage", "side": "south", "stage_m": 100.3, "range_m": (5.0, 25.0)}]
    base = sc.validation_terrain(30, 25, 2.0, lambda x, y: 100.0 + 0.01 * x, n=0.035, segments=segs)
    ot = oracle_terrain(base)
    g = ot.ghosts()
    for side in ("north", "south", "west", "east"):
        assert np.array_equal(g[side], base.ghosts[side]["code"]), side
    codes, par, hyd = base.ghost_codes(), base.ghost_params(), base.ghost_hydro()
    west = slice(2 * 30, 2 * 30 + 25)
    assert list(np.nonzero(codes[west] == tr.GHOST_INFLOW)[0]) == list(range(10, 20))   # floor(21/2)..ceil(39/2)
    assert np.all(par[west][10:20] == 1.0 / 20.0) and np.all(hyd[west][10:20] == 0)
    assert np.all(par[0:30][2:5] == 1.0 / 6.0) and np.all(hyd[0:30][2:5] == 1)
    south = slice(30, 60)
    assert np.all(par[south][2:13] == 100.3 - base.z_ref)
    assert base.has_stage


def test_ensemble_scaling():
    """problem.rs:56-63: member forcing = every rain rate times the rain factor, the inflow
    scale times the inflow factor; the level's parameters are untouched."""
    p = sc.SolverParams(600.0, 60.0, rain=sc.mm_h([[0.0, 20.0], [200.0, 70.0]]))
    p.inflow_scale = 1.5
    q = sc.scaled(p, 1.4, 0.5)
    assert [r for _, r in q.rain] == [r * 1.4 for _, r in p.rain]
    assert [t for t, _ in q.rain] == [t for t, _ in p.rain]
    assert q.inflow_scale == 0.75 and p.inflow_scale == 1.5
