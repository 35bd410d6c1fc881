"""Shorter bands at a grid's end (the tail_rows / tail_band_rows options) and launch
lists split by candidates (list_split), both by default in launches of several rounds of
workgroups: the layout must not change a bit of the run.  Each case sets a layout whose short bands start on a multiple of their height, with
a short last band, and compares against the CPU oracle (the reference's f32 fused kernel
restated) or against the same batch run without short bands."""
from __future__ import annotations

import copy
import math

import numpy as np
import pytest

from conftest import assert_run_equal, oracle_params, oracle_terrain, po
from inverse_terrain_engine_amd import scenario as sc
from inverse_terrain_engine_amd import terrain as tr

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def oracle_run(base, p, edit=None, monitors=()):
    return po.run(oracle_terrain(base), oracle_params(p), precision=32, edit=edit, monitors=monitors)


def _solver(band_rows, tail_rows, tail_band_rows, chunk=6):
    from inverse_terrain_engine_amd.gpu import GpuSolver
    s = GpuSolver(0, band_rows=band_rows, steps_per_chunk=chunk)
    s.set_option("tail_rows", tail_rows)
    s.set_option("tail_band_rows", tail_band_rows)
    return s


def _equivalence_base():
    """West inflow, east transmissive, south stage segment, interior walls (the boundary
    fixture of crates/itr-hydro/tests/equivalence.rs:38-53)."""
    segs = [{"kind": "inflow", "side": "west", "hydrograph": [[0.0, 0.0], [200.0, 1.5], [500.0, 0.2]],
             "range_m": (20.0, 40.0)},
            {"kind": "transmissive", "side": "east"},
            {"kind": "stage", "side": "south", "stage_m": 100.3, "range_m": (5.0, 25.0)}]
    zf = lambda x, y: (100.0 + 0.5 * math.sin(x / 9.0) * math.cos(y / 7.0) + 0.004 * (120.0 - x)
                       + (0.6 if 60.0 <= x < 64.0 else 0.0))
    return sc.validation_terrain(57, 43, 2.0, zf, walls=lambda c, r: 30 <= c < 34 and 10 <= r < 18, n=0.035,
                                 segments=segs)


# (name, nx, ny, level, band_rows, tail_rows, tail_band_rows): 45 rows at 8 -> bands of 8
# to row 24, then 3-row bands (the last one 3 rows); the tail-wave grid (80 columns: the
# last wave marches row groups) and a face-crest level with the short bands' group split.
HLL_CASES = [("signed 70x45", 70, 45, 1, 8, 10, 3),
             ("tail-wave 80x37", 80, 37, 1, 5, 7, 2),
             ("crest 70x41 level 2", 70, 41, 2, 7, 9, 3)]


@pytest.mark.parametrize("case", HLL_CASES, ids=[c[0] for c in HLL_CASES])
def test_tail_bands_hll(case):
    name, nx, ny, level, R, T, R2 = case
    _say(f"case: short end bands, {name}, band_rows {R}, tail_rows {T}, tail_band_rows {R2}")
    s = _solver(R, T, R2)
    fx, fy = nx * level, ny * level
    base = sc.validation_terrain(fx, fy, 2.0, lambda x, y: 4.0 + 0.02 * y + 0.3 * math.sin(x / 11.0) * math.cos(y / 5.0),
                                 walls=lambda c, r: fx - 12 <= c < fx - 9 and fy // 3 <= r < fy // 2, n=0.03,
                                 segments=[{"kind": "transmissive", "side": "south"}], level=level)
    assert base.nx == nx and base.ny == ny
    p = sc.SolverParams(150.0, 30.0, rain=sc.mm_h([[0, 100], [90, 100], [120, 0]]), initial_stage=4.3,
                        infil_capacity=2e-6)
    mon = np.arange(1, nx * ny, 7, dtype=np.uint32)
    g = s.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    s.close()
    assert_run_equal(g, oracle_run(base, p, monitors=mon), name)


@pytest.mark.parametrize("split", [1, 3], ids=["one-list", "three-lists"])
@pytest.mark.parametrize("li", [False, True], ids=["hll", "local-inertial"])
def test_tail_bands_stage_inflow(li, split):
    """Stage and Inflow ghosts (the special kernel, whose ghosts are staged per band) and
    edited candidates, with 6-row bands to row 24 and 4-row bands after; the three
    candidates in one launch list or in three (one per stream)."""
    _say(f"case: short end bands, stage + inflow boundaries, {'local inertial' if li else 'HLL'}, list_split {split}")
    s = _solver(6, 8, 4, chunk=5)
    s.set_option("list_split", split)
    base = _equivalence_base()
    p = sc.SolverParams(900.0, 60.0, infil_capacity=10.0 / 1000.0 / 3600.0,
                        rain=sc.mm_h([[0.0, 0.0], [300.0, 0.0], [400.0, 80.0], [700.0, 0.0]]))
    if li:
        p = copy.copy(p)
        p.kernel = "inertial"
    mon = np.array([c for c in range(57 * 43) if c % 7 == 3], np.uint32)
    dz = np.array([0.05 * (1 + k % 5) for k in range(18)]).reshape(3, 6)
    edits = [None, tr.TerrainEdit(c0=40, r0=5, dz=dz, fill_m3=float(dz.sum()) * 4.0),
             tr.TerrainEdit(c0=2, r0=40, dz=-0.4 * np.ones((3, 10)), cut_m3=12 * 4.0)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    runs = s.run_batch(p, terrs, mon)
    s.close()
    for i, (r, e) in enumerate(zip(runs, edits)):
        o = oracle_run(base, p, edit=None if e is None else e.as_dict(), monitors=mon)
        assert_run_equal(r, o, f"stage+inflow candidate {i}", li=li)


def test_tail_bands_locked_restart():
    """A locked-dt schedule with a failure forced mid-interval: the adaptive restart reads
    the sync snapshot that every band, short ones included, wrote (solver.rs:456-480)."""
    _say("case: short end bands, locked dt with a restart, 4 candidates on 60x48")
    from inverse_terrain_engine_amd.gpu import BatchSession
    nx, ny, dx = 60, 48, 5.0
    rng = sc.SplitMix64(31)
    dem = sc.hills_dem(nx, ny, dx, rng, 6, n_gullies=2)
    geo = tr.GeoTransform(0.0, ny * dx, dx, dx)
    base = tr.TerrainBase.build(tr.FineTerrain.new(dem, None, 0.035, geo), 1, [{"kind": "transmissive", "side": "east"}])
    p = sc.SolverParams(900.0, 30.0, rain=sc.mm_h([[0, 0], [300, 80], [900, 0]]), infil_capacity=1.4e-6)
    dts = po.run(oracle_terrain(base), oracle_params(p), precision=32, want_dts=True)["dts"]
    t = np.cumsum(dts)
    syncs = np.nonzero(np.abs(t / 30.0 - np.round(t / 30.0)) < 1e-9)[0]
    p.locked_dt = dts.copy()
    p.locked_dt[(syncs[7] + syncs[8]) // 2] = 50.0
    edits = [tr.materialize(sc.random_primitives(rng, 3, geo, nx, ny), geo, nx, ny, None, 1.0) for _ in range(4)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    mon = np.arange(0, nx * ny, 11, dtype=np.uint32)
    s = _solver(9, 12, 3, chunk=7)   # 9-row bands to row 36, then 3-row bands
    try:
        sess = BatchSession(s, p, [(base, terrs, mon, None)], None)
        try:
            sess.run()
            res = sess.results()
        finally:
            sess.close()
    finally:
        s.close()
    for ci, e in enumerate(edits):
        r = res[ci]
        assert_run_equal(r, oracle_run(base, p, edit=e.as_dict(), monitors=mon), f"locked cand {ci}")
        assert r.dt_fallback_t == t[syncs[7]]


def test_default_layout_in_a_launch_of_many_rounds():
    """The defaults for a batch whose bands take many rounds of workgroups: its launch list
    is split by candidates into two lists on two streams, and a list that is not split gets
    short end bands.  Every layout gives the same bits: defaults, whole list without short
    bands, whole list with its default short bands (4-row bands, then 2-row ones), three
    lists with an explicit short-band layout, and four lists."""
    w = sc.config2()
    cands = (w.candidates * 8)[:400]
    p = copy.copy(w.params)
    p.duration_s = 600.0
    _say(f"case: {len(cands)} candidates x {w.base.nx}x{w.base.ny}, band_rows 4, launch-list splits and end-band layouts")
    runs = {}
    for tag, split, T, R2 in (("default", 0, -1, 0), ("whole", 1, 0, 0), ("whole, short end bands", 1, -1, 0),
                              ("3 lists, explicit", 3, 9, 2), ("4 lists", 4, 0, 0)):
        s = _solver(4, T, R2, chunk=16)
        s.set_option("list_split", split)
        runs[tag] = s.run_batch(p, cands, w.monitors, want_state=True)
        s.close()
    for tag in runs:
        if tag == "whole":
            continue
        for i, (a, b) in enumerate(zip(runs[tag], runs["whole"])):
            assert a.steps == b.steps, f"{tag} cand {i}: steps"
            for k in ("depth", "qx", "qy", "monitor_max", "monitor_tmax", "ledger"):
                x, y = np.asarray(getattr(a, k)), np.asarray(getattr(b, k))
                assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"{tag} cand {i}: {k} differs"
