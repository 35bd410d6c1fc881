"""Parity of the HIP local-inertial step (the screening physics of coarse fidelity levels,
crates/itr-hydro/src/inertial.rs, chosen by crates/itr-opt/src/problem.rs:178-181) with
the oracle's f32 local-inertial run, through the C ABI: state, steps, t_end, dt
statistics, monitors and the ledger bitwise, except the two cell-by-cell f64 sums
(infiltration, CFL rate) which agree to 1e-12 (conftest.assert_run_equal, li=True)."""
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


def li(p):
    q = copy.copy(p)
    q.kernel = "inertial"
    return q


def oracle_run(base, p, edit=None, monitors=(), observer=None):
    return po.run(oracle_terrain(base), oracle_params(p), precision=32, edit=edit, monitors=monitors,
                  observer=observer)


def _equivalence_base():
    """West inflow, transmissive east, south stage above the bed, interior walls
    (crates/itr-hydro/tests/equivalence.rs:38-53)."""
    segs = [{"kind": "inflow", "side": "west", "hydrograph": [[0.0, 0.0], [200.0, 1.5], [500.0, 0.2]],
             "range_m": (20.0, 40.0)},
            {"kind": "transmissive", "side": "east"},
            {"kind": "stage", "side": "south", "stage_m": 100.3, "range_m": (5.0, 25.0)}]
    zf = lambda x, y: (100.0 + 0.5 * math.sin(x / 9.0) * math.cos(y / 7.0) + 0.004 * (120.0 - x)
                       + (0.6 if 60.0 <= x < 64.0 else 0.0))
    return sc.validation_terrain(57, 43, 2.0, zf, walls=lambda c, r: 30 <= c < 34 and 10 <= r < 18, n=0.035,
                                 segments=segs)


@pytest.mark.parametrize("band_rows", [0, 1, 3])
def test_li_boundaries_walls_edits(band_rows):
    """Every boundary kind (inflow discharge, stage ghost depth, outflow-only transmissive,
    walls), rain and infiltration, and edited candidates in one batch."""
    _say(f"case: local inertial, equivalence fixture, band_rows {band_rows}")
    from inverse_terrain_engine_amd.gpu import GpuSolver
    s = GpuSolver(0, band_rows=band_rows, steps_per_chunk=5)
    base = _equivalence_base()
    p = li(sc.SolverParams(900.0, 60.0, infil_capacity=10.0 / 1000.0 / 3600.0,
                           rain=sc.mm_h([[0.0, 0.0], [300.0, 0.0], [400.0, 80.0], [700.0, 0.0]])))
    mon = np.array([c for c in range(57 * 43) if c % 7 == 3], np.uint32)
    dz = np.array([0.05 * (1 + k % 5) for k in range(18)]).reshape(3, 6)
    edits = [None, tr.TerrainEdit(c0=40, r0=5, dz=dz, fill_m3=float(dz.sum()) * 4.0),
             tr.TerrainEdit(c0=2, r0=40, dz=-0.4 * np.ones((3, 10)), cut_m3=12 * 4.0)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    runs = s.run_batch(p, terrs, mon)
    s.close()
    for i, (r, e) in enumerate(zip(runs, edits)):
        o = oracle_run(base, p, edit=None if e is None else e.as_dict(), monitors=mon)
        assert o["rows"][-1][3] > 0.0 and o["rows"][-1][4] > 0.0 and o["rows"][-1][5] > 0.0
        assert_run_equal(r, o, f"li candidate {i}", li=True)
    assert not np.array_equal(runs[0].depth, runs[2].depth)


@pytest.mark.parametrize("nx,ny", [(1, 5), (5, 1), (2, 3), (60, 7), (61, 9), (62, 4), (122, 6), (123, 11), (200, 3)])
def test_li_shapes_at_wave_boundaries(solver, nx, ny):
    """Row widths around the 61-column wave tiling, single rows and columns."""
    _say(f"case: local inertial shape {nx}x{ny}")
    base = sc.validation_terrain(nx, ny, 2.0, lambda x, y: 5.0 - 0.01 * x + 0.004 * y + 0.2 * math.sin(x / 7.0),
                                 n=0.03, segments=[{"kind": "transmissive", "side": "east"},
                                                   {"kind": "transmissive", "side": "north"}])
    p = li(sc.SolverParams(200.0, 50.0, rain=sc.mm_h([[0, 120], [100, 40], [200, 0]]), initial_depth=0.01,
                           infil_capacity=3e-6))
    mon = np.arange(0, nx * ny, 3, dtype=np.uint32)
    g = solver.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    assert_run_equal(g, oracle_run(base, p, monitors=mon), f"li {nx}x{ny}", li=True)


def test_li_lake_at_rest(solver):
    """inertial.rs (tests):9-25 on the device: a still lake over a wavy bed with a step
    stays at rest, bit for bit the oracle's f32 run."""
    _say("case: local inertial lake at rest 48x36")
    base = sc.validation_terrain(48, 36, 1.0, lambda x, y: 0.3 * math.sin(x / 7.0) * math.cos(y / 5.0)
                                 + (0.4 if x > 24.0 else 0.0), n=0.03)
    p = li(sc.SolverParams(120.0, 10.0, initial_stage=1.5))
    g = solver.run_batch(p, [tr.PreparedTerrain(base)], np.zeros(0, np.uint32))[0]
    assert_run_equal(g, oracle_run(base, p), "li lake", li=True)
    assert max(np.abs(g.qx).max(), np.abs(g.qy).max()) < 1e-5


def test_li_coarse_crest_levels_in_a_ladder(solver):
    """The reference's use: local inertial on the face-crest coarse level, HLL on the
    level-1 grids, all in one resident batch (one launch of each kernel per step)."""
    _say("case: ladder, local inertial on the 45x35 crest level, HLL on 90x70 and 180x140")
    lad = sc.config3(n_inst=7, duration=900.0)
    params = [li(lad.params) if base.crest else lad.params for base, _, _, _ in lad.groups]
    from inverse_terrain_engine_amd.gpu import BatchSession
    sess = BatchSession(solver, lad.params, lad.groups, None, group_params=params)
    try:
        sess.run()
        res = sess.results()
    finally:
        sess.close()
    k = 0
    for (base, cands, mon, _), p in zip(lad.groups, params):
        for _ in cands:
            o = oracle_run(base, p, edit=lad.edits[k].as_dict(), monitors=mon)
            assert_run_equal(res[k], o, f"ladder instance {k} ({base.nx}x{base.ny}, crest={base.crest})",
                             li=base.crest)
            k += 1
    assert any(b.crest for b, _, _, _ in lad.groups)


@pytest.mark.parametrize("use_graph,chunk", [(False, 3), (True, 8)])
def test_li_launch_modes_and_guard(use_graph, chunk):
    """Single launches and graph chunks give the same bits; the exact device guard abort
    (objective.rs:132-144) works on local-inertial candidates."""
    _say(f"case: local inertial launch modes graph={use_graph} chunk={chunk}")
    from inverse_terrain_engine_amd.gpu import GpuSolver
    s = GpuSolver(0, steps_per_chunk=chunk, use_graph=use_graph)
    nx, ny, dx = 60, 48, 5.0
    rng = sc.SplitMix64(21)
    dem = sc.hills_dem(nx, ny, dx, rng, 6, n_gullies=2)
    geo = tr.GeoTransform(0.0, ny * dx, dx, dx)
    base = tr.TerrainBase.build(tr.FineTerrain.new(dem, None, 0.035, geo), 1, [{"kind": "transmissive", "side": "east"}])
    p = li(sc.SolverParams(1200.0, 60.0, rain=sc.mm_h([[0, 0], [300, 80], [1200, 0]]), infil_capacity=1.4e-6))
    mon, gpos, _ = sc.asset_and_guard_cells(nx, ny, [(20, 20, 4, 4)], range(nx - 6, nx))
    base_run = s.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    assert_run_equal(base_run, oracle_run(base, p, monitors=mon), "li baseline", li=True)
    guard = {"baseline_max": base_run.monitor_max, "guard_pos": gpos, "tol": 0.002}
    edits = [tr.materialize(sc.random_primitives(rng, 3, geo, nx, ny), geo, nx, ny, None, 1.0) for _ in range(6)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    runs = s.run_batch(p, terrs, mon, guard=guard)
    s.close()
    for i, (r, e) in enumerate(zip(runs, edits)):
        obs = {"baseline_max": base_run.monitor_max, "guard_pos": gpos, "tol": 0.002, "require": True}
        o = oracle_run(base, p, edit=e.as_dict(), monitors=mon, observer=obs)
        assert_run_equal(r, o, f"li guard cand {i}", li=True)


@pytest.mark.parametrize("control_task", [0, 1])
def test_li_locked_dt_and_restart(solver, control_task):
    """A locked local-inertial schedule with the adaptive restart from the last sync
    snapshot (solver.rs:456-480), which for local inertial includes the boundary faces:
    natural failures on edited terrain and a failure forced mid-interval; the ledger and
    control block on each candidate's first band and on a control task of its own."""
    _say(f"case: local inertial, locked dt with restarts, 2 groups of 5 candidates on 60x48, control_task={control_task}")
    from inverse_terrain_engine_amd.gpu import BatchSession
    nx, ny, dx = 60, 48, 5.0
    rng = sc.SplitMix64(31)
    dem = sc.hills_dem(nx, ny, dx, rng, 6, n_gullies=2)
    geo = tr.GeoTransform(0.0, ny * dx, dx, dx)
    segs = [{"kind": "transmissive", "side": "east"},
            {"kind": "inflow", "side": "west", "hydrograph": [[0.0, 0.0], [300.0, 2.0], [900.0, 0.3]],
             "range_m": (60.0, 120.0)}]
    base = tr.TerrainBase.build(tr.FineTerrain.new(dem, None, 0.035, geo), 1, segs)
    mk = lambda: li(sc.SolverParams(900.0, 30.0, rain=sc.mm_h([[0, 0], [300, 80], [900, 0]]), infil_capacity=1.4e-6))
    dts = po.run(oracle_terrain(base), oracle_params(mk()), precision=32, want_dts=True)["dts"]
    t = np.cumsum(dts)
    syncs = np.nonzero(np.abs(t / 30.0 - np.round(t / 30.0)) < 1e-9)[0]
    pa, pb = mk(), mk()
    pa.locked_dt, pa.cfl_hard = dts, 0.37
    pb.locked_dt = dts.copy()
    pb.locked_dt[(syncs[7] + syncs[8]) // 2] = 50.0
    edits = [tr.materialize(sc.random_primitives(rng, 3, geo, nx, ny), geo, nx, ny, None, 1.0) for _ in range(5)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    mon = np.arange(0, nx * ny, 11, dtype=np.uint32)
    solver.set_option("control_task", control_task)
    try:
        sess = BatchSession(solver, pa, [(base, terrs, mon, None)] * 2, None, group_params=[pa, pb])
        try:
            sess.run()
            res = sess.results()
        finally:
            sess.close()
    finally:
        solver.set_option("control_task", -1)
    falls = []
    for gi, q in enumerate((pa, pb)):
        for ci, e in enumerate(edits):
            r = res[gi * len(edits) + ci]
            assert_run_equal(r, oracle_run(base, q, edit=e.as_dict(), monitors=mon), f"li locked {gi}/{ci}", li=True)
            falls.append((gi, r.dt_fallback_t))
    _say("fallback times:", falls)
    # group 1 falls back at the forced step unless its edits make the check fail earlier
    assert all(0.0 <= f <= t[syncs[7]] for g, f in falls if g == 1)
    assert any(f == t[syncs[7]] for g, f in falls if g == 1) and any(f < 0 for g, f in falls if g == 0)
