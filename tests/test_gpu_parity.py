"""Bitwise parity of the HIP batch evaluator with the CPU oracle (the reference's f32
fused kernel restated), through the C ABI."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import assert_run_equal, bits64, oracle_params, oracle_terrain, po
from inverse_terrain_engine_amd import scenario as sc
from inverse_terrain_engine_amd import terrain as tr

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def oracle_run(base, p, edit=None, monitors=(), observer=None):
    return po.run(oracle_terrain(base), oracle_params(p), precision=32, edit=edit, monitors=monitors,
                  observer=observer)


def test_bench_case_short(solver):
    _say("case: bench_case 64x64, 120 s")
    base = sc.bench_case_terrain(64)
    p = sc.bench_case_params()
    p.duration_s, p.sync_interval_s = 120.0, 30.0
    mon = np.arange(0, 64 * 64, 13, dtype=np.uint32)
    g = solver.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    assert_run_equal(g, oracle_run(base, p, monitors=mon), "bench64")


@pytest.mark.parametrize("nx,ny", [(1, 5), (5, 1), (61, 7), (62, 9), (63, 4), (124, 6), (125, 11), (200, 3)])
def test_shapes_at_wave_boundaries(solver, nx, ny):
    _say(f"case: shape {nx}x{ny}")
    base = sc.validation_terrain(nx, ny, 2.0, lambda x, y: 5.0 - 0.01 * x + 0.004 * y + 0.2 * math.sin(x / 7.0),
                                 n=0.03, segments=[{"kind": "transmissive", "side": "east"}])
    p = sc.SolverParams(200.0, 50.0, rain=sc.mm_h([[0, 120], [100, 40], [200, 0]]), initial_depth=0.01,
                        infil_capacity=3e-6)
    mon = np.arange(0, nx * ny, 3, dtype=np.uint32)
    g = solver.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    assert_run_equal(g, oracle_run(base, p, monitors=mon), f"{nx}x{ny}")


@pytest.mark.parametrize("band_rows", [1, 3, 8, 32])
def test_band_shape_invariance(solver, band_rows):
    _say(f"case: band_rows {band_rows}")
    from inverse_terrain_engine_amd.gpu import GpuSolver
    s = GpuSolver(0, band_rows=band_rows, steps_per_chunk=7)
    base = sc.validation_terrain(70, 45, 2.0, lambda x, y: 3.0 - 0.01 * y + 0.2 * math.sin(x / 9.0),
                                 n=0.035, segments=[{"kind": "transmissive", "side": "south"}])
    p = sc.SolverParams(300.0, 50.0, rain=sc.mm_h([[0, 90], [200, 90], [250, 0]]))
    mon = np.array(sorted([35 * 70 + 5, 20 * 70 + 20, 44 * 70 + 30]), np.uint32)
    g = s.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    s.close()
    assert_run_equal(g, oracle_run(base, p, monitors=mon), f"band_rows={band_rows}")


def test_config1_forward(solver):
    """90x70 at 5 m, 40 min storm, infiltration, transmissive east, assets + guard."""
    for cfl in (0.35, 0.9):
        _say(f"case: config1 cfl={cfl}")
        w = sc.config1(cfl)
        g = solver.run_batch(w.params, w.candidates, w.monitors)[0]
        assert_run_equal(g, oracle_run(w.base, w.params, monitors=w.monitors), f"config1 cfl={cfl}")


def test_batch_matches_single_and_oracle(solver):
    """Edited candidates in one batch: each bitwise equal to its own single run and to
    the oracle (the reference's D2 test, crates/itr-gpu/tests/gpu.rs:81-110, made exact)."""
    _say("case: batch of 6 edited candidates on 48x40")
    base = sc.validation_terrain(48, 40, 2.0, lambda x, y: 3.0 - 0.01 * y + 0.2 * math.sin(x / 9.0),
                                 n=0.035, segments=[{"kind": "transmissive", "side": "south"}])
    p = sc.SolverParams(300.0, 50.0, rain=sc.mm_h([[0, 90], [200, 90], [250, 0]]))
    rng = sc.SplitMix64(11)
    geo = base.fine.geo
    edits = [None] + [tr.materialize(sc.random_primitives(rng, 3, geo, 48, 40), geo, 48, 40, None, 1.0)
                      for _ in range(5)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    mon = np.array(sorted([35 * 48 + 5, 20 * 48 + 20, 39 * 48 + 30]), np.uint32)
    batch = solver.run_batch(p, terrs, mon)
    for i, (t, e) in enumerate(zip(terrs, edits)):
        single = solver.run_batch(p, [t], mon)[0]
        assert_run_equal(batch[i], oracle_run(base, p, edit=None if e is None else e.as_dict(), monitors=mon),
                         f"cand {i}")
        assert np.array_equal(batch[i].depth.view(np.uint32), single.depth.view(np.uint32))
        assert batch[i].steps == single.steps
    assert not np.array_equal(batch[0].depth, batch[1].depth), "edits should change results"


@pytest.mark.parametrize("split", [0, 4], ids=["default-lists", "four-lists"])
def test_guard_abort_matches_oracle(solver, split):
    """Exact device-side guard abort (objective.rs:132-144) on config-2-like candidates,
    also with the candidates in four launch lists (streams)."""
    _say(f"case: guard abort, 10 candidates on 60x48, list_split {split}")
    nx, ny, dx = 60, 48, 5.0
    rng = sc.SplitMix64(21)
    dem = sc.hills_dem(nx, ny, dx, rng, 6, n_gullies=2)
    geo = tr.GeoTransform(0.0, ny * dx, dx, dx)
    base = tr.TerrainBase.build(tr.FineTerrain.new(dem, None, 0.035, geo), 1, [{"kind": "transmissive", "side": "east"}])
    p = sc.SolverParams(1200.0, 60.0, rain=sc.mm_h([[0, 0], [300, 80], [1200, 0]]), infil_capacity=1.4e-6)
    mon, gpos, apos = sc.asset_and_guard_cells(nx, ny, [(20, 20, 4, 4)], range(nx - 6, nx))
    base_run = solver.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    guard = {"baseline_max": base_run.monitor_max, "guard_pos": gpos, "tol": 0.002}
    edits = [tr.materialize(sc.random_primitives(rng, 3, geo, nx, ny), geo, nx, ny, None, 1.0) for _ in range(10)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    solver.set_option("list_split", split)
    try:
        runs = solver.run_batch(p, terrs, mon, guard=guard)
    finally:
        solver.set_option("list_split", 0)
    n_abort = 0
    for i, (r, e) in enumerate(zip(runs, edits)):
        obs = {"baseline_max": base_run.monitor_max, "guard_pos": gpos, "tol": 0.002, "require": True}
        o = oracle_run(base, p, edit=e.as_dict(), monitors=mon, observer=obs)
        assert_run_equal(r, o, f"guard cand {i}")
        n_abort += r.aborted == 1
    _say(f"guard aborts: {n_abort} of 10")


def test_sync_callback_abort(solver):
    """Host sync callback (Observer::on_sync) aborts a candidate at a chosen sync index;
    the result is the state at that sync point, as in the reference run loop."""
    _say("case: sync callback")
    base = sc.bench_case_terrain(40)
    p = sc.bench_case_params()
    p.duration_s, p.sync_interval_s = 300.0, 30.0
    mon = np.arange(0, 1600, 11, dtype=np.uint32)
    seen = []

    def cb(grid, cand, t, step, sync_index, mon_max):
        seen.append((cand, sync_index))
        return 7.5 if (cand == 1 and sync_index == 4) else None

    runs = solver.run_batch(p, [tr.PreparedTerrain(base), tr.PreparedTerrain(base)], mon, on_sync=cb)
    assert runs[0].aborted == 0 and runs[0].sync_count == 10
    assert runs[1].aborted == 2 and runs[1].sync_count == 4 and runs[1].abort_value == 7.5
    q = sc.bench_case_params()
    q.duration_s, q.sync_interval_s = 120.0, 30.0
    o = oracle_run(base, q, monitors=mon)
    assert runs[1].steps == o["steps"]
    assert np.array_equal(runs[1].depth.astype(np.float64).view(np.uint64), o["h"].view(np.uint64))
    assert (1, 4) in seen and (0, 10) in seen


def test_dry_lead_in_accounting(solver):
    """Rain starting after a dry lead-in: dry-row accounting (cell_updates, rows_skipped)
    equals the reference rule, and the skip-free GPU sweep is bitwise the same state."""
    _say("case: dry lead-in")
    base = sc.validation_terrain(57, 43, 2.0, lambda x, y: 100.0 + 0.5 * math.sin(x / 9.0) * math.cos(y / 7.0)
                                 + 0.004 * (120.0 - x), n=0.035, segments=[{"kind": "transmissive", "side": "east"}])
    p = sc.SolverParams(900.0, 60.0, rain=sc.mm_h([[0, 0], [300, 0], [400, 80], [700, 0]]), infil_capacity=2.8e-6)
    mon = np.array([c for c in range(57 * 43) if c % 7 == 3], np.uint32)
    g = solver.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    o = oracle_run(base, p, monitors=mon)
    assert o["rows_skipped"] > 0
    assert_run_equal(g, o, "dry lead-in")


def test_crest_level(solver):
    """Coarse fidelity level (FaceMode::Crest, level 2)."""
    _say("case: crest level 2")
    nx, ny = 64, 48
    geo = tr.GeoTransform(0.0, float(ny), 1.0, 1.0)
    dem = np.array([[0.2 * math.sin(c / 9.0) + (1.0 if c == 30 else 0.0) + 0.01 * r for c in range(nx)]
                    for r in range(ny)])
    base = tr.TerrainBase.build(tr.FineTerrain.new(dem, None, 0.03, geo), 2, [])
    p = sc.SolverParams(120.0, 30.0, rain=sc.mm_h([[0, 60], [120, 0]]), initial_stage=0.4)
    mon = np.arange(0, base.nx * base.ny, 5, dtype=np.uint32)
    g = solver.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    assert_run_equal(g, oracle_run(base, p, monitors=mon), "crest")


def test_walls_and_wide_grid(solver):
    """512 columns (9 waves per row) with interior wall cells."""
    _say("case: 512x24 with walls")
    base = sc.validation_terrain(512, 24, 2.0, lambda x, y: 10.0 + 0.002 * (1024 - x) + 0.3 * math.sin(x / 23.0),
                                 walls=lambda c, r: 200 <= c < 210 and 5 <= r < 15, n=0.035)
    p = sc.SolverParams(60.0, 20.0, rain=sc.mm_h([[0, 80], [60, 80]]), initial_depth=0.02)
    mon = np.arange(0, 512 * 24, 29, dtype=np.uint32)
    g = solver.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    assert_run_equal(g, oracle_run(base, p, monitors=mon), "wide")


@pytest.mark.parametrize("nx,ny,band_rows,level", [(80, 37, 5, 1), (130, 29, 3, 1), (200, 23, 4, 1),
                                                   (70, 41, 7, 2), (126, 33, 6, 2)])
def test_tail_wave_row_groups(nx, ny, band_rows, level):
    """Grids whose last wave updates few columns (8..18 of 62): that wave marches 3 to 8
    row groups of each band side by side.  Band heights not divisible by the group count,
    a short last band, walls on three sides and inside, a lake draining down a slope, both
    face modes, monitors and ledger rows, bitwise against the oracle."""
    _say(f"case: tail-wave row groups {nx}x{ny}, band_rows {band_rows}, level {level}")
    from inverse_terrain_engine_amd.gpu import GpuSolver
    s = GpuSolver(0, band_rows=band_rows, steps_per_chunk=6)
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
    o = oracle_run(base, p, monitors=mon)
    assert_run_equal(g, o, f"tail groups {nx}x{ny} R={band_rows} level {level}")


def _equivalence_base():
    """The boundary fixture of crates/itr-hydro/tests/equivalence.rs:38-53: a west inflow
    segment (hydrograph rising from 0), a transmissive east side, a south stage segment
    above the bed, and interior walls."""
    segs = [{"kind": "inflow", "side": "west", "hydrograph": [[0.0, 0.0], [200.0, 1.5], [500.0, 0.2]],
             "range_m": (20.0, 40.0)},
            {"kind": "transmissive", "side": "east"},
            {"kind": "stage", "side": "south", "stage_m": 100.3, "range_m": (5.0, 25.0)}]
    zf = lambda x, y: (100.0 + 0.5 * math.sin(x / 9.0) * math.cos(y / 7.0) + 0.004 * (120.0 - x)
                       + (0.6 if 60.0 <= x < 64.0 else 0.0))
    return sc.validation_terrain(57, 43, 2.0, zf, walls=lambda c, r: 30 <= c < 34 and 10 <= r < 18, n=0.035,
                                 segments=segs)


@pytest.mark.parametrize("band_rows", [0, 3])
def test_stage_and_inflow_boundaries(band_rows):
    """Stage and Inflow ghosts (boundary.rs:43-47, 66-70), which the reference GPU declines:
    the equivalence fixture with a dry lead-in (dry-row rule with wet ghosts), and edited
    candidates whose edits reach the stage cells (per-candidate ghost depth)."""
    _say(f"case: stage + inflow boundaries, band_rows {band_rows}")
    from inverse_terrain_engine_amd.gpu import GpuSolver
    s = GpuSolver(0, band_rows=band_rows, steps_per_chunk=5)
    base = _equivalence_base()
    p = sc.SolverParams(900.0, 60.0, infil_capacity=10.0 / 1000.0 / 3600.0,
                        rain=sc.mm_h([[0.0, 0.0], [300.0, 0.0], [400.0, 80.0], [700.0, 0.0]]))
    mon = np.array([c for c in range(57 * 43) if c % 7 == 3], np.uint32)
    dz = np.array([0.05 * (1 + k % 5) for k in range(18)]).reshape(3, 6)
    edits = [None, tr.TerrainEdit(c0=40, r0=5, dz=dz, fill_m3=float(dz.sum()) * 4.0),
             tr.TerrainEdit(c0=2, r0=40, dz=-0.4 * np.ones((3, 10)), cut_m3=12 * 4.0)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    runs = s.run_batch(p, terrs, mon)
    s.close()
    for i, (r, e) in enumerate(zip(runs, edits)):
        o = oracle_run(base, p, edit=None if e is None else e.as_dict(), monitors=mon)
        assert o["rows_skipped"] > 0 and o["rows"][-1][3] > 0.0
        assert_run_equal(r, o, f"stage+inflow candidate {i}")
    assert not np.array_equal(runs[0].depth, runs[2].depth)


@pytest.mark.parametrize("scale", [0.5, 2.0])
def test_inflow_scale_and_other_sides(solver, scale):
    """Ensemble inflow scaling (solver.rs:47-48; the reference GPU declines it,
    backend.rs:53-55) with inflow through the north side, stage on the east side and a
    transmissive south: the north / south ghost rows and the east ghost column."""
    _say(f"case: inflow scale {scale}, north inflow, east stage")
    segs = [{"kind": "inflow", "side": "north", "hydrograph": [[0.0, 0.4], [150.0, 2.0], [400.0, 0.0]],
             "range_m": (30.0, 70.0)},
            {"kind": "stage", "side": "east", "stage_m": 50.35, "range_m": (10.0, 60.0)},
            {"kind": "transmissive", "side": "south"}]
    base = sc.validation_terrain(75, 40, 2.0, lambda x, y: 50.0 + 0.003 * (y - 40.0) + 0.15 * math.sin(x / 11.0),
                                 n=0.03, segments=segs)
    p = sc.SolverParams(420.0, 60.0, rain=sc.mm_h([[0.0, 30.0], [420.0, 0.0]]), infil_capacity=1e-6)
    p.inflow_scale = scale
    mon = np.arange(0, 75 * 40, 11, dtype=np.uint32)
    g = solver.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    o = oracle_run(base, p, monitors=mon)
    assert o["rows"][-1][3] > 0.0
    assert_run_equal(g, o, f"inflow scale {scale}")


def test_batch_backend_runs_ensembles(solver):
    """HipBatchBackend.run (BatchBackend<f32>) takes inflow-scaled ensemble members and
    boundary segments instead of declining them."""
    from inverse_terrain_engine_amd.gpu import HipBatchBackend
    base = _equivalence_base()
    p = sc.SolverParams(240.0, 60.0, rain=sc.mm_h([[0.0, 40.0]]))
    p.inflow_scale = 1.4
    mon = np.arange(0, 57 * 43, 17, dtype=np.uint32)
    runs = HipBatchBackend(solver).run(p, [tr.PreparedTerrain(base)], mon)
    o = oracle_run(base, p, monitors=mon)
    assert runs[0].steps == o["steps"] and runs[0].cell_updates == o["steps"] * 57 * 43
    assert np.array_equal(runs[0].monitor_max.view(np.uint64), o["monitor_max"].view(np.uint64))


def test_declines_unsupported(solver):
    """Rows wider than one workgroup (992 columns) decline, as the reference's Err does."""
    from inverse_terrain_engine_amd.gpu import Declined
    base = sc.validation_terrain(1000, 2, 1.0, lambda x, y: 0.0, n=0.03)
    with pytest.raises(Declined):
        solver.run_batch(sc.SolverParams(10.0, 5.0), [tr.PreparedTerrain(base)], np.zeros(0, np.uint32))


@pytest.mark.parametrize("control_task", [0, 1])
def test_fidelity_ladder_one_launch(solver, control_task):
    """Grids of three widths and both face modes (45x35 crest, 90x70 and 180x140 signed)
    in one batch (three launch lists on three streams), with the ledger and control block
    on each candidate's first band or on a control task of its own: every instance
    bitwise equal to the oracle."""
    _say(f"case: fidelity ladder 4/2/1 instances, 900 s, control_task={control_task}")
    lad = sc.config3(n_inst=7, duration=900.0)
    solver.set_option("control_task", control_task)
    try:
        sess = solver.session(lad.params, lad.groups)
        try:
            sess.run()
            res = sess.results()
        finally:
            sess.close()
    finally:
        solver.set_option("control_task", -1)
    k = 0
    for base, cands, mon, _ in lad.groups:
        for _ in cands:
            o = oracle_run(base, lad.params, edit=lad.edits[k].as_dict(), monitors=mon)
            assert_run_equal(res[k], o, f"ladder instance {k} ({base.nx}x{base.ny}, crest={base.crest})")
            k += 1


@pytest.mark.parametrize("use_graph,chunk", [(False, 5), (True, 4), (True, 12)])
def test_launch_modes(use_graph, chunk):
    """Single launches and graph chunks of several sizes give the same bits."""
    _say(f"case: use_graph={use_graph} chunk={chunk}")
    from inverse_terrain_engine_amd.gpu import GpuSolver
    s = GpuSolver(0, steps_per_chunk=chunk, use_graph=use_graph)
    base = sc.bench_case_terrain(48)
    p = sc.bench_case_params()
    p.duration_s, p.sync_interval_s = 90.0, 30.0
    mon = np.arange(0, 48 * 48, 17, dtype=np.uint32)
    g = s.run_batch(p, [tr.PreparedTerrain(base)], mon)[0]
    s.close()
    assert_run_equal(g, oracle_run(base, p, monitors=mon), f"graph={use_graph} chunk={chunk}")


def test_advance_then_run(solver):
    """A resident batch advanced by a step count off the launch grid, then run to the
    end, equals one uninterrupted run."""
    _say("case: advance 13 then run")
    base = sc.bench_case_terrain(40)
    p = sc.bench_case_params()
    p.duration_s, p.sync_interval_s = 90.0, 30.0
    mon = np.arange(0, 1600, 9, dtype=np.uint32)
    sess = solver.session(p, [(base, [tr.PreparedTerrain(base)], mon, None)])
    try:
        st = sess.advance(13)
        assert st["cell_steps"] == 13 * 1600
        sess.run()
        g = sess.results()[0]
    finally:
        sess.close()
    assert_run_equal(g, oracle_run(base, p, monitors=mon), "advance+run")


def test_autotune_then_run_ladder():
    """The autotuner picks band heights per grid shape on a fidelity ladder; a batch
    created afterwards uses them and stays bitwise equal to the oracle."""
    _say("case: autotune on a 14-instance ladder")
    from inverse_terrain_engine_amd.gpu import GpuSolver
    s = GpuSolver(0)
    lad = sc.config3(n_inst=14, duration=600.0)
    tuned = s.autotune(lad.params, lad.groups, trial_steps=8)
    _say(f"tuned band rows {tuned}")
    assert len(tuned) == 3 and all(r >= 1 for r, _ in tuned)
    sess = s.session(lad.params, lad.groups)
    try:
        st = sess.run()
        res = sess.results()
    finally:
        sess.close()
        s.close()
    k = 0
    for base, cands, mon, _ in lad.groups:
        for i in range(len(cands)):
            if i < 2:
                o = oracle_run(base, lad.params, edit=lad.edits[k].as_dict(), monitors=mon)
                assert_run_equal(res[k], o, f"tuned ladder instance {k} ({base.nx}x{base.ny})")
            k += 1


def test_short_sqrt_and_reciprocal_are_ieee(solver):
    """The step's short correctly-rounded sequences equal the IEEE operations on every
    one of the 2^32 f32 inputs (the guarded ranges plus the IEEE fallbacks)."""
    _say("case: exhaustive reciprocal / sqrt check")
    assert solver.math_selfcheck() == (0, 0)


def test_ensemble_members_one_batch(solver):
    """Ensemble members (problem.rs:56-63 scaled forcing; evaluated on the CPU by the
    reference, evaluator.rs:494-560) as grid entries of one batch with their own
    parameters, including a member with another duration and sync interval: every
    (member, candidate) bitwise equal to the oracle run under that member's parameters."""
    _say("case: ensemble of 4 members x 2 candidates, one batch")
    base = _equivalence_base()
    p = sc.SolverParams(600.0, 60.0, infil_capacity=5.0 / 1000.0 / 3600.0,
                        rain=sc.mm_h([[0.0, 20.0], [200.0, 70.0], [500.0, 0.0]]))
    members = [sc.scaled(p, 1.0, 1.0), sc.scaled(p, 1.4, 1.0), sc.scaled(p, 0.6, 2.0), sc.scaled(p, 1.0, 0.5)]
    members[3].duration_s, members[3].sync_interval_s = 420.0, 45.0
    mon = np.array([c for c in range(57 * 43) if c % 5 == 1], np.uint32)
    dz = np.array([0.05 * (1 + k % 5) for k in range(18)]).reshape(3, 6)
    edits = [None, tr.TerrainEdit(c0=40, r0=5, dz=dz, fill_m3=float(dz.sum()) * 4.0)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    runs = solver.run_ensemble(members, terrs, mon)
    for m, mp in enumerate(members):
        for c, e in enumerate(edits):
            o = oracle_run(base, mp, edit=None if e is None else e.as_dict(), monitors=mon)
            assert_run_equal(runs[m][c], o, f"member {m} candidate {c}")
    assert runs[1][0].steps != runs[0][0].steps or not np.array_equal(runs[1][0].depth, runs[0][0].depth)


def test_full_domain_rasters(solver):
    """Full-domain run (crates/itr-cli/src/run.rs:73-93): every cell monitored, peak and
    time-to-peak rasters (NaN on walls; tpeak NaN where the peak is not positive)."""
    _say("case: full-domain peak rasters 70x45")
    base = sc.validation_terrain(70, 45, 2.0, lambda x, y: 3.0 - 0.01 * y + 0.2 * math.sin(x / 9.0),
                                 walls=lambda c, r: 20 <= c < 26 and 10 <= r < 30, n=0.035,
                                 segments=[{"kind": "transmissive", "side": "south"}])
    p = sc.SolverParams(300.0, 50.0, rain=sc.mm_h([[0, 0], [60, 0], [90, 90], [200, 90], [250, 0]]))
    run, peak, tpeak = solver.full_run(p, tr.PreparedTerrain(base))
    mon = np.arange(70 * 45, dtype=np.uint32)
    o = oracle_run(base, p, monitors=mon)
    assert_run_equal(run, o, "full domain")
    wall = base.wall
    assert np.isnan(peak[wall]).all() and np.isnan(tpeak[wall]).all()
    assert np.array_equal(peak[~wall], o["monitor_max"].reshape(45, 70)[~wall])
    pos = (~wall) & (o["monitor_max"].reshape(45, 70) > 0.0)
    assert pos.any() and np.array_equal(tpeak[pos], o["monitor_tmax"].reshape(45, 70)[pos])
    assert np.isnan(tpeak[(~wall) & ~pos]).all()


def _runs_equal(a, b, what):
    """Two GpuRuns bitwise equal on everything a run reports."""
    for k in ("depth", "qx", "qy", "monitor_max", "monitor_tmax", "ledger"):
        x, y = np.asarray(getattr(a, k), np.float64), np.asarray(getattr(b, k), np.float64)
        assert x.shape == y.shape and np.array_equal(x.view(np.uint64), y.view(np.uint64)), f"{what}: {k}"
    for k in ("steps", "cell_updates", "rows_skipped", "sync_count", "aborted"):
        assert getattr(a, k) == getattr(b, k), f"{what}: {k}"
    for k in ("t_end", "dt_min", "dt_max", "dt_mean", "v0", "rel_residual", "volume_m3"):
        assert bits64(getattr(a, k)) == bits64(getattr(b, k)), f"{what}: {k}"


def test_device_designs_equal_host_prepared(solver):
    """Candidate designs materialized on the device (DesignSpace::materialize and
    PreparedTerrain::apply_edit, design.rs:316-385, terrain.rs:460-505) run bitwise like the
    same designs materialized on the host: every primitive kind, a zero-height primitive, an
    empty design, a partial editable mask, an active clamp, primitives at the grid edge; the
    edit volumes are bitwise the host's."""
    _say("case: device designs vs host-prepared terrains, 96x64")
    nx, ny = 96, 64
    base = sc.validation_terrain(nx, ny, 2.0, lambda x, y: 3.0 + 0.008 * (192.0 - x) + 0.3 * math.sin(y / 11.0),
                                 walls=lambda c, r: 40 <= c < 44 and 10 <= r < 30, n=0.035,
                                 segments=[{"kind": "transmissive", "side": "east"}])
    geo = base.fine.geo
    rng = sc.SplitMix64(21)
    kinds = ("berm", "swale", "mound", "cut")
    designs = [sc.random_primitives(rng, k, geo, nx, ny, kinds=kinds) for k in (1, 2, 3, 4, 4, 3)]
    designs.append([])
    zero = sc.random_primitives(rng, 2, geo, nx, ny, kinds=kinds)
    zero[0].height = 0.0
    designs.append(zero)
    edge = sc.random_primitives(rng, 1, geo, nx, ny, kinds=("berm",))
    edge[0].cx, edge[0].cy, edge[0].length, edge[0].width = 188.0, 125.0, 30.0, 8.0
    designs.append(edge)
    editable = np.ones((ny, nx), bool)
    editable[:, :12] = False
    max_dz = 0.35
    p = sc.SolverParams(240.0, 60.0, rain=sc.mm_h([[0, 70], [240, 30]]), infil_capacity=2e-6)
    mon = np.arange(0, nx * ny, 7, dtype=np.uint32)
    runs, vol = solver.run_designs(p, base, designs, mon, editable=editable, max_abs_dz=max_dz)
    edits = [tr.materialize(d, geo, nx, ny, editable, max_dz) for d in designs]
    host = solver.run_batch(p, [tr.PreparedTerrain(base).apply_edit(e) for e in edits], mon)
    assert any(np.abs(e.dz).max() == max_dz for e in edits if e.dz.size), "the clamp should be active"
    for i, (a, b, e) in enumerate(zip(runs, host, edits)):
        _runs_equal(a, b, f"design {i}")
        assert bits64(vol[i, 0]) == bits64(e.cut_m3) and bits64(vol[i, 1]) == bits64(e.fill_m3), f"design {i} volumes"
    assert not np.array_equal(runs[0].depth, runs[1].depth)


@pytest.mark.parametrize("level", [2, 3])
def test_device_designs_on_crest_levels(solver, level):
    """Designs materialized on the device for a face-crest coarse level: the edit on the
    fine grid, the coarse cells (coarse_cell, terrain.rs:357-373) and both crest rises of
    every face (terrain.rs:378-429) built per candidate, bitwise like the host's
    PreparedTerrain::apply_edit; fine grids not divisible by the level, fine walls (one
    coarse cell all walls), an editable mask, an active clamp, primitives at the edge."""
    fnx, fny = 100, 64
    _say(f"case: device designs on level {level} of {fnx}x{fny}")
    walls = lambda c, r: (40 <= c < 46 and 10 <= r < 30) or (c < 3 and r < 3)
    fine_base = sc.validation_terrain(fnx, fny, 2.0, lambda x, y: 3.0 + 0.008 * (200.0 - x) + 0.3 * math.sin(y / 11.0)
                                      + 0.2 * math.cos(x / 5.0), walls=walls, n=0.035,
                                      segments=[{"kind": "transmissive", "side": "east"}])
    fine = fine_base.fine
    base = tr.TerrainBase.build(fine, level, [{"kind": "transmissive", "side": "east"}])
    assert base.crest and base.nx == -(-fnx // level) and base.ny == -(-fny // level)
    geo = fine.geo
    rng = sc.SplitMix64(33 + level)
    kinds = ("berm", "swale", "mound", "cut")
    designs = [sc.random_primitives(rng, k, geo, fnx, fny, kinds=kinds) for k in (1, 2, 3, 4, 3)]
    designs.append([])
    edge = sc.random_primitives(rng, 1, geo, fnx, fny, kinds=("berm",))
    edge[0].cx, edge[0].cy, edge[0].length, edge[0].width = 196.0, 125.0, 30.0, 8.0
    designs.append(edge)
    editable = np.ones((fny, fnx), bool)
    editable[:, :10] = False
    max_dz = 0.35
    p = sc.SolverParams(240.0, 60.0, rain=sc.mm_h([[0, 70], [240, 30]]), infil_capacity=2e-6)
    mon = np.arange(0, base.nx * base.ny, 5, dtype=np.uint32)
    runs, vol = solver.run_designs(p, base, designs, mon, editable=editable, max_abs_dz=max_dz)
    edits = [tr.materialize(d, geo, fnx, fny, editable, max_dz) for d in designs]
    host = solver.run_batch(p, [tr.PreparedTerrain(base).apply_edit(e) for e in edits], mon)
    assert any(np.abs(e.dz).max() == max_dz for e in edits if e.dz.size), "the clamp should be active"
    for i, (a, b, e) in enumerate(zip(runs, host, edits)):
        _runs_equal(a, b, f"level {level} design {i}")
        assert bits64(vol[i, 0]) == bits64(e.cut_m3) and bits64(vol[i, 1]) == bits64(e.fill_m3), f"design {i} volumes"
    assert not np.array_equal(runs[0].depth, runs[1].depth)


@pytest.mark.parametrize("level", [1, 2])
def test_device_designs_with_stage_levels(solver, level):
    """Device designs on levels whose state depends on each candidate's own levels: a Stage
    boundary segment (ghost depth f32(max(stage - z_level, 0)) of the edited boundary cell,
    boundary.rs:43, 66) and initial_stage (init_state, solver.rs:287-312), with edits that
    reach the stage cells; bitwise like host-prepared terrains."""
    fnx, fny = 80, 56
    _say(f"case: device designs with a stage boundary and initial stage, level {level}")
    segs = [{"kind": "stage", "side": "south", "stage_m": 3.35, "range_m": (20.0, 130.0)},
            {"kind": "transmissive", "side": "east"}]
    fb = sc.validation_terrain(fnx, fny, 2.0, lambda x, y: 3.0 + 0.01 * y + 0.3 * math.sin(x / 9.0), n=0.035,
                               walls=lambda c, r: 30 <= c < 34 and 20 <= r < 30, segments=segs)
    base = fb if level == 1 else tr.TerrainBase.build(fb.fine, level, segs)
    geo = fb.fine.geo
    rng = sc.SplitMix64(41 + level)
    designs = [sc.random_primitives(rng, k, geo, fnx, fny, kinds=("berm", "swale", "mound", "cut")) for k in (2, 3, 4)]
    south = sc.random_primitives(rng, 1, geo, fnx, fny, kinds=("mound",))
    south[0].cx, south[0].cy, south[0].length, south[0].width, south[0].height = 60.0, 3.0, 40.0, 10.0, 0.4
    designs.append(south)
    designs.append([])
    p = sc.SolverParams(200.0, 50.0, rain=sc.mm_h([[0, 60], [200, 20]]), initial_stage=3.15, infil_capacity=1e-6)
    mon = np.arange(0, base.nx * base.ny, 6, dtype=np.uint32)
    runs, vol = solver.run_designs(p, base, designs, mon, max_abs_dz=0.5)
    edits = [tr.materialize(d, geo, fnx, fny, None, 0.5) for d in designs]
    host = solver.run_batch(p, [tr.PreparedTerrain(base).apply_edit(e) for e in edits], mon)
    for i, (a, b, e) in enumerate(zip(runs, host, edits)):
        _runs_equal(a, b, f"level {level} stage design {i}")
        assert bits64(vol[i, 0]) == bits64(e.cut_m3) and bits64(vol[i, 1]) == bits64(e.fill_m3), f"design {i} volumes"
    assert not np.array_equal(runs[3].depth, runs[4].depth), "the south edit should change the run"


@pytest.mark.parametrize("control_task", [0, 1])
def test_locked_dt_schedule_and_adaptive_restart(solver, control_task):
    """Baseline-locked dt (solver.rs:452-455) and its fallback: a step whose locked dt
    fails the hard-CFL check restarts adaptively from the last sync snapshot
    (solver.rs:456-480).  Failures from edited terrain (cfl_hard 0.37), forced right after
    a sync step (the snapshot is the state just completed), forced mid-interval (state,
    monitors, ledger and control roll back), and at the first step; each group has its
    own parameters in one resident batch.  Run with the ledger and control block on each
    candidate's first band and on a control task of its own."""
    _say(f"case: locked dt schedules with restarts, 4 groups of 6 candidates on 60x48, control_task={control_task}")
    from inverse_terrain_engine_amd.gpu import BatchSession
    nx, ny, dx = 60, 48, 5.0
    rng = sc.SplitMix64(31)
    dem = sc.hills_dem(nx, ny, dx, rng, 6, n_gullies=2)
    geo = tr.GeoTransform(0.0, ny * dx, dx, dx)
    base = tr.TerrainBase.build(tr.FineTerrain.new(dem, None, 0.035, geo), 1, [{"kind": "transmissive", "side": "east"}])
    mk = lambda: sc.SolverParams(900.0, 30.0, rain=sc.mm_h([[0, 0], [300, 80], [900, 0]]), infil_capacity=1.4e-6)
    dts = po.run(oracle_terrain(base), oracle_params(mk()), precision=32, want_dts=True)["dts"]
    t = np.cumsum(dts)
    syncs = np.nonzero(np.abs(t / 30.0 - np.round(t / 30.0)) < 1e-9)[0]
    pa, pb, pc, pd = mk(), mk(), mk(), mk()
    pa.locked_dt, pa.cfl_hard = dts, 0.37
    pb.locked_dt = dts.copy()
    pb.locked_dt[syncs[5] + 1] = 50.0
    pc.locked_dt = dts.copy()
    pc.locked_dt[(syncs[7] + syncs[8]) // 2] = 50.0
    pd.locked_dt, pd.cfl_hard = dts, 0.0
    edits = [tr.materialize(sc.random_primitives(rng, 3, geo, nx, ny), geo, nx, ny, None, 1.0) for _ in range(6)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    mon = np.arange(0, nx * ny, 11, dtype=np.uint32)
    params = [pa, pb, pc, pd]
    solver.set_option("control_task", control_task)
    try:
        sess = BatchSession(solver, pa, [(base, terrs, mon, None)] * 4, None, group_params=params)
        try:
            sess.run()
            res = sess.results()
        finally:
            sess.close()
    finally:
        solver.set_option("control_task", -1)
    fallbacks = []
    for gi, q in enumerate(params):
        for ci, e in enumerate(edits):
            r = res[gi * len(edits) + ci]
            o = oracle_run(base, q, edit=e.as_dict(), monitors=mon)
            assert_run_equal(r, o, f"locked group {gi} cand {ci}")
            fallbacks.append((gi, r.dt_fallback_t))
    by = {g: [f for gg, f in fallbacks if gg == g] for g in range(4)}
    _say("fallback times:", by)
    assert any(f >= 0 for f in by[0]) and any(f < 0 for f in by[0])
    assert all(f == t[syncs[5]] for f in by[1]) and all(f == t[syncs[7]] for f in by[2])
    assert all(f == 0.0 for f in by[3])
