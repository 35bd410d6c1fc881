"""Subdomain replay on the device (crates/itr-hydro/src/replay.rs, spec §7.11.3): the
tape recorded by the baseline run, and candidates replayed on the region of interest,
bitwise against the oracle's restatement (itself pinned by the reference's replay suite,
crates/itr-hydro/tests/replay.rs)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import assert_run_equal, oracle_params, oracle_terrain, po
from inverse_terrain_engine_amd import scenario as sc
from inverse_terrain_engine_amd import terrain as tr

pytestmark = pytest.mark.gpu

NX, NY = 72, 56
ROI = (20, 14, 30, 26)


def _say(*a):
    print(*a, flush=True)


def _base():
    """crates/itr-hydro/tests/replay.rs:15-21: west inflow segment, transmissive east."""
    segs = [{"kind": "inflow", "side": "west", "hydrograph": [[0.0, 0.0], [200.0, 2.0], [600.0, 0.3]],
             "range_m": (40.0, 70.0)},
            {"kind": "transmissive", "side": "east"}]
    return sc.validation_terrain(NX, NY, 2.0, lambda x, y: 50.0 + 0.004 * (150.0 - x) + 0.3 * math.sin(x / 11.0)
                                 * math.cos(y / 9.0), n=0.035, segments=segs)


def _params():
    return sc.SolverParams(800.0, 40.0, rain=sc.mm_h([[0.0, 0.0], [100.0, 0.0], [200.0, 60.0], [500.0, 0.0]]))


def _in_roi(c, roi=ROI):
    col, row = c % NX, c // NX
    return roi[0] <= col < roi[0] + roi[2] and roi[1] <= row < roi[1] + roi[3]


def _sub(cells, roi=ROI):
    return np.array([((c // NX) - roi[1]) * roi[2] + (c % NX) - roi[0] for c in cells], np.uint32)


@pytest.mark.parametrize("roi", [ROI, (0, 0, 24, 20), (50, 30, 22, 26)])
def test_tape_recorded_on_device(solver, roi):
    """The baseline run records the ring (every step, after the ghost fill) and the band
    (every sync point) exactly as the oracle's Tape does, also for regions on the grid's
    edges, whose ring holds ghost cells (wall, transmissive, inflow)."""
    _say(f"case: record tape, roi {roi}")
    base = _base()
    mon = np.array([c for c in range(NX * NY) if c % 7 == 2], np.uint32)
    tape, run = solver.record_tape(_params(), base, roi, mon)
    o = po.record_tape(oracle_terrain(base), oracle_params(_params()), roi, precision=32, monitors=mon)
    assert tape.steps == o["steps"] and tape.syncs == o["sync_count"]
    assert tape.ring.shape == o["ring"].shape and tape.band.shape == o["band"].shape
    assert np.array_equal(tape.ring.astype(np.float64).view(np.uint64), o["ring"].view(np.uint64)), "ring differs"
    assert np.array_equal(tape.band.astype(np.float64).view(np.uint64), o["band"].view(np.uint64)), "band differs"
    full = po.run(oracle_terrain(base), oracle_params(_params()), precision=32, monitors=mon)
    assert_run_equal(run, full, "recording baseline")


def test_replay_on_region(solver):
    """crates/itr-hydro/tests/replay.rs:85-119 on the device: with the baseline's locked
    schedule, the zero-edit replay equals the baseline on the region bitwise; a small
    interior edit stays valid; a berm across the downstream edge is replay-invalid.  Every
    replay equals the oracle's replay bitwise (state, monitors, ledger, abort sync)."""
    _say("case: replay zero edit / small edit / berm on the 30x26 region")
    base = _base()
    mon = np.array([c for c in range(NX * NY) if _in_roi(c) and c % 5 == 1], np.uint32)
    sub_mon = _sub(mon)
    dts = po.run(oracle_terrain(base), oracle_params(_params()), precision=32, want_dts=True)["dts"]
    tape, run0 = solver.record_tape(_params(), base, ROI, mon)
    pl = _params()
    pl.locked_dt = dts
    c0, r0, w, h = ROI
    edits = [None, tr.TerrainEdit(c0=34, r0=26, dz=np.full((3, 3), 0.02)),
             tr.TerrainEdit(c0=c0 + w - 3, r0=r0, dz=np.full((h, 2), 1.5))]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    runs = solver.run_replay(pl, tape, terrs, sub_mon)
    otape = {"roi": ROI, "ring": tape.ring.astype(np.float64), "band": tape.band.astype(np.float64)}
    for i, (r, e) in enumerate(zip(runs, edits)):
        o = po.run_replay(oracle_terrain(base), oracle_params(pl), otape, precision=32,
                          edit=None if e is None else e.as_dict(), monitors_sub=sub_mon)
        assert_run_equal(r, o, f"replay candidate {i}")
    assert runs[0].aborted == 0 and runs[1].aborted == 0 and runs[2].aborted == 5
    assert np.array_equal(runs[0].depth.view(np.uint32), run0.depth[r0:r0 + h, c0:c0 + w].view(np.uint32))
    assert np.array_equal(runs[0].monitor_max.view(np.uint64), run0.monitor_max.view(np.uint64))
    # the band-height sweep for replay batches picks one of its heights, and a replay at
    # any height gives the same bits
    R = solver.tune_replay_rows(pl, tape, terrs, sub_mon, rows=(3, 7, 26), trial_steps=64)
    assert R in (3, 7, 26)
    for rows in (3, 7, 26):
        _say(f"case: replay at band height {rows}")
        again = solver.run_replay(pl, tape, terrs, sub_mon, band_rows=rows)
        for i, (a, b) in enumerate(zip(again, runs)):
            assert a.aborted == b.aborted and a.steps == b.steps, (rows, i)
            assert np.array_equal(a.depth.view(np.uint32), b.depth.view(np.uint32)), (rows, i)
            assert np.array_equal(a.monitor_max.view(np.uint64), b.monitor_max.view(np.uint64)), (rows, i)


def test_replay_crest_level_with_guard(solver):
    """Replay on a face-crest coarse level (level 2: two rises per face, sub-terrain faces
    copied from each candidate's crest faces), with the device guard abort running inside
    the replay as the reference's GuardObserver does (evaluator.rs:427-431)."""
    _say("case: replay on the level-2 crest grid with a guard")
    fine = _base().fine
    base = tr.TerrainBase.build(fine, 2, _base().segments)
    nx, ny = base.nx, base.ny
    roi = (10, 7, 15, 13)
    inside = lambda c: roi[0] <= c % nx < roi[0] + roi[2] and roi[1] <= c // nx < roi[1] + roi[3]
    mon = np.array([c for c in range(nx * ny) if inside(c) and c % 3 == 0], np.uint32)
    sub_mon = np.array([((c // nx) - roi[1]) * roi[2] + (c % nx) - roi[0] for c in mon], np.uint32)
    p = _params()
    dts = po.run(oracle_terrain(base), oracle_params(p), precision=32, want_dts=True)["dts"]
    tape, run0 = solver.record_tape(p, base, roi, mon)
    o = po.record_tape(oracle_terrain(base), oracle_params(p), roi, precision=32, monitors=mon)
    assert np.array_equal(tape.ring.astype(np.float64).view(np.uint64), o["ring"].view(np.uint64))
    assert np.array_equal(tape.band.astype(np.float64).view(np.uint64), o["band"].view(np.uint64))
    pl = _params()
    pl.locked_dt = dts
    rng = sc.SplitMix64(17)
    geo = fine.geo
    c0f, r0f = 2 * roi[0] + 6, 2 * roi[1] + 6
    geo_w = tr.GeoTransform(geo.origin_x + c0f * geo.dx, geo.origin_y - r0f * geo.dy, geo.dx, geo.dy)
    edits = [tr.materialize(sc.random_primitives(rng, 2, geo_w, 16, 12), geo, fine.nx, fine.ny, None, 1.0)
             for _ in range(6)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    gpos = np.arange(0, len(mon), 4, dtype=np.int32)
    guard = {"baseline_max": run0.monitor_max, "guard_pos": gpos, "tol": 0.0005}
    runs = solver.run_replay(pl, tape, terrs, sub_mon, guard=guard)
    otape = {"roi": roi, "ring": tape.ring.astype(np.float64), "band": tape.band.astype(np.float64)}
    kinds = []
    for i, (r, e) in enumerate(zip(runs, edits)):
        obs = {"baseline_max": run0.monitor_max, "guard_pos": gpos, "tol": 0.0005, "require": True}
        oo = po.run_replay(oracle_terrain(base), oracle_params(pl), otape, precision=32, edit=e.as_dict(),
                           monitors_sub=sub_mon, observer=obs)
        assert_run_equal(r, oo, f"crest replay candidate {i}")
        kinds.append(r.aborted)
    _say("abort kinds:", kinds)
