"""The no-allocation entry point (itr_hip_run_batch_noalloc, every device buffer in the
library's static scratch) against itr_hip_run_batch, bitwise, through the C ABI: two
calls in a row, the second on a smaller batch (the scratch keeps the first call's
contents, so anything not cleared would show), and a batch too large for the scratch."""
from __future__ import annotations

import math

import numpy as np
import pytest

from inverse_terrain_engine_amd import scenario as sc
from inverse_terrain_engine_amd import terrain as tr
from inverse_terrain_engine_amd.gpu import DeviceError

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _bits(a, dt):
    return np.ascontiguousarray(a).view(dt)


def assert_same(a, b, what):
    for k in ("steps", "t_end", "cell_updates", "rows_skipped", "aborted", "abort_cell", "abort_sync", "sync_count",
              "bad", "dt_min", "dt_max", "dt_mean", "lts", "v0", "rel_residual", "abort_value", "volume_m3"):
        va, vb = getattr(a, k), getattr(b, k)
        assert _bits(np.float64(va), np.uint64) == _bits(np.float64(vb), np.uint64), (what, k, va, vb)
    for k in ("depth", "qx", "qy"):
        assert np.array_equal(_bits(getattr(a, k), np.uint32), _bits(getattr(b, k), np.uint32)), (what, k)
    for k in ("monitor_max", "monitor_tmax", "ledger"):
        assert np.array_equal(_bits(getattr(a, k), np.uint64), _bits(getattr(b, k), np.uint64)), (what, k)


def _population(nx, ny, n_cand, seed):
    """A config-2-like population: seeded hills DEM, edited candidates, an asset box and a
    guard strip on the east edge."""
    dx = 5.0
    rng = sc.SplitMix64(seed)
    dem = sc.hills_dem(nx, ny, dx, rng, 6, n_gullies=2)
    geo = tr.GeoTransform(0.0, ny * dx, dx, dx)
    base = tr.TerrainBase.build(tr.FineTerrain.new(dem, None, 0.035, geo), 1,
                                [{"kind": "transmissive", "side": "east"}])
    edits = [tr.materialize(sc.random_primitives(rng, 3, geo, nx, ny), geo, nx, ny, None, 1.0)
             for _ in range(n_cand)]
    terrs = [tr.PreparedTerrain(base)] + [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    mon, gpos, _ = sc.asset_and_guard_cells(nx, ny, [(nx // 3, ny // 3, 4, 4)], range(nx - 6, nx))
    return base, terrs, mon, gpos


def test_noalloc_equals_run_batch_two_calls(solver):
    p = sc.SolverParams(900.0, 60.0, rain=sc.mm_h([[0, 0], [300, 80], [900, 0]]), infil_capacity=1.4e-6)
    # call 1: 120 x 96, 9 candidates, device guard abort against the unedited baseline
    _say("case: no-alloc call 1, 9 candidates of 120x96 with guard")
    base, terrs, mon, gpos = _population(120, 96, 8, 41)
    base_run = solver.run_batch(p, [terrs[0]], mon)[0]
    guard = {"baseline_max": base_run.monitor_max, "guard_pos": gpos, "tol": 0.002}
    ref = solver.run_batch(p, terrs, mon, guard=guard)
    got = solver.run_batch_noalloc(p, [(base, terrs, mon, guard)])
    assert len(got) == len(ref)
    for i, (a, b) in enumerate(zip(got, ref)):
        assert_same(a, b, f"call 1 cand {i}")
    _say(f"call 1: {sum(r.aborted == 1 for r in ref)} guard aborts, steps {[r.steps for r in ref]}")

    # call 2: a smaller batch (48 x 40, 3 candidates) with a host sync callback abort
    _say("case: no-alloc call 2, 3 candidates of 48x40 with a sync-callback abort")
    base2 = sc.validation_terrain(48, 40, 2.0, lambda x, y: 3.0 - 0.01 * y + 0.2 * math.sin(x / 9.0),
                                  n=0.035, segments=[{"kind": "transmissive", "side": "south"}])
    rng = sc.SplitMix64(12)
    geo = base2.fine.geo
    terrs2 = [tr.PreparedTerrain(base2)] + [
        tr.PreparedTerrain(base2).apply_edit(tr.materialize(sc.random_primitives(rng, 2, geo, 48, 40), geo, 48, 40,
                                                            None, 1.0)) for _ in range(2)]
    mon2 = np.array(sorted([35 * 48 + 5, 20 * 48 + 20, 39 * 48 + 30]), np.uint32)
    p2 = sc.SolverParams(300.0, 50.0, rain=sc.mm_h([[0, 90], [200, 90], [250, 0]]))

    def cb(grid, cand, t, step, sync_index, mm):
        return 7.5 if (cand == 1 and sync_index == 2) else None

    ref2 = solver.run_batch(p2, terrs2, mon2, on_sync=cb)
    got2 = solver.run_batch_noalloc(p2, [(base2, terrs2, mon2, None)], on_sync=cb)
    for i, (a, b) in enumerate(zip(got2, ref2)):
        assert_same(a, b, f"call 2 cand {i}")
    assert ref2[1].aborted == 2 and ref2[1].abort_sync == 2


def test_noalloc_rejects_a_batch_beyond_the_scratch(solver):
    _say("case: no-alloc, 8 candidates of 512x512 (beyond the 64 MiB scratch)")
    base = sc.bench_case_terrain(512)
    terrs = [tr.PreparedTerrain(base)] * 8
    p = sc.bench_case_params()
    p.duration_s, p.sync_interval_s = 10.0, 5.0
    with pytest.raises(DeviceError, match="static scratch"):
        solver.run_batch_noalloc(p, [(base, terrs, np.zeros(0, np.uint32), None)], want_state=False)
    # the scratch is still usable after the rejection
    small = sc.bench_case_terrain(32)
    ref = solver.run_batch(p, [tr.PreparedTerrain(small)], np.arange(0, 1024, 7, dtype=np.uint32))
    got = solver.run_batch_noalloc(p, [(small, [tr.PreparedTerrain(small)], np.arange(0, 1024, 7, dtype=np.uint32),
                                        None)])
    assert_same(got[0], ref[0], "after rejection")
