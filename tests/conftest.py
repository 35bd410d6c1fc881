"""Shared test helpers: package / oracle loading and the bitwise comparison of a HIP run
with the CPU oracle."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from __graft_entry__ import load_package  # noqa: E402

pkg = load_package()
import pyoracle as po  # noqa: E402  (test infrastructure: the checker)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def solver():
    from inverse_terrain_engine_amd.gpu import GpuSolver
    s = GpuSolver(0)
    print("device:", s.adapter_info(), flush=True)
    yield s
    s.close()


def oracle_terrain(base):
    """The oracle's own TerrainBase built from the same DEM, walls, roughness, segments."""
    f = base.fine
    segs = [dict(sg) for sg in base.segments]
    for side in ("north", "south", "west", "east") if not segs else ():
        codes = base.ghosts[side]["code"]
        if np.all(codes == 1):
            segs.append({"kind": "transmissive", "side": side})
        elif np.any(codes == 1):
            d = base.dx if side in ("north", "south") else base.dy
            idx = np.nonzero(codes == 1)[0]
            segs.append({"kind": "transmissive", "side": side, "range_m": (idx[0] * d, (idx[-1] + 1) * d)})
    return po.Terrain(f.dem, f.wall, f.n, dx=f.geo.dx, dy=f.geo.dy, origin=(f.geo.origin_x, f.geo.origin_y),
                      level=base.level, segments=segs)


def oracle_params(p):
    return po.SolverParams(p.duration_s, p.sync_interval_s, cfl=p.cfl, dt_max_s=p.dt_max_s, h_dry=p.h_dry,
                           h_eps=p.h_eps, rain=list(p.rain), infil_capacity=p.infil_capacity,
                           initial_depth=p.initial_depth, initial_stage=p.initial_stage, skip_dry=p.skip_dry,
                           locked_dt=p.locked_dt, cfl_hard=p.cfl_hard, inflow_scale=p.inflow_scale,
                           kernel=p.kernel)


def bits64(a):
    return np.ascontiguousarray(np.asarray(a, np.float64)).view(np.uint64)


def assert_run_equal(g, o, what="", li=False):
    """HIP GpuRun vs oracle run dict: bitwise on everything the reference run reports.

    li: a local-inertial run.  The reference adds its infiltration and CFL-rate sums cell
    by cell in f64 (inertial.rs:171-190), the device by bands; those two sums (ledger
    column 5, the residual column 7 that includes it, lts and rel_residual) must then
    agree to 1e-12 relative, everything else bitwise."""
    assert g.steps == o["steps"], f"{what}: steps {g.steps} vs {o['steps']}"
    assert bits64(g.t_end) == bits64(o["t_end"]), f"{what}: t_end {g.t_end!r} vs {o['t_end']!r}"
    for k, ok in (("depth", "h"), ("qx", "qx"), ("qy", "qy")):
        a = bits64(getattr(g, k).astype(np.float64))
        b = bits64(o[ok])
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            r, c = bad[0]
            raise AssertionError(f"{what}: {k} differs at {len(bad)} cells, first (row {r}, col {c}): "
                                 f"{getattr(g, k)[r, c]!r} vs {o[ok][r, c]!r}")
    assert np.array_equal(bits64(g.monitor_max), bits64(o["monitor_max"])), f"{what}: monitor max differs"
    assert np.array_equal(bits64(g.monitor_tmax), bits64(o["monitor_tmax"])), f"{what}: monitor tmax differs"
    assert g.sync_count == o["sync_count"], f"{what}: sync count"
    assert len(g.ledger) == len(o["rows"]), f"{what}: ledger rows {len(g.ledger)} vs {len(o['rows'])}"
    exact_cols = [0, 1, 2, 3, 4, 6] if li else list(range(8))
    if len(g.ledger):
        eq = bits64(g.ledger[:, exact_cols]) == bits64(o["rows"][:, exact_cols])
        assert eq.all(), f"{what}: ledger differs at {np.argwhere(~eq)[:3].tolist()}"
        if li:
            vol = np.abs(o["rows"][:, 1]).max() + np.abs(o["rows"][:, 5]).max() + 1e-300
            for c in (5, 7):
                err = np.abs(g.ledger[:, c] - o["rows"][:, c]).max() / vol
                assert err < 1e-12, f"{what}: ledger column {c} off by {err:.2e} of the volume"
    inexact = ("lts", "rel_residual") if li else ()
    for k in ("dt_min", "dt_max", "dt_mean", "lts", "v0", "rel_residual"):
        if k in inexact:
            a, b = getattr(g, k), o[k]
            assert abs(a - b) <= 1e-12 * max(abs(a), abs(b), 1e-300) or (k == "rel_residual" and abs(a - b) < 1e-12), \
                f"{what}: {k} {a!r} vs {b!r}"
            continue
        assert bits64(getattr(g, k)) == bits64(o[k]), f"{what}: {k} {getattr(g, k)!r} vs {o[k]!r}"
    assert g.cell_updates == o["cell_updates"], f"{what}: cell_updates {g.cell_updates} vs {o['cell_updates']}"
    assert bits64(g.dt_fallback_t) == bits64(o["dt_fallback_t"]), f"{what}: dt_fallback_t {g.dt_fallback_t} vs {o['dt_fallback_t']}"
    assert g.rows_skipped == o["rows_skipped"], f"{what}: rows_skipped"
    kind = {0: 0, 1: 1, 2: 2, 3: 3, 5: 5}[o["aborted"]]
    assert g.aborted == kind, f"{what}: aborted {g.aborted} vs {o['aborted']}"
    if kind in (1, 2, 5):
        assert g.abort_sync == o["abort_sync"], f"{what}: abort sync"
        assert bits64(g.abort_value) == bits64(o["abort_value"]), f"{what}: abort value"
    if kind == 1:
        assert g.abort_cell == o["abort_cell"], f"{what}: abort cell"
