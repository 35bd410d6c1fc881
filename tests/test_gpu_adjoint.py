"""The discrete adjoint through the C ABI (itr_hip_adjoint) against f64 central finite
differences of the oracle on the frozen dt schedule (tests/golden/adjoint_fd_config4.json,
made by tests/golden/make_adjoint_fd.py).  The check uses the 1e-6 m differences where
they agree with the 1e-7 m ones (the objective is differentiable at that scale); the
1e-4 m differences cross reconstruction kinks of mm-thick rain films and are printed
for reference only."""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

from inverse_terrain_engine_amd import objective as ob
from inverse_terrain_engine_amd import scenario as sc
from inverse_terrain_engine_amd import terrain as tr

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
REL_TOL = 1e-3   # relative error bound against f64 central differences


def _say(*a):
    print(*a, flush=True)


def _fixture():
    with open(os.path.join(HERE, "golden", "adjoint_fd_config4.json")) as f:
        return json.load(f)


def _rel(a, b, scale):
    return abs(a - b) / max(abs(b), scale)


def test_adjoint_config4_against_f64_fd(solver):
    fx = _fixture()
    case = sc.config4()
    p = case.params
    p.locked_dt = np.array(fx["locked_dt"])
    _say(f"case: adjoint config4, {len(p.locked_dt)} locked steps")
    res = solver.adjoint(p, case.base, [case.candidate], case.monitors, case.assets, case.tau_m)[0]
    _say({k: v for k, v in res.items() if not isinstance(v, np.ndarray)})
    assert res["steps"] == fx["steps"]
    assert res["replay_mismatches"] == 0, "replayed segments must equal the checkpoints bitwise"
    nx, ny = case.base.nx, case.base.ny
    area = case.base.dx * case.base.dy
    objv = ob.Objective([ob.Asset(np.asarray(a), th, w) for a, th, w in case.assets], case.tau_m, case.earth_weight)
    dz0 = case.edit.dense(nx, ny)
    cut, fill = ob.volumes(dz0, area)
    j = res["j_risk"] + case.earth_weight * objv.j_earth(cut, fill)
    assert abs(j - fx["j0"]) <= 1e-4 * abs(fx["j0"]), (j, fx["j0"])
    g = res["dJ_dz_cell"] + objv.earth_gradient(dz0, area)
    fd6, fd7, fd4 = (fx["fd"][repr(h)] for h in (1e-6, 1e-7, 1e-4))
    scale = 1e-3 * np.abs(np.array(fd6["dJ_dz_cell"])).max()
    worst, checked = 0.0, 0
    for (r, c), v6, v7, v4 in zip(fx["cells"], fd6["dJ_dz_cell"], fd7["dJ_dz_cell"], fd4["dJ_dz_cell"]):
        smooth = _rel(v6, v7, scale) <= 1e-4   # J differentiable at this scale
        e = _rel(g[r, c], v6, scale)
        if smooth:
            worst = max(worst, e)
            checked += 1
        _say(f"cell ({r:3d},{c:3d}) adjoint {g[r, c]: .6e} fd(1e-6) {v6: .6e} rel {e:.2e}"
             f"{'' if smooth else ' (kink within 1e-6: not checked)'}  fd(1e-4) {v4: .6e}")
    jac = tr.materialize_jacobian(case.prims, case.base.fine.geo, nx, ny, None, 1.0)
    for i, (Jp, v6, v7, v4) in enumerate(zip(jac, fd6["dJ_dheight"], fd7["dJ_dheight"], fd4["dJ_dheight"])):
        gh = float(np.sum(g * Jp))
        e = _rel(gh, v6, 1e-9)
        smooth = _rel(v6, v7, 1e-9) <= 1e-3
        if smooth:
            worst = max(worst, e)
            checked += 1
        _say(f"height {i}: adjoint {gh: .8e} fd(1e-6) {v6: .8e} rel {e:.2e}  fd(1e-7) {v7: .8e}  fd(1e-4) {v4: .8e}")
    cost = (res["forward_ms"] + res["replay_ms"] + res["reverse_ms"]) / res["plain_forward_ms"]
    _say(f"worst rel {worst:.2e} over {checked} checked values; cost {cost:.2f}x the plain forward")
    assert checked >= 30
    assert worst <= REL_TOL


def test_adjoint_crest_level_against_f64_fd(solver):
    """The adjoint on a face-crest level (the 60x48 level 2 of the same terrain and design,
    tests/golden/adjoint_fd_config4_level2.json): the rise gradients of every crest face,
    split by the library into crest-line and level gradients and mapped to the fine raster
    by terrain.fine_gradient (block means, crest maxima), against f64 central
    differences of the fine dz at the fine cells of each asset box and its ring, and of
    every primitive height."""
    with open(os.path.join(HERE, "golden", "adjoint_fd_config4_level2.json")) as f:
        fx = json.load(f)
    case = sc.config4(level=2)
    assert case.base.crest
    p = case.params
    p.locked_dt = np.array(fx["locked_dt"])
    _say(f"case: adjoint config4 on the 60x48 crest level, {len(p.locked_dt)} locked steps")
    res = solver.adjoint(p, case.base, [case.candidate], case.monitors, case.assets, case.tau_m)[0]
    _say({k: v for k, v in res.items() if not isinstance(v, np.ndarray)})
    assert res["steps"] == fx["steps"]
    assert res["replay_mismatches"] == 0, "replayed segments must equal the checkpoints bitwise"
    fine = case.base.fine
    nxf, nyf = fine.nx, fine.ny
    area = fine.geo.dx * fine.geo.dy
    objv = ob.Objective([ob.Asset(np.asarray(a), th, w) for a, th, w in case.assets], case.tau_m, case.earth_weight)
    dz0 = case.edit.dense(nxf, nyf)
    cut, fill = ob.volumes(dz0, area)
    j = res["j_risk"] + case.earth_weight * objv.j_earth(cut, fill)
    assert abs(j - fx["j0"]) <= 1e-4 * abs(fx["j0"]), (j, fx["j0"])
    rx, ry = res["dJ_drise_face_x"], res["dJ_drise_face_y"]
    assert np.abs(rx).max() > 0 and np.abs(ry).max() > 0
    g = tr.fine_gradient(case.base, case.edit, res["dJ_dz_cell"], res["dJ_dz_face_x"], res["dJ_dz_face_y"])
    g = g + objv.earth_gradient(dz0, area)
    fd6, fd7, fd4 = (fx["fd"][repr(h)] for h in (1e-6, 1e-7, 1e-4))
    scale = 1e-3 * np.abs(np.array(fd6["dJ_dz_cell"])).max()
    worst, checked, nonzero = 0.0, 0, 0
    for (r, c), v6, v7, v4 in zip(fx["cells"], fd6["dJ_dz_cell"], fd7["dJ_dz_cell"], fd4["dJ_dz_cell"]):
        smooth = _rel(v6, v7, scale) <= 1e-4
        e = _rel(g[r, c], v6, scale)
        if smooth:
            worst = max(worst, e)
            checked += 1
            nonzero += v6 != 0.0
        if v6 != 0.0 or g[r, c] != 0.0:
            _say(f"fine cell ({r:3d},{c:3d}) adjoint {g[r, c]: .6e} fd(1e-6) {v6: .6e} rel {e:.2e}"
                 f"{'' if smooth else ' (kink within 1e-6: not checked)'}  fd(1e-4) {v4: .6e}")
    jac = tr.materialize_jacobian(case.prims, fine.geo, nxf, nyf, None, 1.0)
    for i, (Jp, v6, v7, v4) in enumerate(zip(jac, fd6["dJ_dheight"], fd7["dJ_dheight"], fd4["dJ_dheight"])):
        gh = float(np.sum(g * Jp))
        e = _rel(gh, v6, 1e-9)
        smooth = _rel(v6, v7, 1e-9) <= 1e-3
        if smooth:
            worst = max(worst, e)
            checked += 1
        _say(f"height {i}: adjoint {gh: .8e} fd(1e-6) {v6: .8e} rel {e:.2e}  fd(1e-7) {v7: .8e}  fd(1e-4) {v4: .8e}")
    cost = (res["forward_ms"] + res["replay_ms"] + res["reverse_ms"]) / res["plain_forward_ms"]
    _say(f"worst rel {worst:.2e} over {checked} checked values ({nonzero} nonzero cells); cost {cost:.2f}x the plain forward")
    assert checked >= 100 and nonzero >= 20
    assert worst <= REL_TOL


@pytest.mark.parametrize("kind", ["stage", "inflow", "initial"])
def test_adjoint_special_boundary_against_f64_fd(solver, kind):
    """The adjoint on a grid with a Stage or an Inflow segment, or with an initial stage
    (scenario.adjoint_boundary_case, fixtures tests/golden/adjoint_fd_<kind>.json): the
    ghosts' adjoints fold into their boundary cells as the forward builds the ghosts
    (boundary.rs:43-47, 66-70) -- a Stage ghost's depth f32(max(stage - z, 0)) moves with the
    boundary cell's level, an Inflow ghost's depth is the cell's own above the critical
    depth -- and the initial depth f32(max(h0, stage - z_ref - z, 0)) (solver.rs:292-304)
    moves with each cell's level; against f64 central differences at the segment's boundary
    cells or the pond's shore cells, cells around the assets and the three primitive heights
    (one primitive is a mound on the boundary cells or on the first asset).  With the
    initial stage, cell components below 5% of the largest are held to 3e-2 instead of
    1e-3: the f32 trajectory the adjoint differentiates and the f64 one the differences
    follow part slightly at the pond's shore films (the heights and the dominant components
    agree to 2e-5)."""
    with open(os.path.join(HERE, "golden", f"adjoint_fd_{kind}.json")) as f:
        fx = json.load(f)
    case = sc.adjoint_boundary_case(kind)
    p = case.params
    p.locked_dt = np.array(fx["locked_dt"])
    _say(f"case: adjoint with a {kind} segment, 60x48, {len(p.locked_dt)} locked steps")
    res = solver.adjoint(p, case.base, [case.candidate], case.monitors, case.assets, case.tau_m)[0]
    _say({k: v for k, v in res.items() if not isinstance(v, np.ndarray)})
    assert res["steps"] == fx["steps"]
    assert res["replay_mismatches"] == 0, "replayed segments must equal the checkpoints bitwise"
    nx, ny = case.base.nx, case.base.ny
    area = case.base.dx * case.base.dy
    objv = ob.Objective([ob.Asset(np.asarray(a), th, w) for a, th, w in case.assets], case.tau_m, case.earth_weight)
    dz0 = case.edit.dense(nx, ny)
    cut, fill = ob.volumes(dz0, area)
    j = res["j_risk"] + case.earth_weight * objv.j_earth(cut, fill)
    assert abs(j - fx["j0"]) <= 1e-4 * abs(fx["j0"]), (j, fx["j0"])
    g = res["dJ_dz_cell"] + objv.earth_gradient(dz0, area)
    fd6, fd7, fd4 = (fx["fd"][repr(h)] for h in (1e-6, 1e-7, 1e-4))
    scale = 1e-3 * np.abs(np.array(fd6["dJ_dz_cell"])).max()
    worst, worst_small, checked, boundary = 0.0, 0.0, 0, 0
    big = 50.0 * scale   # 5% of the largest difference
    for (r, c), v6, v7, v4 in zip(fx["cells"], fd6["dJ_dz_cell"], fd7["dJ_dz_cell"], fd4["dJ_dz_cell"]):
        smooth = _rel(v6, v7, scale) <= 1e-4
        e = _rel(g[r, c], v6, scale)
        if smooth:
            if kind == "initial" and abs(v6) < big:
                worst_small = max(worst_small, e)
            else:
                worst = max(worst, e)
            checked += 1
            boundary += c == 0 and v6 != 0.0
        _say(f"cell ({r:3d},{c:3d}) adjoint {g[r, c]: .6e} fd(1e-6) {v6: .6e} rel {e:.2e}"
             f"{'' if smooth else ' (kink within 1e-6: not checked)'}  fd(1e-4) {v4: .6e}")
    jac = tr.materialize_jacobian(case.prims, case.base.fine.geo, nx, ny, None, 1.0)
    for i, (Jp, v6, v7, v4) in enumerate(zip(jac, fd6["dJ_dheight"], fd7["dJ_dheight"], fd4["dJ_dheight"])):
        gh = float(np.sum(g * Jp))
        e = _rel(gh, v6, 1e-9)
        smooth = _rel(v6, v7, 1e-9) <= 1e-3
        if smooth:
            worst = max(worst, e)
            checked += 1
        _say(f"height {i}: adjoint {gh: .8e} fd(1e-6) {v6: .8e} rel {e:.2e}  fd(1e-7) {v7: .8e}"
             f"{'' if smooth else ' (kink within 1e-6: not checked)'}")
    nonzero = sum(1 for v6, v7 in zip(fd6["dJ_dz_cell"], fd7["dJ_dz_cell"]) if v6 != 0.0 and _rel(v6, v7, scale) <= 1e-4)
    _say(f"{kind}: worst rel {worst:.2e} over {checked} checked values ({nonzero} non-zero cells, "
         f"{boundary} on the boundary)")
    if kind == "initial":
        _say(f"initial: components below 5% of the largest: worst rel {worst_small:.2e}")
        assert checked >= 20 and nonzero >= 10 and worst_small <= 3e-2
    else:
        assert checked >= 28 and boundary >= 3
    assert worst <= REL_TOL


def test_autograd_function_backward(solver):
    """The torch.autograd.Function wrapper: J.backward() gives the adjoint gradient, and
    it matches the f64 differences of the committed fixture.  The tensors live on the
    host: the HIP library owns the device work (torch's bundled HIP runtime and the
    system one are not both initialised in one process)."""
    import torch
    from inverse_terrain_engine_amd.autograd import peak_risk_objective
    fx = _fixture()
    case = sc.config4()
    p = case.params
    p.locked_dt = np.array(fx["locked_dt"])
    nx, ny = case.base.nx, case.base.ny
    objv = ob.Objective([ob.Asset(np.asarray(a), th, w) for a, th, w in case.assets], case.tau_m, case.earth_weight)
    dz = torch.tensor(case.edit.dense(nx, ny), dtype=torch.float64, requires_grad=True)
    _say("case: autograd wrapper on config4")
    J = peak_risk_objective(dz, solver, p, case.base, case.monitors, objv)
    J.backward()
    assert abs(J.item() - fx["j0"]) <= 1e-4 * abs(fx["j0"])
    g = dz.grad.cpu().numpy()
    fd6 = fx["fd"][repr(1e-6)]["dJ_dz_cell"]
    scale = 1e-3 * max(abs(v) for v in fd6)
    worst = max(_rel(g[r, c], v, scale) for (r, c), v in zip(fx["cells"], fd6))
    _say(f"autograd worst rel {worst:.2e}")
    assert worst <= REL_TOL


def test_adjoint_declines_local_inertial(solver):
    """The adjoint differentiates the HLL step: a local-inertial grid returns ITR_DECLINED."""
    from inverse_terrain_engine_amd.gpu import Declined
    base = sc.validation_terrain(20, 12, 2.0, lambda x, y: 1.0 + 0.01 * x, n=0.03)
    p = sc.SolverParams(20.0, 10.0)
    p.kernel = "inertial"
    p.locked_dt = [0.1] * 200
    with pytest.raises(Declined):
        solver.adjoint(p, base, [tr.PreparedTerrain(base)], np.array([5], np.uint32), [([0], 0.05, 1.0)], 0.01)


@pytest.mark.parametrize("level", [1, 2])
def test_adjoint_band_march_equals_per_cell(solver, level):
    """The two reverse sweeps (band march, per-cell kernels) give the same bits: gradient of
    every face and cell, J, peaks.  Three candidates on a 130x70 grid (three waves per row,
    two of them holding ghost columns) with interior walls and a transmissive side, and on
    its 65x35 face-crest level 2 (a tail wave of three columns marching row groups)."""
    segs = [{"kind": "transmissive", "side": "east"}]
    base = sc.validation_terrain(130, 70, 2.0, lambda x, y: 5.0 + 0.01 * (260.0 - x) + 0.2 * math.sin(y / 9.0),
                                 walls=lambda c, r: 60 <= c < 64 and 20 <= r < 40, n=0.035, segments=segs)
    fine_geo = base.fine.geo
    if level == 2:
        base = tr.TerrainBase.build(base.fine, 2, segs)
    nx, ny = base.nx, base.ny
    p = sc.SolverParams(300.0, 60.0, rain=sc.mm_h([[0.0, 60.0], [300.0, 20.0]]), cfl=0.5)
    dts = solver.dt_schedule(p, base, tr.PreparedTerrain(base), np.zeros(0, np.uint32), 4000)
    p.locked_dt = np.asarray(dts)
    p.cfl_hard = 1.0
    rng = sc.SplitMix64(11)
    terrs = [tr.PreparedTerrain(base)]
    for _ in range(2):
        prims = sc.random_primitives(rng, 3, fine_geo, 130, 70)
        terrs.append(tr.PreparedTerrain(base).apply_edit(tr.materialize(prims, fine_geo, 130, 70, None, 1.0)))
    L = level
    box1 = {r * nx + c for r in range(30 // L, 36 // L) for c in range(100 // L, 108 // L)}
    box2 = {r * nx + c for r in range(50 // L, 54 // L) for c in range(20 // L, 26 // L)}
    mon = np.array(sorted(box1 | box2), np.uint32)
    assets = [(list(range(0, len(box1))), 0.01, 1.0), (list(range(len(box1), len(mon))), 0.005, 2.0)]
    _say(f"case: band vs per-cell reverse, 3 x {nx}x{ny} (level {level}), {len(dts)} steps")
    out = {}
    for mode in (1, 2):
        solver.set_option("adjoint_mode", mode)
        out[mode] = solver.adjoint(p, base, terrs, mon, assets, 0.01)
    solver.set_option("adjoint_mode", 0)
    for c in range(3):
        a, b = out[1][c], out[2][c]
        assert a["replay_mismatches"] == 0 and b["replay_mismatches"] == 0
        assert a["j_risk"] == b["j_risk"]
        keys = ["dJ_dz_cell", "dJ_dz_face_x", "dJ_dz_face_y"]
        if level == 2:
            keys += ["dJ_drise_face_x", "dJ_drise_face_y"]
        for key in keys:
            assert np.array_equal(a[key].view(np.uint64), b[key].view(np.uint64)), (c, key)
        assert np.abs(a["dJ_dz_cell"]).max() > 0.0


@pytest.mark.parametrize("kind", ["stage", "inflow"])
def test_adjoint_band_march_equals_per_cell_special(solver, kind):
    """Band-march and per-cell replay and reverse give the same gradient bits on a grid with
    a Stage or Inflow segment (special ghosts in the replay, their folds in the reverse)."""
    import copy
    case = sc.adjoint_boundary_case(kind)
    geo = case.base.fine.geo
    nx, ny = case.base.nx, case.base.ny
    rng = sc.SplitMix64(21)
    terrs = [case.candidate]
    for _ in range(2):
        prims = sc.random_primitives(rng, 3, geo, nx, ny)
        terrs.append(tr.PreparedTerrain(case.base).apply_edit(tr.materialize(prims, geo, nx, ny, None, 1.0)))
    p = copy.copy(case.params)
    p.duration_s = 600.0
    dts = solver.dt_schedule(p, case.base, case.candidate, case.monitors, 4000)
    p.locked_dt = np.asarray(dts)
    _say(f"case: band vs per-cell reverse with a {kind} segment, 3 x {nx}x{ny}, {len(dts)} steps")
    out = {}
    for mode in (1, 2):
        solver.set_option("adjoint_mode", mode)
        out[mode] = solver.adjoint(p, case.base, terrs, case.monitors, case.assets, case.tau_m)
    solver.set_option("adjoint_mode", 0)
    for c in range(3):
        a, b = out[1][c], out[2][c]
        assert a["replay_mismatches"] == 0 and b["replay_mismatches"] == 0
        assert a["j_risk"] == b["j_risk"]
        for key in ("dJ_dz_cell", "dJ_dz_face_x", "dJ_dz_face_y"):
            assert np.array_equal(a[key].view(np.uint64), b[key].view(np.uint64)), (c, key)
    assert np.abs(out[1][0]["dJ_dz_cell"][:, 0]).max() > 0.0
