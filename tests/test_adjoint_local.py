"""Host checks of the hand-written local adjoints (inverse-terrain-engine_amd/csrc/itr_adjoint.h):
each block's vector-Jacobian product against double-precision central differences over
operating points that cover its branches.  Compiled with hipcc for the host only; no GPU."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import pytest

from conftest import ROOT

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_local_adjoints_match_central_differences():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "local_check")
        subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O1", "-std=c++17", "-I",
                        os.path.join(ROOT, "inverse-terrain-engine_amd", "csrc"),
                        os.path.join(ROOT, "tests", "adjoint", "local_check.cpp"), "-o", exe],
                       check=True, capture_output=True)
        r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "PASS: 0 mismatches" in r.stdout


def test_adjoint_fixture_is_reproducible():
    """Re-derive two of the committed f64 differences (one cell at 1e-6 m, and the
    objective value) with the oracle."""
    import json
    import numpy as np
    from conftest import oracle_params, oracle_terrain, po
    from inverse_terrain_engine_amd import objective as ob
    from inverse_terrain_engine_amd import scenario as sc
    with open(os.path.join(ROOT, "tests", "golden", "adjoint_fd_config4.json")) as f:
        fx = json.load(f)
    case = sc.config4()
    nx, ny = case.base.nx, case.base.ny
    area = case.base.dx * case.base.dy
    objv = ob.Objective([ob.Asset(np.asarray(a), th, w) for a, th, w in case.assets], case.tau_m, case.earth_weight)
    terr = oracle_terrain(case.base)
    f32 = po.run(terr, oracle_params(case.params), precision=32, edit=case.edit.as_dict(), monitors=case.monitors,
                 want_dts=True)
    assert np.array_equal(f32["dts"], np.array(fx["locked_dt"]))
    q = oracle_params(case.params)
    q.locked_dt = f32["dts"]
    dz0 = case.edit.dense(nx, ny)

    def J(dz):
        o = po.run(terr, q, precision=64, edit={"c0": 0, "r0": 0, "dz": dz}, monitors=case.monitors)
        return objv.j(o["monitor_max"], *ob.volumes(dz, area))

    assert J(dz0) == fx["j0"]
    k = int(np.argmax(np.abs(fx["fd"][repr(1e-6)]["dJ_dz_cell"])))
    r, c = fx["cells"][k]
    dp, dm = dz0.copy(), dz0.copy()
    dp[r, c] += 1e-6
    dm[r, c] -= 1e-6
    assert (J(dp) - J(dm)) / 2e-6 == fx["fd"][repr(1e-6)]["dJ_dz_cell"][k]


def test_fine_gradient_of_a_crest_level():
    """terrain.fine_gradient (the chain rule from a face-crest level's gradients to the fine
    raster) against central differences of the level's own construction: a random linear
    functional of the coarse levels (_coarse, terrain.rs:357-373) and of the faces' crest
    lines (_crest_line_x / _y, terrain.rs:383-397, 410-424), differenced at fine cells of
    the edited window (block means, crest maxima and cells neither touches)."""
    import numpy as np
    from inverse_terrain_engine_amd import scenario as sc
    from inverse_terrain_engine_amd import terrain as tr
    lad = sc.config3(n_inst=7, duration=900.0)
    base = lad.groups[0][0]
    assert base.crest and base.level == 2
    fine, lv, nx, ny = base.fine, base.level, base.nx, base.ny
    edit = lad.edits[0]
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, (ny, nx))
    bx = np.zeros((ny, nx + 1))
    by = np.zeros((ny + 1, nx))
    bx[:, 1:nx] = rng.uniform(-1, 1, (ny, nx - 1))
    by[1:ny, :] = rng.uniform(-1, 1, (ny - 1, nx))
    a[base.wall] = 0.0

    def F(zf):
        z, _, _ = tr._coarse(fine, lv, zf)
        cx = tr._crest_line_x(fine, lv, zf, 0, ny, 1, nx)
        cy = tr._crest_line_y(fine, lv, zf, 1, ny, 0, nx)
        return float(np.sum(a * z) + np.sum(bx[:, 1:nx] * cx) + np.sum(by[1:ny, :] * cy))

    g = tr.fine_gradient(base, edit, a, bx, by)
    z0 = fine.z + edit.dense(fine.nx, fine.ny)
    cells = [(edit.r0 + r, edit.c0 + c) for r in range(0, edit.h, 3) for c in range(0, edit.w, 3)]
    cells += [(0, 0), (fine.ny - 1, fine.nx - 1), (fine.ny // 2, fine.nx // 2)]
    H = 1e-6
    worst, crest_hits = 0.0, 0
    for r, c in cells:
        zp, zm = z0.copy(), z0.copy()
        zp[r, c] += H
        zm[r, c] -= H
        fd = (F(zp) - F(zm)) / (2 * H)
        worst = max(worst, abs(fd - g[r, c]))
        crest_hits += abs(g[r, c] - (0.0 if fine.wall[r, c] else a[r // lv, c // lv] / 4)) > 1e-9
    print(f"{len(cells)} fine cells, worst |fd - chain| {worst:.2e}, {crest_hits} hold a crest maximum")
    assert worst < 1e-6 and crest_hits > 0


def test_crest_fixture_is_reproducible():
    """Re-derive the crest-level fixture's objective value and its largest fine-cell
    difference (1e-6 m) with the oracle (tests/golden/adjoint_fd_config4_level2.json)."""
    import json
    import numpy as np
    from conftest import oracle_params, oracle_terrain, po
    from inverse_terrain_engine_amd import objective as ob
    from inverse_terrain_engine_amd import scenario as sc
    with open(os.path.join(ROOT, "tests", "golden", "adjoint_fd_config4_level2.json")) as f:
        fx = json.load(f)
    case = sc.config4(level=2)
    fine = case.base.fine
    area = fine.geo.dx * fine.geo.dy
    objv = ob.Objective([ob.Asset(np.asarray(a), th, w) for a, th, w in case.assets], case.tau_m, case.earth_weight)
    terr = oracle_terrain(case.base)
    f32 = po.run(terr, oracle_params(case.params), precision=32, edit=case.edit.as_dict(), monitors=case.monitors,
                 want_dts=True)
    assert np.array_equal(f32["dts"], np.array(fx["locked_dt"]))
    q = oracle_params(case.params)
    q.locked_dt = f32["dts"]
    dz0 = case.edit.dense(fine.nx, fine.ny)

    def J(dz):
        o = po.run(terr, q, precision=64, edit={"c0": 0, "r0": 0, "dz": dz}, monitors=case.monitors)
        return objv.j(o["monitor_max"], *ob.volumes(dz, area))

    assert J(dz0) == fx["j0"]
    k = int(np.argmax(np.abs(fx["fd"][repr(1e-6)]["dJ_dz_cell"])))
    r, c = fx["cells"][k]
    dp, dm = dz0.copy(), dz0.copy()
    dp[r, c] += 1e-6
    dm[r, c] -= 1e-6
    assert (J(dp) - J(dm)) / 2e-6 == fx["fd"][repr(1e-6)]["dJ_dz_cell"][k]
