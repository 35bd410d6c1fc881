"""The device's f32 runs of every named workload against the f64 scalar oracle: the f64
ledger totals (final volume, boundary outflow, infiltration), the mass error of each run
and the peak depths at the monitored cells.

The device's f32 state is bitwise the reference's f32 fused kernel (the parity tests); the
f64 oracle restates the reference's solver at T = f64 (crates/itr-hydro/src/solver.rs
with real.rs for f64).  An f32 run and the f64 run of the same candidate part by
rounding, so these are tolerances, not bit checks: each bound is about 100x the largest
difference seen on these workloads and far below what a wrong term in the physics gives
(an error in a flux, a source or a boundary moves volumes and peaks by percent).  With
ITR_F64_REPORT set to a file, each test appends its figures there as JSON lines."""
from __future__ import annotations

import copy
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import oracle_params, oracle_terrain, po
from inverse_terrain_engine_amd import scenario as sc
from inverse_terrain_engine_amd import terrain as tr

pytestmark = pytest.mark.gpu
THREADS = max(1, min(16, os.cpu_count() or 1))

# bounds (see the module docstring); measured worst cases on the five workloads are
# 1.2e-5 (volume), 4.9e-6 (outflow), 9e-9 (infiltration), 5.8e-6 (f32 mass error) and
# 7.5e-6 m (peak depth)
VOL_REL = 1e-3
FLUX_REL = 1e-3
F32_MASS = 1e-4
F64_MASS = 1e-12
PEAK_ABS = 1e-3


def _say(*a):
    print(*a, flush=True)


def f64_many(base, p, edits, monitors):
    """The f64 oracle run of each edit (None = unedited), one per host thread."""
    ot, op = oracle_terrain(base), oracle_params(p)

    def one(e):
        return po.run(ot, op, 64, edit=None if e is None else e.as_dict(), monitors=monitors)
    with ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(one, edits))


def compare(tag, runs, f64s):
    """Worst-case figures over the candidates, asserted against the bounds."""
    w = {"vol_rel": 0.0, "outflow_rel": 0.0, "infil_rel": 0.0, "f32_mass_err": 0.0, "f64_mass_err": 0.0,
         "peak_abs_m": 0.0, "peak_rel": 0.0, "peak_max_m": 0.0, "steps_f32": [], "steps_f64": []}
    for g, o in zip(runs, f64s):
        a, b = g.ledger[-1], o["rows"][-1]
        rel = lambda x, y: abs(x - y) / max(abs(y), 1e-300)
        w["vol_rel"] = max(w["vol_rel"], rel(a[1], b[1]))
        w["outflow_rel"] = max(w["outflow_rel"], rel(a[4], b[4]))
        w["infil_rel"] = max(w["infil_rel"], rel(a[5], b[5]))
        w["f32_mass_err"] = max(w["f32_mass_err"], float(g.rel_residual))
        w["f64_mass_err"] = max(w["f64_mass_err"], float(o["rel_residual"]))
        d = np.abs(g.monitor_max - o["monitor_max"])
        w["peak_abs_m"] = max(w["peak_abs_m"], float(d.max()))
        w["peak_rel"] = max(w["peak_rel"], float((d / np.maximum(o["monitor_max"], 1e-3)).max()))
        w["peak_max_m"] = max(w["peak_max_m"], float(o["monitor_max"].max()))
        w["steps_f32"].append(int(g.steps))
        w["steps_f64"].append(int(o["steps"]))
    rep = {"workload": tag, "candidates": len(runs),
           "steps_f32": [min(w["steps_f32"]), max(w["steps_f32"])],
           "steps_f64": [min(w["steps_f64"]), max(w["steps_f64"])],
           **{k: v for k, v in w.items() if not k.startswith("steps")}}
    _say(json.dumps(rep))
    path = os.environ.get("ITR_F64_REPORT")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps(rep) + "\n")
    assert w["f64_mass_err"] < F64_MASS, rep
    assert w["f32_mass_err"] < F32_MASS, rep
    assert w["vol_rel"] < VOL_REL and w["outflow_rel"] < FLUX_REL and w["infil_rel"] < FLUX_REL, rep
    assert w["peak_abs_m"] < PEAK_ABS, rep


def test_config1_single_forward(solver):
    w = sc.config1()
    _say("case: config1, 90x70, 40 min storm, f32 device vs f64 oracle")
    runs = solver.run_batch(w.params, w.candidates, w.monitors)
    compare("config1: 90x70 at 5 m, 40 min storm", runs, f64_many(w.base, w.params, [None], w.monitors))


def test_config2_population(solver):
    w = sc.config2()
    _say(f"case: config2, {len(w.candidates)} candidates x 120x96, 60 min storm, f32 device vs f64 oracle")
    runs = solver.run_batch(w.params, w.candidates, w.monitors)
    compare("config2: 64 candidates x 120x96 at 5 m, 60 min storm (no guard)", runs,
            f64_many(w.base, w.params, w.edits, w.monitors))


def test_config3_ladder(solver):
    lad = sc.config3()
    _say("case: config3, 256 ladder instances, 40 min storm, f32 device vs f64 oracle")
    sess = solver.session(lad.params, lad.groups)
    try:
        sess.run()
        res = sess.results(want_state=False)
    finally:
        sess.close()
    k = 0
    for base, cands, mon, _ in lad.groups:
        n = len(cands)
        compare(f"config3: {n} x {base.nx}x{base.ny} at {base.dx:g} m{' (face-crest)' if base.crest else ''}",
                res[k:k + n], f64_many(base, lad.params, lad.edits[k:k + n], mon))
        k += n


def test_config4_adjoint_case_forward(solver):
    a = sc.config4()
    _say("case: config4, 120x96 forward, 40 min storm, f32 device vs f64 oracle")
    runs = solver.run_batch(a.params, [a.candidate], a.monitors)
    compare("config4: 120x96 at 5 m, 40 min storm, one design", runs,
            f64_many(a.base, a.params, [a.edit], a.monitors))


def test_config5_first_storm_peak(solver):
    w = sc.config5(n_cand=128, n=512)
    p = copy.copy(w.params)
    p.duration_s = 1800.0   # through the storm's first peak
    k = 4
    _say(f"case: config5, {k} candidates x 512x512, t = 0 - 1800 s, f32 device vs f64 oracle")
    runs = solver.run_batch(p, w.candidates[:k], w.monitors, want_state=False)
    compare(f"config5: {k} of the 128 candidates x 512x512 at 2 m, t = 0 - 1800 s", runs,
            f64_many(w.base, p, w.edits[:k], w.monitors))
