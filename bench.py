"""Throughput of the batched f32 shallow-water step on one MI355X.

A "step" is one time step of every candidate of the batch: one launch of the HIP step
kernel over 128 candidates of a 512x512 grid at 2 m (the large batched-forward
workload; seeded DEM with 40 hills and 8 gullies, 4 primitives per candidate, bimodal
2 h storm), with the guard early abort against the unedited baseline on.  The batch
first runs untimed from t = 0 to about the storm peak (reported as the whole-run figure,
early aborts and host round trips included), then W untimed warmup steps, then exactly
K timed steps with the inputs already resident in HBM.  Prints one JSON line (progress
on stderr).

The cpu_baseline leg times the CPU oracle (the reference's f32 fused kernel restated in
C++, bitwise equal to it) candidate-parallel on the host cores, one candidate per thread
like the reference's evaluator, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, B/s
BYTES_PER_CELL_STEP = 32   # read h, qx, qy, dzx, dzy + write h, qx, qy (f32)
VALU_PEAK = 256 * 2.4e9    # wave64 VALU instructions/s: one per CU per cycle (packed FP32 not used)
COPY_PROBE_BYTES = 1 << 30 # plain device-to-device copy used to calibrate the HBM counters
METRIC = "batched f32 cell-steps/s + % HBM peak; adjoint grad cost vs fwd at tol"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, threads, n_cand, n_steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po  # the checker, timed here as the CPU baseline only
    b = workload.base
    f = b.fine
    segs = [{"kind": "transmissive", "side": "east"}]
    ot = po.Terrain(f.dem, f.wall, f.n, dx=f.geo.dx, dy=f.geo.dy, origin=(f.geo.origin_x, f.geo.origin_y),
                    segments=segs)
    p = workload.params
    op = po.SolverParams(p.duration_s, p.sync_interval_s, cfl=p.cfl, rain=list(p.rain),
                         infil_capacity=p.infil_capacity)
    q_edits = [e.as_dict() for e in workload.edits[:n_cand]]
    # bounded sample: each candidate runs n_steps steps (max_steps)
    params, _keep = po._params(op, 32, 0, n_steps)
    wall, steps, cu = _run_cands(po, ot, params, q_edits, threads, workload.monitors)
    return float(cu.sum()) / wall, wall, int(steps.sum())


def cpu_single_thread_bench_case(n=512, steps=100):
    """The oracle on the reference's own bench case (main.rs:686-694) at n x n, f32, one
    thread, a bounded number of steps: the figure to set beside the published
    single-thread 1.531e8 cell-steps/s (docs/data/bench.json, M4 Pro)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po  # the checker, timed here as the CPU baseline only
    from inverse_terrain_engine_amd import scenario as sc
    base = sc.bench_case_terrain(n)
    p = sc.bench_case_params()
    ot = po.Terrain(base.fine.dem, manning=0.035, dx=2.0)
    t0 = time.perf_counter()
    o = po.run(ot, po.SolverParams(p.duration_s, p.sync_interval_s, initial_depth=0.05, rain=p.rain), precision=32,
               max_steps=steps)
    return o["cell_updates"] / (time.perf_counter() - t0)


def _run_cands(po, ot, params, edits, threads, monitors):
    import ctypes as C
    import numpy as np
    n = len(edits)
    arr = (po.Edit * max(n, 1))()
    keep = []
    for i, ed in enumerate(edits):
        e, dz = po.make_edit(ed)
        keep.append(dz)
        arr[i] = e
    mon = np.ascontiguousarray(monitors, dtype=np.uint32)
    steps = np.zeros(n, np.int64)
    cu = np.zeros(n, np.int64)
    wall = po.lib().orc_run_candidates(ot.h, arr, n, C.byref(params), mon.ctypes.data_as(po._up), len(mon), threads,
                                       steps.ctypes.data_as(po._lp), cu.ctypes.data_as(po._lp))
    return wall, steps, cu


def _pmc_pass(counter, steps, band_rows):
    """One rocprofv3 counter pass over a short run of this benchmark (plus the copy
    probe); returns {kernel name: [per-dispatch values]} or None if it did not run.  The
    pass runs the batch as one launch list (list_split 1), so that every dispatch covers
    every candidate; the per-cell-step figures do not depend on the split."""
    import csv
    import shutil
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    out = tempfile.mkdtemp(prefix="itr_pmc_", dir="/tmp")
    cmd = ["timeout", "-s", "KILL", "150", prof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "pmc",
           "--", sys.executable, os.path.abspath(__file__), "--steps", str(steps), "--warmup", "2", "--no-cpu",
           "--no-traffic", "--copy-probe", "--no-adjoint", "--no-ladder", "--band-rows", str(band_rows),
           "--advance", "0", "--guard-tol", "0", "--no-graph", "--list-split", "1"]
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
    path = None
    for root, _, files in os.walk(out):
        for f in files:
            if f.endswith("counter_collection.csv"):
                path = os.path.join(root, f)
    if path is None:
        print(f"pmc pass {counter} produced no counters (rc {r.returncode}): {r.stderr[-300:]}", file=sys.stderr)
        return None
    vals = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter:
                vals.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    shutil.rmtree(out, ignore_errors=True)
    return vals


def measure_traffic(band_rows, cells, steps=6):
    """HBM bytes and VALU instructions per cell-step of the step kernel, from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch, each scaled by what the same counter reads for a plain
    1 GiB device copy) and SQ_INSTS_VALU (wave64 instructions), in counter passes over a
    short run of the same batch from t = 0 (every candidate running: `cells` cells per
    launch).  Counting the timed window itself would need the untimed advance in every
    pass; the per-cell-step figures do not depend on the window."""
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE", "SQ_INSTS_VALU"):
        v = _pmc_pass(counter, steps, band_rows)
        if v is None:
            return None
        step = [x for k, xs in v.items() if "itr_step_" in k for x in xs]
        copy = [x for k, xs in v.items() if "copyBuffer" in k for x in xs]
        if not step:
            return None
        res[counter] = (sum(step) / len(step), max(copy) if copy else None)
    f_kib, f_copy = res["FETCH_SIZE"]
    w_kib, w_copy = res["WRITE_SIZE"]
    if not f_copy or not w_copy:
        return None
    kf = COPY_PROBE_BYTES / (f_copy * 1024.0)
    kw = COPY_PROBE_BYTES / (w_copy * 1024.0)
    return {"fetch_per_cell_step": f_kib * 1024.0 * kf / cells, "write_per_cell_step": w_kib * 1024.0 * kw / cells,
            "copy_scale_fetch": kf, "copy_scale_write": kw,
            "valu_lane_insts_per_cell_step": res["SQ_INSTS_VALU"][0] * 64 / cells,
            "note": f"counter passes over {steps} steps from t = 0, all candidates running"}


def copy_probe():
    """A plain device-to-device copy of COPY_PROBE_BYTES through the HIP runtime."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    n = C.c_size_t(COPY_PROBE_BYTES)
    ok = hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0
    ok = ok and hip.hipMemset(a, 1, n) == 0 and hip.hipMemcpy(b, a, n, 3) == 0   # 3: device to device
    ok = ok and hip.hipDeviceSynchronize() == 0
    for p in (a, b):
        if p.value:
            hip.hipFree(p)
    if not ok:
        raise RuntimeError("copy probe failed")


def adjoint_config4(solver, sc, ob):
    """The 120x96 adjoint workload: gradient error against the committed f64 central
    differences (1e-6 m steps where they agree with 1e-7 m) and the cost relative to the
    plain forward on the same locked schedule."""
    import numpy as np
    with open(os.path.join(ROOT, "tests", "golden", "adjoint_fd_config4.json")) as f:
        fx = json.load(f)
    case = sc.config4()
    p = case.params
    p.locked_dt = np.array(fx["locked_dt"])
    solver.autotune(p, [(case.base, [case.candidate], case.monitors, None)], trial_steps=16)
    res = solver.adjoint(p, case.base, [case.candidate], case.monitors, case.assets, case.tau_m)[0]
    nx, ny = case.base.nx, case.base.ny
    area = case.base.dx * case.base.dy
    objv = ob.Objective([ob.Asset(np.asarray(a), th, w) for a, th, w in case.assets], case.tau_m, case.earth_weight)
    dz0 = case.edit.dense(nx, ny)
    g = res["dJ_dz_cell"] + objv.earth_gradient(dz0, area)
    fd6, fd7 = fx["fd"][repr(1e-6)], fx["fd"][repr(1e-7)]
    scale = 1e-3 * max(abs(v) for v in fd6["dJ_dz_cell"])
    rel = lambda a, b, sc_: abs(a - b) / max(abs(b), sc_)
    errs = [rel(g[r, c], v6, scale) for (r, c), v6, v7 in zip(fx["cells"], fd6["dJ_dz_cell"], fd7["dJ_dz_cell"])
            if rel(v6, v7, scale) <= 1e-4]
    from inverse_terrain_engine_amd import terrain as tr
    jac = tr.materialize_jacobian(case.prims, case.base.fine.geo, nx, ny, None, 1.0)
    for Jp, v6, v7 in zip(jac, fd6["dJ_dheight"], fd7["dJ_dheight"]):
        if rel(v6, v7, 1e-9) <= 1e-3:
            errs.append(rel(float(np.sum(g * Jp)), v6, 1e-9))
    total = res["forward_ms"] + res["replay_ms"] + res["reverse_ms"]
    return {"workload": "120x96 at 5 m, 40 min storm, one design of 3 primitives, locked dt (CFL 0.9)",
            "steps": res["steps"], "segment_steps": res["segment_steps"], "checkpoints": res["checkpoints"],
            "cost_vs_forward": total / res["plain_forward_ms"], "worst_rel_err_vs_f64_fd": max(errs),
            "values_checked": len(errs), "tol": 1e-3, "replay_mismatches": res["replay_mismatches"],
            "plain_forward_ms": res["plain_forward_ms"], "forward_ms": res["forward_ms"],
            "replay_ms": res["replay_ms"], "reverse_ms": res["reverse_ms"]}


def adjoint_best16(solver, sc, ob, w, res, guard, horizon, t_horizon=1800.0):
    """Gradient refinement of the batch's incumbents: the adjoint of the 16 best feasible
    candidates (not guard-aborted; ranked by the running lower bound of J from the batch's
    monitor maxima, objective.rs:146-149) on the best one's dt schedule from t = 0 to
    t_horizon (through the first storm peak, so that the assets pond), and a central-
    difference check of its primitive-height gradient with device forward runs."""
    import copy
    import numpy as np
    from inverse_terrain_engine_amd import terrain as tr
    base = w.base
    nx, ny = base.nx, base.ny
    objv = ob.Objective([ob.Asset(np.asarray(a), 0.05, 1.0) for a in w.asset_pos], 0.01, 1e-3)
    js = [objv.j(r.monitor_max, e.cut_m3, e.fill_m3) for r, e in zip(res, w.edits)]
    feasible = [i for i, r in enumerate(res) if r.aborted == 0]
    best = sorted(feasible, key=lambda i: js[i])[:16]
    # The gradient check needs a candidate whose primitives reach the assets: the feasible
    # one with a primitive nearest an asset box takes the 16th place if it is not among them.
    geo = base.fine.geo
    centres = [geo.cell_center(c0 + bw / 2.0 - 0.5, r0 + bh / 2.0 - 0.5) for c0, r0, bw, bh in w.asset_boxes]
    dist = lambda i: min(math.hypot(pr.cx - x, pr.cy - y) for pr in w.designs[i] for x, y in centres)
    near = min(feasible, key=dist)
    if near not in best:
        best = best[:15] + [near]
    p = copy.copy(w.params)   # (the workload's own parameters serve the later legs)
    p.duration_s = t_horizon
    p.cfl_hard = 1.0
    sched = solver.dt_schedule(p, base, w.candidates[best[0]], w.monitors, max_steps=horizon)
    p.locked_dt = sched
    p.max_steps = len(sched)
    assets = [(a.positions, a.threshold_m, a.weight) for a in objv.assets]
    cands = [w.candidates[i] for i in best]
    solver.autotune(p, [(base, cands, w.monitors, None)], trial_steps=16)
    out = solver.adjoint(p, base, cands, w.monitors, assets, objv.tau_m)
    r0 = out[0]
    total = r0["forward_ms"] + r0["replay_ms"] + r0["reverse_ms"]
    # d J_risk / d height of the best candidate's primitives: the adjoint through the 4x4
    # quadrature profile, against central differences of device f32 forward runs on the
    # same locked schedule (two steps; f32 limits how small a step can be).
    k_near = best.index(near)
    prims = w.designs[near]
    jac = tr.materialize_jacobian(prims, geo, nx, ny, None, 1.0)
    grad = [float(np.sum(out[k_near]["dJ_dz_cell"] * J)) for J in jac]
    fd = {}
    for step in (1e-2, 3e-3):
        terrs = []
        for k in range(len(prims)):
            for sgn in (1.0, -1.0):
                q = [copy.copy(pr) for pr in prims]
                q[k].height += sgn * step
                terrs.append(tr.PreparedTerrain(base).apply_edit(tr.materialize(q, geo, nx, ny, None, 1.0)))
        runs = solver.run_batch(p, terrs, w.monitors, want_state=False)
        jr = [objv.j_risk(r.monitor_max) for r in runs]
        fd[step] = [(jr[2 * k] - jr[2 * k + 1]) / (2.0 * step) for k in range(len(prims))]
    scale = max(abs(g) for g in grad) or 1.0
    rel = lambda a, b: abs(a - b) / max(abs(b), 1e-3 * scale)
    return {"workload": f"16 best feasible candidates of the batch, {nx}x{ny}, t = 0 - {t_horizon:g} s "
                        f"({len(sched)} steps on the best one's dt schedule)",
            "steps": r0["steps"], "cost_vs_forward": total / r0["plain_forward_ms"],
            "replay_mismatches": r0["replay_mismatches"], "plain_forward_ms": r0["plain_forward_ms"],
            "forward_ms": r0["forward_ms"], "replay_ms": r0["replay_ms"], "reverse_ms": r0["reverse_ms"],
            "cell_steps_per_s_reverse": 16 * nx * ny * r0["steps"] / (r0["reverse_ms"] * 1e-3),
            "best_asset_peaks_m": [float(x) for x in r0["peak"]], "best_j_risk": r0["j_risk"],
            "fd_check": {"what": "d J_risk / d height of the primitives of the candidate with a primitive nearest "
                                 "an asset: adjoint vs central differences of device f32 forward runs on the same "
                                 "locked schedule",
                         "candidate": near, "asset_peaks_m": [float(x) for x in out[k_near]["peak"]],
                         "j_risk": out[k_near]["j_risk"], "nearest_primitive_m": dist(near),
                         "adjoint": grad, "fd_1e-2": fd[1e-2], "fd_3e-3": fd[3e-3],
                         "rel_err_vs_fd_3e-3": [rel(g, f) for g, f in zip(grad, fd[3e-3])],
                         "fd_steps_rel_diff": [rel(a, b) for a, b in zip(fd[1e-2], fd[3e-3])]}}


def prep_leg(solver, sc, tr, w):
    """Candidate preparation of the same batch two ways, each timed up to a resident batch:
    host materialize + apply_edit (design.rs:316-385, terrain.rs:460-505) then upload, or
    the primitives uploaded and materialized on the device.  Both give the same faces
    (bitwise; tests/test_gpu_parity.py)."""
    import numpy as np
    base, geo = w.base, w.base.fine.geo
    groups_mon = w.monitors
    t0 = time.perf_counter()
    edits = [tr.materialize(d, geo, base.nx, base.ny, None, 1.0) for d in w.designs]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    sess = solver.session(w.params, [(base, terrs, groups_mon, None)])
    t_host = time.perf_counter() - t0
    sess.close()
    t0 = time.perf_counter()
    sess = solver.session_designs(w.params, base, w.designs, groups_mon, None, 1.0)
    t_dev = time.perf_counter() - t0
    sess.close()
    # The same designs on the face-crest level 2 (coarse cells and crest rises per face).
    base2 = tr.TerrainBase.build(base.fine, 2, base.segments)
    mon2 = np.unique((w.monitors // base.nx // 2) * base2.nx + (w.monitors % base.nx) // 2).astype(np.uint32)
    t0 = time.perf_counter()
    edits = [tr.materialize(d, geo, base.nx, base.ny, None, 1.0) for d in w.designs]
    terrs = [tr.PreparedTerrain(base2).apply_edit(e) for e in edits]
    sess = solver.session(w.params, [(base2, terrs, mon2, None)])
    t_host2 = time.perf_counter() - t0
    sess.close()
    t0 = time.perf_counter()
    sess = solver.session_designs(w.params, base2, w.designs, mon2, None, 1.0)
    t_dev2 = time.perf_counter() - t0
    sess.close()
    return {"workload": f"{len(w.designs)} candidates x 4 primitives, {base.nx}x{base.ny}",
            "host_prep_s": t_host, "device_prep_s": t_dev, "speedup": t_host / max(t_dev, 1e-9),
            "crest_level2": {"workload": f"the same designs on the {base2.nx}x{base2.ny} face-crest level",
                             "host_prep_s": t_host2, "device_prep_s": t_dev2,
                             "speedup": t_host2 / max(t_dev2, 1e-9)}}


def ladder(solver_cls, sc, steps, warm=600):
    """The fidelity-ladder batch (256 instances: 45x35 crest, 90x70 and 180x140 signed,
    4:2:1) in one call: three launch lists (one per class of grids) on three streams, one
    launch per step each, with band heights from the autotuner.  `warm` untimed steps
    precede the timed ones with no host reads between them (the GPU's clocks drop during
    host work and take tens of milliseconds to come back up)."""
    lad = sc.config3()
    s = solver_cls(0)
    tuned = s.autotune(lad.params, lad.groups, trial_steps=16)
    sess = s.session(lad.params, lad.groups)
    sess.advance(warm)
    st = sess.advance(steps)
    res = sess.results(want_state=False)
    sess.close()
    s.close()
    t = [r.t_end for r in res]
    return {"workload": "256 instances: 146 x 45x35 at 10 m (face-crest), 73 x 90x70 at 5 m, 37 x 180x140 at 2.5 m",
            "cell_steps_per_s": st["cell_steps"] / st["wall_s"], "us_per_step": st["device_ms"] * 1e3 / steps,
            "band_rows": [r for r, _ in tuned], "steps": steps, "warm_steps": warm,
            "t_s_at_end": [float(min(t)), float(max(t))],
            "running_at_end": sum(1 for r in res if r.aborted == 0 and r.t_end < lad.params.duration_s)}


def screening_leg(solver_cls, sc, tr, w, steps, warmup=150, cpu_threads=0, cpu_cands=16, cpu_steps=300, t_wet=1800.0):
    """The same candidates on the face-crest level-2 grid (2x coarser), run with the
    local-inertial screening physics (inertial.rs; problem.rs:178-181) and with HLL.
    Reports device rates and simulated seconds per step (the screening physics takes
    larger steps) in two windows, the early mostly dry one and one at the first storm peak
    (t_wet), the steps both physics take to reach t_wet, and the oracle's local-inertial
    f32 rate on the host cores."""
    import copy
    import numpy as np
    base2 = tr.TerrainBase.build(w.base.fine, 2, w.base.segments)
    terrs = [tr.PreparedTerrain(base2).apply_edit(e) for e in w.edits]
    nx, ny = w.base.nx, w.base.ny
    mon = np.unique((w.monitors // nx // 2) * base2.nx + (w.monitors % nx) // 2).astype(np.uint32)
    out = {"workload": f"{len(terrs)} candidates x {base2.nx}x{base2.ny} face-crest cells at "
                       f"{base2.dx:g} m (level 2 of the forward workload)", "steps": steps}
    def window(sess, n, warm=150):
        # `warm` untimed steps first, with no host reads between them and the timed ones
        # (the GPU's clocks drop during host work and take tens of milliseconds to recover)
        sess.advance(warm)
        t0 = time.perf_counter()
        st = sess.advance(n)
        wall = time.perf_counter() - t0
        t_b = np.array([r.t_end for r in sess.results(want_state=False)])
        kern_s = st["device_ms"] / 1e3 / n
        cells = st["cell_steps"] / n
        sim = st["sim_s"] / len(terrs)   # simulated seconds of the timed steps, per candidate
        return {"t_s_at_end": [float(t_b.min()), float(t_b.max())], "cell_steps_per_s": st["cell_steps"] / wall,
                "us_per_step": kern_s * 1e6, "sim_s_per_step": sim / n,
                "sim_s_per_wall_s": sim / wall, "hbm_frac_40B": 40.0 * cells / kern_s / HBM_PEAK}

    for kernel in ("inertial", "fused"):
        p = copy.copy(w.params)
        p.kernel = kernel
        s = solver_cls(0)
        tuned = s.autotune(p, [(base2, terrs, mon, None)], trial_steps=16)
        sess = s.session(p, [(base2, terrs, mon, None)])
        early = window(sess, steps, warm=warmup)
        # through the first storm peak: wall time to reach t_wet from the start of the run
        n_run, t0 = warmup + steps, time.perf_counter()
        while min(r.t_end for r in sess.results(want_state=False)) < t_wet:
            sess.advance(1000)
            n_run += 1000
        reach = time.perf_counter() - t0
        wet = window(sess, steps)
        sess.close()
        s.close()
        out["local_inertial" if kernel == "inertial" else "hll"] = {
            "band_rows": tuned[0][0], "early": early, "wet": wet,
            "steps_to_t_wet": n_run, "wall_to_t_wet_s": reach}
        say(f"screening {kernel}: {n_run} steps to t = {t_wet:g} s")
    li, hl = out["local_inertial"], out["hll"]
    out["sim_time_speedup_vs_hll"] = {
        "early": li["early"]["sim_s_per_wall_s"] / hl["early"]["sim_s_per_wall_s"],
        "wet": li["wet"]["sim_s_per_wall_s"] / hl["wet"]["sim_s_per_wall_s"],
        "steps_to_t_wet": hl["steps_to_t_wet"] / li["steps_to_t_wet"]}
    if cpu_threads:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po  # the checker, timed here as the CPU baseline only
        f = w.base.fine
        ot = po.Terrain(f.dem, f.wall, f.n, dx=f.geo.dx, dy=f.geo.dy, origin=(f.geo.origin_x, f.geo.origin_y),
                        level=2, segments=list(w.base.segments))
        p = w.params
        op = po.SolverParams(p.duration_s, p.sync_interval_s, cfl=p.cfl, rain=list(p.rain),
                             infil_capacity=p.infil_capacity, kernel="inertial")
        params, _keep = po._params(op, 32, 0, cpu_steps)
        cw, _, cu = _run_cands(po, ot, params, [e.as_dict() for e in w.edits[:cpu_cands]], cpu_threads, mon)
        out["cpu_local_inertial"] = {"value": float(cu.sum()) / cw, "unit": "cell-steps/s", "cores": cpu_threads,
                                     "sample": f"{cpu_cands} candidates x {cpu_steps} steps, one per thread"}
        out["gpu_over_cpu"] = li["early"]["cell_steps_per_s"] / out["cpu_local_inertial"]["value"]
    return out


def replay_leg(solver_cls, sc, tr, n_cand=128, duration=1800.0):
    """Subdomain replay (replay.rs) on the forward workload's terrain and storm: the
    candidates' primitives, the assets and the guard strip lie in a 64x64-cell window, and
    the region of interest is that window with a 48-cell margin.  The baseline records the
    tape; the candidates then run with the baseline's locked schedule on the full domain
    and replayed on the region (replay-invalid ones are the evaluator's escalations).  Both
    runs use band heights tuned for them beforehand (untimed)."""
    import numpy as np
    nx = ny = 512
    dx = 2.0
    rng = sc.SplitMix64(5)
    dem = sc.hills_dem(nx, ny, dx, rng, 40, n_gullies=8, sigma=(15.0, 60.0), noise=0.0)
    geo = tr.GeoTransform(0.0, ny * dx, dx, dx)
    base = tr.TerrainBase.build(tr.FineTerrain.new(dem, None, 0.035, geo), 1, [{"kind": "transmissive", "side": "east"}])
    p = sc.SolverParams(duration, 60.0, rain=sc.mm_h([[0, 0], [1800, 50], [3600, 10], [5400, 90], [7200, 0]]),
                        infil_capacity=5.0 / 1000.0 / 3600.0)
    w0, wn, margin = 224, 64, 48
    mon, _, _ = sc.asset_and_guard_cells(nx, ny, [(248, 248, 8, 8)], [])
    guard = np.array([r * nx + c for r in range(w0, w0 + wn) for c in range(w0 + wn + 12, w0 + wn + 28)], np.uint32)
    mon = np.unique(np.concatenate([mon, guard])).astype(np.uint32)
    roi = (w0 - margin, w0 - margin, wn + 2 * margin, wn + 2 * margin)
    geo_w = tr.GeoTransform(w0 * dx, (ny - w0) * dx, dx, dx)
    edits = [tr.materialize(sc.random_primitives(rng, 4, geo_w, wn, wn), geo, nx, ny, None, 1.0) for _ in range(n_cand)]
    terrs = [tr.PreparedTerrain(base).apply_edit(e) for e in edits]
    c0, r0, rw, rh = roi
    sub_mon = np.array([((c // nx) - r0) * rw + (c % nx) - c0 for c in mon], np.uint32)
    s = solver_cls(0)
    t0 = time.perf_counter()
    dts = s.dt_schedule(p, base, tr.PreparedTerrain(base), mon)
    tape, b0 = s.record_tape(p, base, roi, mon)
    t_rec = time.perf_counter() - t0
    pl = sc.SolverParams(duration, 60.0, rain=p.rain, infil_capacity=p.infil_capacity, locked_dt=dts)
    s.autotune(pl, [(base, terrs, mon, None)], trial_steps=16)   # untimed: band height of the full-domain batch
    t0 = time.perf_counter()
    sess = s.session(pl, [(base, terrs, mon, None)])
    sess.run()
    full = sess.results(want_state=False)
    sess.close()
    t_full = time.perf_counter() - t0
    rows = s.tune_replay_rows(pl, tape, terrs, sub_mon)   # untimed: band height of the replay batch
    t0 = time.perf_counter()
    rep = s.run_replay(pl, tape, terrs, sub_mon, want_state=False, band_rows=rows)
    t_rep = time.perf_counter() - t0
    s.close()
    valid = [r for r in rep if r.aborted == 0]
    same = sum(1 for r, f in zip(rep, full) if r.aborted == 0 and np.array_equal(r.monitor_max, f.monitor_max))
    dmax = max([float(np.abs(r.monitor_max - f.monitor_max).max()) for r, f in zip(rep, full) if r.aborted == 0] or [0.0])
    cs_full = sum(r.cell_updates for r in full)
    cs_rep = sum(r.cell_updates for r in rep)
    return {"workload": f"{n_cand} candidates x 4 primitives in a 64x64 window of the 512x512 terrain, "
                        f"{duration:g} s of the storm, region {rw}x{rh} cells ({100.0 * rw * rh / (nx * ny):.1f}% of the grid)",
            "steps": int(b0.steps), "record_s": t_rec, "full_domain_s": t_full, "replay_s": t_rep,
            "full_cell_steps_per_s": cs_full / t_full, "replay_cell_steps_per_s": cs_rep / t_rep,
            "replay_band_rows": rows,
            "valid": len(valid), "invalid": len(rep) - len(valid), "valid_monitors_equal_full": same,
            "valid_max_peak_depth_diff_m": dmax,
            "speedup_vs_full_domain": t_full / t_rep}


def say(*a):
    """Progress lines on stderr (the JSON result is the one stdout line)."""
    print(*a, file=sys.stderr, flush=True)


def advance_to(sess, n_steps, chunk=16384, label=""):
    """Run n_steps step launches of a resident batch (no sync callbacks), in chunks that
    each report progress; returns (cell-steps run, wall seconds)."""
    done, cs, t0 = 0, 0.0, time.perf_counter()
    while done < n_steps:
        k = min(chunk, n_steps - done)
        st = sess.advance(k)
        done += k
        cs += st["cell_steps"]
        say(f"{label}: {done}/{n_steps} steps, {cs / (time.perf_counter() - t0) / 1e9:.1f} G cell-steps/s")
    return cs, time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--cands", type=int, default=128)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--advance", type=int, default=165000,
                    help="untimed steps before the timed window (165000 reach about t = 5400 s, the storm peak)")
    ap.add_argument("--guard-tol", type=float, default=0.01, help="guard early-abort tolerance (m); 0: no guard")
    ap.add_argument("--band-rows", type=int, default=0)
    ap.add_argument("--min-lds", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint legs")
    ap.add_argument("--adjoint-horizon", type=int, default=0,
                    help="steps of the best-16 adjoint (0: to t = 1800 s on the best candidate's schedule)")
    ap.add_argument("--no-ladder", action="store_true", help="skip the fidelity-ladder leg")
    ap.add_argument("--no-screening", action="store_true", help="skip the local-inertial screening leg")
    ap.add_argument("--no-replay", action="store_true", help="skip the subdomain replay leg")
    ap.add_argument("--replay-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--screening-only", action="store_true",
                    help="run only the local-inertial screening leg (profiling) and print it")
    ap.add_argument("--no-autotune", action="store_true", help="use the default band height")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step directly (no hipGraph chunks): for runs under rocprofv3 --kernel-trace, "
                         "whose tracing of graph launches ends the process on this ROCm")
    ap.add_argument("--copy-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--list-split", type=int, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-cands", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=600)
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(candidates would shard across GPUs with no per-step exchange)", file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_package
    load_package()
    from inverse_terrain_engine_amd import scenario as sc
    from inverse_terrain_engine_amd import terrain as tr
    from inverse_terrain_engine_amd.gpu import GpuSolver as _GpuSolver
    import functools
    import numpy as np
    GpuSolver = functools.partial(_GpuSolver, use_graph=not a.no_graph)   # every leg's solvers

    t0 = time.time()
    w = sc.config5(n_cand=a.cands, n=a.n)
    t_prep = time.time() - t0
    if a.replay_only:
        print(json.dumps(replay_leg(GpuSolver, sc, tr)), flush=True)
        return
    if a.screening_only:
        print(json.dumps(screening_leg(GpuSolver, sc, tr, w, a.steps, warmup=a.warmup)), flush=True)
        return
    solver = GpuSolver(0, band_rows=a.band_rows)
    if a.min_lds:
        solver.set_option("min_lds", a.min_lds)
    if a.list_split is not None:
        solver.set_option("list_split", a.list_split)
    # The guard's baseline: the unedited terrain over the whole storm (objective.rs:132-144).
    guard = None
    base_steps = None
    if a.guard_tol > 0:
        tb = time.perf_counter()
        base_run = solver.run_batch(w.params, [tr.PreparedTerrain(w.base)], w.monitors, want_state=False)[0]
        base_steps = int(base_run.steps)
        say(f"guard baseline: {base_steps} steps in {time.perf_counter() - tb:.1f} s")
        guard = {"baseline_max": base_run.monitor_max, "guard_pos": w.guard_pos, "tol": a.guard_tol}
    groups = [(w.base, w.candidates, w.monitors, guard)]
    tuned = None
    if not a.band_rows and not a.no_autotune:
        tuned = solver.autotune(w.params, groups, trial_steps=16)
    sess = solver.session(w.params, groups)
    # The whole run up to the timed window: every step from t = 0 with the guard's early
    # aborts, host round trips between graph chunks included.
    # The library runs step launches as captured graph chunks of 32, 16, 8 and 4 launches
    # from launch counts that are multiples of 4; the untimed advance is padded so that the
    # warmup ends on such a count and the timed steps run as graph chunks.
    n_adv = a.advance + (-(a.advance + a.warmup)) % 4
    # The host reads of the batch's results leave the GPU idle for a while, and the
    # steps after such a pause run slower until its clocks are back up (about 50 steps of
    # this workload).  The results are read before the last untimed steps, so that the
    # warmup and the timed steps follow untimed steps with no host reads between them.
    tail = min(n_adv, 2048)
    whole_cs, whole_wall = advance_to(sess, n_adv - tail, label="advance") if n_adv > tail else (0.0, 0.0)
    mid = sess.results(want_state=False)
    t_a = np.array([r.t_end for r in mid])
    if tail:
        tail_cs, tail_wall = advance_to(sess, tail, label="advance (last untimed steps)")
        whole_cs, whole_wall = whole_cs + tail_cs, whole_wall + tail_wall
    if a.warmup:
        sess.advance(a.warmup)
    # advance() returns after synchronizing the library's stream (HIP event), so the
    # timed region is bracketed by device synchronization on both sides.
    t1 = time.perf_counter()
    st = sess.advance(a.steps)
    wall = time.perf_counter() - t1
    cell_steps = st["cell_steps"]
    value = cell_steps / wall
    kernel_s = st["device_ms"] / 1e3 / max(a.steps, 1)
    cells_per_launch = cell_steps / max(a.steps, 1)
    achieved = BYTES_PER_CELL_STEP * cells_per_launch / kernel_s
    res = sess.results(want_state=False)
    sess.close()
    active = [i for i, r in enumerate(res) if r.aborted == 0 and r.t_end < w.params.duration_s]
    t_b = np.array([r.t_end for r in res])
    adj = adj16 = None
    if not a.no_adjoint and not a.copy_probe:
        from inverse_terrain_engine_amd import objective as ob
        adj = adjoint_config4(solver, sc, ob)
        if a.cands >= 16:
            adj16 = adjoint_best16(solver, sc, ob, w, mid if n_adv else res, guard, a.adjoint_horizon)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched forward with guard early abort: {a.cands} candidates x {a.n}x{a.n} cells at 2 m, "
                        "4 primitives each, seeded DEM (40 hills, 8 gullies), bimodal 2 h hyetograph (50/90 mm/h), "
                        f"infiltration 5 mm/h, guard {a.guard_tol:g} m against the unedited baseline; timed window "
                        f"after {n_adv} untimed steps",
            "candidates": a.cands, "nx": a.n, "ny": a.n, "band_rows": st["band_rows"],
            "waves_per_task": st["waves_per_task"], "launch_lists": st["launch_lists"],
            "device": solver.adapter_info(),
            "autotuned": tuned is not None,
            "hbm_fraction": achieved / HBM_PEAK, "prep_s": round(t_prep, 2),
            # t of the running candidates before the last untimed steps and at the window's end
            "timed_window_t_s": [float(t_a[active].min()), float(t_b[active].max())] if active else None,
            "candidates_running_in_window": len(active),
            "guard_aborted": sum(1 for r in res if r.aborted == 1),
            "guard_baseline_steps": base_steps,
        },
        "whole_run": {
            "what": f"the same batch from t = 0 through the {n_adv} untimed steps (to t = "
                    f"{float(np.mean([r.t_end for r in res])):.0f} s on average at the timed window's end), guard "
                    "aborts and host round trips between graph chunks included",
            "cell_steps_per_s": whole_cs / whole_wall if whole_wall else None, "wall_s": whole_wall,
            "guard_aborted": sum(1 for r in res if r.aborted == 1)},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": None,
            # per step: with the candidates split into launch lists on their own streams, a
            # step is one launch of each list, side by side (HIP events on the library stream
            # after the lists' join)
            "kernel_avg_us": kernel_s * 1e6,
            "bytes_per_cell_step": BYTES_PER_CELL_STEP,
            "cells_per_launch": cells_per_launch,
        },
    }
    if not a.no_ladder and not a.copy_probe:
        out["ladder"] = ladder(GpuSolver, sc, 200)
    if not a.copy_probe:
        out["prep"] = prep_leg(solver, sc, tr, w)
        if not a.no_replay:
            out["replay"] = replay_leg(GpuSolver, sc, tr)
        if not a.no_screening:
            out["screening"] = screening_leg(GpuSolver, sc, tr, w, 200,
                                             cpu_threads=0 if a.no_cpu else (a.cpu_threads or min(16, os.cpu_count() or 1)))
    if adj is not None:
        out["adjoint"] = adj
    if adj16 is not None:
        out["adjoint_best16"] = adj16
    solver.close()
    if a.copy_probe:
        copy_probe()
    if not a.no_traffic:
        t = measure_traffic(st["band_rows"], float(a.cands) * a.n * a.n)
        if t is not None:
            cells = cells_per_launch
            fetch_b, write_b = t["fetch_per_cell_step"] * cells, t["write_per_cell_step"] * cells
            dram = (fetch_b + write_b) / kernel_s
            valu = t["valu_lane_insts_per_cell_step"] * cells / 64.0
            valu_frac = valu / (kernel_s * VALU_PEAK)
            out["roofline"]["traffic"] = fetch_b + write_b
            out["roofline"]["traffic_detail"] = {
                "fetch_bytes": fetch_b, "write_bytes": write_b,
                "bytes_per_cell_step": t["fetch_per_cell_step"] + t["write_per_cell_step"],
                "copy_scale_fetch": t["copy_scale_fetch"], "copy_scale_write": t["copy_scale_write"],
                "dram_gbs": dram / 1e9, "note": t["note"]}
            out["roofline"]["valu"] = {
                "insts_per_launch": valu, "insts_per_cell_step_lane": t["valu_lane_insts_per_cell_step"],
                "peak_insts_per_s": VALU_PEAK, "frac": valu_frac}
            # The resource the kernel is closest to (the metric is priced against HBM; the
            # step also issues VALU instructions at a measured fraction of the peak).
            out["roofline"]["nearest_limit"] = "valu_issue" if valu_frac > dram / HBM_PEAK else "hbm"
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        v, cw, cs = cpu_baseline(w, threads, a.cpu_cands, a.cpu_steps)
        st1 = cpu_single_thread_bench_case()
        out["cpu_baseline"] = {
            "value": v, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "single_thread_bench_case_512": st1, "published_single_thread_m4pro": 1.531e8,
            "sample": f"{a.cpu_cands} candidates of the same workload x {a.cpu_steps} steps, one candidate per "
                      f"thread (C++ restatement of the reference f32 fused kernel), {cw:.1f} s wall, CPU {cpu_model()}",
        }
        out["config"]["gpu_over_cpu"] = value / v
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
