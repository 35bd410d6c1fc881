"""Throughput of the batched f32 shallow-water step on one MI355X.

A "step" is one time step of every candidate of the batch: one launch of the HIP step
kernel over 128 candidates of a 512x512 grid at 2 m (the large batched-forward
workload; seeded DEM with 40 hills and 8 gullies, 4 primitives per candidate, bimodal
2 h storm).  W untimed warmup steps, then exactly K timed steps with the inputs already
resident in HBM.  Prints one JSON line.

The cpu_baseline leg times the CPU oracle (the reference's f32 fused kernel restated in
C++, bitwise equal to it) candidate-parallel on the host cores, one candidate per thread
like the reference's evaluator, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, B/s
BYTES_PER_CELL_STEP = 32   # read h, qx, qy, dzx, dzy + write h, qx, qy (f32)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--cands", type=int, default=128)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--band-rows", type=int, default=0)
    ap.add_argument("--min-lds", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-cands", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=60)
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(candidates would shard across GPUs with no per-step exchange)", file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_package
    load_package()
    from inverse_terrain_engine_amd import scenario as sc
    from inverse_terrain_engine_amd.gpu import GpuSolver

    t0 = time.time()
    w = sc.config5(n_cand=a.cands, n=a.n)
    t_prep = time.time() - t0
    solver = GpuSolver(0, band_rows=a.band_rows)
    if a.min_lds:
        solver.set_option("min_lds", a.min_lds)
    sess = solver.session(w.params, [(w.base, w.candidates, w.monitors, None)])
    if a.warmup:
        sess.advance(a.warmup)
    try:
        import torch
        torch.cuda.synchronize()
    except Exception:
        torch = None
    t1 = time.perf_counter()
    st = sess.advance(a.steps)
    if torch is not None:
        torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    cell_steps = st["cell_steps"]
    value = cell_steps / wall
    kernel_s = st["device_ms"] / 1e3 / max(a.steps, 1)
    cells_per_launch = cell_steps / max(a.steps, 1)
    achieved = BYTES_PER_CELL_STEP * cells_per_launch / kernel_s
    res = sess.results(want_state=False)
    sess.close()

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
            "workload": f"batched forward: {a.cands} candidates x {a.n}x{a.n} cells at 2 m, 4 primitives each, "
                        "seeded DEM (40 hills, 8 gullies), bimodal 2 h hyetograph (50/90 mm/h), infiltration 5 mm/h",
            "candidates": a.cands, "nx": a.n, "ny": a.n, "band_rows": st["band_rows"],
            "waves_per_task": st["waves_per_task"], "device": solver.adapter_info(),
            "hbm_fraction": achieved / HBM_PEAK, "prep_s": round(t_prep, 2),
            "steps_run_per_candidate": int(res[0].steps),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": None,
            "kernel_avg_us": kernel_s * 1e6,
            "bytes_per_cell_step": BYTES_PER_CELL_STEP,
            "cells_per_launch": cells_per_launch,
        },
    }
    solver.close()
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        v, cw, cs = cpu_baseline(w, threads, a.cpu_cands, a.cpu_steps)
        out["cpu_baseline"] = {
            "value": v, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"{a.cpu_cands} candidates of the same workload x {a.cpu_steps} steps, one candidate per "
                      f"thread (C++ restatement of the reference f32 fused kernel), {cw:.1f} s wall, CPU {cpu_model()}",
        }
        out["config"]["gpu_over_cpu"] = value / v
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
