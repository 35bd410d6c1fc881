"""Python host mirror of the reference's GPU batch backend, over the C ABI.

``GpuSolver`` mirrors crates/itr-gpu/src/lib.rs (``new``, ``adapter_info``,
``run_batch`` returning ``GpuRun`` per candidate) and ``BatchBackend.run`` mirrors the
evaluator seam (crates/itr-opt/src/evaluator.rs:127-135, glue crates/itr-cli/src/backend.rs:48-63).
Errors follow the reference: a declined configuration raises ``Declined`` (the
reference's ``Err(String)``), which callers treat as "use the CPU for this level".
Every call goes to the HIP library; there is no CPU fallback in this module.
"""
from __future__ import annotations

import copy
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .scenario import SolverParams
from .terrain import PreparedTerrain, TerrainBase


class Declined(RuntimeError):
    """The backend does not support this configuration (reference: Err(String))."""


class DeviceError(RuntimeError):
    pass


def _check(rc, err):
    if rc == _abi.ITR_OK:
        return
    msg = err.value.decode(errors="replace") if err is not None else ""
    if rc == _abi.ITR_DECLINED:
        raise Declined(msg)
    raise DeviceError(f"{_abi.STATUS_NAMES.get(rc, rc)}: {msg}")


@dataclass
class GpuRun:
    """GpuRun (lib.rs:49-60) + BatchRun (evaluator.rs:118-125) + RunSummary fields."""
    monitor_max: np.ndarray
    monitor_tmax: np.ndarray
    steps: int
    t_end: float
    volume_m3: float
    depth: np.ndarray
    qx: np.ndarray
    qy: np.ndarray
    bad: bool
    cell_updates: int
    rows_skipped: int
    dt_min: float
    dt_max: float
    dt_mean: float
    lts: float
    v0: float
    rel_residual: float
    aborted: int
    abort_cell: int
    abort_sync: int
    abort_value: float
    sync_count: int
    ledger: np.ndarray = field(repr=False, default=None)
    dt_fallback_t: float = -1.0


class _Keep(list):
    def arr(self, a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        self.append(a)
        return a


def _params_c(p: SolverParams, keep: _Keep) -> _abi.SolverParamsC:
    q = _abi.SolverParamsC()
    q.duration_s, q.sync_interval_s, q.cfl, q.dt_max_s = p.duration_s, p.sync_interval_s, p.cfl, p.dt_max_s
    q.h_dry, q.h_eps, q.infil_capacity, q.initial_depth = p.h_dry, p.h_eps, p.infil_capacity, p.initial_depth
    q.has_initial_stage = 0 if p.initial_stage is None else 1
    q.initial_stage = 0.0 if p.initial_stage is None else float(p.initial_stage)
    q.cfl_hard, q.inflow_scale = p.cfl_hard, p.inflow_scale
    rt = keep.arr([r[0] for r in p.rain] or [0.0], np.float64)
    rr = keep.arr([r[1] for r in p.rain] or [0.0], np.float64)
    q.rain_t, q.rain_r, q.n_rain = rt.ctypes.data_as(_abi._dp), rr.ctypes.data_as(_abi._dp), len(p.rain)
    if p.locked_dt is not None:
        ld = keep.arr(p.locked_dt, np.float64)
        q.locked_dt, q.n_locked = ld.ctypes.data_as(_abi._dp), len(ld)
    q.max_steps = int(p.max_steps)
    q.skip_dry = 1 if p.skip_dry else 0
    # KernelKind: the oracle kernel is bitwise equal to the fused one (equivalence.rs),
    # local inertial is the coarse-level screening physics (inertial.rs).
    phys = {"fused": _abi.ITR_PHYSICS_HLL, "oracle": _abi.ITR_PHYSICS_HLL,
            "inertial": _abi.ITR_PHYSICS_LOCAL_INERTIAL}.get(p.kernel)
    if phys is None:
        raise ValueError(f"unknown kernel {p.kernel!r}")
    q.physics = phys
    return q


def _grid_c(base: TerrainBase, terrains, monitors, keep: _Keep, guard=None, need_z=False) -> _abi.GridC:
    g = _abi.GridC()
    nx, ny = base.nx, base.ny
    g.nx, g.ny, g.dx, g.dy, g.g, g.z_ref = nx, ny, base.dx, base.dy, base.g, base.z_ref
    g.crest = 1 if base.crest else 0
    g.wet_count = base.wet_count
    f = lambda a: a.ctypes.data_as(_abi._fp)
    g.wet = f(keep.arr(base.wet, np.float32))
    g.gn2 = f(keep.arr(base.gn2, np.float32))
    codes = base.ghost_codes()
    g.ghost = keep.arr(codes, np.int32).ctypes.data_as(_abi._ip)
    if (codes >= 2).any():
        # Stage / Inflow cells (boundary.rs:43-47, 66-70): per-cell parameter, hydrograph
        # index, and the hydrographs flattened with offsets.
        g.ghost_value = keep.arr(base.ghost_params(), np.float64).ctypes.data_as(_abi._dp)
        g.ghost_hydro = keep.arr(base.ghost_hydro(), np.int32).ctypes.data_as(_abi._ip)
        hyd = [np.asarray(h, np.float64).reshape(-1, 2) for h in base.ghosts.get("hydrographs", [])]
        if hyd:
            start = np.concatenate([[0], np.cumsum([len(h) for h in hyd])]).astype(np.int32)
            pts = np.concatenate(hyd) if start[-1] > 0 else np.zeros((1, 2))
            g.hydro_start = keep.arr(start, np.int32).ctypes.data_as(_abi._ip)
            g.hydro_t = keep.arr(pts[:, 0], np.float64).ctypes.data_as(_abi._dp)
            g.hydro_q = keep.arr(pts[:, 1], np.float64).ctypes.data_as(_abi._dp)
            g.n_hydro = len(hyd)
        need_z = need_z or base.has_stage
    for t in terrains:
        if t.base is not base:
            raise Declined("all candidates must share the same TerrainBase")
    g.n_cand = len(terrains)
    g.dzx = f(keep.arr(np.stack([t.fx0 for t in terrains]) if terrains else np.zeros(1), np.float32))
    g.dzy = f(keep.arr(np.stack([t.fy0 for t in terrains]) if terrains else np.zeros(1), np.float32))
    if base.crest and terrains:
        g.dzx_r = f(keep.arr(np.stack([t.fx1 for t in terrains]), np.float32))
        g.dzy_r = f(keep.arr(np.stack([t.fy1 for t in terrains]), np.float32))
    if need_z and terrains:
        g.z_level = keep.arr(np.stack([t.z_level for t in terrains]), np.float64).ctypes.data_as(_abi._dp)
    mon = keep.arr(np.asarray(monitors, np.uint32) if len(monitors) else np.zeros(1, np.uint32), np.uint32)
    g.mon_cells = mon.ctypes.data_as(_abi._up)
    g.n_mon = len(monitors)
    if guard is not None:
        g.baseline_max = keep.arr(guard["baseline_max"], np.float64).ctypes.data_as(_abi._dp)
        g.guard_pos = keep.arr(guard["guard_pos"], np.int32).ctypes.data_as(_abi._ip)
        g.n_guard = len(guard["guard_pos"])
        g.guard_tol = float(guard["tol"])
    return g


def _design_grid_c(base: TerrainBase, designs, monitors, keep: _Keep, editable=None, max_abs_dz=1e9, guard=None):
    """A grid whose candidates are primitive lists, materialized on the device
    (itr_design) on the fine grid: level-1 signed grids, or face-crest levels >= 2."""
    from .terrain import design_rows
    if (base.level != 1) != bool(base.crest):
        raise Declined("device designs need a level-1 signed or a face-crest coarse grid")
    fine = base.fine
    rows, start, boxes = [], [0], []
    for prims in designs:
        r, box = design_rows(prims, fine.geo, fine.nx, fine.ny)
        rows += r
        start.append(len(rows))
        boxes.append(box)
    d = _abi.DesignC()
    d.prim_start = keep.arr(start, np.int32).ctypes.data_as(_abi._ip)
    d.prim = keep.arr(np.asarray(rows, np.float64).reshape(-1, 12) if rows else np.zeros((1, 12)),
                      np.float64).ctypes.data_as(_abi._dp)
    d.bbox = keep.arr(np.asarray(boxes, np.int32).reshape(-1, 4) if boxes else np.zeros((1, 4)),
                      np.int32).ctypes.data_as(_abi._ip)
    d.z_fine = keep.arr(fine.z, np.float64).ctypes.data_as(_abi._dp)
    d.wall = keep.arr(fine.wall, np.uint8).ctypes.data_as(C.POINTER(C.c_uint8))
    if editable is not None:
        d.editable = keep.arr(np.asarray(editable, bool), np.uint8).ctypes.data_as(C.POINTER(C.c_uint8))
    g = fine.geo
    d.origin_x, d.origin_y, d.dx, d.dy = g.origin_x, g.origin_y, g.dx, g.dy
    d.max_abs_dz, d.z_wall = float(max_abs_dz), float(fine.z_wall)
    d.level, d.fine_nx, d.fine_ny = int(base.level), int(fine.nx), int(fine.ny)
    vol = keep.arr(np.zeros((max(len(designs), 1), 2)), np.float64)
    d.volumes = vol.ctypes.data_as(_abi._dp)
    keep.append(d)
    grid = _grid_c(base, [], monitors, keep, guard)
    grid.n_cand = len(designs)
    grid.dzx = grid.dzy = grid.dzx_r = grid.dzy_r = None
    grid.design = C.pointer(d)
    return grid, vol


@dataclass
class Tape:
    """A subdomain-replay tape (replay.rs:71-90): the region (c0, r0, w, h), the ghost ring
    of every step [steps][2(w+2)+2h][3] and the inner band of every sync point
    [syncs][n_band][3], as f32 (h, qx, qy)."""
    roi: tuple
    ring: np.ndarray
    band: np.ndarray
    steps: int
    syncs: int


def ring_len(roi):
    return 2 * (roi[2] + 2) + 2 * roi[3]


def band_len(roi):
    w, h = roi[2], roi[3]
    return sum(1 for j in range(1, h + 1) for i in range(1, w + 1) if i in (1, w) or j in (1, h))


def _replay_grid_c(tape: Tape, sub_terrains, sub_monitors, keep: _Keep, tol_h, tol_q, guard=None):
    """The region as a grid of its own (sub_terrain, replay.rs:181-241): cells, every face
    on and around it from each candidate's full prepared terrain, the tape as its ghosts."""
    c0, r0, w, h = tape.roi
    base = sub_terrains[0].base
    g = _abi.GridC()
    g.nx, g.ny, g.dx, g.dy, g.g, g.z_ref = w, h, base.dx, base.dy, base.g, base.z_ref
    g.crest = 1 if base.crest else 0
    wet = keep.arr(base.wet[r0:r0 + h, c0:c0 + w], np.float32)
    g.wet_count = int((wet != 0).sum())
    f = lambda a: a.ctypes.data_as(_abi._fp)
    g.wet = f(wet)
    g.gn2 = f(keep.arr(base.gn2[r0:r0 + h, c0:c0 + w], np.float32))
    g.ghost = keep.arr(np.zeros(2 * w + 2 * h), np.int32).ctypes.data_as(_abi._ip)
    for t in sub_terrains:
        if t.base is not base:
            raise Declined("all candidates must share the same TerrainBase")
    g.n_cand = len(sub_terrains)
    g.dzx = f(keep.arr(np.stack([t.fx0[r0:r0 + h, c0:c0 + w + 1] for t in sub_terrains]), np.float32))
    g.dzy = f(keep.arr(np.stack([t.fy0[r0:r0 + h + 1, c0:c0 + w] for t in sub_terrains]), np.float32))
    if base.crest:
        g.dzx_r = f(keep.arr(np.stack([t.fx1[r0:r0 + h, c0:c0 + w + 1] for t in sub_terrains]), np.float32))
        g.dzy_r = f(keep.arr(np.stack([t.fy1[r0:r0 + h + 1, c0:c0 + w] for t in sub_terrains]), np.float32))
    g.z_level = keep.arr(np.stack([t.z_level[r0:r0 + h, c0:c0 + w] for t in sub_terrains]),
                         np.float64).ctypes.data_as(_abi._dp)
    mon = keep.arr(np.asarray(sub_monitors, np.uint32) if len(sub_monitors) else np.zeros(1, np.uint32), np.uint32)
    g.mon_cells = mon.ctypes.data_as(_abi._up)
    g.n_mon = len(sub_monitors)
    if guard is not None:
        g.baseline_max = keep.arr(guard["baseline_max"], np.float64).ctypes.data_as(_abi._dp)
        g.guard_pos = keep.arr(guard["guard_pos"], np.int32).ctypes.data_as(_abi._ip)
        g.n_guard = len(guard["guard_pos"])
        g.guard_tol = float(guard["tol"])
    rp = _abi.ReplayC()
    rp.mode, rp.c0, rp.r0, rp.w, rp.h = 2, c0, r0, w, h
    rp.ring = f(keep.arr(tape.ring, np.float32))
    rp.band = f(keep.arr(tape.band, np.float32))
    rp.ring_steps, rp.band_syncs = tape.steps, tape.syncs
    rp.tol_h, rp.tol_q = float(tol_h), float(tol_q)
    keep.append(rp)
    g.replay = C.pointer(rp)
    return g


class GpuSolver:
    """GpuSolver (lib.rs:62-373) on one MI355X through the C ABI."""

    def __init__(self, device: int = 0, band_rows: int = 0, steps_per_chunk: int = 0, use_graph: bool = True):
        self._L = _abi.load()
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        _check(self._L.itr_hip_create(device, C.byref(h), err, 512), err)
        self._h = h
        for k, v in (("band_rows", band_rows), ("steps_per_chunk", steps_per_chunk),
                     ("use_graph", 1 if use_graph else 0)):
            self._L.itr_hip_set_option(self._h, k.encode(), int(v))
        self._band_rows = int(band_rows)
        self.last_stats = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.itr_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def adapter_info(self) -> str:
        return self._L.itr_hip_device_info(self._h).decode()

    def set_option(self, name: str, value: int):
        _check(self._L.itr_hip_set_option(self._h, name.encode(), int(value)), None)
        if name == "band_rows":
            self._band_rows = int(value)

    def adjoint(self, p: SolverParams, base: TerrainBase, terrains, monitors, assets, tau_m: float):
        """Discrete adjoint of J_risk for each candidate on the locked schedule p.locked_dt.

        ``assets``: list of (positions into ``monitors``, threshold_m, weight).  Returns one
        dict per candidate: j_risk, peak (per asset), dJ_dz_cell (ny, nx), dJ_dz_face_x
        (ny, nx+1), dJ_dz_face_y (ny+1, nx), steps, segment_steps, checkpoints,
        replay_mismatches and the device times of the plain forward, the checkpointed
        forward, the replay and the reverse sweep (ms).  On a face-crest level the face
        entries are the gradients of the faces' crest lines, dJ_dz_cell holds the crest lines
        fixed, and dJ_drise_face_x (ny, nx+1, 2) / dJ_drise_face_y (ny+1, nx, 2) are the
        gradients of the rises (d_l, d_r); terrain.fine_gradient maps them to the fine
        raster."""
        if p.locked_dt is None:
            raise ValueError("the adjoint needs p.locked_dt")
        keep = _Keep()
        q = _params_c(p, keep)
        g = _grid_c(base, terrains, monitors, keep, need_z=p.initial_stage is not None)
        starts = np.zeros(len(assets) + 1, np.int32)
        pos = []
        for i, (ps, _, _) in enumerate(assets):
            pos.extend(int(x) for x in ps)
            starts[i + 1] = len(pos)
        ob = _abi.ObjectiveC()
        ob.n_assets = len(assets)
        ob.asset_start = keep.arr(starts, np.int32).ctypes.data_as(_abi._ip)
        ob.asset_pos = keep.arr(np.array(pos if pos else [0], np.int32), np.int32).ctypes.data_as(_abi._ip)
        ob.threshold_m = keep.arr(np.array([a[1] for a in assets] or [0.0]), np.float64).ctypes.data_as(_abi._dp)
        ob.weight = keep.arr(np.array([a[2] for a in assets] or [0.0]), np.float64).ctypes.data_as(_abi._dp)
        ob.tau_m = float(tau_m)
        n = len(terrains)
        nx, ny = base.nx, base.ny
        outs = (_abi.AdjointOutC * n)()
        bufs = []
        for i in range(n):
            b = {"peak": np.zeros(max(len(assets), 1)), "cell": np.zeros((ny, nx)), "fx": np.zeros((ny, nx + 1)),
                 "fy": np.zeros((ny + 1, nx))}
            if base.crest:
                b["rx"] = np.zeros((ny, nx + 1, 2))
                b["ry"] = np.zeros((ny + 1, nx, 2))
                outs[i].dJ_drise_face_x = b["rx"].ctypes.data_as(_abi._dp)
                outs[i].dJ_drise_face_y = b["ry"].ctypes.data_as(_abi._dp)
            outs[i].peak_m = b["peak"].ctypes.data_as(_abi._dp)
            outs[i].dJ_dz_cell = b["cell"].ctypes.data_as(_abi._dp)
            outs[i].dJ_dz_face_x = b["fx"].ctypes.data_as(_abi._dp)
            outs[i].dJ_dz_face_y = b["fy"].ctypes.data_as(_abi._dp)
            bufs.append(b)
        err = C.create_string_buffer(1024)
        _check(self._L.itr_hip_adjoint(self._h, C.byref(q), C.byref(g), C.byref(ob), outs, err, 1024), err)
        res = []
        for i in range(n):
            o, b = outs[i], bufs[i]
            res.append({"j_risk": o.j_risk, "peak": b["peak"][:len(assets)].copy(), "dJ_dz_cell": b["cell"],
                        "dJ_dz_face_x": b["fx"], "dJ_dz_face_y": b["fy"], "steps": int(o.steps),
                        "segment_steps": o.segment_steps, "checkpoints": o.checkpoints,
                        "replay_mismatches": int(o.replay_mismatches), "plain_forward_ms": o.plain_forward_ms,
                        "forward_ms": o.forward_ms, "replay_ms": o.replay_ms, "reverse_ms": o.reverse_ms})
            if base.crest:
                res[-1]["dJ_drise_face_x"] = b["rx"]
                res[-1]["dJ_drise_face_y"] = b["ry"]
        return res

    def record_tape(self, p: SolverParams, base: TerrainBase, roi, monitors=(), max_steps: int = 100000,
                    max_syncs: int = 10000):
        """The baseline run on the device recording the subdomain-replay tape of `roi`
        (c0, r0, w, h) (problem.rs:306-319).  Returns (Tape, GpuRun of the baseline)."""
        keep = _Keep()
        q = _params_c(p, keep)
        g = _grid_c(base, [PreparedTerrain(base)], monitors, keep, None, p.initial_stage is not None)
        c0, r0, w, h = (int(v) for v in roi)
        ring = np.zeros((max_steps, ring_len(roi), 3), np.float32)
        band = np.zeros((max_syncs, band_len(roi), 3), np.float32)
        ns, nsy = C.c_int64(0), C.c_int64(0)
        rp = _abi.ReplayC()
        rp.mode, rp.c0, rp.r0, rp.w, rp.h = 1, c0, r0, w, h
        rp.ring = ring.ctypes.data_as(_abi._fp)
        rp.band = band.ctypes.data_as(_abi._fp)
        rp.ring_steps, rp.band_syncs = max_steps, max_syncs
        rp.n_steps_out, rp.n_syncs_out = C.pointer(ns), C.pointer(nsy)
        g.replay = C.pointer(rp)
        s = BatchSession(self, p, [(base, [PreparedTerrain(base)], monitors, None)], None, grids=[g], keep=keep)
        try:
            s.run()
            run = s.results()[0]
        finally:
            s.close()
        if ns.value > max_steps or nsy.value > max_syncs:
            raise ValueError("tape capacity exceeded")
        return Tape((c0, r0, w, h), ring[:ns.value].copy(), band[:nsy.value].copy(), ns.value, nsy.value), run

    def run_replay(self, p: SolverParams, tape: Tape, terrains, sub_monitors, tol_h=2e-3, tol_q=1e-3, guard=None,
                   want_state=True, band_rows: int = 0):
        """Candidates replayed on the tape's region (evaluator.rs:427-440): aborted = 5 is
        ReplayInvalid (escalate to a full-domain run).  State arrays are region-sized.
        band_rows > 0 sets the band height for this call (tune_replay_rows finds one); the
        results are the same bits at any height."""
        keep = _Keep()
        g = _replay_grid_c(tape, terrains, sub_monitors, keep, tol_h, tol_q, guard)
        saved = self._band_rows
        if band_rows > 0:
            self._L.itr_hip_set_option(self._h, b"band_rows", int(band_rows))
        try:
            s = BatchSession(self, p, [(None, terrains, sub_monitors, None)], None, grids=[g], keep=keep,
                             shapes=[(tape.roi[2], tape.roi[3])])
            try:
                s.run()
                return s.results(want_state)
            finally:
                s.close()
        finally:
            if band_rows > 0:
                self._L.itr_hip_set_option(self._h, b"band_rows", saved)

    def tune_replay_rows(self, p: SolverParams, tape: Tape, terrains, sub_monitors, rows=(8, 12, 16, 20, 24, 28),
                         trial_steps: int = 512, tol_h=2e-3, tol_q=1e-3, guard=None) -> int:
        """Band height for replay batches like this one: the first trial_steps steps of the
        replay at each height in `rows` (after one warm-up trial), the fastest by device
        time.  The replay kernel carries the tape's ghost ring and band checks, so its best
        height differs from the plain step's (itr_hip_autotune tunes those)."""
        q = copy.copy(p)
        q.max_steps = int(trial_steps)
        best, best_ms = 0, float("inf")
        for i, R in enumerate((rows[0],) + tuple(rows)):
            self.run_replay(q, tape, terrains, sub_monitors, tol_h, tol_q, guard, want_state=False, band_rows=R)
            if i == 0:
                continue   # warm-up
            ms = self.last_stats["device_ms"] if self.last_stats else float("inf")
            if ms < best_ms:
                best, best_ms = R, ms
        return best

    def dt_schedule(self, p: SolverParams, base: TerrainBase, terrain, monitors=(), max_steps: int = 0,
                    capacity: int = 1 << 20) -> np.ndarray:
        """dt of every step of one candidate's free forward (the schedule to lock to)."""
        keep = _Keep()
        q = _params_c(p, keep)
        if max_steps:
            q.max_steps = int(max_steps)
        g = _grid_c(base, [terrain], monitors, keep)
        cap = int(max_steps) if max_steps else int(capacity)
        out = np.zeros(cap)
        n = C.c_int64(0)
        err = C.create_string_buffer(1024)
        _check(self._L.itr_hip_dt_schedule(self._h, C.byref(q), C.byref(g), 0, out.ctypes.data_as(_abi._dp), cap,
                                           C.byref(n), err, 1024), err)
        return out[:n.value].copy()

    def autotune(self, p: SolverParams, groups, trial_steps: int = 16):
        """Choose band heights per grid shape by timing (itr_hip_autotune); later batches
        of this solver use them.  ``groups`` as for ``session``.  Returns [(band_rows,
        us_per_step)] per group."""
        keep = _Keep()
        q = _params_c(p, keep)
        grids = (_abi.GridC * len(groups))()
        for i, (base, terrs, mons, guard) in enumerate(groups):
            grids[i] = _grid_c(base, terrs, mons, keep, guard, p.initial_stage is not None)
        out = (_abi.TuneResultC * len(groups))()
        err = C.create_string_buffer(1024)
        _check(self._L.itr_hip_autotune(self._h, C.byref(q), grids, len(groups), int(trial_steps), out, err, 1024),
               err)
        return [(o.band_rows, o.us_per_step) for o in out]

    def math_selfcheck(self):
        """(reciprocal, sqrt) mismatch counts of the step's short sequences against the
        IEEE operations over all 2^32 inputs; (0, 0) for a correct build."""
        a, b = C.c_uint64(0), C.c_uint64(0)
        err = C.create_string_buffer(512)
        _check(self._L.itr_hip_math_selfcheck(self._h, C.byref(a), C.byref(b), err, 512), err)
        return int(a.value), int(b.value)

    def session(self, p: SolverParams, groups, on_sync=None):
        """Upload a batch once (inputs stay resident); ``groups`` is a list of
        (TerrainBase, [PreparedTerrain], monitors, guard-or-None)."""
        return BatchSession(self, p, groups, on_sync)

    def run_ensemble(self, members, terrains, monitors, guards=None, want_state=True):
        """Every candidate under every ensemble member's parameters in one batch (one grid
        entry per member, each with its own forcing and guard baseline).  The reference
        runs members on the CPU (evaluator.rs:212-215, 494-560), each member a
        ``scaled`` copy of the level's parameters (problem.rs:56-63).  Returns
        runs[member][candidate]."""
        if len(terrains) == 0 or len(members) == 0:
            return [[] for _ in members]
        base = terrains[0].base
        groups = [(base, terrains, monitors, None if guards is None else guards[m]) for m in range(len(members))]
        s = BatchSession(self, members[0], groups, None, group_params=list(members))
        try:
            s.run()
            flat = s.results(want_state)
        finally:
            s.close()
        n = len(terrains)
        return [flat[m * n:(m + 1) * n] for m in range(len(members))]

    def run_designs(self, p: SolverParams, base: TerrainBase, designs, monitors, editable=None, max_abs_dz=1e9,
                    guard=None, want_state=True):
        """Candidates given as primitive lists: DesignSpace::materialize and
        PreparedTerrain::apply_edit (design.rs:316-385, terrain.rs:460-505) run on the
        device, then the batch.  Returns (runs, volumes[n, 2] = cut_m3, fill_m3)."""
        if len(designs) == 0:
            return [], np.zeros((0, 2))
        s = self.session_designs(p, base, designs, monitors, editable, max_abs_dz, guard)
        try:
            s.run()
            runs = s.results(want_state)
        finally:
            s.close()
        return runs, s.volumes

    def session_designs(self, p: SolverParams, base: TerrainBase, designs, monitors, editable=None, max_abs_dz=1e9,
                        guard=None):
        """A resident batch of candidates given as primitive lists, materialized on the
        device; ``session.volumes`` holds each edit's (cut_m3, fill_m3)."""
        keep = _Keep()
        grid, vol = _design_grid_c(base, designs, monitors, keep, editable, max_abs_dz, guard)
        s = BatchSession(self, p, [(base, [None] * len(designs), monitors, guard)], None, grids=[grid], keep=keep)
        s.volumes = vol[:len(designs)].copy()
        return s

    def full_run(self, p: SolverParams, terrain):
        """Full-domain run for the result rasters (crates/itr-cli/src/run.rs:73-93): every
        cell monitored; returns (run, peak, tpeak) with peak NaN on walls and tpeak NaN on
        walls and where the peak is not positive."""
        base = terrain.base
        mon = np.arange(base.nx * base.ny, dtype=np.uint32)
        run = self.run_batch(p, [terrain], mon)[0]
        wall = base.wall.reshape(-1)
        peak = np.where(wall, np.nan, run.monitor_max).reshape(base.ny, base.nx)
        tpeak = np.where(wall | (run.monitor_max <= 0.0), np.nan, run.monitor_tmax).reshape(base.ny, base.nx)
        return run, peak, tpeak

    def run_batch_noalloc(self, p: SolverParams, groups, on_sync=None, want_state=True):
        """itr_hip_run_batch_noalloc: the batch of ``groups`` (as for ``session``) run with
        every device buffer in the library's static scratch (no device allocation); the
        results equal run_batch's bitwise.  Raises DeviceError (out of memory) when the
        batch does not fit the scratch."""
        keep = _Keep()
        q = _params_c(p, keep)
        grids = (_abi.GridC * len(groups))()
        for i, (base, terrs, mons, guard) in enumerate(groups):
            grids[i] = _grid_c(base, terrs, mons, keep, guard, p.initial_stage is not None)
        outs, bufs = _result_buffers(groups, [(g[0].nx, g[0].ny) for g in groups], want_state)
        cb = _sync_fn(on_sync) if on_sync is not None else C.cast(None, _abi.SYNC_FN)
        err = C.create_string_buffer(1024)
        _check(self._L.itr_hip_run_batch_noalloc(self._h, C.byref(q), grids, len(groups), cb, None, outs, err, 1024),
               err)
        return _collect_results(outs, bufs)

    def run_batch(self, p: SolverParams, terrains, monitors, guard=None, on_sync=None, want_state=True):
        """GpuSolver::run_batch: every candidate (sharing one TerrainBase) to its end."""
        if len(terrains) == 0:
            return []
        s = BatchSession(self, p, [(terrains[0].base, terrains, monitors, guard)], on_sync)
        try:
            s.run()
            return s.results(want_state)
        finally:
            s.close()


def _result_buffers(groups, shapes, want_state):
    """itr_run_out entries (grids concatenated) with their host buffers."""
    n = sum(len(g[1]) for g in groups)
    outs = (_abi.RunOutC * max(n, 1))()
    bufs = []
    k = 0
    for (base, terrs, mons, _), (gnx, gny) in zip(groups, shapes):
        for _ in terrs:
            nm = max(len(mons), 1)
            b = {"mm": np.zeros(nm), "mt": np.zeros(nm), "led": np.zeros((4096, 8))}
            if want_state:
                b["h"] = np.zeros((gny, gnx), np.float32)
                b["qx"] = np.zeros((gny, gnx), np.float32)
                b["qy"] = np.zeros((gny, gnx), np.float32)
            o = outs[k]
            o.monitor_max = b["mm"].ctypes.data_as(_abi._dp)
            o.monitor_tmax = b["mt"].ctypes.data_as(_abi._dp)
            o.ledger = b["led"].ctypes.data_as(_abi._dp)
            o.max_ledger = 4096
            if want_state:
                o.depth = b["h"].ctypes.data_as(_abi._fp)
                o.qx = b["qx"].ctypes.data_as(_abi._fp)
                o.qy = b["qy"].ctypes.data_as(_abi._fp)
            bufs.append((b, len(mons)))
            k += 1
    return outs, bufs


def _collect_results(outs, bufs):
    res = []
    for i in range(len(bufs)):
        o = outs[i]
        b, nm = bufs[i]
        res.append(GpuRun(b["mm"][:nm].copy(), b["mt"][:nm].copy(), o.steps, o.t_end, o.volume_m3, b.get("h"),
                          b.get("qx"), b.get("qy"), bool(o.bad), o.cell_updates, o.rows_skipped, o.dt_min,
                          o.dt_max, o.dt_mean, o.lts, o.v0, o.rel_residual, o.aborted, o.abort_cell,
                          o.abort_sync, o.abort_value, o.sync_count, b["led"][:min(o.n_ledger, 4096)].copy(),
                          o.dt_fallback_t))
    return res


def _sync_fn(on_sync):
    """The Observer callback (itr_sync_fn) around a Python function of (grid, cand, t,
    step, sync_index, monitor_max) that returns None / False to continue or an abort value."""
    def cb(user, grid, cand, t, step, sync_index, mon_ptr, n_mon, abort_value):
        mon = np.ctypeslib.as_array(mon_ptr, shape=(n_mon,)).copy() if n_mon > 0 else np.zeros(0)
        r = on_sync(int(grid), int(cand), float(t), int(step), int(sync_index), mon)
        if r is None or r is False:
            return 0
        abort_value[0] = float(r)
        return 1
    return _abi.SYNC_FN(cb)


class BatchSession:
    """A resident batch.  ``group_params``: optional per-group SolverParams (None = ``p``),
    e.g. one group per ensemble member."""

    def __init__(self, solver: GpuSolver, p: SolverParams, groups, on_sync=None, group_params=None, grids=None,
                 keep=None, shapes=None):
        self.solver = solver
        # (nx, ny) of each group's grid (a replay region's grid is not its TerrainBase's)
        self._shapes = shapes if shapes is not None else [(g[0].nx, g[0].ny) for g in groups]
        L = solver._L
        self._L = L
        self._keep = keep if keep is not None else _Keep()
        self.groups = groups
        q = _params_c(p, self._keep)
        prepared = grids
        grids = (_abi.GridC * len(groups))()
        for i, (base, terrs, mons, guard) in enumerate(groups):
            if prepared is not None and prepared[i] is not None:
                grids[i] = prepared[i]
                continue
            gp = group_params[i] if group_params is not None else None
            eff = gp if gp is not None else p
            grids[i] = _grid_c(base, terrs, mons, self._keep, guard, eff.initial_stage is not None)
            if gp is not None:
                gq = _params_c(gp, self._keep)
                self._keep.append(gq)
                grids[i].params = C.pointer(gq)
        self._q = q
        self._grids = grids
        err = C.create_string_buffer(1024)
        b = C.c_void_p()
        _check(L.itr_hip_batch_create(solver._h, C.byref(q), grids, len(groups), C.byref(b), err, 1024), err)
        self._b = b
        self._cb = _sync_fn(on_sync) if on_sync is not None else None

    def run(self):
        err = C.create_string_buffer(1024)
        cb = self._cb if self._cb is not None else C.cast(None, _abi.SYNC_FN)
        _check(self._L.itr_hip_batch_run(self._b, cb, None, err, 1024), err)
        st = _abi.StatsC()
        self._L.itr_hip_batch_stats(self._b, C.byref(st))
        self.stats = {k: getattr(st, k) for k, _ in _abi.StatsC._fields_}
        self.solver.last_stats = self.stats
        return self.stats

    def advance(self, n_steps: int):
        """Run exactly n_steps step launches (timing a resident batch); returns stats."""
        err = C.create_string_buffer(1024)
        _check(self._L.itr_hip_batch_advance(self._b, int(n_steps), err, 1024), err)
        st = _abi.StatsC()
        self._L.itr_hip_batch_stats(self._b, C.byref(st))
        self.stats = {k: getattr(st, k) for k, _ in _abi.StatsC._fields_}
        return self.stats

    def results(self, want_state=True):
        outs, bufs = _result_buffers(self.groups, self._shapes, want_state)
        err = C.create_string_buffer(1024)
        _check(self._L.itr_hip_batch_results(self._b, outs, err, 1024), err)
        return _collect_results(outs, bufs)

    def close(self):
        if getattr(self, "_b", None):
            self._L.itr_hip_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class BatchRun:
    """evaluator.rs:118-125."""
    monitor_max: np.ndarray
    steps: int
    cell_updates: int
    bad: bool


class HipBatchBackend:
    """BatchBackend<f32> over the HIP library (backend.rs:44-64 semantics)."""

    def __init__(self, solver: GpuSolver | None = None):
        self.solver = solver or GpuSolver()

    def name(self) -> str:
        return f"hip ({self.solver.adapter_info()})"

    def run(self, p: SolverParams, terrains, monitors):
        """The reference's wgpu backend declines inflow scaling (backend.rs:53-55) and Stage /
        Inflow boundaries (lib.rs:160-171); this one runs them (ensemble members, boundary
        hydrographs).  cell_updates = steps * cells as backend.rs:56,61 count them."""
        cells = terrains[0].base.nx * terrains[0].base.ny if terrains else 0
        return [BatchRun(r.monitor_max, r.steps, r.steps * cells, r.bad)
                for r in self.solver.run_batch(p, terrains, monitors, want_state=False)]
