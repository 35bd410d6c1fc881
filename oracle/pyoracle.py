"""ctypes front end of the CPU oracle (oracle/itr_oracle.cpp).

TEST INFRASTRUCTURE ONLY: this module is the checker the HIP path is compared with.
Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py may
import it; the product path (the C-ABI library and its Python host mirror) never does.

The oracle restates the reference engine's CPU solver (Rust) in C++; each C++ function
cites the Rust file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libitr_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)
_lp = C.POINTER(C.c_int64)


class Segment(C.Structure):
    _fields_ = [("range_a", C.c_double), ("range_b", C.c_double), ("stage_m", C.c_double),
                ("hydro_t", _dp), ("hydro_q", _dp), ("n_hydro", C.c_int32), ("side", C.c_int32),
                ("kind", C.c_int32), ("has_range", C.c_int32)]


class Edit(C.Structure):
    _fields_ = [("dz", _dp), ("cut_m3", C.c_double), ("fill_m3", C.c_double),
                ("c0", C.c_int32), ("r0", C.c_int32), ("w", C.c_int32), ("h", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("duration_s", C.c_double), ("sync_interval_s", C.c_double), ("cfl", C.c_double),
                ("dt_max_s", C.c_double), ("h_dry", C.c_double), ("h_eps", C.c_double),
                ("infil_capacity", C.c_double), ("initial_depth", C.c_double), ("initial_stage", C.c_double),
                ("cfl_hard", C.c_double), ("inflow_scale", C.c_double), ("rain_t", _dp), ("rain_r", _dp),
                ("locked_dt", _dp), ("n_locked", C.c_int64), ("max_steps", C.c_int64), ("n_rain", C.c_int32),
                ("has_initial_stage", C.c_int32), ("kernel", C.c_int32), ("skip_dry", C.c_int32),
                ("strip_rows", C.c_int32), ("precision", C.c_int32)]


class Start(C.Structure):
    _fields_ = [("h", _dp), ("qx", _dp), ("qy", _dp), ("t", C.c_double), ("dt_next", C.c_double),
                ("smax", C.c_double), ("v0", C.c_double), ("step", C.c_int64), ("sync_index", C.c_int32),
                ("pad_", C.c_int32)]


class Observer(C.Structure):
    _fields_ = [("baseline_max", _dp), ("guard_pos", _ip), ("asset_pos", _ip), ("asset_len", _ip),
                ("asset_thr", _dp), ("asset_w", _dp), ("guard_tol", C.c_double), ("tau", C.c_double),
                ("cut_cost", C.c_double), ("fill_cost", C.c_double), ("earth_weight", C.c_double),
                ("edit_cut_m3", C.c_double), ("edit_fill_m3", C.c_double), ("incumbent", C.c_double),
                ("n_guard", C.c_int32), ("n_assets", C.c_int32), ("require_no_offsite_worsening", C.c_int32),
                ("has_incumbent", C.c_int32)]


class Summary(C.Structure):
    _fields_ = [("t_end", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double), ("dt_mean", C.c_double),
                ("lts", C.c_double), ("v0", C.c_double), ("rel_residual", C.c_double),
                ("abs_residual", C.c_double), ("gross", C.c_double), ("abort_value", C.c_double),
                ("dt_fallback_t", C.c_double), ("steps", C.c_int64), ("cell_updates", C.c_int64),
                ("rows_skipped", C.c_int64), ("sync_count", C.c_int32), ("n_rows", C.c_int32),
                ("aborted", C.c_int32), ("abort_cell", C.c_int32), ("abort_sync", C.c_int32), ("n_dts", C.c_int32)]


def build() -> None:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_terrain_build.restype = C.c_void_p
        L.orc_terrain_build.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, _dp,
                                        _u8p, _dp, C.c_int32, C.POINTER(Segment), C.c_int32]
        L.orc_terrain_free.argtypes = [C.c_void_p]
        L.orc_terrain_info.argtypes = [C.c_void_p, _ip, _dp]
        L.orc_terrain_faces.argtypes = [C.c_void_p, C.POINTER(Edit), C.c_int32] + [_dp] * 7
        L.orc_terrain_ghosts.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip]
        L.orc_run.restype = C.c_int32
        L.orc_run.argtypes = [C.c_void_p, C.POINTER(Edit), C.POINTER(Params), _up, C.c_int32, C.POINTER(Observer),
                              C.POINTER(Start), C.POINTER(Summary), _dp, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp,
                              C.c_int64]
        _ip4 = C.POINTER(C.c_int32)
        _lp1 = C.POINTER(C.c_int64)
        L.orc_record_tape.restype = C.c_int32
        L.orc_record_tape.argtypes = [C.c_void_p, C.POINTER(Params), _ip4, _up, C.c_int32, C.POINTER(Summary), _dp,
                                      C.c_int64, _dp, C.c_int64, _lp1, _lp1]
        L.orc_run_replay.restype = C.c_int32
        L.orc_run_replay.argtypes = [C.c_void_p, C.POINTER(Edit), C.POINTER(Params), _ip4, _dp, C.c_int64, _dp,
                                     C.c_int64, C.c_double, C.c_double, _up, C.c_int32, C.POINTER(Observer),
                                     C.POINTER(Summary), _dp, C.c_int32, _dp, _dp, _dp, _dp, _dp]
        L.orc_run_candidates.restype = C.c_double
        L.orc_run_candidates.argtypes = [C.c_void_p, C.POINTER(Edit), C.c_int32, C.POINTER(Params), _up, C.c_int32,
                                         C.c_int32, _lp, _lp]
        L.orc_materialize.restype = C.c_int64
        L.orc_materialize.argtypes = [_dp, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32,
                                      C.c_int32, _u8p, C.c_double, _ip, _dp, C.c_int64, _dp]
        L.orc_profile.restype = C.c_double
        L.orc_profile.argtypes = [_dp, C.c_double, C.c_double]
        L.orc_hyeto_rate.restype = C.c_double
        L.orc_hyeto_rate.argtypes = [_dp, _dp, C.c_int32, C.c_double]
        L.orc_hyeto_integral.restype = C.c_double
        L.orc_hyeto_integral.argtypes = [_dp, _dp, C.c_int32, C.c_double, C.c_double]
        L.orc_rcbrt_f32.restype = C.c_float
        L.orc_rcbrt_f32.argtypes = [C.c_float]
        L.orc_rcbrt_f64.restype = C.c_double
        L.orc_rcbrt_f64.argtypes = [C.c_double]
        L.orc_fixedsum_f32.restype = C.c_float
        L.orc_fixedsum_f32.argtypes = [C.POINTER(C.c_float), C.c_int64, C.c_int32]
        L.orc_max_slice_f32.restype = C.c_float
        L.orc_max_slice_f32.argtypes = [C.POINTER(C.c_float), C.c_int64]
        L.orc_rmax_f32.restype = C.c_float
        L.orc_rmax_f32.argtypes = [C.c_float, C.c_float]
        L.orc_objective_j.restype = C.c_double
        L.orc_objective_j.argtypes = [C.POINTER(Observer), _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


SIDES = {"north": 0, "south": 1, "west": 2, "east": 3}
KINDS = {"transmissive": 0, "stage": 1, "inflow": 2}


class Terrain:
    """FineTerrain + TerrainBase for one fidelity level (terrain.rs:204-321)."""

    def __init__(self, dem, walls=None, manning=0.035, dx=1.0, dy=None, origin=None, level=1, segments=()):
        dem = np.ascontiguousarray(dem, dtype=np.float64)
        ny, nx = dem.shape
        dy = dx if dy is None else dy
        if origin is None:
            origin = (0.0, ny * dy)
        walls = np.zeros((ny, nx), np.uint8) if walls is None else np.ascontiguousarray(walls, dtype=np.uint8)
        man = np.broadcast_to(np.asarray(manning, dtype=np.float64), (ny, nx)).copy()
        self._keep = []
        segs = (Segment * max(1, len(segments)))()
        for i, s in enumerate(segments):
            kind = KINDS[s["kind"]]
            segs[i].side = SIDES[s["side"]]
            segs[i].kind = kind
            r = s.get("range_m")
            segs[i].has_range = 0 if r is None else 1
            if r is not None:
                segs[i].range_a, segs[i].range_b = float(r[0]), float(r[1])
            segs[i].stage_m = float(s.get("stage_m", 0.0))
            if kind == 2:
                hy = np.asarray(s["hydrograph"], dtype=np.float64)
                t = np.ascontiguousarray(hy[:, 0])
                q = np.ascontiguousarray(hy[:, 1])
                self._keep += [t, q]
                segs[i].hydro_t, segs[i].hydro_q, segs[i].n_hydro = _d(t), _d(q), len(t)
        self.h = lib().orc_terrain_build(nx, ny, origin[0], origin[1], dx, dy, _d(dem), walls.ctypes.data_as(_u8p),
                                         _d(man), level, segs, len(segments))
        info = (C.c_int32 * 4)()
        dinfo = (C.c_double * 5)()
        lib().orc_terrain_info(self.h, info, dinfo)
        self.nx, self.ny, self.crest, self.wet_count = info[0], info[1], bool(info[2]), info[3]
        self.dx, self.dy, self.z_ref, self.z_wall, self.g = list(dinfo)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_terrain_free(self.h)
        except Exception:
            pass

    def faces(self, edit=None, precision=32):
        """Prepared faces (+ wet, gn2) in the padded (ny+2, nx+2) layout, and z_level."""
        P = (self.ny + 2, self.nx + 2)
        out = {k: np.zeros(P) for k in ("fx0", "fx1", "fy0", "fy1", "wet", "gn2")}
        z = np.zeros((self.ny, self.nx))
        e = make_edit(edit)
        lib().orc_terrain_faces(self.h, C.byref(e[0]) if e else None, precision, *[_d(out[k]) for k in
                                ("fx0", "fx1", "fy0", "fy1", "wet", "gn2")], _d(z))
        out["z_level"] = z
        return out

    def ghosts(self):
        n = np.zeros(self.nx, np.int32)
        s = np.zeros(self.nx, np.int32)
        w = np.zeros(self.ny, np.int32)
        e = np.zeros(self.ny, np.int32)
        ip = lambda a: a.ctypes.data_as(_ip)
        lib().orc_terrain_ghosts(self.h, ip(n), ip(s), ip(w), ip(e))
        return {"north": n, "south": s, "west": w, "east": e}


def make_edit(edit):
    """edit: dict(c0, r0, dz (h, w) array, cut_m3, fill_m3) or None -> (Edit, keepalive)."""
    if edit is None:
        return None
    dz = np.ascontiguousarray(edit["dz"], dtype=np.float64)
    h, w = dz.shape
    e = Edit()
    e.dz = _d(dz)
    e.c0, e.r0, e.w, e.h = int(edit["c0"]), int(edit["r0"]), w, h
    e.cut_m3 = float(edit.get("cut_m3", 0.0))
    e.fill_m3 = float(edit.get("fill_m3", 0.0))
    return (e, dz)


@dataclass
class SolverParams:
    """SolverParams (solver.rs:26-49), defaults of validation::params (validation.rs:23-41)."""
    duration_s: float
    sync_interval_s: float
    cfl: float = 0.35
    dt_max_s: float = 5.0
    h_dry: float = 1e-6
    h_eps: float = 1e-5
    rain: list = field(default_factory=list)    # [(t, rate m/s)]
    infil_capacity: float = 0.0                  # m/s
    initial_depth: float = 0.0
    initial_stage: float | None = None
    kernel: str = "fused"
    skip_dry: bool = True
    locked_dt: np.ndarray | None = None
    cfl_hard: float = 0.5
    inflow_scale: float = 1.0


def hyeto_mm_h(pts):
    """Hyetograph::from_mm_h (scenario.rs:534-536)."""
    return [(float(t), float(r) / 1000.0 / 3600.0) for t, r in pts]


def _params(p: SolverParams, precision, strip_rows=0, max_steps=0):
    q = Params()
    q.duration_s, q.sync_interval_s, q.cfl, q.dt_max_s = p.duration_s, p.sync_interval_s, p.cfl, p.dt_max_s
    q.h_dry, q.h_eps, q.infil_capacity, q.initial_depth = p.h_dry, p.h_eps, p.infil_capacity, p.initial_depth
    q.has_initial_stage = 0 if p.initial_stage is None else 1
    q.initial_stage = 0.0 if p.initial_stage is None else p.initial_stage
    q.cfl_hard, q.inflow_scale = p.cfl_hard, p.inflow_scale
    rt = np.ascontiguousarray([r[0] for r in p.rain], dtype=np.float64)
    rr = np.ascontiguousarray([r[1] for r in p.rain], dtype=np.float64)
    q.rain_t, q.rain_r, q.n_rain = _d(rt), _d(rr), len(rt)
    keep = [rt, rr]
    if p.locked_dt is not None:
        ld = np.ascontiguousarray(p.locked_dt, dtype=np.float64)
        keep.append(ld)
        q.locked_dt, q.n_locked = _d(ld), len(ld)
    q.max_steps = max_steps
    q.kernel = {"oracle": 0, "fused": 1, "inertial": 2}[p.kernel]
    q.skip_dry = 1 if p.skip_dry else 0
    q.strip_rows = strip_rows
    q.precision = precision
    return q, keep


def make_observer(obs):
    """obs dict -> (Observer, keepalive).  Keys: baseline_max, guard_pos, tol, require,
    assets [(positions, thr, w)], tau, cut_cost, fill_cost, earth_weight, cut_m3, fill_m3, incumbent."""
    if obs is None:
        return None
    o = Observer()
    keep = []

    def arr(v, dt):
        a = np.ascontiguousarray(v, dtype=dt)
        keep.append(a)
        return a

    bm = arr(obs.get("baseline_max", [0.0]), np.float64)
    gp = arr(obs.get("guard_pos", [0]), np.int32)
    assets = obs.get("assets", [])
    ap = arr([p for a in assets for p in a[0]] or [0], np.int32)
    al = arr([len(a[0]) for a in assets] or [0], np.int32)
    at = arr([a[1] for a in assets] or [0.0], np.float64)
    aw = arr([a[2] for a in assets] or [0.0], np.float64)
    o.baseline_max, o.guard_pos = _d(bm), gp.ctypes.data_as(_ip)
    o.asset_pos, o.asset_len = ap.ctypes.data_as(_ip), al.ctypes.data_as(_ip)
    o.asset_thr, o.asset_w = _d(at), _d(aw)
    o.n_guard = len(obs.get("guard_pos", []))
    o.n_assets = len(assets)
    o.guard_tol = obs.get("tol", 0.0)
    o.require_no_offsite_worsening = 1 if obs.get("require", False) else 0
    o.tau = obs.get("tau", 0.01)
    o.cut_cost, o.fill_cost = obs.get("cut_cost", 1.0), obs.get("fill_cost", 1.0)
    o.earth_weight = obs.get("earth_weight", 0.0)
    o.edit_cut_m3, o.edit_fill_m3 = obs.get("cut_m3", 0.0), obs.get("fill_m3", 0.0)
    inc = obs.get("incumbent")
    o.has_incumbent = 0 if inc is None else 1
    o.incumbent = 0.0 if inc is None else float(inc)
    return o, keep


def run(terrain: Terrain, p: SolverParams, precision=32, edit=None, monitors=(), observer=None, start=None,
        strip_rows=0, max_steps=0, want_dts=False, max_rows=100000):
    """Solver::run_impl on the oracle; returns a dict with the RunSummary fields, ledger
    rows, monitor maxima and the final interior state (f64 arrays holding exact values)."""
    q, keep = _params(p, precision, strip_rows, max_steps)
    mon = np.ascontiguousarray(monitors, dtype=np.uint32)
    n_mon = len(mon)
    e = make_edit(edit)
    ob = make_observer(observer)
    st = None
    if start is not None:
        st = Start()
        hs = [np.ascontiguousarray(start[k], dtype=np.float64).ravel() for k in ("h", "qx", "qy")]
        keep += hs
        st.h, st.qx, st.qy = (_d(a) for a in hs)
        st.t, st.dt_next, st.smax, st.v0 = start["t"], start["dt_next"], start["smax"], start["v0"]
        st.step, st.sync_index = start.get("step", 0), start.get("sync_index", 0)
    s = Summary()
    rows = np.zeros((max_rows, 8))
    mmax = np.zeros(max(n_mon, 1))
    mtmax = np.zeros(max(n_mon, 1))
    ny, nx = terrain.ny, terrain.nx
    h = np.zeros((ny, nx))
    qx = np.zeros((ny, nx))
    qy = np.zeros((ny, nx))
    max_dts = 2_000_000 if want_dts else 0
    dts = np.zeros(max_dts) if want_dts else None
    rc = lib().orc_run(terrain.h, C.byref(e[0]) if e else None, C.byref(q),
                       mon.ctypes.data_as(_up) if n_mon else None, n_mon, C.byref(ob[0]) if ob else None,
                       C.byref(st) if st else None, C.byref(s), _d(rows), max_rows, _d(mmax), _d(mtmax), _d(h),
                       _d(qx), _d(qy), _d(dts), max_dts)
    if rc != 0:
        raise ValueError(f"orc_run failed ({rc})")
    out = {k: getattr(s, k) for k, _ in Summary._fields_}
    out["rows"] = rows[: s.n_rows].copy()
    out["monitor_max"] = mmax[:n_mon].copy()
    out["monitor_tmax"] = mtmax[:n_mon].copy()
    out["h"], out["qx"], out["qy"] = h, qx, qy
    if want_dts:
        out["dts"] = dts[: s.n_dts].copy()
    return out


def roi_ring_len(roi):
    c0, r0, w, h = roi
    return 2 * (w + 2) + 2 * h


def roi_band_len(roi):
    c0, r0, w, h = roi
    return sum(1 for j in range(1, h + 1) for i in range(1, w + 1) if i == 1 or j == 1 or i == w or j == h)


def record_tape(terrain: Terrain, p: SolverParams, roi, precision=32, monitors=()):
    """The baseline run recording the subdomain-replay tape (replay.rs:71-133): ring
    [steps][ring cells][3] and band [syncs][band cells][3] (h, qx, qy)."""
    q, keep = _params(p, precision)
    mon = np.ascontiguousarray(monitors, dtype=np.uint32)
    r4 = (C.c_int32 * 4)(*[int(v) for v in roi])
    s = Summary()
    nr, nb = C.c_int64(0), C.c_int64(0)
    lib().orc_record_tape(terrain.h, C.byref(q), r4, mon.ctypes.data_as(_up) if len(mon) else None, len(mon),
                          C.byref(s), None, 0, None, 0, C.byref(nr), C.byref(nb))
    ring = np.zeros(max(nr.value, 1))
    band = np.zeros(max(nb.value, 1))
    rc = lib().orc_record_tape(terrain.h, C.byref(q), r4, mon.ctypes.data_as(_up) if len(mon) else None, len(mon),
                               C.byref(s), _d(ring), nr.value, _d(band), nb.value, C.byref(nr), C.byref(nb))
    if rc != 0:
        raise ValueError(f"orc_record_tape failed ({rc})")
    out = {k: getattr(s, k) for k, _ in Summary._fields_}
    out["ring"] = ring[: nr.value].reshape(-1, roi_ring_len(roi), 3)
    out["band"] = band[: nb.value].reshape(-1, roi_band_len(roi), 3)
    out["roi"] = tuple(int(v) for v in roi)
    return out


def run_replay(terrain: Terrain, p: SolverParams, tape, precision=32, edit=None, monitors_sub=(), tol_h=2e-3,
               tol_q=1e-3, observer=None, max_rows=100000):
    """A candidate replayed on the ROI grid with the tape's ring as ghosts
    (evaluator.rs:427-440); aborted = 5 is ReplayInvalid.  State arrays are ROI-sized."""
    q, keep = _params(p, precision)
    roi = tape["roi"]
    r4 = (C.c_int32 * 4)(*roi)
    mon = np.ascontiguousarray(monitors_sub, dtype=np.uint32)
    n_mon = len(mon)
    e = make_edit(edit)
    ob = make_observer(observer)
    ring = np.ascontiguousarray(tape["ring"], dtype=np.float64)
    band = np.ascontiguousarray(tape["band"], dtype=np.float64)
    s = Summary()
    rows = np.zeros((max_rows, 8))
    mmax = np.zeros(max(n_mon, 1))
    mtmax = np.zeros(max(n_mon, 1))
    w, h_ = roi[2], roi[3]
    h = np.zeros((h_, w))
    qx = np.zeros((h_, w))
    qy = np.zeros((h_, w))
    rc = lib().orc_run_replay(terrain.h, C.byref(e[0]) if e else None, C.byref(q), r4, _d(ring), ring.shape[0],
                              _d(band), band.shape[0], tol_h, tol_q, mon.ctypes.data_as(_up) if n_mon else None, n_mon,
                              C.byref(ob[0]) if ob else None, C.byref(s), _d(rows), max_rows, _d(mmax), _d(mtmax),
                              _d(h), _d(qx), _d(qy))
    if rc != 0:
        raise ValueError(f"orc_run_replay failed ({rc})")
    out = {k: getattr(s, k) for k, _ in Summary._fields_}
    out["rows"] = rows[: s.n_rows].copy()
    out["monitor_max"] = mmax[:n_mon].copy()
    out["monitor_tmax"] = mtmax[:n_mon].copy()
    out["h"], out["qx"], out["qy"] = h, qx, qy
    return out


def run_candidates(terrain: Terrain, p: SolverParams, edits, threads, monitors=()):
    """Candidate-parallel f32 fused runs (one candidate per thread); returns (wall_s, steps, cell_updates)."""
    q, keep = _params(p, 32)
    n = len(edits)
    arr = (Edit * max(n, 1))()
    for i, ed in enumerate(edits):
        if ed is None:
            arr[i].w = 0
            arr[i].h = 0
            continue
        e, dz = make_edit(ed)
        keep.append(dz)
        arr[i] = e
    mon = np.ascontiguousarray(monitors, dtype=np.uint32)
    steps = np.zeros(n, np.int64)
    cu = np.zeros(n, np.int64)
    wall = lib().orc_run_candidates(terrain.h, arr, n, C.byref(q), mon.ctypes.data_as(_up) if len(mon) else None,
                                    len(mon), threads, steps.ctypes.data_as(_lp), cu.ctypes.data_as(_lp))
    return wall, steps, cu


def hyeto_integral(pts, a, b):
    t = np.ascontiguousarray([x[0] for x in pts], dtype=np.float64)
    r = np.ascontiguousarray([x[1] for x in pts], dtype=np.float64)
    return lib().orc_hyeto_integral(_d(t), _d(r), len(t), a, b)


def hyeto_rate(pts, x):
    t = np.ascontiguousarray([p[0] for p in pts], dtype=np.float64)
    r = np.ascontiguousarray([p[1] for p in pts], dtype=np.float64)
    return lib().orc_hyeto_rate(_d(t), _d(r), len(t), x)


PRIM_KINDS = {"berm": 0, "swale": 1, "cut": 2, "mound": 3}


def materialize(prims, origin, dx, dy, nx, ny, editable=None, max_abs_dz=1e9):
    """DesignSpace::materialize (design.rs:318-385). prims: list of dicts."""
    arr = np.zeros((max(1, len(prims)), 7))
    for i, p in enumerate(prims):
        arr[i] = [PRIM_KINDS[p["kind"]], p["cx"], p["cy"], p.get("length", 0.0), p["width"],
                  p.get("angle_deg", 0.0), p["height"]]
    ed = None if editable is None else np.ascontiguousarray(editable, dtype=np.uint8)
    cap = nx * ny
    dz = np.zeros(cap)
    bbox = (C.c_int32 * 4)()
    cf = np.zeros(2)
    n = lib().orc_materialize(_d(arr), len(prims), origin[0], origin[1], dx, dy, nx, ny,
                              ed.ctypes.data_as(_u8p) if ed is not None else None, max_abs_dz, bbox, _d(dz), cap,
                              _d(cf))
    c0, r0, w, h = list(bbox)
    return {"c0": c0, "r0": r0, "dz": dz[: w * h].reshape(h, w).copy() if n > 0 else np.zeros((0, 0)),
            "cut_m3": cf[0], "fill_m3": cf[1]}
