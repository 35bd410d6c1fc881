"""The search objective on monitor maxima (crates/itr-core/src/objective.rs:70-149):
J = J_risk + lambda * J_earth, J_risk = sum_a w_a tau softplus((peak_a - eta_a) / tau),
J_earth = cut_cost * cut_m3 + fill_cost * fill_m3.

J itself is computed by the library's one implementation (include/itr_objective.h,
through itr_hip_objective_j), which the adjoint driver and the CPU oracle's incumbent
abort also use.  Host-side f64; the library evaluates exp / log1p with the C library
where the reference uses the libm crate, so J agrees with the reference's to the last
ulps of those functions.  The guard part of Objective::evaluate (feasibility against a
baseline, objective.rs:115-129) is host logic here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi


@dataclass
class Asset:
    positions: np.ndarray   # positions into the monitor list
    threshold_m: float
    weight: float = 1.0


@dataclass
class Report:
    """ObjectiveReport (objective.rs:47-68)."""
    j: float
    j_risk: float
    j_earth: float
    peaks: list
    total_exceedance_m: float
    assets_exceeding: int
    guard_max_worsening_m: float = 0.0
    guard_cells_worsened: int = 0
    guard_violation_m: float = 0.0
    feasible: bool = True


@dataclass
class Objective:
    assets: list
    tau_m: float = 0.01
    earth_weight: float = 0.0
    cut_cost: float = 1.0
    fill_cost: float = 1.0
    guard_pos: list = field(default_factory=list)   # positions into the monitor list, guard order
    guard_tolerance_m: float = 0.0
    require_no_offsite_worsening: bool = True

    def _call(self, monitor_max, cut_m3=0.0, fill_m3=0.0):
        L = _abi.load()
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a
        start = [0]
        pos = []
        for a in self.assets:
            pos.extend(int(k) for k in a.positions)
            start.append(len(pos))
        ob = _abi.ObjectiveC()
        ob.n_assets = len(self.assets)
        ob.asset_start = arr(start, np.int32).ctypes.data_as(_abi._ip)
        ob.asset_pos = arr(pos or [0], np.int32).ctypes.data_as(_abi._ip)
        ob.threshold_m = arr([a.threshold_m for a in self.assets] or [0.0], np.float64).ctypes.data_as(_abi._dp)
        ob.weight = arr([a.weight for a in self.assets] or [0.0], np.float64).ctypes.data_as(_abi._dp)
        ob.tau_m = float(self.tau_m)
        ew = _abi.EarthworkC(float(cut_m3), float(fill_m3), float(self.cut_cost), float(self.fill_cost),
                             float(self.earth_weight))
        mm = arr(monitor_max if len(monitor_max) else [0.0], np.float64)
        j, jr = C.c_double(0.0), C.c_double(0.0)
        peaks = np.zeros(max(len(self.assets), 1))
        err = C.create_string_buffer(256)
        rc = L.itr_hip_objective_j(C.byref(ob), C.byref(ew), mm.ctypes.data_as(_abi._dp), len(monitor_max),
                                   C.byref(j), C.byref(jr), peaks.ctypes.data_as(_abi._dp), err, 256)
        if rc != _abi.ITR_OK:
            raise ValueError(err.value.decode(errors="replace"))
        return j.value, jr.value, [float(x) for x in peaks[:len(self.assets)]]

    def peaks(self, monitor_max) -> list:
        """asset_peaks (objective.rs:90-92): max over the asset's cells, from 0."""
        return self._call(monitor_max)[2]

    def j_risk(self, monitor_max) -> float:
        return self._call(monitor_max)[1]

    def j_earth(self, cut_m3: float, fill_m3: float) -> float:
        return self.cut_cost * cut_m3 + self.fill_cost * fill_m3

    def j(self, monitor_max, cut_m3: float, fill_m3: float) -> float:
        """Objective::evaluate's J; with running maxima, running_lower_bound (objective.rs:146-149)."""
        return self._call(monitor_max, cut_m3, fill_m3)[0]

    def evaluate(self, monitor_max, cut_m3: float = 0.0, fill_m3: float = 0.0, baseline_max=None) -> Report:
        """Objective::evaluate (objective.rs:95-129): J, true exceedances and, against a
        baseline's maxima, the guard worsening and feasibility."""
        j, jr, peaks = self._call(monitor_max, cut_m3, fill_m3)
        ex = [max(pk - a.threshold_m, 0.0) for a, pk in zip(self.assets, peaks)]
        rep = Report(j, jr, self.j_earth(cut_m3, fill_m3), peaks, float(sum(ex)), sum(1 for e in ex if e > 0.0))
        if baseline_max is not None:
            worst = 0.0
            for k in self.guard_pos:
                w = float(monitor_max[k]) - float(baseline_max[k])
                worst = max(worst, w)
                if w > self.guard_tolerance_m:
                    rep.guard_cells_worsened += 1
            rep.guard_max_worsening_m = worst
            if self.require_no_offsite_worsening and worst > self.guard_tolerance_m:
                rep.guard_violation_m = worst - self.guard_tolerance_m
        rep.feasible = rep.guard_violation_m == 0.0
        return rep

    def earth_gradient(self, dz: np.ndarray, cell_area: float) -> np.ndarray:
        """d(lambda * J_earth) / d dz per cell; at dz = 0 the central-difference value
        (fill and cut slopes averaged)."""
        up = self.earth_weight * cell_area * self.fill_cost
        dn = -self.earth_weight * cell_area * self.cut_cost
        return np.where(dz > 0.0, up, np.where(dz < 0.0, dn, 0.5 * (up + dn)))


def volumes(dz: np.ndarray, cell_area: float):
    """cut_m3, fill_m3 of a dense edit (design.rs:372-379)."""
    fill = float(np.sum(np.where(dz > 0.0, dz, 0.0)) * cell_area)
    cut = float(-np.sum(np.where(dz > 0.0, 0.0, dz)) * cell_area)
    return cut, fill
