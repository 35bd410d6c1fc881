"""The search objective on monitor maxima (crates/itr-core/src/objective.rs:70-128):
J = J_risk + lambda * J_earth, J_risk = sum_a w_a tau softplus((peak_a - eta_a) / tau),
J_earth = cut_cost * cut_m3 + fill_cost * fill_m3.  Host-side f64, shared by the CPU and
GPU paths; the softplus uses the C library's exp / log1p (the reference uses the libm
crate, so J agrees to the last few ulps, not bitwise)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def softplus(x: float) -> float:
    """Stable log(1 + e^x), objective.rs:71-74."""
    if x > 30.0:
        return x
    if x < -30.0:
        return math.exp(x)
    return math.log1p(math.exp(x))


@dataclass
class Asset:
    positions: np.ndarray   # positions into the monitor list
    threshold_m: float
    weight: float = 1.0


@dataclass
class Objective:
    assets: list
    tau_m: float = 0.01
    earth_weight: float = 0.0
    cut_cost: float = 1.0
    fill_cost: float = 1.0

    def peaks(self, monitor_max) -> list:
        """asset_peaks (objective.rs:90-92): max over the asset's cells, from 0."""
        out = []
        for a in self.assets:
            pk = 0.0
            for k in a.positions:
                pk = max(pk, float(monitor_max[k]))
            out.append(pk)
        return out

    def j_risk(self, monitor_max) -> float:
        j = 0.0
        for a, pk in zip(self.assets, self.peaks(monitor_max)):
            j += a.weight * self.tau_m * softplus((pk - a.threshold_m) / self.tau_m)
        return j

    def j_earth(self, cut_m3: float, fill_m3: float) -> float:
        return self.cut_cost * cut_m3 + self.fill_cost * fill_m3

    def j(self, monitor_max, cut_m3: float, fill_m3: float) -> float:
        return self.j_risk(monitor_max) + self.earth_weight * self.j_earth(cut_m3, fill_m3)

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
