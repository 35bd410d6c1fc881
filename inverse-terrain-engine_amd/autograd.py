"""A torch.autograd.Function around the HIP discrete adjoint, for verification only:
J(dz) = J_risk + lambda * J_earth of one design (dz: bed change per cell, metres), whose
backward is the adjoint's dJ/d dz.  The forward and the gradient both come from one
itr_hip_adjoint call on the locked dt schedule; torch is only the autograd plumbing."""
from __future__ import annotations

import numpy as np
import torch

from .objective import Objective, volumes
from .terrain import PreparedTerrain, TerrainEdit


class PeakRiskObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dz, solver, params, base, monitors, objective: Objective):
        dz_np = dz.detach().to("cpu", torch.float64).numpy()
        area = base.dx * base.dy
        cut, fill = volumes(dz_np, area)
        cand = PreparedTerrain(base).apply_edit(TerrainEdit(0, 0, dz_np, cut, fill))
        assets = [(a.positions, a.threshold_m, a.weight) for a in objective.assets]
        res = solver.adjoint(params, base, [cand], monitors, assets, objective.tau_m)[0]
        j = res["j_risk"] + objective.earth_weight * objective.j_earth(cut, fill)
        grad = res["dJ_dz_cell"] + objective.earth_gradient(dz_np, area)
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(grad)).to(dz.device))
        return dz.new_tensor(j)

    @staticmethod
    def backward(ctx, gout):
        (g,) = ctx.saved_tensors
        return gout * g.to(gout.dtype), None, None, None, None, None


def peak_risk_objective(dz, solver, params, base, monitors, objective: Objective):
    return PeakRiskObjective.apply(dz, solver, params, base, monitors, objective)
