"""torch.autograd front end over the C ABI (optional; torch is plumbing here).

`ElecEnergy.apply(xyz, pole_t, polarity_t, thole_t, model)` returns the
total electrostatic energy em + ep (kcal/mol) of a system whose per-atom
parameters come from per-type tables: pole_t (ntype, 13) local multipoles,
polarity_t (ntype,) polarizabilities (pdamp = polarity^(1/6) for every atom,
as the parameter files tie them), thole_t (ntype,) Thole parameters.  The
backward pass returns dE/dxyz (the analytic gradient) and the per-type sums
of the library's parameter gradients; no second solve is run.

`model` is an `ElecModel`: an Engine, the system dict and the atom -> type
index map.  All tensors are float64 on the CPU; the evaluation itself runs
on the GPU through libpmq.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .engine import Engine


class ElecModel:
    def __init__(self, system: dict, type_index, engine: Engine | None = None, terms: int = L.TERM_MPOLE | L.TERM_POLAR):
        self.system = dict(system)
        self.tidx = np.asarray(type_index, dtype=np.int64)
        self.engine = engine if engine is not None else Engine(0)
        self.terms = terms

    def evaluate(self, xyz, pole_t, polarity_t, thole_t):
        s = self.system
        s["x"], s["y"], s["z"] = (np.ascontiguousarray(xyz[:, a]) for a in range(3))
        s["pole"] = np.ascontiguousarray(pole_t[self.tidx]).ravel()
        alpha = polarity_t[self.tidx]
        s["polarity"] = alpha
        s["pdamp"] = alpha ** (1.0 / 6.0)
        s["thole"] = thole_t[self.tidx]
        eng = self.engine
        eng.set_system(s)
        out = eng.eval(L.GRADIENT, self.terms, L.PLAIN_FP64)
        pg = eng.param_grad()
        return out, pg, alpha


class ElecEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, pole_t, polarity_t, thole_t, model: ElecModel):
        out, pg, alpha = model.evaluate(xyz.detach().cpu().numpy().astype(np.float64),
                                        pole_t.detach().cpu().numpy().astype(np.float64),
                                        polarity_t.detach().cpu().numpy().astype(np.float64),
                                        thole_t.detach().cpu().numpy().astype(np.float64))
        nt = pole_t.shape[0]
        tidx = model.tidx
        dpole = np.zeros((nt, 13))
        np.add.at(dpole, tidx, pg["dpole"])
        dal_atom = pg["dpolarity"] + pg["dpdamp"] * np.where(alpha > 0, alpha ** (-5.0 / 6.0) / 6.0, 0.0)
        dal = np.bincount(tidx, weights=dal_atom, minlength=nt)
        dth = np.bincount(tidx, weights=pg["dthole"], minlength=nt)
        ctx.save_for_backward(torch.from_numpy(out["desum"].copy()), torch.from_numpy(dpole),
                              torch.from_numpy(dal), torch.from_numpy(dth))
        return torch.tensor(out["em"] + out["ep"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        dx, dpole, dal, dth = ctx.saved_tensors
        g = grad_out.to(torch.float64)
        return g * dx, g * dpole, g * dal, g * dth, None
