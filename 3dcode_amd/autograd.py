"""torch.autograd view of the residual and its discrete adjoint.

``residual(solver, pv, frozen)`` is res(pv) of every block (the map
cav3d_rhs_vjp differentiates: cell properties from the EOS, eddy viscosity,
BDF history and metrics held fixed), with pv the concatenation of every
block's ghosted state as a float64 CPU tensor.  Its backward calls
cav3d_rhs_vjp, so torch's reverse mode over an objective built on the
residual runs through the library's adjoint kernels.  Torch is the harness
here; the arithmetic is libcav3d's.

Import torch before libcav3d.so is first loaded in a process (before the
first Solver, eos_host or case generator call): torch carries its own HIP
runtime, and loading it after the library leaves two HIP runtimes in the
process, under which the library's launch-residency queries fail.
"""
from __future__ import annotations

import numpy as np
import torch


def frozen_of(solver):
    """Eddy viscosity and BDF history of the solver's current state, per block."""
    return solver.frozen()


def _split(solver, a, what):
    b = solver.case.blocks
    if what == "pv":
        sizes = [solver.npv * (x.imax + 5) * (x.jmax + 5) * (x.kmax + 5) for x in b]
    else:
        sizes = [solver.npv * (x.imax - 1) * (x.jmax - 1) * (x.kmax - 1) for x in b]
    if a.size != sum(sizes):
        raise ValueError(f"{what}: expected {sum(sizes)} values, got {a.size}")
    return np.split(a, np.cumsum(sizes)[:-1])


class ResidualFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pv, solver, frozen):
        a = pv.detach().cpu().numpy().astype(np.float64, copy=True)
        res = solver.residual_of(_split(solver, a, "pv"), frozen["mut"], frozen["qq"])
        ctx.solver, ctx.frozen, ctx.pv = solver, frozen, a
        return torch.from_numpy(np.concatenate(res))

    @staticmethod
    def backward(ctx, g):
        s = ctx.solver
        # the context's state is the forward's again before the adjoint reads it
        s.residual_of(_split(s, ctx.pv, "pv"), ctx.frozen["mut"], ctx.frozen["qq"])
        gb = g.detach().cpu().numpy().astype(np.float64, copy=True)
        pb, _ = s.rhs_vjp(_split(s, gb, "res"))
        return torch.from_numpy(np.concatenate(pb)), None, None


def residual(solver, pv, frozen):
    return ResidualFunction.apply(pv, solver, frozen)
