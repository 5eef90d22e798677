"""Time stepping of the diffusion problem on the device: the step of
heat::DiffusionDriver::advanceInTime (heat/DiffusionDriver.hpp:88-118) with a
BDF method (base/time/BDF.hpp) and its start-up (base/time/MultiStep.hpp).

Each step assembles K + (a0/dt) M and the RHS f - M sum_{s>=1} (a_s/dt)
u^{n+1-s} on the device (ins_set_mass + ins_assemble), solves, and pushes the
solution into the history.  The history stays on the host as node vectors;
the combination is uploaded once per step.
"""
import numpy as np

from . import abi

# base/time/BDF.hpp:69-99, as written there (including a_3 = +4/3 of BDF4)
_BDF = {1: (1.0, -1.0), 2: (1.5, -2.0, 0.5), 3: (11.0 / 6.0, -3.0, 1.5, -1.0 / 3.0),
        4: (25.0 / 12.0, -4.0, 3.0, 4.0 / 3.0, 1.0 / 4.0)}


def bdf_weights(order, step):
    """LHS weights a_0..a_q of BDF(order) at time step `step` (0-based): the
    first steps use the lower orders that their history allows
    (base/time/MultiStep.hpp:46-63)."""
    while order > 1 and step + 2 < order + 1:
        order -= 1
    return np.array(_BDF[order])


class DiffusionStepper:
    """advanceInTime on an abi.Context that holds a heat::Laplace problem
    (mesh, DoFs, material, loads and pattern already set)."""

    def __init__(self, ctx, prob, order=1, density=1.0, u0=None):
        self.ctx, self.prob, self.order, self.density = ctx, prob, order, density
        nd = prob.n_fnodes * prob.dpn
        u0 = np.zeros(nd) if u0 is None else np.asarray(u0, np.float64).copy()
        self.hist = [u0]        # hist[0] = u^n, hist[1] = u^{n-1}, ...
        self.step = 0

    def advance(self, dt, mode=abi.SOLVE_MIXED, fmt=abi.FMT_SELL, rtol=1e-8,
                precision=abi.FP64):
        """One time step; returns (u^{n+1}, iterations)."""
        a = bdf_weights(self.order, self.step)
        comb = np.zeros_like(self.hist[0])
        for s in range(1, len(a)):
            comb += -a[s] / dt * self.hist[min(s - 1, len(self.hist) - 1)]
        self.ctx.set_mass(self.density, a[0] / dt, comb)
        self.ctx.assemble(precision, abi.ASM_MATRIX | abi.ASM_RHS)
        it, _, _ = self.ctx.solve(mode, fmt, rtol)
        x = self.ctx.get_solution()
        u = np.array(self.prob.g, dtype=np.float64, copy=True)
        act = self.prob.dof_id >= 0
        u[act] = x[self.prob.dof_id[act]]
        self.hist.insert(0, u)
        del self.hist[self.order + 1:]
        self.step += 1
        return u, it
