"""GPU parity of the time-step terms (mass matrix in the system, lifting and
history on the RHS) against the CPU oracle, and the exact discrete decay of
a separable eigenmode under BDF1/BDF2 through the device stepper."""
import math

import numpy as np
import pytest

from insilico_amd import abi, problems as P
from insilico_amd.transient import DiffusionStepper, bdf_weights

pytestmark = pytest.mark.gpu


def _rel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


CASES = {
    "poisson2d-12": lambda: P.poisson2d(12),
    "diffusion3d-6": lambda: P.diffusion3d(6, seed=2),
    "diffusion3dQ2-3": lambda: P.diffusion3d_q2(3, seed=5),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32])
def test_mass_terms_parity(oracle, name, precision):
    prob = CASES[name]()
    rng = np.random.default_rng(4)
    nd = prob.n_fnodes * prob.dpn
    prob.g = 0.1 * rng.standard_normal(nd) * (prob.dof_id < 0)
    hist = rng.standard_normal(nd)
    tr = {"density": 1.3, "mass_factor": 1.5 / 0.01, "hist": hist}
    print("[transient] case", name, precision, flush=True)
    with abi.Context(0) as ctx:
        ctx.load_problem(prob)
        ctx.set_mass(tr["density"], tr["mass_factor"], hist)
        ctx.assemble(precision, abi.ASM_MATRIX | abi.ASM_RHS)
        val, b = ctx.get_csr_values(), ctx.get_rhs()
    prob.meta["transient"] = tr
    _, _, oval, ob, _ = oracle.assemble(prob)
    tol = 1e-12 if precision == abi.FP64 else 2e-6
    assert _rel(val, oval) <= tol
    assert _rel(b, ob) <= tol


def test_bdf_weights_match_oracle(oracle):
    for q in (1, 2, 3, 4):
        for step in range(5):
            assert np.array_equal(bdf_weights(q, step), oracle.bdf_weights(q, step))


@pytest.mark.parametrize("order", [1, 2])
def test_eigenmode_decay_on_device(order):
    n, dt, steps = 16, 0.005, 5
    prob = P.poisson2d(n)
    prob.bf_kind = P.BF_NONE
    h = 1.0 / n
    lam = 2 * 6 / h ** 2 * (1 - math.cos(math.pi * h)) / (2 + math.cos(math.pi * h))
    u0 = np.sin(math.pi * prob.xyz[:, 0]) * np.sin(math.pi * prob.xyz[:, 1])
    y = [1.0]
    with abi.Context(0) as ctx:
        ctx.load_problem(prob)
        st = DiffusionStepper(ctx, prob, order=order, u0=u0)
        for k in range(steps):
            a = bdf_weights(order, k)
            u, _ = st.advance(dt, abi.SOLVE_FP64, abi.FMT_CSR, rtol=1e-13)
            y.append(-sum(a[s] * y[-s] for s in range(1, len(a))) / (a[0] + dt * lam))
            assert np.abs(u - y[-1] * u0).max() < 1e-11


def test_device_stepping_equals_host_stepping():
    """ins_bdf_init / ins_bdf_step (history kept on the device) give the same
    iterates as the host-driven stepper (ins_set_mass with uploaded
    history), step by step, for BDF2 with its start-up."""
    prob = P.diffusion3d(8, seed=7)
    rng = np.random.default_rng(3)
    u0 = 0.01 * rng.standard_normal(prob.n_nodes)
    with abi.Context(0) as c1, abi.Context(0) as c2:
        c1.load_problem(prob)
        c2.load_problem(prob)
        host = DiffusionStepper(c1, prob, order=2, u0=u0)
        c2.bdf_init(2, 1.0, u0)
        for k in range(4):
            u_h, it_h = host.advance(0.01, abi.SOLVE_FP64, abi.FMT_CSR, rtol=1e-12)
            it_d, _ = c2.bdf_step(0.01, abi.FP64, abi.SOLVE_FP64, abi.FMT_CSR, 1e-12)
            u_d = c2.get_dof_values()
            assert it_h == it_d
            assert np.abs(u_d - u_h).max() <= 1e-12 * np.abs(u_h).max()
