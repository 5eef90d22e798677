"""Error behaviour and determinism of the C ABI on the GPU: every failure is
a status code with a message (the reference aborts through VERIFY_MSG,
base/verify.hpp:63-65), never a crash; repeated work gives the same bits."""
import numpy as np
import pytest

from insilico_amd import abi, problems as P

pytestmark = pytest.mark.gpu

ERR = abi.ERRORS_INV


def _status(fn):
    with pytest.raises(abi.InsError) as e:
        fn()
    return e.value.status


def test_calls_out_of_order():
    prob = P.diffusion3d(4, seed=1)
    with abi.Context(0) as c:
        assert _status(lambda: c.set_dofs(1, prob.dof_id, prob.n_active)) == ERR["STATE"]
        c.set_mesh(prob.xyz, prob.conn)
        c.set_dofs(1, prob.dof_id, prob.n_active)
        c.set_material(prob.kernel, prob.nq, prob.coef)
        c.nnz = 0
        assert _status(lambda: c.assemble(abi.FP64)) == ERR["STATE"]      # no pattern
        c.build_pattern()
        assert _status(lambda: c.solve(abi.SOLVE_FP64, abi.FMT_CSR, 1e-8)) == ERR["STATE"]
        assert _status(lambda: c.get_solution()) == ERR["STATE"]
        assert _status(lambda: c.distribute(abi.DIST_SET)) == ERR["STATE"]


def test_bad_ids_and_arguments():
    prob = P.diffusion3d(4, seed=1)
    with abi.Context(0) as c:
        c.set_mesh(prob.xyz, prob.conn)
        bad = prob.dof_id.copy()
        k = np.nonzero(prob.dof_id >= 0)[0][0]        # out of range
        bad[k] = prob.n_active
        assert _status(lambda: c.set_dofs(1, bad, prob.n_active)) == ERR["RANGE"]
        dup = prob.dof_id.copy()
        a, b = np.nonzero(prob.dof_id >= 0)[0][:2]
        dup[b] = dup[a]                               # used twice
        assert _status(lambda: c.set_dofs(1, dup, prob.n_active)) == ERR["RANGE"]
        conn = prob.conn.copy()
        conn[0, 0] = prob.n_nodes                     # node out of range
        assert _status(lambda: c.set_mesh(prob.xyz, conn)) == ERR["RANGE"]
        assert _status(lambda: c.set_mesh(prob.xyz, prob.conn[:, :4])) == ERR["UNSUPPORTED"]
        c.set_mesh(prob.xyz, prob.conn)
        c.set_dofs(1, prob.dof_id, prob.n_active)
        c.build_pattern()
        c.set_material(prob.kernel, prob.nq, prob.coef)
        c.assemble(abi.FP64)
        assert _status(lambda: c.assemble(7)) == ERR["ARG"]
        assert _status(lambda: c.solve(5, abi.FMT_CSR, 1e-8)) == ERR["ARG"]
        assert _status(lambda: c.solve(abi.SOLVE_FP64, abi.FMT_CSR, -1.0)) == ERR["ARG"]


def test_iteration_limit_and_breakdown():
    prob = P.diffusion3d(10, seed=2)
    with abi.Context(0) as c:
        c.load_problem(prob)
        c.assemble(abi.FP64)
        for mode in (abi.SOLVE_FP64, abi.SOLVE_MIXED):
            it, rel, st = c.solve(mode, abi.FMT_SELL, 1e-8, max_it=5, check=False)
            assert st == ERR["NOCONV"] and 0 < rel < 100  # CG residuals need not fall monotonically
            assert np.all(np.isfinite(c.get_solution()))   # last iterate available
        # an indefinite matrix: p.Ap <= 0 is a breakdown, reported, not a crash
        rp, ci = c.get_pattern()
        val = c.get_csr_values()
        diag = np.repeat(np.arange(rp.shape[0] - 1), np.diff(rp)) == ci
        neg = val.copy()
        neg[diag] *= -1.0
        c.set_csr(rp, ci, neg)
        c.set_rhs(np.ones(prob.n_active))
        it, rel, st = c.solve(abi.SOLVE_FP64, abi.FMT_CSR, 1e-8, check=False)
        assert st == ERR["BREAKDOWN"]


def test_repeat_solves_are_bitwise_equal():
    prob = P.elasticity3d(6, seed=3)
    with abi.Context(0) as c:
        c.load_problem(prob)
        out = []
        for _ in range(2):
            c.assemble(abi.FP64)
            v, b = c.get_csr_values(), c.get_rhs()
            res = []
            for mode, fmt in ((abi.SOLVE_FP64, abi.FMT_CSR), (abi.SOLVE_MIXED, abi.FMT_SELL)):
                it, rel, _ = c.solve(mode, fmt, 1e-8)
                res.append((it, rel, c.get_solution()))
            out.append((v, b, res))
        (v0, b0, r0), (v1, b1, r1) = out
        assert np.array_equal(v0, v1) and np.array_equal(b0, b1)
        for (i0, e0, x0), (i1, e1, x1) in zip(r0, r1):
            assert i0 == i1 and e0 == e1 and np.array_equal(x0, x1)


def test_non_monotone_numbering(oracle):
    """DoF ids in a shuffled order: the pattern is that of the permuted
    system (rows sorted), values and RHS match the oracle's."""
    prob = P.diffusion3d(6, seed=4)
    rng = np.random.default_rng(0)
    perm = rng.permutation(prob.n_active).astype(np.int32)
    act = prob.dof_id >= 0
    prob.dof_id = prob.dof_id.copy()
    prob.dof_id[act] = perm[prob.dof_id[act]]
    with abi.Context(0) as c:
        c.load_problem(prob)
        c.assemble(abi.FP64)
        rp, ci = c.get_pattern()
        orp, oci, oval, ob, _ = oracle.assemble(prob)
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
        assert np.max(np.abs(c.get_csr_values() - oval)) <= 1e-12 * np.max(np.abs(oval))
        assert np.max(np.abs(c.get_rhs() - ob)) <= 1e-12 * np.max(np.abs(ob))
        it, rel, _ = c.solve(abi.SOLVE_MIXED, abi.FMT_SELL, 1e-8)
        x = c.get_solution()
        r = ob - oracle.spmv(orp, oci, oval, x)
        assert np.linalg.norm(r) / np.linalg.norm(ob) <= 1e-8 * 1.001


def test_sell_and_csr_agree():
    prob = P.elasticity3d(8, seed=5)
    with abi.Context(0) as c:
        c.load_problem(prob)
        c.assemble(abi.FP64)
        xs = []
        for fmt in (abi.FMT_CSR, abi.FMT_SELL):
            c.solve(abi.SOLVE_FP64, fmt, 1e-10)
            xs.append(c.get_solution())
        assert np.linalg.norm(xs[0] - xs[1]) <= 1e-7 * np.linalg.norm(xs[0])


def test_new_entry_points_reject_bad_use():
    """Degree-2 fields, time-step terms, BDF stepping and error norms report
    misuse as status codes."""
    prob = P.diffusion3d_q2(2, seed=1)
    q1 = P.diffusion3d(3, seed=1)
    with abi.Context(0) as c:
        # field before mesh, wrong node count, out-of-range node, twice
        assert _status(lambda: c.set_field(2, prob.n_fnodes, prob.fconn)) == ERR["STATE"]
        c.set_mesh(prob.xyz, prob.conn)
        assert _status(lambda: c.set_field(3, prob.n_fnodes, prob.fconn)) == ERR["UNSUPPORTED"]
        assert _status(lambda: c.set_field(2, prob.n_fnodes, prob.fconn[:, :9])) == ERR["ARG"]
        bad = prob.fconn.copy()
        bad[0, 0] = prob.n_fnodes
        assert _status(lambda: c.set_field(2, prob.n_fnodes, bad)) == ERR["RANGE"]
        c.set_field(2, prob.n_fnodes, prob.fconn)
        assert _status(lambda: c.set_field(2, prob.n_fnodes, prob.fconn)) == ERR["STATE"]
        # Neumann faces on a degree-2 field: 9 field nodes, vertices first
        faces = np.array([[0, 1, 2, 3]], np.int32)
        assert _status(lambda: c.set_traction(faces, (0.0, 0.0, 1.0))) == ERR["ARG"]
        cell = prob.fconn[0, 26]
        faces9 = np.array([[cell] * 9], np.int32)
        assert _status(lambda: c.set_traction(faces9, (0.0, 0.0, 1.0))) == ERR["ARG"]
        assert _status(lambda: c.set_traction_profile(3, 0.0)) == ERR["ARG"]
        # mass terms and stepping before the DoFs
        assert _status(lambda: c.set_mass(1.0, 10.0)) == ERR["STATE"]
        assert _status(lambda: c.bdf_init(2, 1.0)) == ERR["STATE"]
        c.set_dofs(1, prob.dof_id, prob.n_active)
        assert _status(lambda: c.set_mass(-1.0, 1.0)) == ERR["ARG"]
        assert _status(lambda: c.bdf_step(0.1)) == ERR["STATE"]            # no init
        assert _status(lambda: c.bdf_init(5, 1.0)) == ERR["ARG"]
        assert _status(lambda: c.bdf_init(2, 0.0)) == ERR["ARG"]
        # error norm before a material, bad order / reference
        assert _status(lambda: c.error_norm(0, abi.REF_ZERO)) == ERR["STATE"]
        c.set_material(prob.kernel, prob.nq, prob.coef)
        assert _status(lambda: c.error_norm(2, abi.REF_ZERO)) == ERR["ARG"]
        assert _status(lambda: c.error_norm(0, 7)) == ERR["ARG"]
        assert _status(lambda: c.error_norm(1, abi.REF_KELVIN, [1, 1, 0, 0, 0, 1, 0, 0])) == ERR["UNSUPPORTED"]
        c.build_pattern()
        c.bdf_init(2, 1.0)
        assert _status(lambda: c.bdf_step(-0.1)) == ERR["ARG"]
    with abi.Context(0) as c:
        # mass terms are for heat::Laplace only
        e = P.elasticity3d(2, seed=1)
        c.load_problem(e)
        c.set_mass(1.0, 10.0)
        assert _status(lambda: c.assemble(abi.FP64)) == ERR["UNSUPPORTED"]
        # a Q1 context after a Q2 one on the same handle works again
        c.load_problem(q1)
        c.assemble(abi.FP64)


def test_host_csr_of_another_size_blocks_the_dof_paths(oracle):
    """ins_set_csr installs a system that need not have the DoFs' size: the
    element assembly (and with it the RHS and time-step kernels) and
    ins_distribute then refuse with INS_ERR_STATE before any launch, the
    uploaded system still solves, a vector problem's host CSR keeps the
    identity row order of its slices, and ins_build_pattern brings the DoF
    pattern back, after which the assembly equals the oracle's again."""
    prob = P.elasticity3d(4, seed=5)
    small = P.diffusion3d(3, seed=2)
    srp, sci, sval, sb, _ = oracle.assemble(small)
    with abi.Context(0) as c:
        c.load_problem(prob)
        assert small.n_active < prob.n_active
        c.set_csr(srp, sci, sval)
        assert _status(lambda: c.assemble(abi.FP64)) == ERR["STATE"]
        assert _status(lambda: c.solve(abi.SOLVE_FP64, abi.FMT_CSR, 1e-8)) == ERR["STATE"]  # no rhs
        c.set_rhs(sb)
        for mode, fmt in ((abi.SOLVE_FP64, abi.FMT_CSR), (abi.SOLVE_MIXED, abi.FMT_SELL)):
            _, rel, st = c.solve(mode, fmt, 1e-8)
            assert st == abi.INS_OK and rel <= 1e-8
            x = c.get_solution()
            r = sb - oracle.spmv(srp, sci, sval, x)
            assert np.linalg.norm(r) <= 1.001e-8 * np.linalg.norm(sb)
        assert _status(lambda: c.distribute(abi.DIST_SET)) == ERR["STATE"]
        c.build_pattern()
        c.assemble(abi.FP64)
        rp, ci = c.get_pattern()
        val, b = c.get_csr_values(), c.get_rhs()
    orp, oci, oval, ob, _ = oracle.assemble(prob)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    assert np.max(np.abs(val - oval)) <= 1e-12 * np.max(np.abs(oval))
    assert np.max(np.abs(b - ob)) <= 1e-12 * np.max(np.abs(ob))
