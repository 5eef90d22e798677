"""Serving: batched mixed-precision solves of several requests in shared
launches.  Each request's solution is bit-identical alone, in a mix with
staggered arrivals, and through the single-request ins_solve path, and it
meets the tolerance checked by the oracle's FP64 SpMV."""
import numpy as np
import pytest

from insilico_amd import abi, problems as P

pytestmark = pytest.mark.gpu

REQS = [
    lambda: P.diffusion3d(8, seed=1),
    lambda: P.elasticity3d(4, seed=2),
    lambda: P.diffusion3d(14, seed=3),
    lambda: P.elasticity3d(7, seed=4),
    lambda: P.poisson2d(40),
    lambda: P.diffusion3d(20, seed=5),
]


def _ctx(prob):
    ctx = abi.Context(0)
    ctx.load_problem(prob)
    ctx.assemble(abi.FP64, abi.ASM_MATRIX | abi.ASM_RHS)
    return ctx


def _serve(srv, ctxs, stagger=0, k=8):
    """attach the contexts (every `stagger` steps one more), run to the end;
    returns [(x, iters, relres, status)]"""
    slots = {}
    pending = list(range(len(ctxs)))
    steps = 0
    while pending or srv.running():
        if pending and (stagger == 0 or steps % stagger == 0):
            while pending:
                i = pending.pop(0)
                slots[i] = srv.attach(ctxs[i], 1e-8)
                if stagger:
                    break
        srv.step(k)
        steps += 1
        assert steps < 100000
    srv.step(k)
    out = []
    for i in range(len(ctxs)):
        st, status, it, rr = srv.poll(slots[i])
        assert st == abi.SLOT_FINISHED
        out.append((ctxs[i].get_solution(), it, rr, status))
        srv.detach(slots[i])
    return out


def test_serving_alone_mix_and_single_identical(oracle):
    probs = [f() for f in REQS]
    ctxs = [_ctx(p) for p in probs]
    with abi.Server(0, 16) as srv:
        alone = [_serve(srv, [c])[0] for c in ctxs]
    with abi.Server(0, 16) as srv:
        mix = _serve(srv, ctxs, stagger=0)
    with abi.Server(0, 8) as srv:
        # arrivals spread over the run: requests join while others iterate
        staggered = _serve(srv, ctxs, stagger=3, k=5)
    for i, (p, c) in enumerate(zip(probs, ctxs)):
        it, rr, st = c.solve(abi.SOLVE_MIXED, abi.FMT_SELL, 1e-8)
        single = c.get_solution()
        xa, ia, ra, sa = alone[i]
        assert sa == abi.INS_OK and ra <= 1e-8
        for xm, im, rm, sm in (mix[i], staggered[i]):
            assert sm == abi.INS_OK
            assert im == ia and rm == ra
            assert np.array_equal(xm, xa), (i, np.abs(xm - xa).max())
        assert it == ia
        assert np.array_equal(single, xa), (i, np.abs(single - xa).max())
        orp, oci, oval, ob, _ = oracle.assemble(p)
        rel = np.linalg.norm(ob - oracle.spmv(orp, oci, oval, xa)) / np.linalg.norm(ob)
        assert rel <= 1e-8 * 1.001, (i, rel)
    for c in ctxs:
        c.close()


def test_serving_iteration_limit_and_cancel():
    c1 = _ctx(P.diffusion3d(12, seed=8))
    c2 = _ctx(P.diffusion3d(10, seed=9))
    with abi.Server(0, 4) as srv:
        s1 = srv.attach(c1, 1e-8, max_it=10)
        s2 = srv.attach(c2, 1e-8)
        for _ in range(20):
            srv.step(4)
            if srv.poll(s1)[0] == abi.SLOT_FINISHED:
                break
        st, status, it, rr = srv.poll(s1)
        assert st == abi.SLOT_FINISHED and status == abi.ERRORS_INV["NOCONV"]
        assert it >= 10 and 0 < rr < 1
        x1 = c1.get_solution()
        assert np.all(np.isfinite(x1))
        # cancel the other one while it runs, then the server is empty
        st2, _, _, _ = srv.poll(s2)
        assert st2 == abi.SLOT_RUNNING
        srv.detach(s2)
        srv.detach(s1)
        assert srv.running() == 0
        # slots are reusable: a fresh request on the freed slots converges
        s3 = srv.attach(c2, 1e-8)
        srv.run_until_done(8)
        st3, status3, it3, rr3 = srv.poll(s3)
        assert st3 == abi.SLOT_FINISHED and status3 == abi.INS_OK and rr3 <= 1e-8
        srv.detach(s3)
    c1.close()
    c2.close()


def test_serving_rejects_unassembled():
    c = abi.Context(0)
    c.load_problem(P.diffusion3d(4, seed=1))
    with abi.Server(0, 2) as srv:
        with pytest.raises(abi.InsError) as e:
            srv.attach(c, 1e-8)
        assert e.value.status == abi.ERRORS_INV["STATE"]
    c.close()


def test_context_reuse_for_a_larger_problem():
    """A context set up again for a larger problem (as the serving loop reuses
    its contexts) sizes and zeroes its SELL values anew: its solution equals
    that of a fresh context, bit for bit, through ins_solve and the server.
    ins_prepare(SELL) keeps the SELL form for these small, padding-heavy
    matrices (ins_solve alone would choose CSR), as the server does."""
    small, big = P.elasticity3d(4, seed=6), P.elasticity3d(9, seed=7)
    reused = _ctx(small)
    reused.prepare(abi.SOLVE_MIXED, abi.FMT_SELL)
    reused.solve(abi.SOLVE_MIXED, abi.FMT_SELL, 1e-8)
    reused.load_problem(big)
    reused.assemble(abi.FP64, abi.ASM_MATRIX | abi.ASM_RHS)
    reused.prepare(abi.SOLVE_MIXED, abi.FMT_SELL)
    it_r, rr_r, _ = reused.solve(abi.SOLVE_MIXED, abi.FMT_SELL, 1e-8)
    x_r = reused.get_solution()
    fresh = _ctx(big)
    fresh.prepare(abi.SOLVE_MIXED, abi.FMT_SELL)
    it_f, rr_f, _ = fresh.solve(abi.SOLVE_MIXED, abi.FMT_SELL, 1e-8)
    x_f = fresh.get_solution()
    assert it_r == it_f and np.array_equal(x_r, x_f)
    # and through the server, after a smaller request on the same context
    reused.load_problem(small)
    reused.assemble(abi.FP64, abi.ASM_MATRIX | abi.ASM_RHS)
    b_f, v_f = fresh.get_rhs(), fresh.get_csr_values()
    with abi.Server(0, 2) as srv:
        _serve(srv, [reused])
        reused.load_problem(big)
        reused.assemble(abi.FP64, abi.ASM_MATRIX | abi.ASM_RHS)
        b_r, v_r = reused.get_rhs(), reused.get_csr_values()
        assert np.array_equal(b_r, b_f), np.abs(b_r - b_f).max()
        assert np.array_equal(v_r, v_f), np.abs(v_r - v_f).max()
        (x_s, it_s, rr_s, st_s), = _serve(srv, [reused])
    assert st_s == abi.INS_OK and it_s == it_f, (st_s, it_s, it_f)
    assert np.array_equal(x_s, x_f), (np.abs(x_s - x_f).max(), np.abs(x_f).max(), rr_s, rr_f)
    reused.close()
    fresh.close()


def test_serving_large_batch_equals_alone():
    """Twelve requests of several hundred blocks each in one batch: the
    blocks of a request start far apart in time, and the request must still
    take the same path as alone (its replacement flag is an iteration index
    that no block of the launch that sets it can see)."""
    probs = ([P.elasticity3d(18, seed=20 + i) for i in range(6)] +
             [P.diffusion3d(40, seed=40 + i) for i in range(6)])
    ctxs = [_ctx(p) for p in probs]
    with abi.Server(0, 16) as srv:
        mix = _serve(srv, ctxs, stagger=0, k=8)
    for c, (xm, im, rm, sm) in zip(ctxs, mix):
        assert sm == abi.INS_OK and rm <= 1e-8
        c.prepare(abi.SOLVE_MIXED, abi.FMT_SELL)
        it, rr, st = c.solve(abi.SOLVE_MIXED, abi.FMT_SELL, 1e-8)
        assert it == im and np.array_equal(c.get_solution(), xm)
    for c in ctxs:
        c.close()


def test_serving_large_requests_and_node_blocks_equal_alone(oracle):
    """Requests large enough for the symmetric SpMV's own grid (diffusion
    72^3: 383,688 rows; elasticity 30^3 on node-block slices) batched with
    small ones: every request takes the grid and the partial-sum order it has
    in ins_solve, so the batch gives the same bits, and the node-block
    slices (3 x 3 blocks of the elasticity requests) give a solution the
    oracle's FP64 SpMV confirms."""
    probs = [P.diffusion3d(72, seed=70), P.elasticity3d(30, seed=71), P.diffusion3d(10, seed=72),
             P.elasticity3d(6, seed=73)]
    ctxs = [_ctx(p) for p in probs]
    for c, p in zip(ctxs, probs):
        c.prepare(abi.SOLVE_MIXED, abi.FMT_SELL)
        st = c.stats()
        if p.n_active > 100000:   # small meshes pad too much and keep the full slices
            assert st["sym_block"] == p.dpn and st["sym_values"] > 0
    with abi.Server(0, 8) as srv:
        mix = _serve(srv, ctxs, stagger=2, k=8)
    for c, p, (xm, im, rm, sm) in zip(ctxs, probs, mix):
        assert sm == abi.INS_OK and rm <= 1e-8
        it, rr, st = c.solve(abi.SOLVE_MIXED, abi.FMT_SELL, 1e-8)
        assert st == abi.INS_OK and it == im and np.array_equal(c.get_solution(), xm)
        rp, ci = c.get_pattern()
        val, b = c.get_csr_values(), c.get_rhs()
        r = b - oracle.spmv(rp, ci, val, xm)
        assert np.linalg.norm(r) <= 1.001e-8 * np.linalg.norm(b)
    for c in ctxs:
        c.close()
