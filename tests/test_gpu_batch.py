"""Batched single-level solves (amrmg_batch_*) against solving each problem alone.

In FP64 every stage of the batched V-cycle evaluates, per problem, the same
expressions on the same operands as the single-problem path (the relaxation, the
residual-restriction and prolongation with their boundary ghosts, the bottom
BiCGStab with the reference's reduction orders), so each problem's history and
solution must be bitwise those of AMRMultiGrid on it alone, which the parity
tests hold bitwise equal to the oracle.  The mixed-precision batch must land
within 1e-9 relative of the FP64 solution (ten times the solve tolerance).
"""
import numpy as np
import pytest

import amrmg
import hierarchy as H
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _problem(n, alpha, beta, bc=(1, 1, 1), seed=0, k=(1, 2, 1)):
    dom = H.box((0, 0, 0), (n - 1, n - 1, n - 1))
    h = H.Hierarchy(dom, [H.domain_split(dom, 64, 8)], 1.0 / n, alpha, beta, bc_type_lo=bc, bc_type_hi=(1, 1, 1))
    return h, H.sinusoid_rhs(h, 2000 + seed, k=k)


def _edge_set():
    """Power-of-two single boxes, non-power-of-two boxes (deeper bottoms), a
    64 + 8 split (bottom of 8 boxes), Laplacian and Helmholtz, Neumann faces."""
    specs = [(16, 0.0, 1.0), (24, 0.3, 1.2), (32, 0.0, 1.0), (40, 0.7, 0.6), (72, 0.5, 1.5), (64, 0.0, 1.0),
             (48, 0.9, 1.9), (8, 0.2, 0.7)]
    out = [_problem(n, a, b, seed=q) for q, (n, a, b) in enumerate(specs)]
    out.append(_problem(32, 0.6, 1.1, bc=(0, 1, 0), seed=20, k=(1, 1, 2)))
    return out


def _single(gpu_ctx, h, rhs, sp, rp):
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist, info = g.solve()
    phi = g.get_field(g.PHI)
    g.close()
    return hist, info, phi


def _batch(gpu_ctx, probs, sp, rp, precision=0, tune=False, order=0):
    b = amrmg.BatchSolver(gpu_ctx, [h for h, _ in probs], precision)
    b.setSolverParameters(sp, rp)
    if order:
        b.set_reduction_order(order)
    for q, (_, rhs) in enumerate(probs):
        b.set_rhs(q, rhs)
    if tune:
        b.tune_relax()
    it, ex = b.solve()
    out = [(b.history(q), int(it[q]), int(ex[q]), b.get_phi(q)) for q in range(len(probs))]
    b.close()
    return out


def test_batch_fp64_bitwise_equals_single_solves(gpu_ctx):
    probs = _edge_set()
    sp, rp = H.solver_params()
    res = _batch(gpu_ctx, probs, sp, rp)
    for q, ((h, rhs), (hist, it, ex, phi)) in enumerate(zip(probs, res)):
        hs, info, phis = _single(gpu_ctx, h, rhs, sp, rp)
        assert it == info["iterations"] and ex == info["exit_status"], q
        np.testing.assert_array_equal(hist, hs)
        np.testing.assert_array_equal(phi, phis)
    # and against the oracle directly on every one of them
    for q, (h, rhs) in enumerate(probs):
        phi_o, hist_o, info_o = Oracle(h).solve(rhs, sp, rp)
        assert res[q][1] == info_o["iterations"] and res[q][2] == info_o["exit_status"], q
        np.testing.assert_array_equal(res[q][0], hist_o)
        np.testing.assert_array_equal(res[q][3], phi_o)


def test_batch_tuned_shapes_bitwise(gpu_ctx):
    probs = _edge_set()[:6]
    sp, rp = H.solver_params()
    a = _batch(gpu_ctx, probs, sp, rp)
    b = _batch(gpu_ctx, probs, sp, rp, tune=True)
    for (h1, i1, e1, p1), (h2, i2, e2, p2) in zip(a, b):
        assert (i1, e1) == (i2, e2)
        np.testing.assert_array_equal(h1, h2)
        np.testing.assert_array_equal(p1, p2)


def test_batch_mixed_precision_within_tolerance(gpu_ctx):
    probs = _edge_set()
    sp, rp = H.solver_params()
    ref = _batch(gpu_ctx, probs, sp, rp)
    mp = _batch(gpu_ctx, probs, sp, rp, precision=1)
    for q, ((h1, i1, e1, p1), (h2, i2, e2, p2)) in enumerate(zip(ref, mp)):
        assert e2 & 1 and h2[-1] <= 1e-10 * h2[0], q
        assert abs(i1 - i2) <= 2, q
        err = np.abs(p2 - p1).max() / np.abs(p1).max()
        assert err < 1e-9, (q, err)


def test_batch_c4_sample_bitwise(gpu_ctx):
    """48 problems of the 512-problem workload's distribution (seed 1004)."""
    probs = H.batch_problems(48, seed=1004)
    sp, rp = H.solver_params()
    res = _batch(gpu_ctx, probs, sp, rp)
    for q in range(0, 48, 3):
        h, rhs = probs[q]
        hs, info, phis = _single(gpu_ctx, h, rhs, sp, rp)
        assert res[q][1] == info["iterations"], q
        np.testing.assert_array_equal(res[q][3], phis)
        phi_o, hist_o, info_o = Oracle(h).solve(rhs, sp, rp)
        assert res[q][1] == info_o["iterations"] and res[q][2] == info_o["exit_status"], q
        np.testing.assert_array_equal(res[q][0], hist_o)
        np.testing.assert_array_equal(res[q][3], phi_o)


def test_batch_graph_replays_bitwise_equal_direct_launches(gpu_ctx):
    """V-cycles replayed from the captured HIP graph give bitwise the history and
    solution of direct launches, over two solves in a row (every V-cycle after the
    capture is a replay, with the eager update, residual and norm read-back between
    replays), in FP64 and in mixed precision; the replay count shows the graph ran."""
    probs = _edge_set()[:7]
    sp, rp = H.solver_params()
    gpu_ctx.profile([])  # the graph is bypassed while the profiler is on
    for precision in (0, 1):
        b = amrmg.BatchSolver(gpu_ctx, [h for h, _ in probs], precision)
        b.setSolverParameters(sp, rp)
        for q, (_, rhs) in enumerate(probs):
            b.set_rhs(q, rhs)
        b.set_graph(False)
        it0, ex0 = b.solve()
        direct = [(b.history(q), b.get_phi(q)) for q in range(len(probs))]
        assert b.set_graph(None) == 0
        b.set_graph(True)
        for rep in range(2):
            it, ex = b.solve()
            np.testing.assert_array_equal(it, it0)
            np.testing.assert_array_equal(ex, ex0)
            for q in range(len(probs)):
                np.testing.assert_array_equal(b.history(q), direct[q][0], err_msg=f"prec {precision} rep {rep} q {q}")
                np.testing.assert_array_equal(b.get_phi(q), direct[q][1], err_msg=f"prec {precision} rep {rep} q {q}")
        assert b.set_graph(None) >= 2 * int(np.max(it0)) - 2
        b.close()


def test_batch_two_lanes_bitwise_equal_one_lane(gpu_ctx):
    """lanes=2 puts the problems with bottoms of at least 16^3 cells (edge 72 here:
    an 18^3 bottom) in a lockstep of their own on a second stream, run concurrently
    by a second host thread; every problem's history and solution are bitwise those
    of the single lane, in FP64 and mixed precision, concurrent (profiler off) and
    one lane after the other (profiler on)."""
    probs = _edge_set()
    sp, rp = H.solver_params()
    for precision in (0, 1):
        ref = _batch(gpu_ctx, probs, sp, rp, precision=precision)
        b = amrmg.BatchSolver(gpu_ctx, [h for h, _ in probs], precision, lanes=2)
        nl, per = b.lanes()
        assert nl == 2 and per[4] == 1 and int(per.sum()) == 1, per  # only the edge-72 problem
        b.setSolverParameters(sp, rp)
        for q, (_, rhs) in enumerate(probs):
            b.set_rhs(q, rhs)
        for profiled in (False, True, False):
            gpu_ctx.profile(["gsrb"] if profiled else [])
            it, ex = b.solve()
            for q in range(len(probs)):
                assert it[q] == ref[q][1] and ex[q] == ref[q][2], (precision, profiled, q)
                np.testing.assert_array_equal(b.history(q), ref[q][0])
                np.testing.assert_array_equal(b.get_phi(q), ref[q][3])
        gpu_ctx.profile([])
        b.close()


def test_batch_three_lanes_bitwise_equal_one_lane(gpu_ctx):
    """lanes=3: edge 120 (64 + 56 boxes, a 30^3 bottom) in the third lane, edge 72
    (an 18^3 bottom) in the second, the rest in the first; bitwise equal to one lane."""
    probs = [_problem(120, 0.4, 1.3, seed=31), _problem(72, 0.5, 1.5, seed=32), _problem(24, 0.3, 1.2, seed=33),
             _problem(16, 0.0, 1.0, seed=34)]
    sp, rp = H.solver_params()
    gpu_ctx.profile([])
    ref = _batch(gpu_ctx, probs, sp, rp)
    b = amrmg.BatchSolver(gpu_ctx, [h for h, _ in probs], 0, lanes=3)
    nl, per = b.lanes()
    assert nl == 3 and list(per) == [2, 1, 0, 0], per
    b.setSolverParameters(sp, rp)
    for q, (_, rhs) in enumerate(probs):
        b.set_rhs(q, rhs)
    it, ex = b.solve()
    for q in range(len(probs)):
        assert it[q] == ref[q][1] and ex[q] == ref[q][2], q
        np.testing.assert_array_equal(b.history(q), ref[q][0])
        np.testing.assert_array_equal(b.get_phi(q), ref[q][3])
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_tree_order_same_alone_and_in_batch_and_near_oracle(gpu_ctx, precision):
    """With the tree reduction order in the bottom solves, a problem's result does not
    depend on what else is in the batch (bitwise equal solved in a batch of one), and
    every converged solution is within 1e-10 of the oracle's FP64 solve (relative,
    max norm), with the same iteration count or one more or fewer."""
    probs = _edge_set()
    sp, rp = H.solver_params()
    mix = _batch(gpu_ctx, probs, sp, rp, precision=precision, order=1)
    for q, (h, rhs) in enumerate(probs):
        hist, it, ex, phi = mix[q]
        assert ex & 1, q
        hist1, it1, ex1, phi1 = _batch(gpu_ctx, [(h, rhs)], sp, rp, precision=precision, order=1)[0]
        assert it1 == it and ex1 == ex, q
        np.testing.assert_array_equal(hist1, hist)
        np.testing.assert_array_equal(phi1, phi)
        o = Oracle(h)
        phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
        o.close()
        err = float(np.abs(phi - phi_o).max() / np.abs(phi_o).max())
        assert err <= 1e-10, (q, err)
        assert abs(it - info_o["iterations"]) <= 1, q


def test_batch_reduction_order_argument_checked(gpu_ctx):
    b = amrmg.BatchSolver(gpu_ctx, [_problem(16, 0.0, 1.0)[0]], 0)
    with pytest.raises(amrmg.AMRMGError):
        b.set_reduction_order(2)
    b.close()


def test_batch_quad_tiles_bitwise_equal_pair_tiles(gpu_ctx):
    """Mixed-precision batches: every x-quad tile shape on every depth where it
    applies gives the x-pair march's results bit for bit; FP64 batches and depths
    with an x extent that is not a multiple of 4 reject them."""
    probs = [p for p in _edge_set() if p[0].domain[1][0] - p[0].domain[0][0] + 1 in (16, 32, 48, 64)]
    sp, rp = H.solver_params()
    quad = [s for s, t in enumerate(amrmg.AMRMultiGrid.RELAX_SHAPES) if t[3] == 2]

    def run(shape, kc):
        b = amrmg.BatchSolver(gpu_ctx, [h for h, _ in probs], 1)
        b.setSolverParameters(sp, rp)
        for q, (_, rhs) in enumerate(probs):
            b.set_rhs(q, rhs)
        nd = len(b.tune_relax())
        for d in range(nd):
            b.set_relax_shape(d, 15, 0)
            if shape is not None:
                try:
                    b.set_relax_shape(d, shape, kc)
                except amrmg.AMRMGError:
                    pass   # a depth with an x extent that is not a multiple of 4
        it, ex = b.solve()
        out = [(b.history(q), int(it[q]), int(ex[q]), b.get_phi(q)) for q in range(len(probs))]
        b.close()
        return out

    ref = run(None, 0)
    for shape in quad:
        for kc in (0, 16):
            print(f"batch quad shape {shape} kc {kc}", flush=True)
            got = run(shape, kc)
            for (h1, i1, e1, p1), (h2, i2, e2, p2) in zip(ref, got):
                assert (i1, e1) == (i2, e2), shape
                np.testing.assert_array_equal(h1, h2)
                np.testing.assert_array_equal(p1, p2)
    b = amrmg.BatchSolver(gpu_ctx, [h for h, _ in probs], 0)
    with pytest.raises(amrmg.AMRMGError, match="x-quad"):
        b.set_relax_shape(0, quad[0], 0)
    b.close()
