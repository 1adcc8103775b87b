"""GPU parity on the branches the oracle pins of test_parity_gap_pins.py close:
forced BiCGStab restarts (host-driven and device bottom solvers), INTERPHOMOLINEAR on
levels with boxes one cell wide (operator-level ABI), and QuadCFInterp ghosts on
cubic data over hierarchies that take every stencil branch."""
import numpy as np
import pytest

import amrmg
import levelops as LO
from helpers import ghost_masks, ghosted_random, small_single
from oracle import Oracle
from test_parity_gap_pins import data_on, lshape_pow2, notch, thin_boxes

pytestmark = pytest.mark.gpu

RESTART = dict(imax=200, numRestarts=2, reps=1e-300, hang=1.0)


@pytest.mark.parametrize("device", [False, True])
def test_bicgstab_forced_restarts_bitwise(gpu_ctx, device):
    """hang = 1 forces a restart every second iteration; with two restarts allowed the
    solve ends with exit status 3 after six iterations.  The GPU bottom solve (host-driven
    BiCGStab over device operators, or the one-workgroup device solver) equals the oracle
    bitwise, and so does a fresh solve from the first restart's phi."""
    print(f"restarts device={device}", flush=True)
    h = small_single(n=40, box=40, block=8)
    g, o = amrmg.AMRMultiGrid(gpu_ctx, h), Oracle(h)
    d = o.mg_depths - 1
    res = np.random.default_rng(21).standard_normal(o.mg_cells(d))
    z = np.zeros_like(res)
    g.set_bottom_params(**RESTART)
    o.set_bottom_params(**RESTART)
    cg, sg, ig = g.bottom_solve(z, res, 2, 1.0, 1e-300, device=device)
    co, so, io = o.bottom_solve(z, res, 2, 1.0, 1e-300, iterations=True)
    assert (so, io) == (3, 6)
    assert ig == io
    if not device:
        assert sg == 3
    np.testing.assert_array_equal(cg, co)
    g.set_bottom_params(**dict(RESTART, numRestarts=0))
    o.set_bottom_params(**dict(RESTART, numRestarts=0))
    first_g = g.bottom_solve(z, res, 2, 1.0, 1e-300, device=device)[0]
    first_o = o.bottom_solve(z, res, 2, 1.0, 1e-300)[0]
    np.testing.assert_array_equal(first_g, first_o)
    g.set_bottom_params(**dict(RESTART, numRestarts=1))
    rest_g = g.bottom_solve(first_g, res, 0, 1.0, 1e-300, device=device)[0]
    np.testing.assert_array_equal(rest_g, co)
    g.close()


def test_solve_with_restarting_bottom_bitwise(gpu_ctx):
    """A whole solve whose bottom BiCGStab restarts (hang 0.5, one restart): history and
    solution bitwise the oracle's (device bottom solver in the V-cycles)."""
    h = small_single(n=40, box=40, block=8)
    import hierarchy as H
    rhs = H.sinusoid_rhs(h, 1033)
    g, o = amrmg.AMRMultiGrid(gpu_ctx, h), Oracle(h)
    bp = dict(imax=80, numRestarts=1, reps=1e-12, hang=0.5)
    g.set_bottom_params(**bp)
    o.set_bottom_params(**bp)
    sp, rp = H.solver_params()
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    assert info_g["iterations"] == info_o["iterations"] and info_g["exit_status"] == info_o["exit_status"]
    np.testing.assert_array_equal(hist_g, hist_o)
    np.testing.assert_array_equal(g.get_field(g.PHI), phi_o)
    g.close()


def _thin_level(ctx, h, boxes):
    L = LO.Level(ctx, boxes, h.level_domain(1), merge=True)
    op = LO.AMRPoissonOp(L, h.dx(1), h.dx(0), h.alpha, h.beta, ((1,) * 6, (0.0,) * 6))
    return L, op


@pytest.mark.parametrize("which", ["four_boxes", "one_box"])
def test_interphomolinear_levelops_bitwise(gpu_ctx, which):
    """A level with boxes one cell wide through the operator-level ABI: homogeneousCFInterp
    (INTERPHOMOLINEAR on the packed face, the linear branch of interpOnIVSHomo on the
    irregular one), applyOp / residual (homogeneous) and relax equal the oracle bitwise.
    one_box: a single one-cell-wide box, stored as one array (the one-launch relaxation
    does not take a linear coarse-fine face and keeps the four-launch sequence)."""
    import hierarchy as H
    h = thin_boxes()
    if which == "one_box":
        h = H.Hierarchy(h.domain, [h.levels[0], [h.levels[1][2]]], h.dx0)
    o = Oracle(h)
    L, op = _thin_level(gpu_ctx, h, h.levels[1])
    rng = np.random.default_rng(17)
    g = ghosted_random(h, 1, rng)
    valid, face, cf = ghost_masks(h, 1)
    f = LO.Field(L).upload(g, ghosted=True)
    op.homogeneousCFInterp(f)
    got = f.download(ghosted=True)
    want = o.homog_cf(1, g)
    if which == "four_boxes":      # per-box storage: every ghost cell is the box's own
        np.testing.assert_array_equal(got[cf], want[cf])
    else:                          # one box: the same array
        np.testing.assert_array_equal(got[cf], want[cf])
    n = h.cells(1)
    phi_h, rhs_h = rng.standard_normal(n), rng.standard_normal(n)
    phi, rhs, out = LO.Field(L).upload(phi_h), LO.Field(L).upload(rhs_h), LO.Field(L)
    # the oracle's level operations start from zero ghost cells; a residual with
    # inhomogeneous BCs does not refill the coarse-fine ghosts (AMRPoissonOp::residual
    # interpolates them only when homogeneous), so they are an input: zero here too
    phi_g0 = np.zeros(h.ghosted_cells(1))
    phi_g0[valid] = phi_h
    for homog in (True, False):
        phi.upload(phi_g0, ghosted=True)
        op.applyOp(out, phi, homog)
        np.testing.assert_array_equal(out.download(), o.amrlevel_apply(1, phi_h, homog))
        phi.upload(phi_g0, ghosted=True)
        op.residual(out, phi, rhs, homog)
        np.testing.assert_array_equal(out.download(), o.amrlevel_residual(1, phi_h, rhs_h, homog))
    phi.upload(phi_g0, ghosted=True)
    op.relax(phi, rhs, 3)
    np.testing.assert_array_equal(phi.download(), o.amrlevel_relax(1, phi_h, rhs_h, 3))
    for x in (f, phi, rhs, out):
        x.close()
    op.close()
    L.close()


@pytest.mark.parametrize("hname", ["lshape", "notch"])
def test_quadcf_cubic_data_bitwise(gpu_ctx, hname):
    """QuadCFInterp ghosts for cubic data on hierarchies whose stencils take every
    branch (centred, shifted +/-, dropped order, one to three mixed corners): the GPU
    equals the oracle that the reflection test pins, bit for bit."""
    h = lshape_pow2() if hname == "lshape" else notch()
    g, o = amrmg.AMRMultiGrid(gpu_ctx, h), Oracle(h)
    valid, face, cf = ghost_masks(h, 1)
    for fn in (lambda x, y, z: x ** 3 + y ** 2 * z, lambda x, y, z: y ** 3 - x * y * z,
               lambda x, y, z: z ** 3 + x ** 2 * y):
        coarse, _ = data_on(h, 0, fn)
        _, fine_g = data_on(h, 1, fn)
        start = np.where(cf, 0.0, fine_g)
        np.testing.assert_array_equal(g.quad_cf(1, start, coarse)[cf], o.quad_cf(1, start, coarse)[cf])
    g.close()
