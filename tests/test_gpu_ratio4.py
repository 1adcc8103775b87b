"""Refinement ratio 4 (AMRPoissonOpFactory's refRatios) on the HIP path against the
oracle: QuadCFInterp, homogeneous CF interpolation and AVERAGE at ratio-4 levels,
composite residual and operator (reflux with nRef = 4), and whole solves bitwise in
FP64, on two-level ratio-4 and three-level mixed-ratio hierarchies, merged and per
box; the reference's parabola test (testAMRPoissonOp.H:50-130) solved to the sampled
parabola; mixed precision within 1e-9; the ratios the library rejects."""
import numpy as np
import pytest

import amrmg
import hierarchy as H
from helpers import ghost_masks, ghosted_random, uncovered_mask
from oracle import Oracle

pytestmark = pytest.mark.gpu


def parabola(x, y, z, l=0):
    return x * x + y * y + z * z


def diri(pos, d, s):
    return pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]


def two_level(box=16):
    """[0, 31]^3 base (dx 1/32); level 1 (ratio 4) over coarse [8, 23] x [8, 23] x
    [12, 19]: one rectangle (box = 64) or 16^3 boxes."""
    dom = H.box((0, 0, 0), (31, 31, 31))
    fine = H.box_refine(H.box((8, 8, 12), (23, 23, 19)), 4)
    return H.Hierarchy(dom, [H.domain_split(dom, 16, 8), H.domain_split(fine, box, 4)], 1.0 / 32, ratios=(4,))


def three_level(ratios, bc=(1, 1, 1), alpha=0.0, beta=1.0):
    """[-8, 7]^3 base (dx 1/16); level 1 over coarse [-4, 7] x [-4, 3]^2 and level 2
    over the last 4 coarse cells of level 1 in x, both reaching the high x face."""
    r1, r2 = ratios
    dom = H.box((-8,) * 3, (7,) * 3)
    lev1 = H.box_refine(H.box((-4, -4, -4), (7, 3, 3)), r1)
    d1 = H.box_refine(dom, r1)
    lev2 = H.box_refine(H.box((d1[1][0] - 4 * r1 + 1, -2 * r1, -2 * r1), (d1[1][0], 2 * r1 - 1, 2 * r1 - 1)), r2)
    return H.Hierarchy(dom, [H.domain_split(dom, 8, 4), H.domain_split(lev1, 16, 4), H.domain_split(lev2, 16, 4)],
                       1.0 / 16, alpha, beta, bc_type_lo=bc, bc_type_hi=bc, ratios=ratios)


CASES = {
    "r4_merged": lambda: two_level(64),
    "r4_boxes": lambda: two_level(16),
    "r42": lambda: three_level((4, 2)),
    "r24_helm": lambda: three_level((2, 4), (0, 1, 2), 0.5, 1.2),
    "r44": lambda: three_level((4, 4)),
}


def _pair(ctx, h, precision=0):
    return amrmg.AMRMultiGrid(ctx, h, precision), Oracle(h)


@pytest.mark.parametrize("case", list(CASES))
def test_ratio4_cf_average_bitwise(gpu_ctx, case):
    h = CASES[case]()
    g, o = _pair(gpu_ctx, h)
    rng = np.random.default_rng(61)
    for lev in range(1, h.nlev):
        valid, face, cf = ghost_masks(h, lev)
        phi_g = ghosted_random(h, lev, rng)
        coarse = rng.standard_normal(h.cells(lev - 1))
        np.testing.assert_array_equal(g.quad_cf(lev, phi_g, coarse)[cf], o.quad_cf(lev, phi_g, coarse)[cf])
        np.testing.assert_array_equal(g.homog_cf(lev, phi_g)[cf], o.homog_cf(lev, phi_g)[cf])
        fine = rng.standard_normal(h.cells(lev))
        np.testing.assert_array_equal(g.average(lev, fine), o.average(lev, fine))
    g.close()
    o.close()


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("homog", [False, True])
def test_ratio4_amr_residual_operator_bitwise(gpu_ctx, case, homog):
    h = CASES[case]()
    g, o = _pair(gpu_ctx, h)
    rng = np.random.default_rng(62)
    phi = rng.standard_normal(h.total_cells())
    rhs = rng.standard_normal(h.total_cells())
    g.set_field(g.PHI, phi)
    g.set_field(g.RHS, rhs)
    rg, ng = g.computeAMRResidual(homog)
    ro, no = o.amr_residual(phi, rhs, homog)
    assert ng == no
    np.testing.assert_array_equal(rg, ro)
    unc = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    np.testing.assert_array_equal(g.computeAMROperator(homog)[unc], o.amr_operator(phi, homog)[unc])
    g.close()
    o.close()


@pytest.mark.parametrize("case", list(CASES))
def test_ratio4_solve_bitwise(gpu_ctx, case):
    h = CASES[case]()
    rhs = H.bump_rhs(h, 1061)
    sp, rp = H.solver_params(iterMax=60)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist, info = g.solve()
    phi = g.get_field(g.PHI)
    g.close()
    o = Oracle(h)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    o.close()
    assert info["iterations"] == info_o["iterations"] and info["exit_status"] == info_o["exit_status"]
    assert info["exit_status"] & 1
    np.testing.assert_array_equal(hist, hist_o)
    np.testing.assert_array_equal(phi, phi_o)


@pytest.mark.parametrize("ratios", [(4, 2), (4, 4)])
@pytest.mark.parametrize("precision", [0, 1])
def test_ratio4_parabola_solve(gpu_ctx, ratios, precision):
    h = three_level(ratios, (2, 2, 2), 1.0, 1.0)
    rhs = h.pack(lambda x, y, z, l: h.alpha * parabola(x, y, z) + h.beta * 6.0)
    sp, rp = H.solver_params(eps=1e-11, iterMax=150)
    g = amrmg.AMRMultiGrid(gpu_ctx, h, precision)
    g.set_bc_values(diri)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    _, info = g.solve()
    phi = g.get_field(g.PHI)
    g.close()
    assert info["exit_status"] & 1
    unc = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    assert np.abs(phi - h.pack(parabola))[unc].max() < 1e-9


@pytest.mark.parametrize("case", ["r4_merged", "r42"])
def test_ratio4_mixed_precision_close(gpu_ctx, case):
    h = CASES[case]()
    rhs = H.bump_rhs(h, 1063)
    sp, rp = H.solver_params(iterMax=60)
    g = amrmg.AMRMultiGrid(gpu_ctx, h, 1)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist, info = g.solve()
    phi = g.get_field(g.PHI)
    g.close()
    phi_o, hist_o, _ = Oracle(h).solve(rhs, sp, rp)
    assert hist[0] == hist_o[0]
    assert info["exit_status"] & 1 and hist[-1] <= 1e-10 * hist[0]
    unc = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    assert np.abs(phi - phi_o)[unc].max() <= 1e-9 * np.abs(phi_o)[unc].max()


def test_ratio4_rejected_inputs(gpu_ctx):
    h = two_level(16)
    h.ratios = (3,)
    with pytest.raises(amrmg.AMRMGError, match="ratio"):
        amrmg.AMRMultiGrid(gpu_ctx, h)
    # fine boxes must be coarsenable by the ratio: 6-cell boxes at ratio 4
    dom = H.box((0, 0, 0), (15, 15, 15))
    bad = H.Hierarchy(dom, [H.domain_split(dom, 8, 4), [((16, 16, 16), (21, 21, 21))]], 1.0 / 16, ratios=(4,))
    with pytest.raises(amrmg.AMRMGError, match="coarsenable"):
        amrmg.AMRMultiGrid(gpu_ctx, bad)


@pytest.mark.parametrize("case", ["r42", "r4_boxes"])
def test_ratio4_vc_solve_bitwise(gpu_ctx, case):
    """Variable coefficients (VCAMRPoissonOp2) across ratio-4 levels: reflux with the
    fine faces' b over 16 fine faces per coarse face."""
    h = CASES[case]()
    a, b = H.vc_coefficients(h, seed=9)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.set_vc_coefs(a, b)
    o = Oracle(h)
    o.vc_set(a, b)
    rhs = H.bump_rhs(h, 1065)
    sp, rp = H.solver_params(iterMax=60)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist, info = g.solve()
    phi = g.get_field(g.PHI)
    g.close()
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    o.close()
    assert info["iterations"] == info_o["iterations"] and info["exit_status"] == info_o["exit_status"]
    assert info["exit_status"] & 1
    np.testing.assert_array_equal(hist, hist_o)
    np.testing.assert_array_equal(phi, phi_o)
