"""Periodic domains (BC type 3 on both faces of a direction) on the HIP path against
the oracle: level kernels at every multigrid depth and whole solves bitwise, on
fully periodic and mixed periodic / Dirichlet domains, one level and with a refined
level inside the domain; mixed precision within 1e-9; the inputs it rejects."""
import numpy as np
import pytest

import amrmg
import hierarchy as H
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _single(n, box, alpha, beta, bc):
    dom = H.box((0, 0, 0), (n - 1,) * 3)
    return H.Hierarchy(dom, [H.domain_split(dom, box, 4)], 1.0 / n, alpha, beta, bc_type_lo=bc, bc_type_hi=bc)


def _amr(bc=(3, 1, 3)):
    dom = H.box((0, 0, 0), (31, 31, 31))
    fine = H.domain_split(H.box((8, 12, 8), (47, 51, 39)), 16, 4)
    return H.Hierarchy(dom, [H.domain_split(dom, 16, 4), fine], 1.0 / 32, 0.0, 1.0, bc_type_lo=bc,
                       bc_type_hi=bc)


CASES = {
    "torus_helm": lambda: _single(32, 8, 1.0, -0.05, (3, 3, 3)),
    "channel_lap": lambda: _single(32, 16, 0.0, 1.0, (1, 3, 3)),
    "slab_neumann_helm": lambda: _single(24, 8, 0.7, 1.1, (3, 0, 3)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_periodic_level_kernels_bitwise(gpu_ctx, case):
    h = CASES[case]()
    g, o = amrmg.AMRMultiGrid(gpu_ctx, h), Oracle(h)
    rng = np.random.default_rng(21)
    assert g.mg_depths == o.mg_depths
    for d in range(o.mg_depths):
        n = o.mg_cells(d)
        phi, rhs = rng.standard_normal(n), rng.standard_normal(n)
        for homog in (False, True):
            np.testing.assert_array_equal(g.level_apply(d, phi, homog), o.level_apply(d, phi, homog))
            np.testing.assert_array_equal(g.level_residual(d, phi, rhs, homog), o.level_residual(d, phi, rhs, homog))
        np.testing.assert_array_equal(g.level_relax(d, phi, rhs, 2), o.level_relax(d, phi, rhs, 2))
        if d + 1 < o.mg_depths:
            np.testing.assert_array_equal(g.level_restrict(d, phi, rhs), o.level_restrict(d, phi, rhs))
            ec = rng.standard_normal(o.mg_cells(d + 1))
            np.testing.assert_array_equal(g.level_prolong(d, phi, ec), o.level_prolong(d, phi, ec))
    g.close()
    o.close()


@pytest.mark.parametrize("case", list(CASES) + ["amr_periodic_xz", "amr_periodic_x_helm"])
def test_periodic_solve_bitwise(gpu_ctx, case):
    if case == "amr_periodic_xz":
        h = _amr()
    elif case == "amr_periodic_x_helm":
        h = _amr(bc=(3, 1, 1))
        h.alpha, h.beta = 0.4, 1.2
    else:
        h = CASES[case]()
    rhs = H.bump_rhs(h, 1031)
    sp, rp = H.solver_params()
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


def test_periodic_mixed_precision_within_1e9(gpu_ctx):
    h = _amr()
    rhs = H.bump_rhs(h, 1032)
    sp, rp = H.solver_params()
    g = amrmg.AMRMultiGrid(gpu_ctx, h, precision=1)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist, info = g.solve()
    phi = g.get_field(g.PHI)
    g.close()
    phi_o, _, info_o = Oracle(h).solve(rhs, sp, rp)
    assert info["exit_status"] & 1
    offs = h.level_offsets()
    from helpers import uncovered_mask
    mask = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    err = float(np.abs(phi - phi_o)[mask].max() / np.abs(phi_o)[mask].max())
    assert err <= 1e-9, err


def test_periodic_inputs_rejected(gpu_ctx):
    dom = H.box((0, 0, 0), (31, 31, 31))
    near = H.domain_split(H.box((0, 8, 8), (31, 47, 47)), 16, 4)  # touches the periodic face x = 0
    h = H.Hierarchy(dom, [H.domain_split(dom, 16, 4), near], 1.0 / 32, bc_type_lo=(3, 1, 1), bc_type_hi=(3, 1, 1))
    with pytest.raises(amrmg.AMRMGError):
        amrmg.AMRMultiGrid(gpu_ctx, h)
    one_sided = _single(16, 8, 0.0, 1.0, (1, 1, 1))
    one_sided.bc_type_lo = (3, 1, 1)
    with pytest.raises(amrmg.AMRMGError):
        amrmg.AMRMultiGrid(gpu_ctx, one_sided)
    with pytest.raises(amrmg.AMRMGError):
        amrmg.BatchSolver(gpu_ctx, [_single(16, 8, 0.5, 1.0, (3, 3, 3))], 0)
