"""The quadratic Dirichlet boundary rule (BC type 2, DiriBC order 2) on the HIP
path against the oracle: level operations at every multigrid depth and whole
solves bitwise equal, on one level, on an AMR hierarchy, with variable
coefficients; the batched solver rejects it with an error."""
import numpy as np
import pytest

import amrmg
import hierarchy as H
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _single(n, box, alpha, beta, lo_type, hi_type, lo_val=(0.0, 0.0, 0.0), hi_val=(0.0, 0.0, 0.0), lo=0):
    dom = H.box((lo,) * 3, (lo + n - 1,) * 3)
    return H.Hierarchy(dom, [H.domain_split(dom, box, 8)], 1.0 / n, alpha, beta, bc_type_lo=lo_type,
                       bc_type_hi=hi_type, bc_val_lo=lo_val, bc_val_hi=hi_val)


CASES = {
    "x_quadratic": dict(n=32, box=16, alpha=0.0, beta=1.0, lo_type=(2, 1, 1), hi_type=(2, 1, 1),
                        lo_val=(0.5, 0.0, 0.0), hi_val=(-0.25, 0.0, 0.0)),
    "mixed_faces": dict(n=24, box=8, alpha=0.6, beta=1.2, lo_type=(2, 0, 1), hi_type=(1, 2, 2),
                        lo_val=(0.1, 0.2, 0.0), hi_val=(0.0, -0.3, 0.4), lo=-8),
}


@pytest.mark.parametrize("case", list(CASES))
def test_quadratic_dirichlet_level_ops_and_solve_bitwise(gpu_ctx, case):
    print(f"bc2 {case}", flush=True)
    h = _single(**CASES[case])
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    o = Oracle(h)
    rng = np.random.default_rng(41)
    for d in range(o.mg_depths):
        n = o.mg_cells(d)
        phi, rhs = rng.standard_normal(n), rng.standard_normal(n)
        for homog in (False, True):
            np.testing.assert_array_equal(g.level_apply(d, phi, homog), o.level_apply(d, phi, homog))
            np.testing.assert_array_equal(g.level_residual(d, phi, rhs, homog), o.level_residual(d, phi, rhs, homog))
        np.testing.assert_array_equal(g.level_relax(d, phi, rhs, 2), o.level_relax(d, phi, rhs, 2))
        if d + 1 < o.mg_depths:
            np.testing.assert_array_equal(g.level_restrict(d, phi, rhs), o.level_restrict(d, phi, rhs))
    rhs = H.sinusoid_rhs(h, 1401)
    sp, rp = H.solver_params()
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    assert info_g["exit_status"] & 1
    assert info_g["iterations"] == info_o["iterations"] and info_g["exit_status"] == info_o["exit_status"]
    np.testing.assert_array_equal(hist_g, hist_o)
    np.testing.assert_array_equal(g.get_field(g.PHI), phi_o)
    g.close()


@pytest.mark.parametrize("vc", [False, True])
def test_quadratic_dirichlet_amr_solve_bitwise(gpu_ctx, vc):
    print(f"bc2 amr vc={vc}", flush=True)
    h = H.amr_hierarchy(32, fracs=(0.25, 0.06), max_box=16, block=4, alpha=0.2, beta=1.0)
    h.bc_type_lo = (2, 2, 1)
    h.bc_type_hi = (1, 2, 2)
    h.bc_val_lo = (0.3, 0.0, 0.0)
    rhs = H.bump_rhs(h, 1402)
    sp, rp = H.solver_params()
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    o = Oracle(h)
    if vc:
        a, b = H.vc_coefficients(h, seed=43)
        g.set_vc_coefs(a, b)
        o.vc_set(a, b)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    assert info_g["exit_status"] & 1
    assert info_g["iterations"] == info_o["iterations"]
    np.testing.assert_array_equal(hist_g, hist_o)
    np.testing.assert_array_equal(g.get_field(g.PHI), phi_o)
    g.close()


def test_batch_rejects_quadratic_dirichlet(gpu_ctx):
    h = _single(**CASES["x_quadratic"])
    with pytest.raises(amrmg.AMRMGError, match="quadratic Dirichlet"):
        amrmg.BatchSolver(gpu_ctx, [h])
