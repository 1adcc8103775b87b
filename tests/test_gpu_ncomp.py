"""Multi-component fields (amrmg_solver_create_ncomp: AMRPoissonOp on LevelData with
nComp > 1) on the HIP path against the oracle: whole solves (iterations, exit status,
residual history and solution bitwise) on merged, per-box and ratio-4 hierarchies,
Helmholtz with Neumann faces, boundary value functions with one value per component,
the composite residual and operator, other relaxation modes, the four-launch
relaxation, l_base > 0; one component zero gives the one-component solve bit for bit;
the inputs the library rejects."""
import numpy as np
import pytest

import amrmg
import hierarchy as H
from helpers import comps_pack, comps_unpack, lshape_amr, mixed_ratio_amr, small_amr, uncovered_mask
from oracle import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "amr32": small_amr,
    "lshape": lshape_amr,                                   # level 1 stored per box
    "ratio42": lambda: mixed_ratio_amr((4, 2)),
    "helm_neumann": lambda: small_amr(alpha=0.7, beta=1.3),
}


def _hier(case):
    h = CASES[case]()
    if case == "helm_neumann":
        h.bc_type_lo = (0, 1, 0)
        h.bc_type_hi = (1, 0, 1)
        h.bc_val_lo = (0.25, -0.5, 0.0)
        h.bc_val_hi = (1.0, 0.5, -0.75)
    return h


def _rhs(h, nc, seed):
    return comps_pack(h, [H.bump_rhs(h, seed + 7 * c) * (1 + c) for c in range(nc)])


def _solve_pair(ctx, h, nc, rhs, sp, rp, setup=None, phi0=None, l_max=None, l_base=0):
    g = amrmg.AMRMultiGrid(ctx, h, ncomp=nc)
    o = Oracle(h)
    o.set_ncomp(nc)
    if setup:
        setup(g, o)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    if phi0 is not None:
        g.set_field(g.PHI, phi0)
    hist, info = g.solve(zero_phi=True, l_max=l_max, l_base=l_base)
    phi = g.get_field(g.PHI)
    g.close()
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp, phi0=phi0, zero_phi=True, l_max=l_max, l_base=l_base)
    o.close()
    return (phi, hist, info), (phi_o, hist_o, info_o)


def _same(a, b):
    (phi, hist, info), (phi_o, hist_o, info_o) = a, b
    assert info["iterations"] == info_o["iterations"] and info["exit_status"] == info_o["exit_status"]
    assert info["exit_status"] & 1
    np.testing.assert_array_equal(hist, hist_o)
    np.testing.assert_array_equal(phi, phi_o)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("nc", [2, 3])
def test_ncomp_solve_bitwise(gpu_ctx, case, nc):
    h = _hier(case)
    sp, rp = H.solver_params(iterMax=60)
    _same(*_solve_pair(gpu_ctx, h, nc, _rhs(h, nc, 1201), sp, rp))


@pytest.mark.parametrize("case", ["amr32", "ratio42"])
def test_ncomp_bc_value_functions_bitwise(gpu_ctx, case):
    """Order-2 Dirichlet on x, linear Dirichlet on y, Neumann on z, each component with
    its own values from one function."""
    h = _hier(case)
    h.bc_type_lo = h.bc_type_hi = (2, 1, 0)

    def vals(pos, d, s):
        x, y, z = pos
        return (x * x - y + 0.5 * z, np.sin(3 * x) + d - s, 2.0 * y * z - 1.0)

    def setup(g, o):
        g.set_bc_values(vals)
        o.set_bc_values(vals)

    sp, rp = H.solver_params(iterMax=60)
    _same(*_solve_pair(gpu_ctx, h, 3, _rhs(h, 3, 1203), sp, rp, setup))


@pytest.mark.parametrize("case", ["amr32", "lshape"])
@pytest.mark.parametrize("homog", [False, True])
def test_ncomp_residual_operator_bitwise(gpu_ctx, case, homog):
    h = _hier(case)
    h.bc_val_lo = (0.5, -1.0, 2.0)
    rng = np.random.default_rng(1205)
    phi = rng.standard_normal(h.total_cells() * 3)
    rhs = rng.standard_normal(h.total_cells() * 3)
    g = amrmg.AMRMultiGrid(gpu_ctx, h, ncomp=3)
    g.set_field(g.PHI, phi)
    g.set_field(g.RHS, rhs)
    rg, ng = g.computeAMRResidual(homog)
    lg = g.computeAMROperator(homog)
    g.close()
    o = Oracle(h)
    o.set_ncomp(3)
    ro, no = o.amr_residual(phi, rhs, homog)
    lo = o.amr_operator(phi, homog)
    o.close()
    assert ng == no
    np.testing.assert_array_equal(rg, ro)
    unc = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    for a, b in zip(comps_unpack(h, lg, 3), comps_unpack(h, lo, 3)):
        np.testing.assert_array_equal(a[unc], b[unc])


@pytest.mark.parametrize("mode", [0, 3, 5])
def test_ncomp_relax_modes_bitwise(gpu_ctx, mode):
    h = _hier("amr32")

    def setup(g, o):
        g.set_relax_mode(mode)
        o.set_relax_mode(mode)

    sp, rp = H.solver_params(iterMax=80)
    _same(*_solve_pair(gpu_ctx, h, 2, _rhs(h, 2, 1207), sp, rp, setup))


def test_ncomp_four_launch_relaxation_equals_one_launch(gpu_ctx):
    h = _hier("amr32")
    rhs = _rhs(h, 2, 1209)
    sp, rp = H.solver_params(iterMax=60)
    out = []
    for fused in (True, False):
        gpu_ctx.set_fused_relax(fused)
        try:
            g = amrmg.AMRMultiGrid(gpu_ctx, h, ncomp=2)
            g.setSolverParameters(sp, rp)
            g.set_field(g.RHS, rhs)
            hist, info = g.solve()
            out.append((g.get_field(g.PHI), hist))
            g.close()
        finally:
            gpu_ctx.set_fused_relax(True)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("l_base,l_max", [(1, 2), (2, 2)])
def test_ncomp_lbase_bitwise(gpu_ctx, l_base, l_max):
    h = _hier("amr32")
    rng = np.random.default_rng(1211)
    rhs = _rhs(h, 2, 1213)
    phi0 = rng.standard_normal(h.total_cells() * 2) * 1e-2
    sp, rp = H.solver_params(iterMax=60)
    a, b = _solve_pair(gpu_ctx, h, 2, rhs, sp, rp, phi0=phi0, l_max=l_max, l_base=l_base)
    _same(a, b)


@pytest.mark.parametrize("zero_first", [False, True])
def test_ncomp_zero_component_equals_one_component_solve(gpu_ctx, zero_first):
    h = _hier("lshape")
    f = H.bump_rhs(h, 1215)
    sp, rp = H.solver_params(iterMax=60)
    g1 = amrmg.AMRMultiGrid(gpu_ctx, h)
    g1.setSolverParameters(sp, rp)
    g1.set_field(g1.RHS, f)
    hist1, info1 = g1.solve()
    phi1 = g1.get_field(g1.PHI)
    g1.close()
    g = amrmg.AMRMultiGrid(gpu_ctx, h, ncomp=2)
    g.setSolverParameters(sp, rp)
    z = np.zeros_like(f)
    g.set_field(g.RHS, comps_pack(h, [z, f] if zero_first else [f, z]))
    hist, info = g.solve()
    c0, c1 = comps_unpack(h, g.get_field(g.PHI), 2)
    g.close()
    got, zero = (c1, c0) if zero_first else (c0, c1)
    assert info["iterations"] == info1["iterations"]
    np.testing.assert_array_equal(hist, hist1)
    np.testing.assert_array_equal(got, phi1)
    assert not zero.any()


def test_ncomp_rejected_inputs(gpu_ctx):
    h = small_amr()
    for bad in (0, 17):
        with pytest.raises(amrmg.AMRMGError, match="ncomp"):
            amrmg.AMRMultiGrid(gpu_ctx, h, ncomp=bad)
    with pytest.raises(amrmg.AMRMGError, match="mixed precision"):
        amrmg.AMRMultiGrid(gpu_ctx, h, precision=1, ncomp=2)
    g = amrmg.AMRMultiGrid(gpu_ctx, h, ncomp=2)
    a, b = H.vc_coefficients(h, seed=3)
    with pytest.raises(amrmg.AMRMGError, match="one component"):
        g.set_vc_coefs(a, b)
    with pytest.raises(amrmg.AMRMGError, match="one component"):
        amrmg.HeatSolver(g, "tga")
    with pytest.raises(ValueError):
        g.set_field(g.RHS, np.zeros(h.total_cells()))
    g.close()
