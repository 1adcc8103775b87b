"""Boundary value functions (DiriBC / NeumBC with a BCValueHolder) on the HIP path
against the oracle: inhomogeneous level kernels at every multigrid depth, composite
residual and operator, and whole solves bitwise in FP64, on merged and per-box single
levels and a three-level hierarchy whose refined levels reach the domain boundary;
the reference's parabola test (lib/test/AMRElliptic/testAMRPoissonOp.H:50-130)
solved to the sampled parabola; mixed precision within 1e-9; the level-operator
handles; NULL restoring the constant face values."""
import numpy as np
import pytest

import amrmg
import hierarchy as H
import levelops as LO
from helpers import uncovered_mask
from oracle import Oracle

pytestmark = pytest.mark.gpu


def parabola(x, y, z, l=0):
    return x * x + y * y + z * z


def diri(pos, d, s):          # Parabola_diri
    return pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]


def neum(pos, d, s):          # Parabola_neum
    return 2.0 * pos[d]


def wavy(pos, d, s):          # neither constant nor polynomial, different per face
    return np.sin(3.0 * pos[0] + 0.5 * d) * np.cos(2.0 * pos[1] - s) + 0.25 * pos[2]


def _single(n, box, bc, alpha=0.0, beta=1.0, lo=0):
    dom = H.box((lo,) * 3, (lo + n - 1,) * 3)
    return H.Hierarchy(dom, [H.domain_split(dom, box, 4)], 1.0 / n, alpha, beta, bc_type_lo=bc, bc_type_hi=bc)


def _amr(bc, alpha=0.0, beta=1.0):
    dom = H.box((-8,) * 3, (7,) * 3)
    lev1 = H.domain_split(H.box((-8, -8, -8), (15, 7, 7)), 8, 4)
    lev2 = H.domain_split(H.box((8, -8, -8), (31, 7, 7)), 8, 4)
    return H.Hierarchy(dom, [H.domain_split(dom, 8, 4), lev1, lev2], 1.0 / 16, alpha, beta, bc_type_lo=bc,
                       bc_type_hi=bc)


CASES = {
    # one rectangle of 32^3 stored merged (register-ghost relaxation)
    "merged_diri2_wavy": (lambda: _single(32, 32, (2, 2, 2), 0.0, 1.0, lo=-16), wavy),
    # 8^3 boxes, per-box ghost plans, Neumann and linear Dirichlet faces
    "perbox_mixed_wavy": (lambda: _single(32, 8, (0, 1, 2), 0.6, 1.1), wavy),
    "amr_diri2_parabola": (lambda: _amr((2, 2, 2)), diri),
    "amr_neum_parabola_helm": (lambda: _amr((0, 0, 0), 1.0, 1.0), neum),
    "amr_linear_wavy": (lambda: _amr((1, 0, 1), 0.3, 1.0), wavy),
}


def _pair(ctx, h, fn, precision=0):
    g, o = amrmg.AMRMultiGrid(ctx, h, precision), Oracle(h)
    g.set_bc_values(fn)
    o.set_bc_values(fn)
    return g, o


@pytest.mark.parametrize("case", list(CASES))
def test_bcfunc_level_kernels_bitwise(gpu_ctx, case):
    mk, fn = CASES[case]
    h = mk()
    g, o = _pair(gpu_ctx, h, fn)
    rng = np.random.default_rng(41)
    assert g.mg_depths == o.mg_depths
    for d in range(o.mg_depths):
        n = o.mg_cells(d)
        phi, rhs = rng.standard_normal(n), rng.standard_normal(n)
        for homog in (False, True):
            np.testing.assert_array_equal(g.level_apply(d, phi, homog), o.level_apply(d, phi, homog))
            np.testing.assert_array_equal(g.level_residual(d, phi, rhs, homog), o.level_residual(d, phi, rhs, homog))
    g.close()
    o.close()


@pytest.mark.parametrize("case", list(CASES))
def test_bcfunc_amr_residual_operator_bitwise(gpu_ctx, case):
    mk, fn = CASES[case]
    h = mk()
    g, o = _pair(gpu_ctx, h, fn)
    rng = np.random.default_rng(43)
    phi = rng.standard_normal(h.total_cells())
    rhs = rng.standard_normal(h.total_cells())
    g.set_field(g.PHI, phi)
    g.set_field(g.RHS, rhs)
    rg, ng = g.computeAMRResidual(False)
    ro, no = o.amr_residual(phi, rhs, False)
    assert ng == no
    np.testing.assert_array_equal(rg, ro)
    unc = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    np.testing.assert_array_equal(g.computeAMROperator(False)[unc], o.amr_operator(phi, False)[unc])
    g.close()
    o.close()


@pytest.mark.parametrize("case", list(CASES))
def test_bcfunc_solve_bitwise(gpu_ctx, case):
    mk, fn = CASES[case]
    h = mk()
    rhs = H.bump_rhs(h, 1041)
    sp, rp = H.solver_params()
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.set_bc_values(fn)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist, info = g.solve()
    phi = g.get_field(g.PHI)
    g.close()
    o = Oracle(h)
    o.set_bc_values(fn)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    o.close()
    assert info["iterations"] == info_o["iterations"] and info["exit_status"] == info_o["exit_status"]
    assert info["exit_status"] & 1
    np.testing.assert_array_equal(hist, hist_o)
    np.testing.assert_array_equal(phi, phi_o)


@pytest.mark.parametrize("case", ["amr_diri2_parabola", "amr_neum_parabola_helm"])
@pytest.mark.parametrize("precision", [0, 1])
def test_bcfunc_parabola_solve(gpu_ctx, case, precision):
    """rhs = alpha |x|^2 + 6 beta with the parabola's boundary values: the discrete
    solution is the sampled parabola on the uncovered cells (the rules, QuadCFInterp
    and reflux are exact on quadratics)."""
    mk, fn = CASES[case]
    h = mk()
    rhs = h.pack(lambda x, y, z, l: h.alpha * parabola(x, y, z) + h.beta * 6.0)
    sp, rp = H.solver_params(eps=1e-13, iterMax=40)
    g = amrmg.AMRMultiGrid(gpu_ctx, h, precision)
    g.set_bc_values(fn)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    _, info = g.solve()
    phi = g.get_field(g.PHI)
    g.close()
    assert info["exit_status"] & 1
    exact = h.pack(parabola)
    offs = h.level_offsets()
    for l in range(h.nlev):
        m = uncovered_mask(h, l)
        assert np.abs(phi[offs[l]:offs[l + 1]] - exact[offs[l]:offs[l + 1]])[m].max() < 1e-9


@pytest.mark.parametrize("case", ["merged_diri2_wavy", "amr_linear_wavy"])
def test_bcfunc_mixed_precision_close(gpu_ctx, case):
    mk, fn = CASES[case]
    h = mk()
    rhs = H.bump_rhs(h, 1043)
    sp, rp = H.solver_params()
    g = amrmg.AMRMultiGrid(gpu_ctx, h, 1)
    g.set_bc_values(fn)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist, info = g.solve()
    phi = g.get_field(g.PHI)
    g.close()
    o = Oracle(h)
    o.set_bc_values(fn)
    phi_o, hist_o, _ = o.solve(rhs, sp, rp)
    o.close()
    # the FP64 outer residual stops both at 1e-10 of the same initial residual; the
    # solutions then agree to ten times that on the composite (uncovered) cells
    assert hist[0] == hist_o[0]
    assert info["exit_status"] & 1 and hist[-1] <= 1e-10 * hist[0]
    unc = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    assert np.abs(phi - phi_o)[unc].max() <= 1e-9 * np.abs(phi_o)[unc].max()


def test_bcfunc_null_restores_constants(gpu_ctx):
    lo, hi = (0.25, -1.5, 2.0), (0.75, 3.0, -0.5)
    bc = (1, 0, 2)
    dom = H.box((0, 0, 0), (15, 15, 15))
    h = H.Hierarchy(dom, [H.domain_split(dom, 8, 4)], 1.0 / 16, 0.3, 1.0, bc_type_lo=bc, bc_type_hi=bc,
                    bc_val_lo=lo, bc_val_hi=hi)
    rng = np.random.default_rng(45)
    phi = rng.standard_normal(h.total_cells())
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    ref = g.level_apply(0, phi, False)
    g.set_bc_values(lambda pos, d, s: (hi if s else lo)[d])
    np.testing.assert_array_equal(g.level_apply(0, phi, False), ref)
    g.set_bc_values(wavy)
    assert np.abs(g.level_apply(0, phi, False) - ref).max() > 0.0
    g.set_bc_values(None)
    np.testing.assert_array_equal(g.level_apply(0, phi, False), ref)
    g.close()


def test_bcfunc_level_handles_match_oracle(gpu_ctx):
    """applyOp / residual (inhomogeneous) and the BC functor alone through the level
    handles with a value function, on every multigrid depth, against the oracle."""
    h = _single(32, 16, (2, 0, 1), 0.4, 1.0, lo=-16)
    o = Oracle(h)
    o.set_bc_values(wavy)
    L0 = LO.Level(gpu_ctx, h.levels[0], h.level_domain(0), merge=False)
    op0 = LO.AMRPoissonOp(L0, h.dx(0), -1.0, h.alpha, h.beta,
                          (tuple(h.bc_type_lo) + tuple(h.bc_type_hi), tuple(h.bc_val_lo) + tuple(h.bc_val_hi)))
    op0.setBCValues(wavy)
    ops, levels = [op0], [L0]
    while len(ops) < o.mg_depths:
        ops.append(ops[-1].createCoarser())      # inherits the value function
        levels.append(ops[-1].level)
    rng = np.random.default_rng(47)
    fields = []
    for d in range(o.mg_depths):
        op, L = ops[d], levels[d]
        n = o.mg_cells(d)
        phi_h, rhs_h = rng.standard_normal(n), rng.standard_normal(n)
        phi, rhs, out = LO.Field(L), LO.Field(L), LO.Field(L)
        fields += [phi, rhs, out]
        phi.upload(phi_h)
        rhs.upload(rhs_h)
        op.applyOp(out, phi, False)
        np.testing.assert_array_equal(out.download(), o.level_apply(d, phi_h, False))
        op.residual(out, phi, rhs, False)
        np.testing.assert_array_equal(out.download(), o.level_residual(d, phi_h, rhs_h, False))
    for f in fields:
        f.close()
    for op in reversed(ops[1:]):
        op.close()
    op0.close()
    L0.close()
    o.close()
