"""Variable-coefficient operator (VCAMRPoissonOp2) on the HIP path against the oracle.

The kernels restate GSRBHELMHOLTZVC, VCCOMPUTEOP, VCCOMPUTERES, RESTRICTRESVC and
SUMFACES in the reference's expression order (no FMA contraction), and the
multigrid coefficients are averaged in CoarseAverage / CoarseAverageFace order,
so level operations at every depth and whole FP64 solves are compared bitwise.
Mixed-precision solves are held to 1e-9 relative of the FP64 oracle's solution
(ten times the 1e-10 solve tolerance), as for constant coefficients.
"""
import numpy as np
import pytest

import amrmg
import hierarchy as H
from helpers import lshape_amr, small_amr, small_single, uncovered_mask
from oracle import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "vc32": dict(n=32, box=16, alpha=0.5, beta=1.0),
    "vc32_negidx": dict(n=32, box=16, lo=-16, alpha=0.0, beta=1.0),
    "vc40_ragged": dict(n=40, box=16, alpha=1.5, beta=0.8),
    "vc16_neumann": dict(n=16, box=8, bc=(0, 0, 0), alpha=1.0, beta=1.0),
}


def _pair(gpu_ctx, name, seed=5, precision=0):
    h = small_single(**CASES[name])
    a, b = H.vc_coefficients(h, seed=seed)
    g = amrmg.AMRMultiGrid(gpu_ctx, h, precision=precision)
    g.set_vc_coefs(a, b)
    o = Oracle(h)
    o.vc_set(a, b)
    return h, g, o


@pytest.mark.parametrize("case", list(CASES))
def test_vc_level_kernels_bitwise(gpu_ctx, case):
    print(f"vc level kernels {case}", flush=True)
    h, g, o = _pair(gpu_ctx, case)
    rng = np.random.default_rng(17)
    assert g.mg_depths == o.mg_depths
    for d in range(o.mg_depths):
        n = o.mg_cells(d)
        phi, rhs = rng.standard_normal(n), rng.standard_normal(n)
        for homog in (False, True):
            np.testing.assert_array_equal(g.level_apply(d, phi, homog), o.level_apply(d, phi, homog))
            np.testing.assert_array_equal(g.level_residual(d, phi, rhs, homog), o.level_residual(d, phi, rhs, homog))
        for iters in (1, 2):
            np.testing.assert_array_equal(g.level_relax(d, phi, rhs, iters), o.level_relax(d, phi, rhs, iters))
        if d + 1 < o.mg_depths:
            np.testing.assert_array_equal(g.level_restrict(d, phi, rhs), o.level_restrict(d, phi, rhs))
    g.close()


def test_vc_bottom_solve_bitwise(gpu_ctx):
    print("vc bottom solve", flush=True)
    h, g, o = _pair(gpu_ctx, "vc32")
    d = o.mg_depths - 1
    rng = np.random.default_rng(3)
    res = rng.standard_normal(o.mg_cells(d))
    c0 = np.zeros_like(res)
    cg, sg, _ = g.bottom_solve(c0, res, 2, 1.0, 1e-10)
    co, so = o.bottom_solve(c0, res, 2, 1.0, 1e-10)
    assert sg == so
    np.testing.assert_array_equal(cg, co)
    g.close()


@pytest.mark.parametrize("case", list(CASES))
def test_vc_solve_matches_oracle_bitwise(gpu_ctx, case):
    print(f"vc solve {case}", flush=True)
    h, g, o = _pair(gpu_ctx, case, seed=8)
    rhs = H.sinusoid_rhs(h, 1005, k=(1, 2, 1))
    sp, rp = H.solver_params()
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_g = g.get_field(g.PHI)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    assert info_g["iterations"] == info_o["iterations"]
    assert info_g["exit_status"] == info_o["exit_status"]
    assert info_g["exit_status"] & 1
    np.testing.assert_array_equal(hist_g, hist_o)
    np.testing.assert_array_equal(phi_g, phi_o)
    g.close()


@pytest.mark.parametrize("case", ["vc32", "vc40_ragged"])
def test_vc_mixed_precision_solve(gpu_ctx, case):
    print(f"vc mixed {case}", flush=True)
    h, g, o = _pair(gpu_ctx, case, seed=8, precision=1)
    rhs = H.sinusoid_rhs(h, 1005, k=(1, 2, 1))
    sp, rp = H.solver_params()
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist, info = g.solve()
    phi = g.get_field(g.PHI)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    err = np.abs(phi - phi_o).max() / np.abs(phi_o).max()
    print(f"{case}: mixed {info['iterations']} it, fp64 {info_o['iterations']} it, err {err:.2e}")
    assert hist[0] == hist_o[0]
    assert info["exit_status"] & 1 and hist[-1] <= 1e-10 * hist[0]
    assert abs(info["iterations"] - info_o["iterations"]) <= 2
    assert err < 1e-9, err
    g.close()


def test_vc_unit_coefficients_match_constant_solver(gpu_ctx):
    """a = b = 1 with beta -> -beta is the constant-coefficient operator: the two
    GPU solves agree to rounding (their expression orders differ)."""
    print("vc unit coefficients", flush=True)
    hc = small_single(n=32, box=16, alpha=0.6, beta=1.7)
    hv = small_single(n=32, box=16, alpha=0.6, beta=-1.7)
    rhs = H.sinusoid_rhs(hc, 1006)
    sp, rp = H.solver_params()
    c = amrmg.AMRMultiGrid(gpu_ctx, hc)
    v = amrmg.AMRMultiGrid(gpu_ctx, hv)
    o = Oracle(hv)
    v.set_vc_coefs(np.ones(hv.cells(0)), [np.ones(o.face_cells(0, d)) for d in range(3)])
    out = []
    for s in (c, v):
        s.setSolverParameters(sp, rp)
        s.set_field(s.RHS, rhs)
        s.solve()
        out.append(s.get_field(s.PHI))
        s.close()
    assert np.abs(out[0] - out[1]).max() < 1e-9 * np.abs(out[0]).max()


def test_vc_clear_and_alpha_beta_equal_fresh_solvers(gpu_ctx):
    """Coefficients set on a solver that has solved, then alpha and beta changed
    (lambda recomputed at every depth), give the fresh solver's solve bit for
    bit; clearing them returns to the constant-coefficient solve (device bottom
    solver included)."""
    print("vc reset", flush=True)
    h1 = small_single(n=32, box=16, alpha=0.4, beta=1.3)
    h2 = small_single(n=32, box=16, alpha=0.9, beta=0.6)
    a, b = H.vc_coefficients(h1, seed=12)
    rhs = H.sinusoid_rhs(h1, 1010)
    sp, rp = H.solver_params()

    def run(s):
        s.setSolverParameters(sp, rp)
        s.set_field(s.RHS, rhs)
        hist, info = s.solve()
        return hist, info, s.get_field(s.PHI)

    g = amrmg.AMRMultiGrid(gpu_ctx, h1)
    base = run(g)
    g.set_vc_coefs(a, b)
    run(g)
    g.set_alpha_beta(0.9, 0.6)
    got = run(g)
    f = amrmg.AMRMultiGrid(gpu_ctx, h2)
    f.set_vc_coefs(a, b)
    ref = run(f)
    f.close()
    assert got[1] == ref[1]
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[2], ref[2])
    g.set_alpha_beta(0.4, 1.3)
    g.set_vc_coefs(None, None)
    back = run(g)
    assert back[1] == base[1]
    np.testing.assert_array_equal(back[2], base[2])
    g.close()


@pytest.mark.parametrize("case", ["vc40_ragged", "vc16_neumann"])
def test_vc_fused_relax_shapes_match_two_colour_launches(gpu_ctx, case):
    """The one-launch VC relaxation (every tile shape) against the reference's
    sequence (ghost refresh, red, ghost refresh, black), both on the GPU: level
    relaxations at every depth and whole solves agree bitwise."""
    print(f"vc fused shapes {case}", flush=True)
    h = small_single(**CASES[case])
    a, b = H.vc_coefficients(h, seed=21)
    rhs = H.sinusoid_rhs(h, 1008, k=(2, 1, 1))
    sp, rp = H.solver_params()
    rng = np.random.default_rng(23)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.set_vc_coefs(a, b)
    depths = g.mg_depths
    fields = [(rng.standard_normal(g.mg_cells(d)), rng.standard_normal(g.mg_cells(d))) for d in range(depths)]

    def run(shape, fused):
        gpu_ctx.set_fused_relax(fused)
        g.set_vc_relax_shape(shape)
        lv = [g.level_relax(d, p, r, 2) for d, (p, r) in enumerate(fields)]
        g.setSolverParameters(sp, rp)
        g.set_field(g.RHS, rhs)
        hist, info = g.solve()
        return lv, hist, info, g.get_field(g.PHI)

    try:
        ref = run(-1, False)
        for shape in (-1, 0, 1, 2, 3):
            got = run(shape, True)
            for d in range(depths):
                np.testing.assert_array_equal(got[0][d], ref[0][d])
            assert got[2] == ref[2]
            np.testing.assert_array_equal(got[1], ref[1])
            np.testing.assert_array_equal(got[3], ref[3])
    finally:
        gpu_ctx.set_fused_relax(True)
    g.close()


def _amr(name, alpha=0.0, beta=1.0):
    if name == "amr32":
        return small_amr(alpha=alpha, beta=beta)
    if name == "lshape":
        return lshape_amr(alpha=alpha, beta=beta)
    return H.amr_hierarchy(64, fracs=(0.25, 0.06), max_box=32, block=8, alpha=alpha, beta=beta)


@pytest.mark.parametrize("hier", ["amr32", "lshape"])
@pytest.mark.parametrize("homog", [False, True])
def test_vc_amr_residual_operator_bitwise(gpu_ctx, homog, hier):
    """Composite residual and operator with variable coefficients on every level
    (coarse-fine interpolation, b-weighted refluxing) against the oracle."""
    print(f"vc amr residual {hier} homog={homog}", flush=True)
    h = _amr(hier, 0.3, 1.1)
    a, b = H.vc_coefficients(h, seed=31)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.set_vc_coefs(a, b)
    o = Oracle(h)
    o.vc_set(a, b)
    rng = np.random.default_rng(9)
    phi = rng.standard_normal(h.total_cells())
    rhs = rng.standard_normal(h.total_cells())
    g.set_field(g.PHI, phi)
    g.set_field(g.RHS, rhs)
    rg, ng = g.computeAMRResidual(homog)
    ro, no = o.amr_residual(phi, rhs, homog)
    assert ng == no
    np.testing.assert_array_equal(rg, ro)
    lg = g.computeAMROperator(homog)
    lo = o.amr_operator(phi, homog)
    offs = h.level_offsets()
    for lev in range(h.nlev):
        m = uncovered_mask(h, lev)
        np.testing.assert_array_equal(lg[offs[lev]:offs[lev + 1]][m], lo[offs[lev]:offs[lev + 1]][m])
    g.close()


@pytest.mark.parametrize("hier", ["amr32", "lshape", "amr64"])
def test_vc_amr_solve_bitwise(gpu_ctx, hier):
    print(f"vc amr solve {hier}", flush=True)
    h = _amr(hier)
    a, b = H.vc_coefficients(h, seed=33)
    rhs = H.bump_rhs(h, 1011)
    sp, rp = H.solver_params()
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.set_vc_coefs(a, b)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_g = g.get_field(g.PHI)
    o = Oracle(h)
    o.vc_set(a, b)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    print(f"{hier}: {info_g}, {hist_g[-1] / hist_g[0]:.2e}")
    assert info_g["iterations"] == info_o["iterations"]
    assert info_g["exit_status"] == info_o["exit_status"]
    assert info_g["exit_status"] & 1
    np.testing.assert_array_equal(hist_g, hist_o)
    np.testing.assert_array_equal(phi_g, phi_o)
    g.close()


@pytest.mark.parametrize("variant", ["wcycle", "neumann_mixed_faces", "negidx"])
def test_vc_amr_solve_variants_bitwise(gpu_ctx, variant):
    """W-cycles (numMG = 2), Neumann faces on the low sides (a Helmholtz problem,
    alpha > 0, so it stays definite), and an AMR hierarchy at negative indices."""
    print(f"vc amr variant {variant}", flush=True)
    kw = {}
    if variant == "wcycle":
        h = _amr("amr32")
        kw = dict(numMG=2)
    elif variant == "neumann_mixed_faces":
        h = small_amr(alpha=0.8, beta=1.0)
        h.bc_type_lo = (0, 0, 0)
    else:
        base = H.box((-32, -32, -32), (-1, -1, -1))
        lev1 = H.domain_split(H.box((-48, -48, -48), (-17, -17, -17)), 16, 4)
        h = H.Hierarchy(base, [H.domain_split(base, 16, 4), lev1], 1.0 / 32, 0.2, 1.0)
    a, b = H.vc_coefficients(h, seed=35)
    rhs = H.sinusoid_rhs(h, 1013) if variant == "negidx" else H.bump_rhs(h, 1013)
    sp, rp = H.solver_params(**kw)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.set_vc_coefs(a, b)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_g = g.get_field(g.PHI)
    o = Oracle(h)
    o.vc_set(a, b)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    print(f"{variant}: {info_g}")
    if variant != "wcycle":
        # (numMG = 2 on an AMR hierarchy does not converge, as with constant
        # coefficients: the oracle stops on the hang test and the GPU must follow it)
        assert info_g["exit_status"] & 1
    assert info_g["iterations"] == info_o["iterations"]
    assert info_g["exit_status"] == info_o["exit_status"]
    np.testing.assert_array_equal(hist_g, hist_o)
    np.testing.assert_array_equal(phi_g, phi_o)
    g.close()


def test_vc_full_size_linearity_and_determinism(gpu_ctx):
    """The benchmark's 3-level hierarchy (113.5 M cells) with variable coefficients:
    the solve converges to 1e-10, repeats bitwise, and solving with 2 rhs gives
    exactly 2 phi (scaling by 2 is exact in every operation of the cycle)."""
    print("vc full size", flush=True)
    h = H.amr_hierarchy(256, fracs=(0.25, 0.06), max_box=64, block=8)
    a, b = H.vc_coefficients(h, seed=7)
    rhs = H.bump_rhs(h, 1002)
    sp, rp = H.solver_params()
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.set_vc_coefs(a, b)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist1, info1 = g.solve()
    phi1 = g.get_field(g.PHI)
    assert info1["exit_status"] & 1 and hist1[-1] <= 1e-10 * hist1[0]
    hist2, _ = g.solve()
    np.testing.assert_array_equal(hist2, hist1)
    np.testing.assert_array_equal(g.get_field(g.PHI), phi1)
    g.set_field(g.RHS, 2 * rhs)
    g.solve()
    np.testing.assert_array_equal(g.get_field(g.PHI), 2 * phi1)
    g.close()


def test_vc_amr_mixed_precision_solve(gpu_ctx):
    print("vc amr mixed", flush=True)
    h = _amr("amr32")
    a, b = H.vc_coefficients(h, seed=33)
    rhs = H.bump_rhs(h, 1011)
    sp, rp = H.solver_params()
    g = amrmg.AMRMultiGrid(gpu_ctx, h, precision=1)
    g.set_vc_coefs(a, b)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist, info = g.solve()
    phi = g.get_field(g.PHI)
    o = Oracle(h)
    o.vc_set(a, b)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    m = np.concatenate([uncovered_mask(h, lev) for lev in range(h.nlev)])
    err = np.abs(phi - phi_o)[m].max() / np.abs(phi_o)[m].max()
    print(f"amr32 mixed: {info['iterations']} it, fp64 {info_o['iterations']} it, composite err {err:.2e}")
    assert hist[0] == hist_o[0]
    assert info["exit_status"] & 1 and hist[-1] <= 1e-10 * hist[0]
    assert abs(info["iterations"] - info_o["iterations"]) <= 2
    assert err < 1e-9, err
    g.close()


@pytest.mark.parametrize("hier", ["single", "amr32"])
def test_vc_jacobi_relax_mode_bitwise(gpu_ctx, hier):
    print(f"vc jacobi {hier}", flush=True)
    if hier == "single":
        h = small_single(**CASES["vc32"])
        rhs = H.sinusoid_rhs(h, 1032)
    else:
        h = _amr("amr32")
        rhs = H.bump_rhs(h, 1032)
    a, b = H.vc_coefficients(h, seed=51)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.set_vc_coefs(a, b)
    g.set_relax_mode(4)
    o = Oracle(h)
    o.vc_set(a, b)
    o.set_relax_mode(4)
    rng = np.random.default_rng(16)
    for d in range(o.mg_depths):
        n = o.mg_cells(d)
        phi, r = rng.standard_normal(n), rng.standard_normal(n)
        np.testing.assert_array_equal(g.level_relax(d, phi, r, 2), o.level_relax(d, phi, r, 2))
    sp, rp = H.solver_params(pre=4, post=4)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    assert info_g["exit_status"] & 1
    assert info_g["iterations"] == info_o["iterations"]
    np.testing.assert_array_equal(hist_g, hist_o)
    np.testing.assert_array_equal(g.get_field(g.PHI), phi_o)
    g.close()


@pytest.mark.parametrize("hier", ["single", "amr32"])
def test_vc_harmonic_coarsening_bitwise(gpu_ctx, hier):
    """Harmonic coefficient coarsening on the multigrid depths: every depth's lambda,
    level relaxation, and whole solves (FP64 bitwise; mixed precision within 1e-9)
    equal the oracle's, and differ from the arithmetic coarsening."""
    print(f"vc harmonic {hier}", flush=True)
    if hier == "single":
        h = small_single(**CASES["vc32"])
        rhs = H.sinusoid_rhs(h, 1033)
    else:
        h = _amr("amr32")
        rhs = H.bump_rhs(h, 1033)
    a, b = H.vc_coefficients(h, seed=52, contrast=8.0)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.set_vc_average(True)
    g.set_vc_coefs(a, b)
    o = Oracle(h)
    o.set_vc_average(True)
    o.vc_set(a, b)
    oa = Oracle(h)
    oa.vc_set(a, b)
    rng = np.random.default_rng(17)
    for d in range(o.mg_depths):
        n = o.mg_cells(d)
        phi, r = rng.standard_normal(n), rng.standard_normal(n)
        np.testing.assert_array_equal(g.level_relax(d, phi, r, 2), o.level_relax(d, phi, r, 2))
        if d > 0:
            assert np.abs(o.vc_lambda(d) - oa.vc_lambda(d)).max() > 0
    sp, rp = H.solver_params()
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    assert info_g["exit_status"] & 1
    assert info_g["iterations"] == info_o["iterations"]
    np.testing.assert_array_equal(hist_g, hist_o)
    np.testing.assert_array_equal(g.get_field(g.PHI), phi_o)
    g.close()
    gm = amrmg.AMRMultiGrid(gpu_ctx, h, precision=1)
    gm.set_vc_average(True)
    gm.set_vc_coefs(a, b)
    gm.setSolverParameters(sp, rp)
    gm.set_field(gm.RHS, rhs)
    _, info_m = gm.solve()
    assert info_m["exit_status"] & 1
    pm = gm.get_field(gm.PHI)
    # covered cells hold no part of the composite solution
    m = np.concatenate([uncovered_mask(h, lev) for lev in range(h.nlev)])
    assert np.abs(pm - phi_o)[m].max() / np.abs(phi_o)[m].max() < 1e-9
    gm.close()


def test_vc_rejects_bad_input(gpu_ctx):
    print("vc errors", flush=True)
    h = small_single(n=32, box=16)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    a, b = H.vc_coefficients(h)
    # box 0 is [0, 15]^3 and box 1 starts at z = 16: face (0, 0, 16) normal to z is
    # the last plane of box 0's z faces and the first of box 1's
    assert h.levels[0][0] == ((0, 0, 0), (15, 15, 15)) and h.levels[0][1][0] == (0, 0, 16)
    bad = b[2].copy()
    bad[16 * 16 * 16] += 1.0
    with pytest.raises(amrmg.AMRMGError, match="shared by two boxes"):
        g.set_vc_coefs(a, [b[0], b[1], bad])
    g.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_vc_average_after_coefs_takes_effect(gpu_ctx, precision):
    """Choosing harmonic coarsening after the coefficients are set re-averages the
    multigrid depths at once: the solve equals the one of a solver that chose it
    first (FP64 and mixed precision), and differs from the arithmetic one."""
    h = small_single(**CASES["vc32"])
    rhs = H.sinusoid_rhs(h, 1034)
    a, b = H.vc_coefficients(h, seed=53, contrast=8.0)
    sp, rp = H.solver_params()
    out = []
    for order in ("before", "after", "never"):
        g = amrmg.AMRMultiGrid(gpu_ctx, h, precision=precision)
        if order == "before":
            g.set_vc_average(True)
        g.set_vc_coefs(a, b)
        if order == "after":
            g.set_vc_average(True)
        g.setSolverParameters(sp, rp)
        g.set_field(g.RHS, rhs)
        hist, info = g.solve()
        out.append((hist, g.get_field(g.PHI)))
        g.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert not np.array_equal(out[0][0], out[2][0])
