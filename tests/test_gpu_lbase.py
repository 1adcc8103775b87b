"""AMRMultiGrid::solve(phi, rhs, l_max, l_base) with l_base > 0 on the HIP path
against the oracle: only levels l_base..l_max are solved, the phi of level l_base-1
gives the coarse-fine boundary values, and level l_base cycles with its own
MultiGrid (MGnewOp on that AMR level) down to the BiCGStab bottom.  Histories and
solutions bitwise, on merged and per-box levels and ratio-4 hierarchies; a full
solve afterwards is unchanged; the parabola test on the finer levels alone; the
inputs it rejects."""
import numpy as np
import pytest

import amrmg
import hierarchy as H
from helpers import lshape_amr, mixed_ratio_amr, small_amr, uncovered_mask
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _slices(h):
    offs = h.level_offsets()
    return [slice(offs[l], offs[l + 1]) for l in range(h.nlev)]


CASES = {
    "amr32": small_amr,
    "lshape": lshape_amr,          # level 1 stored per box
    "ratio42": lambda: mixed_ratio_amr((4, 2)),
    "ratio24_helm": lambda: mixed_ratio_amr((2, 4), (0, 1, 2), 0.5, 1.2),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("l_base,l_max", [(1, 2), (1, 1), (2, 2)])
def test_lbase_solve_bitwise(gpu_ctx, case, l_base, l_max):
    h = CASES[case]()
    rng = np.random.default_rng(73)
    rhs = H.bump_rhs(h, 1073)
    sl = _slices(h)
    phi0 = np.zeros(h.total_cells())
    for l in range(h.nlev):        # coarse-fine data below l_base, untouched data above l_max
        if not (l_base <= l <= l_max):
            phi0[sl[l]] = rng.standard_normal(h.cells(l)) * 1e-2
    sp, rp = H.solver_params(iterMax=60)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    g.set_field(g.PHI, phi0)
    hist, info = g.solve(zero_phi=True, l_max=l_max, l_base=l_base)
    phi = g.get_field(g.PHI)
    o = Oracle(h)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp, phi0=phi0, l_max=l_max, l_base=l_base, zero_phi=True)
    assert info["iterations"] == info_o["iterations"] and info["exit_status"] == info_o["exit_status"]
    assert info["exit_status"] & 1
    np.testing.assert_array_equal(hist, hist_o)
    np.testing.assert_array_equal(phi, phi_o)
    for l in range(h.nlev):
        if not (l_base <= l <= l_max):
            np.testing.assert_array_equal(phi[sl[l]], phi0[sl[l]])
    # a whole-hierarchy solve afterwards (the bottom solver binds back to level 0)
    g.set_field(g.PHI, np.zeros(h.total_cells()))
    hist2, info2 = g.solve()
    phi2 = g.get_field(g.PHI)
    g.close()
    phi2_o, hist2_o, info2_o = o.solve(rhs, sp, rp)
    o.close()
    assert info2["iterations"] == info2_o["iterations"]
    np.testing.assert_array_equal(hist2, hist2_o)
    np.testing.assert_array_equal(phi2, phi2_o)


def test_lbase_parabola(gpu_ctx):
    """The coarser level holds the sampled parabola; solving levels 1..2 alone for
    rhs 6 with the parabola's order-2 Dirichlet values returns it (testAMRPoissonOp.H)."""
    h = mixed_ratio_amr((4, 2), (2, 2, 2))

    def diri(pos, d, s):
        return pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]

    exact = h.pack(lambda x, y, z, l: x * x + y * y + z * z)
    rhs = h.pack(lambda x, y, z, l: 6.0 + 0 * x)
    sl = _slices(h)
    phi0 = exact.copy()
    phi0[sl[1].start:] = 0.0
    sp, rp = H.solver_params(eps=1e-11, iterMax=100)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.set_bc_values(diri)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    g.set_field(g.PHI, phi0)
    _, info = g.solve(zero_phi=True, l_base=1)
    phi = g.get_field(g.PHI)
    g.close()
    assert info["exit_status"] & 1
    np.testing.assert_array_equal(phi[sl[0]], exact[sl[0]])
    for l in (1, 2):
        m = uncovered_mask(h, l)
        assert np.abs(phi[sl[l]] - exact[sl[l]])[m].max() < 1e-9, l


def test_lbase_rejected_inputs(gpu_ctx):
    h = small_amr()
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    sp, rp = H.solver_params()
    g.setSolverParameters(sp, rp)
    for lmax, lbase in [(1, 2), (3, 1), (2, -1)]:
        with pytest.raises(amrmg.AMRMGError, match="l_base"):
            g.solve(l_max=lmax, l_base=lbase)
    a, b = H.vc_coefficients(h, seed=3)
    g.set_vc_coefs(a, b)
    with pytest.raises(amrmg.AMRMGError, match="variable coefficients"):
        g.solve(l_base=1)
    g.close()
    m = amrmg.AMRMultiGrid(gpu_ctx, h, precision=1)
    m.setSolverParameters(sp, rp)
    with pytest.raises(amrmg.AMRMGError, match="mixed precision"):
        m.solve(l_base=1)
    m.close()
