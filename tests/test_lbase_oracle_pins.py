"""Pins of the oracle's AMRMultiGrid::solve(phi, rhs, l_max, l_base) with l_base > 0
(AMRMultiGrid.H:895-1124, computeAMRResidualLevel :837-892): only levels
l_base..l_max are solved, level l_base-1's phi gives the coarse-fine boundary values,
and level l_base relaxes with its own MultiGrid (MGnewOp on the AMR level,
AMRPoissonOp.cpp:1923-1989) down to a BiCGStab bottom.

* the parabola of testAMRPoissonOp.H:50-130: with the coarser levels holding the
  sampled parabola and the order-2 Dirichlet parabola values on the domain faces,
  solving only the finer levels for rhs 6 returns the sampled parabola there
  (QuadCFInterp is exact on quadratics), on one and on two levels above the base,
  ratio 2 and 4;
* the composite solution restricted: with level l_base-1 fixed at the converged
  composite solution, the solve of levels l_base..l_max from zero converges to the
  composite solution there;
* levels outside l_base..l_max come back unchanged, bit for bit.
"""
import numpy as np
import pytest

import hierarchy as H
from helpers import mixed_ratio_amr, small_amr, uncovered_mask
from oracle import Oracle


def parabola(x, y, z, l=0):
    return x * x + y * y + z * z


def diri(pos, d, s):
    return pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]


def _slices(h):
    offs = h.level_offsets()
    return [slice(offs[l], offs[l + 1]) for l in range(h.nlev)]


@pytest.mark.parametrize("ratios,l_base,l_max", [((2, 2), 1, 2), ((2, 2), 1, 1), ((2, 2), 2, 2),
                                                 ((4, 2), 1, 2), ((2, 4), 2, 2)])
def test_lbase_parabola(ratios, l_base, l_max):
    h = mixed_ratio_amr(ratios, (2, 2, 2))
    o = Oracle(h)
    o.set_bc_values(diri)
    exact = h.pack(parabola)
    rhs = h.pack(lambda x, y, z, l: 6.0 + 0 * x)
    sl = _slices(h)
    phi0 = exact.copy()
    for l in range(l_base, l_max + 1):
        phi0[sl[l]] = 0.0
    sp, rp = H.solver_params(eps=1e-11, iterMax=100)
    phi, hist, info = o.solve(rhs, sp, rp, phi0=phi0, l_max=l_max, l_base=l_base, zero_phi=True)
    o.close()
    assert info["exit_status"] & 1
    for l in range(h.nlev):
        if l_base <= l <= l_max:
            m = uncovered_mask(h, l) if l < l_max else np.ones(h.cells(l), bool)
            assert np.abs(phi[sl[l]] - exact[sl[l]])[m].max() < 1e-9, l
        else:
            np.testing.assert_array_equal(phi[sl[l]], phi0[sl[l]])


@pytest.mark.parametrize("l_base", [1, 2])
def test_lbase_reproduces_composite_solution(l_base):
    h = small_amr()
    rhs = H.bump_rhs(h, 1071)
    sp, rp = H.solver_params(eps=1e-12, iterMax=60)
    o = Oracle(h)
    comp, _, info = o.solve(rhs, sp, rp)
    assert info["exit_status"] & 1
    sl = _slices(h)
    phi0 = comp.copy()
    for l in range(l_base, h.nlev):
        phi0[sl[l]] = 0.0
    phi, hist, info2 = o.solve(rhs, sp, rp, phi0=phi0, l_base=l_base, zero_phi=True)
    o.close()
    assert info2["exit_status"] & 1
    scale = np.abs(comp).max()
    for l in range(h.nlev):
        if l >= l_base:
            m = uncovered_mask(h, l)
            assert np.abs(phi[sl[l]] - comp[sl[l]])[m].max() < 1e-9 * scale, l
        else:
            np.testing.assert_array_equal(phi[sl[l]], comp[sl[l]])
