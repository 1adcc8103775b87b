"""Pins of the oracle's boundary value functions (DiriBC / NeumBC with a
BCValueHolder, BCFunc.cpp:217-258 and 291-354, the value taken at
getDomainFacePosition, BCFunc.cpp:187-202):

* the parabola of lib/test/AMRElliptic/testAMRPoissonOp.H:50-130 (domain [-N, N-1]^3,
  dx = 1/(2N)): Parabola_diri |x|^2 with the order-2 Dirichlet rule and Parabola_neum
  2 x_dir with the Neumann rule both reproduce quadratics exactly, so the operator of
  the sampled parabola is 2*3 = 6 on every cell, on one level and on a three-level
  hierarchy whose refined levels touch the domain boundary (QuadCFInterp and reflux
  are exact on quadratics as well), and the multigrid solve with that right-hand side
  returns the sampled parabola on the uncovered cells;
* the position convention: the same functions evaluated half a cell inward (the
  valid cell's centre instead of its face) break that exactness;
* dir / side order: a function returning the face constants per (dir, side)
  reproduces the constant-value boundary conditions bitwise.
"""
import numpy as np
import pytest

import hierarchy as H
from helpers import uncovered_mask
from oracle import Oracle

N = 8


def parabola(x, y, z, l=0):
    return x * x + y * y + z * z


def diri(pos, d, s):          # Parabola_diri
    return pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]


def neum(pos, d, s):          # Parabola_neum
    return 2.0 * pos[d]


def _single(bc, alpha=0.0, beta=1.0, box=8):
    dom = H.box((-N,) * 3, (N - 1,) * 3)
    return H.Hierarchy(dom, [H.domain_split(dom, box, 4)], 1.0 / (2 * N), alpha, beta, bc_type_lo=bc,
                       bc_type_hi=bc)


def _amr(bc, alpha=0.0, beta=1.0):
    """[-8, 7]^3 base; level 1 covers x >= -8, y, z in [-8, 7] of its index space
    and touches the high x face; level 2 touches it too."""
    dom = H.box((-N,) * 3, (N - 1,) * 3)
    lev1 = H.domain_split(H.box((-8, -8, -8), (15, 7, 7)), 8, 4)
    lev2 = H.domain_split(H.box((8, -8, -8), (31, 7, 7)), 8, 4)
    return H.Hierarchy(dom, [H.domain_split(dom, 8, 4), lev1, lev2], 1.0 / (2 * N), alpha, beta,
                       bc_type_lo=bc, bc_type_hi=bc)


CASES = {
    "diri2_single": (lambda: _single((2, 2, 2)), diri),
    "neum_single": (lambda: _single((0, 0, 0)), neum),
    "diri2_amr": (lambda: _amr((2, 2, 2)), diri),
    "neum_amr": (lambda: _amr((0, 0, 0)), neum),
}


def _mixed_fn(pos, d, s):
    return neum(pos, d, s) if d == 1 else diri(pos, d, s)


@pytest.mark.parametrize("case", list(CASES) + ["mixed_single"])
def test_parabola_operator_is_six(case):
    if case == "mixed_single":
        h, fn = _single((2, 0, 2)), _mixed_fn
    else:
        mk, fn = CASES[case]
        h = mk()
    o = Oracle(h)
    o.set_bc_values(fn)
    phi = h.pack(parabola)
    # uncovered cells: the reflux registers also sit on coarse cells that another
    # fine box covers (LevelFluxRegister.cpp:169-201), so covered cells are not 6
    unc = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    L = o.amr_operator(phi, homog=False)
    assert np.abs(L - 6.0)[unc].max() < 1e-9
    # the valid cell's centre in place of its face centre: not exact on the boundary
    inward = lambda pos, d, s: fn(tuple(p - (0.5 * (2 * s - 1) * h.dx(0) if q == d else 0.0)
                                        for q, p in enumerate(pos)), d, s)
    o.set_bc_values(inward)
    assert np.abs(o.amr_operator(phi, homog=False) - 6.0)[unc].max() > 0.1
    o.close()


@pytest.mark.parametrize("case", list(CASES))
def test_parabola_solve(case):
    mk, fn = CASES[case]
    h = mk()
    alpha, beta = (0.0, 1.0) if "diri" in case else (1.0, 1.0)   # pure Neumann needs alpha != 0
    h.alpha, h.beta = alpha, beta
    o = Oracle(h)
    o.set_bc_values(fn)
    rhs = h.pack(lambda x, y, z, l: alpha * parabola(x, y, z) + beta * 6.0)
    sp, rp = H.solver_params(eps=1e-13, iterMax=40)
    phi, hist, info = o.solve(rhs, sp, rp)
    o.close()
    assert info["exit_status"] & 1
    exact = h.pack(parabola)
    offs = h.level_offsets()
    for l in range(h.nlev):
        m = uncovered_mask(h, l)
        err = np.abs(phi[offs[l]:offs[l + 1]] - exact[offs[l]:offs[l + 1]])[m]
        assert err.max() < 1e-9, (l, err.max())


def test_face_constants_reproduce_constant_bc():
    lo, hi = (0.25, -1.5, 2.0), (0.75, 3.0, -0.5)
    bc = (1, 0, 2)
    dom = H.box((0, 0, 0), (15, 15, 15))
    h = H.Hierarchy(dom, [H.domain_split(dom, 8, 4)], 1.0 / 16, 0.3, 1.0, bc_type_lo=bc, bc_type_hi=bc,
                    bc_val_lo=lo, bc_val_hi=hi)
    rng = np.random.default_rng(3)
    phi = rng.standard_normal(h.total_cells())
    a = Oracle(h)
    ref = a.amr_operator(phi, homog=False)
    b = Oracle(h)
    b.set_bc_values(lambda pos, d, s: (hi if s else lo)[d])
    np.testing.assert_array_equal(b.amr_operator(phi, homog=False), ref)
    # and the homogeneous operator ignores the function
    np.testing.assert_array_equal(b.amr_operator(phi, homog=True), a.amr_operator(phi, homog=True))
    b.set_bc_values(None)
    np.testing.assert_array_equal(b.amr_operator(phi, homog=False), ref)
    a.close()
    b.close()
