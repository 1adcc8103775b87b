"""Pins of the oracle's refinement ratio 4 (AMRPoissonOpFactory's refRatios):

* the parabola of lib/test/AMRElliptic/testAMRPoissonOp.H:50-130 with its order-2
  Dirichlet and Neumann value functions: QuadCFInterp (nRef = 4), the flux register
  (nRef = 4: fine fluxes scaled by 1/16, registers at coarsen(fine, 4)) and the BC
  rules are exact on quadratics, so the composite operator of the sampled parabola is
  6 on every uncovered cell of two-level ratio-4 and three-level mixed-ratio
  hierarchies whose refined levels reach the domain boundary, and the solve with that
  right-hand side returns the sampled parabola;
* AVERAGE's ratio-4 branch: the restriction of integer-valued data is the exact mean
  of each 4^3 block;
* PROLONG with m = 4: piecewise-constant injection of each coarse value into its 4^3
  children;
* INTERPHOMO with dxCrse = 4 dx: the ghost is the value of the quadratic through zero
  at the coarse cell centre and the two interior fine cells (Lagrange form).
"""
import numpy as np
import pytest

import hierarchy as H
from helpers import ghosted_random, uncovered_mask
from oracle import Oracle


def parabola(x, y, z, l=0):
    return x * x + y * y + z * z


def diri(pos, d, s):
    return pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]


def neum(pos, d, s):
    return 2.0 * pos[d]


def two_level_r4(bc, alpha=0.0, beta=1.0):
    """[-8, 7]^3 base (dx 1/16); level 1 (ratio 4, domain [-32, 31]^3) touches the
    high x face."""
    dom = H.box((-8,) * 3, (7,) * 3)
    lev1 = H.domain_split(H.box((-8, -16, -16), (31, 15, 15)), 16, 4)
    return H.Hierarchy(dom, [H.domain_split(dom, 8, 4), lev1], 1.0 / 16, alpha, beta, bc_type_lo=bc,
                       bc_type_hi=bc, ratios=(4,))


def three_level(bc, ratios, alpha=0.0, beta=1.0):
    dom = H.box((-8,) * 3, (7,) * 3)
    r1, r2 = ratios
    # level 1 covers coarse [-4, 7] x [-4, 3]^2 (touching the high x face)
    lev1_box = H.box_refine(H.box((-4, -4, -4), (7, 3, 3)), r1)
    # level 2 covers level-1 cells x in [2 r1, 6 r1 - 1] ... up to the high x face
    d1 = H.box_refine(dom, r1)
    lev2_coarse = H.box((d1[1][0] - 4 * r1 + 1, -2 * r1, -2 * r1), (d1[1][0], 2 * r1 - 1, 2 * r1 - 1))
    lev2_box = H.box_refine(lev2_coarse, r2)
    return H.Hierarchy(dom, [H.domain_split(dom, 8, 4), H.domain_split(lev1_box, 16, 4),
                             H.domain_split(lev2_box, 16, 4)], 1.0 / 16, alpha, beta,
                       bc_type_lo=bc, bc_type_hi=bc, ratios=ratios)


CASES = {
    "r4_diri2": (lambda: two_level_r4((2, 2, 2)), diri),
    "r4_neum_helm": (lambda: two_level_r4((0, 0, 0), 1.0, 1.0), neum),
    "r42_diri2": (lambda: three_level((2, 2, 2), (4, 2)), diri),
    "r24_diri2": (lambda: three_level((2, 2, 2), (2, 4)), diri),
    "r44_diri2_helm": (lambda: three_level((2, 2, 2), (4, 4), 1.0, 1.0), diri),
}


def _unc(h):
    return np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])


@pytest.mark.parametrize("case", list(CASES))
def test_ratio4_parabola_operator_is_six(case):
    mk, fn = CASES[case]
    h = mk()
    o = Oracle(h)
    o.set_bc_values(fn)
    phi = h.pack(parabola)
    L = o.amr_operator(phi, homog=False) - h.alpha * phi
    unc = _unc(h)
    assert np.abs(L - 6.0 * h.beta)[unc].max() < 1e-8
    o.close()


@pytest.mark.parametrize("case", list(CASES))
def test_ratio4_parabola_solve(case):
    mk, fn = CASES[case]
    h = mk()
    o = Oracle(h)
    o.set_bc_values(fn)
    rhs = h.pack(lambda x, y, z, l: h.alpha * parabola(x, y, z) + h.beta * 6.0)
    sp, rp = H.solver_params(eps=1e-11, iterMax=150)
    phi, hist, info = o.solve(rhs, sp, rp)
    o.close()
    assert info["exit_status"] & 1
    exact = h.pack(parabola)
    unc = _unc(h)
    assert np.abs(phi - exact)[unc].max() < 1e-9


def _coarse_blocks(h, lev, fine):
    """fine (valid-packed on level lev) as a list of (coarse box, array) per box."""
    out = []
    for b, a in h.unpack_level(fine, lev):
        out.append((H.box_coarsen(b, h.ratio(lev)), a))
    return out


def test_ratio4_average_is_block_mean():
    h = two_level_r4((1, 1, 1))
    o = Oracle(h)
    rng = np.random.default_rng(11)
    n1 = h.cells(1)
    fine = rng.integers(-1000, 1000, n1).astype(np.float64)
    full = np.zeros(h.total_cells())
    offs = h.level_offsets()
    full[offs[1]:offs[2]] = fine
    got = o.average(1, fine)
    want = []
    for cb, a in _coarse_blocks(h, 1, full):
        s = H.box_shape(cb)
        m = a.reshape(4, s[0], 4, s[1], 4, s[2], order="F").sum(axis=(0, 2, 4)) / 64.0
        want.append(m.ravel(order="F"))
    np.testing.assert_array_equal(got, np.concatenate(want))
    o.close()


def test_ratio4_prolong_is_injection():
    h = two_level_r4((1, 1, 1))
    o = Oracle(h)
    rng = np.random.default_rng(12)
    offs = h.level_offsets()
    fine = rng.standard_normal(h.cells(1))
    full = np.zeros(h.total_cells())
    full[offs[1]:offs[2]] = fine
    coarse_parts, want = [], []
    for (b, a) in h.unpack_level(full, 1):
        cb = H.box_coarsen(b, 4)
        c = rng.standard_normal(H.box_shape(cb))
        coarse_parts.append(c.ravel(order="F"))
        up = np.repeat(np.repeat(np.repeat(c, 4, 0), 4, 1), 4, 2)
        want.append((a + up).ravel(order="F"))
    got = o.amr_prolong(1, fine, np.concatenate(coarse_parts))
    np.testing.assert_array_equal(got, np.concatenate(want))
    o.close()


def test_ratio4_homogeneous_cf_is_lagrange():
    h = two_level_r4((1, 1, 1))
    o = Oracle(h)
    rng = np.random.default_rng(13)
    g = ghosted_random(h, 1, rng)
    out = o.homog_cf(1, g)
    dx = h.dx(1)
    dxc = 4 * dx
    # nodes along the outward normal: coarse centre (value 0), first and second
    # interior fine cells; evaluated at the ghost
    a, b, c, x = -dxc / 2, dx / 2, 3 * dx / 2, -dx / 2
    wb = (x - a) * (x - c) / ((b - a) * (b - c))
    wc = (x - a) * (x - b) / ((c - a) * (c - b))
    # per box, the ghost faces that lie on the coarse-fine boundary (inside the
    # domain and outside every level-1 box)
    dom = h.level_domain(1)
    boxes = h.levels[1]
    off = 0
    checked = 0
    for bx in boxes:
        gb = H.box_grow(bx, 1)
        shp = H.box_shape(gb)
        n = int(np.prod(shp))
        A_in = g[off:off + n].reshape(shp, order="F")
        A_out = out[off:off + n].reshape(shp, order="F")
        for d in range(3):
            for side in (0, 1):
                face = bx[1][d] + 1 if side else bx[0][d] - 1
                if face < dom[0][d] or face > dom[1][d]:
                    continue
                inward = -1 if side else 1
                t = [range(bx[0][q], bx[1][q] + 1) for q in range(3)]
                t[d] = [face]
                for i in t[0]:
                    for j in t[1]:
                        for k in t[2]:
                            p = (i, j, k)
                            if any(all(b0[0][q] <= p[q] <= b0[1][q] for q in range(3)) for b0 in boxes):
                                continue
                            loc = tuple(p[q] - gb[0][q] for q in range(3))
                            n1 = list(loc); n1[d] += inward
                            n2 = list(loc); n2[d] += 2 * inward
                            want = wb * A_in[tuple(n1)] + wc * A_in[tuple(n2)]
                            assert abs(A_out[loc] - want) <= 1e-13 * (1 + abs(want))
                            checked += 1
        off += n
    assert checked > 0
    o.close()
