"""Pins of the oracle's multi-component fields (AMRPoissonOp on LevelData<FArrayBox>
with nComp > 1; every ChF kernel loops "do n = 0, ncomp-1", e.g.
AMRPoissonOpF.ChF:158, and ConstDiriNeumBC / DiriBC fill every component,
BCFunc.cpp:58-158, 291-354).

* reduction orders, brute force: LevelDataOps::dotProduct (FArrayBox::dotProduct's
  ForAllThisCBNNXC: one running sum per box, component outermost, FArrayBox.cpp:260-280)
  and norm(BoxLayoutData, p) (BoxLayoutData.cpp:23-92 over FArrayBox::sumPow / norm,
  whose ForAllThisCPencil loops components outermost into one per-column array,
  FArrayBox.cpp:110-230), recomputed here with sequential accumulations in that order;
* exactness: with one component zero (rhs 0, homogeneous boundary values), every dot
  product and norm adds exact zeros, so the other component is bit for bit the
  one-component solve and the zero one stays 0; with rhs (f, -f) the second component
  is exactly the negative of the first (negation commutes with every operation);
* convergence to the right components: two quadratics with different order-2
  Dirichlet values per component (a BCValueFunc filling both a_values entries) are
  each reproduced by a two-component solve (testAMRPoissonOp.H:50-130's parabola test,
  per component);
* the composite operator and residual act per component: bit for bit the
  one-component results, the residual's max norm the larger of the two.
"""
import numpy as np
import pytest

import hierarchy as H
from helpers import comps_pack, comps_unpack, lshape_amr, mixed_ratio_amr, small_amr, small_single, uncovered_mask
from oracle import Oracle


def _seq(x):
    """x[0] + x[1] + ... left to right, from 0.0."""
    return float(np.add.accumulate(np.concatenate([[0.0], np.ravel(x)]))[-1])


def _box_segments(boxes, v, nc):
    o = 0
    for b in boxes:
        n = H.box_npts(b)
        yield b, v[o:o + nc * n].reshape(nc, n)
        o += nc * n


def brute_dot(boxes, a, b, nc):
    sums = []
    for (bx, sa), (_, sb) in zip(_box_segments(boxes, a, nc), _box_segments(boxes, b, nc)):
        sums.append(_seq(sa.ravel() * sb.ravel()))
    return _seq(sums)


def brute_norm(boxes, x, nc, p):
    if p == 0:
        return float(np.abs(x).max())
    sums = []
    for bx, seg in _box_segments(boxes, x, nc):
        nx = H.box_shape(bx)[0]
        rows = seg.reshape(-1, nx)                 # pencils: component, then z, then y
        t = np.abs(rows) if p == 1 else rows * rows
        cols = np.add.accumulate(t, axis=0)[-1]    # the first pencil sets, the others add
        sums.append(float(np.add.accumulate(cols)[-1]))
    s = _seq(sums)
    return s if p == 1 else float(np.sqrt(s))


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("depth", [0, 1])
def test_ncomp_dot_and_norms_in_fab_order(nc, depth):
    h = small_single(n=32, box=16)
    o = Oracle(h)
    o.set_ncomp(nc)
    boxes = [H.box_coarsen(b, 2 ** depth) for b in h.levels[0]]
    n = sum(H.box_npts(b) for b in boxes) * nc
    rng = np.random.default_rng(401 + nc + depth)
    a = rng.standard_normal(n) * np.exp(rng.uniform(-6, 6, n))
    b = rng.standard_normal(n)
    assert o.level_dot(depth, a, b) == brute_dot(boxes, a, b, nc)
    for p in (0, 1, 2):
        assert o.level_norm(depth, a, p) == brute_norm(boxes, a, nc, p), p
    # the planted mistakes: components innermost, or a separate column array per component
    inner = np.concatenate([seg.T.ravel() for _, seg in _box_segments(boxes, a * b, nc)])
    assert o.level_dot(depth, a, b) != _seq(inner) or nc == 1
    assert o.level_norm(depth, b, 1) == brute_norm(boxes, b, nc, 1)
    per = _seq([_seq([np.add.accumulate(np.add.accumulate(np.abs(r).reshape(-1, H.box_shape(bx)[0]),
                                                        axis=0)[-1])[-1] for r in seg])
                for bx, seg in _box_segments(boxes, b, nc)])
    assert o.level_norm(depth, b, 1) != per
    o.close()


CASES = {"amr32": small_amr, "lshape": lshape_amr}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("zero_first", [False, True])
def test_ncomp_zero_component_is_exact(case, zero_first):
    h = CASES[case]()
    f = H.bump_rhs(h, 1101)
    z = np.zeros_like(f)
    sp, rp = H.solver_params(iterMax=40)
    o1 = Oracle(h)
    phi1, hist1, info1 = o1.solve(f, sp, rp)
    o1.close()
    o = Oracle(h)
    o.set_ncomp(2)
    rhs = comps_pack(h, [z, f] if zero_first else [f, z])
    phi, hist, info = o.solve(rhs, sp, rp)
    o.close()
    c0, c1 = comps_unpack(h, phi, 2)
    got, zero = (c1, c0) if zero_first else (c0, c1)
    assert info["iterations"] == info1["iterations"] and info["exit_status"] == info1["exit_status"]
    np.testing.assert_array_equal(hist, hist1)
    np.testing.assert_array_equal(got, phi1)
    assert not zero.any()


def test_ncomp_negated_component_is_exact():
    h = small_amr()
    f = H.bump_rhs(h, 1103)
    sp, rp = H.solver_params(iterMax=40)
    o = Oracle(h)
    o.set_ncomp(2)
    phi, hist, info = o.solve(comps_pack(h, [f, -f]), sp, rp)
    o.close()
    c0, c1 = comps_unpack(h, phi, 2)
    assert info["exit_status"] & 1
    np.testing.assert_array_equal(c1, -c0)
    assert np.abs(c0).max() > 0


@pytest.mark.parametrize("ratios", [(2, 2), (4, 2)])
def test_ncomp_two_quadratics_reproduced(ratios):
    """Component 0: x^2 + y^2 + z^2 (rhs 6); component 1: x^2 - 2 y^2 + z^2 + 3 x - y
    + 0.5 (rhs 0); each with its own order-2 Dirichlet values from one value function.
    (No mixed terms: QuadCFStencil drops the mixed derivative where its stencil does
    not fit, QuadCFStencil.cpp:793-869, so xz is not interpolated exactly there.)"""
    h = mixed_ratio_amr(ratios, (2, 2, 2))
    q0 = lambda x, y, z: x * x + y * y + z * z
    q1 = lambda x, y, z: x * x - 2 * y * y + z * z + 3 * x - y + 0.5

    def diri(pos, d, s):
        return (q0(*pos), q1(*pos))

    o = Oracle(h)
    o.set_ncomp(2)
    o.set_bc_values(diri)
    rhs = comps_pack(h, [h.pack(lambda x, y, z, l: 6.0 + 0 * x), h.pack(lambda x, y, z, l: 0.0 * x)])
    sp, rp = H.solver_params(eps=1e-11, iterMax=150)
    phi, hist, info = o.solve(rhs, sp, rp)
    o.close()
    assert info["exit_status"] & 1
    unc = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    c0, c1 = comps_unpack(h, phi, 2)
    assert np.abs(c0 - h.pack(lambda x, y, z, l: q0(x, y, z)))[unc].max() < 1e-9
    assert np.abs(c1 - h.pack(lambda x, y, z, l: q1(x, y, z)))[unc].max() < 1e-9


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("homog", [False, True])
def test_ncomp_operator_and_residual_per_component(case, homog):
    h = CASES[case]()
    h.bc_val_lo = (0.5, -1.0, 2.0)
    rng = np.random.default_rng(1107)
    phis = [rng.standard_normal(h.total_cells()) for _ in range(3)]
    rhss = [rng.standard_normal(h.total_cells()) for _ in range(3)]
    o1 = Oracle(h)
    ops1 = [o1.amr_operator(p, homog) for p in phis]
    res1 = [o1.amr_residual(p, r, homog) for p, r in zip(phis, rhss)]
    o1.close()
    o = Oracle(h)
    o.set_ncomp(3)
    out = comps_unpack(h, o.amr_operator(comps_pack(h, phis), homog), 3)
    res, nrm = o.amr_residual(comps_pack(h, phis), comps_pack(h, rhss), homog)
    o.close()
    unc = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    for c in range(3):
        np.testing.assert_array_equal(out[c][unc], ops1[c][unc])
    for c, rc in enumerate(comps_unpack(h, res, 3)):
        np.testing.assert_array_equal(rc, res1[c][0])
    assert nrm == max(r[1] for r in res1)


def test_ncomp_rejected():
    h = small_amr()
    o = Oracle(h)
    with pytest.raises(RuntimeError, match="ncomp"):
        o.set_ncomp(0)
    with pytest.raises(RuntimeError, match="ncomp"):
        o.set_ncomp(17)
    a, b = H.vc_coefficients(h, seed=3)
    o.vc_set(a, b)
    with pytest.raises(RuntimeError, match="ncomp"):
        o.set_ncomp(2)
    o.close()
