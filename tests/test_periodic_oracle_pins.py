"""Pins of the oracle's periodic domains (BC type 3 on both faces of a direction):
Copier::exchangeDefine with the periodic images of the boxes (Copier.cpp:1464-1541
over the NeighborIterator's images), no BC on periodic faces (BCFunc.cpp:58-158).

- On a fully periodic grid, products of shifted cosines are discrete eigenfunctions
  of the 7-point Laplacian: L cos(w x + p) = (2 cos(w h) - 2) / h^2 cos(w x + p) in
  each direction, so applyOp and a converged Helmholtz solve have closed forms.  A
  ghost left to a boundary rule, or filled from the wrong side, breaks them at the
  faces (the phases make the field asymmetric).
- With periodic y and z and Dirichlet x, the solution of a y-translated rhs is the
  y-translated solution (to the solve tolerance), on layouts of several boxes.
- A refined level within 2 coarse cells of a periodic face is rejected.
"""
import numpy as np
import pytest

import hierarchy as H
from oracle import Oracle


def _periodic_single(n, box, alpha, beta, bc=(3, 3, 3)):
    dom = H.box((0, 0, 0), (n - 1,) * 3)
    return H.Hierarchy(dom, [H.domain_split(dom, box, 4)], 1.0 / n, alpha, beta, bc_type_lo=bc, bc_type_hi=bc)


def _eigen(h, modes=(1, 2, 3), phases=(0.3, -0.7, 0.1)):
    n = h.domain[1][0] + 1
    w = [2 * np.pi * m for m in modes]
    f = h.pack(lambda x, y, z, l: np.cos(w[0] * x + phases[0]) * np.cos(w[1] * y + phases[1])
               * np.cos(w[2] * z + phases[2]))
    hh = 1.0 / n
    lam = sum((2 * np.cos(wi * hh) - 2) / hh ** 2 for wi in w)
    return f, lam


@pytest.mark.parametrize("box", [16, 8])
def test_apply_on_periodic_eigenfunction(box):
    h = _periodic_single(16, box, 0.6, 1.3)
    f, lam = _eigen(h)
    o = Oracle(h)
    out = o.level_apply(0, f, homog=True)
    expect = (0.6 + 1.3 * lam) * f
    assert np.abs(out - expect).max() <= 1e-11 * np.abs(expect).max()
    # the same operator with Dirichlet faces differs at the boundary: the pin sees the ghosts
    hd = _periodic_single(16, box, 0.6, 1.3, bc=(1, 1, 1))
    od = Oracle(hd)
    assert np.abs(od.level_apply(0, f, homog=True) - expect).max() > 1e-3 * np.abs(expect).max()
    o.close()
    od.close()


def test_helmholtz_solve_on_periodic_eigenfunction():
    # alpha I + beta L with alpha > 0, beta < 0 is positive definite on the torus
    h = _periodic_single(32, 16, 1.0, -0.01)
    f, lam = _eigen(h, modes=(1, 1, 2))
    sp, rp = H.solver_params(eps=1e-12, hang=1e-15, iterMax=60)
    o = Oracle(h)
    phi, hist, info = o.solve(f, sp, rp)
    o.close()
    assert info["exit_status"] & 1
    exact = f / (1.0 - 0.01 * lam)
    assert np.abs(phi - exact).max() <= 1e-9 * np.abs(exact).max()


def _grid(h, v):
    """valid-packed single-level field -> (n, n, n) array"""
    n = h.domain[1][0] + 1
    out = np.zeros((n, n, n))
    off = 0
    for (lo, hi) in h.levels[0]:
        shp = tuple(hi[d] - lo[d] + 1 for d in range(3))
        m = int(np.prod(shp))
        out[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = v[off:off + m].reshape(shp, order="F")
        off += m
    return out


def _packed(h, g):
    parts = []
    for (lo, hi) in h.levels[0]:
        parts.append(g[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1].ravel(order="F"))
    return np.concatenate(parts)


def test_mixed_periodic_dirichlet_solve_is_translation_equivariant():
    h = H.Hierarchy(H.box((0, 0, 0), (31, 31, 31)), [H.domain_split(H.box((0, 0, 0), (31, 31, 31)), 16, 4)],
                    1.0 / 32, 0.0, 1.0, bc_type_lo=(1, 3, 3), bc_type_hi=(1, 3, 3))
    rng = np.random.default_rng(5)
    g = rng.standard_normal((32, 32, 32))
    sp, rp = H.solver_params(eps=1e-12, hang=1e-15, iterMax=80)
    o = Oracle(h)
    phi, _, info = o.solve(_packed(h, g), sp, rp)
    phis, _, infos = o.solve(_packed(h, np.roll(g, 5, axis=1)), sp, rp)
    o.close()
    assert info["exit_status"] & 1 and infos["exit_status"] & 1
    a, b = np.roll(_grid(h, phi), 5, axis=1), _grid(h, phis)
    assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max()
    # a shift in x (Dirichlet) is not a symmetry: the pin separates the two kinds of face
    phix, _, _ = Oracle(h).solve(_packed(h, np.roll(g, 5, axis=0)), sp, rp)
    assert np.abs(np.roll(_grid(h, phi), 5, axis=0) - _grid(h, phix)).max() > 1e-3 * np.abs(a).max()


def test_refined_level_near_periodic_face_rejected():
    dom = H.box((0, 0, 0), (31, 31, 31))
    fine_ok = H.domain_split(H.box((8, 8, 8), (47, 47, 47)), 16, 4)     # 4 fine cells from the faces of 0..63
    fine_bad = H.domain_split(H.box((0, 8, 8), (31, 47, 47)), 16, 4)    # touches x = 0
    for fine, ok in ((fine_ok, True), (fine_bad, False)):
        h = H.Hierarchy(dom, [H.domain_split(dom, 16, 4), fine], 1.0 / 32, bc_type_lo=(3, 1, 1),
                        bc_type_hi=(3, 1, 1))
        if ok:
            Oracle(h).close()
        else:
            with pytest.raises(Exception):
                Oracle(h)
